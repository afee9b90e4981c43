"""Benchmark: light curves per hour of the batched float32-compensated TLS search on one GPU.

Workload (synthetic, seeded): 4096 light curves, N log-uniform in [2000, 70000] points,
cadence 2 or 30 min, noise U(1e-4, 2e-3), half with an injected limb-darkened transit,
each curve with its own Ofir period grid, duration grid and template table (default TLS
settings, T0_fit_margin 0.125). A step is one pass of the search over the whole batch:
the persistent chi2-vs-period search kernel over every (curve, period) plus the device
SR/SDE spectra of every curve. Inputs are resident in HBM before the timed region.

Also measured in the same run (reported beside `value`, never as `value`):
  plain    -- the float64 search one curve per launch with host spectra, on a stratified
              sample of curves, extrapolated to the batch by work (sum N' x periods x widths)
  parity   -- on that sample: same best period (SDE peak) and same best width, SDE error
  cpu      -- the FP64 CPU oracle (OpenMP, all host cores) on a bounded sample of curves,
              extrapolated the same way

    python bench.py [--steps K] [--warmup W] [--curves 4096]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def curve_work(plan):
    """sum over periods of N' * (number of searched widths): the search's work unit."""
    ok = plan.k_lo <= plan.k_hi
    wmax = plan.table.widths[np.clip(plan.k_hi, 0, None)]
    nw = np.where(ok, plan.k_hi - plan.k_lo + 1, 0)
    return float(np.sum((plan.n + wmax) * nw * ok))


def algorithmic_bytes(plans, real_bytes=4):
    """Algorithmic HBM bytes of one search launch: per (curve, period) the permutation
    (4 B) and the gathered flux deviation and weight (2 x real_bytes) for each of the
    N' = N + widest-width samples, and 12 B of result per searched width."""
    total = 0.0
    for p in plans:
        ok = p.k_lo <= p.k_hi
        wmax = p.table.widths[np.clip(p.k_hi, 0, None)]
        nw = np.where(ok, p.k_hi - p.k_lo + 1, 0)
        total += float(np.sum(((4 + 2 * real_bytes) * (p.n + wmax) + 12 * nw) * ok))
    return total


def stratified(plans, k, seed=0):
    """k curves spread over the work distribution (sorted by work, evenly spaced)."""
    w = np.array([curve_work(p) for p in plans])
    order = np.argsort(w)
    idx = order[np.linspace(0, len(order) - 1, k).astype(int)]
    return sorted(set(int(i) for i in idx))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--curves", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--plain-sample", type=int, default=24)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the TLS search of a batch runs on a single device "
              "(see DESIGN.md for how it would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from gtls_amd import batch as B
    from gtls_amd import synthetic
    from gtls_amd import search as S
    from gtls_amd import spectra as SP
    from gtls_amd.grid import skip_factor

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)

    t0 = time.time()
    curves = synthetic.batch(args.curves, seed=args.seed)
    plans = B.build_plans(curves, {})
    bp = B.BatchPlan(plans)
    log("built %d plans: %d points, %d periods, %.1f s" % (
        len(plans), int(bp.pt_off[-1]), bp.total_periods, time.time() - t0))
    db = B.DeviceBatch(bp, 0)
    skip = skip_factor(0.125)

    def step(ev=None):
        if ev is not None:
            ev[0].record()
        chi2, _, _ = db.search("fp32c", skip)
        if ev is not None:
            ev[1].record()
        db.spectra(chi2, 3, sync=False)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log("warmup step %d done" % i)

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    search_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    ms_per_step = elapsed * 1e3 / args.steps
    value = args.curves * args.steps / elapsed * 3600.0
    log("timed %d steps: %.1f ms/step, search kernel %.1f ms" % (args.steps, ms_per_step,
                                                                search_ms))

    algo = algorithmic_bytes(plans, 4)
    achieved = algo / (search_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "tls_search_kernel<float>", "algorithmic_bytes": algo}

    bk_b = bp.unpack(db.best_k.cpu().numpy())
    spec = db.spectra(db.chi2, 3)          # finishes host fallbacks of the last step
    pw_b, sde_b = spec["power"], spec["SDE"]
    extra = {"device_spectra_curves": int(np.sum(spec["device"]))}
    npd = bp.total_periods

    total_work = sum(curve_work(p) for p in plans)

    # plain float64 search, one curve per launch, host spectra; parity with the batch
    if not args.no_plain:
        sample = stratified(plans, min(args.plain_sample, len(plans)))
        t_plain, w_plain = 0.0, 0.0
        same_period = same_width = 0
        max_dsde = 0.0
        for ci in sample:
            p = plans[ci]
            dc = S.DeviceCurve(p, 0)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            c, k, s = dc.search(p.periods, p.k_lo, p.k_hi, mode="fp64", skip_point=skip)
            c = c.cpu().numpy()
            _, _, pw, _, sd = SP.spectra(SP.mask_outliers(c), 3)
            t_plain += time.perf_counter() - ts
            w_plain += curve_work(p)
            kk = k.cpu().numpy()
            ib, ip = int(np.argmax(np.asarray(pw_b[ci]))), int(np.argmax(np.asarray(pw)))
            same_period += ib == ip
            same_width += (ib == ip) and bk_b[ci][ib] == kk[ip]
            max_dsde = max(max_dsde, abs(float(sde_b[ci]) - float(sd)) / abs(float(sd)))
        plain_value = len(plans) / (t_plain * total_work / w_plain) * 3600.0
        extra.update({"plain_fp64_curves_per_hour": round(plain_value, 1),
                      "speedup_vs_plain_fp64": round(value / plain_value, 3),
                      "plain_sample_curves": len(sample),
                      "parity_same_best_period": "%d/%d" % (same_period, len(sample)),
                      "parity_same_best_width": "%d/%d" % (same_width, len(sample)),
                      "parity_max_rel_sde_err": float(max_dsde)})
        log("plain fp64: %.1f curves/h (sample %d), parity %s %s, max dSDE %.2e" % (
            plain_value, len(sample), extra["parity_same_best_period"],
            extra["parity_same_best_width"], max_dsde))

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        cores = os.cpu_count() or 1
        order = stratified(plans, min(64, len(plans)))
        t_cpu, w_cpu, used = 0.0, 0.0, 0
        for ci in order:
            p = plans[ci]
            if t_cpu > args.cpu_budget:
                break
            ts = time.perf_counter()
            oracle.search_plan(p, skip_point=skip, nthreads=cores)
            t_cpu += time.perf_counter() - ts
            w_cpu += curve_work(p)
            used += 1
        cpu_value = len(plans) / (t_cpu * total_work / w_cpu) * 3600.0
        cpu = {"value": round(cpu_value, 2), "unit": "light curves/hour", "cores": cores,
               "kind": "port",
               "sample": "%d stratified curves of the batch (%.1f s of FP64 oracle search "
                         "on %d threads), extrapolated by sum N' x periods x widths"
                         % (used, t_cpu, cores)}
        log("cpu oracle: %.2f curves/h over %d curves" % (cpu_value, used))

    out = {
        "metric": "light curves/hour, batched FP32-comp vs plain FP64 on 1 GPU; same peak, "
                  "SDE err<1e-4",
        "value": round(value, 2),
        "unit": "light curves/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (seeded light curves with injected limb-darkened transits)",
        "config": {"workload": "batched TLS search, %d curves, N log-U(2000, 70000), "
                               "cadence 2/30 min, own period grid per curve, "
                               "T0_fit_margin 0.125" % args.curves,
                   "curves": args.curves, "total_points": int(bp.pt_off[-1]),
                   "total_periods": int(npd), "mode": "fp32c", "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
