"""Benchmark: light curves per hour of the batched float32-compensated TLS search on one GPU.

Workload (synthetic, seeded): 4096 light curves, N log-uniform in [2000, 70000] points,
cadence 2 or 30 min, noise U(1e-4, 2e-3), half with an injected limb-darkened transit,
each curve with its own Ofir period grid, duration grid and template table (default TLS
settings, T0_fit_margin 0.125). The set is cut into batches of --batch curves (512 by
default); a step is one pass of the search over one batch: the persistent
chi2-vs-period search kernel over every (curve, period) of the batch plus the device
SR/SDE spectra of its curves. Step i searches batch i mod (curves / batch), so the default
8 steps cover the whole 4096-curve set once. Inputs of every batch are resident in HBM
before the timed region.

Also measured in the same run (reported beside `value`, never as `value`):
  plain    -- the float64 search one curve per launch with host spectra, on a stratified
              sample of curves, extrapolated to the set by window work
  parity   -- on that sample: same best period (SDE peak) and same best width, SDE error
  cpu      -- the FP64 CPU oracle (OpenMP, all host cores) on a bounded sample of curves,
              extrapolated the same way, with the same peak/SDE check
  single   -- with plain: the single-curve configurations (the 19,728-point 2-min curve,
              the 4-year 29.4-min curve, the reference tutorial's 41,816-point shape) on
              their whole default grids, one gtls_search call per mode; with cpu: the
              oracle on the same grids (C2 on 1 % of its periods, extrapolated)

    python bench.py [--steps K] [--warmup W] [--curves 4096] [--batch 512]
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

FP32_MATRIX_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix-core peak
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def window_macs(plan, skip):
    """Useful multiply-accumulates of the window sums of one curve: for every period and
    every searched width d (stride q = max(d // skip, 1)), ceil(N / q) positions of d
    template samples each (a second product per sample for per-sample weights is not
    counted)."""
    w = plan.table.widths.astype(np.int64)
    q = np.maximum(w // skip, 1) if skip < (1 << 30) else np.ones_like(w)
    per_width = -(-plan.n // q) * np.maximum(w, 0)
    c = np.concatenate([[0], np.cumsum(per_width)])
    ok = plan.k_lo <= plan.k_hi
    return float(np.sum((c[plan.k_hi + 1] - c[plan.k_lo])[ok]))


def algorithmic_bytes(plan, real_bytes=4):
    """Algorithmic HBM bytes of one curve's search: per period the times (8 B) read for the
    fold, and per searched sample the gathered flux deviation and weight
    (2 x real_bytes) of the N' = N + widest-width samples, plus 16 B of result."""
    ok = plan.k_lo <= plan.k_hi
    wmax = plan.table.widths[np.clip(plan.k_hi, 0, None)]
    return float(np.sum(((8 + 2 * real_bytes) * plan.n + 2 * real_bytes * wmax + 16) * ok))


def profiled_traffic(args):
    """HBM bytes of one step's search call and of its windows_kernel launches as profiled
    (FETCH_SIZE and WRITE_SIZE, copy-scaled) for this workload, read from the committed
    profile summary; None for other workloads."""
    if (args.curves, args.batch, args.seed) != (4096, 512, 4):
        return None, None
    path = os.path.join(ROOT, "profiles", "search_fp32c_c4_batch0_traffic.md")
    keys = {"traffic_bytes_per_launch:": "search_call",
            "traffic_bytes_windows_kernel:": "windows_kernel"}
    out = {}
    try:
        with open(path) as f:
            for line in f:
                for k, name in keys.items():
                    if line.startswith(k):
                        out[name] = float(line.split(":", 1)[1])
    except OSError:
        return None, None
    return (out or None), os.path.relpath(path, ROOT)


def refine_leg(args, plans, batch0, skip):
    """Gradient refinement of grid peaks: the 8 highest power peaks of each of the first
    `refine_curves` curves of batch 0 seed a smooth limb-darkened transit model each, fitted
    by Adam (`refine_steps` steps, float32 model; loss and gradient from the fused HIP
    kernel). Timed twice: eager steps (with the kernels' HIP events), and a HIP graph of
    one step replayed (Refiner.capture: its two warm-up steps, then the replays, timed)."""
    import torch
    from gtls_amd import refine as RF
    bp, db = batch0
    db.search("fp32c", skip)
    spec = db.spectra(db.chi2, 3)
    bks = bp.unpack(db.best_k.cpu().numpy())
    bss = bp.unpack(db.best_s.cpu().numpy())
    log("refine: seeding from the peaks of %d curves" % min(args.refine_curves, args.batch))
    ts = time.perf_counter()
    curves, thetas, mc, spans = [], [], [], []
    for j in range(min(args.refine_curves, args.batch)):
        p = plans[j]
        th, _ = RF.seed_peaks(p.t, p.y, p.dy, p.periods, spec["power"][j], p.table.widths,
                              bks[j], bss[j], 8)
        curves.append((p.t, p.y, p.dy))
        thetas.append(th)
        mc += [len(curves) - 1] * len(th)
        spans += [float(np.ptp(p.t))] * len(th)
    thetas = np.concatenate(thetas)
    mb = RF.ModelBatch.build(curves, mc, 0)
    t_seed = time.perf_counter() - ts
    log("refine: %d models over %d points seeded in %.1f s" % (mb.n_models, mb.slots, t_seed))
    RF.Refiner(mb, thetas, spans).run(2)          # warm-up (kernels, allocator)
    ref = RF.Refiner(mb, thetas, spans)
    torch.cuda.synchronize()
    RF.set_timing(True)
    ts = time.perf_counter()
    first, last = ref.run(args.refine_steps,
                          progress=lambda i: log("refine: step %d done" % i))
    torch.cuda.synchronize()
    dt = time.perf_counter() - ts
    kt = RF.timing_ms()
    RF.set_timing(False)
    ratio = (last / first).cpu().numpy()
    # the same fit as graph replays of one captured step
    refg = RF.Refiner(mb, thetas, spans, capturable=True)
    with torch.no_grad():
        first_g = refg.loss()
    warm = 2
    g = refg.capture(warmup=warm)
    torch.cuda.synchronize()
    replays = max(1, args.refine_steps - warm)
    ts = time.perf_counter()
    for _ in range(replays):
        g.replay()
    torch.cuda.synchronize()
    dtg = time.perf_counter() - ts
    with torch.no_grad():
        last_g = refg.loss()
    ratio_g = (last_g / first_g).cpu().numpy()
    ms_g = dtg * 1e3 / replays
    slots = mb.slots
    out = {"curves": len(curves), "models": int(mb.n_models), "steps": args.refine_steps,
           "model_points": slots, "seconds": round(ms_g * args.refine_steps / 1e3, 3),
           "seed_seconds": round(t_seed, 2),
           "ms_per_step": round(ms_g, 3),
           "step": "HIP graph replay of one captured step (%d replays timed after %d warm-up "
                   "steps)" % (replays, warm),
           "model_points_per_s": round(slots / (ms_g * 1e-3), 1),
           "curves_per_hour": round(len(curves) / (ms_g * 1e-3 * args.refine_steps) * 3600.0, 1),
           "chi2_ratio_median": float(np.median(ratio_g)),
           "chi2_improved_models": "%d/%d" % (int(np.sum(ratio_g < 1)), len(ratio_g)),
           "eager": {"ms_per_step": round(dt * 1e3 / args.refine_steps, 3),
                     "chi2_ratio_median": float(np.median(ratio)),
                     "chi2_improved_models": "%d/%d" % (int(np.sum(ratio < 1)), len(ratio))}}
    # the kernels against the HBM roofline. The fused loss (gtls_transit_chi2_grouped: point
    # tiles, every model of the tile's curve evaluated on them, then the in-order sum of the
    # tile sums) reads each point's time, flux and weight (24 B) once per call; the flux
    # kernels read 8 B of time and move 4 B of flux (forward: written; backward: the
    # upstream gradient, read) per model point.
    unique_points = sum(len(c[0]) for c in curves)
    ms, calls = kt.get("chi2", (0.0, 0))
    if calls:
        gbs = 24.0 * unique_points * calls / (ms * 1e-3) / 1e9
        out["transit_chi2_grouped"] = {
            "kernels": "transit_chi2_tile_kernel + transit_chi2_sum_kernel",
            "ms_per_call": round(ms / calls, 4), "calls": calls,
            "points_per_call": unique_points, "bytes_per_call": 24 * unique_points,
            "hbm_GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "model_points_per_s": round(slots * calls / (ms * 1e-3), 1)}
    for k, name in (("fwd", "transit_fwd_kernel"), ("bwd", "transit_bwd_kernel")):
        ms, calls = kt.get(k, (0.0, 0))
        if calls:
            gbs = 12.0 * slots * calls / (ms * 1e-3) / 1e9
            out[name] = {"ms_per_call": round(ms / calls, 4), "calls": calls,
                         "algorithmic_GBps": round(gbs, 1),
                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    log("refine: %d models, %d steps: eager %.2f ms/step, graph replays %.3f ms/step, "
        "median chi2 ratio %.4f" % (mb.n_models, args.refine_steps, dt * 1e3 / args.refine_steps,
                                    out["ms_per_step"], out["chi2_ratio_median"]))
    return out


# Single-curve workloads searched one curve per call (the plain path): key, description,
# synthetic generator, share of the period grid the CPU oracle times (extrapolated by
# window work when below 1), and published context.
SINGLE_CURVES = (
    ("C1", "19,728-point 2-min curve (27.4 d, one gap), whole default grid", "single_2min",
     1.0, None),
    ("C2", "4-year 29.4-min curve (17 gaps, red noise), whole default grid", "long_30min",
     0.01, None),
    ("tutorial_shape", "41,816 points at 58.85 s over 30.77 d, whole default grid (2,868 "
     "periods of 0.601-15.386 d), the shape of the reference tutorial's Kepler search",
     "kepler_sc_month", 1.0,
     "the reference tutorial (tutorial/1-keplerTest.ipynb:102-162) reports about 1.7 s for "
     "this search with GTLS 0.4.2 on an unstated GPU and about 7 s for TLS on 32 CPU threads"),
)


def single_curve_plans():
    """Host plans (default grids) of the single-curve workloads, with their planning time."""
    from gtls_amd import search as S
    from gtls_amd import synthetic
    from gtls_amd.validate import validate_args, validate_inputs
    out = []
    for key, what, fn, frac, context in SINGLE_CURVES:
        t, y, dy, truth = getattr(synthetic, fn)()
        t, y, dy = validate_inputs(t, y, dy)
        ts = time.perf_counter()
        plan = S.make_plan(t, y, dy, validate_args({}))
        out.append(dict(key=key, what=what, plan=plan, truth=truth, frac=frac,
                        context=context, plan_s=time.perf_counter() - ts))
    return out


def single_gpu_leg(singles, skip):
    """Every single-curve workload's whole grid in one gtls_search call per mode (FP64,
    FP32c), after one untimed call; wall-clock seconds of the call including the upload
    of the per-period arrays. Returns (JSON figures, device results per key)."""
    import torch
    from gtls_amd import search as S
    res, arrays = {}, {}
    for sc in singles:
        p = sc["plan"]
        dc = S.DeviceCurve(p, 0)
        dc.search(p.periods, p.k_lo, p.k_hi, mode="fp32c", skip_point=skip)
        torch.cuda.synchronize()
        r = {"what": sc["what"], "points": int(p.n), "periods": int(len(p.periods)),
             "widths": int(len(p.table.widths)),
             "period_range_d": [round(float(p.periods.min()), 4),
                                round(float(p.periods.max()), 4)],
             "host_planning_s": round(sc["plan_s"], 3)}
        got = {}
        for mode in ("fp64", "fp32c"):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            c, k, s = dc.search(p.periods, p.k_lo, p.k_hi, mode=mode, skip_point=skip)
            torch.cuda.synchronize()
            r[mode + "_search_s"] = round(time.perf_counter() - ts, 4)
            got[mode] = tuple(x.cpu().numpy() for x in (c, k, s))
        # useful window FLOPs and algorithmic bytes of the call, as for the batched set
        r["fp32c_window_tflops"] = round(2 * window_macs(p, skip) / r["fp32c_search_s"] / 1e12, 2)
        r["fp32c_algorithmic_GBps"] = round(algorithmic_bytes(p) / r["fp32c_search_s"] / 1e9, 1)
        c64, k64, s64 = got["fp64"]
        c32, k32, s32 = got["fp32c"]
        ok = np.isfinite(c64)
        r["fp32c_max_rel_chi2_vs_fp64"] = float(np.max(np.abs(c32[ok] - c64[ok]) / c64[ok]))
        r["fp32c_same_index_as_fp64"] = "%d/%d" % (
            int(np.sum(((k32 == k64) & (s32 == s64))[ok])), int(ok.sum()))
        r["best_period_d"] = {m: float(p.periods[int(np.nanargmin(got[m][0]))])
                              for m in ("fp64", "fp32c")}
        if sc["truth"] is not None:
            r["injected_period_d"] = float(sc["truth"]["period"])
        if sc["context"]:
            r["context"] = sc["context"]
        res[sc["key"]] = r
        arrays[sc["key"]] = got["fp64"]
        log("single %s: %d points, %d periods: fp64 %.3f s, fp32c %.3f s, best %.4f d" % (
            sc["key"], p.n, len(p.periods), r["fp64_search_s"], r["fp32c_search_s"],
            r["best_period_d"]["fp32c"]))
        del dc
        torch.cuda.synchronize()
    return res, arrays


def single_cpu_leg(singles, skip, cores, gpu_res, gpu_arrays):
    """The FP64 CPU oracle on every single-curve workload (its whole grid, or an evenly
    spaced share of its periods extrapolated by window work), with the device FP64 result
    checked on the same periods."""
    import dataclasses
    from oracle import oracle
    res = {}
    for sc in singles:
        p = sc["plan"]
        n_per = len(p.periods)
        m = max(1, int(round(sc["frac"] * n_per)))
        sel = np.unique(np.linspace(0, n_per - 1, m).astype(np.int64))
        per, klo, khi = p.periods[sel], p.k_lo[sel], p.k_hi[sel]
        ts = time.perf_counter()
        c, k, s = oracle.search_plan(p, skip_point=skip, periods=per, k_lo=klo, k_hi=khi,
                                     nthreads=cores)
        dt = time.perf_counter() - ts
        sub = dataclasses.replace(p, periods=per, k_lo=klo, k_hi=khi)
        full_s = dt * window_macs(p, skip) / max(window_macs(sub, skip), 1.0)
        r = {"sample_periods": int(len(sel)), "sample_s": round(dt, 3),
             "full_grid_s": round(full_s, 3), "extrapolated": bool(len(sel) < n_per),
             "threads": cores}
        if sc["key"] in gpu_arrays:
            c64, k64, s64 = (x[sel] for x in gpu_arrays[sc["key"]])
            ok = np.isfinite(c)
            r["gpu_fp64_same_index"] = "%d/%d" % (
                int(np.sum(((k64 == k) & (s64 == s))[ok])), int(ok.sum()))
            r["gpu_fp64_max_rel_chi2"] = float(np.max(np.abs(c64[ok] - c[ok]) / c[ok]))
            g = gpu_res[sc["key"]]
            r["gpu_fp64_speedup"] = round(full_s / g["fp64_search_s"], 2)
            r["gpu_fp32c_speedup"] = round(full_s / g["fp32c_search_s"], 2)
        res[sc["key"]] = r
        log("cpu single %s: %d of %d periods in %.2f s -> %.2f s whole grid" % (
            sc["key"], len(sel), n_per, dt, full_s))
    return res


def stratified(work, k):
    """k curve indices spread evenly over the sorted work distribution."""
    order = np.argsort(work, kind="stable")
    idx = order[np.linspace(0, len(order) - 1, min(k, len(order))).astype(int)]
    return list(dict.fromkeys(int(i) for i in idx))


def peak_check(sde_a, pw_a, bk_a, c_b, k_b, oversampling=3):
    """Best period / best width / SDE of a search (c_b, k_b) against the batched result."""
    from gtls_amd import spectra as SP
    _, _, pw, _, sd = SP.spectra(SP.mask_outliers(np.asarray(c_b, np.float64)), oversampling)
    ia, ib = int(np.argmax(np.asarray(pw_a))), int(np.argmax(np.asarray(pw)))
    same_p = ia == ib
    same_w = same_p and int(bk_a[ia]) == int(np.asarray(k_b)[ib])
    return same_p, same_w, abs(float(sde_a) - float(sd)) / abs(float(sd))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def spectra_of(chi2, oversampling=3):
    from gtls_amd import spectra as SP
    _, _, pw, _, sd = SP.spectra(SP.mask_outliers(np.asarray(chi2, np.float64)), oversampling)
    return np.asarray(pw), float(sd)


def compare_searches(bp, a, b, oversampling=3):
    """Per curve of a batch: searches a and b (chi2, best_k per period, packed order) give the
    same best period (SDE peak) and the same best width there; max relative chi2 difference
    over all periods and max relative SDE difference."""
    ca, ka = bp.unpack(a[0]), bp.unpack(a[1])
    cb, kb = bp.unpack(b[0]), bp.unpack(b[1])
    same_p = same_w = 0
    max_rel = max_sde = 0.0
    for i in range(bp.n_curves):
        ok = np.isfinite(cb[i])
        if ok.any():
            max_rel = max(max_rel, float(np.max(np.abs(ca[i][ok] - cb[i][ok]) / np.abs(cb[i][ok]))))
        pa, sa = spectra_of(ca[i], oversampling)
        pb, sb = spectra_of(cb[i], oversampling)
        ia, ib = int(np.argmax(pa)), int(np.argmax(pb))
        same_p += ia == ib
        same_w += ia == ib and int(ka[i][ia]) == int(kb[i][ib])
        max_sde = max(max_sde, abs(sa - sb) / abs(sb))
    return same_p, same_w, max_rel, max_sde


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--curves", type=int, default=4096)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--plain-sample", type=int, default=24)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work")
    ap.add_argument("--weighted-curves", type=int, default=64,
                    help="curves of the per-sample-dy leg (0: skip)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-refine", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tuning", default=None, metavar="TOEP_Q,CHUNK_MB,WG_PER_CU[,QUAD_Q]",
                    help="apply this setting instead of the autotuner's choice")
    ap.add_argument("--no-full-check", action="store_true",
                    help="skip the float64 search of a whole batch (parity of every curve)")
    ap.add_argument("--eager", action="store_true",
                    help="time eager search calls instead of graph replays")
    ap.add_argument("--refine-curves", type=int, default=256)
    ap.add_argument("--refine-steps", type=int, default=200)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scale for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the search runs each batch on a single device "
              "and there is no multi-GPU figure to report; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.curves % args.batch:
        print("--curves must be a multiple of --batch", file=sys.stderr)
        sys.exit(2)

    import torch

    from gtls_amd import batch as B
    from gtls_amd import search as S
    from gtls_amd import synthetic
    from gtls_amd import tune
    from gtls_amd.grid import skip_factor
    from gtls_amd.validate import validate_inputs

    if args.copy_calibration:
        assert torch.cuda.is_available(), "bench.py needs a GPU"
        # a plain copy of 2^28 float32 (1 GiB read, 1 GiB written), three launches
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30,
                          "launches": 3}), flush=True)
        return
    skip = skip_factor(0.125)

    # host planning (grids, templates, width windows) before the GPU runtime starts
    t0 = time.time()
    curves = synthetic.batch(args.curves, seed=args.seed)
    tp = time.perf_counter()
    plans = B.build_plans(curves, {})
    t_plan = time.perf_counter() - tp
    del curves
    workers = min(16, os.cpu_count() or 1)
    log("planned %d curves in %.1f s on %d processes (%.2f ms per curve)" % (
        len(plans), t_plan, workers, t_plan * 1e3 / len(plans)))

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n_batches = args.curves // args.batch
    batches = []
    for b in range(n_batches):
        bp = B.BatchPlan(plans[b * args.batch:(b + 1) * args.batch])
        batches.append((bp, B.DeviceBatch(bp, 0)))
    macs = np.array([window_macs(p, skip) for p in plans])
    algo = np.array([algorithmic_bytes(p, 4) for p in plans])
    log("built %d plans in %d batches: %d points, %d periods, %.3g window MACs, %.1f s" % (
        len(plans), n_batches, sum(p.n for p in plans), sum(len(p.periods) for p in plans),
        macs.sum(), time.time() - t0))

    tuning = None
    if args.tuning:
        setting = tuple(int(x) for x in args.tuning.split(","))
        if len(setting) not in (3, 4):
            print("--tuning takes TOEP_Q,CHUNK_MB,WG_PER_CU[,QUAD_Q]", file=sys.stderr)
            sys.exit(2)
        if len(setting) == 3:
            setting = setting + (tune.DEFAULT[3],)
        tune.set_tuning(0, *setting)
        tuning = {"choice": {"toep_q": setting[0], "chunk_mb": setting[1],
                             "wg_per_cu": setting[2], "quad_q": setting[3]},
                  "source": "--tuning"}
    elif not args.no_tune:
        ts = time.perf_counter()
        choice, times = tune.autotune(batches[0][1], skip, n_sample=32, repeats=1)
        tuning = {"choice": {"toep_q": choice[0], "chunk_mb": choice[1],
                             "wg_per_cu": choice[2], "quad_q": choice[3]},
                  "candidates_ms": {"%d/%d/%d/%d" % k: round(v, 2) for k, v in times.items()},
                  "seconds": round(time.perf_counter() - ts, 1),
                  "sample": "32 curves of batch 0 spread over its sizes, one timed call per "
                            "setting; coordinate search (workgroups/CU, chunk budget, "
                            "Toeplitz bound, operand layout threshold in turn)"}
        log("autotune: toep_q %d, chunk %d MiB, %d workgroups/CU, quad_q %d (%.1f s)" % (
            choice + (tuning["seconds"],)))

    # one graph per batch, all over one workspace (replayed in stream order), each with
    # the library's per-chunk kernel events captured
    graphs = []
    ws = None
    if not args.eager:
        need = max(db.workspace_bytes("fp32c", skip) for _, db in batches)
        ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        for bp, db in batches:
            graphs.append(db.capture("fp32c", skip, ws=ws, timing=True))
        log("captured %d graphs over a %.1f GiB workspace" % (len(graphs), need / 2 ** 30))

    def step(i, ev=None):
        bp, db = batches[i % n_batches]
        if ev is not None:
            ev[0].record()
        if graphs:
            graphs[i % n_batches].replay()
            chi2 = db.chi2
        else:
            chi2, _, _ = db.search("fp32c", skip)
        if ev is not None:
            ev[1].record()
        db.spectra(chi2, 3, sync=False)

    def step_kernel_ms(i):
        if graphs:
            return B.graph_kernel_ms(graphs[i % n_batches])
        return S.kernel_ms(0)

    if not graphs:
        S.set_kernel_timing(0, True)
    for i in range(args.warmup):
        step(i)
        torch.cuda.synchronize()
        step_kernel_ms(i)
        log("warmup step %d done" % i)

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    kern_ms = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(i, evs[i])
        torch.cuda.synchronize()          # a progress line per finished step
        kern_ms.append(step_kernel_ms(i))
        log("step %d done" % i)
    elapsed = time.perf_counter() - t_start
    if not graphs:
        S.set_kernel_timing(0, False)
    prep_ms = np.array([k[0][0] for k in kern_ms])
    win_ms = np.array([k[0][1] for k in kern_ms])
    red_ms = np.array([k[0][2] for k in kern_ms])
    chunks = np.array([k[1] for k in kern_ms])
    search_ms = np.array([a.elapsed_time(b) for a, b in evs])
    ms_per_step = elapsed * 1e3 / args.steps
    n_done = args.batch * args.steps
    value = n_done / elapsed * 3600.0
    log("timed %d steps: %.1f ms/step, search %.1f ms on average (prep %.1f, windows %.1f)" % (
        args.steps, ms_per_step, search_ms.mean(), prep_ms.mean(), win_ms.mean()))

    # per step: the useful window FLOPs (2 per MAC) and algorithmic bytes of its batch
    step_curves = [np.arange((i % n_batches) * args.batch, (i % n_batches + 1) * args.batch)
                   for i in range(args.steps)]
    flops = np.array([2.0 * macs[c].sum() for c in step_curves])
    bytes_ = np.array([algo[c].sum() for c in step_curves])
    # the dominant kernel: windows_kernel (all launches of a step: one per period chunk),
    # useful window FLOPs over its HIP-event time
    achieved = float(np.mean(flops / (win_ms * 1e-3)) / 1e12)
    pipe_achieved = float(np.mean(flops / (search_ms * 1e-3)) / 1e12)
    hbm_achieved = float(np.mean(bytes_ / (search_ms * 1e-3)) / 1e9)
    traffic, traffic_src = profiled_traffic(args)
    roofline = {"bound": "mfma", "achieved": round(achieved, 3),
                "peak": FP32_MATRIX_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_MATRIX_PEAK_TFLOPS, 5),
                "traffic": traffic.get("windows_kernel") if traffic else None,
                "traffic_source": (traffic_src + " (batch 0)") if traffic_src else None,
                "kernel": "windows_kernel: every launch of a step (one per period chunk)",
                "peak_note": "useful float32-equivalent window FLOPs against the FP32 "
                             "matrix peak; the tiles run on 16x16x32 f16 MFMAs over "
                             "(hi, lo) pairs (2.5 PFLOP/s dense f16 peak)",
                "timing": "HIP events the library records around every chunk's kernels on "
                          "the launch stream, " + ("captured in each batch's graph and "
                                                   "recorded by every replay" if graphs
                                                   else "eager calls"),
                "launch_ms": round(float(win_ms.mean()), 3),
                "launch_ms_per_step": [round(float(x), 3) for x in win_ms],
                "launches_per_step": [int(x) for x in chunks],
                "algorithmic_flops_per_launch": float(flops.mean()),
                "search_call": {
                    "kernels": "gtls_search_batched: prep_kernel + windows_kernel + "
                               "reduce_kernel over every period chunk",
                    "launch_ms": round(float(search_ms.mean()), 3),
                    "launch_ms_per_step": [round(float(x), 3) for x in search_ms],
                    "prep_kernel_ms": round(float(prep_ms.mean()), 3),
                    "reduce_kernel_ms": round(float(red_ms.mean()), 3),
                    "achieved_tflops": round(pipe_achieved, 3),
                    "frac": round(pipe_achieved / FP32_MATRIX_PEAK_TFLOPS, 5),
                    "traffic": traffic.get("search_call") if traffic else None,
                    "algorithmic_bytes_per_launch": float(bytes_.mean()),
                    "algorithmic_hbm_GBps": round(hbm_achieved, 1),
                    "algorithmic_hbm_frac": round(hbm_achieved / HBM_PEAK_GBS, 5)}}

    # batched results of every searched batch (the last step of each batch wrote them)
    res = {}
    for b, (bp, db) in enumerate(batches[:min(args.steps, n_batches)]):
        spec = db.spectra(db.chi2, 3)
        bk = bp.unpack(db.best_k.cpu().numpy())
        for j in range(args.batch):
            ci = b * args.batch + j
            res[ci] = (spec["SDE"][j], spec["power"][j], bk[j])
    extra = {"batch_curves": args.batch, "batches_in_set": n_batches,
             "step": "graph replay" if graphs else "eager search call",
             "host_planning_ms_per_curve": round(t_plan * 1e3 / len(plans), 3),
             "host_planning_note": "grids, templates and width windows on %d host processes, "
                                   "outside value (as in the plain leg)" % workers,
             "tuning": tuning}
    done = np.array(sorted(res))
    total_macs = macs[done].sum()

    if not args.no_full_check:
        # every curve of batch 0: the batched float32 search against the batched float64 one
        bp0, db0 = batches[0]
        if graphs:
            graphs[0].replay()
        else:
            db0.search("fp32c", skip)
        a = (db0.chi2.cpu().numpy().copy(), db0.best_k.cpu().numpy().copy())
        ts = time.perf_counter()
        db0.search("fp64", skip)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - ts
        b = (db0.chi2.cpu().numpy().copy(), db0.best_k.cpu().numpy().copy())
        sp, sw, mrel, msde = compare_searches(bp0, a, b)
        extra["full_batch_check"] = {
            "curves": bp0.n_curves, "against": "batched float64 search of the same batch",
            "same_best_period": "%d/%d" % (sp, bp0.n_curves),
            "same_best_width": "%d/%d" % (sw, bp0.n_curves),
            "max_rel_chi2_err": mrel, "max_rel_sde_err": msde,
            "fp64_batched_curves_per_hour": round(bp0.n_curves / t64 * 3600.0, 1)}
        log("full batch 0: best period %d/%d, width %d/%d, max rel chi2 %.2e, max dSDE %.2e"
            % (sp, bp0.n_curves, sw, bp0.n_curves, mrel, msde))
        if graphs:
            graphs[0].replay()          # batch 0's float32 results back in place
            torch.cuda.synchronize()

    if args.weighted_curves > 0:
        # per-sample dy (sigma U(0.5, 1.5)) on the first curves of the set: the weighted
        # path (two sample planes, float64 window means)
        rng = np.random.default_rng(args.seed + 100)
        k = min(args.weighted_curves, len(plans))
        wc = []
        for p in plans[:k]:
            sig = float(np.std(p.y))
            wc.append((p.t, p.y, sig * rng.uniform(0.5, 1.5, p.n)))
        wplans = B.build_plans(wc, {})
        wbp = B.BatchPlan(wplans)
        wdb = B.DeviceBatch(wbp, 0)
        wdb.search("fp32c", skip)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        wdb.search("fp32c", skip)
        e1.record()
        torch.cuda.synchronize()
        tw = e0.elapsed_time(e1) * 1e-3
        a = (wdb.chi2.cpu().numpy().copy(), wdb.best_k.cpu().numpy().copy())
        wdb.search("fp64", skip)
        torch.cuda.synchronize()
        b = (wdb.chi2.cpu().numpy().copy(), wdb.best_k.cpu().numpy().copy())
        sp, sw, mrel, msde = compare_searches(wbp, a, b)
        wm = float(np.sum([window_macs(p, skip) for p in wplans]))
        extra["weighted"] = {
            "curves": k, "dy": "std(flux) x U(0.5, 1.5) per sample",
            "search_seconds": round(tw, 3),
            "curves_per_hour": round(k / tw * 3600.0, 1),
            "window_tflops": round(2 * wm / tw / 1e12, 3),
            "same_best_period_vs_fp64": "%d/%d" % (sp, k),
            "same_best_width_vs_fp64": "%d/%d" % (sw, k),
            "max_rel_chi2_err_vs_fp64": mrel, "max_rel_sde_err_vs_fp64": msde}
        log("weighted: %d curves in %.2f s, best period %d/%d vs fp64, max rel chi2 %.2e" % (
            k, tw, sp, k, mrel))

    singles, single_res, single_arrays = [], {}, {}
    if not args.no_plain:
        sample = [int(done[i]) for i in stratified(macs[done], args.plain_sample)]
        t_plain = w_plain = 0.0
        same_p = same_w = 0
        max_dsde = 0.0
        for ci in sample:
            p = plans[ci]
            dc = S.DeviceCurve(p, 0)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            c, k, _ = dc.search(p.periods, p.k_lo, p.k_hi, mode="fp64", skip_point=skip)
            c, k = c.cpu().numpy(), k.cpu().numpy()
            from gtls_amd import spectra as SP
            SP.spectra(SP.mask_outliers(c), 3)
            t_plain += time.perf_counter() - ts
            w_plain += macs[ci]
            sp, sw, d = peak_check(*res[ci], c, k)
            same_p += sp
            same_w += sw
            max_dsde = max(max_dsde, d)
        plain_value = len(done) / (t_plain * total_macs / w_plain) * 3600.0
        extra.update({"plain_fp64_curves_per_hour": round(plain_value, 1),
                      "speedup_vs_plain_fp64": round(value / plain_value, 3),
                      "plain_sample_curves": len(sample),
                      "parity_vs_plain_same_best_period": "%d/%d" % (same_p, len(sample)),
                      "parity_vs_plain_same_best_width": "%d/%d" % (same_w, len(sample)),
                      "parity_vs_plain_max_rel_sde_err": float(max_dsde)})
        log("plain fp64: %.1f curves/h (sample %d), parity %d/%d %d/%d, max dSDE %.2e" % (
            plain_value, len(sample), same_p, len(sample), same_w, len(sample), max_dsde))
        # the single-curve configurations, one curve per call
        singles = single_curve_plans()
        single_res, single_arrays = single_gpu_leg(singles, skip)
        extra["single_curve"] = single_res

    if not args.no_refine:
        extra["refine"] = refine_leg(args, plans, batches[0], skip)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        cores = os.cpu_count() or 1
        cand = [int(done[i]) for i in stratified(macs[done], 64)]
        # smallest first to learn the rate, then the spread sample within the budget
        first = min(cand, key=lambda i: macs[i])
        t_cpu = w_cpu = 0.0
        used = []
        same_p = same_w = 0
        max_dsde = 0.0
        for ci in [first] + [i for i in cand if i != first]:
            if used:
                rate = t_cpu / w_cpu
                if t_cpu + rate * macs[ci] > args.cpu_budget:
                    continue
            p = plans[ci]
            ts = time.perf_counter()
            c, k, _ = oracle.search_plan(p, skip_point=skip, nthreads=cores)
            t_cpu += time.perf_counter() - ts
            w_cpu += macs[ci]
            used.append(ci)
            sp, sw, d = peak_check(*res[ci], c, k)
            same_p += sp
            same_w += sw
            max_dsde = max(max_dsde, d)
        cpu_value = len(done) / (t_cpu * total_macs / w_cpu) * 3600.0
        cpu = {"value": round(cpu_value, 2), "unit": "light curves/hour", "cores": cores,
               "kind": "port", "cpu_model": cpu_model(),
               "sample": "%d curves of the set spread over its work distribution "
                         "(N %d-%d; %.1f s of FP64 oracle search on %d threads), "
                         "extrapolated by window MACs" % (
                             len(used), min(plans[i].n for i in used),
                             max(plans[i].n for i in used), t_cpu, cores)}
        if singles:
            cpu["single_curve"] = single_cpu_leg(singles, skip, cores, single_res,
                                                 single_arrays)
        extra.update({"parity_vs_cpu_same_best_period": "%d/%d" % (same_p, len(used)),
                      "parity_vs_cpu_same_best_width": "%d/%d" % (same_w, len(used)),
                      "parity_vs_cpu_max_rel_sde_err": float(max_dsde)})
        log("cpu oracle: %.2f curves/h over %d curves, parity %d/%d %d/%d, max dSDE %.2e" % (
            cpu_value, len(used), same_p, len(used), same_w, len(used), max_dsde))

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
        "dtype": "fp32+fp64 (f16 hi/lo pairs on the matrix cores)",
        "data": "synthetic (seeded light curves with injected limb-darkened transits)",
        "config": {"workload": "batched TLS search, %d-curve set in batches of %d, "
                               "N log-U(2000, 70000), cadence 2/30 min, own period grid "
                               "per curve, T0_fit_margin 0.125" % (args.curves, args.batch),
                   "curves": args.curves, "batch": args.batch,
                   "curves_searched": n_done, "mode": "fp32c", "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)

    # teardown in order: graphs and device buffers, then the library contexts, so that the
    # process leaves nothing queued or allocated on the device when it exits
    from gtls_amd import _ffi
    del graphs, batches, ws
    S._WORKSPACES.clear()
    torch.cuda.synchronize()
    _ffi.Context.close_all()
    torch.cuda.empty_cache()
    log("done")


if __name__ == "__main__":
    main()
