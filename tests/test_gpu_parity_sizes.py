"""Parity with the FP64 CPU oracle at the sizes of the benchmark configurations, through
the C ABI (bar in tests/parity_util.py):

  * the 4-year 29.4-min curve (N ~ 65,000, ~1.9e5 periods): a seeded 1,500-period sample of
    its full grid plus the periods around the injected 45.3 d;
  * batches over the batched workload's size range (2,000-70,000 points, 2- and 30-min
    cadences, two 30-min curves above 40,000 points), with a common dy and with per-sample
    dy, each curve on a 400-period sample of its own grid plus its injected period;
  * a curve with a sample time repeated 300 times, so that one phase bucket exceeds the
    in-LDS window sort and the prep kernel takes its global sort form;
  * a 70-day 20-s curve (N = 302,400) on periods whose widths exceed 32,768 samples.
"""
import numpy as np
import pytest

from gtls_amd import batch as B
from gtls_amd import search as S
from gtls_amd import synthetic
from gtls_amd.grid import skip_factor
from gtls_amd.validate import validate_args, validate_inputs

from parity_util import check, report

pytestmark = pytest.mark.gpu

SKIP = skip_factor(0.125)


def period_sample(periods, k, seed, around=None, n_around=5):
    rng = np.random.default_rng(seed)
    sel = set(rng.choice(len(periods), size=min(k, len(periods)), replace=False).tolist())
    if around is not None:
        i = int(np.argmin(np.abs(periods - around)))
        sel |= set(range(max(0, i - n_around), min(len(periods), i + n_around + 1)))
    return np.array(sorted(sel))


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_long_curve_period_sample_matches_oracle(gpu, mode):
    t, y, dy, truth = synthetic.long_30min()
    t, y, dy = validate_inputs(t, y, dy)
    plan = S.make_plan(t, y, dy, validate_args({}))
    assert plan.n > 60000 and len(plan.periods) > 150000
    sel = period_sample(plan.periods, 1500, 17, around=truth["period"])
    per, klo, khi = plan.periods[sel], plan.k_lo[sel], plan.k_hi[sel]
    c, k, s = [x.cpu().numpy() for x in S.DeviceCurve(plan).search(per, klo, khi, mode=mode,
                                                                   skip_point=SKIP)]
    worst, n_same, n_ok = check(plan, mode, SKIP, per, klo, khi, c, k, s, "C2")
    report("C2-long-30min", mode, worst, n_same, n_ok, "n=%d" % plan.n)
    i = int(np.nanargmin(c))
    assert per[i] == pytest.approx(truth["period"], rel=2e-3)


def c4_range_curves(weighted, seed=31):
    """Eight curves across the batched workload's range: N from 2,100 to 69,000 (three at
    2-min and five at 30-min cadence), noise U(1e-4, 2e-3), a transit in every other one;
    with `weighted`, dy varies per sample (sigma times U(0.5, 1.5))."""
    rng = np.random.default_rng(seed)
    spec = [(2100, 2.0), (5200, 30.0), (12000, 2.0), (24000, 30.0), (41000, 30.0),
            (52000, 30.0), (69000, 2.0), (66000, 30.0)]
    out = []
    for j, (n, cad_min) in enumerate(spec):
        cad = cad_min / 1440.0
        sigma = rng.uniform(1e-4, 2e-3)
        t = synthetic.T_START + np.arange(n) * cad
        y = 1.0 + rng.normal(0.0, sigma, n)
        truth = None
        if j % 2 == 0:
            P = rng.uniform(0.8, 0.3 * n * cad)
            depth = rng.uniform(2e-4, 5e-3)
            t0 = synthetic.T_START + rng.uniform(0, P)
            y *= synthetic.transit_flux(t, P, t0, synthetic.solar_t14(P), depth)
            truth = P
        dy = sigma * rng.uniform(0.5, 1.5, n) if weighted else None
        out.append((t, y, dy, truth))
    return out


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_batch_over_the_size_range_matches_oracle(gpu, mode, weighted):
    curves = c4_range_curves(weighted)
    plans = []
    for j, (t, y, dy, truth) in enumerate(curves):
        t, y, dy = validate_inputs(t, y, dy)
        full = S.make_plan(t, y, dy, validate_args({}))
        sel = period_sample(full.periods, 400, 100 + j, around=truth)
        plans.append(S.make_plan(t, y, dy, validate_args({}), periods=full.periods[sel]))
    assert sum(p.n > 40000 and np.median(np.diff(p.t)) > 0.02 for p in plans) >= 2
    bp = B.BatchPlan(plans)
    db = B.DeviceBatch(bp)
    db.search(mode, SKIP)
    chi2 = bp.unpack(db.chi2.cpu().numpy())
    bk = bp.unpack(db.best_k.cpu().numpy())
    bs = bp.unpack(db.best_s.cpu().numpy())
    worst, n_same, n_ok = 0.0, 0, 0
    for i, p in enumerate(plans):
        w, ns, no = check(p, mode, SKIP, p.periods, p.k_lo, p.k_hi, chi2[i], bk[i], bs[i],
                          "curve %d (n=%d)" % (i, p.n))
        worst, n_same, n_ok = max(worst, w), n_same + ns, n_ok + no
    report("C4-range-%s" % ("weighted" if weighted else "common-dy"), mode, worst, n_same,
           n_ok, "curves=%d" % len(plans))


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_repeated_times_take_the_global_sort(gpu, mode):
    """300 copies of one sample time: their phases fall into one bucket at every period,
    above the window sort's bucket bound (256 keys), so every period is sorted by the prep
    kernel's global form; ties are ordered by sample index, as in a stable argsort."""
    rng = np.random.default_rng(41)
    n = 2500
    t = 1000.0 + np.arange(n) / 720.0
    t = np.sort(np.concatenate([t, np.full(300, t[1234])]))
    y = 1.0 + rng.normal(0, 5e-4, t.size)
    y *= synthetic.transit_flux(t, 0.9, 1000.2, 0.07, 3e-3)
    t, y, dy = validate_inputs(t, y, None)
    plan = S.make_plan(t, y, dy, validate_args({}))
    sel = np.arange(0, len(plan.periods), 4)
    per, klo, khi = plan.periods[sel], plan.k_lo[sel], plan.k_hi[sel]
    c, k, s = [x.cpu().numpy() for x in S.DeviceCurve(plan).search(per, klo, khi, mode=mode,
                                                                   skip_point=SKIP)]
    worst, n_same, n_ok = check(plan, mode, SKIP, per, klo, khi, c, k, s, "repeated")
    report("repeated-times-global-sort", mode, worst, n_same, n_ok)


def fast_cadence_curve(seed=53):
    """A 70-day curve at 20-s cadence (N = 302,400) with a 23-day transit. Its widest
    durations span up to 36,288 samples (0.12 N), past 32,768."""
    rng = np.random.default_rng(seed)
    n = 302400
    t = synthetic.T_START + np.arange(n) * (20.0 / 86400.0)
    y = 1.0 + rng.normal(0.0, 1e-3, n)
    y *= synthetic.transit_flux(t, 23.0, synthetic.T_START + 7.0, synthetic.solar_t14(23.0),
                                2e-3)
    return validate_inputs(t, y, None)


def wide_width_sample(plan, k=12, seed=61):
    """k random periods, those around the injected 23 d, and k spread over the periods whose
    widest width exceeds 32,768 samples."""
    wide = np.nonzero(plan.table.widths[plan.k_hi] > 32768)[0]
    spread = wide[np.linspace(0, len(wide) - 1, k).astype(int)]
    return np.union1d(period_sample(plan.periods, k, seed, around=23.0, n_around=2), spread)


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_widths_past_32768_samples_match_oracle(gpu, mode):
    """Widths wider than 32,768 samples: the zeroed template tail after the last template
    row and the zero block of the chunked rows are sized from the widest width of the
    search, and the sort takes about 120 windows per period (N = 302,400)."""
    t, y, dy = fast_cadence_curve()
    plan = S.make_plan(t, y, dy, validate_args({}))
    assert plan.table.widths.max() > 36000
    sel = wide_width_sample(plan)
    per, klo, khi = plan.periods[sel], plan.k_lo[sel], plan.k_hi[sel]
    assert np.sum(plan.table.widths[khi] > 32768) >= 12
    c, k, s = [x.cpu().numpy() for x in S.DeviceCurve(plan).search(per, klo, khi, mode=mode,
                                                                   skip_point=SKIP)]
    worst, n_same, n_ok = check(plan, mode, SKIP, per, klo, khi, c, k, s, "wide widths")
    report("wide-widths-20s", mode, worst, n_same, n_ok, "n=%d" % plan.n)
    i = int(np.nanargmin(c))
    assert per[i] == pytest.approx(23.0, rel=2e-3)
