"""Parity of the HIP search with the FP64 CPU oracle, through the C ABI.

Bar: the phase permutation is bit-identical to a stable argsort; in FP64 mode every
period's best (width, position) index equals the oracle's and chi2 agrees to 1e-9
relative (the engine's bar is 1e-6); in FP32-compensated mode chi2 agrees to 2e-6
relative and any period whose best index differs from the oracle's is a near tie: the
oracle's own statistic at the device's index is within 2e-6 relative of its minimum.
"""
import numpy as np
import pytest

from gtls_amd import search as S
from gtls_amd import synthetic
from gtls_amd.grid import skip_factor
from gtls_amd.validate import validate_args, validate_inputs
from oracle import oracle

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-9
FP32_RTOL = 2e-6


def plan_for(t, y, dy=None, periods=None, **kw):
    t, y, dy = validate_inputs(t, y, dy)
    return S.make_plan(t, y, dy, validate_args(kw), periods=periods)


def small_curve(n, seed, cadence=1 / 720, depth=3e-3, period=1.37):
    rng = np.random.default_rng(seed)
    t = 1000.0 + np.arange(n) * cadence
    y = 1.0 + rng.normal(0, 5e-4, n)
    y *= synthetic.transit_flux(t, period, 1000.3, 0.08, depth)
    return t, y


def assert_parity(plan, mode, skip, periods=None, k_lo=None, k_hi=None, device=0):
    periods = plan.periods if periods is None else periods
    k_lo = plan.k_lo if k_lo is None else k_lo
    k_hi = plan.k_hi if k_hi is None else k_hi
    dc = S.DeviceCurve(plan, device)
    c_g, k_g, s_g = [x.cpu().numpy() for x in dc.search(periods, k_lo, k_hi, mode=mode,
                                                            skip_point=skip)]
    c_o, k_o, s_o = oracle.search_plan(plan, skip_point=skip, periods=periods, k_lo=k_lo,
                                       k_hi=k_hi)
    empty = k_lo > k_hi
    assert np.all(np.isnan(c_g[empty])) and np.all(k_g[empty] == -1)
    ok = ~empty
    rel = np.abs(c_g[ok] - c_o[ok]) / np.abs(c_o[ok])
    print("mode %s skip %d: %d periods, max rel chi2 err %.3g" % (mode, skip, ok.sum(),
                                                                 rel.max() if rel.size else 0))
    tb = plan.table
    if mode == "fp64":
        assert rel.max(initial=0) < FP64_RTOL
        bad = np.nonzero((k_g != k_o) | (s_g != s_o))[0]
        assert bad.size == 0, (bad[:10], k_g[bad[:10]], k_o[bad[:10]], s_g[bad[:10]], s_o[bad[:10]])
    else:
        assert rel.max(initial=0) < FP32_RTOL
        bad = np.nonzero(ok & ((k_g != k_o) | (s_g != s_o)))[0]
        for i in bad:
            table, cmin, *_ = oracle.period_table(plan.t, plan.y, plan.dy, periods[i],
                                                  tb.widths, tb.S, tb.offsets, tb.overshoot,
                                                  k_lo[i], k_hi[i], skip_point=skip)
            at = table[k_g[i] - k_lo[i], s_g[i]]
            assert (at - cmin) <= FP32_RTOL * abs(cmin), (i, at, cmin)
        assert bad.size <= max(2, ok.sum() // 100)
    return c_g, k_g, s_g


def test_fold_sort_bit_exact(gpu):
    rng = np.random.default_rng(1)
    cases = [
        (np.sort(rng.uniform(1000, 1030, 5000)), rng.uniform(0.5, 12, 40)),
        (1000 + np.arange(20000) / 720.0, np.array([0.5, 1.0, 1.5, 2.0, 3.71, 7.0])),  # ties
        (rng.uniform(-40, 40, 3000), rng.uniform(0.4, 9, 20)),                        # t < 0
        (1000 + np.arange(300) / 48.0, np.array([0.5, 0.75, 1.0])),                   # small n
        # periods above the span: part of the phase circle is empty, so a window of the
        # in-LDS sort spans thousands of empty buckets
        (np.sort(rng.uniform(1000, 1010, 40000)), np.array([25.0, 31.0, 3.3])),
        # periods that are multiples of the cadence: ~139 and ~69 equal phases per bucket
        (1000 + np.arange(50000) / 720.0, np.array([0.5, 0.25, 0.5 + 1e-9])),
        # 8 distinct times repeated 600 times: 600-key buckets
        (np.repeat(1000 + np.arange(8) / 48.0, 600), np.array([0.5, 0.77])),
        # one time repeated: a single bucket of 5000 equal phases, ordered by index
        (np.full(5000, 1234.5), np.array([0.7, 3.1])),
        # times below 2^-960 (the reciprocal form of the fold would not apply: the dividing
        # global form), zeros of both signs, and periods past 2^1000
        (np.concatenate([[-0.0, 0.0, 3e-310, 1e-300], 1000 + np.arange(4000) / 720.0]),
         np.array([0.5, 1.37, 2.0 ** 1001])),
        (1000 + np.arange(4000) / 720.0, np.array([2.0 ** 1001, 2.0 ** -1001])),
    ]
    for t, periods in cases:
        got = S.fold_sort(t, periods).cpu().numpy()
        for i, P in enumerate(periods):
            x = t / P
            want = np.argsort(x - np.trunc(x), kind="stable")
            assert np.array_equal(got[i], want), (len(t), P)
            assert np.array_equal(got[i], oracle.fold_sort(t, P))


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
@pytest.mark.parametrize("margin", [0.125, 0.0])
def test_search_parity_small_curves(gpu, mode, margin):
    for n, seed in ((600, 1), (2500, 2)):
        t, y = small_curve(n, seed)
        plan = plan_for(t, y)
        assert_parity(plan, mode, skip_factor(margin))


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_search_parity_with_dy_and_gaps(gpu, mode):
    rng = np.random.default_rng(7)
    n = 4000
    t = 1000.0 + np.arange(n) / 48.0
    keep = (t < 1030) | (t > 1034)
    t = t[keep]
    y = 1 + rng.normal(0, 1e-3, t.size)
    y *= synthetic.transit_flux(t, 5.3, 1002.0, 0.2, 4e-3)
    dy = rng.uniform(5e-4, 2e-3, t.size)
    plan = plan_for(t, y, dy)
    assert_parity(plan, mode, 8)


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_search_parity_single_2min_full_grid(gpu, mode):
    """The 19,728-point 2-min curve on its whole period grid at the default T0_fit_margin
    0.125 (the float32 mode of it is the precision configuration), and on every 25th period
    with no skip (T0_fit_margin 0)."""
    from parity_util import check, report
    t, y, dy, truth = synthetic.single_2min()
    plan = plan_for(t, y, dy)
    skip = skip_factor(0.125)
    c, k, s = [x.cpu().numpy() for x in S.DeviceCurve(plan).search(
        plan.periods, plan.k_lo, plan.k_hi, mode=mode, skip_point=skip)]
    worst, n_same, n_ok = check(plan, mode, skip, plan.periods, plan.k_lo, plan.k_hi, c, k, s,
                                "C1")
    report("C1-2min" if mode == "fp64" else "C3-2min-precision", mode, worst, n_same, n_ok,
           "n=%d skip=%d" % (plan.n, skip))
    sel = np.arange(0, len(plan.periods), 25)
    per, klo, khi = plan.periods[sel], plan.k_lo[sel], plan.k_hi[sel]
    none = skip_factor(0.0)
    c, k, s = [x.cpu().numpy() for x in S.DeviceCurve(plan).search(per, klo, khi, mode=mode,
                                                                   skip_point=none)]
    worst, n_same, n_ok = check(plan, mode, none, per, klo, khi, c, k, s, "C1 no skip")
    report("C1-2min-no-skip" if mode == "fp64" else "C3-2min-no-skip", mode, worst, n_same,
           n_ok, "n=%d skip=none" % plan.n)


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_tiny_time_takes_the_dividing_form(gpu, mode):
    """A curve with a time below 2^-960: the windowed sort's reciprocal fold does not
    apply, so every period of it is sorted by the global form, which divides."""
    t, y = small_curve(1200, 5)
    periods = plan_for(t, y).periods
    plan = plan_for(np.concatenate([[1e-300], t]), np.concatenate([[1.0], y]),
                    periods=periods)
    sel = np.arange(0, len(plan.periods), 7)
    assert_parity(plan, mode, 8, plan.periods[sel], plan.k_lo[sel], plan.k_hi[sel])


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_times_before_zero_match_oracle(gpu, mode):
    """Times on both sides of zero (reachable through the C ABI and make_plan; the
    reference's input cleaning drops them): folds with phases in (-1, 1), the bucket map
    of negative times, against the oracle's fold of the same times."""
    t, y = small_curve(1500, 9)
    t = t - 1000.5   # about -0.5 .. 1.6 d
    dy = np.full(t.size, float(np.std(y)))
    plan = S.make_plan(t, y, dy, validate_args({}))
    assert plan.t.min() < 0 < plan.t.max()
    sel = np.arange(0, len(plan.periods), 5)
    assert_parity(plan, mode, 8, plan.periods[sel], plan.k_lo[sel], plan.k_hi[sel])


def test_edge_cases(gpu):
    t, y = small_curve(400, 3)
    plan = plan_for(t, y)
    periods = plan.periods[:12].copy()
    k_lo = plan.k_lo[:12].copy()
    k_hi = plan.k_hi[:12].copy()
    k_lo[3], k_hi[3] = 5, 4          # empty width set -> NaN
    k_hi[5] = k_lo[5]                # single width
    for mode in ("fp64", "fp32c"):
        assert_parity(plan, mode, 8, periods, k_lo, k_hi)
        assert_parity(plan, mode, skip_factor(0), periods, k_lo, k_hi)
    # constant flux: every gate closed -> chi2 = N at the first position
    flat = plan_for(t, np.ones_like(t) * 1.0 + 0.0, np.full(t.size, 1e-3))
    dc = S.DeviceCurve(flat)
    c, k, s = [x.cpu().numpy() for x in dc.search(flat.periods[:5], flat.k_lo[:5],
                                                  flat.k_hi[:5])]
    assert np.all(c == t.size) and np.all(k == flat.k_lo[:5]) and np.all(s == 0)


@pytest.mark.parametrize("skip", [2, 3, 5, 8, 16, 100])
@pytest.mark.parametrize("weighted", [False, True])
def test_search_parity_skip_strides(gpu, skip, weighted):
    """Every T0 stride (register-blocked windows: segment and remainder bookkeeping) on
    both paths: common weight (dy constant) and per-sample weights."""
    t, y = small_curve(1500, 11)
    dy = np.random.default_rng(4).uniform(4e-4, 9e-4, t.size) if weighted else None
    plan = plan_for(t, y, dy)
    sel = np.arange(0, len(plan.periods), 3)
    for mode in ("fp64", "fp32c"):
        assert_parity(plan, mode, skip, plan.periods[sel], plan.k_lo[sel], plan.k_hi[sel])
