"""Host-side pins of the spectra (src/gputls/stats.py:95-122, helpers.py:91-106), the
single-period statistics helpers (src/gputls/stats.py:233-488) and the candidate
selection (src/gputls/core.py:904-915), by closed forms and brute force.
"""
import numpy as np
import numpy.ma as ma
import pytest

from gtls_amd import spectra as SP
from gtls_amd import stats as ST
from gtls_amd.api import candidate_indices, nearest_indices


def brute_running_median(x, k):
    x = np.asarray(x, float)
    med = np.array([np.median(x[i:i + k]) for i in range(len(x) - k + 1)])
    front = (len(x) - len(med)) // 2
    back = len(x) - len(med) - front
    return np.concatenate([np.full(front, med[0]), med, np.full(back, med[-1])])


@pytest.mark.parametrize("n,k", [(200, 91), (57, 7), (91, 91), (1000, 61)])
def test_running_median_matches_brute_force(n, k):
    x = np.random.default_rng(n + k).normal(size=n)
    assert np.allclose(SP.running_median(x, k), brute_running_median(x, k), rtol=0, atol=0)


def test_running_median_ignores_masked_values():
    x = np.arange(30, dtype=float)
    m = ma.array(x, mask=np.zeros(30, bool))
    m.mask[10] = True
    got = SP.running_median(m, 5)
    # window 8..12 without the masked 10: median of (8, 9, 11, 12) = 10
    assert got[10] == 10.0
    # windows without masked values are plain medians
    assert got[3] == 3.0 and got[20] == 20.0


def test_median_kernel_is_odd():
    assert SP.median_kernel(3) == 91
    assert SP.median_kernel(2) == 61
    assert SP.median_kernel(5) == 151


@pytest.mark.parametrize("P", [150, 1000, 183])
def test_single_dip_sde_closed_form(P):
    """chi2 flat except one lower value: SR is a constant a except 1 at the dip, so
    SDE_raw = (1 - mean)/std = sqrt(P - 1) for any depth, with or without the running
    median detrend (P <= 182 and P > 182 for oversampling 3)."""
    chi2 = np.full(P, 1000.0)
    dip = P // 3
    chi2[dip] = 700.0
    SR, power_raw, power, SDE_raw, SDE = SP.spectra(chi2, 3)
    assert np.allclose(SR[dip], 1.0) and np.allclose(np.delete(SR, dip), 0.7)
    assert SDE_raw == pytest.approx(np.sqrt(P - 1), rel=1e-12)
    assert SDE == pytest.approx(np.sqrt(P - 1), rel=1e-12)
    assert int(np.argmax(power_raw)) == dip and power_raw[dip] == pytest.approx(SDE_raw, rel=1e-12)
    assert int(np.argmax(power)) == dip and power[dip] == pytest.approx(SDE, rel=1e-12)


def test_detrend_removes_a_smooth_trend():
    """A slow linear trend in SR is taken out by the running median: the dip stays the
    peak of the detrended power while the raw power peaks elsewhere."""
    P = 2000
    x = np.linspace(0, 1, P)
    chi2 = 1000.0 / (1 + 0.5 * x)          # SR rises along the grid
    dip = 300
    chi2[dip] *= 0.9
    _, power_raw, power, _, _ = SP.spectra(chi2, 3)
    assert int(np.argmax(power)) == dip
    assert int(np.argmax(power_raw)) != dip


def test_mask_outliers():
    x = np.array([1.0, 2.0, 3.0, 199.0, 2.0])
    assert not SP.mask_outliers(x).mask.any()            # 199 < 100 * median (2)
    x[3] = 201.0
    assert list(SP.mask_outliers(x).mask) == [False, False, False, True, False]


def test_all_transit_times_is_the_grid_within_the_span():
    t = np.linspace(100.0, 130.0, 5000)
    tt = ST.all_transit_times(101.3, t, 2.5)
    assert np.allclose(np.diff(tt), 2.5) and tt[0] == 101.3 and tt[-1] < 130.0
    assert tt[-1] + 2.5 >= 130.0
    # a T0 before the data starts one period later
    tt2 = ST.all_transit_times(98.8, t, 2.5)
    assert tt2[0] == pytest.approx(101.3)


def test_transit_mask_brute_force():
    rng = np.random.default_rng(3)
    t = np.sort(rng.uniform(0, 50, 3000))
    P, dur, T0 = 3.7, 0.2, 1.1
    centers = T0 + P * np.arange(-2, 20)
    dist = np.min(np.abs(t[:, None] - centers[None, :]), axis=1)
    assert np.array_equal(ST.transit_mask(t, P, dur, T0), dist < 0.5 * dur)


def test_duration_days_of_uniform_phases():
    """Uniform sampling: the raw duration (fraction of the period) maps to raw x P days,
    to one sample."""
    P = 2.0
    t = np.arange(0, 200.0, 1 / 720)
    d = ST.duration_days(t, P, 0.3, 0.05)
    assert d == pytest.approx(0.05 * P, abs=2 * P / len(t) * 1.0 + 1e-9)


def test_pink_noise_of_white_noise():
    """For white noise the mean per-window std over sqrt(width) is ~ sigma / sqrt(width)."""
    x = np.random.default_rng(5).normal(0, 2.0, 20000)
    assert ST.pink_noise(x, 25) == pytest.approx(2.0 / 5.0, rel=0.03)


def test_candidate_indices_top_then_long_periods():
    periods = np.linspace(0.5, 20.0, 1000)
    power = np.random.default_rng(9).normal(size=1000)
    idx = candidate_indices(periods, power, 100, 100)
    assert len(idx) == 200 and len(set(idx)) == 200
    order = np.argsort(-power, kind="stable")
    assert idx[:100] == list(order[:100])
    nxt = idx[100:]
    assert all(periods[i] > 1 for i in nxt)
    # the next 100 are the best remaining long-period ones, in power order
    rest = [i for i in order if i not in set(order[:100]) and periods[i] > 1][:100]
    assert nxt == rest
    assert nearest_indices([0.5, 20.0, 10.0], periods) == [0, 999, int(np.argmin(abs(periods - 10)))]


def test_refine_period_box_transit_closed_forms():
    """refine_period end to end on a box transit of known period, T0, duration and depth,
    from the oracle's best (width, T0) of that period: transit times P apart across the
    span, T0 and duration to a few cadences, and depth = delta * overshoot * (box samples in
    the window / window samples)."""
    from gtls_amd import search as S
    from gtls_amd.validate import validate_args
    from oracle import oracle
    rng = np.random.default_rng(3)
    dt = 2 / 1440
    t = 1000 + np.arange(6000) * dt
    P, T0, D, delta = 1.7, 1000.5, 0.1, 0.01
    ph = ((t - T0 + P / 2) % P) - P / 2
    inbox = np.abs(ph) < D / 2
    y = 1 - delta * inbox + rng.normal(0, 1e-5, t.size)
    dy = np.full(t.size, 1e-3)
    plan = S.make_plan(t, y, dy, validate_args({}), periods=[P])
    tb = plan.table
    _, _, k, s = oracle.period_table(plan.t, plan.y, plan.dy, P, tb.widths, tb.S, tb.offsets,
                                     tb.overshoot, plan.k_lo[0], plan.k_hi[0], skip_point=8)
    r = ST.refine_period(plan, P, k, s)
    tt = np.asarray(r["transitTimes"])
    assert np.allclose(np.diff(tt), P, rtol=0, atol=1e-9)
    assert t.min() <= tt[0] < t.min() + P and tt[-1] < t.max()
    assert len(tt) == int(np.floor((t.max() - tt[0]) / P)) + 1
    dT0 = ((r["T0"] - T0 + P / 2) % P) - P / 2
    assert abs(dT0) <= 4 * dt
    assert abs(r["duration"] - D) <= 4 * dt
    width = int(tb.widths[k])
    assert r["durationPoints"] == width and width >= inbox.sum()
    # the window's mean holds n_in box samples of depth delta: depth / (delta overshoot)
    # times the width is an integer, at most the folded box and short of it only by the
    # samples of a few-cadence T0 offset
    n_in = r["depth"] / (delta * tb.overshoot[k]) * width
    assert abs(n_in - round(n_in)) < 0.1
    assert inbox.sum() - 4 * len(tt) <= round(n_in) <= inbox.sum()
    for key in ("snr", "snrPink", "snrFit", "snrFitPink"):
        assert np.isfinite(r[key]) and r[key] > 0
