"""The SNR fields of refine_period (src/gputls/core.py:1436-1496, src/gputls/stats.py:
433-488) on a constructed light curve whose phase-sorted flux is known in closed form.

The curve has m phase bins per period P and R cycles, one sample per bin and cycle at
t = (j + 1/4) P / m + 10 P. Bins k0 .. k0 + L - 1 hold a box transit of depth delta; every
other bin k holds 1 + sigma (-1)^k. After the fold, the sorted flux is R copies of each bin
in bin order, so for the window of L R samples starting at sorted position k0 R:
  * depth = delta * overshoot (the window mean is 1 - delta);
  * the samples outside the window (with the wrap-around copy of the first samples) take
    the two values 1 +- sigma, n+ and n- times, so their std is 2 sigma sqrt(n+ n-) / n and
    snrFit = (1 - depth) sqrt(L R) / std;
  * with L and m even those samples alternate in blocks of R, so every sliding mean over
    L R of them is exactly 1: the red noise is zero and snrFitPink = snrFit;
  * each transit's time window holds box samples only (all but one bin of the box), so snr
    is delta over the std of the out-of-transit flux (three values in known counts) times
    sqrt(in-transit count), and snrPink is delta over the pink noise of that flux times
    sqrt(number of transits); the pink noise itself is pinned on an alternating sequence.
"""
import numpy as np
import pytest

from gtls_amd import search as S
from gtls_amd import stats as ST
from gtls_amd.validate import validate_args

P = 2.0
DELTA, SIGMA = 4e-3, 1e-3


def build(m, R):
    j = np.arange(m * R)
    t = (j + 0.25) * (P / m) + 10 * P
    return t, j % m


def choose(m, R):
    """A plan of the curve and a width of its template table that is L R samples with L
    even, placed at bins k0 .. k0 + L - 1 clear of the wrap-around copy."""
    t, _ = build(m, R)
    plan = S.make_plan(t, np.ones_like(t), np.full(t.size, SIGMA), validate_args({}),
                       periods=[P])
    tb = plan.table
    wrap_bins = -(-tb.wmax_even // R) + 1
    for k, w in enumerate(tb.widths):
        w = int(w)
        if w % R == 0 and (w // R) % 2 == 0 and 4 <= w // R <= m // 4:
            k0 = wrap_bins + (wrap_bins % 2)
            if k0 + w // R < m - 2:
                return k, w // R, k0
    return None


@pytest.mark.parametrize("m,R", [(400, 4), (600, 3), (512, 8)])
def test_snr_fields_closed_forms(m, R):
    pick = choose(m, R)
    assert pick is not None, "no template width fits the construction"
    k, L, k0 = pick
    t, kbin = build(m, R)
    box = (kbin >= k0) & (kbin < k0 + L)
    y = np.where(box, 1 - DELTA, 1 + SIGMA * (-1.0) ** kbin)
    plan = S.make_plan(t, y, np.full(t.size, SIGMA), validate_args({}), periods=[P])
    tb = plan.table
    assert int(tb.widths[k]) == L * R
    r = ST.refine_period(plan, P, k, k0 * R)
    w = L * R
    depth = DELTA * tb.overshoot[k]
    assert r["depth"] == pytest.approx(depth, rel=1e-12)
    # counts of +sigma and -sigma among the sorted positions outside the window, the
    # wrap-around copy of the first tb.wmax_even sorted samples included
    n = m * R
    pos = np.concatenate([np.arange(0, k0 * R), np.arange(k0 * R + w, n + tb.wmax_even)])
    bins = (pos % n) // R
    n_plus = int(np.sum(bins % 2 == 0))
    n_minus = int(np.sum(bins % 2 == 1))
    assert n_plus + n_minus == n + tb.wmax_even - w
    std_out = 2 * SIGMA * np.sqrt(n_plus * n_minus) / (n_plus + n_minus)
    snr_fit = (1 - depth) * np.sqrt(w) / std_out
    assert r["snrFit"] == pytest.approx(snr_fit, rel=1e-9)
    assert r["snrFitPink"] == pytest.approx(snr_fit, rel=1e-9)
    # transits: R of them, P apart; the in-transit selection holds box samples only, up to
    # one bin at each edge
    tt = np.asarray(r["transitTimes"])
    assert len(tt) == R and np.allclose(np.diff(tt), P)
    _, _, count = ST.intransit_stats(t, y, tt, r["duration"])
    assert np.all(np.abs(count - L) <= 1), count
    odd, even, _ = ST.intransit_stats(t, y, tt, r["duration"])
    sel = np.concatenate([odd, even])
    assert np.all(sel == 1 - DELTA)              # box samples only
    n_in = float(sel.size)
    # the out-of-transit flux (transit_mask over twice the raw duration, pinned in
    # test_spectra_stats) holds the three values 1 + sigma, 1 - sigma, 1 - delta (box
    # samples at the mask's edge) in counts a, b, c: its std in closed form
    ootr = y[~ST.transit_mask(t, P, 2 * tb.durations[k], r["T0"])]
    vals = np.array([1 + SIGMA, 1 - SIGMA, 1 - DELTA])
    cnt = np.array([np.sum(ootr == v) for v in vals], dtype=float)
    assert cnt.sum() == ootr.size
    mu = cnt @ vals / cnt.sum()
    std_ootr = np.sqrt(cnt @ (vals - mu) ** 2 / cnt.sum())
    assert r["snr"] == pytest.approx(DELTA / std_ootr * np.sqrt(n_in), rel=1e-9)
    # snrPink: the per-transit deficit (delta, box samples only) over the pink noise of the
    # out-of-transit flux at the mean transit width, times sqrt(number of transits)
    wd = int(np.mean(count))
    assert r["snrPink"] == pytest.approx(DELTA / ST.pink_noise(ootr, wd) * np.sqrt(R),
                                         rel=1e-12)


@pytest.mark.parametrize("width", [6, 7, 24, 25])
def test_pink_noise_of_an_alternating_sequence(width):
    """+-sigma alternating: a window of even width has std sigma, of odd width
    sigma sqrt(1 - 1/width^2), so pink noise = that / sqrt(width)."""
    x = 1 + SIGMA * (-1.0) ** np.arange(5000)
    want = SIGMA * (1.0 if width % 2 == 0 else np.sqrt(1 - 1 / width ** 2)) / np.sqrt(width)
    assert ST.pink_noise(x, width) == pytest.approx(want, rel=1e-9)
