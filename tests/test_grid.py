"""gtls_amd.grid against the reference's own grid outputs (tests/golden/grid_vectors.npz,
written by tests/golden/make_grid_vectors.py) and closed-form properties."""
import os
import warnings

import numpy as np
import pytest

from gtls_amd import constants as C
from gtls_amd import grid

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "grid_vectors.npz"))


def n_cases():
    return len([k for k in GOLD.files if k.endswith("_in")])


@pytest.mark.parametrize("i", range(7))
def test_period_and_duration_grid_bit_exact(i):
    R, M, span, pmin, pmax, os_, ntm = GOLD["case%d_in" % i]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p = grid.period_grid(R_star=R, M_star=M, time_span=span, period_min=pmin,
                             period_max=pmax, oversampling_factor=int(os_),
                             n_transits_min=int(ntm))
    assert p.size == int(GOLD["case%d_n_periods" % i][0])
    idx = GOLD["case%d_index" % i]
    assert np.array_equal(p[idx], GOLD["case%d_periods" % i])
    d = grid.duration_grid(p, shortest=1.0 / 1000, log_step=1.1)
    assert np.array_equal(np.asarray(d), GOLD["case%d_durations" % i])
    d2 = grid.duration_grid(p, shortest=1.0 / 1000, log_step=1.05)
    assert np.array_equal(np.asarray(d2), GOLD["case%d_durations_step105" % i])


def test_t14_bit_exact():
    for R, M, P, small, want in GOLD["t14"]:
        assert grid.T14(R, M, P, small=bool(small)) == want


def test_width_window_closed_form():
    """[lo, hi] are floor/ceil of T14(R=0.05, M=0.05, small) * N and
    T14(R=4, M=1, +2 R_jup) * N * (L/P + 1)/(L/P), both fractions capped at 0.15."""
    P = np.array([0.5, 1.0, 3.71, 13.7, 45.3, 300.0])
    L, N = 1461.0, 65000
    lo, hi = grid.width_window(P, L, N)
    for i, p in enumerate(P):
        s = p * 86400.0
        gm_min = np.pi * C.G * 0.05 * C.M_SUN
        gm_max = np.pi * C.G * 1.0 * C.M_SUN
        fmin = min(0.15, 0.05 * C.R_SUN * (4 * s / gm_min) ** (1 / 3) / s)
        fmax = min(0.15, (4.0 * C.R_SUN + 2 * C.R_JUP) * (4 * s / gm_max) ** (1 / 3) / s)
        corr = (L / p + 1) / (L / p)
        assert abs(lo[i] - fmin * N) < 1.0 + 1e-6 * N
        assert abs(hi[i] - fmax * N * corr) < 1.0 + 1e-6 * N
        assert lo[i] <= hi[i]


def test_width_ranges_are_inclusive_index_windows():
    widths = np.array([3, 5, 8, 13, 21, 34], dtype=np.int32)
    lo = np.array([0, 4, 5, 9, 22, 40, 13])
    hi = np.array([100, 12, 5, 13, 33, 50, 13])
    k_lo, k_hi = grid.width_ranges(widths, lo, hi)
    for a, b, kl, kh in zip(lo, hi, k_lo, k_hi):
        sel = np.nonzero((widths >= a) & (widths <= b))[0]
        if sel.size:
            assert (kl, kh) == (sel[0], sel[-1])
        else:
            assert kl > kh


def test_skip_factor():
    assert grid.skip_factor(0.125) == 8
    assert grid.skip_factor(0.01) == 100
    assert grid.skip_factor(0) == C.SKIP_NONE


def test_periodogram_is_the_grid_power_searches():
    """gtls.periodogram (main.py:186-206) returns the period grid power() searches for the
    same kwargs (in period_grid's own descending order; the search keeps it ascending), and
    the user's periods unchanged when given."""
    from gtls_amd.api import gtls
    from gtls_amd.search import make_plan
    from gtls_amd.validate import validate_args
    rng = np.random.default_rng(3)
    t = 1000.0 + np.arange(4000) / 144.0
    y = 1.0 + rng.normal(0.0, 1e-3, t.size)
    g = gtls(t, y, verbose=False)
    kw = dict(oversampling_factor=2, period_min=0.7)
    p = g.periodogram(**kw)
    plan = make_plan(g.t, g.y, g.dy, validate_args(kw))
    assert len(p) > 100 and np.all(np.diff(p) < 0)
    assert np.array_equal(p[::-1], plan.periods)
    user = np.array([1.1, 2.2, 3.3])
    assert np.array_equal(g.periodogram(periods=user), user)
