"""Pins of the FP64 CPU oracle (oracle/tls_oracle.c) on the CPU.

Each check is chosen so that a plausible restatement mistake fails it: an off-by-one in a
prefix-sum index or in the out-of-transit lookup, a wrong sign of the edge correction, a
mis-scaled template, an unstable sort, a wrong skip stride or argmin order.
"""
import numpy as np
import pytest

from oracle import oracle

SKIP_NONE = oracle.SKIP_NONE


def stable_perm(t, P):
    x = t / P
    return np.argsort(x - np.trunc(x), kind="stable")


def brute_stat(t, y, dy, P, d, S_row, os_, s, skip, depth_min, wmax):
    """The TLS statistic of one (width, position) by direct sums, no prefix sums:
    in-transit residual against the scaled template + residual of every other patched
    sample against 1 - residual of the wrapped tail (the edge correction)."""
    n = len(t)
    perm = stable_perm(t, P)
    idx = perm[np.arange(n + wmax) % n]
    Y = y[idx]
    W = 1.0 / dy[idx] ** 2
    win = np.arange(s, s + d)
    mean = 1.0 - Y[win].mean()
    if not (mean > depth_min and s % skip == 0):
        return float(n)
    r = 2.0 * mean * os_
    model = 1.0 - S_row[:d] * r
    inside = np.sum((Y[win] - model) ** 2 * W[win])
    mask = np.ones(n + wmax, dtype=bool)
    mask[win] = False
    outside = np.sum((Y[mask] - 1.0) ** 2 * W[mask])
    edge = np.sum((Y[n:] - 1.0) ** 2 * W[n:])
    return inside + outside - edge


def random_case(rng, n=61, widths=(3, 5, 8, 12, 20), dip=True):
    t = np.sort(rng.uniform(1.0, 30.0, n))
    y = 1.0 + rng.normal(0, 1e-3, n)
    if dip:
        y[rng.choice(n, 6, replace=False)] -= 3e-3
    dy = rng.uniform(5e-4, 2e-3, n)
    widths = np.asarray(widths, dtype=np.int32)
    offs = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    S = rng.uniform(0.0, 0.5, int(offs[-1]))
    overshoot = rng.uniform(0.9, 1.3, len(widths))
    return t, y, dy, widths, S, offs, overshoot


@pytest.mark.parametrize("skip", [1, 2, 8, SKIP_NONE])
def test_period_table_matches_direct_sums(skip):
    rng = np.random.default_rng(11 + (skip % 97))
    t, y, dy, widths, S, offs, os_ = random_case(rng)
    wmax = int(widths.max()) + int(widths.max()) % 2
    for P in (0.731, 2.9, 7.77):
        table, chi2, bk, bs = oracle.period_table(t, y, dy, P, widths, S, offs, os_, 0,
                                                  len(widths) - 1, skip_point=skip,
                                                  depth_min=1e-5)
        for k, d in enumerate(widths):
            q = d // skip if d > skip else 1
            for s in range(len(t)):
                want = brute_stat(t, y, dy, P, int(d), S[offs[k]:offs[k + 1]], os_[k], s, q,
                                  1e-5, wmax)
                assert table[k, s] == pytest.approx(want, rel=1e-10, abs=1e-9), (P, k, s)
        flat = table.ravel()
        assert chi2 == flat.min()
        assert bk * len(t) + bs == int(np.argmin(flat))


def test_search_is_table_minimum_and_respects_ranges():
    rng = np.random.default_rng(5)
    t, y, dy, widths, S, offs, os_ = random_case(rng, n=80)
    periods = np.array([0.9, 1.3, 2.2, 4.1, 6.0])
    k_lo = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    k_hi = np.array([2, 3, 4, 4, 3], dtype=np.int32)  # last one empty
    chi2, bk, bs = oracle.search(t, y, dy, periods, widths, S, offs, os_, k_lo, k_hi,
                                 skip_point=8)
    for i, P in enumerate(periods):
        if k_lo[i] > k_hi[i]:
            assert np.isnan(chi2[i]) and bk[i] == -1 and bs[i] == -1
            continue
        table, c, k, s = oracle.period_table(t, y, dy, P, widths, S, offs, os_, k_lo[i],
                                             k_hi[i], skip_point=8)
        assert chi2[i] == c and bk[i] == k and bs[i] == s
        assert k_lo[i] <= bk[i] <= k_hi[i]


def test_fold_sort_is_stable_argsort_with_ties_and_negative_times():
    rng = np.random.default_rng(3)
    t = np.sort(rng.uniform(0.5, 100.0, 500))
    for P in (0.37, 1.0, 3.3, 17.0):
        assert np.array_equal(oracle.fold_sort(t, P), stable_perm(t, P))
    # exact resonance: cadence divides the period, every phase repeats many times
    tc = np.arange(2000) * 0.25
    for P in (1.0, 2.5, 0.75):
        ph = tc / P - np.trunc(tc / P)
        assert len(np.unique(ph)) < 100  # thousands of samples on a few dozen keys
        assert np.array_equal(oracle.fold_sort(tc, P), stable_perm(tc, P))
    tn = rng.uniform(-50, 50, 300)
    assert np.array_equal(oracle.fold_sort(tn, 3.7), stable_perm(tn, 3.7))


def test_constant_flux_closes_every_gate():
    n = 50
    t = np.linspace(1, 20, n)
    y = np.ones(n)
    dy = np.full(n, 1e-3)
    widths = np.array([4, 6], dtype=np.int32)
    offs = np.array([0, 4, 10], dtype=np.int64)
    S = np.full(10, 0.4)
    chi2, bk, bs = oracle.search(t, y, dy, [2.0, 3.0], widths, S, offs, [1.0, 1.0],
                                 [0, 1], [1, 1], skip_point=SKIP_NONE)
    assert np.all(chi2 == n)
    assert list(bk) == [0, 1] and list(bs) == [0, 0]


def test_box_template_reduces_to_box_least_squares():
    """S = 0.5, overshoot = 1: the model in the window is its own mean, so the in-transit
    residual is the weighted scatter about the window mean (the BLS box step)."""
    rng = np.random.default_rng(8)
    n = 70
    t = np.sort(rng.uniform(0, 10, n))
    y = 1 + rng.normal(0, 1e-3, n)
    y[10:20] -= 4e-3
    dy = np.full(n, 1e-3)
    d = 9
    widths = np.array([d], dtype=np.int32)
    offs = np.array([0, d], dtype=np.int64)
    P = 4.4
    table, *_ = oracle.period_table(t, y, dy, P, widths, np.full(d, 0.5), offs, [1.0], 0, 0,
                                    skip_point=SKIP_NONE, depth_min=-1.0)
    Y = y[stable_perm(t, P)]
    Ye = np.concatenate([Y, Y[: d + 1]])
    w = 1 / 1e-3 ** 2
    flat = np.sum((1 - Y) ** 2) * w
    for s in range(n):
        win = Ye[s:s + d]
        box = flat - np.sum((1 - win) ** 2) * w + np.sum((win - win.mean()) ** 2) * w
        assert table[0, s] == pytest.approx(box, rel=1e-10)


def test_injected_box_found_at_exact_width_and_position():
    """Noise-free box transit of exactly one width of phase-sorted samples: the statistic
    is zero exactly where the window of that width covers the dip."""
    n, P, d, first = 240, 2.0, 12, 100
    t = 0.05 + np.arange(n) * (10.0 / n)
    perm = stable_perm(t, P)
    y = np.ones(n)
    y[perm[first:first + d]] = 1 - 2e-3
    dy = np.full(n, 1e-3)
    widths = np.array([6, 12, 24], dtype=np.int32)
    offs = np.concatenate([[0], np.cumsum(widths)]).astype(np.int64)
    S = np.full(int(offs[-1]), 0.5)
    chi2, bk, bs = oracle.search(t, y, dy, [P], widths, S, offs, [1.0, 1.0, 1.0], [0], [2],
                                 skip_point=SKIP_NONE)
    assert chi2[0] == pytest.approx(0.0, abs=1e-9)
    assert bk[0] == 1 and bs[0] == first


def test_skip_positions_are_a_subset():
    rng = np.random.default_rng(21)
    t, y, dy, widths, S, offs, os_ = random_case(rng, n=90, widths=(9, 16, 24, 32))
    periods = np.array([1.1, 2.3, 3.9])
    lo = np.zeros(3, np.int32)
    hi = np.full(3, 3, np.int32)
    c_all, *_ = oracle.search(t, y, dy, periods, widths, S, offs, os_, lo, hi,
                              skip_point=SKIP_NONE)
    c_8, *_ = oracle.search(t, y, dy, periods, widths, S, offs, os_, lo, hi, skip_point=8)
    assert np.all(c_8 >= c_all)
    table8, *_ = oracle.period_table(t, y, dy, periods[0], widths, S, offs, os_, 0, 3,
                                     skip_point=8)
    for k, d in enumerate(widths):
        q = d // 8
        assert np.all(table8[k, 1::q][(np.arange(1, len(t))[::q] % q) != 0] == len(t))
        off = np.arange(len(t)) % q != 0
        assert np.all(table8[k, off] == len(t))


def test_zero_width_rows_are_sentinels():
    rng = np.random.default_rng(2)
    t, y, dy, _, _, _, _ = random_case(rng, n=30)
    widths = np.array([0, 4], dtype=np.int32)
    offs = np.array([0, 0, 4], dtype=np.int64)
    table, chi2, bk, bs = oracle.period_table(t, y, dy, 1.7, widths, np.full(4, 0.3), offs,
                                              [1.0, 1.0], 0, 1, skip_point=8)
    assert np.all(table[0] == 30)
