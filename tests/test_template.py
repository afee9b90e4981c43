"""Pins of the engine's limb-darkened occultation model and template table.

batman (the reference's template source) is not available, so the model is pinned by the
mathematics: the uniform-source closed form (Mandel & Agol 2002, eq. 1), adaptive
quadrature of the same radial integral, a brute-force 2-D integration over the stellar
disk, and the small-planet limit. The template pipeline is pinned by its invariants.
"""
import numpy as np
import pytest
from scipy import integrate

from gtls_amd import constants as C
from gtls_amd import template as T


def uniform_closed_form(z, p):
    if z >= 1 + p:
        return 1.0
    if z <= 1 - p:
        return 1 - p * p
    if z <= p - 1:
        return 0.0
    k0 = np.arccos((p * p + z * z - 1) / (2 * p * z))
    k1 = np.arccos((1 - p * p + z * z) / (2 * z))
    lam = (p * p * k0 + k1 - 0.5 * np.sqrt(4 * z * z - (1 + z * z - p * p) ** 2)) / np.pi
    return 1 - lam


def quad_flux(z, p, u1, u2):
    def I(r):
        mu = np.sqrt(max(0.0, 1 - r * r))
        return 1 - u1 * (1 - mu) - u2 * (1 - mu) ** 2

    def theta(r):
        if r <= p - z:
            return np.pi
        if r >= z + p or r <= z - p:
            return 0.0
        return np.arccos(np.clip((r * r + z * z - p * p) / (2 * r * z), -1, 1))

    pts = [x for x in (abs(z - p), z + p) if 0 < x < 1]
    b, _ = integrate.quad(lambda r: I(r) * 2 * r * theta(r), 0, 1, points=pts or None,
                          epsabs=1e-15, epsrel=1e-13, limit=500)
    return 1 - b / (np.pi * (1 - u1 / 3 - u2 / 6))


def test_uniform_source_closed_form():
    for p in (0.01, 0.05, 0.1, 0.3):
        zs = np.concatenate([np.linspace(1e-3, 1 + p - 1e-9, 60), [1 - p, p, 1.0]])
        got = T.occultation_flux(zs, p, (0.0, 0.0), "quadratic")
        want = np.array([uniform_closed_form(z, p) for z in zs])
        assert np.max(np.abs(got - want)) < 1e-11


@pytest.mark.parametrize("u", [(0.4804, 0.1867), (0.45, 0.20), (0.9, -0.2), (0.3, 0.0)])
def test_quadratic_against_adaptive_quadrature(u):
    for p in (0.02, 0.1, 0.25):
        zs = np.concatenate([np.linspace(0.01, 1 + p - 1e-6, 25), [1 - p, 1 + p - 1e-8, p]])
        got = T.occultation_flux(zs, p, u, "quadratic")
        want = np.array([quad_flux(z, p, *u) for z in zs])
        deficit = np.maximum(1 - want, 1e-30)
        assert np.max(np.abs(got - want) / deficit) < 1e-9


def test_brute_force_disk_integration():
    """Sum of I over a fine Cartesian grid of the stellar disk, the planet masked out."""
    u1, u2, p, n = 0.4804, 0.1867, 0.15, 3000
    x = (np.arange(n) + 0.5) / n * 2 - 1
    X, Y = np.meshgrid(x, x)
    R2 = X * X + Y * Y
    on = R2 < 1
    mu = np.sqrt(np.where(on, 1 - R2, 0))
    I = np.where(on, 1 - u1 * (1 - mu) - u2 * (1 - mu) ** 2, 0)
    total = I.sum()
    for z in (0.0, 0.3, 0.8, 0.95, 1.05):
        blocked = I[(X - z) ** 2 + Y ** 2 < p * p].sum()
        want = 1 - blocked / total
        got = T.occultation_flux([z], p, (u1, u2), "quadratic")[0]
        assert abs(got - want) < 2e-5 * p * p / 0.0225 + 1e-6


def test_small_planet_limit():
    """p -> 0: deficit = p^2 I(mu(z)) / (1 - u1/3 - u2/6)."""
    u1, u2, p = 0.45, 0.2, 1e-3
    for z in (0.0, 0.5, 0.9):
        mu = np.sqrt(1 - z * z)
        want = p * p * (1 - u1 * (1 - mu) - u2 * (1 - mu) ** 2) / (1 - u1 / 3 - u2 / 6)
        got = 1 - T.occultation_flux([z], p, (u1, u2), "quadratic")[0]
        assert got == pytest.approx(want, rel=3e-3)


def test_no_overlap_is_exactly_one():
    got = T.occultation_flux([1.2, 1.0300001, 5.0], 0.03)
    assert np.all(got == 1.0)


def test_interp_linear_matches_reference_bracketing():
    x = np.linspace(-0.5, 0.5, 11)
    y = x ** 2
    xn = np.array([-0.5, -0.43, 0.0, 0.31, 0.5, 0.6, -0.55])
    got = T.interp_linear(xn, x, y)
    for v, g in zip(xn, got):
        i = int(np.clip(np.searchsorted(x, v, "right") - 1, 0, len(x) - 2))
        th = (v - x[i]) / (x[i + 1] - x[i])
        assert g == (1 - th) * y[i] + th * y[i + 1]


def test_template_table_invariants():
    from gtls_amd.grid import duration_grid, period_grid
    n = 19728
    periods = period_grid(1.0, 1.0, 28.4)
    durs = duration_grid(periods)
    tb, overview, rows = T.build_templates(n, durs)
    w = tb.widths
    assert np.all(np.diff(w) > 0)
    assert w[-1] == tb.wmax_even or w[-1] + 1 == tb.wmax_even
    assert tb.wmax_even % 2 == 0
    assert len(tb.S) == int(w.astype(np.int64).sum())
    for k in range(tb.n_widths):
        row = tb.S[tb.offsets[k]:tb.offsets[k + 1]]
        assert np.all(row >= 0) and np.all(row <= 0.5 + 1e-12)
        assert row.max() == pytest.approx(0.5, abs=1e-3)
        sig = 1 - row[: len(tb.rows[k])]
        assert tb.overshoot[k] == pytest.approx(1 / (2 - np.mean(sig) / np.min(sig)), rel=1e-12)
    # the trimmed rows fill their widths (the template occupies exactly its width)
    assert all(len(r) == d for r, d in zip(tb.rows, w))


# --- the resampling, trimming and overshoot steps (src/gputls/transit.py:45-160) fed with
# analytic reference profiles whose rows and overshoot are known in closed form


def test_box_reference_rows_and_overshoot():
    """A box reference (unit depth everywhere in transit): every row is the box at the
    signal depth, width_in_samples = int(duration / max * samples) samples long, with
    overshoot 1 (mean = min)."""
    from gtls_amd import constants as C
    from gtls_amd.template import cache_rows
    samples = 1200
    durations = np.array([0.01, 0.013, 0.02, 0.035, 0.05])
    ov, rows = cache_rows(durations, samples, np.zeros(samples))
    for k, d in enumerate(durations):
        occ = int(d / durations.max() * samples)
        assert ov["width_in_samples"][k] == occ
        assert len(rows[k]) == occ
        assert np.allclose(rows[k], 1 - C.SIGNAL_DEPTH, rtol=0, atol=1e-15)
        assert ov["overshoot"][k] == pytest.approx(1.0, abs=1e-15)


@pytest.mark.parametrize("h", [40, 125, 300])
def test_triangle_reference_rows_and_overshoot(h):
    """A V-shaped reference r(x) = |2 x| on x in [-1/2, 1/2] (unit depth at the centre),
    sampled on an odd grid so that its linear interpolant is the V itself: the row of
    2h + 1 occupied samples is 1 - (1 - r_j) D at r_j = |j - h| / h without its two end
    samples (r = 1, above the trimming cut), so mean(1 - r) = h / (2h - 1) over the row and
    overshoot = 1 / (2 - (1 - D h / (2h - 1)) / (1 - D))."""
    from gtls_amd import constants as C
    from gtls_amd.template import cache_rows
    samples = 2001
    ref = np.abs(2 * np.linspace(-0.5, 0.5, samples))
    occ = 2 * h + 1
    dmax = 0.1
    dur = dmax * (occ + 0.5) / samples          # int(dur / dmax * samples) = occ
    ov, rows = cache_rows(np.array([dur, dmax]), samples, ref)
    D = C.SIGNAL_DEPTH
    j = np.arange(1, occ - 1)
    want = 1 - (1 - np.abs(j - h) / h) * D
    assert ov["width_in_samples"][0] == occ
    assert len(rows[0]) == occ - 2
    assert np.allclose(rows[0], want, rtol=0, atol=1e-12)
    over = 1 / (2 - (1 - D * h / (2 * h - 1)) / (1 - D))
    assert ov["overshoot"][0] == pytest.approx(over, rel=1e-12)
