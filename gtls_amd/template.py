"""Limb-darkened transit templates for the search.

The reference builds its templates with batman (src/gputls/transit.py:8-160), which is
not part of this engine. Here the occultation of a limb-darkened star by an opaque disk
is computed by the engine's own model: the blocked flux is the radial integral

    B(z, p) = int_0^1 I(r) * 2 r * theta(r) dr,
    theta(r) = arccos((r^2 + z^2 - p^2) / (2 r z))  inside the overlap annulus,

with theta = pi on the part of the star the planet covers completely (closed form) and a
tanh-sinh quadrature on the partial annulus [|z - p|, min(1, z + p)], whose endpoint
square-root singularities it resolves. Limb-darkening laws: uniform, linear, quadratic.

The steps that turn the reference light curve into the per-width template rows (trim,
resample, depth rescale, overshoot) follow src/gputls/transit.py:8-160 step by step, so
that the template table is the one the reference would build from the same light curve.
"""
from __future__ import annotations

import functools

import numpy as np

from . import constants as C

# tanh-sinh nodes on (-1, 1): x_j = tanh(pi/2 sinh(t_j)), with 1 - |x_j| kept exactly
_TS_H = 1.0 / 32.0
_TS_T = np.arange(-4.0, 4.0 + _TS_H / 2, _TS_H)
_TS_U = 0.5 * np.pi * np.sinh(_TS_T)
_TS_X = np.tanh(_TS_U)
_TS_ONE_MINUS = 1.0 / (np.exp(np.abs(_TS_U)) * np.cosh(_TS_U))  # 1 - |x|
_TS_W = _TS_H * 0.5 * np.pi * np.cosh(_TS_T) / np.cosh(_TS_U) ** 2


def _ld_coeffs(u, law):
    """Quadratic-form coefficients (u1, u2) of I(mu) = 1 - u1 (1-mu) - u2 (1-mu)^2."""
    law = law.lower()
    u = list(np.atleast_1d(np.asarray(u, dtype=np.float64)))
    if law == "uniform":
        return 0.0, 0.0
    if law == "linear":
        return float(u[0]), 0.0
    if law == "quadratic":
        if len(u) != 2:
            raise ValueError("quadratic limb darkening needs two coefficients")
        return float(u[0]), float(u[1])
    raise ValueError("unsupported limb darkening law: %r (uniform, linear, quadratic)" % law)


def _intensity(r, u1, u2):
    mu = np.sqrt(np.maximum(0.0, 1.0 - r * r))
    om = 1.0 - mu
    return 1.0 - u1 * om - u2 * om * om


def _full_disk_integral(R1, u1, u2):
    """int_0^R1 I(r) 2 pi r dr in closed form (mu substitution)."""
    mu1 = np.sqrt(np.maximum(0.0, 1.0 - R1 * R1))

    def g(m):
        return (m * m / 2 - u1 * (m * m / 2 - m ** 3 / 3)
                - u2 * (m * m / 2 - 2 * m ** 3 / 3 + m ** 4 / 4))

    return 2 * np.pi * (g(1.0) - g(mu1))


def occultation_flux(z, p, u=C.DEFAULT_U, law=C.DEFAULT_LIMB_DARK):
    """Relative flux of a star with limb darkening `law`/`u` occulted by a dark disk of
    radius ratio `p` at projected separation `z` (both in stellar radii)."""
    u1, u2 = _ld_coeffs(u, law)
    z = np.abs(np.atleast_1d(np.asarray(z, dtype=np.float64)))
    p = float(p)
    total = np.pi * (1.0 - u1 / 3.0 - u2 / 6.0)
    blocked = np.zeros_like(z)

    # fully covered disk r < p - z
    full = z < p
    if np.any(full):
        blocked[full] += _full_disk_integral(np.minimum(1.0, p - z[full]), u1, u2)

    # partial annulus |z - p| < r < min(1, z + p)
    ra = np.abs(z - p)
    rb = np.minimum(1.0, z + p)
    part = (ra < rb) & (z > 0)
    if np.any(part):
        a = ra[part][:, None]
        b = rb[part][:, None]
        zz = z[part][:, None]
        half = 0.5 * (b - a)
        # node positions measured from the nearer endpoint to keep precision
        left = _TS_X < 0
        r = np.where(left, a + half * _TS_ONE_MINUS, b - half * _TS_ONE_MINUS)
        arg = (r * r + zz * zz - p * p) / (2.0 * r * zz)
        theta = np.arccos(np.clip(arg, -1.0, 1.0))
        f = _intensity(r, u1, u2) * 2.0 * r * theta
        blocked[part] += np.sum(f * _TS_W, axis=1) * half[:, 0]
    return 1.0 - blocked / total


def _sky_separation(t, t0, per, a, inc, ecc, w):
    """Projected star-planet separation in stellar radii and a front-side flag."""
    inc_r = np.radians(inc)
    w_r = np.radians(w)
    if ecc == 0:
        f_conj = np.pi / 2 - w_r
        f = 2 * np.pi * (t - t0) / per + f_conj
        r = a * np.ones_like(f)
    else:
        f_conj = np.pi / 2 - w_r
        E_conj = 2 * np.arctan(np.sqrt((1 - ecc) / (1 + ecc)) * np.tan(f_conj / 2))
        M_conj = E_conj - ecc * np.sin(E_conj)
        M = 2 * np.pi * (t - t0) / per + M_conj
        E = M.copy()
        for _ in range(50):
            E = E - (E - ecc * np.sin(E) - M) / (1 - ecc * np.cos(E))
        f = 2 * np.arctan(np.sqrt((1 + ecc) / (1 - ecc)) * np.tan(E / 2))
        r = a * (1 - ecc * np.cos(E))
    s = np.sin(w_r + f)
    d = r * np.sqrt(np.maximum(0.0, 1.0 - s * s * np.sin(inc_r) ** 2))
    return d, s > 0


def transit_lightcurve(t, t0, per, rp, a, inc, ecc, w, u, limb_dark):
    """Model flux of one planet on a Keplerian orbit (mid-transit at t0)."""
    t = np.asarray(t, dtype=np.float64)
    d, front = _sky_separation(t, t0, per, a, inc, ecc, w)
    flux = np.ones_like(d)
    m = front & (d < 1.0 + rp)
    if np.any(m):
        flux[m] = occultation_flux(d[m], rp, u, limb_dark)
    return flux


def interp_linear(x_new, x, y):
    """Linear interpolation with the reference's bracketing and weights
    (src/gputls/interpolation.py:8-58): index i with x[i] <= z < x[i+1], clamped to
    [0, n-2] (ends extrapolate), value (1 - theta) * y[i] + theta * y[i+1]."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    x_new = np.asarray(x_new, dtype=np.float64)
    n = len(x)
    i = np.clip(np.searchsorted(x, x_new, side="right") - 1, 0, n - 2)
    theta = (x_new - x[i]) / (x[i + 1] - x[i])
    return (1 - theta) * y[i] + theta * y[i + 1]


@functools.lru_cache(maxsize=16)
def _supersampled(per, rp, a, inc, ecc, w, u, limb_dark):
    """Model light curve on the reference's unit-width supersample grid."""
    t = np.linspace(-0.5, 0.5, C.SUPERSAMPLE_SIZE)
    flux = transit_lightcurve(t, 0.0, per, rp, a, inc, ecc, w, u, limb_dark)
    t.setflags(write=False)
    flux.setflags(write=False)
    return t, flux


def reference_transit(samples, per, rp, a, inc, ecc, w, u, limb_dark):
    """Unit-width, unit-depth template of `samples` points (src/gputls/transit.py:8-42)."""
    t, flux = _supersampled(float(per), float(rp), float(a), float(inc), float(ecc),
                            float(w), tuple(float(v) for v in np.atleast_1d(u)), limb_dark)
    first = int(np.argmax(flux < 1))
    in_flux = flux[first: -first + 1]
    in_time = t[first: -first + 1]
    x_new = np.linspace(t[first], t[-first - 1], int(samples))
    down = interp_linear(x_new, in_time, in_flux)
    lo = np.min(down)
    return (lo - down) / (lo - 1)


def fractional_transit(duration, maxwidth, depth, samples, reference):
    """Template of a fractional width and depth (src/gputls/transit.py:45-95)."""
    ref_time = np.linspace(-0.5, 0.5, samples)
    occupied = int((duration / maxwidth) * samples)
    x_new = np.linspace(-0.5, 0.5, occupied)
    y_new = interp_linear(x_new, ref_time, reference)
    pad = np.ones(int((samples - occupied) * 0.5))
    res = np.concatenate([pad, y_new, pad])
    if res.size < samples:
        res = np.append(res, np.ones(1))
    return 1 - ((1 - res) * depth)


def get_cache(durations, maxwidth_in_samples, per, rp, a, inc, ecc, w, u, limb_dark,
              verbose=False):
    """Template rows for every duration (src/gputls/transit.py:98-160).

    Returns (overview, rows): overview is a structured array with fields duration,
    width_in_samples and overshoot; rows[k] is the trimmed template flux of duration k.
    """
    ref = reference_transit(maxwidth_in_samples, per, rp, a, inc, ecc, w, u, limb_dark)
    return cache_rows(durations, maxwidth_in_samples, ref)


def cache_rows(durations, maxwidth_in_samples, ref):
    """get_cache's steps after the reference template `ref` (unit depth, in-transit part
    only, maxwidth_in_samples points; src/gputls/transit.py:129-160): per duration the
    fractional-width template at SIGNAL_DEPTH, its width in samples, the row trimmed to
    the samples below 1 - NUMERICAL_STABILITY_CUTOFF, and overshoot = 1 / (2 - mean / min)
    of that row."""
    n = np.size(durations)
    overview = np.zeros(n, dtype=[("duration", "f8"), ("width_in_samples", "i8"),
                                  ("overshoot", "f8")])
    dmax = np.max(durations)
    rows = []
    for k, dur in enumerate(durations):
        scaled = fractional_transit(dur, dmax, C.SIGNAL_DEPTH, maxwidth_in_samples, ref)
        overview["duration"][k] = dur
        overview["width_in_samples"][k] = int((dur / dmax) * maxwidth_in_samples)
        inside = np.where(scaled < (1 - C.NUMERICAL_STABILITY_CUTOFF))
        sig = scaled[np.min(inside): np.max(inside) + 1]
        rows.append(sig)
        overview["overshoot"][k] = 1 / (2 - np.mean(sig) / np.min(sig))
    arr = np.empty(n, dtype=object)
    arr[:] = rows
    return overview, arr


class TemplateTable:
    """Search-ready template table: unique ascending widths, one row per width.

    Row k holds S_k[i] = 1 - signal_k[i] for i < widths[k] (signal rows shorter than the
    width are padded with signal 0, i.e. S = 1, as the reference's padding does,
    src/gputls/core.py:753-754). Unique widths keep the first duration that produced
    them (np.unique(..., return_index=True), src/gputls/core.py:605-607).
    """

    def __init__(self, overview, rows):
        widths, first = np.unique(overview["width_in_samples"], return_index=True)
        self.widths = widths.astype(np.int32)
        self.durations = overview["duration"][first].astype(np.float64)
        self.overshoot = overview["overshoot"][first].astype(np.float64)
        self.rows = [np.asarray(rows[i], dtype=np.float64) for i in first]
        wmax = int(widths.max())
        self.wmax_even = wmax + (wmax % 2)
        offs = np.zeros(len(widths) + 1, dtype=np.int64)
        offs[1:] = np.cumsum(self.widths.astype(np.int64))
        self.offsets = offs
        S = np.ones(int(offs[-1]), dtype=np.float64)
        for k, row in enumerate(self.rows):
            d = int(self.widths[k])
            m = min(d, len(row))
            S[offs[k]: offs[k] + m] = 1.0 - row[:m]
        self.S = S

    @property
    def n_widths(self):
        return len(self.widths)


def build_templates(n_points, durations, per=C.DEFAULT_PERIOD, rp=C.DEFAULT_RP,
                    a=C.DEFAULT_A, inc=C.DEFAULT_INC, ecc=C.DEFAULT_ECC, w=C.DEFAULT_W,
                    u=C.DEFAULT_U, limb_dark=C.DEFAULT_LIMB_DARK):
    """Templates for a light curve of `n_points` samples (src/gputls/main.py:44-60)."""
    maxwidth = int(np.max(durations) * n_points)
    if maxwidth % 2 != 0:
        maxwidth += 1
    overview, rows = get_cache(durations, maxwidth, per, rp, a, inc, ecc, w, u, limb_dark)
    return TemplateTable(overview, rows), overview, rows
