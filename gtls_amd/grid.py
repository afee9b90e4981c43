"""Period grid, duration grid and the per-period window of template widths.

`period_grid` is Ofir (2014, A&A 561, A138) optimal frequency sampling, `duration_grid` a
log-spaced set of fractional durations, `T14` the maximum transit duration. These are host
inputs to the search; the float operations are written in the same order as the reference
(src/gputls/grid.py:7-157) so that the generated grids are bit-identical to it.

`width_window` restates the per-period bounds kernel of the reference
(src/gputls/GPUFun.cu:45-84, called at src/gputls/core.py:652-655) and `width_ranges`
turns those bounds into an inclusive index range into the sorted width table, which is
the per-period width set that both the device search and the oracle consume.
"""
from __future__ import annotations

import warnings

import numpy as np

from . import constants as C


def T14(R_s, M_s, P, upper_limit=C.FRACTIONAL_TRANSIT_DURATION_MAX, small=False):
    """Maximum transit duration as a fraction of the period (src/gputls/grid.py:7-29)."""
    P = P * C.SECONDS_PER_DAY
    R_s = C.R_SUN * R_s
    M_s = C.M_SUN * M_s
    scale = ((4 * P) / (np.pi * C.G * M_s)) ** (1 / 3)
    radius = R_s if small else R_s + 2 * C.R_JUP
    frac = radius * scale / P
    return upper_limit if frac > upper_limit else frac


def duration_grid(periods, shortest=None, log_step=C.DURATION_GRID_STEP):
    """Log-spaced fractional durations (src/gputls/grid.py:32-53).

    `shortest` is accepted for signature compatibility; the reference does not use it.
    """
    longest = T14(R_s=C.R_STAR_MAX, M_s=C.M_STAR_MAX, P=min(periods), small=False)
    d = T14(R_s=C.R_STAR_MIN, M_s=C.M_STAR_MIN, P=max(periods), small=True)
    out = [d]
    while d * log_step < longest:
        d = d * log_step
        out.append(d)
    out.append(longest)
    return out


def period_grid(R_star, M_star, time_span, period_min=0, period_max=float("inf"),
                oversampling_factor=C.OVERSAMPLING_FACTOR, n_transits_min=C.N_TRANSITS_MIN):
    """Ofir (2014) optimal period grid in days (src/gputls/grid.py:56-157)."""
    # unphysical stellar parameters are clamped exactly as the reference does
    # (including its R_star < 0.01 -> 0.1 replacement, src/gputls/grid.py:71-78)
    if R_star < 0.01:
        warnings.warn("R_star was unphysical (%r); using 0.1 for period_grid" % (R_star,))
        R_star = 0.1
    if R_star > 10000:
        warnings.warn("R_star was unphysical (%r); using 10000 for period_grid" % (R_star,))
        R_star = 10000
    if M_star < 0.01:
        warnings.warn("M_star was unphysical (%r); using 0.01 for period_grid" % (M_star,))
        M_star = 0.01
    if M_star > 1000:
        warnings.warn("M_star was unphysical (%r); using 1000 for period_grid" % (M_star,))
        M_star = 1000

    R = R_star * C.R_SUN
    M = M_star * C.M_SUN
    span = time_span * C.SECONDS_PER_DAY
    pi = np.pi

    f_min = n_transits_min / span
    f_max = 1.0 / (2 * pi) * np.sqrt(C.G * M / (3 * R) ** 3)
    A = ((2 * pi) ** (2.0 / 3) / pi * R / (C.G * M) ** (1.0 / 3)
         / (span * oversampling_factor))
    Cc = f_min ** (1.0 / 3) - A / 3.0
    n_opt = (f_max ** (1.0 / 3) - f_min ** (1.0 / 3) + A / 3) * 3 / A
    X = np.arange(n_opt) + 1
    f_x = (A / 3 * X + Cc) ** 3
    periods = (1 / f_x) / C.SECONDS_PER_DAY

    keep = np.where(np.logical_and(periods > period_min, periods <= period_max))
    n_keep = np.size(periods[keep])
    if n_keep > 10 ** 6:
        warnings.warn("period_grid generates a very large grid (%d)" % n_keep)
    if n_keep < C.MINIMUM_PERIOD_GRID_SIZE:
        if span < 5 * C.SECONDS_PER_DAY:
            span = 5 * C.SECONDS_PER_DAY
        warnings.warn("period_grid defaults to R_star=1 and M_star=1: too few periods")
        return period_grid(R_star=1, M_star=1, time_span=span / C.SECONDS_PER_DAY)
    return periods[keep]


def width_window(periods, time_span, n_points):
    """Per-period [lo, hi] template width bounds in samples (src/gputls/GPUFun.cu:45-84).

    The time span enters as a float32 value (the dtype the reference's worker passes,
    src/gputls/_worker.py:106) and the arithmetic is float64.
    """
    P = np.asarray(periods, dtype=np.float64)
    L = np.float64(np.float32(time_span))
    n_transits = L / P
    correction = (n_transits + 1.0) / n_transits
    P_sec = P * C.SECONDS_PER_DAY
    rs_min = C.R_SUN * C.WIN_R_STAR_MIN
    rs_max = C.R_SUN * C.WIN_R_STAR_MAX
    fac_min = (4.0 * P_sec) / (C.WIN_PI_GM_MIN * C.WIN_SCALE)
    fac_max = (4.0 * P_sec) / (C.WIN_PI_GM_MAX * C.WIN_SCALE)
    t14_min = rs_min * np.power(fac_min, 1.0 / 3.0)
    t14_max = (rs_max + C.R_JUP * 2.0) * np.power(fac_max, 1.0 / 3.0)
    d_min = np.minimum(t14_min / P_sec, C.WIN_FRACTIONAL_MAX)
    d_max = np.minimum(t14_max / P_sec, C.WIN_FRACTIONAL_MAX)
    lo = np.floor(d_min * n_points).astype(np.int64)
    hi = np.ceil(d_max * n_points * correction).astype(np.int64)
    return lo, hi


def width_ranges(widths, lo, hi):
    """Inclusive index range [k_lo, k_hi] of `widths` (ascending) inside [lo, hi].

    An empty range is returned as k_lo > k_hi. This is the per-period width set of
    src/gputls/GPUFun.cu:89-103, evaluated for each period on its own rather than as
    the union over a memory-sized batch (src/gputls/core.py:693-719).
    """
    w = np.asarray(widths, dtype=np.int64)
    k_lo = np.searchsorted(w, lo, side="left").astype(np.int32)
    k_hi = (np.searchsorted(w, hi, side="right") - 1).astype(np.int32)
    return k_lo, k_hi


def skip_factor(T0_fit_margin):
    """Window position stride divisor SKIP (src/gputls/core.py:597-601)."""
    if T0_fit_margin == 0:
        return C.SKIP_NONE
    return int(1 / T0_fit_margin)
