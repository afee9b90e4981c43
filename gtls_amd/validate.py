"""Input cleaning and keyword validation of the search API.

Behaviour follows src/gputls/validate.py:9-203 and src/gputls/helpers.py:16-59: samples
with a NaN, non-positive or infinite time, flux or error are dropped; a given dy is
normalised to mean 1, a missing dy becomes the constant std(y); keyword arguments get the
reference's defaults and the same range checks and error messages' meaning.
"""
from __future__ import annotations

import multiprocessing
import warnings

import numpy as np

from . import constants as C


def cleaned_array(t, y, dy=None):
    """Drop samples where any of t, y (, dy) is NaN, None, <= 0 or infinite, over the
    first len(y) samples (the reference walks range(len(y)), src/gputls/helpers.py:16-59;
    a shorter t or dy is an error)."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    cols = [t, y] if dy is None else [t, y, np.asarray(dy, dtype=np.float64)]
    if any(c.ndim != 1 for c in cols):
        raise ValueError("Inputs (t, y, dy) must be 1-dimensional")
    n = len(y)
    if any(len(c) < n for c in cols):
        raise ValueError("Arrays (t, y, dy) must be of the same dimensions")
    keep = np.ones(n, dtype=bool)
    for c in cols:
        c = c[:n]
        keep &= np.isfinite(c) & (c > 0)
    out = tuple(np.array(c[:n][keep], dtype=float) for c in cols)
    return out


def impact_to_inclination(b, semimajor_axis):
    return np.degrees(np.arccos(b / semimajor_axis))


def validate_inputs(t, y, dy):
    if dy is None:
        t, y = cleaned_array(t, y)
    else:
        t, y, dy = cleaned_array(t, y, dy)
        dy = dy / np.mean(dy)
    if np.size(t) == 0 or max(t) - min(t) <= 0:
        raise ValueError("Time duration must positive")
    if np.size(y) < 3 or np.size(t) < 3:
        raise ValueError("Too few values in data set")
    if np.mean(y) > 1.01 or np.mean(y) < 0.99:
        warnings.warn("The mean flux should be normalized to 1, but it was found to be "
                      + str(np.mean(y)))
    if min(y) < 0:
        raise ValueError("Flux values must be positive")
    if max(y) >= float("inf"):
        raise ValueError("Flux values must be finite")
    if dy is None:
        dy = np.full(len(y), np.std(y))
    if np.size(t) != np.size(y) or np.size(t) != np.size(dy):
        raise ValueError("Arrays (t, y, dy) must be of the same dimensions")
    return t, y, dy


class SearchArgs:
    """Validated keyword arguments of power() with the reference's defaults."""


def validate_args(kwargs):
    a = SearchArgs()
    for key in kwargs:
        if key not in C.VALID_PARAMETERS:
            warnings.warn("Ignoring unknown parameter: " + str(key))
    g = kwargs.get
    a.GPUDeviceID = g("GPUDeviceID", 0)
    a.GPUDeviceIDs = g("GPUDeviceIDs", None)
    a.legacy = g("legacy", False)
    a.fast = g("fast", False)
    a.bar_location = g("bar_location", 0)
    a.mode = g("mode", "fp64")
    a.show_progress_bar = g("show_progress_bar", True)
    a.transit_depth_min = g("transit_depth_min", C.TRANSIT_DEPTH_MIN)
    a.R_star = g("R_star", C.R_STAR)
    a.M_star = g("M_star", C.M_STAR)
    a.oversampling_factor = g("oversampling_factor", C.OVERSAMPLING_FACTOR)
    a.period_max = g("period_max", float("inf"))
    a.period_min = g("period_min", 0)
    a.periods = g("periods", [])
    a.n_transits_min = g("n_transits_min", C.N_TRANSITS_MIN)
    a.R_star_min = g("R_star_min", C.R_STAR_MIN)
    a.R_star_max = g("R_star_max", C.R_STAR_MAX)
    a.M_star_min = g("M_star_min", C.M_STAR_MIN)
    a.M_star_max = g("M_star_max", C.M_STAR_MAX)
    a.duration_grid_step = g("duration_grid_step", C.DURATION_GRID_STEP)
    a.use_threads = g("use_threads", multiprocessing.cpu_count())
    a.per = g("per", C.DEFAULT_PERIOD)
    a.rp = g("rp", C.DEFAULT_RP)
    a.a = g("a", C.DEFAULT_A)
    a.T0_fit_margin = g("T0_fit_margin", C.T0_FIT_MARGIN)
    if "b" in kwargs:
        a.b = kwargs["b"]
        a.inc = impact_to_inclination(a.b, a.a)
    else:
        a.inc = g("inc", C.DEFAULT_INC)
    a.ecc = g("ecc", C.DEFAULT_ECC)
    a.w = g("w", C.DEFAULT_W)
    a.u = g("u", C.DEFAULT_U)
    a.limb_dark = g("limb_dark", C.DEFAULT_LIMB_DARK)
    a.transit_template = g("transit_template", "default")

    # template selection (src/gputls/validate.py:110-145)
    if a.transit_template == "default":
        a.per, a.rp, a.a, a.inc = C.DEFAULT_PERIOD, C.DEFAULT_RP, C.DEFAULT_A, C.DEFAULT_INC
    elif a.transit_template == "grazing":
        a.b = C.GRAZING_B
        a.inc = impact_to_inclination(a.b, a.a)
    elif a.transit_template == "box":
        a.per, a.rp, a.a, a.b, a.inc = (C.BOX_PERIOD, C.BOX_RP, C.BOX_A, C.BOX_B,
                                        C.BOX_INC)
        a.u, a.limb_dark = C.BOX_U, C.BOX_LIMB_DARK
    else:
        raise ValueError('Unknown transit_template. Known values: "default", "grazing", "box"')

    def positive(name, v):
        if v <= 0 or v >= float("inf"):
            raise ValueError(name + " must be positive")

    positive("R_star", a.R_star)
    if a.R_star_min > a.R_star:
        raise ValueError("R_star_min <= R_star is required")
    positive("R_star_min", a.R_star_min)
    if a.R_star_max < a.R_star:
        raise ValueError("R_star_max >= R_star is required")
    positive("R_star_max", a.R_star_max)
    positive("M_star", a.M_star)
    if a.M_star_min > a.M_star:
        raise ValueError("M_star_min <= M_star is required")
    positive("M_star_min", a.M_star_min)
    if a.M_star_max < a.M_star:
        raise ValueError("M_star_max >= M_star required")
    positive("M_star_max", a.M_star_max)
    if a.period_min < 0:
        raise ValueError("period_min >= 0 required")
    if a.period_min >= a.period_max:
        raise ValueError("period_min < period_max required")
    if not isinstance(a.n_transits_min, int) or a.n_transits_min < 1:
        raise ValueError("n_transits_min must be an integer value >= 1")
    if not isinstance(a.use_threads, int) or a.use_threads < 1:
        raise ValueError("use_threads must be an integer value >= 1")
    if a.T0_fit_margin < 0:
        a.T0_fit_margin = 0
    elif a.T0_fit_margin > 0.125:
        a.T0_fit_margin = 0.125
    if a.mode not in ("fp64", "fp32c"):
        raise ValueError('mode must be "fp64" or "fp32c"')
    return a
