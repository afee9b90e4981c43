"""`gtls(t, y, dy).power(**kwargs)` -- the search API of the reference
(src/gputls/main.py:11-206, results in src/gputls/results.py:1-24), running the chi2 search
on the MI355X through libgtls_amd.so.

Flow of power() (src/gputls/core.py:568-999):
  1. full-grid search with the T0 position stride of T0_fit_margin  -> raw chi2
  2. outlier mask and spectra (SR, power, SDE)                       -> raw_power
     (fast=True stops here and returns (periods, power))
  3. re-search without stride of the top-100 periods and the next 100 above 1 d,
     spectra again, best period
  4. re-search of 1/2, 1, 2, 2/3, 3/2 times that period, spectra, final period and SDE
  5. depth, T0, duration, transit times and SNRs of the final period
"""
from __future__ import annotations

import warnings

import numpy as np
import numpy.ma as ma

from . import constants as C
from .grid import period_grid, skip_factor
from .search import DeviceCurve, make_plan
from .spectra import mask_outliers, spectra
from .stats import refine_period
from .validate import validate_args, validate_inputs

HARMONICS = (0.5, 1, 2, 2 / 3, 3 / 2)


class gtlsResult:
    """Result of a search (field names of src/gputls/results.py:4-24)."""

    def __init__(self, **fields):
        for k, v in fields.items():
            setattr(self, k, v)

    def __repr__(self):
        return "gtlsResult(period=%r, SDE=%r, depth=%r, T0=%r, duration=%r)" % (
            getattr(self, "period", None), getattr(self, "SDE", None),
            getattr(self, "depth", None), getattr(self, "T0", None),
            getattr(self, "duration", None))


def candidate_indices(periods, power, n_top=100, n_next=100):
    """Top-n_top periods by power, then the next n_next with period > 1 d
    (src/gputls/core.py:904-915); ties keep grid order."""
    order = np.argsort(-np.asarray(power), kind="stable")
    top = list(order[:n_top])
    taken = set(top)
    nxt = [i for i in order if i not in taken and periods[i] > 1][:n_next]
    return top + nxt


def nearest_indices(targets, periods):
    periods = np.asarray(periods)
    return [int(np.argmin(np.abs(periods - v))) for v in targets]


def run_power(plan, args, search):
    """The steps of power() after the plan (src/gputls/core.py:568-999), with `search`
    doing the chi2 search: search(period_indices, skip_point) -> (chi2, best width index,
    best T0 index) per period, numpy. power() passes the device search; any search with
    the same contract (the FP64 oracle, in the tests) gives the same pipeline. Returns
    (periods, power) with args.fast, else the gtlsResult."""
    periods = plan.periods
    skip = skip_factor(args.T0_fit_margin)
    raw_chi2, _, _ = search(np.arange(len(periods)), skip)
    chi2 = mask_outliers(raw_chi2)
    periods_m = ma.array(periods, mask=ma.getmaskarray(chi2))
    _, _, power, _, _ = spectra(chi2, args.oversampling_factor)
    raw_power = power.copy()
    if args.fast:
        return periods_m, power

    cand = candidate_indices(periods, power)
    again, _, _ = search(np.asarray(cand), C.SKIP_NONE)
    chi2[cand] = again
    _, _, power, _, _ = spectra(chi2, args.oversampling_factor)
    best = cand[int(np.argmax(np.asarray(power)[cand]))]

    harm = nearest_indices([periods[best] * r for r in HARMONICS], periods)
    again, hk, hs = search(np.asarray(harm), C.SKIP_NONE)
    chi2[harm] = again
    _, _, power, SDE_raw, SDE = spectra(chi2, args.oversampling_factor)
    j = int(np.argmax(np.asarray(power)[harm]))
    best = harm[j]
    period = float(periods[best])
    ref = refine_period(plan, period, int(hk[j]), int(hs[j]))
    return gtlsResult(
        periods=periods_m, period=period, rawDuration=ref["rawDuration"],
        durationPoints=ref["durationPoints"], rawDurations=list(plan.overview["duration"]),
        duration=ref["duration"], depth=ref["depth"], T0=ref["T0"], SDE=SDE,
        SDE_raw=SDE_raw, chi2=chi2, transitTimes=ref["transitTimes"], power=power,
        snr=ref["snr"], snrPink=ref["snrPink"], snrFit=ref["snrFit"],
        snrFitPink=ref["snrFitPink"], raw_power=raw_power, raw_chi2=raw_chi2,
        possiblePeriodsIndices=cand, possiblePeriods=[float(periods[i]) for i in cand],
        bestWidthIndex=int(hk[j]), bestT0Index=int(hs[j]))


class gtls:
    """Transit least squares of limb-darkened transit models on one light curve."""

    def __init__(self, t, y, dy=None, verbose=True):
        self.t, self.y, self.dy = validate_inputs(t, y, dy)
        self.verbose = verbose

    def _device_id(self, args):
        if args.GPUDeviceIDs is not None and len(args.GPUDeviceIDs) > 1:
            warnings.warn("one GPU per search: using GPUDeviceIDs[0]")
            return int(args.GPUDeviceIDs[0])
        return int(args.GPUDeviceID)

    def power(self, **kwargs):
        args = validate_args(kwargs)
        if self.verbose:
            print(C.TLS_VERSION)
        plan = make_plan(self.t, self.y, self.dy, args, periods=args.periods)
        self.plan = plan
        periods = plan.periods
        if self.verbose:
            print("Searching %d data points, %d periods from %.3f to %.3f days (%s)"
                  % (plan.n, len(periods), periods.min(), periods.max(), args.mode))
        dc = DeviceCurve(plan, self._device_id(args))
        depth_min = args.transit_depth_min

        def search(idx, skip_point):
            c, k, s = dc.search(periods[idx], plan.k_lo[idx], plan.k_hi[idx], mode=args.mode,
                                skip_point=skip_point, depth_min=depth_min)
            return c.cpu().numpy(), k.cpu().numpy(), s.cpu().numpy()

        out = run_power(plan, args, search)
        if args.fast:
            return out
        self.period = out.period
        self.rawDurations = out.rawDurations
        return out

    def periodogram(self, **kwargs):
        """The period grid power() would search (src/gputls/main.py:186-206)."""
        args = validate_args(kwargs)
        if len(args.periods):
            return args.periods
        return period_grid(R_star=args.R_star, M_star=args.M_star,
                           time_span=np.max(self.t) - np.min(self.t),
                           period_min=args.period_min, period_max=args.period_max,
                           oversampling_factor=args.oversampling_factor,
                           n_transits_min=args.n_transits_min)
