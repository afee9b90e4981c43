"""End to end on the GPU: gtls(t, y, dy).power() -- the full-grid search, spectra, the
candidate and harmonic re-search without stride and the refinement of the final period --
on an injected transit (the 2-min curve with P = 3.71 d) and on the TESS light curve the
reference's scripts load; the final period's chi2 is checked against the oracle's search
of that period without stride, and the refinement of grid peaks runs on the result."""
import os

import numpy as np
import pytest

from gtls_amd import fits, synthetic
from gtls_amd.api import gtls
from gtls_amd.grid import skip_factor
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = ("periods", "period", "rawDuration", "durationPoints", "rawDurations", "duration",
          "depth", "T0", "SDE", "chi2", "transitTimes", "power", "snr", "snrPink", "snrFit",
          "snrFitPink", "raw_power", "raw_chi2", "possiblePeriodsIndices", "possiblePeriods")
FILE = os.path.join(os.path.dirname(__file__), "golden",
                    "tess2022302161335-s0058-0000000021132157-0247-s_lc.fits")


def check_final_chi2(model, res):
    plan = model.plan
    g = int(np.argmin(np.abs(plan.periods - res.period)))
    c, _, _ = oracle.search_plan(plan, skip_point=skip_factor(0.0), periods=plan.periods[g:g + 1],
                                 k_lo=plan.k_lo[g:g + 1], k_hi=plan.k_hi[g:g + 1])
    assert float(np.asarray(res.chi2)[g]) == pytest.approx(float(c[0]), rel=2e-6)


def test_power_recovers_injected_transit(gpu):
    t, y, dy, truth = synthetic.single_2min()
    model = gtls(t, y, dy, verbose=False)
    res = model.power()
    for f in FIELDS:
        assert hasattr(res, f), f
    assert res.period == pytest.approx(truth["period"], rel=2e-3)
    assert res.SDE > 9
    assert res.depth == pytest.approx(truth["depth"], rel=0.35)
    assert len(res.power) == len(res.periods) == len(res.chi2)
    # transit times are spaced by the period and T0 falls on the injected ephemeris
    tt = np.asarray(res.transitTimes)
    assert np.allclose(np.diff(tt), res.period, rtol=1e-9)
    dt = (res.T0 - truth["t0"] + truth["period"] / 2) % truth["period"] - truth["period"] / 2
    assert abs(dt) < 0.05
    check_final_chi2(model, res)
    # fast mode: the first-pass spectrum
    periods, power = gtls(t, y, dy, verbose=False).power(fast=True)
    assert len(periods) == len(power) == len(res.periods)
    assert np.allclose(np.asarray(power), np.asarray(res.raw_power))


def test_power_on_tess_light_curve(gpu):
    t, y, dy, hdr = fits.read_lightcurve(FILE)
    model = gtls(t, y, dy, verbose=False)
    res = model.power(R_star=float(hdr["RADIUS"]))
    assert np.isfinite(res.SDE) and res.SDE > 0
    assert np.asarray(res.periods).min() <= res.period <= np.asarray(res.periods).max()
    assert len(res.possiblePeriods) > 0
    check_final_chi2(model, res)
