"""Full-size checks (the 4-year 29.4-min curve: N ~ 65,000, ~1.9e5 trial periods) through
properties that do not need the CPU oracle at this size: the float32 matrix-core search and
the float64 search are two independent arithmetic paths over the same grid, so their
per-period chi2 must agree to the float32 bar, their best indices may differ only at near
ties, and their spectra must give the same peak and an SDE within 1e-4; every chi2 lies
below the sentinel N; and the strongest peak is the injected 45.3-d transit."""
import numpy as np
import pytest

from gtls_amd import search as S
from gtls_amd import spectra as SP
from gtls_amd import synthetic
from gtls_amd.grid import skip_factor
from gtls_amd.validate import validate_args, validate_inputs

pytestmark = pytest.mark.gpu


def test_full_grid_fp32c_matches_fp64_on_the_long_curve(gpu):
    t, y, dy, truth = synthetic.long_30min()
    t, y, dy = validate_inputs(t, y, dy)
    plan = S.make_plan(t, y, dy, validate_args({}))
    assert plan.n > 60000 and len(plan.periods) > 150000
    dc = S.DeviceCurve(plan)
    skip = skip_factor(0.125)
    out = {}
    for mode in ("fp32c", "fp64"):
        c, k, s = [x.cpu().numpy() for x in dc.search(plan.periods, plan.k_lo, plan.k_hi,
                                                        mode=mode, skip_point=skip)]
        out[mode] = (c, k, s)
    c32, k32, s32 = out["fp32c"]
    c64, k64, s64 = out["fp64"]
    ok = plan.k_lo <= plan.k_hi
    assert np.all(np.isfinite(c64[ok])) and np.all(np.isfinite(c32[ok]))
    assert np.all(c64[ok] <= plan.n * (1 + 1e-12))
    rel = np.abs(c32[ok] - c64[ok]) / np.abs(c64[ok])
    assert rel.max() < 2e-6, rel.max()
    differ = np.mean((k32[ok] != k64[ok]) | (s32[ok] != s64[ok]))
    assert differ < 0.01, differ
    sde = {}
    peak = {}
    for mode, (c, _, _) in out.items():
        _, _, power, _, sd = SP.spectra(SP.mask_outliers(c), 3)
        sde[mode] = sd
        peak[mode] = int(np.argmax(np.asarray(power)))
    assert peak["fp32c"] == peak["fp64"]
    assert abs(sde["fp32c"] - sde["fp64"]) / sde["fp64"] < 1e-4
    assert plan.periods[peak["fp64"]] == pytest.approx(truth["period"], rel=2e-3)
