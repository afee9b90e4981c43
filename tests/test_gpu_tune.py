"""Autotuning of the float32 search: every value of every tuning axis is timed, the fastest
setting seen is applied, and a search under any setting agrees with the default one to the
float32 bar."""
import numpy as np
import pytest
import torch

from gtls_amd import batch as B
from gtls_amd import synthetic
from gtls_amd import tune

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_candidate_and_keeps_results(gpu):
    curves = synthetic.batch(24, seed=5, n_min=2000, n_max=12000)
    db = B.DeviceBatch(B.BatchPlan(B.build_plans(curves, {})), 0)
    c0 = db.search("fp32c", 8)[0].clone()
    torch.cuda.synchronize()
    try:
        choice, times = tune.autotune(db, n_sample=12, repeats=1)
        # every value of every axis was timed, at most one setting per value and axis
        for axis, values in enumerate((tune.TOEP_Q, tune.CHUNK_MB, tune.WG_PER_CU, tune.QUAD_Q)):
            assert {s[axis] for s in times} == set(values)
        assert len(times) <= (len(tune.TOEP_Q) + len(tune.CHUNK_MB) + len(tune.WG_PER_CU)
                              + len(tune.QUAD_Q))
        assert all(np.isfinite(v) and v > 0 for v in times.values())
        assert choice in times and times[choice] == min(times.values())
        for q, wg in ((q, wg) for q in tune.TOEP_Q for wg in tune.WG_PER_CU):
            tune.set_tuning(0, q, 0, wg)
            c = db.search("fp32c", 8)[0]
            torch.cuda.synchronize()
            ok = torch.isfinite(c0)
            rel = ((c - c0).abs() / c0.abs())[ok]
            assert float(rel.max()) < 2e-6, (q, wg)
        # the operand load layout changes nothing in the results
        tune.set_tuning(0, 8, 0, 0, 96)
        c1 = db.search("fp32c", 8)[0].clone()
        for quad in (1, 64, 1 << 20):
            tune.set_tuning(0, 8, 0, 0, quad)
            c = db.search("fp32c", 8)[0]
            torch.cuda.synchronize()
            assert torch.equal(torch.nan_to_num(c, nan=-1.0), torch.nan_to_num(c1, nan=-1.0)), quad
    finally:
        tune.set_tuning(0, *tune.DEFAULT)


def test_out_of_range_tuning_is_refused(gpu):
    """gtls_set_tuning refuses a Toeplitz bound outside [0, 24], a negative or unshiftable
    chunk budget, and negative workgroup counts or layout thresholds; the setting applied
    before stays in force."""
    from gtls_amd import _ffi
    tune.set_tuning(0, *tune.DEFAULT)
    try:
        for bad in ((25, 0, 0, 0), (-1, 0, 0, 0), (8, -1, 0, 0), (8, 1 << 43, 0, 0),
                    (8, 0, -1, 0), (8, 0, 0, -1)):
            with pytest.raises(_ffi.GtlsError):
                tune.set_tuning(0, *bad)
            assert tune.applied(0) == tune.DEFAULT, bad
    finally:
        tune.set_tuning(0, *tune.DEFAULT)
