"""Batched search (one call over many curves) and device spectra, through
the C ABI, against the FP64 oracle per curve and the host spectra."""

import numpy as np
import pytest

from gtls_amd import batch as B
from gtls_amd import search as S
from gtls_amd import spectra as SP
from gtls_amd import synthetic

pytestmark = pytest.mark.gpu


def small_batch(n_curves=12, seed=9):
    return synthetic.batch(n_curves, seed=seed, n_min=300, n_max=6000)


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_batched_matches_oracle_per_curve(gpu, mode):
    curves = small_batch()
    plans = B.build_plans(curves, {}, workers=1)
    bp = B.BatchPlan(plans)
    db = B.DeviceBatch(bp)
    db.search(mode, 8)
    chi2 = bp.unpack(db.chi2.cpu().numpy())
    bk = bp.unpack(db.best_k.cpu().numpy())
    bs = bp.unpack(db.best_s.cpu().numpy())
    # the shared bar (tests/parity_util.py): index-exact in FP64, near ties only in FP32c
    from parity_util import check, report
    worst, n_same, n_ok = 0.0, 0, 0
    for i, p in enumerate(plans):
        w, ns, no = check(p, mode, 8, p.periods, p.k_lo, p.k_hi, chi2[i], bk[i], bs[i],
                          "curve %d (n=%d)" % (i, p.n))
        worst, n_same, n_ok = max(worst, w), n_same + ns, n_ok + no
    report("batched-small", mode, worst, n_same, n_ok, "curves=%d" % len(plans))


def test_batched_equals_single_curve_launches(gpu):
    """The same curve gives the same bits in the batch and alone."""
    curves = small_batch(6, seed=3)
    plans = B.build_plans(curves, {}, workers=1)
    bp = B.BatchPlan(plans)
    db = B.DeviceBatch(bp)
    db.search("fp32c", 8)
    chi2 = bp.unpack(db.chi2.cpu().numpy())
    for i, p in enumerate(plans):
        c, k, s = S.DeviceCurve(p).search(p.periods, p.k_lo, p.k_hi, mode="fp32c")
        assert np.array_equal(c.cpu().numpy(), chi2[i])


def test_device_spectra_match_host(gpu):
    curves = small_batch(10, seed=5)
    plans = B.build_plans(curves, {}, workers=1)
    bp = B.BatchPlan(plans)
    db = B.DeviceBatch(bp)
    chi2, _, _ = db.search("fp64", 8)
    res = db.spectra(chi2, 3)
    c_all = bp.unpack(chi2.cpu().numpy())
    assert res["device"].sum() >= bp.n_curves // 2
    for i in range(bp.n_curves):
        _, p_raw, p_det, s_raw, s_det = SP.spectra(SP.mask_outliers(c_all[i]), 3)
        if not res["device"][i]:   # host path: identical by construction
            continue
        assert np.allclose(res["power_raw"][i], np.asarray(p_raw), rtol=1e-12, atol=1e-12)
        assert np.allclose(res["power"][i], np.asarray(p_det), rtol=1e-10, atol=1e-10)
        assert abs(res["SDE_raw"][i] - s_raw) <= 1e-12 * abs(s_raw)
        assert abs(res["SDE"][i] - s_det) <= 1e-10 * abs(s_det)
        assert int(np.argmax(res["power"][i])) == int(np.argmax(np.asarray(p_det)))
    print("device spectra on %d of %d curves" % (res["device"].sum(), bp.n_curves))


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_batched_without_skip_matches_oracle(gpu, mode):
    """The batched call with no T0 skip (T0_fit_margin 0, the form of the candidates'
    re-search) on every 10th period of six curves, against the oracle at the shared bar
    (tests/parity_util.py: index-exact in FP64, near ties only in FP32c)."""
    from gtls_amd.grid import skip_factor
    from gtls_amd.validate import validate_args, validate_inputs
    from parity_util import check, report
    none = skip_factor(0.0)
    plans = []
    for (t, y, dy, _) in synthetic.batch(6, seed=21, n_min=300, n_max=6000):
        t, y, dy = validate_inputs(t, y, dy)
        full = S.make_plan(t, y, dy, validate_args({}))
        plans.append(S.make_plan(t, y, dy, validate_args({}), periods=full.periods[::10]))
    bp = B.BatchPlan(plans)
    db = B.DeviceBatch(bp)
    db.search(mode, none)
    chi2 = bp.unpack(db.chi2.cpu().numpy())
    bk = bp.unpack(db.best_k.cpu().numpy())
    bs = bp.unpack(db.best_s.cpu().numpy())
    worst, n_same, n_ok = 0.0, 0, 0
    for i, p in enumerate(plans):
        w, ns, no = check(p, mode, none, p.periods, p.k_lo, p.k_hi, chi2[i], bk[i], bs[i],
                          "curve %d (n=%d)" % (i, p.n))
        worst, n_same, n_ok = max(worst, w), n_same + ns, n_ok + no
    report("batched-no-skip", mode, worst, n_same, n_ok, "curves=%d" % len(plans))


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_batch_mixing_common_and_per_sample_dy_matches_oracle(gpu, mode):
    """One call over curves of both weight paths (every other curve with per-sample dy),
    each on every 4th period, against the oracle at the shared bar."""
    from gtls_amd.validate import validate_args, validate_inputs
    from parity_util import check, report
    rng = np.random.default_rng(77)
    plans = []
    for j, (t, y, dy, _) in enumerate(small_batch(8, seed=13)):
        if j % 2:
            dy = np.std(y) * rng.uniform(0.5, 1.5, len(t))
        t, y, dy = validate_inputs(t, y, dy)
        full = S.make_plan(t, y, dy, validate_args({}))
        plans.append(S.make_plan(t, y, dy, validate_args({}), periods=full.periods[::4]))
    assert len({bool(np.ptp(p.dy) > 0) for p in plans}) == 2
    bp = B.BatchPlan(plans)
    db = B.DeviceBatch(bp)
    db.search(mode, 8)
    chi2 = bp.unpack(db.chi2.cpu().numpy())
    bk = bp.unpack(db.best_k.cpu().numpy())
    bs = bp.unpack(db.best_s.cpu().numpy())
    worst, n_same, n_ok = 0.0, 0, 0
    for i, p in enumerate(plans):
        w, ns, no = check(p, mode, 8, p.periods, p.k_lo, p.k_hi, chi2[i], bk[i], bs[i],
                          "curve %d (n=%d)" % (i, p.n))
        worst, n_same, n_ok = max(worst, w), n_same + ns, n_ok + no
    report("batched-mixed-dy", mode, worst, n_same, n_ok, "curves=%d" % len(plans))
