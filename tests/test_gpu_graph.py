"""The batched search captured into a HIP graph: every replay, with eager device work and
synchronisations between replays, reproduces the eager search bit for bit."""
import numpy as np
import pytest
import torch

from gtls_amd import batch as B
from gtls_amd import search as S
from gtls_amd import synthetic

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager_search(gpu):
    curves = synthetic.batch(12, seed=11, n_min=2000, n_max=6000)
    db = B.DeviceBatch(B.BatchPlan(B.build_plans(curves, {})), 0)
    for mode in ("fp32c", "fp64"):
        want = [x.clone() for x in db.search(mode, 8)]
        torch.cuda.synchronize()
        g = db.capture(mode, 8)
        other = B.DeviceBatch(B.BatchPlan(B.build_plans(curves[:3], {})), 0)
        for r in range(3):
            db.chi2.fill_(float("nan"))
            db.best_k.fill_(-7)
            db.best_s.fill_(-7)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            got = (db.chi2, db.best_k, db.best_s)
            assert np.array_equal(got[0].cpu().numpy(), want[0].cpu().numpy(), equal_nan=True)
            assert torch.equal(got[1], want[1]) and torch.equal(got[2], want[2]), (mode, r)
            other.search(mode, 8)            # eager work between replays
            torch.cuda.synchronize()


def test_two_graphs_keep_their_own_workspaces(gpu):
    """Graphs captured from two batches, with an eager search that needs a larger shared
    workspace run between the replays: each graph writes only into the workspace it owns,
    so both keep reproducing their eager results."""
    small = B.DeviceBatch(B.BatchPlan(B.build_plans(
        synthetic.batch(4, seed=21, n_min=2000, n_max=3000), {})), 0)
    large = B.DeviceBatch(B.BatchPlan(B.build_plans(
        synthetic.batch(6, seed=22, n_min=5000, n_max=9000), {})), 0)
    bigger = B.DeviceBatch(B.BatchPlan(B.build_plans(
        synthetic.batch(3, seed=23, n_min=15000, n_max=20000), {})), 0)
    want = {}
    for name, db in (("small", small), ("large", large)):
        want[name] = [x.clone() for x in db.search("fp32c", 8)]
    torch.cuda.synchronize()
    graphs = {"small": small.capture("fp32c", 8), "large": large.capture("fp32c", 8)}
    assert graphs["small"].workspace.data_ptr() != graphs["large"].workspace.data_ptr()
    for r in range(2):
        for name, db in (("small", small), ("large", large)):
            db.chi2.fill_(float("nan"))
            torch.cuda.synchronize()
            graphs[name].replay()
            torch.cuda.synchronize()
            assert np.array_equal(db.chi2.cpu().numpy(), want[name][0].cpu().numpy(),
                                  equal_nan=True), (name, r)
            assert torch.equal(db.best_k, want[name][1]) and torch.equal(db.best_s, want[name][2])
            bigger.search("fp32c", 8)    # grows (reallocates) the shared workspace
            torch.cuda.synchronize()


def test_graph_kernel_timing_reads_each_replay(gpu):
    """A graph captured with timing records the library's per-chunk events on every replay:
    the figures are positive, their chunk count is the eager call's, and their sum stays
    within the replay's own wall time."""
    curves = synthetic.batch(10, seed=31, n_min=3000, n_max=9000)
    db = B.DeviceBatch(B.BatchPlan(B.build_plans(curves, {})), 0)
    S.set_kernel_timing(0, True)
    db.search("fp32c", 8)
    torch.cuda.synchronize()
    (_, _, _), n_eager = S.kernel_ms(0)
    S.set_kernel_timing(0, False)
    g = db.capture("fp32c", 8, timing=True)
    for r in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        (pm, wm, rm), n = B.graph_kernel_ms(g)
        assert n == n_eager > 0, (n, n_eager)
        assert pm > 0 and wm > 0 and rm >= 0, (pm, wm, rm)
        assert pm + wm + rm <= e0.elapsed_time(e1) * 1.05 + 0.05, r


def test_capture_keeps_the_tuning_of_the_calling_stream(gpu):
    """A graph captured after tune.set_tuning runs with that setting (its own library
    context inherits it): the replay equals the eager search under the same setting bit
    for bit, and that setting's results differ from the default one's (another tile form
    for strides 8-23), so a capture on the defaults would show."""
    from gtls_amd import tune
    curves = synthetic.batch(12, seed=11, n_min=2000, n_max=6000)
    db = B.DeviceBatch(B.BatchPlan(B.build_plans(curves, {})), 0)
    try:
        tune.set_tuning(0, *tune.DEFAULT)
        default = db.search("fp32c", 8)[0].clone()
        tune.set_tuning(0, 24, 0, 0, 1)
        want = [x.clone() for x in db.search("fp32c", 8)]
        torch.cuda.synchronize()
        assert not np.array_equal(default.cpu().numpy(), want[0].cpu().numpy(), equal_nan=True)
        g = db.capture("fp32c", 8)
        db.chi2.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(db.chi2.cpu().numpy(), want[0].cpu().numpy(), equal_nan=True)
        assert torch.equal(db.best_k, want[1]) and torch.equal(db.best_s, want[2])
    finally:
        tune.set_tuning(0, *tune.DEFAULT)
