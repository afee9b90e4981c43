"""Smooth transit model of the peak refinement: the torch restatement against closed forms
and finite differences (CPU), the HIP forward/backward kernels against the torch model,
its autograd and float64 central differences (GPU), and an Adam refinement that recovers
an injected transit (GPU).
"""
import numpy as np
import pytest
import torch

from gtls_amd import refine as R

U = (0.4804, 0.1867)


def theta_row(P=3.0, T0=1.2, D=0.15, depth=2e-3, u=U):
    q1, q2 = R.q_from_u(*u)
    return torch.tensor([[P, T0, D, depth, q1, q2]], dtype=torch.float64)


def test_q_u_round_trip():
    q1, q2 = R.q_from_u(0.45, 0.2)
    assert np.allclose(R.u_from_q(q1, q2), (0.45, 0.2), rtol=0, atol=1e-15)


def test_torch_model_closed_forms():
    th = theta_row()
    P, T0, D, depth = 3.0, 1.2, 0.15, 2e-3
    u1, u2 = U
    # centre of every transit: 1 - depth; half a period away: 1
    t = torch.tensor([[T0, T0 + 7 * P, T0 - 2 * P, T0 + P / 2]], dtype=torch.float64)
    f = R.transit_model_torch(th, t)[0]
    assert torch.allclose(f[:3], torch.full((3,), 1 - depth, dtype=torch.float64), rtol=0, atol=1e-15)
    assert f[3] == 1.0
    # inside the disk the profile is the quadratic law at mu = sqrt(1 - x^2)
    for x in (0.3, 0.6, 0.8):
        tt = torch.tensor([[T0 + x * D / 2]], dtype=torch.float64)
        mu = np.sqrt(1 - x * x)
        want = 1 - depth * (1 - u1 * (1 - mu) - u2 * (1 - mu) ** 2)
        assert float(R.transit_model_torch(th, tt)[0, 0]) == pytest.approx(want, abs=depth * 1e-6)
    # the smoothed edge: half the limb intensity at |x| = 1, monotone through the edge
    edge = torch.tensor([[T0 + D / 2]], dtype=torch.float64)
    assert float(R.transit_model_torch(th, edge)[0, 0]) == pytest.approx(
        1 - depth * 0.5 * (1 - u1 * (1 - np.sqrt(np.log(2) / R.KAPPA))
                           - u2 * (1 - np.sqrt(np.log(2) / R.KAPPA)) ** 2), rel=1e-12)
    xs = torch.linspace(0.9, 1.2, 200, dtype=torch.float64)
    fe = R.transit_model_torch(th, (T0 + xs * D / 2)[None, :])[0]
    assert torch.all(torch.diff(fe) >= 0)


def test_torch_model_gradient_matches_central_differences():
    """Pins the derivative chain the kernels restate: autograd of the torch model against
    float64 central differences (h = 1e-6 relative) of a chi2 loss."""
    rng = np.random.default_rng(3)
    t = torch.tensor(np.sort(rng.uniform(0, 30, 4000)), dtype=torch.float64)
    th0 = theta_row(P=2.7, T0=0.9, D=0.2, depth=3e-3)
    y = R.transit_model_torch(theta_row(P=2.7004, T0=0.93, D=0.18, depth=2.5e-3), t[None])[0]
    y = y + torch.tensor(rng.normal(0, 5e-4, t.numel()), dtype=torch.float64)

    def loss(th):
        return ((y - R.transit_model_torch(th, t[None])[0]) ** 2).sum() * 4e6

    th = th0.clone().requires_grad_(True)
    loss(th).backward()
    g = th.grad[0].numpy()
    for k in range(6):
        h = 1e-6 * abs(float(th0[0, k]))
        tp, tm = th0.clone(), th0.clone()
        tp[0, k] += h
        tm[0, k] -= h
        fd = (float(loss(tp)) - float(loss(tm))) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-4, abs=1e-6 * np.abs(g).max()), k


# ------------------------------------------------------------------------------ GPU


def model_batch(n_curves=3, n_per_curve=2, seed=0, n=3000, gpu=None):
    rng = np.random.default_rng(seed)
    curves, thetas, truths = [], [], []
    for c in range(n_curves):
        t = np.sort(rng.uniform(100.0, 100.0 + 40.0, n + 500 * c))
        P, T0, D, depth = 2.0 + 1.3 * c, 100.7 + 0.1 * c, 0.12 + 0.02 * c, 2e-3 * (1 + c)
        truth = theta_row(P, T0, D, depth)
        y = R.transit_model_torch(truth, torch.tensor(t)[None])[0].numpy()
        dy = np.full(t.size, 4e-4)
        y = y + rng.normal(0, 4e-4, t.size)
        curves.append((t, y, dy))
        truths.append(truth[0].numpy())
        for j in range(n_per_curve):
            th = truth[0].numpy().copy()
            th[0] *= 1 + 2e-5 * (j + 1)
            th[1] += 0.01 * (j + 1)
            th[2] *= 0.9
            th[3] *= 0.8
            th[4] = 0.3
            th[5] = 0.4
            thetas.append(th)
    mc = np.repeat(np.arange(n_curves), n_per_curve)
    mb = R.ModelBatch.build(curves, mc, 0)
    return mb, np.asarray(thetas), curves, mc, np.asarray(truths)


def torch_flux(theta, curves, mc, dtype=torch.float64):
    out = []
    for m, c in enumerate(mc):
        t = torch.tensor(curves[c][0], device="cuda")
        out.append(R.transit_model_torch(theta[m:m + 1].to(dtype), t[None])[0])
    return torch.cat(out)


@pytest.mark.gpu
def test_transit_forward_matches_torch_model(gpu):
    mb, th, curves, mc, _ = model_batch()
    theta = torch.tensor(th, device="cuda")
    want = torch_flux(theta, curves, mc)
    f64 = R.transit_fwd(mb, theta, "fp64")
    f32 = R.transit_fwd(mb, theta, "fp32c")
    assert f32.dtype == torch.float32 and f64.dtype == torch.float64
    assert float((f64 - want).abs().max()) < 1e-13
    assert float((f32.double() - want).abs().max()) < 2e-6


@pytest.mark.gpu
def test_transit_backward_matches_autograd_and_central_differences(gpu):
    mb, th, curves, mc, _ = model_batch(seed=1)
    theta = torch.tensor(th, device="cuda")
    # upstream gradient of the chi2 loss at theta
    thr = theta.clone().requires_grad_(True)
    f = torch_flux(thr, curves, mc)
    loss = (mb.w * (mb.y - f) ** 2).sum()
    loss.backward()
    want = thr.grad
    gflux = (-2 * mb.w * (mb.y - f.detach()))
    got64 = R.transit_bwd(mb, theta, gflux, "fp64")
    scale = want.abs().amax(dim=1, keepdim=True)
    assert float(((got64 - want).abs() / scale).max()) < 1e-9
    got32 = R.transit_bwd(mb, theta, gflux.float(), "fp32c")
    assert float(((got32 - want).abs() / scale).max()) < 2e-3
    # float64 central differences of the kernel's own forward, h = 1e-6 relative, with one
    # Richardson step (h and h / 2) against the O(h^2) error of the sharp period direction
    def kloss(tt):
        ff = R.transit_fwd(mb, tt, "fp64")
        return R.chi2_per_model(ff, mb)

    def central(k, h):
        tp, tm = theta.clone(), theta.clone()
        tp[:, k] += h
        tm[:, k] -= h
        return (kloss(tp) - kloss(tm)) / (2 * h)

    for k in range(6):
        h = 1e-6 * theta[:, k].abs()
        fd = (4 * central(k, h / 2) - central(k, h)) / 3
        err = (fd - got64[:, k]).abs() / scale[:, 0]
        assert float(err.max()) < 1e-5, (k, fd, got64[:, k])


@pytest.mark.gpu
def test_fused_chi2_matches_the_flux_path(gpu):
    mb, th, curves, mc, _ = model_batch(seed=5)
    theta = torch.tensor(th, device="cuda")
    # float64 reference: the torch model, its loss and autograd gradient
    thr = theta.clone().requires_grad_(True)
    f = torch_flux(thr, curves, mc)
    per_model = R.chi2_per_model(f, mb)
    per_model.sum().backward()
    want_c, want_g = per_model.detach(), thr.grad
    scale = want_g.abs().amax(dim=1, keepdim=True)
    c64, g64 = R.transit_chi2(mb, theta, "fp64")
    assert float(((c64 - want_c).abs() / want_c).max()) < 1e-12
    assert float(((g64 - want_g).abs() / scale).max()) < 1e-9
    c64n, none = R.transit_chi2(mb, theta, "fp64", grad=False)
    assert none is None and float(((c64n - c64).abs() / c64).max()) < 1e-13
    # float32 model: the same flux values as the forward kernel's
    c32, g32 = R.transit_chi2(mb, theta, "fp32c")
    flux_c32 = R.chi2_per_model(R.transit_fwd(mb, theta, "fp32c"), mb)
    assert float(((c32 - flux_c32).abs() / flux_c32).max()) < 1e-7
    assert float(((c32 - want_c).abs() / want_c).max()) < 1e-4
    assert float(((g32 - want_g).abs() / scale).max()) < 2e-3
    c32n, _ = R.transit_chi2(mb, theta, "fp32c", grad=False)
    assert float(((c32n - c32).abs() / c32).max()) < 1e-7
    # through autograd: the same gradient as the flux kernels' chain
    ta = theta.clone().requires_grad_(True)
    R.chi2_fused(ta, mb, "fp64").sum().backward()
    assert float(((ta.grad - want_g).abs() / scale).max()) < 1e-9


@pytest.mark.gpu
def test_grouped_chi2_matches_the_per_model_kernel(gpu):
    """Curves of exactly one point tile, one tile and a point, several tiles, and one with no
    model; models listed out of curve order: the grouped kernel (each point read once) and
    the one-workgroup-per-model kernel agree for every model, in both modes."""
    from gtls_amd import _ffi
    tp = int(_ffi.load().gtls_transit_tile_points())
    rng = np.random.default_rng(11)
    sizes = [tp, tp + 1, 3 * tp + 517, 700, 1500]
    curves, thetas, mc = [], [], []
    for c, n in enumerate(sizes):
        t = np.sort(rng.uniform(200.0, 230.0, n))
        truth = theta_row(1.7 + 0.4 * c, 200.4 + 0.05 * c, 0.1 + 0.01 * c, 3e-3)
        y = R.transit_model_torch(truth, torch.tensor(t)[None])[0].numpy()
        curves.append((t, y + rng.normal(0, 5e-4, n), rng.uniform(3e-4, 8e-4, n)))
        if c == 3:
            continue   # a curve without models
        for j in range(1 + (c % 3) * 2):
            th = truth[0].numpy().copy()
            th[0] *= 1 + 1e-5 * j
            th[1] += 0.004 * j
            th[4], th[5] = 0.2 + 0.1 * j, 0.35
            thetas.append(th)
            mc.append(c)
    perm = rng.permutation(len(mc))
    mc = np.asarray(mc)[perm]
    theta = torch.tensor(np.asarray(thetas)[perm], device="cuda")
    mb = R.ModelBatch.build(curves, mc, 0)
    assert mb.model_order is not None and mb.max_models == 5
    for mode, tol in (("fp64", 1e-12), ("fp32c", 1e-11)):
        c1, g1 = R.transit_chi2(mb, theta, mode, grouped=False)
        c2, g2 = R.transit_chi2(mb, theta, mode)
        scale = g1.abs().amax(dim=1, keepdim=True)
        assert float(((c2 - c1).abs() / c1).max()) < tol, mode
        assert float(((g2 - g1).abs() / scale).max()) < tol, mode
        c3, none = R.transit_chi2(mb, theta, mode, grad=False)
        assert none is None and float(((c3 - c1).abs() / c1).max()) < tol, mode


@pytest.mark.gpu
def test_autograd_function_chains_through_torch(gpu):
    mb, th, curves, mc, _ = model_batch(seed=2)
    theta = torch.tensor(th, device="cuda", requires_grad=True)
    f = R.transit_flux(theta, mb, "fp64")
    R.chi2_per_model(f, mb).sum().backward()
    theta2 = torch.tensor(th, device="cuda", requires_grad=True)
    f2 = torch_flux(theta2, curves, mc)
    (mb.w * (mb.y - f2) ** 2).sum().backward()
    scale = theta2.grad.abs().amax(dim=1, keepdim=True)
    assert float(((theta.grad - theta2.grad).abs() / scale).max()) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [True, False])
def test_refinement_recovers_injected_transits(gpu, fused):
    mb, th, curves, mc, truths = model_batch(seed=4, n=6000)
    spans = np.array([np.ptp(curves[c][0]) for c in mc])
    ref = R.Refiner(mb, th, spans, mode="fp32c", fused=fused)
    first, last = ref.run(200)
    assert torch.all(last < first)
    got = ref.theta().detach().cpu().numpy()
    for m, c in enumerate(mc):
        P, T0, D, depth = truths[c][:4]
        assert abs(got[m, 0] - P) < 0.25 * D * P / spans[m]
        dt = (got[m, 1] - T0 + P / 2) % P - P / 2
        assert abs(dt) < 0.1 * D
        assert got[m, 2] == pytest.approx(D, rel=0.15)
        assert got[m, 3] == pytest.approx(depth, rel=0.15)


@pytest.mark.gpu
def test_captured_step_replays_match_eager_steps(gpu):
    """A HIP graph of one refinement step (Refiner.capture) replayed 6 times gives the
    eager steps' parameters and losses bit for bit; between replays the process runs eager
    work of its own and synchronises (every replay is checked)."""
    mb, th, curves, mc, _ = model_batch(seed=4, n=6000)
    spans = np.array([np.ptp(curves[c][0]) for c in mc])
    a = R.Refiner(mb, th, spans, mode="fp32c", capturable=True)
    want_z, want_loss = [], []
    for i in range(8):
        loss = a.step()
        torch.cuda.synchronize()
        want_z.append(a.z.detach().clone())
        want_loss.append(loss.clone())
    b = R.Refiner(mb, th, spans, mode="fp32c", capturable=True)
    g = b.capture(warmup=2)
    torch.cuda.synchronize()
    assert torch.equal(b.z.detach(), want_z[1])
    for i in range(2, 8):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(g.loss, want_loss[i]), i
        assert torch.equal(b.z.detach(), want_z[i]), i
        # eager work between replays
        other = R.chi2_fused(b.theta().detach(), mb, "fp32c")
        torch.cuda.synchronize()
        assert torch.all(torch.isfinite(other))
