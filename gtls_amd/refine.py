"""Gradient refinement of grid peaks with a smooth limb-darkened transit model.

Each grid peak of a light curve (period, best width and T0 position of the search) seeds a
model with parameters theta = (P, T0, duration, depth, q1, q2); the models are fitted to
their curves by Adam on the weighted chi2 sum_i w_i (y_i - f_i)^2. The per-model chi2 and
its gradient come from one fused HIP kernel pass (gtls_transit_chi2_grouped: each point
is read once and every model of its curve evaluated on it; wrapped as the
torch.autograd.Function `TransitChi2`); for losses of the caller's own, the model flux of
every point and its backward run in gtls_transit_fwd / gtls_transit_bwd (`TransitModel`).
All three kernels are in csrc/transit_model.h; the parameter transform is ordinary torch
code.

The reference has no such step: it reports the grid values refined on the host
(src/gputls/stats.py:233-488). `transit_model_torch` restates the model in torch
operations (any dtype) for checking the kernels' values and gradients.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _ffi
from .search import MODES, _ctx, _device, _torch

KAPPA = 64.0     # edge sharpness of the model (csrc/transit_model.h: kTransitKappa)
LOG_CUT = 60.0   # kappa s below -LOG_CUT: outside the model (kLogCut)


def q_from_u(u1, u2):
    """Kipping's (q1, q2) of quadratic limb darkening (u1, u2)."""
    s = u1 + u2
    return s * s, u1 / (2.0 * s)


def u_from_q(q1, q2):
    sq = np.sqrt(q1)
    return 2.0 * sq * q2, sq * (1.0 - 2.0 * q2)


def transit_model_torch(theta, t, kappa=KAPPA):
    """The model of csrc/transit_model.h in torch operations: theta [..., 6] against times
    t [..., N] (broadcast), in theta's dtype (the fold in float64)."""
    torch = _torch()
    P, T0, D, delta, q1, q2 = [theta[..., k:k + 1] for k in range(6)]
    td = t.to(torch.float64)
    e = td - T0.to(torch.float64)
    n = torch.round(e / P.to(torch.float64)).detach()
    tau = (e - n * P.to(torch.float64)).to(theta.dtype)
    x = 2 * tau / D
    ks_raw = kappa * (1 - x * x)
    ks = torch.clamp(ks_raw, min=-LOG_CUT)   # keeps the unused branch's gradients finite
    W = torch.sigmoid(ks)
    mu = torch.sqrt(torch.nn.functional.softplus(ks) / kappa)
    sq1 = torch.sqrt(q1)
    u1, u2 = 2 * sq1 * q2, sq1 * (1 - 2 * q2)
    om = 1 - mu
    f = 1 - delta * W * (1 - u1 * om - u2 * om * om)
    return torch.where(ks_raw < -LOG_CUT, torch.ones_like(f), f)


@dataclass
class ModelBatch:
    """Device layout of a set of models: model m fits curve model_curve[m] and owns the
    flux slots [out_off[m], out_off[m + 1])."""
    dev: object
    t: object            # float64 [total points], curves back to back
    pt_off: object       # int64 [n_curves + 1]
    model_curve: object  # int32 [n_models]
    out_off: object      # int64 [n_models + 1]
    y: object            # float64 [total slots]: flux of the point behind each slot
    w: object            # float64 [total slots]: its weight 1 / dy^2
    max_points: int
    n_models: int
    y_pt: object = None  # float64 [total points]: flux per point (like t)
    w_pt: object = None  # float64 [total points]: weight per point
    # the models grouped by curve for gtls_transit_chi2_grouped: grouped slot k holds model
    # model_order[k] (None when the models already come curve by curve), model_rank its
    # inverse; per curve its model and point-tile offsets; part, the tile-sum scratch
    model_order: object = None
    model_rank: object = None
    model_off: object = None   # int64 [n_curves + 1]
    tile_off: object = None    # int64 [n_curves + 1]
    n_curves: int = 0
    n_tiles: int = 0
    max_models: int = 0
    part: object = None        # float64 [n_tiles * (max_models * 7 + 1)]

    @classmethod
    def build(cls, curves, model_curve, device=0):
        """curves: list of (t, y, dy) host arrays; model_curve: curve index of each model."""
        torch = _torch()
        dev = _device(device)
        model_curve = np.asarray(model_curve, np.int32)
        n = np.array([len(c[0]) for c in curves], np.int64)
        pt_off = np.concatenate([[0], np.cumsum(n)])
        t = np.concatenate([np.asarray(c[0], np.float64) for c in curves])
        yc = [np.asarray(c[1], np.float64) for c in curves]
        wc = [1.0 / np.asarray(c[2], np.float64) ** 2 for c in curves]
        nm = n[model_curve]
        out_off = np.concatenate([[0], np.cumsum(nm)])
        y = np.concatenate([yc[c] for c in model_curve]) if len(model_curve) else np.zeros(0)
        w = np.concatenate([wc[c] for c in model_curve]) if len(model_curve) else np.zeros(0)
        f64 = dict(dtype=torch.float64, device=dev)
        y_pt = np.concatenate(yc) if yc else np.zeros(0)
        w_pt = np.concatenate(wc) if wc else np.zeros(0)
        # grouping by curve (stable), point tiles and the tile-sum scratch
        n_curves = len(curves)
        if len(model_curve) and (model_curve.min() < 0 or model_curve.max() >= n_curves):
            raise ValueError("model_curve holds a curve index outside [0, %d)" % n_curves)
        counts = np.bincount(model_curve, minlength=n_curves) if n_curves else np.zeros(0, int)
        order = np.argsort(model_curve, kind="stable")
        grouped = np.array_equal(order, np.arange(len(order)))
        model_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        tp = int(_ffi.load().gtls_transit_tile_points())
        tile_off = np.concatenate([[0], np.cumsum((n + tp - 1) // tp)]).astype(np.int64)
        n_tiles = int(tile_off[-1])
        max_models = int(counts.max(initial=0))
        i64 = dict(dtype=torch.int64, device=dev)
        return cls(dev, torch.as_tensor(t, **f64), torch.as_tensor(pt_off, device=dev),
                   torch.as_tensor(model_curve, device=dev), torch.as_tensor(out_off, device=dev),
                   torch.as_tensor(y, **f64),
                   torch.as_tensor(w, **f64), int(n.max(initial=0)), len(model_curve),
                   torch.as_tensor(y_pt, **f64), torch.as_tensor(w_pt, **f64),
                   None if grouped else torch.as_tensor(order, **i64),
                   None if grouped else torch.as_tensor(np.argsort(order), **i64),
                   torch.as_tensor(model_off, **i64), torch.as_tensor(tile_off, **i64),
                   n_curves, n_tiles, max_models,
                   torch.empty(max(1, n_tiles * (max_models * 7 + 1)), **f64))

    @property
    def slots(self):
        return int(self.y.numel())


# optional HIP-event timing of the kernels: {"fwd": [(start, end), ...], "bwd": [...],
# "chi2": [...]}
_TIMING = None


def set_timing(on=True):
    """Record torch (HIP) events around every gtls_transit_fwd / _bwd / _chi2 call from now
    on."""
    global _TIMING
    _TIMING = {"fwd": [], "bwd": [], "chi2": []} if on else None


def timing_ms():
    """Summed milliseconds and call counts of the recorded forward and backward kernels
    (waits for them), then clears the records."""
    torch = _torch()
    out = {}
    if _TIMING is None:
        return out
    torch.cuda.synchronize()
    for k, evs in _TIMING.items():
        out[k] = (sum(a.elapsed_time(b) for a, b in evs), len(evs))
        evs.clear()
    return out


def _timed(kind, fn):
    if _TIMING is None:
        return fn()
    torch = _torch()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    r = fn()
    b.record()
    _TIMING[kind].append((a, b))
    return r


def _flux_dtype(mode):
    torch = _torch()
    return torch.float64 if mode == "fp64" else torch.float32


def transit_fwd(mb: ModelBatch, theta, mode="fp32c"):
    """Model flux of every slot (float32 in fp32c mode, float64 in fp64 mode)."""
    torch = _torch()
    theta = theta.detach().to(torch.float64).contiguous()
    if theta.shape != (mb.n_models, 6) or theta.device != mb.dev:
        raise ValueError("theta must be a float64 [n_models, 6] tensor on the batch's device")
    flux = torch.empty(mb.slots, dtype=_flux_dtype(mode), device=mb.dev)
    ctx = _ctx(mb.dev)
    code = _timed("fwd", lambda: _ffi.load().gtls_transit_fwd(
        ctx.handle, MODES[mode], mb.n_models, _ffi.ptr(theta), _ffi.ptr(mb.model_curve),
        _ffi.ptr(mb.pt_off), _ffi.ptr(mb.t), _ffi.ptr(mb.out_off), mb.max_points,
        _ffi.ptr(flux)))
    _ffi.check(ctx.handle, code)
    return flux


def transit_bwd(mb: ModelBatch, theta, grad_flux, mode="fp32c"):
    """grad_theta [n_models, 6] (float64) of sum(grad_flux * flux)."""
    torch = _torch()
    theta = theta.detach().to(torch.float64).contiguous()
    g = grad_flux.detach().to(_flux_dtype(mode)).contiguous()
    if g.numel() != mb.slots:
        raise ValueError("grad_flux has %d slots, the batch %d" % (g.numel(), mb.slots))
    out = torch.empty((mb.n_models, 6), dtype=torch.float64, device=mb.dev)
    ctx = _ctx(mb.dev)
    code = _timed("bwd", lambda: _ffi.load().gtls_transit_bwd(
        ctx.handle, MODES[mode], mb.n_models, _ffi.ptr(theta), _ffi.ptr(mb.model_curve),
        _ffi.ptr(mb.pt_off), _ffi.ptr(mb.t), _ffi.ptr(mb.out_off), _ffi.ptr(g),
        _ffi.ptr(out)))
    _ffi.check(ctx.handle, code)
    return out


def transit_chi2(mb: ModelBatch, theta, mode="fp32c", grad=True, grouped=True):
    """Per-model chi2 = sum_i w_i (y_i - f_i)^2 over the model's curve (float64 [n_models])
    and, when grad, its gradient with respect to theta (float64 [n_models, 6]; else None),
    in one fused pass that writes no flux slots: by default gtls_transit_chi2_grouped (every
    point read once, all models of its curve evaluated on it), with grouped=False
    gtls_transit_chi2 (one workgroup per model)."""
    torch = _torch()
    theta = theta.detach().to(torch.float64).contiguous()
    if theta.shape != (mb.n_models, 6) or theta.device != mb.dev:
        raise ValueError("theta must be a float64 [n_models, 6] tensor on the batch's device")
    if mb.y_pt is None or mb.w_pt is None:
        raise ValueError("the batch has no per-point flux and weights (ModelBatch.build sets them)")
    chi2 = torch.empty(mb.n_models, dtype=torch.float64, device=mb.dev)
    g = torch.empty((mb.n_models, 6), dtype=torch.float64, device=mb.dev) if grad else None
    ctx = _ctx(mb.dev)
    lib = _ffi.load()
    if not grouped or mb.model_off is None:
        code = _timed("chi2", lambda: lib.gtls_transit_chi2(
            ctx.handle, MODES[mode], mb.n_models, _ffi.ptr(theta), _ffi.ptr(mb.model_curve),
            _ffi.ptr(mb.pt_off), _ffi.ptr(mb.t), _ffi.ptr(mb.y_pt), _ffi.ptr(mb.w_pt),
            _ffi.ptr(chi2), _ffi.ptr(g) if grad else None))
        _ffi.check(ctx.handle, code)
        return chi2, g
    th = theta if mb.model_order is None else theta[mb.model_order].contiguous()
    code = _timed("chi2", lambda: lib.gtls_transit_chi2_grouped(
        ctx.handle, MODES[mode], mb.n_curves, mb.n_models, mb.max_models, _ffi.ptr(th),
        _ffi.ptr(mb.model_off), _ffi.ptr(mb.pt_off), _ffi.ptr(mb.tile_off), mb.n_tiles,
        _ffi.ptr(mb.t), _ffi.ptr(mb.y_pt), _ffi.ptr(mb.w_pt), _ffi.ptr(mb.part),
        mb.part.numel(), _ffi.ptr(chi2), _ffi.ptr(g) if grad else None))
    _ffi.check(ctx.handle, code)
    if mb.model_order is not None:
        chi2 = chi2[mb.model_rank]
        g = g[mb.model_rank] if grad else None
    return chi2, g


def _chi2_function():
    torch = _torch()

    class TransitChi2(torch.autograd.Function):
        """Per-model chi2 of the model against its curve; the gradient comes from the same
        fused pass as the value."""

        @staticmethod
        def forward(ctx, theta, mb, mode):
            chi2, g = transit_chi2(mb, theta, mode, grad=ctx.needs_input_grad[0])
            ctx.save_for_backward(g if g is not None else chi2)
            return chi2

        @staticmethod
        def backward(ctx, grad_chi2):
            (g,) = ctx.saved_tensors
            return (grad_chi2[:, None] * g).to(grad_chi2.dtype), None, None

    return TransitChi2


_TC = None


def chi2_fused(theta, mb: ModelBatch, mode="fp32c"):
    """Differentiable per-model chi2 (the fused kernel): the same value as
    chi2_per_model(transit_flux(theta, mb, mode), mb)."""
    global _TC
    if _TC is None:
        _TC = _chi2_function()
    return _TC.apply(theta, mb, mode)


def _function():
    torch = _torch()

    class TransitModel(torch.autograd.Function):
        """flux = model(theta) on the HIP kernels; backward through gtls_transit_bwd."""

        @staticmethod
        def forward(ctx, theta, mb, mode):
            ctx.save_for_backward(theta)
            ctx.mb = mb
            ctx.mode = mode
            return transit_fwd(mb, theta, mode)

        @staticmethod
        def backward(ctx, grad_flux):
            (theta,) = ctx.saved_tensors
            g = transit_bwd(ctx.mb, theta, grad_flux, ctx.mode)
            return g.to(theta.dtype), None, None

    return TransitModel


_TM = None


def transit_flux(theta, mb: ModelBatch, mode="fp32c"):
    """Differentiable model flux of every slot of the batch."""
    global _TM
    if _TM is None:
        _TM = _function()
    return _TM.apply(theta, mb, mode)


def chi2_per_model(flux, mb: ModelBatch):
    """sum over each model's points of w (y - f)^2, float64 (a segmented sum over the
    models' contiguous slot ranges)."""
    torch = _torch()
    r = mb.y - flux.to(torch.float64)
    return torch.segment_reduce(mb.w * r * r, "sum", offsets=mb.out_off)


def _logit(p):
    p = np.clip(p, 1e-6, 1 - 1e-6)
    return np.log(p / (1 - p))


class Refiner:
    """Adam on unconstrained offsets z of the seeds theta0:
        P = P0 + sP zP, T0 = T0_0 + sT zT, D = D0 exp(zD / 4), depth = d0 exp(zd / 2),
        q1 = sigmoid(z1), q2 = sigmoid(z2),
    with sP = D0 P0 / (4 span) (a quarter duration of drift over the baseline) and
    sT = D0 / 4. With fused (the default) the loss and its gradient come from one kernel
    pass per step (gtls_transit_chi2); otherwise from the flux kernels and torch's
    segmented sum. With capturable, the optimizer keeps its step counts on the device, so
    that capture() can record a whole step in a HIP graph."""

    def __init__(self, mb: ModelBatch, theta0, spans, mode="fp32c", lr=0.05, fused=True,
                 capturable=False):
        torch = _torch()
        self.mb = mb
        self.mode = mode
        self.fused = fused
        th = np.asarray(theta0, np.float64).reshape(-1, 6)
        spans = np.maximum(np.asarray(spans, np.float64), 1e-9)
        self.base = torch.as_tensor(th[:, :4], dtype=torch.float64, device=mb.dev)
        sc = np.stack([th[:, 2] * th[:, 0] / (4 * spans), th[:, 2] / 4], axis=1)
        self.scale = torch.as_tensor(sc, dtype=torch.float64, device=mb.dev)
        z = np.zeros((len(th), 6))
        z[:, 4] = _logit(th[:, 4])
        z[:, 5] = _logit(th[:, 5])
        self.z = torch.tensor(z, dtype=torch.float64, device=mb.dev, requires_grad=True)
        self.opt = torch.optim.Adam([self.z], lr=lr, fused=True, capturable=capturable)
        self.capturable = capturable

    def theta(self):
        torch = _torch()
        z = self.z
        return torch.stack([
            self.base[:, 0] + self.scale[:, 0] * z[:, 0],
            self.base[:, 1] + self.scale[:, 1] * z[:, 1],
            self.base[:, 2] * torch.exp(0.25 * z[:, 2]),
            self.base[:, 3] * torch.exp(0.5 * z[:, 3]),
            torch.sigmoid(z[:, 4]),
            torch.sigmoid(z[:, 5]),
        ], dim=1)

    def loss(self):
        if self.fused:
            return chi2_fused(self.theta(), self.mb, self.mode)
        return chi2_per_model(transit_flux(self.theta(), self.mb, self.mode), self.mb)

    def step(self):
        self.opt.zero_grad(set_to_none=True)
        per_model = self.loss()
        per_model.sum().backward()
        self.opt.step()
        return per_model.detach()

    def capture(self, warmup=2):
        """A HIP graph (torch.cuda.CUDAGraph) of one step: replay() runs the parameter
        transform, the fused loss and its gradient, the backward and the Adam update with
        no host work; the graph's `loss` tensor holds the per-model chi2 of the parameters
        the replayed step started from. Needs fused and capturable. The `warmup` steps run
        first, as ordinary steps on the capture stream (they create that stream's library
        context and the optimizer state before the capture, so nothing is allocated
        outside torch's graph pool while recording)."""
        torch = _torch()
        if not (self.fused and self.capturable):
            raise ValueError("capture() needs a Refiner built with fused=True, capturable=True")
        if _TIMING is not None:
            raise RuntimeError("switch kernel timing off (set_timing(False)) before capture()")
        dev = self.mb.dev
        s = torch.cuda.Stream(dev)
        s.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(s):
            for _ in range(max(1, warmup)):
                self.step()
        torch.cuda.current_stream(dev).wait_stream(s)
        torch.cuda.synchronize(dev)
        g = torch.cuda.CUDAGraph()
        self.opt.zero_grad(set_to_none=True)
        with torch.cuda.graph(g, stream=s):
            loss = self.loss()
            loss.sum().backward()
            self.opt.step()
        g.loss = loss
        return g

    def run(self, steps=200, progress=None):
        """Adam steps; returns the per-model chi2 before the first and after the last.
        progress(i) is called after every 50th step."""
        first = self.step()
        for i in range(1, steps):
            self.step()
            if progress is not None and (i + 1) % 50 == 0:
                progress(i + 1)
        with _torch().no_grad():
            last = self.loss()
        return first, last


def seed_peaks(t, y, dy, periods, power, widths, best_k, best_s, n_peaks=8, device=0):
    """theta seeds of the n_peaks highest local maxima of `power` (a curve's search):
    period P, T0 at the centre of the best window in the phase order of the fold
    x - trunc(x), x = t / P, duration = its phase extent times P, depth = the window's
    mean of 1 - y, q1/q2 of the default limb darkening. Returns (theta [k, 6], indices)."""
    torch = _torch()
    from . import constants as C
    power = np.asarray(power, np.float64)
    p = np.where(np.isfinite(power), power, -np.inf)
    left = np.concatenate([[-np.inf], p[:-1]])
    right = np.concatenate([p[1:], [-np.inf]])
    peaks = np.nonzero((p >= left) & (p > right) & np.isfinite(p))[0]
    peaks = peaks[np.argsort(-p[peaks], kind="stable")][:n_peaks]
    dev = _device(device)
    n = len(t)
    q1, q2 = q_from_u(*C.DEFAULT_U)
    if len(peaks) == 0:
        return np.zeros((0, 6)), peaks
    td = torch.as_tensor(np.asarray(t, np.float64), device=dev)
    yd = torch.as_tensor(np.asarray(y, np.float64), device=dev)
    P = torch.as_tensor(np.asarray(periods, np.float64)[peaks], device=dev)
    d = np.asarray(widths, np.int64)[np.asarray(best_k)[peaks].astype(np.int64)]
    s0 = np.asarray(best_s, np.int64)[peaks]
    x = td[None, :] / P[:, None]
    ph = x - torch.trunc(x)                                   # [k, N]
    order = torch.sort(ph, dim=1, stable=True).indices
    dmax = int(d.max())
    j = torch.as_tensor(s0, device=dev)[:, None] + torch.arange(dmax, device=dev)[None, :]
    valid = torch.arange(dmax, device=dev)[None, :] < torch.as_tensor(d, device=dev)[:, None]
    src = torch.gather(order, 1, j % n)                       # window samples, phase order
    wph = torch.gather(ph, 1, src) + (j >= n).to(torch.float64)
    lo = wph[:, 0]
    hi = torch.where(valid, wph, torch.full_like(wph, -1.0)).amax(dim=1)
    dep = ((1.0 - yd[src]) * valid).sum(dim=1) / valid.sum(dim=1)
    lo, hi, dep, Ph = [v.cpu().numpy() for v in (lo, hi, dep, P)]
    tmin = float(np.min(t))
    out = []
    for i in range(len(peaks)):
        mid = 0.5 * (lo[i] + hi[i])
        T0 = mid * Ph[i] + np.floor(tmin / Ph[i]) * Ph[i]
        D = max((hi[i] - lo[i]) * Ph[i], Ph[i] / n)
        out.append([Ph[i], T0, D, max(float(dep[i]), 1e-6), q1, q2])
    return np.asarray(out, np.float64).reshape(-1, 6), peaks
