"""Host orchestration of the device search: search plans and calls into the C ABI.

A `CurvePlan` holds everything one light curve needs on the device -- cleaned samples,
period grid, template table, per-period width ranges -- built on the host the way the
reference builds it (src/gputls/main.py:18-60, src/gputls/core.py:568-690). The chi2
search itself runs in libgtls_amd.so; this module only moves arrays and checks codes.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import numpy as np

from . import _ffi
from . import constants as C
from .grid import duration_grid, period_grid, skip_factor, width_ranges, width_window
from .template import TemplateTable, build_templates

MODES = {"fp64": _ffi.MODE_FP64, "fp32c": _ffi.MODE_FP32C}


def _torch():
    import torch
    return torch


@dataclass
class CurvePlan:
    t: np.ndarray
    y: np.ndarray
    dy: np.ndarray
    periods: np.ndarray        # ascending, days
    table: TemplateTable
    overview: np.ndarray       # template metadata of every duration (reference layout)
    rows: np.ndarray           # template rows of every duration
    k_lo: np.ndarray           # per period, inclusive width index range
    k_hi: np.ndarray

    @property
    def n(self):
        return len(self.t)

    @property
    def span(self):
        return float(np.max(self.t) - np.min(self.t))

    def ranges(self, periods):
        lo, hi = width_window(periods, self.span, self.n)
        return width_ranges(self.table.widths, lo, hi)


def make_plan(t, y, dy, args, periods=None):
    """Period grid, durations, templates and width ranges of one cleaned light curve."""
    span = float(np.max(t) - np.min(t))
    if periods is None or len(periods) == 0:
        periods = period_grid(R_star=args.R_star, M_star=args.M_star, time_span=span,
                              period_min=args.period_min, period_max=args.period_max,
                              oversampling_factor=args.oversampling_factor,
                              n_transits_min=args.n_transits_min)
    periods = np.sort(np.asarray(periods, dtype=np.float64))
    durations = duration_grid(periods, shortest=1 / len(t), log_step=args.duration_grid_step)
    table, overview, rows = build_templates(len(t), durations, per=args.per, rp=args.rp,
                                            a=args.a, inc=args.inc, ecc=args.ecc, w=args.w,
                                            u=args.u, limb_dark=args.limb_dark)
    lo, hi = width_window(periods, span, len(t))
    k_lo, k_hi = width_ranges(table.widths, lo, hi)
    return CurvePlan(np.asarray(t, np.float64), np.asarray(y, np.float64),
                     np.asarray(dy, np.float64), periods, table, overview, rows, k_lo, k_hi)


def _device(device):
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("gtls_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", int(device))


def _ctx(dev):
    torch = _torch()
    stream = torch.cuda.current_stream(dev).cuda_stream
    return _ffi.Context.get(dev.index, stream)


def _ptr(x):
    return ct.c_void_p(x.data_ptr())


def set_kernel_timing(device=0, on=True):
    """Record HIP events around the search kernels of later searches on torch's current
    stream of `device` (gtls_set_timing)."""
    dev = _device(device)
    ctx = _ctx(dev)
    _ffi.check(ctx.handle, _ffi.load().gtls_set_timing(ctx.handle, 1 if on else 0))


def kernel_ms(device=0):
    """(prep, windows, reduce) milliseconds of the search kernels recorded since the last
    call, and the number of period chunks (gtls_kernel_ms; waits for them to finish)."""
    dev = _device(device)
    ctx = _ctx(dev)
    ms = (ct.c_double * 3)()
    n = _ffi.load().gtls_kernel_ms(ctx.handle, ms)
    if n < 0:
        _ffi.check(ctx.handle, int(n))
    return (ms[0], ms[1], ms[2]), int(n)


_WORKSPACES = {}


def workspace(dev, nbytes):
    """Search workspace shared by every search on (device, current stream): searches on
    one stream run in order, so they can reuse one buffer; it grows to the largest need."""
    torch = _torch()
    key = (dev.index, torch.cuda.current_stream(dev).cuda_stream)
    ws = _WORKSPACES.get(key)
    if ws is None or ws.numel() < nbytes:
        _WORKSPACES.pop(key, None)
        ws = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        _WORKSPACES[key] = ws
    return ws


class DeviceCurve:
    """Device-resident arrays of one curve (reused across repeated searches)."""

    def __init__(self, plan: CurvePlan, device=0):
        torch = _torch()
        dev = _device(device)
        self.dev = dev
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        tb = plan.table
        self.n = plan.n
        self.t = torch.as_tensor(plan.t, **f64)
        self.y = torch.as_tensor(plan.y, **f64)
        self.dy = torch.as_tensor(plan.dy, **f64)
        self.widths = torch.as_tensor(tb.widths, **i32)
        self.overshoot = torch.as_tensor(tb.overshoot, **f64)
        self.tmpl_off = torch.as_tensor(tb.offsets[:-1], dtype=torch.int64, device=dev)
        self.tmpl = torch.as_tensor(tb.S if len(tb.S) else np.ones(1), **f64)
        self.total_tmpl = int(tb.offsets[-1])
        self.n_widths = tb.n_widths
        self.max_width = int(tb.widths.max())
        self.negative_time = int(np.min(plan.t) < 0)

    def workspace(self, ctx, mode, n_periods):
        lib = _ffi.load()
        need = lib.gtls_search_workspace_bytes_1(ctx.handle, mode, self.n, n_periods,
                                                 self.n_widths, self.total_tmpl,
                                                 self.max_width)
        if need < 0:
            _ffi.check(ctx.handle, int(need))
        return workspace(self.dev, need)

    def search(self, periods, k_lo, k_hi, mode="fp64", skip_point=8,
               depth_min=C.TRANSIT_DEPTH_MIN, sync=True):
        """chi2, best width index, best position per period (torch tensors on device)."""
        torch = _torch()
        m = MODES[mode]
        dev = self.dev
        per = torch.as_tensor(np.asarray(periods, np.float64), dtype=torch.float64, device=dev)
        klo = torch.as_tensor(np.asarray(k_lo, np.int32), dtype=torch.int32, device=dev)
        khi = torch.as_tensor(np.asarray(k_hi, np.int32), dtype=torch.int32, device=dev)
        npd = per.numel()
        chi2 = torch.empty(npd, dtype=torch.float64, device=dev)
        bk = torch.empty(npd, dtype=torch.int32, device=dev)
        bs = torch.empty(npd, dtype=torch.int32, device=dev)
        ctx = _ctx(dev)
        ws = self.workspace(ctx, m, npd)
        code = _ffi.load().gtls_search(
            ctx.handle, m, _ptr(self.t), _ptr(self.y), _ptr(self.dy), self.n, _ptr(per), npd,
            _ptr(klo), _ptr(khi), _ptr(self.widths), self.n_widths, self.max_width,
            _ptr(self.overshoot), _ptr(self.tmpl_off), _ptr(self.tmpl), self.total_tmpl,
            int(skip_point), float(depth_min), self.negative_time, _ptr(ws), ws.numel(),
            _ptr(chi2), _ptr(bk), _ptr(bs))
        _ffi.check(ctx.handle, code)
        if sync:
            torch.cuda.synchronize(dev)
        return chi2, bk, bs


def fold_sort(t, periods, device=0):
    """Stable phase permutations of t for every period (int32 [n_periods, n])."""
    torch = _torch()
    dev = _device(device)
    tt = torch.as_tensor(np.asarray(t, np.float64), dtype=torch.float64, device=dev)
    pp = torch.as_tensor(np.asarray(periods, np.float64), dtype=torch.float64, device=dev)
    n, npd = tt.numel(), pp.numel()
    out = torch.empty((npd, n), dtype=torch.int32, device=dev)
    ctx = _ctx(dev)
    lib = _ffi.load()
    need = lib.gtls_fold_sort_workspace_bytes(ctx.handle, n)
    if need < 0:
        _ffi.check(ctx.handle, int(need))
    ws = torch.empty(int(need), dtype=torch.uint8, device=dev)
    code = lib.gtls_fold_sort(ctx.handle, _ptr(tt), n, _ptr(pp), npd,
                              int(np.min(t) < 0), _ptr(ws), ws.numel(), _ptr(out))
    _ffi.check(ctx.handle, code)
    torch.cuda.synchronize(dev)
    return out


def search_plan(plan: CurvePlan, mode="fp64", T0_fit_margin=C.T0_FIT_MARGIN,
                depth_min=C.TRANSIT_DEPTH_MIN, device=0):
    """Full-grid chi2 of one plan; numpy (chi2, best_k, best_s)."""
    dc = DeviceCurve(plan, device)
    chi2, bk, bs = dc.search(plan.periods, plan.k_lo, plan.k_hi, mode=mode,
                             skip_point=skip_factor(T0_fit_margin), depth_min=depth_min)
    return chi2.cpu().numpy(), bk.cpu().numpy(), bs.cpu().numpy()
