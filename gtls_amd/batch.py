"""Batched search: many light curves, each with its own period grid and template table, in
one gtls_search_batched call per size group.

Curves are packed in CSR layout (points, periods, widths, template samples), largest
first so that the device work queue hands out the longest items early, and the
per-period outputs are mapped back to the caller's curve order.
"""
from __future__ import annotations

import ctypes as ct
import multiprocessing
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import _ffi
from . import constants as C
from .search import MODES, _ctx, _device, _ptr, make_plan, workspace
from .validate import validate_args, validate_inputs


def _plan_one(item):
    t, y, dy, kw = item
    t, y, dy = validate_inputs(t, y, dy)
    return make_plan(t, y, dy, validate_args(kw))


def build_plans(curves, kwargs=None, workers=None):
    """CurvePlans of (t, y, dy) tuples, built in a process pool. The workers are started
    fresh (spawn), not forked: a fork of a process whose GPU runtime is initialised copies
    its threads' locks in whatever state they hold, and a worker can then block forever."""
    kwargs = kwargs or {}
    items = [(c[0], c[1], c[2], kwargs) for c in curves]
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(items) < 8:
        return [_plan_one(it) for it in items]
    with ProcessPoolExecutor(max_workers=workers,
                             mp_context=multiprocessing.get_context("spawn")) as ex:
        return list(ex.map(_plan_one, items, chunksize=max(1, len(items) // (4 * workers))))


class BatchPlan:
    """Host CSR arrays of a list of CurvePlans."""

    def __init__(self, plans):
        self.plans = plans
        n = len(plans)
        # longest curves first: the work queue hands out the largest items early, and
        # consecutive curves have similar lengths (size groups, DeviceBatch)
        npts = np.array([p.n for p in plans], dtype=np.int64)
        self.order = np.argsort(-npts, kind="stable")      # packed position -> curve
        ps = [plans[i] for i in self.order]
        self.n_curves = n
        self.pt_off = np.zeros(n + 1, np.int64)
        self.per_off = np.zeros(n + 1, np.int64)
        self.wid_off = np.zeros(n + 1, np.int64)
        self.pt_off[1:] = np.cumsum([p.n for p in ps])
        self.per_off[1:] = np.cumsum([len(p.periods) for p in ps])
        self.wid_off[1:] = np.cumsum([p.table.n_widths for p in ps])
        self.t = np.concatenate([p.t for p in ps])
        self.y = np.concatenate([p.y for p in ps])
        self.dy = np.concatenate([p.dy for p in ps])
        self.periods = np.concatenate([p.periods for p in ps])
        self.k_lo = np.concatenate([p.k_lo for p in ps]).astype(np.int32)
        self.k_hi = np.concatenate([p.k_hi for p in ps]).astype(np.int32)
        self.widths = np.concatenate([p.table.widths for p in ps]).astype(np.int32)
        self.overshoot = np.concatenate([p.table.overshoot for p in ps])
        tbase = np.cumsum([0] + [len(p.table.S) for p in ps[:-1]])
        self.tmpl_off = np.concatenate([p.table.offsets[:-1] + b for p, b in zip(ps, tbase)])
        self.tmpl = np.concatenate([p.table.S for p in ps])
        self.max_points = int(max(p.n for p in ps))
        self.max_width = int(self.widths.max())
        ok = self.k_lo <= self.k_hi
        self.max_period_widths = int(max(1, np.max((self.k_hi - self.k_lo + 1)[ok], initial=1)))
        self.negative_time = int(np.min(self.t) < 0)

    @property
    def total_periods(self):
        return int(self.per_off[-1])

    def unpack(self, arr):
        """Per-curve arrays (caller order) of a packed per-period array."""
        out = [None] * self.n_curves
        for pos, ci in enumerate(self.order):
            out[ci] = arr[self.per_off[pos]:self.per_off[pos + 1]]
        return out


def size_groups(bp, ratio=2.0):
    """Contiguous ranges [c0, c1) of packed curves (longest first) in which no curve is
    more than `ratio` times shorter than the first: each range is searched with its own
    point capacity, so short curves get chunks of many more periods."""
    n = np.diff(bp.pt_off)
    groups, c0 = [], 0
    for c in range(1, len(n) + 1):
        if c == len(n) or n[c] * ratio < n[c0]:
            groups.append((c0, c))
            c0 = c
    return groups


class DeviceBatch:
    """Device-resident BatchPlan; search() enqueues one persistent search launch per size
    group of curves (size_groups)."""

    MAX_TIMED_CHUNKS = 4096       # period chunks a timed graph can hold stamps for

    def __init__(self, bp: BatchPlan, device=0):
        import torch
        self.bp = bp
        dev = _device(device)
        self.dev = dev

        def f64(a):
            return torch.as_tensor(np.ascontiguousarray(a, np.float64), device=dev)

        def i64(a):
            return torch.as_tensor(np.ascontiguousarray(a, np.int64), device=dev)

        def i32(a):
            return torch.as_tensor(np.ascontiguousarray(a, np.int32), device=dev)

        self.arrays = dict(
            pt_off=i64(bp.pt_off), t=f64(bp.t), y=f64(bp.y), dy=f64(bp.dy),
            per_off=i64(bp.per_off), periods=f64(bp.periods), k_lo=i32(bp.k_lo),
            k_hi=i32(bp.k_hi), wid_off=i64(bp.wid_off), widths=i32(bp.widths),
            overshoot=f64(bp.overshoot), tmpl_off=i64(bp.tmpl_off), tmpl=f64(bp.tmpl))
        npd = bp.total_periods
        self.chi2 = torch.empty(npd, dtype=torch.float64, device=dev)
        self.best_k = torch.empty(npd, dtype=torch.int32, device=dev)
        self.best_s = torch.empty(npd, dtype=torch.int32, device=dev)
        # per size group: offsets re-based to the group's first curve
        self.groups = []
        for c0, c1 in size_groups(bp):
            p0, p1 = int(bp.per_off[c0]), int(bp.per_off[c1])
            w0, w1 = int(bp.wid_off[c0]), int(bp.wid_off[c1])
            t0 = int(bp.pt_off[c0])
            tm0 = int(bp.tmpl_off[w0]) if w1 > w0 else 0
            tm1 = int(bp.tmpl_off[w1]) if w1 < len(bp.tmpl_off) else len(bp.tmpl)
            ok = bp.k_lo[p0:p1] <= bp.k_hi[p0:p1]
            kw = (bp.k_hi[p0:p1] - bp.k_lo[p0:p1] + 1)[ok]
            self.groups.append(dict(
                c0=c0, c1=c1, p0=p0, p1=p1, w0=w0, t0=t0, tm0=tm0,
                pt_off=i64(bp.pt_off[c0:c1 + 1] - t0), per_off=i64(bp.per_off[c0:c1 + 1] - p0),
                wid_off=i64(bp.wid_off[c0:c1 + 1] - w0),
                tmpl_off=i64(bp.tmpl_off[w0:w1] - tm0),
                total_points=int(bp.pt_off[c1]) - t0, total_periods=p1 - p0,
                total_widths=w1 - w0, total_tmpl=tm1 - tm0,
                max_points=int(np.max(np.diff(bp.pt_off[c0:c1 + 1]))),
                max_width=int(bp.widths[w0:w1].max()),
                max_period_widths=int(max(1, kw.max(initial=1)))))

    def batch_struct(self, skip_point, depth_min, g):
        """The C batch of size group g: re-based offsets, data pointers into the batch's
        arrays at the group's first curve / period / width / template sample."""
        a, bp = self.arrays, self.bp
        b = _ffi.GtlsBatch()
        b.n_curves = g["c1"] - g["c0"]
        b.pt_off = g["pt_off"].data_ptr()
        b.per_off = g["per_off"].data_ptr()
        b.wid_off = g["wid_off"].data_ptr()
        b.tmpl_off = g["tmpl_off"].data_ptr()
        for name, base in (("t", g["t0"]), ("y", g["t0"]), ("dy", g["t0"]),
                           ("periods", g["p0"]), ("k_lo", g["p0"]), ("k_hi", g["p0"]),
                           ("widths", g["w0"]), ("overshoot", g["w0"]), ("tmpl", g["tm0"])):
            x = a[name]
            setattr(b, name, x.data_ptr() + base * x.element_size())
        b.skip_point = int(skip_point)
        b.depth_min = float(depth_min)
        b.negative_time = bp.negative_time
        b.total_points = g["total_points"]
        b.total_periods = g["total_periods"]
        b.total_widths = g["total_widths"]
        b.total_tmpl = g["total_tmpl"]
        b.max_points = g["max_points"]
        b.max_width = g["max_width"]
        b.max_period_widths = g["max_period_widths"]
        return b

    def spectra(self, chi2, oversampling_factor=C.OVERSAMPLING_FACTOR, sync=True):
        """Device SR/SDE spectra of every curve (gtls_spectra_batched); curves the device
        hands back (status 1) are recomputed on the host. Returns per-curve lists
        (caller order) of power, power_raw, SDE and SDE_raw."""
        import torch
        from . import spectra as SP
        bp = self.bp
        npd = bp.total_periods
        if not hasattr(self, "_spec"):
            dev = self.dev
            lib = _ffi.load()
            ctx = _ctx(dev)
            self._spec = dict(
                pr=torch.empty(npd, dtype=torch.float64, device=dev),
                pw=torch.empty(npd, dtype=torch.float64, device=dev),
                sde_raw=torch.empty(bp.n_curves, dtype=torch.float64, device=dev),
                sde=torch.empty(bp.n_curves, dtype=torch.float64, device=dev),
                status=torch.empty(bp.n_curves, dtype=torch.int32, device=dev),
                ws=torch.empty(int(lib.gtls_spectra_workspace_bytes(ctx.handle, npd)),
                               dtype=torch.uint8, device=dev))
        sp = self._spec
        lib = _ffi.load()
        ctx = _ctx(self.dev)
        code = lib.gtls_spectra_batched(
            ctx.handle, bp.n_curves, _ptr(self.arrays["per_off"]), _ptr(chi2), npd,
            SP.median_kernel(oversampling_factor), _ptr(sp["ws"]), sp["ws"].numel(),
            _ptr(sp["pr"]), _ptr(sp["pw"]), _ptr(sp["sde_raw"]), _ptr(sp["sde"]),
            _ptr(sp["status"]))
        _ffi.check(ctx.handle, code)
        if not sync:
            return None
        inv = np.argsort(bp.order)
        status = sp["status"].cpu().numpy()[inv]
        power = bp.unpack(sp["pw"].cpu().numpy())
        power_raw = bp.unpack(sp["pr"].cpu().numpy())
        sde = sp["sde"].cpu().numpy()[inv].astype(object)
        sde_raw = sp["sde_raw"].cpu().numpy()[inv].astype(object)
        if np.any(status != 0):
            c_all = bp.unpack(chi2.cpu().numpy())
            for i in np.nonzero(status != 0)[0]:
                _, pr, pw, sr, sd = SP.spectra(SP.mask_outliers(c_all[i]), oversampling_factor)
                power[i], power_raw[i], sde[i], sde_raw[i] = pw, pr, sd, sr
        return dict(power=power, power_raw=power_raw, SDE=sde, SDE_raw=sde_raw,
                    device=status == 0)

    def capture(self, mode="fp32c", skip_point=8, depth_min=C.TRANSIT_DEPTH_MIN, ws=None,
                timing=False):
        """A HIP graph (torch.cuda.CUDAGraph) of search(): replay() re-runs every size
        group's search call into chi2 / best_k / best_s with no host work.

        The graph gets a workspace of its own, allocated before the capture and kept alive
        with the graph (attribute `workspace`), so no later search or capture can free or
        share the memory its kernels write; `ws` passes one instead (the caller then keeps
        it alive and must not replay graphs that share it concurrently). The capture runs
        with a library context of its own (attribute `gtls_ctx`), given the tuning last
        applied (tune.set_tuning) to the calling stream's context; with `timing`, its
        per-chunk kernel events are captured too and graph_kernel_ms(g) reads the figures
        of the last replay (wall-clock stamps written by one-thread kernels between the
        chunks' kernels, gtls_set_graph_timing). Every kernel of the call is captured (no
        memset nodes)."""
        import torch
        from . import tune
        setting = tune.applied(self.dev.index)
        s = torch.cuda.Stream(self.dev)
        s.wait_stream(torch.cuda.current_stream(self.dev))
        ctx = _ffi.Context(self.dev.index, s.cuda_stream)
        tune.apply_to(ctx.handle, setting)
        if ws is None:
            ws = torch.empty(self.workspace_bytes(mode, skip_point, depth_min, ctx=ctx),
                             dtype=torch.uint8, device=self.dev)
        stamps = None
        if timing:
            stamps = torch.zeros(4 * self.MAX_TIMED_CHUNKS, dtype=torch.int64, device=self.dev)
            _ffi.check(ctx.handle, _ffi.load().gtls_set_graph_timing(
                ctx.handle, _ptr(stamps), self.MAX_TIMED_CHUNKS))
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.search(mode, skip_point, depth_min, ws=ws, ctx=ctx)
        g.workspace = ws
        g.gtls_ctx = ctx
        g.stamps = stamps
        return g

    def workspace_bytes(self, mode="fp32c", skip_point=8, depth_min=C.TRANSIT_DEPTH_MIN,
                        ctx=None):
        """Workspace bytes of search(): the largest need over the size groups."""
        m = MODES[mode]
        ctx = ctx or _ctx(self.dev)
        lib = _ffi.load()
        need = 0
        for g in self.groups:
            b = self.batch_struct(skip_point, depth_min, g)
            nb = lib.gtls_search_workspace_bytes(ctx.handle, m, ct.byref(b))
            if nb < 0:
                _ffi.check(ctx.handle, int(nb))
            need = max(need, int(nb))
        return need

    def search(self, mode="fp32c", skip_point=8, depth_min=C.TRANSIT_DEPTH_MIN, ws=None,
               ctx=None):
        """Enqueue the batched search on torch's current stream (no synchronisation): one
        gtls_search_batched call per size group, each writing its slice of the outputs.
        `ws` (a uint8 device tensor of at least workspace_bytes()) is the workspace; by
        default the one shared by searches on the current stream. `ctx`: the library
        context (default: the one of the current stream)."""
        m = MODES[mode]
        ctx = ctx or _ctx(self.dev)
        lib = _ffi.load()
        structs = [self.batch_struct(skip_point, depth_min, g) for g in self.groups]
        need = 0
        for b in structs:
            nb = lib.gtls_search_workspace_bytes(ctx.handle, m, ct.byref(b))
            if nb < 0:
                _ffi.check(ctx.handle, int(nb))
            need = max(need, int(nb))
        if ws is None:
            ws = workspace(self.dev, need)
        elif ws.numel() < need:
            raise ValueError("workspace of %d bytes, the search needs %d" % (ws.numel(), need))
        for g, b in zip(self.groups, structs):
            if b.total_periods == 0:
                continue
            code = lib.gtls_search_batched(
                ctx.handle, m, ct.byref(b), _ptr(ws), ws.numel(),
                ct.c_void_p(self.chi2.data_ptr() + 8 * g["p0"]),
                ct.c_void_p(self.best_k.data_ptr() + 4 * g["p0"]),
                ct.c_void_p(self.best_s.data_ptr() + 4 * g["p0"]))
            _ffi.check(ctx.handle, code)
        return self.chi2, self.best_k, self.best_s


def graph_kernel_ms(g):
    """(prep, windows, reduce) milliseconds of the last replay of a graph captured with
    DeviceBatch.capture(timing=True), and its number of period chunks."""
    ms = (ct.c_double * 3)()
    n = _ffi.load().gtls_kernel_ms(g.gtls_ctx.handle, ms)
    if n < 0:
        _ffi.check(g.gtls_ctx.handle, int(n))
    return (ms[0], ms[1], ms[2]), int(n)
