"""Autotuning of the float32 batched search on the device it runs on.

The tunables are those of gtls_set_tuning: the stride below which a width takes the
Toeplitz matrix-core tiles instead of the diagonal-sum tiles, the buffer budget of one
period chunk (which sets how many kernel tails a call pays), the resident workgroups per
CU of the persistent search kernels (waves per SIMD: one or two), and the stride from which
the diagonal-sum tiles of a common weight load their sample operands in the 8-row x 128-byte
layout (QUAD_Q; the largest value keeps every stride on the direct layout). The K group of
the tiles is fixed at 16 samples by the matrix instruction (16x16x32 over sample pairs).
`autotune` times search calls of a sample batch (curves spread over the batch's sizes) with
the library's per-kernel HIP events, one axis at a time (workgroups per CU, then the chunk
budget, then the Toeplitz bound, then the load layout threshold, each axis over all its
values with the others at their best so far), and applies the fastest setting to the
context of the device's current stream; later searches from this process on that stream
use it. Every setting gives the same results bit for bit except the Toeplitz bound (another
tile form: float32 sums in another order).
"""
from __future__ import annotations

import numpy as np

from . import _ffi
from . import constants as C
from .search import _ctx, _device, kernel_ms, set_kernel_timing

TOEP_Q = (8, 12, 16, 24)
CHUNK_MB = (8 << 10, 64 << 10)
WG_PER_CU = (1, 2)
QUAD_Q = (64, 96, 128, 1 << 20)
DEFAULT = (8, 64 << 10, 0, 96)

# the last setting applied through set_tuning, per (device, stream) context
_APPLIED = {}


def set_tuning(device=0, toep_q=8, chunk_mb=0, wg_per_cu=0, quad_q=0):
    """Apply a setting to the context of `device`'s current torch stream (quad_q 0: the
    library's default threshold)."""
    dev = _device(device)
    ctx = _ctx(dev)
    setting = (int(toep_q), int(chunk_mb), int(wg_per_cu), int(quad_q))
    _ffi.check(ctx.handle, _ffi.load().gtls_set_tuning(ctx.handle, *setting))
    _APPLIED[(ctx.device, ctx.stream)] = setting


def applied(device=0):
    """The setting last applied to the context of `device`'s current torch stream, or None
    (the library's defaults)."""
    dev = _device(device)
    ctx = _ctx(dev)
    return _APPLIED.get((ctx.device, ctx.stream))


def apply_to(handle, setting):
    """Apply a setting (from applied()) to another library context."""
    if setting is not None:
        _ffi.check(handle, _ffi.load().gtls_set_tuning(handle, *setting))


def sample_batch(db, n=32):
    """A DeviceBatch of up to n curves of db's plan, spread over its curve sizes."""
    from .batch import BatchPlan, DeviceBatch
    bp = db.bp
    order = np.argsort([p.n for p in bp.plans], kind="stable")
    pick = order[np.linspace(0, len(order) - 1, min(n, len(order))).astype(int)]
    return DeviceBatch(BatchPlan([bp.plans[i] for i in sorted(set(pick.tolist()))]), db.dev.index)


def autotune(db, skip_point=8, depth_min=C.TRANSIT_DEPTH_MIN, n_sample=32, repeats=1):
    """Coordinate search over (toep_q, chunk_mb, wg_per_cu, quad_q): after one untimed call,
    each axis in turn (WG_PER_CU, CHUNK_MB, TOEP_Q, QUAD_Q) is timed over all of its values
    with the others at their best so far (the best of `repeats` calls per setting), and the
    fastest setting seen is applied. Returns (best setting, {setting: milliseconds})."""
    import torch
    sb = sample_batch(db, n_sample)
    dev = db.dev.index
    times = {}

    def timed(setting):
        if setting not in times:
            set_tuning(dev, *setting)
            best = np.inf
            for _ in range(repeats):
                sb.search("fp32c", skip_point, depth_min)
                torch.cuda.synchronize(db.dev)
                (pm, wm, rm), _ = kernel_ms(dev)
                best = min(best, pm + wm + rm)
            times[setting] = best
        return times[setting]

    cur = (TOEP_Q[0], CHUNK_MB[-1], WG_PER_CU[-1], DEFAULT[3])
    set_kernel_timing(dev, True)
    try:
        set_tuning(dev, *cur)
        sb.search("fp32c", skip_point, depth_min)   # kernels loaded, workspace allocated
        torch.cuda.synchronize(db.dev)
        kernel_ms(dev)
        for axis, values in ((2, WG_PER_CU), (1, CHUNK_MB), (0, TOEP_Q), (3, QUAD_Q)):
            cands = [cur[:axis] + (v,) + cur[axis + 1:] for v in values]
            cur = min(cands, key=timed)
    finally:
        set_kernel_timing(dev, False)
    choice = min(times, key=times.get)
    set_tuning(dev, *choice)
    return choice, times
