"""ctypes wrapper of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, as
the checker of the device search and as its CPU baseline. The engine (gtls_amd) never
imports it.

Pins of each function (tests/test_oracle.py):
  fold_sort      -- numpy argsort(kind="stable") of x - trunc(x), duplicate-key cases
  period_table   -- direct-sum brute force of the TLS statistic (no prefix sums) on small
                    inputs; closed forms: constant flux -> chi2 = N (every gate closed),
                    box template -> the least-squares box step, skip subset ordering
  search         -- equals period_table's minimum per period; injected-transit recovery
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libtls_oracle.so")
SKIP_NONE = 0x7F800000

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build as _build
            _build.build(verbose=False)
        if not os.path.exists(LIB):
            raise ImportError("oracle not built: run `python oracle/build.py`")
        lib = ct.CDLL(LIB)
        d = ct.POINTER(ct.c_double)
        i32 = ct.POINTER(ct.c_int32)
        i64 = ct.POINTER(ct.c_int64)
        lib.tls_oracle_fold_sort.argtypes = [d, ct.c_int64, ct.c_double, i32]
        lib.tls_oracle_fold_sort.restype = None
        lib.tls_oracle_search.argtypes = [d, d, d, ct.c_int64, d, ct.c_int64, i32, ct.c_int32,
                                          ct.c_int32, d, i64, d, i32, i32, ct.c_int64,
                                          ct.c_double, ct.c_int32, d, i32, i32]
        lib.tls_oracle_search.restype = ct.c_int
        lib.tls_oracle_period_table.argtypes = [d, d, d, ct.c_int64, ct.c_double, i32,
                                                ct.c_int32, d, i64, d, ct.c_int32, ct.c_int32,
                                                ct.c_int64, ct.c_double, d, d, i32, i32]
        lib.tls_oracle_period_table.restype = ct.c_int
        _lib = lib
    return _lib


def _p(a, ctype):
    return a.ctypes.data_as(ct.POINTER(ctype))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def fold_sort(t, period):
    t = _f64(t)
    out = np.empty(t.size, dtype=np.int32)
    load().tls_oracle_fold_sort(_p(t, ct.c_double), t.size, float(period), _p(out, ct.c_int32))
    return out


def _table_args(widths, S, offsets, overshoot):
    widths = _i32(widths)
    S = _f64(S) if len(S) else np.ones(1)
    offsets = _i64(offsets[: len(widths)])
    overshoot = _f64(overshoot)
    wmax = int(widths.max())
    wmax += wmax % 2
    return widths, S, offsets, overshoot, wmax


def search(t, y, dy, periods, widths, S, offsets, overshoot, k_lo, k_hi, skip_point=8,
           depth_min=1e-5, nthreads=0):
    """chi2, best width index, best position for every period (float64 CPU)."""
    t, y, dy, periods = _f64(t), _f64(y), _f64(dy), _f64(periods)
    widths, S, offsets, overshoot, wmax = _table_args(widths, S, offsets, overshoot)
    k_lo, k_hi = _i32(k_lo), _i32(k_hi)
    n, npd = t.size, periods.size
    chi2 = np.empty(npd)
    bk = np.empty(npd, dtype=np.int32)
    bs = np.empty(npd, dtype=np.int32)
    d, i32, i64 = ct.c_double, ct.c_int32, ct.c_int64
    rc = load().tls_oracle_search(
        _p(t, d), _p(y, d), _p(dy, d), n, _p(periods, d), npd, _p(widths, i32),
        widths.size, wmax, _p(S, d), _p(offsets, i64), _p(overshoot, d), _p(k_lo, i32),
        _p(k_hi, i32), int(skip_point), float(depth_min), int(nthreads), _p(chi2, d),
        _p(bk, i32), _p(bs, i32))
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return chi2, bk, bs


def period_table(t, y, dy, period, widths, S, offsets, overshoot, k_lo, k_hi, skip_point=8,
                 depth_min=1e-5):
    """Full statistic table [k_hi - k_lo + 1, n] of one period and its argmin."""
    t, y, dy = _f64(t), _f64(y), _f64(dy)
    widths, S, offsets, overshoot, wmax = _table_args(widths, S, offsets, overshoot)
    n = t.size
    rows = max(0, int(k_hi) - int(k_lo) + 1)
    table = np.empty((rows, n))
    chi2 = np.empty(1)
    bk = np.empty(1, dtype=np.int32)
    bs = np.empty(1, dtype=np.int32)
    d, i32, i64 = ct.c_double, ct.c_int32, ct.c_int64
    rc = load().tls_oracle_period_table(
        _p(t, d), _p(y, d), _p(dy, d), n, float(period), _p(widths, i32), wmax, _p(S, d),
        _p(offsets, i64), _p(overshoot, d), int(k_lo), int(k_hi), int(skip_point),
        float(depth_min), _p(table, d), _p(chi2, d), _p(bk, i32), _p(bs, i32))
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return table, float(chi2[0]), int(bk[0]), int(bs[0])


def search_plan(plan, skip_point=8, depth_min=1e-5, nthreads=0, periods=None, k_lo=None,
                k_hi=None):
    """Oracle search of a gtls_amd CurvePlan (same inputs the device search gets)."""
    tb = plan.table
    periods = plan.periods if periods is None else periods
    k_lo = plan.k_lo if k_lo is None else k_lo
    k_hi = plan.k_hi if k_hi is None else k_hi
    return search(plan.t, plan.y, plan.dy, periods, tb.widths, tb.S, tb.offsets, tb.overshoot,
                  k_lo, k_hi, skip_point=skip_point, depth_min=depth_min, nthreads=nthreads)
