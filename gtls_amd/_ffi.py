"""ctypes binding of libgtls_amd.so (the C ABI of include/gtls_amd.h).

The library is the only compute path of the engine: if it is missing or cannot be loaded
every search call raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os
import threading

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgtls_amd.so")

GTLS_OK = 0
GTLS_E_SHAPE = -1001
GTLS_E_WORKSPACE = -1002
GTLS_E_MODE = -1003
GTLS_E_DEVICE = -1004
MODE_FP64 = 0
MODE_FP32C = 1
SKIP_NONE = 0x7F800000

# every symbol include/gtls_amd.h declares
EXPORTED = (
    "gtls_create", "gtls_destroy", "gtls_set_stream", "gtls_last_error", "gtls_search_slots",
    "gtls_search_workspace_bytes", "gtls_search_batched", "gtls_search_workspace_bytes_1",
    "gtls_search", "gtls_fold_sort_workspace_bytes", "gtls_fold_sort",
    "gtls_spectra_workspace_bytes", "gtls_spectra_batched", "gtls_spectra_batched_noalloc",
    "gtls_transit_fwd",
    "gtls_transit_bwd", "gtls_transit_chi2", "gtls_set_timing", "gtls_kernel_ms", "gtls_set_tuning",
    "gtls_set_graph_timing", "gtls_transit_tile_points", "gtls_transit_chi2_grouped",
)

P = ct.c_void_p
I32 = ct.c_int32
I64 = ct.c_int64
F64 = ct.c_double


class GtlsBatch(ct.Structure):
    _fields_ = [
        ("n_curves", I32),
        ("pt_off", P), ("t", P), ("y", P), ("dy", P),
        ("per_off", P), ("periods", P), ("k_lo", P), ("k_hi", P),
        ("wid_off", P), ("widths", P), ("overshoot", P), ("tmpl_off", P), ("tmpl", P),
        ("skip_point", I64), ("depth_min", F64), ("negative_time", I32),
        ("total_points", I64), ("total_periods", I64), ("total_widths", I64),
        ("total_tmpl", I64), ("max_points", I32), ("max_width", I32),
        ("max_period_widths", I32),
    ]


class GtlsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gtls error %d: %s" % (code, msg))
        self.code = code


_lib = None
_lock = threading.Lock()


def _declare(lib):
    sig = {
        "gtls_create": (ct.c_int, [ct.c_int, P, ct.POINTER(P)]),
        "gtls_destroy": (None, [P]),
        "gtls_set_stream": (ct.c_int, [P, P]),
        "gtls_last_error": (ct.c_char_p, [P]),
        "gtls_search_slots": (I32, [P]),
        "gtls_search_workspace_bytes": (I64, [P, I32, ct.POINTER(GtlsBatch)]),
        "gtls_search_batched": (ct.c_int, [P, I32, ct.POINTER(GtlsBatch), P, I64, P, P, P]),
        "gtls_search_workspace_bytes_1": (I64, [P, I32, I64, I64, I32, I64, I32]),
        "gtls_search": (ct.c_int, [P, I32, P, P, P, I64, P, I64, P, P, P, I32, I32, P, P, P,
                                   I64, I64, F64, I32, P, I64, P, P, P]),
        "gtls_fold_sort_workspace_bytes": (I64, [P, I64]),
        "gtls_fold_sort": (ct.c_int, [P, P, I64, P, I64, I32, P, I64, P]),
        "gtls_spectra_workspace_bytes": (I64, [P, I64]),
        "gtls_spectra_batched": (ct.c_int, [P, I32, P, P, I64, I32, P, I64, P, P, P, P, P]),
        "gtls_spectra_batched_noalloc": (ct.c_int, [P, I32, P, P, I64, I32, P, P, P, P, P]),
        "gtls_transit_fwd": (ct.c_int, [P, I32, I64, P, P, P, P, P, I64, P]),
        "gtls_transit_bwd": (ct.c_int, [P, I32, I64, P, P, P, P, P, P, P]),
        "gtls_transit_chi2": (ct.c_int, [P, I32, I64, P, P, P, P, P, P, P, P]),
        "gtls_set_timing": (ct.c_int, [P, I32]),
        "gtls_kernel_ms": (I64, [P, ct.POINTER(F64)]),
        "gtls_set_tuning": (ct.c_int, [P, I32, I64, I32, I32]),
        "gtls_set_graph_timing": (ct.c_int, [P, P, I64]),
        "gtls_transit_tile_points": (I64, []),
        "gtls_transit_chi2_grouped": (ct.c_int, [P, I32, I32, I64, I32, P, P, P, P, I64, P, P, P,
                                                 P, I64, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib_path():
    return _LIB_PATH


def load():
    """The loaded library; raises if it was not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                raise ImportError("libgtls_amd.so is not built: run `python -m gtls_amd.build`")
            lib = ct.CDLL(_LIB_PATH)
            _declare(lib)
            _lib = lib
    return _lib


def check(ctx, code):
    if code != GTLS_OK:
        msg = load().gtls_last_error(ctx)
        raise GtlsError(code, msg.decode() if msg else "")
    return code


def ptr(tensor):
    """Device pointer of a torch tensor (None -> NULL)."""
    return None if tensor is None else ct.c_void_p(tensor.data_ptr())


class Context:
    """One gtls_ctx per (device, stream)."""

    _cache = {}

    def __init__(self, device, stream_handle):
        lib = load()
        h = P()
        code = lib.gtls_create(int(device), P(stream_handle), ct.byref(h))
        if code != GTLS_OK:
            raise GtlsError(code, "gtls_create failed for device %d" % device)
        self.handle = h
        self.device = device
        self.stream = stream_handle

    @classmethod
    def get(cls, device, stream_handle):
        key = (int(device), int(stream_handle or 0))
        ctx = cls._cache.get(key)
        if ctx is None:
            ctx = cls(device, stream_handle)
            cls._cache[key] = ctx
        return ctx

    def slots(self):
        return load().gtls_search_slots(self.handle)

    @classmethod
    def close_all(cls):
        """Destroys every cached context (gtls_destroy); the caller has finished all work on
        their streams. Later calls create new ones."""
        lib = load()
        for ctx in cls._cache.values():
            lib.gtls_destroy(ctx.handle)
        cls._cache.clear()
