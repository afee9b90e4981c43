"""gtls_spectra_batched_noalloc (the spectra without a workspace, for callers that cannot
allocate) against gtls_spectra_batched, through the C ABI: bit-identical outputs on two
calls in a row, the second with fewer curves and periods than the first (the shared
device buffer keeps the first call's contents), and a clear error past its size limit."""
import ctypes as ct

import numpy as np
import pytest

from gtls_amd import _ffi

pytestmark = pytest.mark.gpu

MAX_PERIODS = 1 << 22


def chi2_batch(sizes, seed):
    """Per-curve chi2 vectors: noise around N with one dip, some curves below the
    detrending length (2 * 91 periods) and one with a large outlier (host fallback)."""
    rng = np.random.default_rng(seed)
    out = []
    for j, n in enumerate(sizes):
        c = 5000.0 * (1 + 0.01 * rng.standard_normal(n))
        c[rng.integers(n)] *= 0.8
        if j == 1:
            c[0] *= 150.0
        out.append(c)
    per_off = np.concatenate([[0], np.cumsum([len(c) for c in out])]).astype(np.int64)
    return np.concatenate(out), per_off


def run(ctx, lib, torch, chi2, per_off, noalloc, kernel=91):
    dev = torch.device("cuda", 0)
    n_curves = len(per_off) - 1
    total = int(per_off[-1])
    c = torch.as_tensor(chi2, device=dev)
    po = torch.as_tensor(per_off, device=dev)
    pr = torch.full((total,), float("nan"), dtype=torch.float64, device=dev)
    pw = torch.full((total,), float("nan"), dtype=torch.float64, device=dev)
    sr = torch.full((n_curves,), float("nan"), dtype=torch.float64, device=dev)
    sd = torch.full((n_curves,), float("nan"), dtype=torch.float64, device=dev)
    st = torch.full((n_curves,), -9, dtype=torch.int32, device=dev)
    p = _ffi.ptr
    if noalloc:
        code = lib.gtls_spectra_batched_noalloc(ctx.handle, n_curves, p(po), p(c), total, kernel,
                                                p(pr), p(pw), p(sr), p(sd), p(st))
    else:
        nb = lib.gtls_spectra_workspace_bytes(ctx.handle, total)
        ws = torch.empty(int(nb), dtype=torch.uint8, device=dev)
        code = lib.gtls_spectra_batched(ctx.handle, n_curves, p(po), p(c), total, kernel, p(ws),
                                        ws.numel(), p(pr), p(pw), p(sr), p(sd), p(st))
    _ffi.check(ctx.handle, code)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in (pr, pw, sr, sd, st)]


def test_noalloc_spectra_match_the_workspace_entry_point(gpu):
    import torch
    lib = _ffi.load()
    ctx = _ffi.Context.get(0, torch.cuda.current_stream(0).cuda_stream)
    for sizes, seed in (((3000, 800, 150, 20000, 5000), 1), ((900, 400, 60), 2)):
        chi2, per_off = chi2_batch(sizes, seed)
        want = run(ctx, lib, torch, chi2, per_off, noalloc=False)
        got = run(ctx, lib, torch, chi2, per_off, noalloc=True)
        for a, b in zip(want, got):
            assert np.array_equal(a, b, equal_nan=True)
        assert np.sum(want[4] == 0) >= 2       # detrended on the device
    # past the buffer: refused before any launch
    big = MAX_PERIODS + 1
    code = lib.gtls_spectra_batched_noalloc(ctx.handle, 1, ct.c_void_p(1), ct.c_void_p(1), big,
                                            91, ct.c_void_p(1), ct.c_void_p(1), ct.c_void_p(1),
                                            ct.c_void_p(1), ct.c_void_p(1))
    assert code == _ffi.GTLS_E_WORKSPACE
    assert b"4194304" in lib.gtls_last_error(ctx.handle)
