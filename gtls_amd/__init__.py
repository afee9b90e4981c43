"""gtls_amd -- MI355X-native Transit Least Squares search.

    from gtls_amd import gtls
    result = gtls(t, y, dy).power()            # full search, result.period, result.SDE, ...
    periods, power = gtls(t, y).power(fast=True)

The chi2-vs-period search runs in hand-written HIP kernels for gfx950
(gtls_amd/libgtls_amd.so, C ABI in include/gtls_amd.h); the host builds grids and
templates with NumPy and moves arrays with PyTorch.
"""
from .api import gtls, gtlsResult  # noqa: F401
from .constants import VERSION as __version__  # noqa: F401
