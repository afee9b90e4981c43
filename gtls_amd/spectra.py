"""Signal residue, SDE and the detrended power spectrum of a chi2-vs-period vector.

Restates src/gputls/stats.py:95-122 with the running median of src/gputls/helpers.py:91-106
and the outlier mask of src/gputls/core.py:890-896, in float64 on the host. The
arithmetic steps are kept in the reference's order so that SDE values agree to rounding.
"""
from __future__ import annotations

import numpy as np
import numpy.ma as ma

from . import constants as C


def running_median(data, kernel):
    """Sliding median of odd width `kernel`, padded with the first/last value to the
    input length (src/gputls/helpers.py:91-106)."""
    data = ma.asarray(data)
    n = len(data)
    windows = np.lib.stride_tricks.sliding_window_view(np.arange(n), kernel)
    med = ma.median(data[windows], axis=1)
    front = int((n - len(med)) * 0.5)
    back = (n - len(med)) - front
    med = np.append(np.full(front, med[0]), med)
    med = np.append(med, np.full(back, med[-1]))
    return med


def median_kernel(oversampling_factor):
    k = oversampling_factor * C.SDE_MEDIAN_KERNEL_SIZE
    if k % 2 == 0:
        k = k + 1
    return k


def spectra(chi2, oversampling_factor):
    """(SR, power_raw, power, SDE_raw, SDE) of a chi2 vector (src/gputls/stats.py:95-122)."""
    SR = np.min(chi2) / chi2
    SDE_raw = (1 - np.mean(SR)) / np.std(SR)
    power_raw = SR - np.mean(SR)
    power_raw = power_raw * (SDE_raw / np.max(power_raw))
    kernel = median_kernel(oversampling_factor)
    if len(power_raw) > 2 * kernel:
        power = power_raw - running_median(power_raw, kernel)
        power = power - np.mean(power)
        SDE = np.max(power / np.std(power))
        power = power * (SDE / np.max(power))
    else:
        power = power_raw
        SDE = SDE_raw
    return SR, power_raw, power, SDE_raw, SDE


def mask_outliers(raw_chi2):
    """chi2 values above 100x the median are masked (src/gputls/core.py:890-896)."""
    med = np.median(raw_chi2)
    return ma.array(raw_chi2, mask=raw_chi2 > (100 * med))
