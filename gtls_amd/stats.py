"""Statistics of the best period: depth, T0, transit times, duration in days and the SNRs.

Host float64 restatement of the reference's single-period refinement
(src/gputls/core.py:1418-1500) and the helpers it calls (src/gputls/stats.py:197-488,
src/gputls/helpers.py:61-65). The inputs -- best width index and best T0 position of the
period -- come from the device search of that one period without position skipping
(src/gputls/core.py:1406-1425).
"""
from __future__ import annotations

import numpy as np


def fold_cpu(t, y, dy, period):
    """Host fold and sort (src/gputls/core.py:16-20): phase = (t mod P) / P."""
    phase = (t % period) / period
    rank = np.argsort(phase)
    return t[rank], y[rank], dy[rank]


def transit_mask(t, period, duration, T0):
    return np.abs((t - T0 + 0.5 * period) % period - 0.5 * period) < 0.5 * duration


def all_transit_times(T0, t, period):
    """Mid-transit times within t (src/gputls/stats.py:233-250)."""
    tt = [T0 + period] if T0 < min(t) else [T0]
    prev = tt[0]
    end = np.min(t) + (np.max(t) - np.min(t))
    while True:
        nxt = prev + period
        if nxt < end:
            tt.append(nxt)
            prev = nxt
        else:
            break
    return tt


def duration_days(times, period, T0, raw_duration):
    """Duration in days from the phase-sorted samples (src/gputls/stats.py:252-275)."""
    T0c = T0 + period / 2
    phases = (times - T0c) / period - np.floor((times - T0c) / period)
    ps = phases[np.argsort(phases)]
    start = int(np.abs(ps - 0.5).argmin())
    end = start + int(np.array(raw_duration) * len(times))
    end = min(end, len(ps) - 1)  # the reference indexes past the end here for P-long transits
    return (ps[end] - 0.5) * period


def intransit_stats(t, y, transit_times, duration):
    """Per-transit counts and odd/even in-transit flux (src/gputls/stats.py:351-424)."""
    odd, even = [], []
    count = np.zeros(len(transit_times))
    for i, mid in enumerate(transit_times):
        lo, hi = mid - 0.5 * duration, mid + 0.5 * duration
        if np.isnan(lo) or np.isnan(hi):
            sel = np.array([], dtype=float)
        else:
            sel = y[(t > lo) & (t < hi)]
        count[i] = sel.size
        (even if i % 2 == 0 else odd).append(sel)
    odd = np.concatenate(odd) if odd else np.array([])
    even = np.concatenate(even) if even else np.array([])
    return odd, even, count


def pink_noise(data, width):
    std = 0
    n = len(data) - width + 1
    for i in range(n):
        std += np.std(data[i:i + width])
    return std / (n * width ** 0.5)


def snr_stats(t, y, period, duration, T0, transit_times, duration_days_, per_transit_count):
    """Per-transit SNR and pink-noise SNR (src/gputls/stats.py:433-488)."""
    snr = np.zeros(len(transit_times))
    snr_pink = np.zeros(len(transit_times))
    flux_ootr = y[~transit_mask(t, period, 2 * duration, T0)]
    try:
        pink = pink_noise(flux_ootr, int(np.mean(per_transit_count)))
    except Exception:
        pink = np.nan
    std = np.std(flux_ootr) if len(flux_ootr) > 0 else np.nan
    for i, mid in enumerate(transit_times):
        lo, hi = mid - 0.5 * duration_days_, mid + 0.5 * duration_days_
        if np.isnan(lo) or np.isnan(hi):
            sel = np.array([])
        else:
            sel = y[(t > lo) & (t < hi)]
        mean_flux = np.mean(sel) if sel.size > 0 else np.nan
        try:
            snr_pink[i] = (1 - mean_flux) / pink
            if sel.size > 0 and not np.isnan(std):
                snr[i] = (1 - mean_flux) / (std / sel.size ** 0.5)
            else:
                snr[i] = 0
                snr_pink[i] = 0
        except Exception:
            snr[i] = 0
            snr_pink[i] = 0
    return snr, snr_pink


def refine_period(plan, period, width_index, t0_index):
    """Depth, T0, transit times, duration and SNRs of the best period.

    `width_index` is the curve-local index of the best width and `t0_index` the best
    position in the device's phase order (src/gputls/core.py:1418-1500).
    Returns a dict with the reference's result fields.
    """
    t, y, dy = plan.t, plan.y, plan.dy
    tb = plan.table
    n = len(t)
    width = int(tb.widths[width_index])
    raw_duration = float(tb.durations[width_index])
    overshoot = float(tb.overshoot[width_index])

    best_t, best_y, _ = fold_cpu(t, y, dy, period)
    wrap = tb.wmax_even
    patched = np.concatenate((best_y, best_y[:wrap]))
    s = int(t0_index)
    mean_in = patched[s:s + width].mean()
    depth = float((1 - mean_in) * overshoot)
    outside = np.concatenate((patched[0:s], patched[s + width:]))
    if s > n - 1:
        s -= n
    sd_out = np.std(outside)
    snr_fit = (1 - depth) * (width ** 0.5) / sd_out
    cs = np.cumsum(outside)
    slide = (cs[width:] - cs[:-width]) / width
    red = np.std(slide)

    Tx = best_t[s]
    T0 = Tx - int((Tx - min(t)) / period) * period - period
    transit_times = all_transit_times(T0, t, period)
    snr_fit_pink = (1 - depth) / ((sd_out ** 2 / width) + (red ** 2 / len(transit_times))) ** 0.5

    dur_days = duration_days(t, period, T0, raw_duration)
    T0 = T0 + dur_days / 2
    transit_times = np.asarray(transit_times) + dur_days / 2
    if T0 < min(t):
        T0 = T0 + period

    odd, even, count = intransit_stats(t, y, transit_times, dur_days)
    snr_tr, snr_pink_tr = snr_stats(t, y, period, raw_duration, T0, transit_times, dur_days,
                                    count)
    all_in = np.concatenate([odd, even])
    flux_ootr = y[~transit_mask(t, period, 2 * raw_duration, T0)]
    snr = ((1 - np.mean(all_in)) / np.std(flux_ootr)) * len(all_in) ** 0.5
    snr_pink = np.mean(snr_pink_tr) * (len(transit_times) ** 0.5)
    return dict(rawDuration=raw_duration, durationPoints=width, duration=dur_days,
                depth=depth, T0=T0, transitTimes=transit_times, snr=snr, snrPink=snr_pink,
                snrFit=snr_fit, snrFitPink=snr_fit_pink)
