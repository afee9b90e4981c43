"""Seeded synthetic light curves of the benchmark workloads.

Every curve is flux = 1 + noise with an optional injected limb-darkened transit computed
by the engine's own occultation model (template.occultation_flux). Times start at
t = 1000 d so that the fold sees realistic magnitudes of t / P.

  single_2min():  27.4 d at 2-min cadence with a 1-d gap in the middle (19,728 points),
                  white noise 8e-4, P = 3.71 d, depth 1.2e-3, T14 = 2.4 h, u = (0.45, 0.20)
  long_30min():   4 years at 29.4-min cadence with 17 gaps of 1-5 d (~65,000 points), AR(1)
                  red noise (phi 0.7, sigma 3e-4) plus white 2e-4, P = 45.3 d, depth 5e-4,
                  T14 = 6.1 h
  kepler_sc_month(): one month at 58.85-s cadence (41,816 points, 30.772 d, one gap):
                  the shape of the reference tutorial's Kepler search (2,868 periods)
  batch():        n curves, N log-uniform in [2000, 70000], cadence 2 or 30 min, noise
                  U(1e-4, 2e-3), half with a transit (P U(0.8, 60) d, depth U(2e-4, 5e-3),
                  T14 of a central transit across a solar-density star)
"""
from __future__ import annotations

import numpy as np

from . import constants as C
from .template import occultation_flux

T_START = 1000.0
LD = (0.45, 0.20)


def transit_flux(t, period, t0, t14, depth, u=LD):
    """Central (b = 0) quadratic limb-darkened transit of the given depth and duration on a
    circular orbit: the radius ratio gives the centre deficit `depth`, the semi-major axis
    the first-to-fourth contact duration `t14`."""
    u1, u2 = u
    p = np.sqrt(depth * (1 - u1 / 3 - u2 / 6))
    a = (1 + p) / np.sin(np.pi * t14 / period)
    phase = ((t - t0) / period + 0.5) % 1.0 - 0.5
    z = a * np.abs(np.sin(2 * np.pi * phase))
    flux = np.ones_like(t)
    m = (np.abs(phase) < 0.25) & (z < 1 + p)
    if np.any(m):
        flux[m] = occultation_flux(z[m], p, u, "quadratic")
    return flux


def single_2min(seed_noise=1, seed_t0=2):
    cad = 1.0 / 720.0
    half = int(round(13.7 * 720))
    t = np.concatenate([T_START + np.arange(half) * cad,
                        T_START + 14.7 + np.arange(half) * cad])
    noise = np.random.default_rng(seed_noise).normal(0.0, 8e-4, t.size)
    P = 3.71
    t0 = T_START + np.random.default_rng(seed_t0).uniform(0, P)
    y = 1.0 + noise
    y *= transit_flux(t, P, t0, 2.4 / 24, 1.2e-3)
    return t, y, None, dict(period=P, t0=t0, t14=0.1, depth=1.2e-3)


def long_30min(seed=3):
    rng = np.random.default_rng(seed)
    cad = 29.4 / 1440.0
    span = 1461.0
    t = T_START + np.arange(int(span / cad)) * cad
    keep = np.ones(t.size, dtype=bool)
    starts = np.sort(rng.uniform(0, span - 5, 17))
    for s in starts:
        keep &= ~((t >= T_START + s) & (t < T_START + s + rng.uniform(1, 5)))
    t = t[keep]
    phi, sig_red, sig_white = 0.7, 3e-4, 2e-4
    eps = rng.normal(0.0, sig_red * np.sqrt(1 - phi ** 2), t.size)
    red = np.empty(t.size)
    red[0] = rng.normal(0.0, sig_red)
    for i in range(1, t.size):
        red[i] = phi * red[i - 1] + eps[i]
    y = 1.0 + red + rng.normal(0.0, sig_white, t.size)
    P, t14, depth = 45.3, 6.1 / 24, 5e-4
    t0 = T_START + rng.uniform(0, P)
    y *= transit_flux(t, P, t0, t14, depth)
    return t, y, None, dict(period=P, t0=t0, t14=t14, depth=depth)


def kepler_sc_month(seed=5):
    """The shape of the reference tutorial's Kepler curve (tutorial/1-keplerTest.ipynb:
    41,816 points searched over 2,868 periods of 0.601-15.386 d with the default grid): one
    month at Kepler's 58.85-s short cadence (30.772 d) with one 2.3-d gap, white noise
    3e-4, a transit at P = 2.47 d of depth 1e-3."""
    rng = np.random.default_rng(seed)
    cad = 58.85 / 86400.0
    n_full = int(round(30.772 / cad)) + 1
    t = T_START + np.arange(n_full) * cad
    keep = np.ones(n_full, dtype=bool)
    keep[20000:20000 + n_full - 41816] = False
    t = t[keep]
    P, depth = 2.47, 1e-3
    t0 = T_START + rng.uniform(0, P)
    y = (1.0 + rng.normal(0.0, 3e-4, t.size)) * transit_flux(t, P, t0, solar_t14(P), depth)
    return t, y, None, dict(period=P, t0=t0, t14=solar_t14(P), depth=depth)


def solar_t14(period):
    """Central-transit duration (days) of a small planet around a Sun-like star."""
    P = period * C.SECONDS_PER_DAY
    a = (C.G * C.M_SUN * P ** 2 / (4 * np.pi ** 2)) ** (1 / 3)
    return period / np.pi * np.arcsin(min(1.0, C.R_SUN / a))


def batch(n_curves=4096, seed=4, n_min=2000, n_max=70000):
    """List of (t, y, None, truth) for the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_curves):
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        cad = (2.0 if rng.random() < 0.5 else 30.0) / 1440.0
        sigma = rng.uniform(1e-4, 2e-3)
        t = T_START + np.arange(n) * cad
        y = 1.0 + rng.normal(0.0, sigma, n)
        truth = None
        if rng.random() < 0.5:
            P = rng.uniform(0.8, 60.0)
            depth = rng.uniform(2e-4, 5e-3)
            t14 = solar_t14(P)
            t0 = T_START + rng.uniform(0, P)
            y *= transit_flux(t, P, t0, t14, depth)
            truth = dict(period=P, t0=t0, t14=t14, depth=depth)
        out.append((t, y, None, truth))
    return out
