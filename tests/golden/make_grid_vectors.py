"""Writes tests/golden/grid_vectors.npz from the reference's own grid functions.

Runs only where the reference checkout exists (/root/reference). It imports the
reference's pure-NumPy modules src/gputls/grid.py and constants.py under a package name
without executing the package __init__ (which imports CuPy), calls period_grid,
duration_grid and T14 on a set of inputs, and stores inputs and outputs as arrays.
The reference's exact float sequence of the Ofir period grid is not fixed by any closed
form, so these vectors are what pins gtls_amd.grid bit for bit.

    python tests/golden/make_grid_vectors.py
"""
import importlib
import os
import sys
import types
import warnings

import numpy as np

REF = "/root/reference/src/gputls"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "grid_vectors.npz")

CASES = [
    # R_star, M_star, time_span, period_min, period_max, oversampling, n_transits_min
    (1.0, 1.0, 28.4, 0, float("inf"), 3, 2),
    (1.0, 1.0, 27.4, 0, float("inf"), 3, 2),
    (1.0, 1.0, 1461.0, 0.5, 700.0, 3, 2),
    (0.5, 0.6, 90.0, 0, float("inf"), 5, 3),
    (2.0, 1.3, 12.0, 1.0, 5.0, 2, 2),
    (1.0, 1.0, 3.0, 0, float("inf"), 3, 2),   # too few periods: default grid fallback
    (0.005, 0.005, 40.0, 0, float("inf"), 3, 2),  # clamped stellar parameters
]


def load_reference_grid():
    pkg = types.ModuleType("refgtls")
    pkg.__path__ = [REF]
    sys.modules["refgtls"] = pkg
    return importlib.import_module("refgtls.grid")


def main():
    grid = load_reference_grid()
    data = {}
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for i, (R, M, span, pmin, pmax, os_, ntm) in enumerate(CASES):
            p = grid.period_grid(R_star=R, M_star=M, time_span=span, period_min=pmin,
                                 period_max=pmax, oversampling_factor=os_,
                                 n_transits_min=ntm)
            data["case%d_in" % i] = np.array([R, M, span, pmin, pmax, os_, ntm], dtype=float)
            p = np.asarray(p, dtype=np.float64)
            idx = np.arange(p.size)
            if p.size > 4000:  # keep the fixture small: an even subsample plus both ends
                idx = np.unique(np.concatenate([idx[:200], idx[-200:],
                                                np.linspace(0, p.size - 1, 2000).astype(int)]))
            data["case%d_n_periods" % i] = np.array([p.size])
            data["case%d_index" % i] = idx
            data["case%d_periods" % i] = p[idx]
            d = grid.duration_grid(p, shortest=1.0 / 1000, log_step=1.1)
            data["case%d_durations" % i] = np.asarray(d, dtype=np.float64)
            d2 = grid.duration_grid(p, shortest=1.0 / 1000, log_step=1.05)
            data["case%d_durations_step105" % i] = np.asarray(d2, dtype=np.float64)
        Ps = np.array([0.3, 0.6, 1.0, 3.71, 13.7, 45.3, 700.0])
        t14 = []
        for R, M in ((1.0, 1.0), (0.13, 0.1), (3.5, 1.0)):
            for small in (False, True):
                for P in Ps:
                    t14.append([R, M, P, float(small), grid.T14(R, M, P, small=small)])
        data["t14"] = np.array(t14)
    np.savez(OUT, **data)
    print("wrote", OUT, len(data), "arrays")


if __name__ == "__main__":
    main()
