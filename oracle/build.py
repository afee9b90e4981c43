"""Builds the FP64 CPU oracle (test infrastructure) into oracle/libtls_oracle.so.

    python oracle/build.py

gcc -O2 -fopenmp, no -march flags: plain x86-64 IEEE float64 without FMA contraction.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libtls_oracle.so")
SRC = os.path.join(HERE, "tls_oracle.c")


def build(verbose=True):
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
           "-Wall", "-o", tmp, SRC, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("gcc failed building the oracle")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
