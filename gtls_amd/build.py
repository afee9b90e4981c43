"""Builds the gfx950 shared library libgtls_amd.so in the package directory.

    python -m gtls_amd.build

hipcc cross-compiles for gfx950 on a machine without a GPU. Packed float32 instructions
(v_pk_add/mul/fma_f32) are switched off for the device code (-packed-fp32-ops): some of
their forms have given wrong results on these GPUs without any error, so the compiler
must not form them from ordinary float code.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libgtls_amd.so")
SOURCES = [os.path.join(PKG, "csrc", "gtls_abi.hip")]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def command(out=LIB):
    return [
        hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "csrc"),
        "-o", out, *SOURCES,
    ]


def build(verbose=True):
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libgtls_amd.so")
    # hipcc's host pass does not know the packed-fp32 feature flag; that warning is benign
    noise = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if verbose and noise:
        sys.stderr.write("\n".join(noise) + "\n")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
