"""The fold's quotient (gtls_amd/csrc/fold_math.h) equals IEEE float64 division bit for bit.

The fold of src/gputls/GPUFun.cu:28-39 divides every time by the period; the engine's
windowed sort forms the same correctly rounded quotient from the period's reciprocal with
two correction steps (Markstein's theorem for the last). The check compiles that header for
the host and compares it with the division on random times and periods, on quotients a
tiny fraction of an ulp from a rounding midpoint, and on a 2-minute cadence against a fine
period grid; on zeros, infinite and NaN times, overflowing ratios and periods at the
range's ends it compares the fold phases x - trunc(x), wherever the form applies (the
small-time test, checked on its boundaries, sends the rest to the dividing form). It also
counts the misroundings of the plain product t * RN(1/P), which must be nonzero, so that
the sample holds inputs on which a dropped correction step would show.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "gtls_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fold_quotient_equals_division(tmp_path):
    exe = str(tmp_path / "fold_quotient_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC,
                    os.path.join(HERE, "fold_quotient_check.cc"), "-o", exe, "-lquadmath"],
                   check=True)
    out = subprocess.run([exe, "4000000"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    counts = dict(zip(out.stdout.split()[0::2], map(int, out.stdout.split()[1::2])))
    assert counts["random"] == counts["near_midpoint"] == counts["cadence"] == 0
    assert counts["special"] == 0 and counts["naive_product_wrong"] > 0
