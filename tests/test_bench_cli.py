"""bench.py's command line on the CPU: the multi-GPU refusal and the argument checks exit
before anything touches a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=300, env=e)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 1" in r.stderr
    assert r.stdout == ""          # no JSON line


def test_world_size_above_one_is_refused():
    r = _run(env={"WORLD_SIZE": "4"})
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_curves_must_fill_whole_batches():
    r = _run("--curves", "100", "--batch", "64")
    assert r.returncode == 2 and "multiple of --batch" in r.stderr


def test_single_curve_workloads_have_their_stated_shapes():
    """The single-curve legs search the stated workloads: the 19,728-point 2-min curve, the
    4-year 29.4-min curve (~1.9e5 periods up to 730 d) and the reference tutorial's shape
    (41,816 points, 2,868 periods of 0.601-15.386 d on the default grid)."""
    import warnings
    sys.path.insert(0, ROOT)
    import bench
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        got = {sc["key"]: sc["plan"] for sc in bench.single_curve_plans()}
    assert got["C1"].n == 19728
    assert got["C2"].n > 65000 and len(got["C2"].periods) > 150000
    tut = got["tutorial_shape"]
    assert tut.n == 41816 and len(tut.periods) == 2868
    assert round(float(tut.periods.min()), 3) == 0.601
    assert round(float(tut.periods.max()), 3) == 15.386
