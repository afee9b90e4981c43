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
