"""bench.py's argument handling (CPU): a request for more than one GPU and an
unknown path or a non-positive --cycle are refused with exit status 2
before any device work."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_refused():
    p = run("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_unknown_path_refused():
    for paths in ("fast,turbo", "first-order", ","):
        p = run("--paths", paths)
        assert p.returncode == 2, paths
        assert "--paths" in p.stderr


def test_bad_cycle_refused():
    for cycle in ("0", "-2"):
        p = run("--cycle", cycle)
        assert p.returncode == 2, cycle
        assert "--cycle" in p.stderr
