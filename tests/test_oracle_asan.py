"""Host AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU oracle.

`make -C oracle asan` builds oracle/_build/asan_check with g++
-fsanitize=address,undefined (host code only; the oracle has no GPU code):
every table accessor and operator of oracle/isg_oracle.cpp at 32x32 and 64x64
pixels, plane waves and a finite-radius circle, BiCGSTAB forward and adjoint
and two DBIM iterations.  Any out-of-bounds access, use after free, leak or
undefined arithmetic aborts the run.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_clean_under_asan_ubsan():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", OMP_NUM_THREADS="4")
    p = subprocess.run([os.path.join(ROOT, "oracle", "_build", "asan_check")], capture_output=True, text=True,
                       timeout=600, env=env)
    assert p.returncode == 0, p.stderr[-3000:]
    assert "ASAN CHECK OK" in p.stdout
    assert "runtime error" not in p.stderr
