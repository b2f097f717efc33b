"""Operator set-up on the device (setup_mlfma.cpp:453-631) against the host
set-up.

Device set-up (isg_config.setup = 1, the default) computes the near blocks,
the P2M/L2P patterns, the shifts and the translation operators with kernels;
every other GPU parity test already runs on those tables against the CPU
oracle (FP64 products within 1e-12 at 64^2 .. 512^2).  Here the same operator
is built both ways at 64^2, 256^2 and 1024^2 (levels 4, 6, 8; the 1024^2
translation sums reach order 259 at arguments up to 1365) and the FP64
products must agree to 1e-12, the FP32 operator's to 1e-6 (its tables are the
same FP64 values rounded; a last-bit difference of an FP64 entry can flip a
rounding).  The set-up times of both are printed.
"""
import time

import numpy as np
import pytest

from conftest import rand_c, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("level", [4, 6, 8])
def test_device_setup_matches_host_setup(level):
    import isg
    print("device vs host operator set-up, level", level, flush=True)
    t = {}
    s = {}
    for setup in (0, 1):
        t0 = time.perf_counter()
        s[setup] = isg.Solver(level=level, numrx=16, max_rhs=2, precision=1, setup=setup, tuning=False)
        t[setup] = time.perf_counter() - t0
    print("context creation: host set-up %.3f s, device set-up %.3f s" % (t[0], t[1]), flush=True)
    X = rand_c(np.random.default_rng(level), (2, s[0].N))
    y0 = s[0].mlfma(X[0])
    y1 = s[1].mlfma(X[0])
    e64 = relerr(y1, y0)
    f0 = s[0].mlfma_many(X)
    f1 = s[1].mlfma_many(X)
    e32 = max(relerr(f1[k], f0[k]) for k in range(2))
    print("FP64 product difference %.2e, FP32 %.2e" % (e64, e32), flush=True)
    assert e64 < 1e-12
    assert e32 < 1e-6
    for v in s.values():
        v.close()
