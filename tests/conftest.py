"""Shared test set-up: import paths, the `gpu` marker, oracle fixtures."""
import os
import sys

import numpy as np
import pytest

# torch (device buffers for the *_dev entry points in a few GPU tests) must
# load its HIP runtime before libisg.so starts HIP in this process.
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT =os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "inverse-solver-gpu_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libisg.so")


def rand_c(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def relerr(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


_ORACLES = {}


def oracle_for(level, numrx=16, **kw):
    """Cached CPU oracle (test infrastructure) for one configuration."""
    from oracle.oracle import Oracle
    key = (level, numrx, tuple(sorted(kw.items())))
    if key not in _ORACLES:
        _ORACLES[key] = Oracle(level, numrx=numrx, **kw)
    return _ORACLES[key]
