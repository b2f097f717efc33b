"""C ABI and host logic of libisg.so, without a GPU.

  * the library loads and exports every function include/isg.h declares;
  * defaults equal the reference's compile-time constants (mom.cpp:4-32,
    setup_mlfma.cpp:3-4);
  * without a HIP device, creating a context fails loudly (no CPU fallback);
  * the seeded phantoms match an independent splitmix64 restatement and
    their closed forms.
"""
import numpy as np
import pytest

import isg


def gpu_present():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def test_exports_every_declared_symbol():
    L = isg.load()
    names = isg.declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_defaults_are_reference_constants():
    c = isg.default_config()
    assert (c.res, c.box, c.level, c.beta, c.ninter) == (0.1, 8, 7, 5, 30)
    assert (c.toli, c.iti, c.regular) == (1e-4, 300, 0.0)
    assert c.max_rhs == 1 and c.precision == 0 and c.mixed_margin == 0.5
    assert c.graph == 1 and c.radius == 0.0 and c.products == 0 and c.solve_start == 0


@pytest.mark.skipif(gpu_present(), reason="checks the no-device error path")
def test_no_device_fails_loudly():
    with pytest.raises(isg.IsgError, match="no HIP device"):
        isg.Solver(level=4)


@pytest.mark.skipif(gpu_present(), reason="checks argument validation before device use")
def test_bad_config_rejected():
    with pytest.raises(isg.IsgError):
        isg.Solver(level=4, box=4)
    with pytest.raises(isg.IsgError):
        isg.Solver(level=4, precision=7)
    with pytest.raises(isg.IsgError):
        isg.Solver(level=4, precision=1, mixed_margin=0.0)
    with pytest.raises(isg.IsgError):
        isg.Solver(level=4, precision=0, products=1)
    with pytest.raises(isg.IsgError):
        isg.Solver(level=4, precision=1, products=-1)
    with pytest.raises(isg.IsgError):
        isg.Solver(level=4, solve_start=2)
    for beta in (0, -3, 33):  # a non-positive beta would make the sample-count rule NaN
        with pytest.raises(isg.IsgError, match="beta must be"):
            isg.Solver(level=4, beta=beta)


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & (2 ** 64 - 1)

    def next(self):
        M = 2 ** 64 - 1
        self.s = (self.s + 0x9E3779B97F4A7C15) & M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    def uniform(self, a, b):
        return a + (b - a) * ((self.next() >> 11) * 2.0 ** -53)


def test_blobs_match_restatement():
    side, res, n, seed = 48, 0.1, 3, 1
    got = isg.phantom_blobs(side, n, seed, res=res)
    D = side * res
    k, l = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    X, Y = -D / 2 + (l + 0.5) * res, D / 2 - (k + 0.5) * res
    rng = SplitMix64(seed)
    ref = np.zeros((side, side), complex)
    for _ in range(n):
        cx, cy = rng.uniform(-0.4 * D, 0.4 * D), rng.uniform(-0.4 * D, 0.4 * D)
        sg, ar, ai = rng.uniform(0.3, 1.0), rng.uniform(0.1, 0.5), rng.uniform(0.01, 0.05)
        ref += (ar + 1j * ai) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sg * sg))
    assert np.abs(got - ref).max() < 1e-12
    # closed form: imaginary/real ratio of each blob within the drawn ranges
    assert 0 < got.real.max() <= 0.5 * n


def test_ellipses_match_restatement():
    side, res, n, seed = 64, 0.1, 5, 3
    got = isg.phantom_ellipses(side, n, seed, res=res)
    D = side * res
    k, l = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    X, Y = -D / 2 + (l + 0.5) * res, D / 2 - (k + 0.5) * res
    rng = SplitMix64(seed)
    ref = np.zeros((side, side), complex)
    for _ in range(n):
        cx, cy = rng.uniform(-0.4 * D, 0.4 * D), rng.uniform(-0.4 * D, 0.4 * D)
        a, b = rng.uniform(0.02 * D, 0.08 * D), rng.uniform(0.02 * D, 0.08 * D)
        th, v = rng.uniform(0, np.pi), rng.uniform(0.05, 0.8)
        u = ((X - cx) * np.cos(th) + (Y - cy) * np.sin(th)) / a
        w = (-(X - cx) * np.sin(th) + (Y - cy) * np.cos(th)) / b
        ref[u * u + w * w <= 1] = v + 0.02j
    assert np.array_equal(got, ref)
    vals = np.unique(got[got != 0].real)
    assert np.all((vals >= 0.05) & (vals <= 0.8))
    assert np.all(got[got != 0].imag == 0.02)


def test_raster_tree_roundtrip():
    rng = np.random.default_rng(0)
    side = 32
    r2t = rng.permutation(side * side).astype(np.int32)
    img = rng.standard_normal((side, side)) + 1j * rng.standard_normal((side, side))
    assert np.array_equal(isg.tree_to_raster(isg.raster_to_tree(img, r2t), r2t, side), img)
