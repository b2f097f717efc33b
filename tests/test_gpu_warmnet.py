"""Warm-start network kernels on the GPU against the CPU oracle.

Workload: 2 images of 64x64 (and one 128x64, and one 48x80 whose levels end
in partial 8x8 tiles) pixels with 8 input channels, random inputs, seeded
He-normal weights.  The oracle reproduces the GPU's rounding points (both
operands of every convolution rounded to BF16); the remaining differences are
FP32 versus float64 accumulation, and the BF16 roundings they occasionally
flip downstream (one BF16 ulp is 3.9e-3 relative; a flipped rounding in an
early layer moves the output by ~1e-3 of its norm).  Bars: output 1e-2
relative (norm) -- a wrong tile, halo or channel mapping moves whole 8x8
tiles and fails it by orders of magnitude --, loss 1e-4, per-layer weight
gradient 1e-2, Adam step 1e-5 given the GPU's own gradient.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

SCALES = np.array([3.0, 3.0, 1.0, 1.0, 1.0, 1.0, 2.0, 2.0, 0.05], np.float32)


@pytest.fixture(scope="module")
def net():
    import isg
    n = isg.WarmNet(max_batch=2, max_pixels=128 * 64, seed=11)
    n.scales = SCALES
    return n


def data(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape + (8,)).astype(np.float32)
    t = (0.05 * rng.standard_normal(shape + (2,))).astype(np.float32)
    return x, t


@pytest.mark.parametrize("shape", [(2, 64, 64), (1, 128, 64), (1, 48, 80)])
def test_forward_matches_oracle(net, shape):
    from oracle import warmnet_oracle as wo
    print("warm-start network forward", shape, flush=True)
    x, _ = data(shape, 1)
    y = net.forward(x)
    ref = wo.forward(net.params, x, SCALES, rnd=True)["y"]
    err = relerr(y, ref)
    print("forward relative error %.3e" % err)
    assert err < 1e-2


def test_gradient_matches_oracle(net):
    from oracle import warmnet_oracle as wo
    print("warm-start network gradient", flush=True)
    x, t = data((2, 64, 64), 2)
    loss = net.loss_grad(x, t)
    g = net.grads()
    lref, gref = wo.loss_grad(net.params, x, t, SCALES, rnd=True)
    assert abs(loss - lref) < 1e-4 * lref
    offs, _ = wo.offsets()
    for li, (a, b) in enumerate(offs):
        cout = wo.LAYERS[li][2]
        e = relerr(g[a:b + cout], gref[a:b + cout])
        print("layer %d gradient relative error %.3e" % (li, e))
        assert e < 1e-2, li


def test_adam_step_and_save_load(net, tmp_path):
    import isg
    from oracle import warmnet_oracle as wo
    x, t = data((2, 64, 64), 3)
    p0 = net.params
    net.params = p0  # resets the optimiser state
    net.loss_grad(x, t)
    g = net.grads()
    net.adam(1e-3)
    p1 = net.params
    ref = wo.adam(p0.copy(), g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    np.testing.assert_allclose(p1, ref, rtol=1e-5, atol=1e-7)
    path = tmp_path / "w.bin"
    net.save(path)
    other = isg.WarmNet(max_batch=1, max_pixels=64 * 64, seed=99, path=path)
    np.testing.assert_array_equal(other.params, p1)
    np.testing.assert_array_equal(other.scales, SCALES)
    np.testing.assert_allclose(other.forward(x[:1]), net.forward(x[:1]), rtol=0, atol=0)
    net.params = p0
