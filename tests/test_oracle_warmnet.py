"""Pin the CPU oracle of the warm-start network (oracle/warmnet_oracle.py).

  * BF16 rounding: round-to-nearest-even on hand-picked values;
  * convolution: against a direct loop over pixels, taps and channels;
  * loss gradient: central finite differences of the exact (float64, no
    rounding) loss along random directions, one layer at a time, so a wrong
    index, transpose or sign in any layer's backward fails;
  * Adam: the first step moves every weight by lr * sign(grad).
"""
import numpy as np
import pytest

from oracle import warmnet_oracle as wo


def test_bf16_round_to_nearest_even():
    one = 1.0
    vals = np.array([one, one + 2.0 ** -9, one + 3 * 2.0 ** -9, one + 2.0 ** -8, one + 3 * 2.0 ** -8, -2.5, 3.0e-3],
                    np.float32)
    got = wo.bf16(vals)
    # the BF16 spacing at 1 is 2^-7: 1 + 2^-9 rounds down, 1 + 3*2^-9 rounds up,
    # 1 + 2^-8 and 1 + 3*2^-8 are ties and go to the even neighbour (1 and 1 + 2^-6)
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[2] == 1.0 + 2.0 ** -7 and got[3] == 1.0
    assert got[4] == 1.0 + 2.0 ** -6
    assert got[5] == -2.5
    assert abs(got[6] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8


@pytest.mark.parametrize("ks", [1, 3])
def test_conv_matches_direct_loop(ks):
    rng = np.random.default_rng(ks)
    n, H, W, cin, cout = 2, 5, 6, 3, 4
    x = rng.standard_normal((n, H, W, cin))
    w = rng.standard_normal((ks * ks, cout, cin))
    got = wo.conv(x, w, ks, rnd=False)
    r = ks // 2
    ref = np.zeros((n, H, W, cout))
    for b in range(n):
        for y in range(H):
            for xx in range(W):
                for t in range(ks * ks):
                    yy, xs = y + t // ks - r, xx + t % ks - r
                    if 0 <= yy < H and 0 <= xs < W:
                        ref[b, y, xx] += w[t] @ x[b, yy, xs]
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_space_to_depth_layout():
    """s2d puts input pixel (2Y + dy, 2X + dx) channel c at position (Y, X)
    channel 8 (2 dy + dx) + c (direct loop), and d2s inverts it."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 6, 4, 8))
    y = wo.s2d(x)
    assert y.shape == (2, 3, 2, 32)
    for n in range(2):
        for Y in range(3):
            for X in range(2):
                for dy in range(2):
                    for dx in range(2):
                        for c in range(8):
                            assert y[n, Y, X, 8 * (2 * dy + dx) + c] == x[n, 2 * Y + dy, 2 * X + dx, c]
    np.testing.assert_array_equal(wo.d2s(y, 8), x)
    z = rng.standard_normal((1, 3, 2, 16))
    out = wo.d2s(z, 2)
    for Y in range(3):
        for X in range(2):
            for s in range(4):
                for c in range(2):
                    assert out[0, 2 * Y + s // 2, 2 * X + s % 2, c] == z[0, Y, X, 2 * s + c]


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(7)
    n, H, W = 1, 32, 32
    x = rng.standard_normal((n, H, W, 8))
    t = rng.standard_normal((n, H, W, 2))
    scales = np.array([1.5, 1.5, 1.0, 1.0, 0.8, 0.8, 1.2, 1.2, 0.7])
    p = np.zeros(wo.num_params())
    offs, _ = wo.offsets()
    for (ks, cin, cout, _, _), (a, b) in zip(wo.LAYERS, offs):
        p[a:b] = rng.standard_normal(b - a) * np.sqrt(2.0 / (ks * ks * cin))
        p[b:b + cout] = 0.05 * rng.standard_normal(cout)
    loss, g = wo.loss_grad(p, x, t, scales, rnd=False)
    assert loss > 0
    eps = 1e-7  # small enough that no ReLU input of this draw changes sign within +-eps v
    for li, (a, b) in enumerate(offs):
        cout = wo.LAYERS[li][2]
        v = np.zeros_like(p)
        v[a:b + cout] = rng.standard_normal(b + cout - a)
        lp, _ = wo.loss_grad(p + eps * v, x, t, scales, rnd=False)
        lm, _ = wo.loss_grad(p - eps * v, x, t, scales, rnd=False)
        fd = (lp - lm) / (2 * eps)
        an = float(g @ v)
        assert abs(fd - an) <= 1e-5 * abs(an) + 1e-10, (li, fd, an)


def test_adam_first_step_is_signed_lr():
    rng = np.random.default_rng(3)
    w = rng.standard_normal(100).astype(np.float32)
    g = rng.standard_normal(100)
    m1, m2 = np.zeros(100, np.float32), np.zeros(100, np.float32)
    w1 = wo.adam(w.copy(), g, m1, m2, 1, 1e-3)
    np.testing.assert_allclose(w1 - w, -1e-3 * np.sign(g), rtol=1e-4, atol=1e-9)
