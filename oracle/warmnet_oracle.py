"""CPU oracle of the warm-start network -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU leg may import this
module; the product runs the HIP kernels of
inverse-solver-gpu_amd/csrc/isg_warmnet.hip and never calls this code.

The network has no counterpart in the reference (whose total-field solves
restart from zero, mom.cpp:324); this is a plain numpy restatement of the
same 4-level encoder-decoder, written independently of the kernels:

  input (n, H, W, 8) times per-channel scales, space-to-depth by 2:
  position (Y, X) holds channel 8 s + c = input (2Y + s // 2, 2X + s % 2, c);
  L0 conv3 32->32, avg-pool, L1 conv3 32->64, avg-pool, L2 conv3 64->128,
  avg-pool, L3 conv3 128->256, L4 conv1 256->128, L5 conv3(e2 + up(u3)),
  L6 conv1 128->64, L7 conv3(e1 + up(u2)), L8 conv1 64->32,
  L9 conv3(e0 + up(u1)), L10 conv1 32->16; output channels 2 s + c (s < 4,
  c < 2) go back depth-to-space to (n, H, W, 2), times an output scale;
  ReLU after every 3x3 convolution; up = nearest 2x.

Parameters: per layer W[ks*ks][cout][cin] then bias[cout], flat float32.
`rnd=True` reproduces the GPU's numerics: both operands of every
convolution (forward, input gradient, weight gradient) rounded to BF16
(round to nearest even), sums in float64 here (FP32 on the GPU).
`rnd=False` is exact float64 arithmetic, used for finite-difference checks.

Pinned by: central finite differences of the loss with rnd=False
(tests/test_oracle_warmnet.py), the convolution against a direct 4-loop sum,
and a 1x1-identity network that reproduces its input.
"""
import numpy as np

LAYERS = [  # ks, cin, cout, level, relu
    (3, 32, 32, 0, 1), (3, 32, 64, 1, 1), (3, 64, 128, 2, 1), (3, 128, 256, 3, 1),
    (1, 256, 128, 3, 0), (3, 128, 128, 2, 1), (1, 128, 64, 2, 0), (3, 64, 64, 1, 1),
    (1, 64, 32, 1, 0), (3, 32, 32, 0, 1), (1, 32, 16, 0, 0),
]
IN_CH, OUT_CH, S2D = 8, 2, 2


def offsets():
    out, off = [], 0
    for ks, cin, cout, _, _ in LAYERS:
        wo = off
        off += ks * ks * cout * cin
        out.append((wo, off))
        off += cout
    return out, off


def num_params():
    return offsets()[1]


def unpack(params):
    """List of (W[tap][cout][cin], b[cout]) views into the flat array."""
    offs, _ = offsets()
    res = []
    for (ks, cin, cout, _, _), (wo, bo) in zip(LAYERS, offs):
        res.append((params[wo:bo].reshape(ks * ks, cout, cin), params[bo:bo + cout]))
    return res


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def s2d(x):
    """(n, H, W, C) -> (n, H/2, W/2, 4 C), channel C s + c with s = 2 dy + dx."""
    n, H, W, C = x.shape
    return x.reshape(n, H // 2, 2, W // 2, 2, C).transpose(0, 1, 3, 2, 4, 5).reshape(n, H // 2, W // 2, 4 * C)


def d2s(y, C):
    """Inverse of s2d for the first 4 C channels: (n, h, w, >= 4 C) -> (n, 2h, 2w, C)."""
    n, h, w, _ = y.shape
    return y[..., :4 * C].reshape(n, h, w, 2, 2, C).transpose(0, 1, 3, 2, 4, 5).reshape(n, 2 * h, 2 * w, C)


def up2(u):
    return np.repeat(np.repeat(u, 2, axis=1), 2, axis=2)


def pool(x, mode="avg"):
    s = x[:, 0::2, 0::2] + x[:, 1::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 1::2]
    return 0.25 * s if mode == "avg" else s


def conv_in(x, u=None, m=None):
    v = x if u is None else x + up2(u)
    if m is not None:
        v = np.where(m > 0, v, 0.0)
    return v


def conv(inp, Wt, ks, rnd):
    """out[n,y,x,co] = sum_{tap,ci} inp[n,y+dy,x+dx,ci] Wt[tap,co,ci], zero padding."""
    a = bf16(inp) if rnd else inp
    w = bf16(Wt) if rnd else Wt
    n, H, W, _ = a.shape
    r = ks // 2
    ap = np.pad(a, ((0, 0), (r, r), (r, r), (0, 0)))
    out = np.zeros((n, H, W, Wt.shape[1]))
    for t in range(ks * ks):
        ky, kx = divmod(t, ks)
        out += ap[:, ky:ky + H, kx:kx + W, :] @ w[t].T
    return out


def flip_t(Wt, ks):
    """Input-gradient weights: Wg[t'][ci][co] = W[ks*ks-1-t'][co][ci]."""
    return np.ascontiguousarray(Wt[::-1].transpose(0, 2, 1))


def wgrad(dpre, inp, ks, rnd):
    a = bf16(dpre) if rnd else dpre
    x = bf16(inp) if rnd else inp
    n, H, W, cout = a.shape
    r = ks // 2
    xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0)))
    gw = np.zeros((ks * ks, cout, x.shape[3]))
    for t in range(ks * ks):
        ky, kx = divmod(t, ks)
        gw[t] = a.reshape(-1, cout).T @ xp[:, ky:ky + H, kx:kx + W, :].reshape(-1, x.shape[3])
    return gw, a.reshape(-1, cout).sum(axis=0)


def forward(params, x8, scales, rnd=True):
    """x8: (n, H, W, 8).  Returns the activations dict; 'y' = (n, H, W, 2)
    output times the output scale."""
    P = unpack(np.asarray(params, dtype=np.float64))
    x = s2d(np.asarray(x8, dtype=np.float64) * np.asarray(scales[:IN_CH], dtype=np.float64)[None, None, None, :])
    A = {"x": x}

    def layer(i, inp):
        ks, _, _, _, relu = LAYERS[i]
        W_, b = P[i]
        v = conv(inp, W_, ks, rnd) + b
        return np.maximum(v, 0.0) if relu else v

    A["e0"] = layer(0, x)
    A["p0"] = pool(A["e0"])
    A["e1"] = layer(1, A["p0"])
    A["p1"] = pool(A["e1"])
    A["e2"] = layer(2, A["p1"])
    A["p2"] = pool(A["e2"])
    A["e3"] = layer(3, A["p2"])
    A["u3"] = layer(4, A["e3"])
    A["d2"] = layer(5, conv_in(A["e2"], A["u3"]))
    A["u2"] = layer(6, A["d2"])
    A["d1"] = layer(7, conv_in(A["e1"], A["u2"]))
    A["u1"] = layer(8, A["d1"])
    A["d0"] = layer(9, conv_in(A["e0"], A["u1"]))
    A["out"] = layer(10, A["d0"])
    A["y"] = d2s(A["out"], OUT_CH) * scales[IN_CH]
    return A


def loss_grad(params, x8, target, scales, rnd=True):
    """Mean squared error of the (depth-to-space) output against
    target / out_scale and its gradient with respect to the flat parameters."""
    P = unpack(np.asarray(params, dtype=np.float64))
    A = forward(params, x8, scales, rnd)
    k = S2D * S2D * OUT_CH
    d = A["out"][..., :k] - s2d(np.asarray(target, dtype=np.float64) / scales[IN_CH])
    loss = float(np.mean(d ** 2))
    g_out = np.zeros_like(A["out"])
    g_out[..., :k] = 2.0 * d / d.size
    offs, npar = offsets()
    G = np.zeros(npar)

    def put(i, gw, gb):
        wo, bo = offs[i]
        G[wo:bo] = gw.ravel()
        G[bo:bo + LAYERS[i][2]] = gb

    def back(i, g, out, inp):
        """Weight gradient of layer i and the gradient of its input."""
        ks, _, _, _, relu = LAYERS[i]
        dpre = np.where(out > 0, g, 0.0) if relu else g
        gw, gb = wgrad(dpre, inp, ks, rnd)
        put(i, gw, gb)
        return conv(dpre, flip_t(P[i][0], ks), ks, rnd)

    g_d0 = back(10, g_out, A["out"], A["d0"])
    g_in0 = back(9, g_d0, A["d0"], conv_in(A["e0"], A["u1"]))
    g_d1 = back(8, pool(g_in0, "sum"), A["u1"], A["d1"])
    g_in1 = back(7, g_d1, A["d1"], conv_in(A["e1"], A["u2"]))
    g_d2 = back(6, pool(g_in1, "sum"), A["u2"], A["d2"])
    g_in2 = back(5, g_d2, A["d2"], conv_in(A["e2"], A["u3"]))
    g_e3 = back(4, pool(g_in2, "sum"), A["u3"], A["e3"])
    g_p2 = back(3, g_e3, A["e3"], A["p2"])
    g_e2 = g_in2 + 0.25 * up2(g_p2)
    g_p1 = back(2, g_e2, A["e2"], A["p1"])
    g_e1 = g_in1 + 0.25 * up2(g_p1)
    g_p0 = back(1, g_e1, A["e1"], A["p0"])
    g_e0 = g_in0 + 0.25 * up2(g_p0)
    back(0, g_e0, A["e0"], A["x"])
    return loss, G


def adam(params, grads, m1, m2, step, lr):
    """One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) in float32, as the GPU."""
    f = np.float32
    g = grads.astype(f)
    m1[:] = f(0.9) * m1 + f(0.1) * g
    m2[:] = f(0.999) * m2 + f(0.001) * g * g
    c1 = f(1) - f(0.9) ** step
    c2 = f(1) - f(0.999) ** step
    params[:] = params - f(lr) * (m1 / c1) / (np.sqrt(m2 / c2) + f(1e-8))
    return params
