"""Pin the oracle's quadtree and operator tables (setup_mlfma.cpp:70-695).

Every table is checked against what the mathematics fixes, not against a
retyped formula: brute-force geometry, Lagrange identities, the transpose
relation of the anterpolators, closed-form plane waves and the addition
theorem for the translation operators.
"""
import numpy as np
import pytest
from scipy import special

from conftest import oracle_for


@pytest.fixture(scope="module")
def orc():
    return oracle_for(5, numrx=64)


def grid_xy(orc, i):
    """(X, Y) box coordinates of every cluster of level i from its centre."""
    a = orc.clusize[i]
    D = orc.clusize[0]
    c = orc.centres(i)
    X = np.rint((c.real + D / 2) / a - 0.5).astype(int)
    Y = np.rint((D / 2 - c.imag) / a - 0.5).astype(int)
    return X, Y


def test_raster_map_and_positions(orc):
    um = orc.unkmap()
    assert np.array_equal(np.sort(um), np.arange(orc.N))
    pos = orc.pos()
    side, res = orc.side, orc.res
    D = side * res
    k, l = np.divmod(np.arange(orc.N), side)
    ref = (-D / 2 + (l + 0.5) * res) + 1j * (D / 2 - (k + 0.5) * res)
    assert np.abs(pos[um] - ref).max() < 1e-12


def test_morton_tree(orc):
    """child = 4 parent + 2 row + col, row bit first (setup_mlfma.cpp:137)."""
    for i in range(1, orc.level):
        X, Y = grid_xy(orc, i)
        Xp, Yp = grid_xy(orc, i - 1)
        m = np.arange(4 ** i)
        assert np.array_equal(X, 2 * Xp[m // 4] + (m % 4) % 2)
        assert np.array_equal(Y, 2 * Yp[m // 4] + (m % 4) // 2)


def test_near_slots_and_far_lists(orc):
    for i in range(1, orc.level):
        X, Y = grid_xy(orc, i)
        nb = 2 ** i
        where = {(x, y): m for m, (x, y) in enumerate(zip(X, Y))}
        far = orc.far_list(i)
        near = orc.near_list(i)
        for m in range(4 ** i):
            par_nb = {(X[m] // 2 + dx, Y[m] // 2 + dy) for dx in (-1, 0, 1) for dy in (-1, 0, 1)}
            cand = {where[(2 * px + cx, 2 * py + cy)] for (px, py) in par_nb for cx in (0, 1) for cy in (0, 1)
                    if 0 <= px < nb // 2 and 0 <= py < nb // 2}
            adj = {n for n in cand if abs(X[n] - X[m]) <= 1 and abs(Y[n] - Y[m]) <= 1}
            assert set(far[m][far[m] >= 0]) == cand - adj
            assert set(near[m][near[m] >= 0]) == adj
            if i == orc.level - 1:  # leaf slots are geometric: 3 (dY+1) + (dX+1)
                for s in range(9):
                    x, y = X[m] + s % 3 - 1, Y[m] + s // 3 - 1
                    want = where.get((x, y), -1) if 0 <= x < nb and 0 <= y < nb else -1
                    assert near[m][s] == want


def test_near_blocks_are_green(orc):
    blk = orc.near_blocks()
    near = orc.near_list(orc.level - 1)
    pos = orc.pos()
    rng = np.random.default_rng(1)
    for s in range(9):
        b = next(m for m in range(near.shape[0]) if near[m][s] >= 0)
        n = near[b][s]
        for t, u in rng.integers(0, 64, (12, 2)):
            assert abs(blk[s, t, u] - orc.green(pos[b * 64 + t], pos[n * 64 + u])) < 1e-15


def test_p2m_l2p_plane_waves(orc):
    S = orc.numsamp[-1]
    pos = orc.pos()[:64]
    c0 = orc.centres(orc.level - 1)[0]
    ang = 2 * np.pi * np.arange(S) / S
    ud = np.cos(ang)[:, None] * (pos - c0).real[None, :] + np.sin(ang)[:, None] * (pos - c0).imag[None, :]
    coeff = orc.res ** 2 * np.exp(-1j * orc.k0 * ud)
    basis = 0.25j / S * np.exp(1j * orc.k0 * ud).T
    assert np.abs(orc.coeff() - coeff).max() < 1e-14
    assert np.abs(orc.basis() - basis).max() < 1e-14


def test_shifts(orc):
    for i in range(orc.level - 1):
        S = orc.numsamp[i]
        mul, loc = orc.shift(i)
        cp = orc.centres(i)[0]
        ch = orc.centres(i + 1)[:4]
        ang = 2 * np.pi * np.arange(S) / S
        for c in range(4):
            rho = cp - ch[c]
            ph = orc.k0 * (np.cos(ang) * rho.real + np.sin(ang) * rho.imag)
            assert np.abs(mul[c] - np.exp(1j * ph)).max() < 1e-13
            assert np.abs(loc[c] - np.exp(-1j * ph)).max() < 1e-13


@pytest.mark.parametrize("i", [0, 1, 2, 3])
def test_interpolators_lagrange(orc, i):
    """Rows are Lagrange weights: they reproduce polynomials of degree < 30 in
    the unwrapped angle coordinate and interpolate band-limited patterns."""
    Sp, Sc = orc.numsamp[i], orc.numsamp[i + 1]
    w, idx = orc.interp(i)
    assert np.allclose(w.sum(1), 1.0, atol=1e-9)
    m = np.arange(Sp)
    centre = np.ceil(m * (Sc / Sp)).astype(int)
    xt = (centre[:, None] + np.arange(30)[None, :] - 15) / Sc
    assert np.array_equal(idx, np.mod(centre[:, None] + np.arange(30)[None, :] - 15, Sc))
    for p in (1, 2, 5):
        assert np.abs((w * xt ** p).sum(1) - (m / Sp) ** p).max() < 1e-7
    # band-limited pattern: far below the child's Nyquist
    q = Sc // 8
    f = lambda th: np.cos(q * th) + 0.3j * np.sin((q // 2) * th)
    fc = f(2 * np.pi * np.arange(Sc) / Sc)
    approx = (w * fc[idx]).sum(1)
    assert np.abs(approx - f(2 * np.pi * m / Sp)).max() < 1e-6


@pytest.mark.parametrize("i", [1, 2, 3, 4])
def test_anterpolators_are_scaled_transpose(orc, i):
    Sp, Sc = orc.numsamp[i - 1], orc.numsamp[i]
    w, idx = orc.interp(i - 1)
    W = np.zeros((Sp, Sc))
    for m in range(Sp):
        np.add.at(W[m], idx[m], w[m])
    aw, aidx = orc.anterp(i)
    A = np.zeros((Sc, Sp))
    for j in range(Sc):
        np.add.at(A[j], aidx[j], aw[j])
    assert np.abs(A - W.T * Sc / Sp).max() < 1e-12
    # nonzero taps run contiguously modulo Sp (setup_mlfma.cpp:619-630)
    for j in range(Sc):
        nz = np.nonzero(aw[j])[0]
        first = aidx[j][0]
        assert np.array_equal(aidx[j][nz], np.mod(first + nz, Sp))


@pytest.mark.parametrize("i", [2, 3, 4])
def test_translation_addition_theorem(orc, i):
    """(1/S) sum_k e^{ik u_k.(r-c_m)} T_(m,n)[k] e^{-ik u_k.(r'-c_n)} -> H0^(1)(k|r-r'|)."""
    S = orc.numsamp[i]
    T = orc.trans(i)
    far, traid = orc.far_list(i), orc.traid(i)
    cen = orc.centres(i)
    a = orc.clusize[i]
    rng = np.random.default_rng(i)
    ang = 2 * np.pi * np.arange(S) / S
    u = np.cos(ang) + 1j * np.sin(ang)
    dot = lambda v: u.real * v.real + u.imag * v.imag
    worst = 0
    for m in rng.integers(0, 4 ** i, 6):
        for s in np.nonzero(far[m] >= 0)[0][:8]:
            n = far[m][s]
            r = cen[m] + complex(*rng.uniform(-a / 2, a / 2, 2))
            rp = cen[n] + complex(*rng.uniform(-a / 2, a / 2, 2))
            val = np.mean(np.exp(1j * orc.k0 * dot(r - cen[m])) * T[traid[m][s]] *
                          np.exp(-1j * orc.k0 * dot(rp - cen[n])))
            ref = special.hankel1(0, orc.k0 * abs(r - rp))
            worst = max(worst, abs(val - ref) / abs(ref))
    assert worst < 1e-4


def test_receiver_interpolators(orc):
    R, S0 = orc.numrx, orc.numsamp[0]
    w, idx, aw, aidx = orc.rx_interp()
    assert np.allclose(w.sum(1), 1.0, atol=1e-9)
    W = np.zeros((R, S0))
    for m in range(R):
        np.add.at(W[m], idx[m], w[m])
    A = np.zeros((S0, R))
    for j in range(S0):
        np.add.at(A[j], aidx[j], aw[j])
    assert np.abs(A - W.T * S0 / R).max() < 1e-12
