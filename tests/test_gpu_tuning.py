"""Autotuner: every kernel variant it can choose is an exact restatement of
the same operator, so each forced variant must reproduce the oracle's FP64
matvec (1e-12 relative, as the built-in choices do); the tuned choices
survive a save / load round trip.

Workload: 128x128 (level 5) and 256x256 (level 6) pixels, 4 right-hand sides
(batch class 0) and 16 (class 1), seeded random inputs.
"""
import numpy as np
import pytest

from conftest import oracle_for, rand_c, relerr

pytestmark = pytest.mark.gpu

M2L_TS = [4, 8, 8, 16, 8, 16, 8, 8, 8, 8]  # tile width of each M2L variant
M2L_VARIANTS_F64 = [0, 1, 2, 3, 6, 8, 9]


@pytest.fixture(scope="module")
def isg():
    import isg as m
    return m


def check(s, o, nrhs, tol=1e-12):
    X = np.stack([rand_c(np.random.default_rng(10 + r), o.N) for r in range(nrhs)])
    Y = s.mlfma_many(X)
    for r in range(nrhs):
        assert relerr(Y[r], o.mlfma(X[r])) < tol


@pytest.mark.parametrize("level", [5, 6])
def test_forced_m2l_variants_match_oracle(isg, level):
    o = oracle_for(level)
    s = isg.Solver(level=level, max_rhs=4)
    for v in M2L_VARIANTS_F64:
        print("m2l variant", v, "level", level, flush=True)
        for i in range(2, level):
            if (1 << i) >= M2L_TS[v]:
                s.tuning_set(0, 0, "m2l", i, v)
        check(s, o, 4)


def test_forced_m2l_variants_fp32(isg):
    """Every M2L variant of the FP32 operator (4, 5 and 7 exist only there)
    within the mixed-precision bar of the FP64 oracle."""
    level = 6
    o = oracle_for(level)
    s = isg.Solver(level=level, max_rhs=4, precision=1)
    X = np.stack([rand_c(np.random.default_rng(50 + r), o.N) for r in range(4)])
    ref = [o.mlfma(X[r]) for r in range(4)]
    for v in range(10):
        print("fp32 m2l variant", v, flush=True)
        for i in range(2, level):
            s.tuning_set(1, 0, "m2l", i, v if (1 << i) >= M2L_TS[v] else -1)
        Y = s.mlfma_many(X)
        for r in range(4):
            e = relerr(Y[r], ref[r])
            assert 1e-12 < e < 1e-5, (v, e)


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_forced_interp_kinds_match_oracle(isg, kind):
    level = 6
    o = oracle_for(level)
    s = isg.Solver(level=level, max_rhs=4)
    print("interp kind", kind, flush=True)
    for i in range(0, level - 1):
        s.tuning_set(0, 0, "m2m", i, kind)  # kinds a level lacks fall back to the built-in rule
        s.tuning_set(0, 0, "l2l", i, kind)
    for w in (16, 32):
        s.tuning_set(0, 0, "near", 0, w)
        s.tuning_set(0, 0, "p2m", 0, w)
        check(s, o, 4)


@pytest.mark.parametrize("level", [5, 6])
def test_fused_leaf_matches_oracle(isg, level):
    """Leaf M2L fused with the last L2L (tuning stage "leaf" = 1): the FP64
    operator to 1e-12 of the oracle with every upper-level interpolation
    kind, the FP32 operator within the mixed-precision bar."""
    o = oracle_for(level)
    s64 = isg.Solver(level=level, max_rhs=4)
    s32 = isg.Solver(level=level, max_rhs=4, precision=1)
    s64.tuning_set(0, 0, "leaf", 0, 1)
    s32.tuning_set(1, 0, "leaf", 0, 1)
    assert s64.tuning_get(0, 0, "leaf") == 1 and s32.tuning_get(1, 0, "leaf") == 1
    X = np.stack([rand_c(np.random.default_rng(70 + r), o.N) for r in range(4)])
    ref = [o.mlfma(X[r]) for r in range(4)]
    for kind in (-1, 0, 1):
        print("fused leaf, level", level, "upper l2l kind", kind, flush=True)
        for i in range(0, level - 2):
            s64.tuning_set(0, 0, "l2l", i, kind)
            s32.tuning_set(1, 0, "l2l", i, kind)
        Y = s64.mlfma_many(X)
        Y32 = s32.mlfma_many(X)
        for r in range(4):
            assert relerr(Y[r], ref[r]) < 1e-12
            assert 1e-12 < relerr(Y32[r], ref[r]) < 1e-5


@pytest.mark.parametrize("herm", [False, True])
def test_fp32_tile_width_64(isg, herm):
    """64-box near-field and P2M tiles (FP32 operator only): the batched
    product within the mixed-precision bar of the oracle, and a mixed
    BiCGSTAB solve, whose near-field launches carry the fused post steps,
    to the FP64 residual of the tolerance."""
    level, tol = 5, 1e-6
    o = oracle_for(level)
    s = isg.Solver(level=level, max_rhs=4, toli=tol, iti=300, precision=1)
    for st in ("near", "p2m"):
        s.tuning_set(1, 0, st, 0, 64)
        assert s.tuning_get(1, 0, st) == 64
    X = np.stack([rand_c(np.random.default_rng(90 + r), o.N) for r in range(4)])
    Y = s.mlfma_many(X)
    for r in range(4):
        assert 1e-12 < relerr(Y[r], o.mlfma(X[r])) < 1e-5
    pos = o.pos()
    g = 0.4 * np.exp(-np.abs(pos) ** 2 / (2 * 2.5 ** 2)) * (1 + 0.05j) * o.k0 ** 2
    b = np.exp(1j * o.k0 * (0.6 * pos.real + 0.8 * pos.imag))
    x, info = s.bicgs(g, b, herm=herm)
    assert info["status"] == 0, info
    ax = x - np.conj(g * o.mlfma(np.conj(x))) if herm else x - o.mlfma(g * x)
    assert np.linalg.norm(b - ax) / np.linalg.norm(b) < tol * 1.05


def test_autotune_save_load(isg, tmp_path):
    level = 6
    o = oracle_for(level)
    s = isg.Solver(level=level, max_rhs=16, precision=1)
    print("autotune level 6, 16 rhs", flush=True)
    s.autotune(16, reps=2)
    p = tmp_path / "tune.txt"
    s.tuning_save(p)
    # the context's batched matvec runs the FP32 operator with the tuned choices
    X = np.stack([rand_c(np.random.default_rng(30 + r), o.N) for r in range(16)])
    Y = s.mlfma_many(X)
    for r in range(16):
        e = relerr(Y[r], o.mlfma(X[r]))
        assert 1e-12 < e < 1e-5, e
    # an FP64 context loading the same file runs the tuned FP64 choices
    t = isg.Solver(level=level, max_rhs=16)
    t.tuning_load(p)
    for prec in (0, 1):
        for st in ("near", "p2m", "leaf"):
            assert t.tuning_get(prec, 1, st) == s.tuning_get(prec, 1, st)
        for st in ("m2l", "m2m", "l2l"):
            for i in range(level):
                assert t.tuning_get(prec, 1, st, i) == s.tuning_get(prec, 1, st, i)
    check(t, o, 16)
