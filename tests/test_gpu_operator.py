"""GPU parity of the MLFMA operator, receivers and BiCGSTAB against the CPU oracle.

FP64 path: every stage is compared with the oracle on the same seeded input;
tolerance 1e-12 relative (both are FP64; summation order and FMA contraction
differ).
"""
import numpy as np
import pytest

from conftest import oracle_for, rand_c, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def isg():
    import isg as m
    return m


@pytest.mark.parametrize("level", [3, 4, 5, 6])
def test_mlfma_matches_oracle(isg, level):
    print("mlfma parity level", level, flush=True)
    o = oracle_for(level)
    s = isg.Solver(level=level, numrx=16)
    rng = np.random.default_rng(100 + level)
    x = rand_c(rng, o.N)
    r_ref = o.mlfma(x)
    r = s.mlfma(x)
    assert relerr(r, r_ref) < TOL
    # stage checkpoints (solve.cpp:218-245): multipoles and locals per level
    for i in range(2, level):
        assert relerr(s.stage(0, i), o.stage(0, i)) < TOL, ("multipoles", i)
        assert relerr(s.stage(1, i), o.stage(1, i)) < TOL, ("locals", i)


def test_mlfma_level7_matches_oracle(isg):
    """512x512: the two leaf levels take the dense matrix-core interpolation path."""
    o = oracle_for(7)
    s = isg.Solver(level=7, numrx=16)
    rng = np.random.default_rng(7)
    x = rand_c(rng, o.N)
    assert relerr(s.mlfma(x), o.mlfma(x)) < TOL
    for i in range(2, 7):
        assert relerr(s.stage(1, i), o.stage(1, i)) < TOL, ("locals", i)


@pytest.mark.parametrize("level,nrhs", [(5, 4), (6, 16)])
def test_mlfma_batched_matches_oracle(isg, level, nrhs):
    """Batched passes (every launch shape that depends on the RHS count)."""
    print("batched mlfma parity", level, nrhs, flush=True)
    o = oracle_for(level)
    s = isg.Solver(level=level, numrx=16, max_rhs=nrhs)
    rng = np.random.default_rng(300 + level)
    X = rand_c(rng, (nrhs, o.N))
    R = s.mlfma_many(X)
    for k in (0, nrhs // 2, nrhs - 1):
        assert relerr(R[k], o.mlfma(X[k])) < TOL, k


def test_geometry_matches_oracle(isg):
    o = oracle_for(5)
    s = isg.Solver(level=5, numrx=16)
    pos, r2t = s.geometry()
    assert np.array_equal(r2t, o.unkmap())
    assert np.abs(pos - o.pos()).max() < 1e-12


@pytest.mark.parametrize("level,numrx", [(4, 16), (5, 64)])
def test_receivers_match_oracle(isg, level, numrx):
    o = oracle_for(level, numrx=numrx)
    s = isg.Solver(level=level, numrx=numrx)
    rng = np.random.default_rng(7)
    x = rand_c(rng, o.N)
    y = rand_c(rng, numrx)
    assert relerr(s.aggregate(x), o.aggregate(x)) < TOL
    assert relerr(s.aggregateh(y), o.aggregateh(y)) < TOL
    assert relerr(s.mlfmah(x), o.mlfmah(x)) < TOL


@pytest.mark.parametrize("herm", [False, True])
def test_bicgs_matches_oracle(isg, herm):
    level = 5
    o = oracle_for(level)
    s = isg.Solver(level=level, numrx=16, toli=1e-6, iti=300)
    rng = np.random.default_rng(11)
    pos = o.pos()
    contrast = 0.3 * np.exp(-np.abs(pos) ** 2 / (2 * 2.0 ** 2)) * (1 + 0.1j)
    g = contrast * o.k0 ** 2
    b = np.exp(1j * o.k0 * pos.real)
    x_ref, info_ref = o.bicgs(g, b, herm=herm, tol=1e-6, maxit=300)
    x, info = s.bicgs(g, b, herm=herm)
    assert info["status"] == 0 and info_ref["status"] == 0
    assert abs(info["iters"] - info_ref["iters"]) <= 1
    assert relerr(x, x_ref) < 1e-9
