"""Mixed precision on the GPU: FP32 operator, FP64 solver scalars and vectors.

Bars (stated here, checked against the CPU oracle on seeded inputs):
  * FP32 MLFMA matvec within 1e-5 relative of the FP64 oracle matvec;
  * mixed BiCGSTAB: FP64 true relative residual below the tolerance (the
    solver's own defect correction), solution within 10x tolerance of the FP64
    oracle solution;
  * mixed + batched DBIM on the 64x64 configuration: reconstructed contrast
    within 1e-4 relative of the oracle's after 3 iterations.
"""
import numpy as np
import pytest

from conftest import oracle_for, rand_c, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def isg():
    import isg as m
    return m


@pytest.mark.parametrize("level,nrhs", [(4, 1), (6, 4), (7, 2)])
def test_fp32_matvec_within_1e5(isg, level, nrhs):
    print("fp32 matvec level", level, flush=True)
    o = oracle_for(level)
    s = isg.Solver(level=level, numrx=16, precision=1, max_rhs=nrhs)
    rng = np.random.default_rng(500 + level)
    X = rand_c(rng, (nrhs, o.N))
    R = s.mlfma_many(X)
    for k in range(nrhs):
        e = relerr(R[k], o.mlfma(X[k]))
        assert e < 1e-5, e
        assert e > 1e-12  # really computed in FP32


@pytest.mark.parametrize("level,r0,nr", [(4, 0, 4096), (5, 2048, 1024), (6, 4096, 4096)])
def test_matvec_vs_direct_mom(isg, level, r0, nr):
    """Both operators against the dense direct MoM product (the matrix the
    MLFMA approximates, sampled by the oracle).  Measured on MI355X: FP64
    MLFMA 3e-7 .. 5.4e-7, FP32 operator 1.4e-6 .. 2.5e-6, i.e. the FP32
    rounding is about 4x the MLFMA's own approximation error (with three- or
    four-product complex MACs alike).  Bars: FP64 < 1e-6, FP32 < 5e-6; every
    mixed solve is certified against the FP64 operator.  Level 5 compares a
    1024-row window; level 6 (256^2) a 64x64-pixel sub-window (8x8 leaf boxes,
    Morton-contiguous rows 4096..8191)."""
    print("matvec vs direct MoM, level", level, flush=True)
    o = oracle_for(level)
    s64 = isg.Solver(level=level, numrx=16, max_rhs=2)
    s32 = isg.Solver(level=level, numrx=16, max_rhs=2, precision=1)
    X = rand_c(np.random.default_rng(700 + level), (2, o.N))
    R64, R32 = s64.mlfma_many(X), s32.mlfma_many(X)
    for k in range(2):
        d = o.direct(X[k], r0, nr)
        e64 = relerr(R64[k][r0:r0 + nr], d)
        e32 = relerr(R32[k][r0:r0 + nr], d)
        print("  rhs", k, "fp64 error", e64, "fp32 error", e32, flush=True)
        assert 1e-12 < e64 < 1e-6, e64
        assert e32 < 5e-6, e32


@pytest.mark.parametrize("herm", [False, True])
def test_mixed_defect_correction(isg, herm):
    """A tolerance below the FP32 operator's accuracy (1e-8 on a strong
    scatterer): the FP32 solve's recursive residual meets its target but the
    FP64 check does not, so at least one defect-correction round runs, and the
    final FP64 residual (recomputed with the oracle) meets the tolerance."""
    level, tol = 5, 1e-8
    o = oracle_for(level)
    s = isg.Solver(level=level, numrx=16, toli=tol, iti=300, precision=1)
    pos = o.pos()
    g = 0.4 * np.exp(-np.abs(pos) ** 2 / (2 * 2.5 ** 2)) * (1 + 0.05j) * o.k0 ** 2
    b = np.exp(1j * o.k0 * (0.6 * pos.real + 0.8 * pos.imag))
    x, info = s.bicgs(g, b, herm=herm)
    assert info["status"] == 0, info
    assert s.counters()["refinements"] >= 1
    ax = x - np.conj(g * o.mlfma(np.conj(x))) if herm else x - o.mlfma(g * x)
    assert np.linalg.norm(b - ax) / np.linalg.norm(b) < tol * 1.05


@pytest.mark.parametrize("herm", [False, True])
def test_mixed_bicgs_reaches_fp64_residual(isg, herm):
    level, tol = 5, 1e-6
    o = oracle_for(level)
    s = isg.Solver(level=level, numrx=16, toli=tol, iti=300, precision=1)
    pos = o.pos()
    contrast = 0.4 * np.exp(-np.abs(pos) ** 2 / (2 * 2.5 ** 2)) * (1 + 0.05j)
    g = contrast * o.k0 ** 2
    b = np.exp(1j * o.k0 * (0.6 * pos.real + 0.8 * pos.imag))
    x, info = s.bicgs(g, b, herm=herm)
    assert info["status"] == 0, info
    # independent FP64 check of the residual with the oracle operator
    if herm:
        ax = x - np.conj(g * o.mlfma(np.conj(x)))
    else:
        ax = x - o.mlfma(g * x)
    true_res = np.linalg.norm(b - ax) / np.linalg.norm(b)
    assert true_res < tol * 1.05, true_res
    x_ref, _ = o.bicgs(g, b, herm=herm, tol=tol, maxit=300)
    assert relerr(x, x_ref) < 10 * tol * 10


@pytest.mark.parametrize("products", [1, 0])
def test_mixed_batched_dbim_matches_oracle(isg, products):
    """products=1: the derivative source G y on the FP32 operator; 0: on the
    FP64 one.  The gradient's G^H c is never computed separately: it is the
    FP64 product of the adjoint solve's final residual check, so the FP64
    products of an iteration are exactly the checks, plus numtx when
    products == 0."""
    from oracle.oracle import Dbim as OrcDbim
    print("mixed dbim parity, products", products, flush=True)
    level, ntx, nrx, tol, reg = 4, 16, 16, 1e-6, float(np.sqrt(1e-3))
    o = oracle_for(level, numrx=nrx)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    ref = OrcDbim(o, ntx, otrue, regular=reg, tol=tol, maxit=300)
    s = isg.Solver(level=level, numrx=nrx, numtx=ntx, toli=tol, iti=300, regular=reg, max_rhs=ntx, precision=1,
                   products=products)
    d = isg.Dbim(s, otrue, field_start=1)
    for _ in range(3):
        ref.iterate()
        c0 = s.counters()
        st = d.iterate()
        c1 = s.counters()
        assert st["not_converged"] == 0
        # FP64 products: the true-residual check of every solved system, once
        # more per defect-correction round (each covers the batch of ntx), plus
        # the derivative source when products == 0
        fp64 = (c1["matvecs"] - c1["matvecs_fp32"]) - (c0["matvecs"] - c0["matvecs_fp32"])
        checks = (c1["solves"] - c0["solves"]) + ntx * (c1["refinements"] - c0["refinements"])
        assert fp64 == checks + (ntx if products == 0 else 0), (fp64, checks, c0, c1)
    assert relerr(d.get("o"), ref.get("o")) < 1e-4


def test_fp32_vector_matvec(isg):
    """isg_mlfma_f32_dev: FP32 operator on complex64 device vectors, 8
    right-hand sides at 128x128, within the FP32 bar of the oracle."""
    import torch
    level = 5
    o = oracle_for(level)
    s = isg.Solver(level=level, max_rhs=8, precision=1)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((8, o.N)) + 1j * rng.standard_normal((8, o.N))
    x = torch.from_numpy(X.astype(np.complex64)).cuda()
    r = torch.empty_like(x)
    torch.cuda.synchronize()
    s.mlfma_f32_dev(x.data_ptr(), r.data_ptr(), 8)
    s.synchronize()
    R = r.cpu().numpy()
    for k in range(8):
        e = relerr(R[k], o.mlfma(X[k].astype(np.complex64).astype(np.complex128)))
        assert 1e-12 < e < 1e-5, e
