"""Split-FP16 near-field kernel (k_near_h16, tuning value 128 of the FP32
operator's near stage) against the FP32 kernel and the FP64 oracle.

The kernel carries every FP32 operand as two FP16 numbers (hi + lo, 22
significant bits) under power-of-two scales, so its products differ from the
FP32 kernel's by a few 2^-22 of their norm.  Bars: the FP32 operator with the
split near field within 1e-5 relative of the FP64 oracle (the FP32 bar), and
within 2e-6 of the same operator with the FP32 near kernel; a wrong table,
halo or scale mapping moves whole boxes and fails both by orders of
magnitude.  Workload: 128x128 (level 5) and 256x256 (level 6) pixels, 5
right-hand sides (an odd count leaves the last workgroup's pair half empty)
whose magnitudes differ by 1e16 (per-workgroup data scales), seeded.
"""
import numpy as np
import pytest

from conftest import oracle_for, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def isg():
    import isg as m
    return m


def fp32_products(s, X):
    import torch
    x = torch.from_numpy(X.astype(np.complex64)).cuda()
    r = torch.empty_like(x)
    torch.cuda.synchronize()
    s.mlfma_f32_dev(x.data_ptr(), r.data_ptr(), X.shape[0])
    s.synchronize()
    return r.cpu().numpy().astype(np.complex128)


@pytest.mark.parametrize("level", [5, 6])
def test_split_near_matches_fp32_kernel_and_oracle(isg, level):
    print("split near level", level, flush=True)
    o = oracle_for(level)
    nrhs = 5
    s = isg.Solver(level=level, max_rhs=nrhs, precision=1)
    rng = np.random.default_rng(70 + level)
    X = rng.standard_normal((nrhs, o.N)) + 1j * rng.standard_normal((nrhs, o.N))
    X[1] *= 1e8
    X[2] *= 1e-8
    # a field decaying by 1e-4 across the domain: elements far below the
    # workgroup's largest one
    pos = o.pos()
    X[3] *= np.exp(-9.2 * (pos.real - pos.real.min()) / np.ptp(pos.real))
    X = X.astype(np.complex64).astype(np.complex128)
    s.tuning_set(1, 0, "near", 0, 32)
    R32 = fp32_products(s, X)
    s.tuning_set(1, 0, "near", 0, 128)
    assert s.tuning_get(1, 0, "near", 0) == 128
    R16 = fp32_products(s, X)
    for k in range(nrhs):
        ref = o.mlfma(X[k])
        e16, e32, d = relerr(R16[k], ref), relerr(R32[k], ref), relerr(R16[k], R32[k])
        print("rhs %d: split %.2e  fp32 %.2e  split vs fp32 %.2e" % (k, e16, e32, d))
        assert e16 < 1e-5, (k, e16)
        assert 0 < d < 2e-6, (k, d)


def test_split_near_mixed_bicgs_reaches_fp64_residual(isg):
    """The solver's post-step epilogues (modes 1 and 2) in the split kernel:
    a mixed solve with it meets the FP64 true residual (recomputed with the
    oracle), forward and adjoint."""
    level, tol = 5, 1e-6
    o = oracle_for(level)
    s = isg.Solver(level=level, numrx=16, toli=tol, iti=300, precision=1)
    s.tuning_set(1, 0, "near", 0, 128)
    pos = o.pos()
    g = 0.4 * np.exp(-np.abs(pos) ** 2 / (2 * 2.5 ** 2)) * (1 + 0.05j) * o.k0 ** 2
    b = np.exp(1j * o.k0 * (0.6 * pos.real + 0.8 * pos.imag))
    for herm in (False, True):
        x, info = s.bicgs(g, b, herm=herm)
        assert info["status"] == 0, info
        ax = x - np.conj(g * o.mlfma(np.conj(x))) if herm else x - o.mlfma(g * x)
        res = np.linalg.norm(b - ax) / np.linalg.norm(b)
        print("herm %d: %d iterations, FP64 residual %.2e" % (herm, info["iters"], res))
        assert res < tol * 1.05, res


def test_split_near_batched_dbim_matches_oracle(isg):
    """Batched mixed DBIM (16 right-hand sides with deflation: workgroup pairs
    whose second system has converged) with the split near field, within 1e-4
    of the oracle's contrast after 3 iterations."""
    from oracle.oracle import Dbim as OrcDbim
    level, ntx, nrx, tol, reg = 4, 16, 16, 1e-6, float(np.sqrt(1e-3))
    o = oracle_for(level, numrx=nrx)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    ref = OrcDbim(o, ntx, otrue, regular=reg, tol=tol, maxit=300)
    s = isg.Solver(level=level, numrx=nrx, numtx=ntx, toli=tol, iti=300, regular=reg, max_rhs=ntx, precision=1)
    for cls in (0, 1):
        s.tuning_set(1, cls, "near", 0, 128)
    d = isg.Dbim(s, otrue, field_start=1)
    for _ in range(3):
        ref.iterate()
        st = d.iterate()
        assert st["not_converged"] == 0
    e = relerr(d.get("o"), ref.get("o"))
    print("contrast relative difference %.2e" % e)
    assert e < 1e-4
