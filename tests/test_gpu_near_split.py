"""Split-FP16 kernels of the FP32 operator (near tuning value 128: k_near_h16,
256: k_near_h16w for the near field + L2P; P2M value 128: k_p2m_h16) against
the FP32 kernels and the FP64 oracle.

The kernels carry every FP32 operand as two FP16 numbers (hi + lo, 22
significant bits) under power-of-two scales and multiply in the
four-product complex form.  Bars: the FP32 operator with split stages within
1e-5 relative of the FP64 oracle (the FP32 bar) and no further from it than
with the FP32 kernels (x 1.1); within 5e-6 of the product with the FP32
kernels (the split near field moves it by 5-7e-7; the split P2M by ~2e-6,
mostly the FP32 P2M's own rounding: its three-product form cancels in the
imaginary part, and its error is amplified by the leaf translations, so the
split product is the closer one to the oracle).  A wrong table, halo or
scale mapping moves whole boxes and fails by orders of magnitude.  Workload: 128x128 (level 5) and 256x256 (level 6) pixels, 5
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

    def products(near, p2m):
        s.tuning_set(1, 0, "near", 0, near)
        s.tuning_set(1, 0, "p2m", 0, p2m)
        assert s.tuning_get(1, 0, "near", 0) == near and s.tuning_get(1, 0, "p2m", 0) == p2m
        return fp32_products(s, X)

    R32 = products(32, 32)
    refs = [o.mlfma(X[k]) for k in range(nrhs)]
    for near, p2m in ((128, 32), (32, 128), (128, 128), (256, 128)):
        R16 = products(near, p2m)
        for k in range(nrhs):
            e16, e32, d = relerr(R16[k], refs[k]), relerr(R32[k], refs[k]), relerr(R16[k], R32[k])
            print("near %d p2m %d rhs %d: split %.2e  fp32 %.2e  split vs fp32 %.2e" % (near, p2m, k, e16, e32, d))
            assert e16 < 1e-5 and e16 < 1.1 * e32, (near, p2m, k, e16, e32)
            assert 0 < d < 5e-6, (near, p2m, k, d)


def test_split_near_mixed_bicgs_reaches_fp64_residual(isg):
    """The solver's post-step epilogues (modes 1 and 2) in the split kernels
    (pairs on 64-box tiles and single systems on 128-box tiles; split P2M):
    a batched mixed solve of 5 systems (plane waves from 5 directions, so
    they converge at different iterations and deflation masks workgroup
    pairs half off) meets the FP64 true residual (recomputed with the
    oracle) in every system, forward and adjoint."""
    import torch
    level, tol, nrhs = 5, 1e-6, 5
    o = oracle_for(level)
    pos = o.pos()
    g = 0.4 * np.exp(-np.abs(pos) ** 2 / (2 * 2.5 ** 2)) * (1 + 0.05j) * o.k0 ** 2
    ang = np.linspace(0, 2 * np.pi, nrhs, endpoint=False) + 0.3
    B = np.stack([np.exp(1j * o.k0 * (np.cos(a) * pos.real + np.sin(a) * pos.imag)) for a in ang])
    g_d = torch.from_numpy(g).cuda()
    B_d = torch.from_numpy(B).cuda()
    for near, herm in ((128, False), (128, True), (256, False), (256, True)):
        s = isg.Solver(level=level, numrx=16, toli=tol, iti=300, precision=1, max_rhs=nrhs)
        s.tuning_set(1, 0, "near", 0, near)
        s.tuning_set(1, 0, "p2m", 0, 128)
        X_d = torch.zeros_like(B_d)
        torch.cuda.synchronize()
        _, infos = s.bicgs_dev(X_d.data_ptr(), g_d.data_ptr(), B_d.data_ptr(), nrhs, herm, 0, True)
        s.synchronize()
        X = X_d.cpu().numpy()
        for k in range(nrhs):
            assert infos[k]["status"] == 0, infos[k]
            x = X[k]
            ax = x - np.conj(g * o.mlfma(np.conj(x))) if herm else x - o.mlfma(g * x)
            res = np.linalg.norm(B[k] - ax) / np.linalg.norm(B[k])
            print("near %d herm %d rhs %d: %d iterations, FP64 residual %.2e" % (near, herm, k, infos[k]["iters"], res))
            assert res < tol * 1.05, (k, res)


def test_split_near_batched_dbim_matches_oracle(isg):
    """Batched mixed DBIM (16 right-hand sides with deflation: workgroup pairs
    whose second system has converged) with the split near field and P2M, within 1e-4
    of the oracle's contrast after 3 iterations."""
    from oracle.oracle import Dbim as OrcDbim
    level, ntx, nrx, tol, reg = 4, 16, 16, 1e-6, float(np.sqrt(1e-3))
    o = oracle_for(level, numrx=nrx)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    ref = OrcDbim(o, ntx, otrue, regular=reg, tol=tol, maxit=300)
    s = isg.Solver(level=level, numrx=nrx, numtx=ntx, toli=tol, iti=300, regular=reg, max_rhs=ntx, precision=1)
    for cls in (0, 1):
        s.tuning_set(1, cls, "near", 0, 128)
        s.tuning_set(1, cls, "p2m", 0, 128)
    d = isg.Dbim(s, otrue, field_start=1)
    for _ in range(3):
        ref.iterate()
        st = d.iterate()
        assert st["not_converged"] == 0
    e = relerr(d.get("o"), ref.get("o"))
    print("contrast relative difference %.2e" % e)
    assert e < 1e-4
