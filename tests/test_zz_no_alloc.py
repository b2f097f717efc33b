"""isg_bicgs_dev_static (the solve that allocates no device memory) against
isg_bicgs_dev on the same inputs, through the C ABI.

128x128 pixels (level 5), a weak Gaussian contrast, plane-wave right-hand
sides, tolerance 1e-6, FP64 and mixed precision.  Two calls in a row on one
context -- 8 systems forward from zero, then 3 adjoint systems from a non-zero
start (a smaller input, on scratch that still holds the first call's
contents) -- must give the same iteration counts and the same solutions as the
allocating entry point (the same kernels in the same order: agreement to
1e-14, in practice bitwise).
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

LEVEL, TOL = 5, 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_static_scratch_solve_matches_allocating_solve(precision):
    import torch
    import isg
    print("static-scratch bicgs vs bicgs_dev, precision", precision, flush=True)
    s = isg.Solver(level=LEVEL, numrx=16, max_rhs=8, precision=precision, toli=TOL, iti=300)
    pos, _ = s.geometry()
    k0 = 2 * np.pi
    g = 0.05 * np.exp(-np.abs(pos - pos.mean()) ** 2 / (2 * 3.0 ** 2)) * (1 + 0.05j) * k0 ** 2
    ang = 2 * np.pi * np.arange(8) / 8
    B = np.exp(1j * k0 * (np.cos(ang)[:, None] * pos.real[None, :] + np.sin(ang)[:, None] * pos.imag[None, :]))
    rng = np.random.default_rng(7)
    X0 = 0.1 * (rng.standard_normal((3, s.N)) + 1j * rng.standard_normal((3, s.N)))
    g_d = torch.from_numpy(np.ascontiguousarray(g)).cuda()
    B_d = torch.from_numpy(np.ascontiguousarray(B)).cuda()
    cases = [  # (nrhs, herm, start)
        (8, False, None),
        (3, True, X0),
    ]
    for nrhs, herm, start in cases:
        out = {}
        for kind in ("static", "dev", "static2"):
            X_d = (torch.zeros((nrhs, s.N), dtype=torch.complex128, device="cuda") if start is None
                   else torch.from_numpy(np.ascontiguousarray(start)).cuda())
            torch.cuda.synchronize()
            call = s.bicgs_dev if kind == "dev" else s.bicgs_dev_static
            rc, infos = call(X_d.data_ptr(), g_d.data_ptr(), B_d.data_ptr(), nrhs, herm, 0, start is None)
            s.synchronize()
            assert rc == 0, (kind, infos)
            out[kind] = (X_d.cpu().numpy(), [i["iters"] for i in infos])
        print(nrhs, "systems, herm", herm, "iterations", out["dev"][1], flush=True)
        for kind in ("static", "static2"):
            assert out[kind][1] == out["dev"][1], (kind, out[kind][1], out["dev"][1])
            for k in range(nrhs):
                assert relerr(out[kind][0][k], out["dev"][0][k]) <= 1e-14, (kind, k)
    s.close()
