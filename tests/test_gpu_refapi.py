"""The reference-named shim (csrc/isg_refapi.cpp) executed on the GPU.

tests/_build/refapi_driver is a driver in the reference's style: it defines
mom.cpp:4-22's globals, lets setup_mlfma fill pos and unkmap, and calls
setup_mlfma / setup_gpu / setup_bicgs (mom.cpp:184-190), then mlfma, mlfmah,
aggregate, aggregateh and a forward and an adjoint bicgs from zero
(mom.cpp:250-350, setup_born.cpp:144,148) through the shim on 64x64 pixels
(level 4).  Every result is compared with the CPU oracle: operators to 1e-12,
solutions to 1e-9 with BiCGSTAB iteration counts (the reference's printed
line, bicgs_gpu.cu:220) within one of the oracle's.  With ISG_PRECISION=1 the
shim's solves run on the mixed-precision path: their solutions meet the same
FP64 tolerance, so they agree with the oracle's to the tolerance's order.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import oracle_for, rand_c, relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_build", "refapi_driver")
LEVEL, NRX, TOL = 4, 16, 1e-6


def rdbin(path, dtype, n):
    return np.fromfile(path, dtype=dtype, count=n)


@pytest.mark.parametrize("precision", [0, 1])
def test_reference_driver_through_shim(tmp_path, precision):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/_build/refapi_driver is not built (__graft_entry__.build())")
    import isg
    print("reference-style driver through the shim, ISG_PRECISION=%d" % precision, flush=True)
    o = oracle_for(LEVEL, numrx=NRX)
    N = o.N
    rng = np.random.default_rng(31)
    x = rand_c(rng, N)
    r = rand_c(rng, NRX)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    g = otrue * o.k0 ** 2
    b = np.exp(1j * o.k0 * o.pos().real)
    for name, v in (("x", x), ("r", r), ("g", g), ("b", b)):
        np.ascontiguousarray(v, np.complex128).tofile(str(tmp_path / (name + ".bin")))
    env = dict(os.environ, ISG_PRECISION=str(precision))
    p = subprocess.run([DRIVER, str(tmp_path), str(LEVEL), str(NRX), repr(TOL)], capture_output=True, text=True,
                       timeout=100, env=env)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0 and "DRIVER DONE" in p.stdout, (p.returncode, p.stderr)
    assert "MLFMAH CHANGED ITS INPUT" not in p.stdout
    assert "NOT CONVERGED" not in p.stdout

    # globals filled by setup_mlfma
    assert np.array_equal(rdbin(tmp_path / "unkmap.bin", np.int32, N), o.unkmap())
    assert relerr(rdbin(tmp_path / "pos.bin", np.complex128, N), o.pos()) < 1e-15

    for name, ref in (("mlfma", o.mlfma(x)), ("mlfmah", o.mlfmah(x)), ("aggregate", o.aggregate(x)),
                      ("aggregateh", o.aggregateh(r))):
        got = rdbin(tmp_path / (name + ".bin"), np.complex128, ref.size)
        assert relerr(got, ref) < 1e-12, (name, relerr(got, ref))

    its = [int(m.group(2)) for m in re.finditer(r"NUMBER OF (\d) ITERATIONS: (\d+) \((\d+)\)", p.stdout)]
    assert len(its) == 2, p.stdout
    for k, herm in enumerate((False, True)):
        xr, info = o.bicgs(g, b, herm=herm, tol=TOL, maxit=300)
        got = rdbin(tmp_path / ("bicgs_h.bin" if herm else "bicgs_f.bin"), np.complex128, N)
        e = relerr(got, xr)
        if precision == 0:
            assert e < 1e-9, (herm, e)
            assert abs(its[k] - info["iters"]) <= 1, (herm, its[k], info["iters"])
        else:
            assert e < 1e-5, (herm, e)
            # certified in FP64: the true residual of the shim's solution meets the tolerance
            ax = got - (np.conj(g) * o.mlfmah(got) if herm else o.mlfma(g * got))
            assert np.linalg.norm(b - ax) / np.linalg.norm(b) < TOL * 1.05
