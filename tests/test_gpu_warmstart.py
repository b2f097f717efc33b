"""Total-field warm starts of the DBIM loop on the GPU.

64x64 pixels (level 4), 16 plane-wave transmitters, 16 receivers, 3 seeded
blobs (seed 1), tolerance 1e-6, regular = sqrt(1e-3), 16 right-hand sides per
batch, 3 DBIM iterations.  Every start must reproduce the oracle's contrast
(zero start, mom.cpp:324) within 1e-4 relative: a start changes only the
iterate BiCGSTAB begins from, never the tolerance it stops at.
"""
import numpy as np
import pytest

from conftest import oracle_for, relerr

pytestmark = pytest.mark.gpu

LEVEL, NTX, NRX, TOL, REG = 4, 16, 16, 1e-6, float(np.sqrt(1e-3))


@pytest.fixture(scope="module")
def case():
    import isg
    from oracle.oracle import Dbim as OrcDbim
    o = oracle_for(LEVEL, numrx=NRX)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    ref = OrcDbim(o, NTX, otrue, regular=REG, tol=TOL, maxit=300)
    for _ in range(3):
        ref.iterate()
    return isg, otrue, ref.get("o")


def run(isg, otrue, field_start, precision=0):
    s = isg.Solver(level=LEVEL, numrx=NRX, numtx=NTX, toli=TOL, iti=300, regular=REG, max_rhs=NTX,
                   precision=precision)
    d = isg.Dbim(s, otrue, field_start=field_start)
    iters = 0
    for _ in range(3):
        st = d.iterate()
        assert st["not_converged"] == 0
        iters += st["solve_iters"]
    return d.get("o"), iters


@pytest.mark.parametrize("precision", [0, 1])
def test_first_order_start_matches_oracle_and_saves_iterations(case, precision):
    isg, otrue, o_ref = case
    print("first-order field start, precision", precision, flush=True)
    o_prev, it_prev = run(isg, otrue, 1, precision)
    o_lin, it_lin = run(isg, otrue, 2, precision)
    assert relerr(o_prev, o_ref) < 1e-4
    assert relerr(o_lin, o_ref) < 1e-4
    assert it_lin < it_prev, (it_lin, it_prev)
