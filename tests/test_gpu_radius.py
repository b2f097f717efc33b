"""Finite-radius geometry on the GPU against the oracle: line-source
transmitters and receivers on a circle of 10 wavelengths (the 64x64
configuration's geometry), 16 transmitters / 16 receivers, 3 seeded blobs.

Bars: receiver operator and its adjoint 1e-12 (dense FP64 sums, different
summation order); incident fields 1e-12; DBIM contrast after 3 iterations
1e-4 (as for the reference geometry); mixed-precision batched DBIM 1e-4.
"""
import numpy as np
import pytest

from conftest import oracle_for, rand_c, relerr

pytestmark = pytest.mark.gpu

LEVEL, NTX, NRX, TOL, REG, RADIUS = 4, 16, 16, 1e-6, float(np.sqrt(1e-3)), 10.0


@pytest.fixture(scope="module")
def isg():
    import isg as m
    return m


def test_receivers_match_oracle(isg):
    print("finite-radius receivers, level 5", flush=True)
    o = oracle_for(5, numrx=NRX, radius=RADIUS)
    s = isg.Solver(level=5, numrx=NRX, radius=RADIUS)
    x = rand_c(np.random.default_rng(1), o.N)
    assert relerr(s.aggregate(x), o.aggregate(x)) < 1e-12
    r = rand_c(np.random.default_rng(2), NRX)
    assert relerr(s.aggregateh(r), o.aggregateh(r)) < 1e-12


@pytest.mark.parametrize("batch,precision", [(1, 0), (16, 1), (6, 1)])
def test_dbim_matches_oracle(isg, batch, precision):
    from oracle.oracle import Dbim as OrcDbim
    print("finite-radius DBIM batch", batch, "precision", precision, flush=True)
    o = oracle_for(LEVEL, numrx=NRX, radius=RADIUS)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    ref = OrcDbim(o, NTX, otrue, regular=REG, tol=TOL, maxit=300)
    s = isg.Solver(level=LEVEL, numrx=NRX, numtx=NTX, toli=TOL, iti=300, regular=REG, max_rhs=batch,
                   precision=precision, radius=RADIUS)
    d = isg.Dbim(s, otrue, field_start=2 if batch > 1 else 0)
    assert relerr(d.get("inc"), ref.get("inc")) < 1e-12
    assert relerr(d.get("mes"), ref.get("mes")) < 1e-5
    for _ in range(3):
        ref.iterate()
        st = d.iterate()
        assert st["not_converged"] == 0
    assert relerr(d.get("o"), ref.get("o")) < 1e-4
