"""Pin the oracle's DBIM loop (mom.cpp:222-400).

  * the gradient od is 1/(2 k0^2) times the gradient of the misfit
    F(o) = sum_t |aggregate(k0^2 o phi_t(o)) - mes_t|^2 (phi_t the total field):
    checked by central finite differences along random directions;
  * the line-search step alpha minimises the linearised misfit along ot;
  * the misfit falls every iteration on the 64x64 configuration.
"""
import numpy as np
import pytest

from conftest import oracle_for, relerr


def misfit(o, otree, inc, mes, ntx, tol=1e-12):
    k0 = o.k0
    g = otree * k0 * k0
    F = 0.0
    for t in range(ntx):
        phi, _ = o.bicgs(g, inc[t * o.N:(t + 1) * o.N], tol=tol, maxit=1000)
        F += np.sum(np.abs(o.aggregate(g * phi) - mes[t * o.numrx:(t + 1) * o.numrx]) ** 2)
    return F


def test_gradient_finite_differences():
    from oracle.oracle import Dbim
    o = oracle_for(3, numrx=16)
    ntx = 4
    pos = o.pos()
    otrue = 0.2 * np.exp(-np.abs(pos - (0.5 + 0.3j)) ** 2 / 0.8) * (1 + 0.1j)
    d = Dbim(o, ntx, otrue, regular=0.0, tol=1e-12, maxit=1000)
    d.iterate()  # move away from o = 0
    ocur, od = d.get("o"), d.get("od")
    inc, mes = d.get("inc"), d.get("mes")
    rng = np.random.default_rng(0)
    for _ in range(2):
        dirn = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)) * 0.05
        eps = 1e-4
        fd = (misfit(o, ocur + eps * dirn, inc, mes, ntx) - misfit(o, ocur - eps * dirn, inc, mes, ntx)) / (2 * eps)
        an = 2 * o.k0 ** 2 * np.real(np.vdot(od, dirn))
        assert abs(fd - an) < 1e-5 * abs(an), (fd, an)


def test_misfit_decreases_64():
    from oracle.oracle import Dbim
    import isg
    o = oracle_for(4, numrx=16)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    d = Dbim(o, 16, otrue, regular=float(np.sqrt(1e-3)), tol=1e-6, maxit=300)
    hist = [d.hist()[0]]
    for _ in range(3):
        alpha = d.iterate()
        assert alpha > 0
        hist.append(d.hist()[0])
    assert np.all(np.diff(hist) < 0), hist
    assert hist[-1] < 0.1


def test_first_order_field_start():
    """The first-order total-field start used by the DBIM driver: with
    A(o) = I - G diag(k0^2 o), tot(o) = A^-1 inc and the derivative field
    dtot = A^-1 G(k0^2 ot tot) (mom.cpp:284-290), tot(o - a ot) equals
    tot(o) - a dtot up to O(a^2): halving a quarters the error, and the error is
    far below the change of the field itself."""
    o = oracle_for(3, numrx=16)
    k0 = o.k0
    pos = o.pos()
    rng = np.random.default_rng(5)
    ocur = 0.15 * np.exp(-np.abs(pos - (0.3 - 0.2j)) ** 2 / 0.5) * (1 + 0.1j)
    ot = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)) * 0.05
    inc = np.exp(1j * k0 * pos.real)
    tot, _ = o.bicgs(ocur * k0 * k0, inc, tol=1e-13, maxit=1000)
    dtot, _ = o.bicgs(ocur * k0 * k0, o.mlfma(k0 * k0 * ot * tot), tol=1e-13, maxit=1000)
    errs = []
    for a in (0.2, 0.1):
        exact, _ = o.bicgs((ocur - a * ot) * k0 * k0, inc, tol=1e-13, maxit=1000)
        err = np.linalg.norm(exact - (tot - a * dtot))
        assert err < 0.05 * np.linalg.norm(exact - tot), (a, err)
        errs.append(err)
    assert 3.0 < errs[0] / errs[1] < 5.0, errs
