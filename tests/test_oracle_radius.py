"""Pin the oracle's finite-radius geometry (line-source transmitters, point
receivers on a circle): the DBIM configurations quote 16 transmitters and 16
receivers on a circle of 10 wavelengths at 64x64 pixels.

  * receiver operator against an independent numpy/scipy direct sum of
    (i/4) H0^(1)(k0 |rho_m - r_n|) res^2 x_n;
  * back-projection = exact adjoint: <A x, r> = <x, A^H r>;
  * incident field = (i/4) H0^(1)(k0 |r - rho_t|) (scipy);
  * far-field limit: on a circle of 2e4 wavelengths the receiver operator
    equals the far-field pattern of the reference geometry times the
    asymptotic Hankel factor (i/4) sqrt(2 / (pi k0 R)) exp(i (k0 R - pi/4));
  * DBIM gradient by central finite differences (the misfit now built from
    the direct-sum receivers and line sources).
"""
import numpy as np
import pytest
from scipy.special import hankel1

from conftest import oracle_for, rand_c


def circle(R, count):
    a = 2 * np.pi * np.arange(count) / count
    return R * (np.cos(a) + 1j * np.sin(a))


def test_receivers_direct_sum_and_adjoint():
    o = oracle_for(3, numrx=16, radius=10.0)
    x = rand_c(np.random.default_rng(1), o.N)
    r = o.aggregate(x)
    rho = circle(10.0, 16)
    K = 0.25j * hankel1(0, o.k0 * np.abs(rho[:, None] - o.pos()[None, :])) * o.res ** 2
    np.testing.assert_allclose(r, K @ x, rtol=1e-12, atol=1e-14)
    rr = rand_c(np.random.default_rng(2), 16)
    lhs = np.vdot(rr, o.aggregate(x))
    rhs = np.vdot(o.aggregateh(rr), x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_line_source_incident_field():
    from oracle.oracle import Dbim
    o = oracle_for(3, numrx=16, radius=10.0)
    d = Dbim(o, 4, np.zeros(o.N, complex), tol=1e-8)
    inc = d.get("inc").reshape(4, o.N)
    rho = circle(10.0, 4)
    ref = 0.25j * hankel1(0, o.k0 * np.abs(o.pos()[None, :] - rho[:, None]))
    np.testing.assert_allclose(inc, ref, rtol=1e-12, atol=1e-15)


def test_far_field_limit():
    """The far-field pattern of the reference geometry is the coefficient of
    the outgoing cylindrical wave: on a circle of R = 2e4 wavelengths the
    finite-radius receivers equal it times the large-argument form of
    (i/4) H0^(1)(k0 R), up to the O(k0 D^2 / R) phase error of a 3.2-wavelength
    domain and the interpolation error of the pattern."""
    R = 2.0e4
    far = oracle_for(3, numrx=16)
    fin = oracle_for(3, numrx=16, radius=R)
    x = rand_c(np.random.default_rng(3), far.N)
    a = fin.aggregate(x)
    k = far.k0
    h = 0.25j * np.sqrt(2 / (np.pi * k * R)) * np.exp(1j * (k * R - np.pi / 4))
    b = h * far.aggregate(x)
    assert np.linalg.norm(a - b) < 2e-3 * np.linalg.norm(b)


def misfit(o, otree, inc, mes, ntx, tol=1e-12):
    k0 = o.k0
    g = otree * k0 * k0
    F = 0.0
    for t in range(ntx):
        phi, _ = o.bicgs(g, inc[t * o.N:(t + 1) * o.N], tol=tol, maxit=1000)
        F += np.sum(np.abs(o.aggregate(g * phi) - mes[t * o.numrx:(t + 1) * o.numrx]) ** 2)
    return F


def test_gradient_finite_differences_radius():
    from oracle.oracle import Dbim
    o = oracle_for(3, numrx=16, radius=10.0)
    ntx = 4
    pos = o.pos()
    otrue = 0.2 * np.exp(-np.abs(pos - (0.5 + 0.3j)) ** 2 / 0.8) * (1 + 0.1j)
    d = Dbim(o, ntx, otrue, regular=0.0, tol=1e-12, maxit=1000)
    d.iterate()
    ocur, od = d.get("o"), d.get("od")
    inc, mes = d.get("inc"), d.get("mes")
    rng = np.random.default_rng(1)
    dirn = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)) * 0.05
    eps = 1e-4
    fd = (misfit(o, ocur + eps * dirn, inc, mes, ntx) - misfit(o, ocur - eps * dirn, inc, mes, ntx)) / (2 * eps)
    an = 2 * o.k0 ** 2 * np.real(np.vdot(od, dirn))
    assert abs(fd - an) < 1e-5 * abs(an), (fd, an)
