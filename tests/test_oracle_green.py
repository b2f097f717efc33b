"""Pin the oracle's Green's-function pixel integral and Hankel functions.

integrate.cpp:15-41 and :7-9.  Off the diagonal the reference uses the
midpoint rule (i/4) res^2 H0^(1)(k0 d); on the diagonal it integrates the
small-argument model J0 + i(2/pi)(ln(k0 r/2) + gamma) analytically over the
pixel.  Both are checked against independent evaluations (scipy Hankel
functions; 2-D Gauss-Legendre quadrature of the model with the log
singularity split off by polar integration), so a sign, factor or log-term
slip fails here.
"""
import numpy as np
import pytest
from scipy import special

from conftest import oracle_for


@pytest.fixture(scope="module")
def orc():
    return oracle_for(3)


def test_hankel_matches_scipy():
    from oracle.oracle import lib
    import ctypes
    L = lib()
    out = np.empty(2)
    for order in (0, 1, 2, 5, 17, 40):
        for x in (0.3, 1.0, 7.5, 30.0, 120.0):
            L.orc_hn(order, x, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            ref = special.hankel1(order, x)
            assert abs(complex(out[0], out[1]) - ref) <= 1e-12 * max(1.0, abs(ref)), (order, x)


@pytest.mark.parametrize("dx,dy", [(1, 0), (0, 3), (2, -5), (7, 7), (15, 0)])
def test_offdiagonal_is_midpoint_hankel(orc, dx, dy):
    res, k0 = orc.res, orc.k0
    pt = complex(0.05, -0.05)
    pb = pt + complex(dx * res, -dy * res)
    d = abs(pb - pt)
    ref = 0.25j * res * res * special.hankel1(0, k0 * d)
    assert abs(orc.green(pt, pb) - ref) < 1e-13 * abs(ref)


def test_self_term_matches_quadrature(orc):
    """(i/4) [res^2 J0(0) + i (2/pi) integral over the pixel of (ln(k r / 2) + gamma)]:
    the regular part at the midpoint, the logarithm integrated (integrate.cpp:22-35)."""
    res, k0 = orc.res, orc.k0
    h = res / 2
    smooth = res * res * special.j0(0.0) + 1j * (2 / np.pi) * (np.log(k0 / 2) + np.euler_gamma) * res * res
    # log part by polar integration over the square (the r-integral of r ln r
    # is exact in closed form, the angle integral by Gauss-Legendre)
    # int_square ln r dA = 8 * int_0^{pi/4} int_0^{h/cos t} r ln r dr dt
    tg, tw = np.polynomial.legendre.leggauss(60)
    t = (tg + 1) * np.pi / 8
    R = h / np.cos(t)
    inner = R ** 2 / 2 * np.log(R) - R ** 2 / 4
    logint = 8 * np.sum(tw * inner) * np.pi / 8
    ref = 0.25j * (smooth + 1j * (2 / np.pi) * logint)
    got = orc.green(complex(0.05, -0.05), complex(0.05, -0.05))
    assert abs(got - ref) < 1e-10 * abs(ref)
