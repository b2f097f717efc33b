"""Pin the oracle's operators and solver against independent computations.

  mlfma       vs the dense direct MoM product (the matrix the near blocks sample)
  aggregate   vs the closed-form far field sum_n res^2 exp(-i k0 u.r_n) x_n
  aggregateh  adjoint of aggregate: <aggregate x, y> = <x, aggregateh y>
  mlfmah      conj(G conj x)
  bicgs       vs numpy.linalg.solve on the dense MLFMA operator
"""
import numpy as np
import pytest

from conftest import oracle_for, rand_c, relerr


@pytest.mark.parametrize("level", [3, 4])
def test_mlfma_vs_direct_mom(level):
    o = oracle_for(level)
    rng = np.random.default_rng(level)
    x = rand_c(rng, o.N)
    err = relerr(o.mlfma(x), o.direct(x))
    # beta = 5 sampling: MLFMA error well below 1e-5, but not roundoff
    assert 1e-12 < err < 1e-5, err


def test_mlfma_vs_direct_window_128():
    """128x128 grid, direct product on a 16-row window of targets."""
    o = oracle_for(5)
    rng = np.random.default_rng(5)
    x = rand_c(rng, o.N)
    rows = slice(2048, 2048 + 1024)
    err = relerr(o.mlfma(x)[rows], o.direct(x, 2048, 1024))
    assert err < 1e-5, err


def test_aggregate_closed_form_and_adjoint():
    o = oracle_for(4, numrx=24)
    rng = np.random.default_rng(3)
    x = rand_c(rng, o.N)
    y = rand_c(rng, 24)
    pos = o.pos()
    ang = 2 * np.pi * np.arange(24) / 24
    ff = np.array([np.sum(o.res ** 2 * np.exp(-1j * o.k0 * (np.cos(a) * pos.real + np.sin(a) * pos.imag)) * x)
                   for a in ang])
    ax = o.aggregate(x)
    assert relerr(ax, ff) < 1e-12
    lhs = np.vdot(y, ax)
    rhs = np.vdot(o.aggregateh(y), x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_mlfmah_identity():
    o = oracle_for(4)
    rng = np.random.default_rng(4)
    x = rand_c(rng, o.N)
    assert relerr(o.mlfmah(x), np.conj(o.mlfma(np.conj(x)))) < 1e-15


@pytest.mark.parametrize("herm", [False, True])
def test_bicgs_vs_dense_solve(herm):
    o = oracle_for(3)
    N = o.N
    G = np.empty((N, N), complex)
    e = np.zeros(N, complex)
    for j in range(N):
        e[:] = 0
        e[j] = 1
        G[:, j] = o.mlfma(e)
    pos = o.pos()
    g = 0.3 * np.exp(-np.abs(pos) ** 2 / 2.0) * (1 + 0.1j) * o.k0 ** 2
    A = np.eye(N) - G @ np.diag(g)
    if herm:
        A = np.eye(N) - np.diag(np.conj(g)) @ G.conj()
    b = np.exp(1j * o.k0 * pos.real)
    x_ref = np.linalg.solve(A, b)
    x, info = o.bicgs(g, b, herm=herm, tol=1e-10, maxit=500)
    assert info["status"] == 0
    assert relerr(x, x_ref) < 1e-8
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9


def test_bicgs_zero_rhs_and_iteration_limit():
    """b = 0 exits at once (bicgs_gpu.cu:148-150); max_it = 1 reports non-convergence."""
    o = oracle_for(3)
    pos = o.pos()
    g = 0.3 * np.exp(-np.abs(pos) ** 2 / 2.0) * o.k0 ** 2
    x, info = o.bicgs(g, np.zeros(o.N, complex), tol=1e-6)
    assert info["iters"] == 0 and np.all(x == 0)
    x, info = o.bicgs(g, np.exp(1j * o.k0 * pos.real), tol=1e-12, maxit=1)
    assert info["status"] == 1 and info["iters"] == 1
