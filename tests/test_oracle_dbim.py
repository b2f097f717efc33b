"""Pin the oracle's DBIM loop (mom.cpp:222-400).

  * the gradient od is 1/(2 k0^2) times the gradient of the regularised misfit
    F(o) = sum_t |aggregate(k0^2 o phi_t(o)) - mes_t|^2 + k0^2 numtx regular^2 |o|^2
    (phi_t the total field; the reference adds regular^2 o once per
    transmitter, mom.cpp:267-269, :357-360): central finite differences along
    random directions, with and without the regulariser;
  * the step length alpha (mom.cpp:304-316) minimises |mesd - alpha J ot|^2,
    J the Frechet derivative of the forward map o -> aggregate(k0^2 o phi(o))
    built here by central finite differences of full forward solves, and the
    nonlinear misfit is higher on both sides of it;
  * Polak-Ribiere (mom.cpp:383-399): in the Born limit (a contrast so weak
    that the misfit is quadratic) nonlinear CG with this alpha is linear CG on
    the normal equations, so successive directions are J^H J-conjugate;
  * each of these checks fails when a scale error is planted in the oracle's
    alpha, beta or regulariser term (Dbim.mutate);
  * the misfit falls every iteration on the 64x64 configuration.
Left open: where the conjugate sits in beta's numerator (od conj(od - save)
vs conj(od) (od - save)); the two differ only in the imaginary part, which a
quadratic misfit leaves zero.
"""
import numpy as np
import pytest

from conftest import oracle_for, relerr


def forward_data(o, otree, inc, ntx, tol=1e-13):
    """Synthetic data of contrast otree for every transmitter: an independent
    forward solve (tight tolerance) and the receiver operator."""
    k0 = o.k0
    g = otree * k0 * k0
    out = []
    for t in range(ntx):
        phi, _ = o.bicgs(g, inc[t * o.N:(t + 1) * o.N], tol=tol, maxit=2000)
        out.append(o.aggregate(g * phi))
    return np.concatenate(out)


def misfit(o, otree, inc, mes, ntx, regular=0.0, tol=1e-12):
    F = float(np.sum(np.abs(forward_data(o, otree, inc, ntx, tol) - mes) ** 2))
    return F + o.k0 ** 2 * ntx * regular ** 2 * float(np.vdot(otree, otree).real)


def frechet(o, otree, v, inc, ntx, scale):
    """J v by central differences of the forward map; the step is 1e-3 of
    `scale` (the contrast's size) along v."""
    eps = 1e-3 * scale / np.linalg.norm(v)
    return (forward_data(o, otree + eps * v, inc, ntx) - forward_data(o, otree - eps * v, inc, ntx)) / (2 * eps)


def two_blobs(o, amp):
    pos = o.pos()
    return (amp * np.exp(-np.abs(pos - (0.5 + 0.3j)) ** 2 / 0.8) * (1 + 0.1j)
            + 0.7 * amp * np.exp(-np.abs(pos - (-0.6 - 0.4j)) ** 2 / 0.3))


def gradient_fd_error(regular, reg_scale=1.0):
    """Largest relative gap between 2 k0^2 Re<od, v> and the central difference
    of the regularised misfit along random v, after one DBIM iteration."""
    from oracle.oracle import Dbim
    o = oracle_for(3, numrx=16)
    ntx = 4
    pos = o.pos()
    otrue = 0.2 * np.exp(-np.abs(pos - (0.5 + 0.3j)) ** 2 / 0.8) * (1 + 0.1j)
    d = Dbim(o, ntx, otrue, regular=regular, tol=1e-12, maxit=1000)
    d.mutate(reg=reg_scale)
    d.iterate()  # move away from o = 0; the gradient is recomputed with the planted scale
    ocur, od = d.get("o"), d.get("od")
    inc, mes = d.get("inc"), d.get("mes")
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(2):
        dirn = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)) * 0.05
        eps = 1e-4
        fd = (misfit(o, ocur + eps * dirn, inc, mes, ntx, regular)
              - misfit(o, ocur - eps * dirn, inc, mes, ntx, regular)) / (2 * eps)
        an = 2 * o.k0 ** 2 * np.real(np.vdot(od, dirn))
        worst = max(worst, abs(fd - an) / abs(an))
    return worst


def test_gradient_finite_differences():
    assert gradient_fd_error(0.0) < 1e-5


def test_gradient_finite_differences_regularised():
    """regular = 0.3: the Tikhonov term carries about a fifth of the directional
    derivative, so a wrong factor (regular instead of regular^2, or a missing
    factor numtx) moves it far outside the bar."""
    assert gradient_fd_error(0.3) < 1e-5
    assert gradient_fd_error(0.3, reg_scale=1.0 / 4) > 1e-2   # added once, not per transmitter
    assert gradient_fd_error(0.3, reg_scale=1.0 / 0.3) > 1e-2  # regular instead of regular^2


def step_length_check(alpha_scale=1.0):
    """Second DBIM iteration (o != 0, a Polak-Ribiere direction) at 32x32:
    the oracle's alpha against the minimiser of |mesd - a J ot|^2 with J from
    finite differences; returns (relative gap, nonlinear misfits at
    alpha (1 - 0.3, 1, 1 + 0.3))."""
    from oracle.oracle import Dbim
    o = oracle_for(3, numrx=16)
    ntx = 4
    otrue = two_blobs(o, 0.05)
    d = Dbim(o, ntx, otrue, regular=0.0, tol=1e-13, maxit=2000)
    d.iterate()
    d.mutate(alpha=alpha_scale)
    ocur, ot, mesd = d.get("o"), d.get("ot"), d.get("mesd")
    inc, mes = d.get("inc"), d.get("mes")
    alpha = d.iterate()
    J = frechet(o, ocur, ot, inc, ntx, np.linalg.norm(ocur))
    a_fd = np.real(np.vdot(J, mesd)) / np.real(np.vdot(J, J))
    F = [misfit(o, ocur - a * alpha * ot, inc, mes, ntx) for a in (0.7, 1.0, 1.3)]
    return abs(alpha - a_fd) / abs(a_fd), F


def test_step_length_minimises_linearised_misfit():
    gap, F = step_length_check()
    assert gap < 1e-6, gap
    assert F[1] < F[0] and F[1] < F[2], F


def test_step_length_pin_detects_scale_error():
    gap, _ = step_length_check(alpha_scale=1.01)
    assert gap > 1e-3, gap


def conjugacy(beta_scale=1.0):
    """Born limit (contrast 1e-6): |<J ot_1, J ot_0>| / (|J ot_1| |J ot_0|) for
    the first two directions of the oracle's loop."""
    from oracle.oracle import Dbim
    o = oracle_for(3, numrx=16)
    ntx = 4
    otrue = two_blobs(o, 1e-6)
    d = Dbim(o, ntx, otrue, regular=0.0, tol=1e-13, maxit=2000)
    d.mutate(beta=beta_scale)
    ot0, inc = d.get("ot"), d.get("inc")
    d.iterate()
    o1, ot1 = d.get("o"), d.get("ot")
    scale = np.linalg.norm(otrue)
    J0 = frechet(o, o1, ot0, inc, ntx, scale)
    J1 = frechet(o, o1, ot1, inc, ntx, scale)
    return abs(np.vdot(J1, J0)) / (np.linalg.norm(J0) * np.linalg.norm(J1))


def test_polak_ribiere_conjugate_in_born_limit():
    assert conjugacy() < 1e-4


def test_polak_ribiere_pin_detects_scale_error():
    assert conjugacy(beta_scale=1.2) > 1e-2
    assert conjugacy(beta_scale=0.8) > 1e-2


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
