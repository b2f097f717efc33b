"""GPU DBIM against the CPU oracle on the 64x64 configuration.

64x64 pixels of lambda/10 (level 4), 16 plane-wave transmitters, 16 far-field
receivers, 3 seeded Gaussian blobs (seed 1), BiCGSTAB tolerance 1e-6,
regular = sqrt(1e-3) (Tikhonov alpha = 1e-3), 3 DBIM iterations.
Bar: reconstructed contrast within 1e-4 relative of the oracle's after the
fixed number of iterations; misfit history within 1e-4 relative.
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
    phantom = isg.phantom_blobs(o.side, 3, 1)
    otrue = isg.raster_to_tree(phantom, o.unkmap())
    ref = OrcDbim(o, NTX, otrue, regular=REG, tol=TOL, maxit=300)
    hist_ref = [ref.hist()]
    for _ in range(3):
        ref.iterate()
        hist_ref.append(ref.hist())
    return isg, otrue, ref, np.array(hist_ref)


@pytest.mark.parametrize("batch,precision,field_start,solve_start,margin",
                         [(1, 0, 0, 0, 0.5), (16, 0, 0, 0, 0.5), (6, 0, 0, 0, 0.5), (6, 1, 2, 1, 0.9)])
def test_dbim_matches_oracle(case, batch, precision, field_start, solve_start, margin):
    """batch 6 leaves a ragged last batch (6 + 6 + 4 transmitters); the last
    case also runs every option of the fast path on it (mixed precision with
    the benchmark's solve margin, first-order field start, previous-solution
    adjoint/derivative starts)."""
    isg, otrue, ref, hist_ref = case
    print("dbim parity batch", batch, "precision", precision, "starts", field_start, solve_start, "margin", margin,
          flush=True)
    s = isg.Solver(level=LEVEL, numrx=NRX, numtx=NTX, toli=TOL, iti=300, regular=REG, max_rhs=batch,
                   precision=precision, solve_start=solve_start, mixed_margin=margin)
    d = isg.Dbim(s, otrue, field_start=field_start)
    st0 = d.setup_stats
    hist = [[st0["misfit"], st0["object_error"], st0["object_cost"], st0["gradient_norm2"], st0["object_norm2"]]]
    for _ in range(3):
        st = d.iterate()
        assert st["not_converged"] == 0
        hist.append([st["misfit"], st["object_error"], st["object_cost"], st["gradient_norm2"], st["object_norm2"]])
    hist = np.array(hist)
    assert relerr(d.get("mes"), ref.get("mes")) < 1e-5
    assert relerr(d.get("o"), ref.get("o")) < 1e-4
    np.testing.assert_allclose(hist[:, [0, 1, 3, 4]], hist_ref[:, [0, 1, 3, 4]], rtol=1e-4, atol=0)
    # the misfit must fall every iteration (mom.cpp:277-400 is a descent method)
    assert np.all(np.diff(hist[:, 0]) < 0)


def test_dbim_checkpoint_resume(case):
    """3 iterations == 2 iterations + checkpoint {o, od, ot, mesd, tot} + resume + 1."""
    isg, otrue, ref, _ = case
    kw = dict(level=LEVEL, numrx=NRX, numtx=NTX, toli=TOL, iti=300, regular=REG, max_rhs=16)
    a = isg.Dbim(isg.Solver(**kw), otrue)
    for _ in range(2):
        a.iterate()
    ck = {k: a.get(k) for k in ("o", "od", "ot", "mesd", "tot")}
    a.iterate()
    b = isg.Dbim(isg.Solver(**kw), otrue)
    for k, v in ck.items():
        b.set(k, v)
    b.iterate()
    assert relerr(b.get("o"), a.get("o")) < 1e-12


@pytest.mark.parametrize("field_start,solve_start", [(0, 0), (2, 1)])
def test_dbim_restart_repeats_iterations(case, field_start, solve_start):
    """isg_dbim_restart returns to the state after setup: iterations 1 and 2
    after a restart reproduce the first run's iterations 1 and 2 bitwise
    (also with the first-order start and the previous-solution starts, whose
    stores the restart drops)."""
    isg, otrue, ref, _ = case
    kw = dict(level=LEVEL, numrx=NRX, numtx=NTX, toli=TOL, iti=300, regular=REG, max_rhs=16, precision=1,
              solve_start=solve_start)
    d = isg.Dbim(isg.Solver(**kw), otrue, field_start=field_start)
    first = []
    for _ in range(2):
        st = d.iterate()
        first.append((d.get("o"), st["misfit"], st["solve_iters"]))
    d.iterate()
    d.restart()
    for k in range(2):
        st = d.iterate()
        assert np.array_equal(d.get("o"), first[k][0]), k
        assert st["misfit"] == first[k][1] and st["solve_iters"] == first[k][2], k


def test_dbim_checkpoint_resume_previous_solution_start(case):
    """With solve_start 1 a restored state drops the stored adjoint and
    derivative solutions (they belong to another history), so the resumed
    iteration starts those solves from zero: the same contrast as the
    uninterrupted run to the solves' tolerance, not bitwise."""
    isg, otrue, ref, _ = case
    print("checkpoint/resume with previous-solution starts", flush=True)
    kw = dict(level=LEVEL, numrx=NRX, numtx=NTX, toli=TOL, iti=300, regular=REG, max_rhs=16, solve_start=1)
    a = isg.Dbim(isg.Solver(**kw), otrue)
    for _ in range(2):
        a.iterate()
    ck = {k: a.get(k) for k in ("o", "od", "ot", "mesd", "tot")}
    a.iterate()
    b = isg.Dbim(isg.Solver(**kw), otrue)
    for k, v in ck.items():
        b.set(k, v)
    st = b.iterate()
    assert st["not_converged"] == 0
    e = relerr(b.get("o"), a.get("o"))
    assert e < 1e-5, e


def test_cg_files(case, tmp_path):
    """cg_<k>.bin / cg_ref.bin are the raster-order gathers of the contrast
    (setup_born.cpp:115-121, 209-218) and read back as a restart."""
    isg, otrue, _, _ = case
    kw = dict(level=LEVEL, numrx=NRX, numtx=NTX, toli=TOL, iti=300, regular=REG, max_rhs=16)
    s = isg.Solver(**kw)
    _, r2t = s.geometry()
    side = 8 << (LEVEL - 1)
    d = isg.Dbim(s, otrue)
    d.write_cg(tmp_path / "cg_ref.bin", "true")
    st = d.iterate()
    isg.hist_append(tmp_path / "hist.txt", st, create=True)
    d.write_cg(tmp_path / "cg_1.bin")
    np.testing.assert_array_equal(isg.read_raster_bin(tmp_path / "cg_ref.bin", side),
                                  isg.tree_to_raster(otrue, r2t, side))
    np.testing.assert_array_equal(isg.read_raster_bin(tmp_path / "cg_1.bin", side),
                                  isg.tree_to_raster(d.get("o"), r2t, side))
    e = isg.Dbim(isg.Solver(**kw), otrue)
    e.read_cg(tmp_path / "cg_1.bin")
    np.testing.assert_array_equal(e.get("o"), d.get("o"))
    assert len((tmp_path / "hist.txt").read_text().splitlines()) == 1


def test_synthesis_progress_reports_each_batch(case):
    """isg_set_progress: the measurement synthesis calls back once per batch of
    transmitters (16 transmitters in batches of 6: 6, 12, 16 of 16), and the
    measurements equal those of a run without the callback bitwise."""
    isg, otrue, _, _ = case
    kw = dict(level=LEVEL, numrx=NRX, numtx=NTX, toli=TOL, iti=300, regular=REG, max_rhs=6)
    s = isg.Solver(**kw)
    calls = []
    s.set_progress(lambda done, total: calls.append((done, total)))
    d = isg.Dbim(s, otrue)
    assert calls == [(6, NTX), (12, NTX), (16, NTX)]
    plain = isg.Dbim(isg.Solver(**kw), otrue)
    np.testing.assert_array_equal(d.get("mes"), plain.get("mes"))
