"""The reference's file formats (setup_born.cpp:105-121, 171-242; mom.cpp:220).

CPU only: phantom text round trip (the reader multiplies by the scale the
reference hard-codes, 0.1), hist.txt lines byte-identical to C's "%e" format
of the five metrics, raw raster files.  The cg_<k>.bin path through a DBIM
context is covered in tests/test_gpu_dbim.py.
"""
import numpy as np

import isg


def test_phantom_text_round_trip(tmp_path):
    side = 64
    ph = isg.phantom_ellipses(side, 5, 7)
    p = tmp_path / "phantom_64.txt"
    isg.write_phantom_txt(p, ph, scale=0.1)
    lines = p.read_text().splitlines()
    assert len(lines) == side * side
    assert lines[0] == "%f" % (ph.real.ravel()[0] / 0.1)
    back = isg.read_phantom_txt(p, side, scale=0.1)
    np.testing.assert_allclose(back.real, ph.real, atol=1e-6 * 0.1 + 1e-7)
    assert np.all(back.imag == 0)


def test_hist_lines(tmp_path):
    p = tmp_path / "hist.txt"
    rows = [dict(misfit=0.5, object_error=0.25, object_cost=1e-3, gradient_norm2=12.5, object_norm2=3.0),
            dict(misfit=0.1234567, object_error=2e-10, object_cost=0.0, gradient_norm2=1e20, object_norm2=7.0)]
    isg.hist_append(p, rows[0], create=True)
    isg.hist_append(p, rows[1])
    want = "".join("%e %e %e %e %e\n" % (r["misfit"], r["object_error"], r["object_cost"], r["gradient_norm2"],
                                         r["object_norm2"]) for r in rows)
    assert p.read_text() == want


def test_raster_bin_round_trip(tmp_path):
    side = 32
    rng = np.random.default_rng(0)
    v = rng.standard_normal((side, side)) + 1j * rng.standard_normal((side, side))
    p = tmp_path / "cg_3.bin"
    lib = isg.load()
    a = np.ascontiguousarray(v.ravel())
    assert lib.isg_raster_bin_write(str(p).encode(), side * side, a.ctypes.data_as(isg._D)) == 0
    assert p.stat().st_size == 16 * side * side
    np.testing.assert_array_equal(isg.read_raster_bin(p, side), v)
    np.testing.assert_array_equal(np.fromfile(p, dtype=np.complex128).reshape(side, side), v)
