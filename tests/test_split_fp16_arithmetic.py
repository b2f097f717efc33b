"""The split-FP16 arithmetic of the FP32 operator's near field, L2P and P2M
kernels, restated in numpy (CPU): operands scaled by a power of two so that
the largest component lies in [2^13, 2^14), split into hi = rne16(x) and
lo = rne16(x - hi), products formed as hi.hi + hi.lo + lo.hi (accumulated
here in float64; the kernels' FP32 accumulation adds about k 2^-24 relative,
below the bar), complex products in the four-product form.  Checks:
  * the scale exponent puts every component below 2^14 (FP16's largest
    finite value is 65504 < 2^16) and the largest one at or above 2^13;
  * hi + lo equals x to 2^-22 of |x| once the scaled component is at least
    2^-3 (at least 2^-16 of the largest), and to 2^-25 in scaled units (at
    most 2^-38 of the largest) below that, where lo is subnormal;
  * a 64 x 64 complex block product (the near field's slot operator) is
    within 1e-6 relative (norm) of the float64 product for operands spanning
    eight decades, and a dropped lo.hi term (a plausible slip) fails the bar,
    as a data scale three powers of two too large overflows FP16 into Inf.
"""
import numpy as np


def scale_exp(m):
    """Exponent e with m 2^e in [2^13, 2^14) (0 for m = 0 or not finite),
    clamped to [-126, 126] -- as h16_scale_exp."""
    if not (m > 0) or not np.isfinite(m):
        return 0
    _, e = np.frexp(np.float32(m))  # m = f 2^e, f in [0.5, 1)
    return int(np.clip(14 - e, -126, 126))


def split(x):
    x = np.asarray(x, np.float32)
    hi = x.astype(np.float16)
    lo = (x - hi.astype(np.float32)).astype(np.float16)
    return hi, lo


def split_matmul(A, B, drop_cross=False):
    """Complex A (m x k) @ B (k x n) the way the kernels do it."""
    ea = scale_exp(max(np.abs(A.real).max(), np.abs(A.imag).max()))
    eb = scale_exp(max(np.abs(B.real).max(), np.abs(B.imag).max()))
    sa, sb = np.float32(2.0 ** ea), np.float32(2.0 ** eb)
    parts = {}
    for name, M, s in (("a", A, sa), ("b", B, sb)):
        for comp, v in (("r", M.real), ("i", M.imag)):
            h, lo = split(np.float32(v) * s)
            parts[name + comp] = (h.astype(np.float64), lo.astype(np.float64))

    def prod(x, y):
        (xh, xl), (yh, yl) = parts[x], parts[y]
        p = xh @ yh + xh @ yl
        if not drop_cross:
            p = p + xl @ yh
        return p

    cr = prod("ar", "br") - prod("ai", "bi")
    ci = prod("ar", "bi") + prod("ai", "br")
    return (cr + 1j * ci) / (float(sa) * float(sb))


def test_scale_exponent_range():
    rng = np.random.default_rng(0)
    for m in np.concatenate([10.0 ** rng.uniform(-30, 30, 200), [1.0, 2.0 ** 13, 2.0 ** 14, 65504.0]]):
        e = scale_exp(m)
        v = np.float32(m) * np.float32(2.0 ** e)
        assert 2.0 ** 13 <= v < 2.0 ** 14, (m, e, v)
    assert scale_exp(0.0) == 0 and scale_exp(np.inf) == 0 and scale_exp(np.nan) == 0


def test_split_precision():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(100000) * 10.0 ** rng.uniform(-12, 0, 100000)
    e = scale_exp(np.abs(x).max())
    xs = np.float32(x) * np.float32(2.0 ** e)
    hi, lo = split(xs)
    back = hi.astype(np.float64) + lo.astype(np.float64)
    err = np.abs(back - xs.astype(np.float64))
    big = np.abs(xs) >= 2.0 ** -3
    assert np.all(err[big] <= 2.0 ** -22 * np.abs(xs[big]) * 1.0001)
    assert np.all(err[~big] <= 2.0 ** -25 * 1.0001)
    assert big.any() and (~big).any()
    assert np.all(np.isfinite(hi.astype(np.float64)))


def test_block_product_accuracy_and_plausible_slips():
    rng = np.random.default_rng(2)
    k = 64
    # data spanning eight decades across pixels, a table with a 1e3 dynamic range
    A = (rng.standard_normal((32, k)) + 1j * rng.standard_normal((32, k))) * 10.0 ** rng.uniform(-8, 0, (32, k))
    B = (rng.standard_normal((k, 64)) + 1j * rng.standard_normal((k, 64))) * 10.0 ** rng.uniform(-3, 0, (k, 64))
    A = A.astype(np.complex64).astype(np.complex128)
    B = B.astype(np.complex64).astype(np.complex128)
    ref = A @ B
    rel = np.linalg.norm(split_matmul(A, B) - ref) / np.linalg.norm(ref)
    assert rel < 1e-6, rel
    # dropping one cross term loses the lo half of one operand: ~2^-11
    rel_drop = np.linalg.norm(split_matmul(A, B, drop_cross=True) - ref) / np.linalg.norm(ref)
    assert rel_drop > 1e-5, rel_drop
    # a data scale three powers of two too large takes the largest components past FP16's 65504
    e = scale_exp(max(np.abs(A.real).max(), np.abs(A.imag).max()))
    with np.errstate(over="ignore"):
        hi, _ = split(np.float32(A.real) * np.float32(2.0 ** (e + 3)))
    assert not np.all(np.isfinite(hi.astype(np.float64)))
