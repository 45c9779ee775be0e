"""The kernels form k / 255.0f (palette colours and RGBA8 texels,
icbc.h:1656) as q = k * r, r = 1/255 rounded, corrected by one fused
multiply-add residual step: fma(fma(-q, 255, k), r, q).  This checks, in
exact rational arithmetic, that the formula gives the correctly rounded
quotient -- the value of the reference's division -- for every k in 0..255,
and that the uncorrected product alone would not."""
from fractions import Fraction

import numpy as np


def _round_f32(x: Fraction) -> np.float32:
    """x rounded to the nearest float32 (ties to even), exactly."""
    c = np.float32(float(x))
    best = None
    for v in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        d = abs(Fraction(float(v)) - x)
        even = int(np.array(v, np.float32).view(np.uint32)) & 1 == 0
        if best is None or d < best[0] or (d == best[0] and even):
            best = (d, v)
    return np.float32(best[1])


def _fma(a, b, c) -> np.float32:
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def test_corrected_product_is_the_correctly_rounded_quotient():
    r = np.float32(1) / np.float32(255)
    plain_wrong = 0
    for k in range(256):
        want = _round_f32(Fraction(k, 255))
        assert want == np.float32(k) / np.float32(255)  # numpy's float32 division agrees
        q = np.float32(k) * r
        plain_wrong += q != want
        got = _fma(_fma(-q, np.float32(255), np.float32(k)), r, q)
        assert got == want, k
    assert plain_wrong > 0  # the correction step is needed
