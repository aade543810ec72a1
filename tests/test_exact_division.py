"""The narrow phase's division-free x / 3 and x / 27 (div_exact in ccc_libm.h): q = RN(x z),
t = RN(D q - x), RN(q - t z) with z = RN(1 / D), emulated here with exact rational
arithmetic (each fma is the exact value rounded once), equals the IEEE division for finite
x of every exponent, subnormals included.  The GPU test in test_gpu_device_math.py checks
the device code on the same input classes and on the special values.
"""
from fractions import Fraction

import numpy as np
import pytest


def _rn(v: Fraction) -> float:
    return v.numerator / v.denominator     # int / int is correctly rounded (subnormals too)


def _div_exact(x: float, d: float) -> float:
    z = 1.0 / d
    q = _rn(Fraction(x) * Fraction(z))
    t = _rn(Fraction(q) * Fraction(d) - Fraction(x))
    return _rn(Fraction(q) - Fraction(t) * Fraction(z))


def test_reciprocals_are_within_2_pow_minus_54():
    for d in (3, 27):
        assert abs(Fraction(1.0 / d) * d - 1) == Fraction(1, 2**54)


@pytest.mark.parametrize("d", [3.0, 27.0])
def test_sequence_equals_division(d):
    rng = np.random.default_rng(int(d))
    n = 4000
    u = rng.integers(0, 2**62, n, dtype=np.uint64) * np.uint64(4) + rng.integers(0, 4, n, dtype=np.uint64)
    x = u.view(np.float64)
    x = x[np.isfinite(x) & (x != 0)]
    e = np.arange(1, 2047, 7, dtype=np.uint64)
    tops = ((e << np.uint64(52)) | np.uint64(0xFFFFFFFFFFFFF)).view(np.float64)
    tops = tops[np.isfinite(tops)]
    sub = (rng.integers(1, 2**52, 500, dtype=np.uint64)).view(np.float64)
    near = rng.uniform(-10, 10, 2000)
    for v in np.concatenate([x, tops, -tops, sub, near, [5e-324, 1.7976931348623157e308]]):
        v = float(v)
        got, want = _div_exact(v, d), v / d
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (v.hex(), got, want)
