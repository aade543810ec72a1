"""Coefficients of the portable cbrt / acos / sin / cos used by the cubic solver.

Each table is a near-minimax polynomial (Chebyshev interpolation, mpmath.chebyfit at 60
digits) rounded to double, listed from the constant term up:
  ASIN : P(t) = (asin(z) - z) / z^3, t = z^2 in [0, 1/4]     (asin(z) = z + z^3 P(z^2))
  SIN  : S(t) = (sin(y) - y) / y^3,  t = y^2 in [0, 1.1]     (sin(y) = y + y^3 S(y^2))
  COS  : C(t) = (cos(y) - 1) / y^2,  t = y^2 in [0, 1.1]     (cos(y) = 1 + y^2 C(y^2))
  RCBRT: m^(-1/3) on [1/2, 1], the starting value of the inverse-cube-root iteration.
Run it to print the tables as C hex literals (both the kernels' header and the CPU
restatement carry the same values)."""
import mpmath as mp

mp.mp.dps = 60


def _asin_p(t):
    if t == 0:
        return mp.mpf(1) / 6
    z = mp.sqrt(t)
    return (mp.asin(z) - z) / z ** 3


def _sin_s(t):
    if t == 0:
        return -mp.mpf(1) / 6
    y = mp.sqrt(t)
    return (mp.sin(y) - y) / y ** 3


def _cos_c(t):
    if t == 0:
        return -mp.mpf(1) / 2
    y = mp.sqrt(t)
    return (mp.cos(y) - 1) / t


TABLES = {
    "ASIN": (_asin_p, [0, mp.mpf(1) / 4], 14),
    "SIN": (_sin_s, [0, mp.mpf("1.1")], 7),
    "COS": (_cos_c, [0, mp.mpf("1.1")], 8),
    "RCBRT": (lambda m: m ** (-mp.mpf(1) / 3), [mp.mpf(1) / 2, 1], 5),
}


def tables():
    out = {}
    for name, (f, iv, n) in TABLES.items():
        poly, err = mp.chebyfit(f, iv, n, error=True)
        out[name] = ([float(c) for c in reversed(poly)], float(err))
    return out


if __name__ == "__main__":
    for name, (cs, err) in tables().items():
        print(f"// {name}: {len(cs)} terms, interpolation error {err:.2e}")
        print("{" + ", ".join(c.hex().replace("0x1.", "0x1.") for c in cs) + "}")
