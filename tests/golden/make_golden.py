"""Known-answer vectors for the CCD narrow phase, computed with exact rational arithmetic
(sympy, 50-digit root isolation), independently of the oracle and of the kernels.

The reference ships no test vectors, so these fixtures come from constructed motions:
  * linear_dropped: a vertex moves straight through a static triangle.  Only one row of
    the coplanarity determinant depends on t, so the polynomial is linear and the reference
    computes its root but never stores it (collisions_continuous.cu:532-536): no collision.
  * vertex_face / edge_edge: the same crossings with the other triangle deforming, so the
    polynomial is cubic.  Its real roots are isolated exactly; a root is a collision when
    the vertex lies inside the triangle (or the two edges cross) at that time, decided at
    50 digits.  Only roots at least 1e-6 away from every decision boundary are kept.
  * coefficient cases: random tetrahedra whose cubic coefficients are expanded exactly.

Run:  python tests/golden/make_golden.py   (writes tests/golden/ccd_known_answers.json)
"""
import json
import os
import random
from fractions import Fraction as F

import mpmath
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
mpmath.mp.dps = 50
T = sp.Symbol("t")
EDGES = [(0, 1), (0, 2), (1, 2)]


def tet_rows(pos, dlt):
    return [[sp.Rational(F(pos[i][k]) - F(pos[3][k])) + T * sp.Rational(F(dlt[i][k]) - F(dlt[3][k]))
             for k in range(3)] for i in range(3)]


def exact_coeffs(pos, dlt):
    """Coefficients (c3, c2, c1, c0) of det[(v_i - v3) + t (dv_i - dv3)], i = 0..2."""
    poly = sp.Poly(sp.expand(sp.Matrix(tet_rows(pos, dlt)).det()), T)
    return [poly.coeff_monomial(T ** k) for k in (3, 2, 1, 0)]


def corners(type_):
    if type_ < 9:
        a, b = EDGES[type_ // 3], EDGES[type_ % 3]
        return [(0, a[0]), (0, a[1]), (1, b[0]), (1, b[1])]
    if type_ < 12:
        return [(0, 0), (0, 1), (0, 2), (1, type_ - 9)]
    return [(0, type_ - 12), (1, 0), (1, 1), (1, 2)]


def mp(v):
    return mpmath.mpf(F(v).numerator) / F(v).denominator


def at(p, d, t):
    return [mp(p[k]) + t * mp(d[k]) for k in range(3)]


def sub(a, b):
    return [a[k] - b[k] for k in range(3)]


def dot(a, b):
    return sum(a[k] * b[k] for k in range(3))


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def margins(type_, P):
    """Decision quantities of filterCollisions at the root (must lie in [0, 1])."""
    if type_ < 9:
        da, db, dc = sub(P[1], P[0]), sub(P[3], P[2]), sub(P[2], P[0])
        cab = cross(da, db)
        s = dot(cross(dc, db), cab) / dot(cab, cab)
        q = dot(cross(dc, da), cab) / dot(cab, cab)
        return [s, q]
    v0, v1, v2, vt = (P[0], P[1], P[2], P[3]) if type_ < 12 else (P[1], P[2], P[3], P[0])
    e1, e2, ep = sub(v0, v1), sub(v2, v1), sub(vt, v1)
    d00, d01, d11, d20, d21 = dot(e1, e1), dot(e1, e2), dot(e2, e2), dot(ep, e1), dot(ep, e2)
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    return [v, w, 1 - v - w]


def expected(t0, t1, type_):
    pos = [t0[s][c] for s, c in corners(type_)]
    dlt = [[t1[s][c][k] - t0[s][c][k] for k in range(3)] for s, c in corners(type_)]
    coeffs = exact_coeffs(pos, dlt)
    poly = sp.Poly(sum(c * T ** (3 - i) for i, c in enumerate(coeffs)), T)
    roots = []
    if poly.degree() >= 2:
        for r in sp.Poly(poly, T).real_roots():
            rv = mpmath.mpf(sp.N(r, 60))
            if rv < 0 or rv > 1:
                continue
            P = [at(pos[k], dlt[k], rv) for k in range(4)]
            m = margins(type_, P)
            dist = min(min(abs(x), abs(x - 1)) for x in m)
            if dist < 1e-6 or min(rv, 1 - rv) < 1e-6:
                return None      # too close to a decision boundary for a fixture
            if all(0 <= x <= 1 for x in m):
                roots.append(float(rv))
    return {"coeffs": [float(c) for c in coeffs], "roots": sorted(roots)}


def vertex_face_case(toi, seed, deform):
    """Triangle y near the plane z = 0, vertex 0 of triangle x moving along -z through it."""
    rnd = random.Random(seed)
    y0 = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
    y1 = [[v[k] + deform * (rnd.random() - 0.5) for k in range(3)] for v in y0]
    px, py = 0.2 + 0.1 * rnd.random(), 0.2 + 0.1 * rnd.random()
    z0 = 0.5 + rnd.random()
    dz = -z0 / toi
    x0 = [[px, py, z0], [px + 2.0, py, z0 + 3.0], [px, py + 2.0, z0 + 3.0]]
    x1 = [[px, py, z0 + dz], [px + 2.0, py, z0 + 3.0], [px, py + 2.0, z0 + 3.0]]
    return [x0, y0], [x1, y1], 12


def edge_edge_case(toi, seed, deform):
    """Edge (0,1) of y along the x axis; edge (0,1) of x parallel to the y axis, moving
    along -z across it."""
    rnd = random.Random(seed)
    y0 = [[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, -3.0, -1.0]]
    y1 = [[v[k] + (deform * (rnd.random() - 0.5) if i < 2 else 0.0) for k in range(3)] for i, v in enumerate(y0)]
    a = -0.5 - 0.2 * rnd.random()
    z0 = 0.3 + rnd.random()
    dz = -z0 / toi
    x0 = [[0.1, a, z0], [0.1, a + 2.0, z0], [3.0, 0.5, z0 + 2.0]]
    x1 = [[0.1, a, z0 + dz], [0.1, a + 2.0, z0 + dz], [3.0, 0.5, z0 + 2.0]]
    return [x0, y0], [x1, y1], 0


def main():
    cases = []
    for i, toi in enumerate([0.125, 0.25, 0.5, 0.75, 0.9375]):
        for maker, name in ((vertex_face_case, "vertex_face"), (edge_edge_case, "edge_edge")):
            t0, t1, ty = maker(toi, i, 0.0)
            cases.append({"kind": "linear_dropped", "geometry": name, "type": ty, "t0": t0, "t1": t1,
                          "expect": {"roots": []}})
            for seed in range(40):
                t0, t1, ty = maker(toi, 1000 * i + seed, 0.3)
                e = expected(t0, t1, ty)
                if e is not None and e["roots"]:
                    cases.append({"kind": name, "type": ty, "t0": t0, "t1": t1, "expect": e})
                    break
    rnd = random.Random(7)
    coeff_cases = []
    for _ in range(40):
        pos = [[rnd.uniform(-1, 1) for _ in range(3)] for _ in range(4)]
        dlt = [[rnd.uniform(-0.5, 0.5) for _ in range(3)] for _ in range(4)]
        c = exact_coeffs(pos, dlt)
        coeff_cases.append({"pos": pos, "dlt": dlt, "coeffs": [float(v) for v in c], "coeffs_exact": [str(v) for v in c]})
    out = {"ccd_cases": cases, "coefficient_cases": coeff_cases}
    with open(os.path.join(HERE, "ccd_known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(cases)} CCD cases and {len(coeff_cases)} coefficient cases")


if __name__ == "__main__":
    main()
