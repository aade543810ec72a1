"""Known-answer vectors for the CCD narrow phase and the impulses, computed with exact
rational arithmetic (sympy, 50-digit root isolation), independently of the oracle and of
the kernels.

The reference ships no test vectors, so these fixtures come from constructed motions.

ccd_known_answers.json
  * linear_dropped: a vertex moves straight through a static triangle.  Only one row of
    the coplanarity determinant depends on t, so the polynomial is linear and the reference
    computes its root but never stores it (collisions_continuous.cu:532-536): no collision.
  * vertex_face / edge_edge: the same crossings with the other triangle deforming, so the
    polynomial is cubic.  Its real roots are isolated exactly; a root is a collision when
    the vertex lies inside the triangle (or the two edges cross) at that time, decided at
    50 digits.  Only roots at least 1e-6 away from every decision boundary are kept.
  * coefficient cases: random tetrahedra whose cubic coefficients are expanded exactly.

ccd_typed_cases.json -- one scene per sub-primitive type (0..14), three variants each, in
which exactly that sub-primitive pair collides, at a time of impact t* fixed by the
construction.  Every coordinate, mass and velocity is a dyadic rational with few bits, so
the inputs are exact in float32 as well as float64.
  * vertex-face (types 9+k: face of triangle 0, vertex k of triangle 1; types 12+k: vertex
    k of triangle 0, face of triangle 1).  The face lies in the plane z = 0 and scales about
    C = (1, 1, 0) by diag(1 + t/2, 1 + t/4, 1), so its area is A0 (1 + t/2)(1 + t/4).  The
    moving triangle translates rigidly with velocity (wx, wy, -V); its vertex k is at height
    V t* at t = 0 and passes the face point with barycentric coordinates (b0, b1, b2) at
    t*.  Coplanarity polynomial: A0 (1 + t/2)(1 + t/4) (V t* - V t): roots t*, -2, -4.
  * edge-edge (types 3a+b: edge E[a] of triangle 0, edge E[b] of triangle 1,
    E = {(0,1), (0,2), (1,2)}).  Edge b of triangle 1 lies on the line y = 1, z = 0 and
    scales about C by (1 + t/2); edge a of triangle 0 is parallel to the y axis at height
    V (t* - t), scales by (1 + t/4) about its point at parameter s*, and crosses edge b at
    parameters (s*, u*) at t*.  Polynomial: height x length_a x length_b: roots t*, -2, -4.
  Every scene is checked exactly over all 15 types: the only accepted collision is the
  constructed one.

  Expected impulses (collisions_impulses.cu:162-543, :936-1067), derived from the
  construction, not from positions: at t* the contact weights are the constructed (b0, b1,
  b2) or (s*, u*); the unit normal is +z for vertex-face (from the face to the vertex at
  t = 0), -z for edge-edge (from edge a, above, towards edge b), and the relative normal
  velocity is nv = -V in every case.  With the interpolated masses mpx (first primitive)
  and mpy (second), mf0 = mpx / (mpx + mpy):
      first primitive's corners:  n * nv * (1 - mf0)
      second primitive's corners: n * (-nv) * mf0
  and for types 12-14 the normal is inverted after nv is computed.  Position corrections
  dx = impulse * (1 - max(t* - offset, 0)), offset = clamp(2 * 2e-3 / V, 0, 1).  Each of the
  four corners counts 1.

Run:  python tests/golden/make_golden.py   (writes both JSON files next to this script)
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


# ---- typed scenes: one constructed collision per sub-primitive type -----------------------

C = (F(1), F(1), F(0))
ALPHA, BETA, GAMMA = F(1, 2), F(1, 4), F(1, 4)
TWO_THICKNESS = F(4, 1000)          # 2 * THICKNESS (collisions_impulses.cu:12-14, :975)


def vadd(a, b):
    return tuple(a[k] + b[k] for k in range(3))


def vsub(a, b):
    return tuple(a[k] - b[k] for k in range(3))


def vscale(a, s):
    return tuple(a[k] * s for k in range(3))


def face_scale(p, t):
    """The vertex-face scenes' face motion: scaling about C by diag(1 + t/2, 1 + t/4, 1)."""
    return (C[0] + (1 + ALPHA * t) * (p[0] - C[0]), C[1] + (1 + BETA * t) * (p[1] - C[1]), p[2])


def vf_scene(k, face_is_x, tstar, bary, V, w, swap):
    """Returns (t0, t1) [2][3][3] of Fractions: the face (triangle 0 if face_is_x, else 1) and
    the rigidly moving triangle whose vertex k passes the face point `bary` at tstar."""
    base = [(F(0), F(0), F(0)), (F(4), F(0), F(0)), (F(0), F(4), F(0))]
    if swap:                                    # the other winding: the normal starts as -z
        base = [base[0], base[2], base[1]]
    face0 = base
    face1 = [face_scale(p, 1) for p in base]
    q = (F(0), F(0), F(0))
    for i in range(3):
        q = vadd(q, vscale(face_scale(base[i], tstar), bary[i]))
    vel = (w[0], w[1], -V)
    vk0 = vsub(q, vscale(vel, tstar))           # height V * tstar at t = 0
    offsets = [(F(2), F(1, 2), F(3)), (F(1, 4), F(2), F(5, 2))]
    mov0 = [None] * 3
    mov0[k] = vk0
    for j, off in zip([j for j in range(3) if j != k], offsets):
        mov0[j] = vadd(vk0, off)
    mov1 = [vadd(p, vel) for p in mov0]
    if face_is_x:
        return [face0, mov0], [face1, mov1]
    return [mov0, face0], [mov1, face1]


def ee_scene(a, b, tstar, s_star, u_star, V, w):
    """Edge E[a] of triangle 0 crosses edge E[b] of triangle 1 at parameters (s*, u*) at t*."""
    i0, i1 = EDGES[a]
    i2 = 3 - i0 - i1
    j0, j1 = EDGES[b]
    j2 = 3 - j0 - j1

    def ys(p, t):       # isotropic in-plane scaling about C
        return (C[0] + (1 + ALPHA * t) * (p[0] - C[0]), C[1] + (1 + ALPHA * t) * (p[1] - C[1]), p[2])

    y0 = [None] * 3
    y0[j0] = (C[0] - 2, C[1], F(0))
    y0[j1] = (C[0] + 2, C[1], F(0))
    y0[j2] = (C[0] + F(1, 2), C[1] - 1, F(-2))
    y1 = [ys(p, 1) for p in y0]
    pa, pb = ys(y0[j0], tstar), ys(y0[j1], tstar)
    cross_pt = vadd(pa, vscale(vsub(pb, pa), u_star))
    l0 = F(2)
    lstar = l0 * (1 + GAMMA * tstar)
    e_y = (F(0), F(1), F(0))
    h0s = vsub(cross_pt, vscale(e_y, s_star * lstar))
    h1s = vadd(cross_pt, vscale(e_y, (1 - s_star) * lstar))
    d0 = (w[0], w[1] - s_star * l0 * GAMMA, -V)
    d1 = (w[0], w[1] + (1 - s_star) * l0 * GAMMA, -V)
    x0 = [None] * 3
    x1 = [None] * 3
    x0[i0] = vsub(h0s, vscale(d0, tstar))
    x0[i1] = vsub(h1s, vscale(d1, tstar))
    x1[i0] = vadd(x0[i0], d0)
    x1[i1] = vadd(x0[i1], d1)
    mid = vscale(vadd(x0[i0], x0[i1]), F(1, 2))
    x0[i2] = vadd(mid, (F(1, 2), F(0), F(3)))
    x1[i2] = vadd(x0[i2], (w[0], w[1], -V))
    return [x0, y0], [x1, y1]


def closed_form_impulses(type_, tstar, weights, V, masses):
    """dv, dx [2][3][3] and cnt [2][3] (Fractions / ints) of the one constructed collision."""
    if type_ < 9:
        a, b = EDGES[type_ // 3], EDGES[type_ % 3]
        s, u = weights
        mpx = masses[0][a[0]] * (1 - s) + masses[0][a[1]] * s
        mpy = masses[1][b[0]] * (1 - u) + masses[1][b[1]] * u
        n = (F(0), F(0), F(-1))
        first, second = [(0, a[0]), (0, a[1])], [(1, b[0]), (1, b[1])]
        invert = False
    elif type_ < 12:
        k = type_ - 9
        mpx = sum(masses[0][i] * weights[i] for i in range(3))
        mpy = masses[1][k]
        n = (F(0), F(0), F(1))
        first, second = [(0, 0), (0, 1), (0, 2)], [(1, k)]
        invert = False
    else:
        k = type_ - 12
        mpx = masses[0][k]
        mpy = sum(masses[1][i] * weights[i] for i in range(3))
        n = (F(0), F(0), F(1))
        first, second = [(0, k)], [(1, 0), (1, 1), (1, 2)]
        invert = True
    nv = -V
    mf0 = mpx / (mpx + mpy)
    im0, im1 = nv * (1 - mf0), -nv * mf0
    if invert:
        n = vscale(n, -1)
    offset = min(max(TWO_THICKNESS / V, F(0)), F(1))
    keep = 1 - max(tstar - offset, F(0))
    dv = [[[F(0)] * 3 for _ in range(3)] for _ in range(2)]
    dx = [[[F(0)] * 3 for _ in range(3)] for _ in range(2)]
    cnt = [[0] * 3 for _ in range(2)]
    for group, mag in ((first, im0), (second, im1)):
        for tri, c in group:
            imp = vscale(n, mag)
            for d in range(3):
                dv[tri][c][d] += imp[d]
                dx[tri][c][d] += imp[d] * keep
            cnt[tri][c] += 1
    return dv, dx, cnt


def fl(x):
    """Fractions -> floats (nested lists)."""
    if isinstance(x, (list, tuple)):
        return [fl(v) for v in x]
    return float(x)


def exact_in_f32(x):
    import numpy as np
    a = np.array(fl(x), np.float64)
    return bool((a.astype(np.float32).astype(np.float64) == a).all())


TSTARS = [F(1, 2), F(3, 8), F(5, 8)]
VS = [F(1, 2), F(3, 4), F(1, 2)]
WS = [(F(1, 8), F(-1, 16)), (F(0), F(0)), (F(-3, 16), F(1, 8))]
BARYS = [(F(1, 4), F(3, 8), F(3, 8)), (F(1, 2), F(1, 4), F(1, 4)), (F(3, 8), F(1, 8), F(1, 2))]
NUDGES = [(F(0), F(0)), (F(1, 16), F(0)), (F(0), F(1, 16)), (F(-1, 16), F(1, 32)), (F(1, 32), F(-1, 16))]
EE_PARAMS = [(F(3, 8), F(3, 8)), (F(5, 8), F(1, 4)), (F(1, 2), F(11, 16))]
MASSES = [
    [[F(3, 2), F(1), F(5, 2)], [F(2), F(7, 4), F(1, 2)]],
    [[F(1, 4), F(3), F(2)], [F(5, 4), F(1, 2), F(9, 4)]],
    [[F(2), F(5, 8), F(1)], [F(3, 8), F(3), F(3, 2)]],
]


def typed_cases():
    cases = []
    for type_ in range(15):
        for var in range(3):
            tstar, V, masses = TSTARS[var], VS[var], MASSES[(var + type_) % 3]
            # the first in-plane velocity whose scene has no root of any type near a decision
            # boundary and no collision besides the constructed one
            for nudge in NUDGES:
                w = (WS[var][0] + nudge[0], WS[var][1] + nudge[1])
                if type_ < 9:
                    s_star, u_star = EE_PARAMS[var]
                    t0, t1 = ee_scene(type_ // 3, type_ % 3, tstar, s_star, u_star, V, w)
                    weights = (s_star, u_star)
                else:
                    k = (type_ - 9) % 3
                    weights = BARYS[(var + k) % 3]
                    t0, t1 = vf_scene(k, type_ < 12, tstar, weights, V, w, swap=(var == 1))
                found = []
                for ty in range(15):
                    e = expected(t0, t1, ty)
                    if e is None:
                        found = None
                        break
                    found += [(ty, r) for r in e["roots"]]
                if found == [(type_, float(tstar))]:
                    break
            else:
                raise AssertionError(f"no clean scene for type {type_} variant {var}")
            dv, dx, cnt = closed_form_impulses(type_, tstar, weights, V, masses)
            assert exact_in_f32(t0) and exact_in_f32(t1) and exact_in_f32(masses)
            cases.append({
                "type": type_, "variant": var, "t0": fl(t0), "t1": fl(t1), "masses": fl(masses),
                "construction": {"tstar": str(tstar), "V": str(V), "weights": [str(v) for v in weights],
                                 "velocity_xy": [str(v) for v in w]},
                "expect": {"collisions": [[0, 1, type_]], "roots": [float(tstar)],
                           "dv": fl(dv), "dx": fl(dx), "cnt": cnt},
            })
    return cases


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
    typed = typed_cases()
    with open(os.path.join(HERE, "ccd_typed_cases.json"), "w") as f:
        json.dump({"typed_cases": typed}, f, indent=None, separators=(",", ":"))
        f.write("\n")
    print(f"wrote {len(cases)} CCD cases, {len(coeff_cases)} coefficient cases and {len(typed)} typed cases")


if __name__ == "__main__":
    main()
