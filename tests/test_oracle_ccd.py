"""Pinning the oracle's continuous narrow phase and impulses: exact coefficient expansions,
multiprecision roots, the solver's branches (including the dropped linear root), the
known-answer fixtures in tests/golden, hand-derived impulses and the portable math
functions.  CPU only."""
import json
import os
from fractions import Fraction as F

import mpmath
import numpy as np
import pytest

from tests import cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ccd_known_answers.json")
mpmath.mp.dps = 40


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_coefficients_match_exact_expansion(oracle, golden):
    """figures.cu:150-281 against the exact determinant expansion (sympy): a dropped term,
    wrong sign or swapped operand shows up far above rounding."""
    for c in golden["coefficient_cases"]:
        tet = np.array(c["pos"] + c["dlt"], np.float64)
        got = oracle.cubic_coeffs(tet)
        exact = [F(s) for s in c["coeffs_exact"]]
        scale = max(abs(float(e)) for e in exact)
        for g, e in zip(got, exact):
            assert abs(F(float(g)) - e) <= F(1, 10 ** 13) * F(scale), (got, c["coeffs"])


def test_golden_ccd_cases(oracle, golden):
    for c in golden["ccd_cases"]:
        t0, t1 = np.array(c["t0"]), np.array(c["t1"])
        coll, roots, _ = oracle.ccd(t0, t1, np.array([[0, 1]]))
        got = sorted(float(r) for cc, r in zip(coll, roots) if cc[2] == c["type"])
        want = c["expect"]["roots"]
        assert len(got) == len(want), (c["kind"], got, want)
        for g, w in zip(got, want):
            assert abs(g - w) <= 1e-9 * max(abs(w), 1e-300), (c["kind"], g, w)


def test_solver_matches_mpmath(oracle):
    """Random cubics with three real roots and with one: every root the solver reports in
    [0, 1] is a root of the polynomial (mpmath, 40 digits) to 1e-9."""
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(3000):
        r = rng.uniform(-0.5, 1.5, 3)
        if rng.random() < 0.5:
            coef = np.poly(r) * rng.uniform(0.1, 10)
        else:
            coef = rng.normal(size=4)
        sl = oracle.solve_cubic(coef)
        ref = mpmath.polyroots([mpmath.mpf(float(v)) for v in coef], maxsteps=200, extraprec=200)
        real = sorted(float(mpmath.re(z)) for z in ref if abs(mpmath.im(z)) < 1e-20)
        for s in sl:
            if 0 <= s <= 1:
                checked += 1
                assert min(abs(s - x) for x in real) <= 1e-9 * max(1.0, abs(s)), (coef, sl, real)
    assert checked > 1000


def test_solver_branches(oracle):
    # linear (c3 = c2 = 0): the reference computes -d/c but never stores it
    assert (oracle.solve_cubic([0.0, 0.0, 2.0, -1.0]) == -1).all()
    # quadratic branch triggers on |c3| < 1e-16 absolute: (t - 0.25)(t - 0.75)
    sl = oracle.solve_cubic([1e-17, 1.0, -1.0, 0.1875])
    assert sorted(sl[:2]) == [0.25, 0.75] and sl[2] == -1
    # a tiny but non-zero c3 above the threshold goes through the cubic formulas
    assert not (oracle.solve_cubic([2e-16, 1.0, -1.0, 0.1875]) == [0.75, 0.25, -1]).all()
    # triple root (t - 0.5)^3 through the float k path: f = g = h = 0
    sl = oracle.solve_cubic([1.0, -1.5, 0.75, -0.125])
    assert sl[0] == np.float64(np.float32(0.5)) and (sl[1:] == -1).all()
    # roots outside [-1e-16, 1 + 1e-16] are not stored
    sl = oracle.solve_cubic(np.poly([-0.5, 1.5, 2.0]))
    assert (sl == -1).all()


def test_flatten_window_and_duplicates(oracle):
    """A double root fills two slots (collisions_continuous.cu:537-547: quadratic branch with
    D = 0 exactly, x1 = x2); the flatten keeps 0 <= r <= 1 only (:743-763), so it emits it
    twice (tests/test_oracle_typed.py checks the two rows end to end)."""
    # (t - 1/4)^2: c3 = 0, D = 0.25 - 4 * 0.0625 = 0 exactly
    assert oracle.solve_cubic([0.0, 1.0, -0.5, 0.0625]).tolist() == [0.25, 0.25, -1.0]
    # roots at -1/2 and 3/2 are stored only inside [-1e-16, 1 + 1e-16]; 1/2 is kept
    assert oracle.solve_cubic([0.0, 1.0, -0.0, -0.25]).tolist() == [0.5, -1.0, -1.0]
    sl = oracle.solve_cubic([0.0, 4.0, -4.0, -3.0])     # 4 (t - 3/2)(t + 1/2)
    assert (sl == -1).all()


def test_dd_cube_is_rounded_cube(oracle):
    """The double-double product's leading term is the rounded exact cube."""
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.normal(size=2000), 10.0 ** rng.uniform(-80, 80, 500)])
    for x in xs:
        exact = F(float(x)) ** 3
        got = F(oracle.dd_cube_lead(float(x)))
        rounded = F(float(exact))
        assert got == rounded


def _real_cbrt(v):
    return mpmath.sign(v) * mpmath.cbrt(abs(v))


@pytest.mark.parametrize("name,fn,lo,hi", [
    ("cbrt", _real_cbrt, 1e-3, 1e3), ("acos", mpmath.acos, -1.0, 1.0),
    ("sin", mpmath.sin, 0.0, np.pi / 3), ("cos", mpmath.cos, 0.0, np.pi / 3)])
def test_portable_math_within_one_ulp(oracle, name, fn, lo, hi):
    rng = np.random.default_rng(3)
    x = rng.uniform(lo, hi, 4000)
    if name == "cbrt":
        x = np.concatenate([x, 10.0 ** rng.uniform(-300, 300, 1000), -x[:100]])
    if name == "acos":
        x = np.concatenate([x, 1 - 10.0 ** rng.uniform(-16, -1, 300), -1 + 10.0 ** rng.uniform(-16, -1, 300), [-1.0, 1.0, 0.5, -0.5]])
    y = oracle.libm(name, x)
    ref = np.array([float(fn(mpmath.mpf(float(v)))) for v in x])
    ulps = np.abs(y - ref) / np.spacing(np.abs(ref))
    assert ulps.max() <= 1.0, (name, ulps.max())


def test_portable_math_special_values(oracle):
    assert np.isnan(oracle.libm("acos", [1.5])[0]) and np.isnan(oracle.libm("acos", [np.nan])[0])
    assert oracle.libm("cbrt", [0.0])[0] == 0.0 and oracle.libm("cbrt", [np.inf])[0] == np.inf
    assert oracle.libm("acos", [1.0])[0] == 0.0
    assert oracle.libm("cbrtf", [8.0])[0] == 2.0


# ---- impulses -------------------------------------------------------------------------

def test_vertex_face_impulse_hand_case(oracle):
    """Vertex 0 of triangle x falls straight through a static triangle y (type 12) with unit
    masses; the face is tilted slightly over the step so that the cubic is not linear.
    Derivation (collisions_impulses.cu:196-241, :936-1050): n points from the face to the
    vertex at t0 (+z), nv = <dvt - dp, n> ~ -2, both mass fractions 1/2, so the vertex gets
    +n * |nv|/2 after the inverse flip and each face corner -n * |nv|/2."""
    t0 = np.array([[[0.25, 0.25, 1.0], [3.0, 0.25, 4.0], [0.25, 3.0, 4.0]],
                   [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]])
    t1 = t0.copy()
    t1[0, 0, 2] = -1.0
    t1[1, 1] = [1.0 + 1e-3, 0.0, 1e-3]       # the face deforms in and out of its plane,
    t1[1, 2] = [0.0, 1.0 - 1e-3, -1e-3]      # so t appears in more than one determinant column
    m = np.ones((2, 3))
    coll, roots, _ = oracle.ccd(t0, t1, np.array([[0, 1]]))
    hit = [(tuple(c), r) for c, r in zip(coll, roots) if c[2] == 12]
    assert len(hit) == 1
    root = hit[0][1]
    assert abs(root - 0.5) < 1e-3
    dv, dx, cnt = oracle.impulses(t0, t1, m, np.array([[0, 1, 12]]), np.array([root]))
    assert cnt[0].tolist() == [1, 0, 0] and cnt[1].tolist() == [1, 1, 1]
    assert dv[0, 0, 2] > 0.99 and abs(dv[0, 0, 0]) < 1e-2 and abs(dv[0, 0, 1]) < 1e-2
    for k in range(3):
        assert dv[1, k, 2] < -0.99
    # dx = imp * (1 - max(root - offset, 0)) with offset = 4e-3 / |nv| ~ 2e-3
    nv = -2.0 * np.linalg.norm(dv[0, 0])
    offset = min(max(0.004 / -nv, 0.0), 1.0)
    np.testing.assert_allclose(dx[0, 0], dv[0, 0] * (1 - max(root - offset, 0.0)), rtol=1e-12)
    assert (dv[0, 1:] == 0).all()


def test_edge_edge_equal_mass_conservation(oracle):
    """With equal corner masses both mass fractions are 1/2, so the four corner impulses of
    an edge-edge collision sum to zero (n * (im0 + im0 + im1 + im1), im1 = -im0)."""
    t0, t1, _ = cases.random_soup(300, seed=21)
    m = np.ones((300, 3))
    bb, tree, _, _ = oracle.bvh(t0, t1)
    pairs, _ = oracle.traverse(bb, tree)
    coll, roots, _ = oracle.ccd(t0, t1, pairs)
    ee = coll[:, 2] < 9
    assert ee.sum() > 10
    for c, r in zip(coll[ee][:200], roots[ee][:200]):
        dv, _, _ = oracle.impulses(t0, t1, m, c[None], np.array([r]))
        s = dv.reshape(-1, 3).sum(0)
        if np.isfinite(s).all():
            assert np.abs(s).max() <= 1e-12 * max(np.abs(dv).max(), 1e-300)


def test_impulses_skip_padding(oracle):
    t0, t1, m = cases.random_soup(100, seed=2)
    bb, tree, _, _ = oracle.bvh(t0, t1)
    pairs, _ = oracle.traverse(bb, tree)
    coll, roots, _ = oracle.ccd(t0, t1, pairs)
    a = oracle.impulses(t0, t1, m, coll, roots)
    pad_c = np.concatenate([coll, -np.ones((5, 3), np.int64)])
    pad_r = np.concatenate([roots, -np.ones(5)])
    b = oracle.impulses(t0, t1, m, pad_c, pad_r)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
