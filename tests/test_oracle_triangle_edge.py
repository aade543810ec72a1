"""Pinning the oracle's triangle-edge queries (collisions_static.cu:345-943, icontour.cu):
hand cases with known piercing parameters, the shared-vertex loop rules, an independent
Moller-Trumbore edge/triangle intersection in numpy, and closed-form row counts of the
candidate queries.  CPU only."""
import numpy as np
import pytest

from tests import cases

FACE = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])


def _rows(r):
    return sorted(map(tuple, np.asarray(r).tolist()))


def test_edge_pierces_face_at_known_parameter(oracle):
    # edge v0v1 of triangle 0 runs from z = -0.25 to z = 0.75 through (0.2, 0.2, 0): t = 0.25;
    # edge v2v0 crosses z = 0 at (0.8, 0.8), outside the face.  Triangle 0 lies in the plane
    # x = y, which the face's edge v1v2 (cid 4) crosses inside it at t = 0.5.
    x = np.array([[0.2, 0.2, -0.25], [0.2, 0.2, 0.75], [2.0, 2.0, 0.5]])
    tris = np.stack([x, FACE])
    rows, coefs = oracle.triangle_edge(tris, [[0, 1]], "collisions")
    got = {tuple(r): float(c) for r, c in zip(rows.tolist(), coefs)}
    assert set(got) == {(0, 1, 0, -1), (0, 1, 4, -1)}
    assert got[(0, 1, 0, -1)] == pytest.approx(0.25, rel=1e-15)
    assert got[(0, 1, 4, -1)] == pytest.approx(0.5, rel=1e-15)
    # v1v2 crosses z = 0 at parameter 0.1 / 0.6 (point 0.3167), v2v0 at 0.5 / 0.75 (point
    # 0.4333), both inside the face
    x2 = np.array([[0.2, 0.2, -0.25], [0.2, 0.2, -0.1], [0.9, 0.9, 0.5]])
    rows, coefs = oracle.triangle_edge(np.stack([x2, FACE]), [[0, 1]], "collisions")
    got = dict(((int(r[2]), float(c)) for r, c in zip(rows, coefs)))
    assert set(got) == {1, 2}
    assert got[2] == pytest.approx(0.5 / 0.75, rel=1e-14)
    assert got[1] == pytest.approx(0.1 / 0.6, rel=1e-14)


def test_face_edges_pierce_the_other_triangle(oracle):
    """cid 3..5 are the edges of the second triangle against the first."""
    x = np.array([[0.2, 0.2, -0.25], [0.2, 0.2, 0.75], [2.0, 2.0, 0.5]])
    rows, coefs = oracle.triangle_edge(np.stack([FACE, x]), [[0, 1]], "collisions")
    got = {tuple(r): float(c) for r, c in zip(rows.tolist(), coefs)}
    assert set(got) == {(0, 1, 1, -1), (0, 1, 3, -1)}
    assert got[(0, 1, 3, -1)] == pytest.approx(0.25, rel=1e-15)
    assert got[(0, 1, 1, -1)] == pytest.approx(0.5, rel=1e-15)


def test_parallel_and_outside_edges_emit_nothing(oracle):
    par = np.array([[0.1, 0.1, 0.5], [0.6, 0.1, 0.5], [0.1, 0.6, 0.5]])       # parallel plane
    out = np.array([[2.0, 2.0, -1.0], [2.0, 2.0, 1.0], [3.0, 2.0, 0.0]])      # misses the face
    for x in (par, out):
        rows, _ = oracle.triangle_edge(np.stack([x, FACE]), [[0, 1]], "collisions")
        assert rows.shape[0] == 0


def test_shared_vertex_rules(oracle):
    """Collisions reject any shared vertex (allow 0).  Candidates allow one: for edges of x
    the one starting at the shared corner is a loop row (loop id 0), the one ending there
    emits nothing, the third a plain row; for edges of y only the edge away from the corner
    emits a row."""
    y = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    x = np.array([[0.0, 0.0, 0.0], [0.3, 0.3, 0.5], [0.3, 0.3, -0.5]])      # x.v0 == y.v0
    tris = np.stack([x, y])
    assert oracle.triangle_edge(tris, [[0, 1]], "collisions")[0].shape[0] == 0
    rows = oracle.triangle_edge(tris, [[0, 1]], "candidates")
    assert _rows(rows) == [(0, 1, 0, 0), (0, 1, 1, -1), (0, 1, 4, -1)]
    pairs, masks = oracle.triangle_edge(tris, [[0, 1]], "candidates2")
    assert pairs.tolist() == [[0, 1]]
    assert masks.tolist() == [[0b10011, 0b1]]


def _moller_trumbore(e0, e1, tri):
    """Independent edge/triangle intersection: solve e0 + t (e1 - e0) = a + u (b - a) + v (c - a)."""
    a, b, c = tri
    d = e1 - e0
    m = np.stack([-d, b - a, c - a], axis=1)
    if abs(np.linalg.det(m)) < 1e-14:
        return None
    t, u, v = np.linalg.solve(m, e0 - a)
    return t, u, v


def test_collisions_match_independent_intersection(oracle):
    """Every (pair, edge) is checked with an independent linear solve; cases within 1e-9 of
    a boundary (t, u, v or 1-u-v near 0) are skipped."""
    checked = 0
    for seed in range(3):
        t0, _, _ = cases.random_soup(300, seed=seed)
        bb, tree, _, _ = oracle.bvh(t0)
        pairs, _ = oracle.traverse(bb, tree, 0.0)
        ok = oracle.static_collisions(t0, pairs, 0.0)          # SAT pairs, no shared vertex
        rows, coefs = oracle.triangle_edge(t0, pairs, "collisions")
        got = {(int(r[0]), int(r[1]), int(r[2])): float(c) for r, c in zip(rows, coefs)}
        for x, y in ok:
            for cid in range(6):
                t1, t2 = (t0[x], t0[y]) if cid < 3 else (t0[y], t0[x])
                e = cid % 3
                sol = _moller_trumbore(t1[e], t1[(e + 1) % 3], t2)
                if sol is None:
                    continue
                t, u, v = sol
                margins = np.array([t, 1 - t, u, v, 1 - u - v])
                if np.abs(margins).min() < 1e-9:
                    continue
                hit = (margins > 0).all()
                key = (int(x), int(y), cid)
                assert (key in got) == hit, (key, sol)
                if hit:
                    assert got[key] == pytest.approx(t, rel=1e-9, abs=1e-12)
                    checked += 1
    assert checked > 50


def test_candidate_row_counts_closed_form(oracle):
    """Candidates: 6 rows for a pair with no shared vertex, 3 (one loop) with one shared
    vertex; candidates2 masks carry the same rows as bits."""
    t0, _, _ = cases.shared_vertex_soup(200)
    bb, tree, _, _ = oracle.bvh(t0)
    pairs, _ = oracle.traverse(bb, tree, 0.0)
    rows = oracle.triangle_edge(t0, pairs, "candidates")
    p2, masks = oracle.triangle_edge(t0, pairs, "candidates2")
    assert p2.shape[0] > 100
    per = {}
    for r in rows:
        per.setdefault((int(r[0]), int(r[1])), []).append((int(r[2]), int(r[3])))
    n_shared = 0
    for (x, y), (mx, my) in zip(p2.tolist(), masks.tolist()):
        a, b = t0[x], t0[y]
        shared = sum(any((a[i] == b[j]).all() for j in range(3)) for i in range(3))
        got = per.get((x, y), [])
        assert len(got) == (6 if shared == 0 else 3), (x, y, shared, got)
        assert sum(1 for _, l in got if l == 0) == (0 if shared == 0 else 1)
        assert sorted(c for c, _ in got) == [c for c in range(6) if mx >> c & 1]
        assert [c for c, l in sorted(got) if l == 0] == [c for c in range(6) if my >> c & 1]
        n_shared += shared == 1
    assert n_shared > 20


def test_candidates_reject_two_shared_vertices(oracle):
    t0, _, _ = cases.shared_vertex_soup(200)
    bb, tree, _, _ = oracle.bvh(t0)
    pairs, _ = oracle.traverse(bb, tree, 0.0)
    p2, _ = oracle.triangle_edge(t0, pairs, "candidates2")
    for x, y in p2.tolist():
        a, b = t0[x], t0[y]
        assert sum(any((a[i] == b[j]).all() for j in range(3)) for i in range(3)) <= 1
