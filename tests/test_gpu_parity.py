"""Parity of the HIP path (through the torch extension, which calls the C ABI) against the
CPU oracle on seeded inputs.  Trees, boxes, pair and collision lists and roots are compared
bit-exactly (lists as sorted multisets); impulse sums, accumulated with fp64 atomics in a
different order, within 1e-9 relative."""
import numpy as np
import pytest

from tests import cases
from tests.conftest import sorted_rows, valid_rows

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _t(a, dev="cuda"):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _np(t):
    return t.detach().cpu().numpy()


def _check_close(gpu, ref, name):
    scale = max(float(np.abs(ref).max()) if ref.size else 0.0, 1e-300)
    err = np.abs(gpu - ref)
    bad = err > RTOL * np.abs(ref) + 1e-12 * scale
    assert not bad.any(), f"{name}: {int(bad.sum())} entries differ, max err {err.max():.3e} (scale {scale:.3e})"


SCENES = {
    "folded20": lambda: cases.folded(20),
    "garment": lambda: cases.garment(),
    "random300": lambda: cases.random_soup(300),
    "shared200": lambda: cases.shared_vertex_soup(200),
}


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("motion", [False, True])
def test_bvh_bit_exact(ccc, oracle, name, motion):
    t0, t1, _ = SCENES[name]()
    bb, tree = ccc.bvh_motion(_t(t0[None]), _t(t1[None])) if motion else ccc.bvh(_t(t0[None]))
    rb, rt, codes, ids = oracle.bvh(t0, t1 if motion else None)
    assert (codes == 0).all()
    np.testing.assert_array_equal(_np(tree)[0], rt)
    np.testing.assert_array_equal(_np(bb)[0], rb)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 1000, 4097])
def test_bvh_sizes(ccc, oracle, n):
    t0, t1, _ = cases.random_soup(n, seed=n)
    bb, tree = ccc.bvh_motion(_t(t0[None]), _t(t1[None]))
    rb, rt, _, _ = oracle.bvh(t0, t1)
    np.testing.assert_array_equal(_np(tree)[0], rt)
    np.testing.assert_array_equal(_np(bb)[0], rb)


def test_bvh_batch_and_fp32(ccc, oracle):
    a = cases.random_soup(257, seed=1, dtype=np.float32)
    b = cases.random_soup(257, seed=2, dtype=np.float32)
    t0 = np.stack([a[0], b[0]])
    t1 = np.stack([a[1], b[1]])
    bb, tree = ccc.bvh_motion(_t(t0), _t(t1))
    for i in range(2):
        rb, rt, _, _ = oracle.bvh(t0[i], t1[i])
        np.testing.assert_array_equal(_np(tree)[i], rt)
        np.testing.assert_array_equal(_np(bb)[i], rb)


def _ccd_ref(oracle, t0, t1):
    rb, rt, _, _ = oracle.bvh(t0, t1)
    pairs, _ = oracle.traverse(rb, rt, 0.0)
    coll, roots, counts = oracle.ccd(t0, t1, pairs)
    return rb, rt, coll, roots, counts


def _compare_collisions(gc, gr, coll, roots):
    g = valid_rows(gc)
    gr = gr[: g.shape[0]].reshape(-1)
    assert g.shape[0] == coll.shape[0], f"{g.shape[0]} GPU collisions vs {coll.shape[0]} oracle"
    if coll.shape[0] == 0:
        return
    ka = np.column_stack([g, gr])
    kb = np.column_stack([coll, roots])
    ia = np.lexsort((gr, g[:, 2], g[:, 1], g[:, 0]))
    ib = np.lexsort((roots, coll[:, 2], coll[:, 1], coll[:, 0]))
    np.testing.assert_array_equal(ka[ia][:, :3].astype(np.int64), kb[ib][:, :3].astype(np.int64))
    # the oracle and the kernels share the portable cbrt/acos/sin/cos: roots agree bit for bit
    np.testing.assert_array_equal(ka[ia][:, 3], kb[ib][:, 3])


@pytest.mark.parametrize("name", list(SCENES))
def test_continuous_collisions(ccc, oracle, name):
    t0, t1, _ = SCENES[name]()
    rb, rt, coll, roots, counts = _ccd_ref(oracle, t0, t1)
    mc = int(np.ceil(counts["candidates"] / len(t0))) + 1
    mcol = int(np.ceil(max(counts["collisions"], 1) / len(t0))) + 1
    gc, gr = ccc.find_collisions_continuous(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), mc, mcol)
    K, M, V, R = ccc.last_counts()[0]
    assert (K, M, V, R) == (counts["candidates"], counts["subprims"], counts["roots"], counts["collisions"])
    _compare_collisions(_np(gc)[0], _np(gr)[0], coll, roots)


@pytest.mark.parametrize("name", list(SCENES))
def test_compute_impulses(ccc, oracle, name):
    t0, t1, m = SCENES[name]()
    rb, rt, coll, roots, counts = _ccd_ref(oracle, t0, t1)
    dv, dx, cnt = oracle.impulses(t0, t1, m, coll, roots)
    mc = int(np.ceil(counts["candidates"] / len(t0))) + 1
    g = ccc.compute_impulses(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), _t(m[None]), mc, 16)
    assert ccc.last_counts()[0][3] == counts["collisions"]
    np.testing.assert_array_equal(_np(g[2])[0], cnt)
    _check_close(_np(g[0])[0], dv, "dv")
    _check_close(_np(g[1])[0], dx, "dx")


def test_garment_full_size_first_pass(ccc, oracle):
    """The garment-on-body scene at full size (23k-triangle body + 99k-triangle sleeve): one
    impulse pass on the GPU against the oracle's multithreaded pass over the same soups:
    counts equal, cnt exact, dv / dx within 1e-9 relative."""
    from cccollisions_amd import scenes
    sc = scenes.garment_on_body()
    t0, t1 = sc.soups()
    m = sc.corner_masses()
    dv, dx, cnt, counts = oracle.impulse_pass(t0, t1, m, nthreads=16)
    assert counts["collisions"] > 100000
    bb, tree = ccc.bvh_motion(_t(t0[None]), _t(t1[None]))
    mc = int(np.ceil(counts["candidates"] / len(t0))) + 1
    g = ccc.compute_impulses(bb, tree, _t(t0[None]), _t(t1[None]), _t(m[None]), mc, 16)
    K, M, V, R = ccc.last_counts()[0]
    assert (K, M, V, R) == (counts["candidates"], counts["subprims"], counts["roots"], counts["collisions"])
    np.testing.assert_array_equal(_np(g[2])[0], cnt)
    for a, b, name in ((_np(g[0])[0], dv, "dv"), (_np(g[1])[0], dx, "dx")):
        np.testing.assert_array_equal(np.isfinite(a), np.isfinite(b))
        fin = np.isfinite(b)
        _check_close(np.where(fin, a, 0.0), np.where(fin, b, 0.0), name)


PACKED_MESHES = [
    lambda: cases.folded(8, seed=1),
    lambda: cases.random_soup(300, seed=8),
    lambda: cases.random_soup(1, seed=9),
    lambda: cases.folded(14, seed=3),
    lambda: cases.shared_vertex_soup(60),
    lambda: cases.random_soup(2, seed=10),
]


def test_packed_meshes_match_single_mesh_oracle(ccc, oracle):
    """compute_impulses_packed over meshes of different sizes (1..~400 triangles): each
    mesh's cnt equals its own single-mesh oracle bit for bit, dv/dx to 1e-9, and the batch
    counts are the sums of the per-mesh counts (no pair crosses a mesh)."""
    meshes = [f() for f in PACKED_MESHES]
    offs = np.cumsum([0] + [len(m[0]) for m in meshes]).tolist()
    t0 = np.concatenate([m[0] for m in meshes])
    t1 = np.concatenate([m[1] for m in meshes])
    ms = np.concatenate([m[2] for m in meshes])
    dv, dx, cnt = ccc.compute_impulses_packed(_t(t0), _t(t1), _t(ms), offs, 512)
    K, M, V, R = ccc.last_counts()[0]
    tot = np.zeros(4, np.int64)
    for i, (a0, a1, am) in enumerate(meshes):
        lo, hi = offs[i], offs[i + 1]
        rdv, rdx, rcnt, _, _, rc = oracle.compute_impulses(*oracle.bvh(a0, a1)[:2], a0, a1, am)
        tot += [rc["candidates"], rc["subprims"], rc["roots"], rc["collisions"]]
        np.testing.assert_array_equal(_np(cnt)[lo:hi], rcnt)
        _check_close(_np(dv)[lo:hi], rdv, f"dv mesh {i}")
        _check_close(_np(dx)[lo:hi], rdx, f"dx mesh {i}")
    assert (K, M, V, R) == tuple(tot)
    assert R > 0


def test_packed_alone_equals_in_the_mix(ccc):
    """Each mesh alone (plain compute_impulses on its own BVH) and inside the packed batch:
    same counts of impulses per corner, dv/dx within 1e-9."""
    meshes = [f() for f in PACKED_MESHES]
    offs = np.cumsum([0] + [len(m[0]) for m in meshes]).tolist()
    cat = [np.concatenate([m[k] for m in meshes]) for k in range(3)]
    dv, dx, cnt = ccc.compute_impulses_packed(_t(cat[0]), _t(cat[1]), _t(cat[2]), offs, 512)
    for i, (a0, a1, am) in enumerate(meshes):
        bb, tree = ccc.bvh_motion(_t(a0[None]), _t(a1[None]))
        sdv, sdx, scnt = ccc.compute_impulses(bb, tree, _t(a0[None]), _t(a1[None]), _t(am[None]), 512)
        lo, hi = offs[i], offs[i + 1]
        np.testing.assert_array_equal(_np(cnt)[lo:hi], _np(scnt)[0])
        _check_close(_np(dv)[lo:hi], _np(sdv)[0], f"dv mesh {i}")
        _check_close(_np(dx)[lo:hi], _np(sdx)[0], f"dx mesh {i}")


def test_packed_rejects_bad_offsets(ccc):
    t = _t(np.zeros((4, 3, 3)))
    m = _t(np.ones((4, 3)))
    with pytest.raises(RuntimeError, match="offsets"):
        ccc.compute_impulses_packed(t, t, m, [0, 2, 3])
    with pytest.raises(RuntimeError, match="strictly increasing"):
        ccc.compute_impulses_packed(t, t, m, [0, 2, 2, 4])


def test_impulses_from_collisions_with_padding(ccc, oracle):
    t0, t1, m = cases.folded(20)
    rb, rt, coll, roots, counts = _ccd_ref(oracle, t0, t1)
    cap = coll.shape[0] + 17
    c = -np.ones((1, cap, 3), np.int64)
    r = -np.ones((1, cap, 1))
    c[0, : coll.shape[0]] = coll
    r[0, : coll.shape[0], 0] = roots
    g = ccc.compute_impulses_from_collisions(_t(c), _t(r), _t(t0[None]), _t(t1[None]), _t(m[None]))
    dv, dx, cnt = oracle.impulses(t0, t1, m, coll, roots)
    np.testing.assert_array_equal(_np(g[2])[0], cnt)
    _check_close(_np(g[0])[0], dv, "dv")
    _check_close(_np(g[1])[0], dx, "dx")


def test_candidate_capacity_grows(ccc, oracle):
    t0, t1, _ = cases.folded(20)
    rb, rt, coll, roots, counts = _ccd_ref(oracle, t0, t1)
    assert counts["candidates"] > len(t0)          # default capacity 1 per triangle overflows
    with pytest.warns(UserWarning):
        gc, gr = ccc.find_collisions_continuous(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), 1, 64)
    _compare_collisions(_np(gc)[0], _np(gr)[0], coll, roots)


def test_output_truncation_warns(ccc, oracle):
    t0, t1, _ = cases.folded(20)
    rb, rt, coll, roots, counts = _ccd_ref(oracle, t0, t1)
    assert counts["collisions"] > len(t0)
    with pytest.warns(UserWarning):
        gc, gr = ccc.find_collisions_continuous(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), 64, 1)
    g = valid_rows(_np(gc)[0])
    assert g.shape[0] == len(t0)
    ref = {tuple(r) for r in coll}
    assert all(tuple(r) in ref for r in g)


def test_batched_continuous(ccc, oracle):
    a = cases.random_soup(300, seed=11)
    b = cases.random_soup(300, seed=12)
    t0 = np.stack([a[0], b[0]])
    t1 = np.stack([a[1], b[1]])
    bb, tree = ccc.bvh_motion(_t(t0), _t(t1))
    gc, gr = ccc.find_collisions_continuous(bb, tree, _t(t0), _t(t1), 64, 64)
    for i in range(2):
        _, _, coll, roots, _ = _ccd_ref(oracle, t0[i], t1[i])
        _compare_collisions(_np(gc)[i], _np(gr)[i], coll, roots)


def test_fp32_continuous(ccc, oracle):
    t0, t1, m = cases.random_soup(300, seed=7, dtype=np.float32)
    rb, rt, coll, roots, counts = _ccd_ref(oracle, t0, t1)
    gc, gr = ccc.find_collisions_continuous(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), 64, 64)
    g = sorted_rows(valid_rows(_np(gc)[0]))
    np.testing.assert_array_equal(g, sorted_rows(coll))


def test_fp32_queries(ccc, oracle):
    """fp32 soups through the other queries: pair lists, collision rows and roots bit-exact
    against the oracle's fp32 instantiation (the cubic is solved in double for both dtypes,
    then rounded to float); impulses within 1e-5 relative (fp32 sums in another atomic
    order: a few float ulps)."""
    t0, t1, m = cases.random_soup(300, seed=7, dtype=np.float32)
    m = m.astype(np.float32)
    rb, rt, coll, roots, counts = _ccd_ref(oracle, t0, t1)
    assert counts["collisions"] > 0
    gc, gr = ccc.find_collisions_continuous(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), 64, 64)
    _compare_collisions(_np(gc)[0], _np(gr)[0], coll, roots)
    dv, dx, cnt = oracle.impulses(t0, t1, m, coll, roots)
    g = ccc.compute_impulses(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), _t(m[None]), 64, 16)
    np.testing.assert_array_equal(_np(g[2])[0], cnt)
    for a, b in ((_np(g[0])[0], dv), (_np(g[1])[0], dx)):
        fin = np.isfinite(b)
        np.testing.assert_array_equal(np.isfinite(a), fin)
        scale = float(np.abs(b[fin]).max())
        np.testing.assert_allclose(a[fin], b[fin], rtol=1e-5, atol=1e-6 * scale)
    ref = oracle.find_collisions(rb, rt, t0, 1e-3)
    gs = ccc.find_collisions(_t(rb[None]), _t(rt[None]), _t(t0[None]), 1e-3, 64)
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(gs)[0])), sorted_rows(ref))
    pairs, _ = oracle.traverse(rb, rt, 0.0)
    gm = ccc.find_mincol_cand(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), 64)
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(gm)[0])), sorted_rows(oracle.mincol(t0, t1, pairs)))
    tri = cases.proximity_fold(24, gap=1e-3).astype(np.float32)
    pb, pt, _, _ = oracle.bvh(tri)
    rows, pc = oracle.find_proximity(pb, pt, tri, 2e-3)
    assert pc["collisions"] > 0
    mc = int(np.ceil(max(pc["candidates"], pc["collisions"]) / len(tri))) + 1
    gp = ccc.find_proximity(_t(pb[None]), _t(pt[None]), _t(tri[None]), 2e-3, mc)[0]
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(gp)[0])), sorted_rows(rows))


def test_partial(ccc, oracle):
    t0, t1, m = cases.random_soup(300, seed=9)
    rb, rt, _, _ = oracle.bvh(t0, t1)
    rng = np.random.default_rng(0)
    mask = rng.random(300) < 0.3
    pairs = oracle.traverse_partial(rb, rt, mask)
    coll, roots, counts = oracle.ccd(t0, t1, pairs)
    import torch
    tc = torch.from_numpy(mask[None].copy()).cuda()
    gc, gr = ccc.find_collisions_continuous_partial(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), tc, 64, 64)
    assert ccc.last_counts()[0][0] == counts["candidates"]
    _compare_collisions(_np(gc)[0], _np(gr)[0], coll, roots)
    dv, dx, cnt = oracle.impulses(t0, t1, m, coll, roots)
    g = ccc.compute_impulses_partial(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), _t(m[None]), tc, 64, 16)
    np.testing.assert_array_equal(_np(g[2])[0], cnt)
    _check_close(_np(g[0])[0], dv, "dv")


@pytest.mark.parametrize("gap", [1e-3, 5e-3])
def test_proximity(ccc, oracle, gap):
    tri = cases.proximity_fold(24, gap=gap)
    rb, rt, _, _ = oracle.bvh(tri)
    rows, counts = oracle.find_proximity(rb, rt, tri, 2e-3)
    mc = int(np.ceil(max(counts["candidates"], counts["collisions"]) / len(tri))) + 1
    g = ccc.find_proximity(_t(rb[None]), _t(rt[None]), _t(tri[None]), 2e-3, mc)[0]
    K, M, V, R = ccc.last_counts()[0]
    assert (K, M, R) == (counts["candidates"], counts["subprims"], counts["collisions"])
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(g)[0])), sorted_rows(rows))


def test_proximity_config1_shape(ccc, oracle):
    from cccollisions_amd import scenes
    v, f = scenes.proximity_cloth(64, seed=1)
    tri = v[f]
    bb, tree = ccc.bvh(_t(tri[None]))
    rb, rt, _, _ = oracle.bvh(tri)
    np.testing.assert_array_equal(_np(tree)[0], rt)
    rows, counts = oracle.find_proximity(rb, rt, tri, 2e-3)
    g = ccc.find_proximity(bb, tree, _t(tri[None]), 2e-3, 16)[0]
    K, M, V, R = ccc.last_counts()[0]
    assert (K, M, R) == (counts["candidates"], counts["subprims"], counts["collisions"])
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(g)[0])), sorted_rows(rows))


@pytest.mark.parametrize("thr", [0.0, 1e-3])
def test_static_sat(ccc, oracle, thr):
    t0, _, _ = cases.random_soup(400, seed=4)
    rb, rt, _, _ = oracle.bvh(t0)
    ref = oracle.find_collisions(rb, rt, t0, thr)
    g = ccc.find_collisions(_t(rb[None]), _t(rt[None]), _t(t0[None]), thr, 64)
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(g)[0])), sorted_rows(ref))


TE_SCENES = {
    "random300": lambda dt: cases.random_soup(300, seed=7, dtype=dt)[0],
    "shared200": lambda dt: cases.shared_vertex_soup(200, dtype=dt)[0],
    "fold": lambda dt: cases.proximity_fold(16, dtype=dt),
}


def _te_rows(a, width):
    a = np.asarray(a).reshape(-1, width)
    return a[a[:, 0] >= 0]


@pytest.mark.parametrize("name", list(TE_SCENES))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_triangle_edge_queries(ccc, oracle, name, dtype):
    """All three triangle-edge queries: row multisets bit-exact, edge coefficients (per row)
    bit-exact, pair/mask rows of candidates2 bit-exact; threshold is ignored as in the
    reference."""
    t0 = TE_SCENES[name](dtype)
    n = t0.shape[0]
    rb, rt, _, _ = oracle.bvh(t0)
    # capacity for every candidate pair and every row (the shared-vertex soup needs ~190 rows
    # per triangle)
    pairs, _ = oracle.traverse(rb, rt, 0.0)
    mc = max(64, 6 * int(np.ceil(pairs.shape[0] / n)) + 1)
    g, gc = ccc.find_triangle_edge_collisions(_t(rb[None]), _t(rt[None]), _t(t0[None]), 0.5, mc)
    rows, coefs = oracle.find_triangle_edge(rb, rt, t0, "collisions")
    gr = _np(g)[0]
    keep = gr[:, 0] >= 0
    got = sorted_rows(np.concatenate([gr[keep], _np(gc)[0][keep].view(np.int64 if dtype == np.float64 else np.int32)
                                      .astype(np.int64)], axis=1))
    want = sorted_rows(np.concatenate([rows, coefs.reshape(-1, 1).view(np.int64 if dtype == np.float64 else np.int32)
                                       .astype(np.int64)], axis=1))
    np.testing.assert_array_equal(got, want)
    assert (_np(gc)[0][~keep] == -1).all()

    g = ccc.find_triangle_edge_candidates(_t(rb[None]), _t(rt[None]), _t(t0[None]), 0.0, mc)
    ref = oracle.find_triangle_edge(rb, rt, t0, "candidates")
    np.testing.assert_array_equal(sorted_rows(_te_rows(_np(g)[0], 4)), sorted_rows(ref))

    gp, gm = ccc.find_triangle_edge_candidates2(_t(rb[None]), _t(rt[None]), _t(t0[None]), 0.0, mc)
    rp, rm = oracle.find_triangle_edge(rb, rt, t0, "candidates2")
    gp, gm = _np(gp)[0], _np(gm)[0]
    keep = gp[:, 0] >= 0
    np.testing.assert_array_equal(sorted_rows(np.concatenate([gp[keep], gm[keep]], axis=1)),
                                  sorted_rows(np.concatenate([rp, rm.astype(np.int64)], axis=1)))
    assert (gm[~keep] == 0).all() and keep.sum() == rp.shape[0]
    assert ref.shape[0] > 0 and n > 0


def test_triangle_edge_batch_and_truncation(ccc, oracle):
    a = cases.random_soup(300, seed=7)[0]
    b = cases.shared_vertex_soup(300)[0]
    t0 = np.stack([a, b])
    bbs, trees = [], []
    for i in range(2):
        rb, rt, _, _ = oracle.bvh(t0[i])
        bbs.append(rb)
        trees.append(rt)
    refs = [oracle.find_triangle_edge(bbs[i], trees[i], t0[i], "candidates") for i in range(2)]
    kmax = max(oracle.traverse(bbs[i], trees[i], 0.0)[0].shape[0] for i in range(2))
    mc = max(6 * int(np.ceil(kmax / 300)) + 1, int(np.ceil(max(r.shape[0] for r in refs) / 300)) + 1)
    g = ccc.find_triangle_edge_candidates(_t(np.stack(bbs)), _t(np.stack(trees)), _t(t0), 0.0, mc)
    for i in range(2):
        np.testing.assert_array_equal(sorted_rows(_te_rows(_np(g)[i], 4)), sorted_rows(refs[i]))
    # 1 row per triangle of capacity: the shared-vertex soup has more candidate rows
    ref = refs[1]
    assert ref.shape[0] > 300
    with pytest.warns(UserWarning):
        g = ccc.find_triangle_edge_candidates(_t(bbs[1][None]), _t(trees[1][None]), _t(t0[1][None]), 0.0, 1)
    got = _te_rows(_np(g)[0], 4)
    assert got.shape[0] == 300
    want = {tuple(r) for r in ref.tolist()}
    assert all(tuple(r) in want for r in got.tolist())


def test_mincol(ccc, oracle):
    t0, t1, _ = cases.random_soup(300, seed=3)
    rb, rt, _, _ = oracle.bvh(t0, t1)
    pairs, _ = oracle.traverse(rb, rt, 0.0)
    ref = oracle.mincol(t0, t1, pairs)
    mc = int(np.ceil(pairs.shape[0] / 300)) + 1
    g = ccc.find_mincol_cand(_t(rb[None]), _t(rt[None]), _t(t0[None]), _t(t1[None]), mc)
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(g)[0])), sorted_rows(ref))


def test_static_pairs_have_no_roots(ccc, oracle):
    """t1 == t0: every cubic is c3 = c2 = c1 = 0 and falls in the linear branch, whose root
    the reference never stores (collisions_continuous.cu:532-536)."""
    t0, _, m = cases.folded(16)
    bb, tree = ccc.bvh_motion(_t(t0[None]), _t(t0[None]))
    gc, gr = ccc.find_collisions_continuous(bb, tree, _t(t0[None]), _t(t0[None]), 64, 16)
    K, M, V, R = ccc.last_counts()[0]
    assert K > 0 and M > 0 and V == 0 and R == 0
    assert (_np(gc) == -1).all() and (_np(gr) == -1).all()


def _shuffled_leaf_tree(rb, rt, seed):
    """A valid wire tree over the same leaves whose subtrees no longer cover contiguous
    position ranges: leaf rows are permuted (each keeps its box and triangle id, its parent
    now points at the new row) and the internal boxes are refit."""
    n = (rt.shape[0] + 1) // 2
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)                       # leaf row n-1+k moves to row n-1+perm[k]
    bb, tree = rb.copy(), rt.copy()
    for k in range(n):
        src, dst = n - 1 + k, n - 1 + perm[k]
        bb[dst], tree[dst] = rb[src], rt[src]
    for i in range(n - 1):
        for c in (1, 2):
            if rt[i, c] >= n - 1:
                tree[i, c] = n - 1 + perm[rt[i, c] - (n - 1)]
    order = sorted(range(n - 1), key=lambda i: -_depth(tree, i))
    for i in order:
        l, r = tree[i, 1], tree[i, 2]
        bb[i, :3] = np.minimum(bb[l, :3], bb[r, :3])
        bb[i, 3:] = np.maximum(bb[l, 3:], bb[r, 3:])
    return bb, tree


def _depth(tree, i):
    d = 0
    while tree[i, 0] >= 0:
        i, d = tree[i, 0], d + 1
    return d


def test_unordered_tree_takes_the_climb(ccc, oracle):
    """Trees from other builders need not cover contiguous ranges; the rightmost leaves then
    come from the atomic climb and the pairs still equal the oracle's on the same tree."""
    t0, t1, _ = cases.random_soup(300, seed=13)
    rb, rt, _, _ = oracle.bvh(t0, t1)
    bb, tree = _shuffled_leaf_tree(rb, rt, 3)
    pairs, _ = oracle.traverse(bb, tree, 0.0)
    ref_pairs, _ = oracle.traverse(rb, rt, 0.0)
    np.testing.assert_array_equal(sorted_rows(pairs), sorted_rows(ref_pairs))   # topology independent
    coll, roots, counts = oracle.ccd(t0, t1, pairs)
    mc = int(np.ceil(counts["candidates"] / 300)) + 1
    gc, gr = ccc.find_collisions_continuous(_t(bb[None]), _t(tree[None]), _t(t0[None]), _t(t1[None]), mc, 64)
    assert ccc.last_counts()[0][0] == counts["candidates"]
    _compare_collisions(_np(gc)[0], _np(gr)[0], coll, roots)


def _left_spine_tree(t0, t1, depth):
    """A valid wire tree whose left spine is `depth` internal nodes long, each with a
    two-leaf subtree as its right child (the bottom one has two): descending the spine
    keeps every right subtree pending, so a traversal's stack grows with the depth.
    Leaves (n = 2 depth + 2) are numbered in depth-first order; boxes are refit."""
    n = 2 * depth + 2
    assert t0.shape[0] == n
    rows = 2 * n - 1
    tree = np.full((rows, 4), -1, np.int64)
    spine = list(range(depth))                   # S_i = row i, S_0 the root
    pair = [depth + i for i in range(depth + 1)]  # P_i = row depth + i
    for i in range(depth - 1):
        tree[spine[i], 1], tree[spine[i + 1], 0] = spine[i + 1], spine[i]
    tree[spine[depth - 1], 1], tree[pair[depth], 0] = pair[depth], spine[depth - 1]
    for i in range(depth):
        tree[spine[i], 2], tree[pair[i], 0] = pair[i], spine[i]
    k = 0
    for i in range(depth, -1, -1):               # depth-first leaf order: P_depth first
        for c in (1, 2):
            leaf = n - 1 + k
            tree[pair[i], c], tree[leaf, 0], tree[leaf, 3] = leaf, pair[i], k
            k += 1
    bb = np.zeros((rows, 6))
    for k in range(n):
        v = np.concatenate([t0[k], t1[k]])
        bb[n - 1 + k, :3], bb[n - 1 + k, 3:] = v.min(0), v.max(0)
    for i in list(reversed(pair)) + list(reversed(spine)):   # children before parents
        l, r = tree[i, 1], tree[i, 2]
        bb[i, :3] = np.minimum(bb[l, :3], bb[r, :3])
        bb[i, 3:] = np.maximum(bb[l, 3:], bb[r, 3:])
    return bb, tree


def _cluster(n, seed):
    """n small moving triangles whose boxes all contain the point c: every pair overlaps."""
    rng = np.random.default_rng(seed)
    c = np.full(3, 0.05)
    t0 = np.empty((n, 3, 3))
    t0[:, 0] = c - rng.uniform(0.01, 0.05, (n, 3))
    t0[:, 1] = c + rng.uniform(0.01, 0.05, (n, 3))
    t0[:, 2] = c + rng.uniform(-0.05, 0.05, (n, 3))
    t1 = t0 + rng.normal(0.0, 0.01, (n, 3, 3))
    return t0, t1


@pytest.mark.parametrize("depth", [24, 60])
def test_deep_tree_reruns_with_the_deep_stack(ccc, oracle, depth):
    """A left spine keeps one subtree pending per two levels of the four-wide traversal: 24
    levels fit the short stack (24 entries), 60 overflow it below the frontier, and the
    traversal runs again with the deep stack.  Both find every pair the oracle finds."""
    n = 2 * depth + 2
    t0, t1 = _cluster(n, seed=21)
    bb, tree = _left_spine_tree(t0, t1, depth)
    pairs, total = oracle.traverse(bb, tree, 0.0)
    assert total == n * (n - 1) // 2
    np.testing.assert_array_equal(sorted_rows(pairs), sorted_rows(oracle.bruteforce_pairs(bb, tree, 0.0)))
    coll, roots, counts = oracle.ccd(t0, t1, pairs)
    gc, gr = ccc.find_collisions_continuous(_t(bb[None]), _t(tree[None]), _t(t0[None]), _t(t1[None]), 64, 64)
    assert ccc.last_counts()[0][0] == counts["candidates"] == total
    _compare_collisions(_np(gc)[0], _np(gr)[0], coll, roots)


def test_tree_deeper_than_the_deep_stack_raises(ccc, oracle):
    """Beyond the deep stack (an 80-level spine: 36 pending subtrees below the start points
    with four-wide nodes, more than the 32 entries) the call fails with an error instead of
    returning a partial pair list."""
    t0, t1 = _cluster(162, seed=22)
    bb, tree = _left_spine_tree(t0, t1, 80)
    with pytest.raises(RuntimeError, match="deeper than the traversal stack"):
        ccc.find_collisions_continuous(_t(bb[None]), _t(tree[None]), _t(t0[None]), _t(t1[None]), 64, 64)


def _renumbered_internal_tree(rb, rt, seed):
    """The same tree with its internal rows other than the root (row 0) renumbered at random:
    the nodes of a subtree no longer lie close together in the row order."""
    n = (rt.shape[0] + 1) // 2
    rng = np.random.default_rng(seed)
    remap = np.arange(2 * n - 1)                     # old row -> new row
    remap[1:n - 1] = 1 + rng.permutation(n - 2)
    bb, tree = np.empty_like(rb), np.empty_like(rt)
    bb[remap], tree[remap] = rb, rt
    tree[:, 0] = np.where(tree[:, 0] >= 0, remap[np.maximum(tree[:, 0], 0)], tree[:, 0])
    tree[:n - 1, 1:3] = remap[tree[:n - 1, 1:3]]
    return bb, tree


def test_far_apart_subtree_nodes_take_the_deep_stack(ccc, oracle):
    """The short traversal stack holds 16-bit offsets from the start point's node; in a tree
    of 40,000 leaves with internal rows numbered at random those offsets do not fit, the
    traversal runs again with the deep stack, and the pairs still equal the oracle's."""
    n = 40000
    rng = np.random.default_rng(31)
    c = rng.random((n, 3))
    # triangles numbered in Z order of their centres: the reference's tree over triangle
    # indices then has tight boxes
    cell = np.minimum((c * 1024).astype(np.int64), 1023)
    key = np.zeros(n, np.int64)
    for bit in range(10):
        for a in range(3):
            key |= ((cell[:, a] >> bit) & 1) << (3 * bit + a)
    t0 = c[np.argsort(key, kind="stable")][:, None, :] + 0.01 * rng.standard_normal((n, 3, 3))
    t1 = t0 + 0.005 * rng.standard_normal((n, 3, 3))
    rb, rt, _, _ = oracle.bvh(t0, t1)
    bb, tree = _renumbered_internal_tree(rb, rt, 4)
    # premise: below the start points (depth 7), internal rows lie more than 32767 rows
    # away from their start point's row
    level, far = [0], 0
    for _ in range(7):
        level = [c for r in level for c in (tree[r, 1], tree[r, 2]) if r < n - 1] + [r for r in level if r >= n - 1]
    for r0 in level:
        todo = [r0]
        while todo:
            r = todo.pop()
            if r < n - 1:
                far = max(far, abs(int(r) - int(r0)))
                todo += [tree[r, 1], tree[r, 2]]
    assert far > 32767
    pairs, _ = oracle.traverse(bb, tree, 0.0)
    np.testing.assert_array_equal(sorted_rows(pairs), sorted_rows(oracle.traverse(rb, rt, 0.0)[0]))
    coll, roots, counts = oracle.ccd(t0, t1, pairs)
    assert counts["candidates"] > n and counts["collisions"] > 0
    mc = int(np.ceil(counts["candidates"] / n)) + 8
    gc, gr = ccc.find_collisions_continuous(_t(bb[None]), _t(tree[None]), _t(t0[None]), _t(t1[None]), mc, 4096)
    assert ccc.last_counts()[0][0] == counts["candidates"]
    _compare_collisions(_np(gc)[0], _np(gr)[0], coll, roots)


def test_single_triangle_queries(ccc, oracle):
    """N = 1 (the reference cannot build it): a one-leaf tree, no pairs, padded outputs and
    zero impulses from every query."""
    t0, t1, m = cases.random_soup(1, seed=3)
    m = np.ones((1, 3))
    bb, tree = ccc.bvh_motion(_t(t0[None]), _t(t1[None]))
    rb, rt, _, _ = oracle.bvh(t0, t1)
    np.testing.assert_array_equal(_np(tree)[0], rt)
    gc, gr = ccc.find_collisions_continuous(bb, tree, _t(t0[None]), _t(t1[None]), 4, 4)
    assert (_np(gc) == -1).all() and (_np(gr) == -1).all()
    assert ccc.last_counts()[0][0] == 0
    dv, dx, cnt = ccc.compute_impulses(bb, tree, _t(t0[None]), _t(t1[None]), _t(m[None]), 4, 4)
    assert not _np(dv).any() and not _np(dx).any() and not _np(cnt).any()
    assert (_np(ccc.find_collisions(bb, tree, _t(t0[None]), 0.0, 4)) == -1).all()
    [close] = ccc.find_proximity(bb, tree, _t(t0[None]), 2e-3, 4)
    assert (_np(close) == -1).all()


def test_queries_without_candidates(ccc, oracle):
    """Triangles far apart: no candidate pair, so every narrow phase runs over an empty list
    (each launches a grid of idle blocks) and the outputs stay padded."""
    t0, t1, m = cases.random_soup(5, seed=4)
    t0 = t0 * 1e-3 + 10.0 * np.arange(5)[:, None, None]
    t1 = t0 + 1e-4
    bb, tree = ccc.bvh_motion(_t(t0[None]), _t(t1[None]))
    assert oracle.traverse(*oracle.bvh(t0, t1)[:2], 0.0)[1] == 0
    assert (_np(ccc.find_collisions(bb, tree, _t(t0[None]), 0.0, 4)) == -1).all()
    gc, gr = ccc.find_collisions_continuous(bb, tree, _t(t0[None]), _t(t1[None]), 4, 4)
    assert (_np(gc) == -1).all() and ccc.last_counts()[0][0] == 0
    assert (_np(ccc.find_mincol_cand(bb, tree, _t(t0[None]), _t(t1[None]), 4)) == -1).all()
    coll, coeff = ccc.find_triangle_edge_collisions(bb, tree, _t(t0[None]), 0.0, 4)
    assert (_np(coll) == -1).all()
    assert (_np(ccc.find_triangle_edge_candidates(bb, tree, _t(t0[None]), 0.0, 4)) == -1).all()
    pairs, masks = ccc.find_triangle_edge_candidates2(bb, tree, _t(t0[None]), 0.0, 4)
    assert (_np(pairs) == -1).all()
    dv, dx, cnt = ccc.compute_impulses(bb, tree, _t(t0[None]), _t(t1[None]), _t(m[None]), 4, 4)
    assert not _np(cnt).any()


def test_bad_tree_raises(ccc, oracle):
    t0, t1, _ = cases.random_soup(50, seed=1)
    rb, rt, _, _ = oracle.bvh(t0, t1)
    bad = rt.copy()
    bad[3, 1] = 5     # child whose parent field does not point back
    with pytest.raises(RuntimeError, match="malformed BVH tree"):
        ccc.find_collisions_continuous(_t(rb[None]), _t(bad[None]), _t(t0[None]), _t(t1[None]), 64, 16)
