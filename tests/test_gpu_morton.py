"""Fitted-box Morton trees on the GPU (set_morton_mode 1: 30-bit codes, 2: 63-bit codes)
against the oracle's fitted modes, the radix sort against numpy's stable argsort of the
same codes, and the queries on fitted trees against the reference (compat) tree: the same
candidate, sub-primitive, root and collision counts, the same collision sets and roots,
equal cnt and dv / dx within 1e-9 relative."""
import numpy as np
import pytest

from tests import cases
from tests.conftest import sorted_rows, valid_rows
from tests.test_gpu_parity import _check_close, _compare_collisions, _np, _t
from tests.test_oracle_bvh import fitted_codes

pytestmark = pytest.mark.gpu


@pytest.fixture
def morton(ccc):
    """Sets the Morton mode of the current device's context and restores the reference's."""
    def set_mode(mode):
        ccc.set_morton_mode(mode)
    yield set_mode
    ccc.set_morton_mode(0)


def _shuffled_folded(n, seed):
    t0, t1, m = cases.folded(n, seed=seed)
    perm = np.random.default_rng(seed).permutation(t0.shape[0])
    return t0[perm], t1[perm], m[perm]


SCENES = {
    "random300": lambda: cases.random_soup(300),
    "folded20": lambda: cases.folded(20),
    "garment": lambda: cases.garment(),
    "shuffled24": lambda: _shuffled_folded(24, 3),
    "random300_f32": lambda: cases.random_soup(300, seed=2, dtype=np.float32),
}


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("motion", [False, True])
def test_fitted_tree_bit_exact(ccc, oracle, morton, mode, name, motion):
    t0, t1, _ = SCENES[name]()
    morton(mode)
    bb, tree = ccc.bvh_motion(_t(t0[None]), _t(t1[None])) if motion else ccc.bvh(_t(t0[None]))
    rb, rt, codes, ids = oracle.bvh(t0, t1 if motion else None, morton_mode=mode)
    assert len(np.unique(codes)) > 1
    np.testing.assert_array_equal(_np(tree)[0], rt)
    np.testing.assert_array_equal(_np(bb)[0], rb)


@pytest.mark.parametrize("mode,bits", [(1, 10), (2, 21)])
def test_radix_sort_on_real_keys(ccc, morton, mode, bits):
    """100k triangles (49 sort tiles): the leaf order is numpy's stable argsort of the codes
    (computed with numpy, so the sort is checked independently of the oracle), and a batch
    of two soups sorts each element by its own scene box."""
    a = cases.random_soup(100_000, seed=21)
    b = cases.random_soup(100_000, seed=22)
    morton(mode)
    _, tree = ccc.bvh_motion(_t(np.stack([a[0], b[0]])), _t(np.stack([a[1], b[1]])))
    for i, sc in enumerate((a, b)):
        want = np.argsort(np.array(fitted_codes(sc[0], sc[1], bits), dtype=np.uint64), kind="stable")
        n = sc[0].shape[0]
        np.testing.assert_array_equal(_np(tree)[i][n - 1:, 3], want)


def test_compat_mode_keeps_the_reference_tree(ccc, oracle, morton):
    t0, t1, _ = _shuffled_folded(16, 5)
    morton(2)
    morton(0)
    _, tree = ccc.bvh_motion(_t(t0[None]), _t(t1[None]))
    _, rt, codes, _ = oracle.bvh(t0, t1)
    assert (codes == 0).all()
    np.testing.assert_array_equal(_np(tree)[0], rt)


def test_bad_morton_mode_raises(ccc):
    with pytest.raises(RuntimeError):
        ccc.set_morton_mode(3)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("name", ["folded20", "shuffled24", "garment"])
def test_queries_on_fitted_trees(ccc, oracle, morton, mode, name):
    """compute_impulses and find_collisions_continuous on a fitted tree equal the oracle's
    results on the reference tree; static proximity and SAT give the same row sets."""
    t0, t1, m = SCENES[name]()
    rb, rt, _, _ = oracle.bvh(t0, t1)
    pairs, _ = oracle.traverse(rb, rt, 0.0)
    coll, roots, counts = oracle.ccd(t0, t1, pairs)
    dv, dx, cnt = oracle.impulses(t0, t1, m, coll, roots)
    morton(mode)
    T0, T1, M = _t(t0[None]), _t(t1[None]), _t(m[None])
    bb, tree = ccc.bvh_motion(T0, T1)
    assert not np.array_equal(_np(tree)[0], rt)
    mc = int(np.ceil(counts["candidates"] / len(t0))) + 1
    mcol = int(np.ceil(max(counts["collisions"], 1) / len(t0))) + 1
    gc, gr = ccc.find_collisions_continuous(bb, tree, T0, T1, mc, mcol)
    assert tuple(ccc.last_counts()[0]) == (counts["candidates"], counts["subprims"], counts["roots"], counts["collisions"])
    _compare_collisions(_np(gc)[0], _np(gr)[0], coll, roots)
    g = ccc.compute_impulses(bb, tree, T0, T1, M, mc, 16)
    assert ccc.last_counts()[0][3] == counts["collisions"]
    np.testing.assert_array_equal(_np(g[2])[0], cnt)
    for a, b, nm in ((_np(g[0])[0], dv, "dv"), (_np(g[1])[0], dx, "dx")):
        fin = np.isfinite(b)
        np.testing.assert_array_equal(np.isfinite(a), fin)
        _check_close(np.where(fin, a, 0.0), np.where(fin, b, 0.0), nm)
    # static queries on a static fitted tree
    sb, st = ccc.bvh(T0)
    rsb, rst, _, _ = oracle.bvh(t0)
    rows, pc = oracle.find_proximity(rsb, rst, t0, 2e-3)
    mp = int(np.ceil(max(pc["candidates"], pc["collisions"], 1) / len(t0))) + 1
    gp = ccc.find_proximity(sb, st, T0, 2e-3, mp)[0]
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(gp)[0])), sorted_rows(rows))
    sat = oracle.find_collisions(rsb, rst, t0, 0.0)
    spairs, _ = oracle.traverse(rsb, rst, 0.0)
    ms = int(np.ceil(max(len(spairs), 1) / len(t0))) + 1
    gs = ccc.find_collisions(sb, st, T0, 0.0, ms)
    np.testing.assert_array_equal(sorted_rows(valid_rows(_np(gs)[0])), sorted_rows(sat))
