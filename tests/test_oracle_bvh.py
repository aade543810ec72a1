"""Pinning the oracle's BVH build (src/build_bvh.cu) with closed forms and independent
constructions.  CPU only."""
import numpy as np
import pytest

from tests import cases


def index_radix_tree(n):
    """The binary radix tree over the keys 0..n-1 built from its definition: a node covering
    the sorted leaf range [lo, hi] splits after the last key whose prefix agrees with lo on
    the highest bit where lo and hi differ; internal nodes take Karras' numbering (a node's
    left child is internal node `split` or leaf `split`, its right child split + 1).
    Returns (left, right) row indices of the n-1 internal nodes."""
    left = np.full(max(n - 1, 0), -1, np.int64)
    right = np.full(max(n - 1, 0), -1, np.int64)
    stack = [(0, n - 1, 0)] if n > 1 else []
    while stack:
        lo, hi, idx = stack.pop()
        top = (lo ^ hi).bit_length() - 1
        split = ((hi >> top) << top) - 1          # last index with bit `top` equal to lo's
        left[idx] = (n - 1 + split) if split == lo else split
        right[idx] = (n - 1 + split + 1) if split + 1 == hi else split + 1
        if split > lo:
            stack.append((lo, split, split))
        if split + 1 < hi:
            stack.append((split + 1, hi, split + 1))
    return left, right


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 31, 64, 100, 1000, 4097, 7938])
def test_compat_tree_is_index_radix_tree(oracle, n):
    t0, t1, _ = cases.random_soup(n, seed=n)
    bb, tree, codes, ids = oracle.bvh(t0, t1)
    assert (codes == 0).all(), "scene box [-MAX, MAX] makes every Morton code 0"
    np.testing.assert_array_equal(ids, np.arange(n))
    l, r = index_radix_tree(n)
    np.testing.assert_array_equal(tree[: n - 1, 1], l)
    np.testing.assert_array_equal(tree[: n - 1, 2], r)
    # parents point back, root has none, leaves carry ids, internal ids are -1
    if n > 1:
        assert tree[0, 0] == -1
        for i in range(n - 1):
            assert tree[tree[i, 1], 0] == i and tree[tree[i, 2], 0] == i
        assert (tree[: n - 1, 3] == -1).all()
    np.testing.assert_array_equal(tree[n - 1:, 3], np.arange(n))
    assert (tree[n - 1:, 1:3] == -1).all()


def test_literal_search_loops_match_textbook(oracle):
    """build_bvh.cu:100-123 re-tests the first step of each binary search and keeps testing
    t = 1 in the split loop; for monotone prefixes the result equals the textbook search."""
    for n in list(range(2, 300)) + [511, 512, 513, 1000, 2047, 3000]:
        codes = np.zeros(n, np.uint32)
        ids = np.arange(n, dtype=np.int32)
        for i in range(n - 1):
            assert oracle.karras_node(codes, ids, i) == oracle.karras_node(codes, ids, i, literal=True)
    rng = np.random.default_rng(0)
    for n in (17, 100, 777):
        raw = rng.integers(0, 64, n).astype(np.uint32)              # many duplicate keys
        ids = np.argsort(raw, kind="stable").astype(np.int32)       # as after the stable sort
        codes = raw[ids]
        for i in range(n - 1):
            assert oracle.karras_node(codes, ids, i) == oracle.karras_node(codes, ids, i, literal=True)


@pytest.mark.parametrize("motion", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_boxes(oracle, motion, dtype):
    t0, t1, _ = cases.random_soup(500, seed=3, dtype=dtype)
    bb, tree, _, ids = oracle.bvh(t0, t1 if motion else None)
    n = t0.shape[0]
    lo = t0.min(axis=1)
    hi = t0.max(axis=1)
    if motion:
        lo = np.minimum(lo, t1.min(axis=1))
        hi = np.maximum(hi, t1.max(axis=1))
    np.testing.assert_array_equal(bb[n - 1:, :3], lo[ids])
    np.testing.assert_array_equal(bb[n - 1:, 3:], hi[ids])
    for i in range(n - 1):
        l, r = tree[i, 1], tree[i, 2]
        np.testing.assert_array_equal(bb[i, :3], np.minimum(bb[l, :3], bb[r, :3]))
        np.testing.assert_array_equal(bb[i, 3:], np.maximum(bb[l, 3:], bb[r, 3:]))


def _interleave(x, y, z):
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def test_morton_bits(oracle):
    """bvh.cu:54-62: quantise to [0, 1023] (clamped, truncated) and interleave x, y, z."""
    rng = np.random.default_rng(1)
    pts = np.concatenate([rng.random((500, 3)), rng.uniform(-0.5, 1.5, (200, 3)),
                          [[0, 0, 0], [1, 1, 1], [np.nan, 0.5, 0.5], [1023.5 / 1024, 0, 0]]])
    for x, y, z in pts:
        q = [0 if not np.isfinite(v) else int(min(max(v * 1024.0, 0.0), 1023.0)) for v in (x, y, z)]
        assert oracle.morton3d(x, y, z) == _interleave(*q)


def test_fitted_scene_box_codes_sorted(oracle):
    """With a fitted scene box (not the reference's) the codes differ and the sort is
    stable: equal codes keep increasing ids."""
    t0, t1, _ = cases.random_soup(2000, seed=5)
    _, _, codes, ids = oracle.bvh(t0, t1, morton_mode=1)
    assert (np.diff(codes.astype(np.int64)) >= 0).all()
    eq = np.diff(codes.astype(np.int64)) == 0
    assert (np.diff(ids)[eq] > 0).all()
    assert len(np.unique(codes)) > 100


def radix_tree_over_keys(keys):
    """The binary radix tree over sorted distinct integer keys, from its definition: a node
    covering [lo, hi] splits after the last key that agrees with keys[lo] on the highest bit
    where keys[lo] and keys[hi] differ; Karras' numbering as in index_radix_tree."""
    n = len(keys)
    left = np.full(max(n - 1, 0), -1, np.int64)
    right = np.full(max(n - 1, 0), -1, np.int64)
    stack = [(0, n - 1, 0)] if n > 1 else []
    while stack:
        lo, hi, idx = stack.pop()
        top = (keys[lo] ^ keys[hi]).bit_length() - 1
        split = lo
        while split + 1 < hi and not (keys[split + 1] >> top) & 1:
            split += 1
        left[idx] = (n - 1 + split) if split == lo else split
        right[idx] = (n - 1 + split + 1) if split + 1 == hi else split + 1
        if split > lo:
            stack.append((lo, split, split))
        if split + 1 < hi:
            stack.append((split + 1, hi, split + 1))
    return left, right


def _interleave_bits(x, y, z, bits):
    code = 0
    for b in range(bits):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def fitted_codes(t0, t1, bits):
    """Morton codes with the scene box fitted to the (swept) leaf boxes, computed in the
    soups' own dtype with numpy: centroid (v0 + v1 + v2) / 3 (the mean of the t0 and t1
    centroids when moving), normalised by the scene box, scaled by 2^bits, clamped to
    [0, 2^bits - 1] (NaN -> 0), truncated and interleaved."""
    dt = t0.dtype.type
    lo, hi = t0.min(axis=1), t0.max(axis=1)
    c = (t0[:, 0] + t0[:, 1] + t0[:, 2]) / dt(3.0)
    if t1 is not None:
        lo, hi = np.minimum(lo, t1.min(axis=1)), np.maximum(hi, t1.max(axis=1))
        c = (c + (t1[:, 0] + t1[:, 1] + t1[:, 2]) / dt(3.0)) / dt(2.0)
    smin, smax = lo.min(axis=0), hi.max(axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        u = (c - smin) / (smax - smin)
        q = np.fmin(np.fmax(u * dt(2.0 ** bits), dt(0.0)), dt(2.0 ** bits - 1))
    q = q.astype(np.int64)
    return [_interleave_bits(int(a), int(b), int(d), bits) for a, b, d in q]


def test_morton63_bits(oracle):
    """63-bit codes: 21 bits per axis, quantised and interleaved like the 30-bit ones."""
    rng = np.random.default_rng(4)
    pts = np.concatenate([rng.random((500, 3)), rng.uniform(-0.5, 1.5, (200, 3)),
                          [[0, 0, 0], [1, 1, 1], [np.nan, 0.5, 0.5], [1 - 2.0 ** -22, 0, 0]]])
    for x, y, z in pts:
        q = [0 if not np.isfinite(v) else int(min(max(v * 2.0 ** 21, 0.0), 2.0 ** 21 - 1)) for v in (x, y, z)]
        assert oracle.morton63(x, y, z) == _interleave_bits(*q, 21)


@pytest.mark.parametrize("mode,bits", [(1, 10), (2, 21)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("scene", ["random", "folded", "flat"])
def test_fitted_morton_tree(oracle, mode, bits, dtype, scene):
    """Fitted-box modes: codes from numpy, leaf order = numpy's stable argsort of the codes,
    internal nodes = the radix tree over the sorted (code, id) keys (what the Karras search
    with the id tie-break computes), leaf boxes in sorted order."""
    if scene == "random":
        t0, t1, _ = cases.random_soup(1500, seed=6, dtype=dtype)
    elif scene == "folded":
        t0, t1, _ = cases.folded(20, seed=7, dtype=dtype)
    else:   # static flat grid: the scene box has zero extent in z (0/0 -> NaN -> 0)
        t0, _, _ = cases.folded(12, seed=1, dtype=dtype)
        t0 = t0.copy()
        t0[..., 2] = 0.25
        t1 = None
    bb, tree, codes, ids = oracle.bvh(t0, t1, morton_mode=mode)
    n = t0.shape[0]
    want = np.array(fitted_codes(t0, t1, bits), dtype=np.uint64)
    order = np.argsort(want, kind="stable")
    np.testing.assert_array_equal(ids, order)
    np.testing.assert_array_equal(codes, want[order])
    assert len(np.unique(codes)) > n // 10
    keys = [(int(c) << 32) | int(i) for c, i in zip(codes, ids)]
    l, r = radix_tree_over_keys(keys)
    np.testing.assert_array_equal(tree[: n - 1, 1], l)
    np.testing.assert_array_equal(tree[: n - 1, 2], r)
    np.testing.assert_array_equal(tree[n - 1:, 3], ids)
