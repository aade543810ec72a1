"""The synthetic scenes behind the benchmark configurations.  CPU only."""
import numpy as np

from cccollisions_amd import scenes


def test_garment_on_body_sizes_and_motion():
    sc = scenes.garment_on_body()
    body = 12 * 2 * 31 * 31
    assert sc.num_triangles == body + 2 * 223 * 223          # ~24k body + ~100k sleeve
    t0, t1 = sc.soups()
    np.testing.assert_array_equal(t0[:body], t1[:body])         # the body is static
    moved = np.linalg.norm(t1[body:] - t0[body:], axis=-1)
    assert moved.max() > 0 and moved.max() < 0.1


def test_garment_collisions_involve_the_sleeve(oracle):
    """Body-body pairs have t1 = t0: their cubics are linear and the reference drops linear
    roots, so every collision has a sleeve triangle."""
    sc = scenes.garment_on_body(8, 24, 23)
    body = 12 * 2 * 7 * 7
    t0, t1 = sc.soups()
    rb, rt, _, _ = oracle.bvh(t0, t1)
    pairs, _ = oracle.traverse(rb, rt, 0.0)
    coll, roots, counts = oracle.ccd(t0, t1, pairs)
    assert counts["collisions"] > 100
    assert (coll[:, 1] >= body).all()                          # y = max index is a sleeve triangle


def test_batched_cloths_sizes():
    sims = scenes.batched_cloths(8, seed=4)
    sizes = [s.num_triangles for s in sims]
    assert min(sizes) >= 1500 and max(sizes) <= 210000
    assert len(set(sizes)) > 4
