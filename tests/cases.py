"""Seeded inputs shared by the parity tests (triangle soups as numpy arrays)."""
import numpy as np

from cccollisions_amd import scenes


def folded(n=20, seed=2, dtype=np.float64, **kw):
    sc = scenes.folded_cloth(n, seed=seed, **kw)
    t0, t1 = sc.soups()
    m = sc.corner_masses()
    return t0.astype(dtype), t1.astype(dtype), m.astype(dtype)


def random_soup(n=300, seed=0, motion=0.05, dtype=np.float64):
    """Unstructured triangles in the unit cube with random motion and masses."""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 1, 3))
    t0 = c + 0.06 * rng.standard_normal((n, 3, 3))
    t1 = t0 + motion * rng.standard_normal((n, 3, 3))
    m = rng.uniform(0.5, 2.0, (n, 3))
    return t0.astype(dtype), t1.astype(dtype), m.astype(dtype)


def shared_vertex_soup(n=200, seed=5, dtype=np.float64):
    """Triangles drawn from a small vertex pool so that many pairs share vertices exactly."""
    rng = np.random.default_rng(seed)
    pool0 = rng.random((60, 3))
    pool1 = pool0 + 0.05 * rng.standard_normal((60, 3))
    f = np.stack([rng.choice(60, 3, replace=False) for _ in range(n)])
    return pool0[f].astype(dtype), pool1[f].astype(dtype), np.ones((n, 3), dtype)


def proximity_fold(n=24, seed=3, gap=1e-3, dtype=np.float64):
    """Static folded cloth: the two layers are `gap` apart, so the proximity list is not empty."""
    sc = scenes.folded_cloth(n, seed=seed, gap=gap)
    return sc.x0[sc.faces].astype(dtype)


def garment(body_grid=8, rings=24, segments=23, seed=3, dtype=np.float64):
    """Small garment-on-body scene: static capsule body + a sleeve moving inward."""
    sc = scenes.garment_on_body(body_grid, rings, segments, seed=seed)
    t0, t1 = sc.soups()
    return t0.astype(dtype), t1.astype(dtype), sc.corner_masses().astype(dtype)
