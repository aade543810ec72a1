"""cccollisions_amd -- MI355X-native drop-in for the `cccollisions` PyTorch extension.

    import cccollisions_amd as cccollisions
    bboxes, tree = cccollisions.bvh_motion(tri0, tri1)
    dv, dx, cnt = cccollisions.compute_impulses(bboxes, tree, tri0, tri1, masses)

The 14 entry points keep the reference's names, argument names, defaults, output shapes
and dtypes (src/cccollisions.cpp:426-468).  They run hand-written gfx950 HIP kernels
(libccc_amd.so, C ABI in include/cccollisions_amd.h) on the current HIP stream.  Importing
fails loudly if the native extension has not been built; there is no fallback.
"""
from __future__ import annotations

import torch  # noqa: F401  (loads the HIP runtime the extension links against)

try:
    from . import _cccollisions as _ext
except ImportError as e:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "cccollisions_amd native extension is not built; run `python cccollisions_amd/build.py` "
        f"(or __graft_entry__.build()): {e}") from e

bvh = _ext.bvh
bvh_motion = _ext.bvh_motion
find_collisions = _ext.find_collisions
find_triangle_edge_collisions = _ext.find_triangle_edge_collisions
find_triangle_edge_candidates = _ext.find_triangle_edge_candidates
find_triangle_edge_candidates2 = _ext.find_triangle_edge_candidates2
find_collisions_continuous = _ext.find_collisions_continuous
find_collisions_continuous_partial = _ext.find_collisions_continuous_partial
find_proximity = _ext.find_proximity
find_mincol_cand = _ext.find_mincol_cand
compute_impulses = _ext.compute_impulses
compute_impulses_partial = _ext.compute_impulses_partial
compute_impulses_from_collisions = _ext.compute_impulses_from_collisions
compute_impulses_packed = _ext.compute_impulses_packed
set_fixed_capacity = _ext.set_fixed_capacity
set_pair_chunk = _ext.set_pair_chunk
set_morton_mode = _ext.set_morton_mode
last_counts = _ext.last_counts
set_profiling = _ext.set_profiling
kernel_times = _ext.kernel_times
Sim = _ext.Sim

ENTRY_POINTS = (
    "bvh", "bvh_motion", "find_collisions", "find_triangle_edge_collisions",
    "find_triangle_edge_candidates", "find_triangle_edge_candidates2",
    "find_collisions_continuous", "find_collisions_continuous_partial", "find_proximity",
    "find_mincol_cand", "compute_impulses", "compute_impulses_partial",
    "compute_impulses_from_collisions",
)

# beyond the reference's API: many meshes of different sizes in one pass of launches
EXTENSIONS = ("compute_impulses_packed", "Sim")

__all__ = list(ENTRY_POINTS) + list(EXTENSIONS) + ["set_fixed_capacity", "set_pair_chunk", "set_morton_mode", "last_counts", "set_profiling",
                                                    "kernel_times"]
