"""Host-side checks that run without a GPU: the C ABI library loads and exports every
declared symbol, the extension exposes the reference's entry points with the reference's
argument names and defaults, inputs are validated, and the product never imports the
oracle.  CPU only."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cccollisions_amd.h")
LIB = os.path.join(ROOT, "cccollisions_amd", "libccc_amd.so")

# src/cccollisions.cpp:426-468: name -> argument names with defaults
REFERENCE_API = {
    "bvh": ["triangles"],
    "bvh_motion": ["triangles", "triangles_next"],
    "find_collisions": ["bboxes", "tree", "triangles", "threshold=0.0", "max_collisions=16"],
    "find_triangle_edge_collisions": ["bboxes", "tree", "triangles", "threshold=0.0", "max_collisions=16"],
    "find_triangle_edge_candidates": ["bboxes", "tree", "triangles", "threshold=0.0", "max_collisions=16"],
    "find_triangle_edge_candidates2": ["bboxes", "tree", "triangles", "threshold=0.0", "max_collisions=16"],
    "find_collisions_continuous": ["bboxes", "tree", "triangles", "triangles_next",
                                   "max_candidates_per_triangle=16", "max_collisions_per_triangle=16"],
    "find_collisions_continuous_partial": ["bboxes", "tree", "triangles", "triangles_next", "triangles_to_check",
                                           "max_candidates_per_triangle=16", "max_collisions_per_triangle=16"],
    "find_proximity": ["bboxes", "tree", "triangles", "threshold", "max_candidates_per_triangle=16"],
    "find_mincol_cand": ["bboxes", "tree", "triangles", "triangles_next", "max_collisions=16"],
    "compute_impulses": ["bboxes", "tree", "triangles", "triangles_next", "triangles_mass",
                         "max_candidates_per_triangle=16", "max_collisions_per_triangle=16"],
    "compute_impulses_partial": ["bboxes", "tree", "triangles", "triangles_next", "triangles_mass",
                                 "triangles_to_check", "max_candidates_per_triangle=16",
                                 "max_collisions_per_triangle=16"],
    "compute_impulses_from_collisions": ["collisions", "roots", "triangles", "triangles_next", "triangles_mass"],
}


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ccc_[a-z0-9_]+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 17
    lib = ctypes.CDLL(LIB)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    lib.ccc_abi_version.restype = ctypes.c_int
    assert lib.ccc_abi_version() == 1
    lib.ccc_status_string.restype = ctypes.c_char_p
    assert lib.ccc_status_string(4) == b"malformed BVH tree"
    assert lib.ccc_status_string(0x200) == b"ok (output truncated)"


def test_c_abi_rejects_bad_arguments_without_gpu_work():
    lib = ctypes.CDLL(LIB)
    lib.ccc_bvh.restype = ctypes.c_int32
    lib.ccc_last_error.restype = ctypes.c_char_p
    # NULL context is refused before any device call
    st = lib.ccc_bvh(None, None, ctypes.c_int64(1), ctypes.c_int64(10), 1, None, None, None)
    assert st == 1 and b"context" in lib.ccc_last_error()


def _signature(fn):
    first = fn.__doc__.strip().splitlines()[0]
    inside = first[first.index("(") + 1: first.rindex(")")]
    args = []
    for part in re.split(r",\s*(?![^\[]*\])", inside):
        name = part.split(":")[0].strip()
        default = part.split("=")[1].strip() if "=" in part else None
        args.append(name if default is None else f"{name}={default}")
    return args


def test_extension_matches_reference_api():
    import cccollisions_amd as ccc
    for name, args in REFERENCE_API.items():
        got = _signature(getattr(ccc, name))
        assert got == args, (name, got)


def test_inputs_must_be_device_tensors():
    import torch
    import cccollisions_amd as ccc
    t = torch.zeros((1, 4, 3, 3), dtype=torch.float64)
    with pytest.raises(RuntimeError, match="triangles must be a CUDA tensor"):
        ccc.bvh(t)
    bb = torch.zeros((1, 7, 6), dtype=torch.float64)
    tree = torch.zeros((1, 7, 4), dtype=torch.int64)
    with pytest.raises(RuntimeError, match="bboxes must be a CUDA tensor"):
        ccc.compute_impulses(bb, tree, t, t, torch.zeros((1, 4, 3), dtype=torch.float64))


def test_triangle_edge_queries_validate_inputs():
    import torch
    import cccollisions_amd as ccc
    t = torch.zeros((1, 4, 3, 3), dtype=torch.float64)
    bb = torch.zeros((1, 7, 6), dtype=torch.float64)
    tree = torch.zeros((1, 7, 4), dtype=torch.int64)
    for name in ("find_triangle_edge_collisions", "find_triangle_edge_candidates", "find_triangle_edge_candidates2"):
        with pytest.raises(RuntimeError, match="bboxes must be a CUDA tensor"):
            getattr(ccc, name)(bb, tree, t)


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "cccollisions_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle|ccc_oracle|libccc_oracle", text, re.M), f


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_scene_sizes_and_caller_update():
    from cccollisions_amd import scenes
    assert scenes.folded_cloth(256).num_triangles == 130050
    v, f = scenes.proximity_cloth(64)
    assert f.shape[0] == 7938 and v.shape[0] == 4096
    faces = np.array([[0, 1, 2], [1, 2, 3]])
    x1 = np.zeros((4, 3))
    dx = np.zeros((2, 3, 3))
    cnt = np.zeros((2, 3), np.int32)
    dx[0, 1] = [1, 0, 0]; cnt[0, 1] = 1          # vertex 1 from triangle 0
    dx[1, 0] = [3, 0, 0]; cnt[1, 0] = 1          # vertex 1 from triangle 1
    dx[1, 2] = [np.nan, 0, 0]; cnt[1, 2] = 1     # vertex 3: non-finite, ignored
    out = scenes.apply_corrections(faces, x1, dx, cnt)
    np.testing.assert_array_equal(out[1], [2, 0, 0])
    np.testing.assert_array_equal(out[3], [0, 0, 0])


def test_cpu_baseline_pass_equals_composed_oracle(oracle):
    from tests import cases
    t0, t1, m = cases.folded(12)
    dv, dx, cnt, counts = oracle.impulse_pass(t0, t1, m, nthreads=4)
    rdv, rdx, rcnt, coll, roots, rc = oracle.compute_impulses(*oracle.bvh(t0, t1)[:2], t0, t1, m)
    assert counts == rc
    np.testing.assert_array_equal(cnt, rcnt)
    fin = np.isfinite(rdv)
    np.testing.assert_allclose(dv[fin], rdv[fin], rtol=1e-12, atol=1e-14)
