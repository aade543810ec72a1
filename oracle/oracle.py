"""ctypes wrapper of the CPU oracle (oracle/_build/libccc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, to check and time the GPU path against.  The product
package (``cccollisions_amd``) never imports it.

Every function takes/returns numpy arrays in the reference's layouts:
triangle soups ``[N, 3, 3]``, BVH ``bboxes [2N-1, 6]`` / ``tree [2N-1, 4]`` int64,
collisions ``[R, 3]`` int64 ``(x, y, type)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libccc_oracle.so")
_lib = None


class Counts(C.Structure):
    _fields_ = [("candidates", C.c_int64), ("subprims", C.c_int64),
                ("roots", C.c_int64), ("collisions", C.c_int64)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_I64 = C.c_int64


def _declare(L):
    for sfx, real in (("_f64", C.c_double), ("_f32", C.c_float)):
        getattr(L, "orc_bvh" + sfx).argtypes = [_P, _P, _I64, C.c_int, _P, _P, _P, _P]
        getattr(L, "orc_bvh" + sfx).restype = C.c_int
        getattr(L, "orc_traverse" + sfx).argtypes = [_P, _P, _I64, real, _P, _I64]
        getattr(L, "orc_traverse" + sfx).restype = _I64
        getattr(L, "orc_traverse_partial" + sfx).argtypes = [_P, _P, _I64, _P, _P, _I64]
        getattr(L, "orc_traverse_partial" + sfx).restype = _I64
        getattr(L, "orc_bruteforce_pairs" + sfx).argtypes = [_P, _P, _I64, real, _P, _I64]
        getattr(L, "orc_bruteforce_pairs" + sfx).restype = _I64
        getattr(L, "orc_ccd" + sfx).argtypes = [_P, _P, _I64, _P, _I64, _P, _P, _I64, C.POINTER(Counts)]
        getattr(L, "orc_ccd" + sfx).restype = C.c_int
        getattr(L, "orc_impulses" + sfx).argtypes = [_P, _P, _P, _I64, _P, _P, _I64, _P, _P, _P]
        getattr(L, "orc_impulses" + sfx).restype = C.c_int
        getattr(L, "orc_proximity" + sfx).argtypes = [_P, _I64, _P, _I64, real, _P, _I64, C.POINTER(Counts)]
        getattr(L, "orc_proximity" + sfx).restype = C.c_int
        getattr(L, "orc_static_collisions" + sfx).argtypes = [_P, _I64, _P, _I64, real, _P, _I64]
        getattr(L, "orc_static_collisions" + sfx).restype = _I64
        getattr(L, "orc_triangle_edge" + sfx).argtypes = [_P, _I64, _P, _I64, C.c_int, _P, _P, _P, _I64]
        getattr(L, "orc_triangle_edge" + sfx).restype = _I64
        getattr(L, "orc_mincol" + sfx).argtypes = [_P, _P, _I64, _P, _I64, _P, _I64]
        getattr(L, "orc_mincol" + sfx).restype = _I64
        getattr(L, "orc_cubic_coeffs" + sfx).argtypes = [_P, _P]
        getattr(L, "orc_solve_cubic" + sfx).argtypes = [_P, _P]
    L.orc_dd_cube_lead.argtypes = [C.c_double]
    L.orc_dd_cube_lead.restype = C.c_double
    for f in ("orc_karras_node", "orc_karras_node_literal"):
        getattr(L, f).argtypes = [_P, _P, _I64, _I64, C.POINTER(_I64), C.POINTER(_I64)]
        getattr(L, f).restype = C.c_int
    L.orc_morton3d_f64.argtypes = [C.c_double, C.c_double, C.c_double]
    L.orc_morton3d_f64.restype = C.c_uint32
    L.orc_morton63_f64.argtypes = [C.c_double, C.c_double, C.c_double]
    L.orc_morton63_f64.restype = C.c_uint64
    L.orc_libm_eval.argtypes = [C.c_int, _P, _P, _I64]
    L.orc_impulse_pass_f64.argtypes = [_P, _P, _P, _I64, C.c_int, _I64, _P, _P, _P, C.POINTER(Counts)]
    L.orc_impulse_pass_f64.restype = C.c_int


def _sfx(a):
    if a.dtype == np.float64:
        return "_f64", C.c_double
    if a.dtype == np.float32:
        return "_f32", C.c_float
    raise TypeError(f"unsupported dtype {a.dtype}")


def _c(a, dtype=None):
    a = np.ascontiguousarray(a if dtype is None else a.astype(dtype, copy=False))
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data


def bvh(tris, tris_next=None, morton_mode=0):
    """Reference `bvh` / `bvh_motion` for one mesh: returns (bboxes, tree, codes, ids).
    morton_mode 0 is the reference (infinite scene box, all codes 0); 1 and 2 fit the scene
    box to the leaf boxes and use 30-bit / 63-bit codes (uint64 either way)."""
    tris = _c(tris)
    n = tris.shape[0]
    sfx, _ = _sfx(tris)
    if tris_next is not None:
        tris_next = _c(tris_next, tris.dtype)
    bb = np.empty((2 * n - 1, 6), tris.dtype)
    tree = np.empty((2 * n - 1, 4), np.int64)
    codes = np.empty(n, np.uint64)
    ids = np.empty(n, np.int32)
    rc = getattr(lib(), "orc_bvh" + sfx)(_ptr(tris), _ptr(tris_next), n, morton_mode,
                                         _ptr(bb), _ptr(tree), _ptr(codes), _ptr(ids))
    if rc != 0:
        raise RuntimeError(f"orc_bvh failed: {rc}")
    return bb, tree, codes, ids


def traverse(bboxes, tree, threshold=0.0, cap=None):
    """Reference traversal (ordering on): candidate pairs [K, 2] (min id, max id)."""
    bboxes = _c(bboxes)
    tree = _c(tree, np.int64)
    n = (bboxes.shape[0] + 1) // 2
    sfx, real = _sfx(bboxes)
    f = getattr(lib(), "orc_traverse" + sfx)
    if cap is None:
        total = f(_ptr(bboxes), _ptr(tree), n, real(threshold), None, 0)
        if total < 0:
            raise RuntimeError("traversal stack overflow")
        cap = total
    out = np.empty((max(cap, 1), 2), np.int64)
    total = f(_ptr(bboxes), _ptr(tree), n, real(threshold), _ptr(out), cap)
    return out[:min(total, cap)], total


def traverse_partial(bboxes, tree, to_check):
    bboxes = _c(bboxes)
    tree = _c(tree, np.int64)
    tc = _c(np.asarray(to_check).astype(np.uint8))
    n = (bboxes.shape[0] + 1) // 2
    sfx, _ = _sfx(bboxes)
    f = getattr(lib(), "orc_traverse_partial" + sfx)
    total = f(_ptr(bboxes), _ptr(tree), n, _ptr(tc), None, 0)
    out = np.empty((max(total, 1), 2), np.int64)
    f(_ptr(bboxes), _ptr(tree), n, _ptr(tc), _ptr(out), total)
    return out[:total]


def bruteforce_pairs(bboxes, tree, threshold=0.0):
    bboxes = _c(bboxes)
    tree = _c(tree, np.int64)
    n = (bboxes.shape[0] + 1) // 2
    sfx, real = _sfx(bboxes)
    f = getattr(lib(), "orc_bruteforce_pairs" + sfx)
    total = f(_ptr(bboxes), _ptr(tree), n, real(threshold), None, 0)
    out = np.empty((max(total, 1), 2), np.int64)
    f(_ptr(bboxes), _ptr(tree), n, real(threshold), _ptr(out), total)
    return out[:total]


def ccd(tris, tris_next, pairs, cap=None):
    """Narrow phase of find_collisions_continuous: (collisions [R,3], roots [R], counts)."""
    tris = _c(tris)
    tris_next = _c(tris_next, tris.dtype)
    pairs = _c(np.asarray(pairs).reshape(-1, 2), np.int64)
    n = tris.shape[0]
    sfx, _ = _sfx(tris)
    f = getattr(lib(), "orc_ccd" + sfx)
    cnt = Counts()
    if cap is None:
        f(_ptr(tris), _ptr(tris_next), n, _ptr(pairs), pairs.shape[0], None, None, 0, C.byref(cnt))
        cap = cnt.collisions
    coll = np.empty((max(cap, 1), 3), np.int64)
    roots = np.empty(max(cap, 1), tris.dtype)
    f(_ptr(tris), _ptr(tris_next), n, _ptr(pairs), pairs.shape[0], _ptr(coll), _ptr(roots), cap, C.byref(cnt))
    r = min(cnt.collisions, cap)
    return coll[:r], roots[:r], cnt.as_dict()


def impulses(tris, tris_next, masses, collisions, roots):
    """compute_impulses_from_collisions for one mesh: (dv [N,3,3], dx [N,3,3], cnt [N,3])."""
    tris = _c(tris)
    tris_next = _c(tris_next, tris.dtype)
    masses = _c(masses, tris.dtype)
    coll = _c(np.asarray(collisions).reshape(-1, 3), np.int64)
    roots = _c(np.asarray(roots).reshape(-1), tris.dtype)
    n = tris.shape[0]
    sfx, _ = _sfx(tris)
    dv = np.zeros((n, 3, 3), tris.dtype)
    dx = np.zeros((n, 3, 3), tris.dtype)
    cnt = np.zeros((n, 3), np.int32)
    getattr(lib(), "orc_impulses" + sfx)(_ptr(tris), _ptr(tris_next), _ptr(masses), n, _ptr(coll),
                                         _ptr(roots), coll.shape[0], _ptr(dv), _ptr(dx), _ptr(cnt))
    return dv, dx, cnt


def proximity(tris, pairs, threshold):
    """Narrow phase of find_proximity: (rows [R,3], counts).  `threshold` is used as given."""
    tris = _c(tris)
    pairs = _c(np.asarray(pairs).reshape(-1, 2), np.int64)
    n = tris.shape[0]
    sfx, real = _sfx(tris)
    f = getattr(lib(), "orc_proximity" + sfx)
    cnt = Counts()
    f(_ptr(tris), n, _ptr(pairs), pairs.shape[0], real(threshold), None, 0, C.byref(cnt))
    cap = cnt.collisions
    out = np.empty((max(cap, 1), 3), np.int64)
    f(_ptr(tris), n, _ptr(pairs), pairs.shape[0], real(threshold), _ptr(out), cap, C.byref(cnt))
    return out[:cap], cnt.as_dict()


def static_collisions(tris, pairs, threshold):
    tris = _c(tris)
    pairs = _c(np.asarray(pairs).reshape(-1, 2), np.int64)
    n = tris.shape[0]
    sfx, real = _sfx(tris)
    f = getattr(lib(), "orc_static_collisions" + sfx)
    total = f(_ptr(tris), n, _ptr(pairs), pairs.shape[0], real(threshold), None, 0)
    out = np.empty((max(total, 1), 2), np.int64)
    f(_ptr(tris), n, _ptr(pairs), pairs.shape[0], real(threshold), _ptr(out), total)
    return out[:total]


TE_MODES = {"collisions": 0, "candidates": 1, "candidates2": 2}


def triangle_edge(tris, pairs, mode):
    """Triangle-edge rows over candidate pairs (collisions_static.cu:345-943, icontour.cu).
    mode "collisions" -> (rows [R,4], coefs [R]); "candidates" -> rows [R,4];
    "candidates2" -> (pairs [R,2], masks [R,2] int32)."""
    tris = _c(tris)
    pairs = _c(np.asarray(pairs).reshape(-1, 2), np.int64)
    n = tris.shape[0]
    sfx, _ = _sfx(tris)
    m = TE_MODES[mode]
    f = getattr(lib(), "orc_triangle_edge" + sfx)
    total = f(_ptr(tris), n, _ptr(pairs), pairs.shape[0], m, None, None, None, 0)
    width = 2 if m == 2 else 4
    rows = np.empty((max(total, 1), width), np.int64)
    coefs = np.empty(max(total, 1), tris.dtype)
    masks = np.empty((max(total, 1), 2), np.int32)
    f(_ptr(tris), n, _ptr(pairs), pairs.shape[0], m, _ptr(rows), _ptr(coefs) if m == 0 else None,
      _ptr(masks) if m == 2 else None, total)
    if m == 0:
        return rows[:total], coefs[:total]
    if m == 1:
        return rows[:total]
    return rows[:total], masks[:total]


def mincol(tris, tris_next, pairs):
    tris = _c(tris)
    tris_next = _c(tris_next, tris.dtype)
    pairs = _c(np.asarray(pairs).reshape(-1, 2), np.int64)
    n = tris.shape[0]
    sfx, _ = _sfx(tris)
    f = getattr(lib(), "orc_mincol" + sfx)
    total = f(_ptr(tris), _ptr(tris_next), n, _ptr(pairs), pairs.shape[0], None, 0)
    out = np.empty((max(total, 1), 3), np.int64)
    f(_ptr(tris), _ptr(tris_next), n, _ptr(pairs), pairs.shape[0], _ptr(out), total)
    return out[:total]


def cubic_coeffs(tet):
    """tet: [8, 3] = 4 positions then 4 deltas.  Returns (c3, c2, c1, c0)."""
    tet = _c(np.asarray(tet).reshape(8, 3))
    sfx, _ = _sfx(tet)
    out = np.empty(4, tet.dtype)
    getattr(lib(), "orc_cubic_coeffs" + sfx)(_ptr(tet), _ptr(out))
    return out


def solve_cubic(coef):
    coef = _c(np.asarray(coef).reshape(4))
    sfx, _ = _sfx(coef)
    out = np.empty(3, coef.dtype)
    getattr(lib(), "orc_solve_cubic" + sfx)(_ptr(coef), _ptr(out))
    return out


def dd_cube_lead(b):
    return lib().orc_dd_cube_lead(float(b))


def karras_node(codes, ids, i, literal=False):
    codes = _c(codes, np.uint64)
    ids = _c(ids, np.int32)
    l, r = C.c_int64(), C.c_int64()
    f = lib().orc_karras_node_literal if literal else lib().orc_karras_node
    f(_ptr(codes), _ptr(ids), codes.shape[0], i, C.byref(l), C.byref(r))
    return l.value, r.value


def morton3d(x, y, z):
    return lib().orc_morton3d_f64(x, y, z)


def morton63(x, y, z):
    return lib().orc_morton63_f64(x, y, z)


LIBM_FUNCS = {"cbrt": 0, "acos": 1, "sin": 2, "cos": 3, "cbrtf": 4}


def libm(name, x):
    """Portable cbrt/acos/sin/cos used by the oracle's cubic solver (ccc_libm_ref.h)."""
    x = np.ascontiguousarray(np.asarray(x, np.float64).ravel())
    y = np.empty_like(x)
    lib().orc_libm_eval(LIBM_FUNCS[name], _ptr(x), _ptr(y), x.size)
    return y


def impulse_pass(tris, tris_next, masses, nthreads=0, query_stride=1):
    """CPU baseline: bvh_motion + compute_impulses for one mesh (OpenMP over query leaves;
    query_stride > 1 runs only every query_stride-th query leaf)."""
    tris = _c(tris, np.float64)
    tris_next = _c(tris_next, np.float64)
    masses = _c(masses, np.float64)
    n = tris.shape[0]
    dv = np.zeros((n, 3, 3))
    dx = np.zeros((n, 3, 3))
    cnt = np.zeros((n, 3), np.int32)
    c = Counts()
    rc = lib().orc_impulse_pass_f64(_ptr(tris), _ptr(tris_next), _ptr(masses), n, nthreads,
                                    query_stride, _ptr(dv), _ptr(dx), _ptr(cnt), C.byref(c))
    if rc != 0:
        raise RuntimeError(f"orc_impulse_pass failed: {rc}")
    return dv, dx, cnt, c.as_dict()


# ---- composite reference pipelines (one mesh) -------------------------------------------

def find_collisions_continuous(bboxes, tree, tris, tris_next):
    pairs, _ = traverse(bboxes, tree, 0.0)
    return ccd(tris, tris_next, pairs)


def compute_impulses(bboxes, tree, tris, tris_next, masses):
    coll, roots, counts = find_collisions_continuous(bboxes, tree, tris, tris_next)
    dv, dx, cnt = impulses(tris, tris_next, masses, coll, roots)
    return dv, dx, cnt, coll, roots, counts


def find_proximity(bboxes, tree, tris, threshold):
    thr = float(np.float32(threshold))        # the binding's threshold is a C++ float
    pairs, _ = traverse(bboxes, tree, thr)
    return proximity(tris, pairs, thr)


def find_collisions(bboxes, tree, tris, threshold=0.0):
    thr = float(np.float32(threshold))
    pairs, _ = traverse(bboxes, tree, thr)
    return static_collisions(tris, pairs, thr)


def find_triangle_edge(bboxes, tree, tris, mode):
    """The three triangle-edge queries end to end: traversal and SAT at threshold 0 (the
    bindings' threshold argument is not used, collisions_static.cu:681, :733)."""
    pairs, _ = traverse(bboxes, tree, 0.0)
    return triangle_edge(tris, pairs, mode)
