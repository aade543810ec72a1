"""ccc_compute_impulses_noalloc (workspace carved from the library's static scratch array,
no device allocation) against ccc_compute_impulses, both called through the C ABI.

Two calls in a row on the same GPU, the second on a smaller cloth: the scratch array keeps
the first call's contents, so the second call must clear what it relies on.  cnt and the
stage counts must be equal, dv / dx within 1e-9 relative (fp64 atomic summation order).
"""
import ctypes
import os

import numpy as np
import pytest

from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cccollisions_amd", "libccc_amd.so")
SCRATCH_BYTES = 256 << 20
CCC_ERR_NO_MEMORY = 3

pytestmark = pytest.mark.gpu


class Counts(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int64), ("subprims", ctypes.c_int64), ("roots", ctypes.c_int64),
                ("collisions", ctypes.c_int64)]


def _bind(lib, name):
    fn = getattr(lib, name)
    P, I64, I32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    fn.argtypes = [P, P, P, P, P, P, I64, I64, I32, I32, I32, P, P, P, ctypes.POINTER(Counts), ctypes.c_uint32, P]
    fn.restype = ctypes.c_int32
    return fn


@pytest.fixture(scope="module")
def capi(ccc):
    import torch
    lib = ctypes.CDLL(LIB)
    lib.ccc_context_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]
    lib.ccc_context_create.restype = ctypes.c_int32
    lib.ccc_context_destroy.argtypes = [ctypes.c_void_p]
    lib.ccc_last_error.restype = ctypes.c_char_p
    ctx = ctypes.c_void_p()
    st = lib.ccc_context_create(ctypes.byref(ctx), torch.cuda.current_device(), None, None, None)
    assert st == 0, lib.ccc_last_error()
    yield lib, ctx, _bind(lib, "ccc_compute_impulses"), _bind(lib, "ccc_compute_impulses_noalloc")
    lib.ccc_context_destroy(ctx)


def _call(capi, fn, bb, tree, T0, T1, M, maxc):
    import torch
    lib, ctx, _, _ = capi
    B, n = T0.shape[0], T0.shape[1]
    dv = torch.full_like(T0, 7.0)      # outputs start dirty: the call must clear them
    dx = torch.full_like(T0, -3.0)
    cnt = torch.full((B, n, 3), 5, dtype=torch.int32, device=T0.device)
    counts = (Counts * B)()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = fn(ctx, bb.data_ptr(), tree.data_ptr(), T0.data_ptr(), T1.data_ptr(), M.data_ptr(), B, n,
            1 if T0.dtype == torch.float64 else 0, maxc, 16, dv.data_ptr(), dx.data_ptr(), cnt.data_ptr(), counts, 0, stream)
    torch.cuda.synchronize()
    return st, dv, dx, cnt, [(c.candidates, c.subprims, c.roots, c.collisions) for c in counts], lib.ccc_last_error()


def _inputs(ccc, scenes_list):
    import torch
    dev = torch.device("cuda:0")
    T0 = torch.from_numpy(np.stack([s[0] for s in scenes_list])).to(dev)
    T1 = torch.from_numpy(np.stack([s[1] for s in scenes_list])).to(dev)
    M = torch.from_numpy(np.stack([s[2] for s in scenes_list])).to(dev)
    bb, tree = ccc.bvh_motion(T0, T1)
    return bb, tree, T0, T1, M


def _compare(ref, got):
    st_r, dv_r, dx_r, cnt_r, counts_r, _ = ref
    st_g, dv_g, dx_g, cnt_g, counts_g, err = got
    assert st_g == st_r == 0, err
    assert counts_g == counts_r
    assert counts_g[0][3] > 0, "the scene must produce collisions"
    assert np.array_equal(cnt_g.cpu().numpy(), cnt_r.cpu().numpy())
    for a, b in ((dv_g, dv_r), (dx_g, dx_r)):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        # collisions whose parameters are NaN are accepted (as in the reference) and make
        # their corners' sums NaN: the same corners on both sides, the rest within 1e-9
        fin = np.isfinite(b)
        np.testing.assert_array_equal(np.isfinite(a), fin)
        a, b = np.where(fin, a, 0.0), np.where(fin, b, 0.0)
        scale = np.abs(b).max()
        assert scale > 0
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * scale)


def test_noalloc_matches_compute_impulses_on_two_calls_in_a_row(ccc, capi):
    _, _, plain, noalloc = capi
    big = cases.folded(40, seed=4)
    small = cases.folded(14, seed=9)
    for sc, maxc in ((big, 256), (small, 128)):
        args = _inputs(ccc, [(sc[0], sc[1], sc[2])])
        print(f"noalloc: {args[2].shape[1]} triangles, max_candidates {maxc}", flush=True)
        ref = _call(capi, plain, *args, maxc)
        got = _call(capi, noalloc, *args, maxc)
        _compare(ref, got)


def test_noalloc_batch_and_candidate_regrowth(ccc, capi):
    """Two batch elements, and a candidate capacity (1 per triangle) that the first traversal
    exceeds: the re-run takes the released pair buffer's place in the scratch array."""
    _, _, plain, noalloc = capi
    a, b = cases.folded(18, seed=2), cases.folded(18, seed=3)
    args = _inputs(ccc, [(a[0], a[1], a[2]), (b[0], b[1], b[2])])
    print(f"noalloc batch: 2 x {args[2].shape[1]} triangles", flush=True)
    ref = _call(capi, plain, *args, 1)
    got = _call(capi, noalloc, *args, 1)
    assert ref[0] & 0x100 and got[0] & 0x100, "both calls report the exceeded candidate capacity"
    _compare((ref[0] & 0xff,) + ref[1:], (got[0] & 0xff,) + got[1:])


def test_noalloc_rejects_a_workspace_beyond_the_scratch_array(ccc, capi):
    _, _, _, noalloc = capi
    sc = cases.folded(14, seed=9)
    args = _inputs(ccc, [(sc[0], sc[1], sc[2])])
    n = args[2].shape[1]
    maxc = SCRATCH_BYTES // (8 * n) + 1          # the candidate pairs alone exceed the array
    st, *_, err = _call(capi, noalloc, *args, maxc)
    assert st & 0xff == CCC_ERR_NO_MEMORY
    assert b"scratch" in err
