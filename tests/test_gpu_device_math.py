"""The narrow phase's device arithmetic against the oracle's, bit for bit, called through the
C ABI test hook ccc_debug_device_math: the portable cbrt / acos / sin / cos / cbrtf on wide
ranges and on special values (zeros, infinities, NaN, |x| > 1 for acos, subnormals), and
the cubic solver (cubic_prepare + cubic_roots) on random cubics with three real roots and
with one, on each branch of the reference's solveCubic (collisions_continuous.cu:511-613):
linear, quadratic (|c3| < 1e-16, including a double root), triple root, trigonometric and
Cardano forms, and coefficients with infinities and NaN.  The oracle's side is pinned
against mpmath and hand cases in tests/test_oracle_ccd.py.  The kernels' replacements of
IEEE operations -- division by 3 and 27, sqrt without rescaling, the edge-edge filter's
division-free parameter test -- are compared with numpy's operations.
"""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cccollisions_amd", "libccc_amd.so")
FUNCS = ["cbrt", "acos", "sin", "cos", "cbrtf"]

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dmath(ccc):
    import torch
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.ccc_context_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, P, P, P]
    lib.ccc_context_create.restype = ctypes.c_int32
    lib.ccc_context_destroy.argtypes = [P]
    lib.ccc_last_error.restype = ctypes.c_char_p
    lib.ccc_debug_device_math.argtypes = [P, ctypes.c_int, P, P, ctypes.c_int64, P]
    lib.ccc_debug_device_math.restype = ctypes.c_int32
    ctx = P()
    assert lib.ccc_context_create(ctypes.byref(ctx), torch.cuda.current_device(), None, None, None) == 0

    def run(what, x, per=1):
        x = torch.from_numpy(np.ascontiguousarray(x, np.float64).ravel()).cuda()
        y = torch.full_like(x, 12345.0)
        n = x.numel() // per
        torch.cuda.synchronize()
        st = lib.ccc_debug_device_math(ctx, what, x.data_ptr(), y.data_ptr(), n, None)
        assert st == 0, lib.ccc_last_error()
        torch.cuda.synchronize()
        return y.cpu().numpy()

    yield lib, ctx, run
    lib.ccc_context_destroy(ctx)


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def _inputs(name):
    rng = np.random.default_rng(11)
    special = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 0.5, -0.5, 5e-324, -5e-324, 2.2250738585072014e-308,
               1.7976931348623157e308, -1.7976931348623157e308]
    if name in ("cbrt", "cbrtf"):
        x = np.concatenate([10.0 ** rng.uniform(-300, 300, 20000) * rng.choice([-1, 1], 20000), rng.normal(size=20000)])
    elif name == "acos":
        x = np.concatenate([rng.uniform(-1, 1, 20000), 1 - 10.0 ** rng.uniform(-16, -1, 2000),
                            -1 + 10.0 ** rng.uniform(-16, -1, 2000), [1.0000000000000002, -1.0000000000000002, 2.0, -3.0]])
    else:
        x = np.concatenate([rng.uniform(0, np.pi / 3, 20000), rng.uniform(-4, 4, 2000)])
    return np.concatenate([x, special])


@pytest.mark.parametrize("what,name", list(enumerate(FUNCS)))
def test_portable_math_bit_exact(dmath, oracle, what, name):
    _, _, run = dmath
    x = _inputs(name)
    got = run(what, x)
    want = oracle.libm(name, x)
    bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
    assert bad.size == 0, (name, x[bad[:5]], got[bad[:5]], want[bad[:5]])


def _oracle_kept(oracle, coef):
    """The oracle's roots kept by the narrow phase: its slots in [0, 1], in order."""
    sl = oracle.solve_cubic(coef)
    return [float(s) for s in sl if 0 <= s <= 1]


def test_cubic_roots_bit_exact(dmath, oracle):
    _, _, run = dmath
    rng = np.random.default_rng(12)
    cases = []
    for _ in range(4000):
        r = rng.uniform(-0.5, 1.5, 3)
        cases.append(np.poly(r) * rng.uniform(0.1, 10) if rng.random() < 0.5 else rng.normal(size=4))
    # coefficients at the scale of near-coplanar cloth items (tiny, noisy)
    for _ in range(2000):
        cases.append(rng.normal(size=4) * 10.0 ** rng.uniform(-20, -8))
    cases += [
        [0.0, 0.0, 2.0, -1.0],              # linear: computed, never stored
        [1e-17, 1.0, -1.0, 0.1875],         # quadratic branch: (t - 1/4)(t - 3/4)
        [0.0, 1.0, -0.5, 0.0625],           # double root 1/4, kept twice
        [0.0, 1.0, -0.0, -0.25],            # quadratic: 1/2 kept, -1/2 dropped
        [2e-16, 1.0, -1.0, 0.1875],         # just above the quadratic threshold
        [1.0, -1.5, 0.75, -0.125],          # triple root 1/2 through the float path
        [1.0, -3.0, 3.0, -1.0],             # triple root 1
        np.poly([-0.5, 1.5, 2.0]),          # all roots outside [0, 1]
        np.poly([0.0, 0.5, 1.0]),           # roots on both window edges
        [0.0, 0.0, 0.0, 0.0],
        [np.nan, 1.0, 1.0, 1.0], [1.0, np.inf, 1.0, 1.0], [1.0, 1.0, 1.0, -np.inf],
    ]
    coef = np.array(cases, np.float64)
    got = run(5, coef, per=4).reshape(-1, 4)
    for i, c in enumerate(coef):
        want = _oracle_kept(oracle, c)
        k = int(got[i, 0])
        assert k == len(want), (c, got[i], want)
        assert _same_bits(got[i, 1:1 + k], want), (c, got[i], want)


def _division_inputs(rng, n):
    """Doubles of every exponent and sign (random bit patterns), values near 1, subnormals,
    mantissas of all ones at every exponent, and the special values."""
    u = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    anyb = u.view(np.float64)
    near1 = rng.uniform(-4, 4, n)
    sub = (u & np.uint64(0x800FFFFFFFFFFFFF)).view(np.float64)
    e = np.arange(2047, dtype=np.uint64)
    tops = ((e << np.uint64(52)) | np.uint64(0xFFFFFFFFFFFFF)).view(np.float64)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 1.7976931348623157e308, 3.0, 27.0, 1.0])
    return np.concatenate([anyb, near1, sub, tops, -tops, special])


@pytest.mark.parametrize("what,div", [(6, 3.0), (7, 27.0)])
def test_exact_constant_division(dmath, what, div):
    """The solver's division-free x / 3 and x / 27 give the division's bits on every input
    class, NaN payloads included."""
    _, _, run = dmath
    x = _division_inputs(np.random.default_rng(13 + what), 200000)
    got = run(what, x)
    want = x / div
    bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
    assert bad.size == 0, (x[bad[:5]], got[bad[:5]], want[bad[:5]])


def test_exact_division_by_three_on_angles(dmath):
    """The guard-free variant, as the trigonometric form uses it (angles in [0, pi], NaN)."""
    _, _, run = dmath
    rng = np.random.default_rng(17)
    x = np.concatenate([rng.uniform(0, np.pi, 200000), 10.0 ** rng.uniform(-300, 0, 20000), [0.0, np.pi, np.nan]])
    got = run(8, x)
    want = x / 3.0
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_sqrt_without_rescaling(dmath):
    """sqrt on the arguments pm_acos gives it (0, 2^-54 and up, negative, NaN): sqrt's bits."""
    _, _, run = dmath
    rng = np.random.default_rng(19)
    x = np.concatenate([10.0 ** rng.uniform(-230, 308, 100000), rng.uniform(0, 0.5, 100000),
                        2.0 ** -54 * (1 + rng.uniform(0, 1, 1000)), [0.0, -0.0, np.inf, 2.0 ** -767, 0.25, 0.5]])
    got = run(9, x)
    assert np.array_equal(got.view(np.uint64), np.sqrt(x).view(np.uint64))
    bad = run(9, np.array([-1.0, -np.inf, np.nan, -1e-300]))
    assert np.isnan(bad).all()


def _ratio_pairs(rng, dtype):
    """(n, d) pairs around every decision of the edge-edge parameter test: quotients within a
    few ulps of 1 (the 1 + u tie included) and of 0 (tiny negative n that round to -0), random
    pairs of every exponent, and the special operands.  d is +0, NaN, +inf or at least the
    smallest value the test decides without dividing."""
    fin = np.finfo(dtype)
    dlo = dtype(2.0 ** -960) if dtype == np.float64 else dtype(2.0 ** -100)
    m = 20000
    emax = 300 if dtype == np.float64 else 30
    d = (10.0 ** rng.uniform(-emax, emax, m)).astype(dtype)
    d = np.maximum(d, dlo)
    n = []
    dd = []
    u = dtype(fin.eps / 2)
    for k in (-3, -2, -1, 0, 1, 2, 3):
        n.append((d * (dtype(1) + dtype(k) * u)).astype(dtype))
        dd.append(d)
    one_tie = (d.astype(np.float64) * (1 + float(u))).astype(dtype)        # RN(d (1 + u))
    n += [one_tie, np.nextafter(one_tie, dtype(np.inf)), np.nextafter(one_tie, dtype(0))]
    dd += [d, d, d]
    # tiny negative numerators over large denominators: n / d rounds to -0 or to -tiny
    tiny = fin.smallest_subnormal
    big_d = (2.0 ** rng.integers(-2, 60, m)).astype(dtype)
    for j in (1, 2, 3, 4, 5):
        n.append(np.full(m, -tiny * j, dtype))
        dd.append(big_d)
    n.append((rng.uniform(-2, 2, m) * d).astype(dtype))
    dd.append(d)
    n.append((10.0 ** rng.uniform(-emax, emax, m) * rng.choice([-1, 1], m)).astype(dtype))
    dd.append(d)
    sp_n = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-30, -1e-30], dtype)
    sp_d = np.array([0.0, np.inf, np.nan, 1.0, dlo, fin.max], dtype)
    gn, gd = np.meshgrid(sp_n, sp_d)
    n.append(gn.ravel())
    dd.append(gd.ravel())
    return np.concatenate(n), np.concatenate(dd)


@pytest.mark.parametrize("what,dtype", [(10, np.float64), (11, np.float32)])
def test_edge_edge_parameter_test_without_division(dmath, what, dtype):
    """unit_ratio(n, d) equals !(s < 0 || s > 1) for s = n / d rounded in the dtype."""
    _, _, run = dmath
    n, d = _ratio_pairs(np.random.default_rng(what), dtype)
    with np.errstate(all="ignore"):
        s = n / d
    want = ~((s < 0) | (s > 1))
    # x = [n, d]: the hook reads x[i] and x[n + i] for i < n (per=2 passes n = x.size / 2)
    got = run(what, np.concatenate([n.astype(np.float64), d.astype(np.float64)]), per=2)[: n.size] == 1.0
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (n[bad[:5]], d[bad[:5]], s[bad[:5]], got[bad[:5]])


def test_hook_rejects_bad_arguments(dmath):
    lib, ctx, _ = dmath
    assert lib.ccc_debug_device_math(ctx, 12, None, None, 0, None) != 0
    assert lib.ccc_debug_device_math(ctx, 0, None, None, 4, None) != 0
    assert lib.ccc_debug_device_math(ctx, 0, None, None, 0, None) == 0
