"""Pinning the oracle's CCD and impulses for every sub-primitive type against the
constructed known answers of tests/golden/ccd_typed_cases.json (exact rational scenes; the
expected roots, collision labels and impulses follow from the construction, see
tests/golden/make_golden.py), plus the solver's edge cases: the float32 root window and
duplicate roots.  CPU only."""
import json
import os

import numpy as np
import pytest

from tests import cases

TYPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ccd_typed_cases.json")
# root and impulse tolerances per dtype (float32: coefficients and filter in float)
TOL = {np.float64: 1e-12, np.float32: 2e-5}


def typed_cases():
    with open(TYPED) as f:
        return json.load(f)["typed_cases"]


CASES = typed_cases()
IDS = [f"type{c['type']}-v{c['variant']}" for c in CASES]


def arrays(c, dtype):
    return (np.array(c["t0"], dtype), np.array(c["t1"], dtype), np.array(c["masses"], dtype))


def check_impulses(dv, dx, cnt, expect, tol, what=""):
    edv, edx = np.array(expect["dv"]), np.array(expect["dx"])
    scale = np.abs(edv).max()
    np.testing.assert_array_equal(cnt, np.array(expect["cnt"]), err_msg=f"cnt {what}")
    np.testing.assert_allclose(dv, edv, rtol=tol, atol=tol * scale, err_msg=f"dv {what}")
    np.testing.assert_allclose(dx, edx, rtol=tol, atol=tol * scale, err_msg=f"dx {what}")


def test_every_type_covered():
    types = sorted({c["type"] for c in CASES})
    assert types == list(range(15))
    assert all(sum(c["type"] == t for c in CASES) == 3 for t in types)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("idx", range(len(CASES)), ids=IDS)
def test_typed_collision(oracle, idx, dtype):
    """The oracle reports exactly the constructed (0, 1, type) collision at t*."""
    c = CASES[idx]
    t0, t1, _ = arrays(c, dtype)
    coll, roots, counts = oracle.ccd(t0, t1, np.array([[0, 1]]))
    assert coll.tolist() == c["expect"]["collisions"], (coll.tolist(), c["expect"])
    np.testing.assert_allclose(roots, c["expect"]["roots"], rtol=TOL[dtype])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("idx", range(len(CASES)), ids=IDS)
def test_typed_impulses(oracle, idx, dtype):
    """Impulses of the constructed collision: corners by the type's table, masses
    interpolated by the constructed contact weights (unequal per corner), normal +z / -z
    (inverted for types 12-14), nv = -V, dx scaled by 1 - max(t* - offset, 0)."""
    c = CASES[idx]
    t0, t1, m = arrays(c, dtype)
    coll, roots, _ = oracle.ccd(t0, t1, np.array([[0, 1]]))
    dv, dx, cnt = oracle.impulses(t0, t1, m, coll, roots)
    check_impulses(dv, dx, cnt, c["expect"], TOL[dtype], IDS[idx])


@pytest.mark.parametrize("idx", range(0, len(CASES), 3), ids=IDS[::3])
def test_typed_impulses_detect_wrong_table(oracle, idx):
    """The fixtures tell the types apart: the same collision scattered under any other
    type's corner table (a swapped row) or with the corner masses permuted differs from the
    known answer."""
    c = CASES[idx]
    t0, t1, m = arrays(c, np.float64)
    edv = np.array(c["expect"]["dv"])
    for other in range(15):
        if other == c["type"]:
            continue
        dv, _, cnt = oracle.impulses(t0, t1, m, np.array([[0, 1, other]]), np.array(c["expect"]["roots"]))
        assert not (np.allclose(dv, edv, rtol=1e-6, atol=1e-9) and (cnt == np.array(c["expect"]["cnt"])).all()), other
    for perm in ([1, 2, 0], [2, 0, 1]):
        dv, _, _ = oracle.impulses(t0, t1, m[:, perm], np.array(c["expect"]["collisions"]), np.array(c["expect"]["roots"]))
        assert not np.allclose(dv, edv, rtol=1e-6, atol=1e-9), perm


def test_fp32_root_rounded_to_one_is_kept(oracle):
    """collisions_continuous.cu:448-464: add_root<float> takes `float root`, so a double root
    in (1, 1 + 2^-24) becomes 1.0f before the window test and is kept; the flatten keeps
    r <= 1 (:743-763) and the filter runs at t = 1."""
    t0, t1, exact = cases.unit_root_scene_f32(oracle)
    assert 1.0 < exact < 1.0 + 2.0 ** -24
    d = (t1 - t0).astype(np.float32)
    tet = np.concatenate([[t0[0, 0], t0[1, 0], t0[1, 1], t0[1, 2]], [d[0, 0], d[1, 0], d[1, 1], d[1, 2]]])
    sl = oracle.solve_cubic(oracle.cubic_coeffs(tet.astype(np.float32)))
    assert np.float32(1.0) in sl.tolist(), sl
    coll, roots, _ = oracle.ccd(t0, t1, np.array([[0, 1]]))
    hit = [(tuple(cc), float(r)) for cc, r in zip(coll.tolist(), roots) if cc[2] == 12]
    assert hit == [((0, 1, 12), 1.0)], hit


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_double_root_emitted_twice(oracle, dtype):
    """An exactly quadratic vertex-face polynomial with the double root 1/4 (D = 0): both root
    slots hold 1/4, the flatten emits a row per slot (:743-763) and the NaN of the degenerate
    face passes the filter (:615-637): two (0, 1, 12) rows at 0.25."""
    t0, t1 = cases.double_root_scene(dtype)
    d = (t1 - t0).astype(dtype)
    tet = np.concatenate([[t0[0, 0], t0[1, 0], t0[1, 1], t0[1, 2]], [d[0, 0], d[1, 0], d[1, 1], d[1, 2]]])
    coef = oracle.cubic_coeffs(tet.astype(dtype))
    assert coef[0] == 0 and coef[2] * coef[2] == 4 * coef[1] * coef[3]
    assert sorted(oracle.solve_cubic(coef).tolist()) == [-1.0, 0.25, 0.25]
    coll, roots, counts = oracle.ccd(t0, t1, np.array([[0, 1]]))
    rows = [float(r) for cc, r in zip(coll.tolist(), roots) if cc[2] == 12]
    assert rows == [0.25, 0.25], rows
