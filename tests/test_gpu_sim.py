"""Resident simulations (cccollisions_amd.Sim, C ABI ccc_sim_*): steps of impulse passes
(bvh_motion + compute_impulses + the caller's vertex update) captured as HIP graphs.

* the vertex update equals scenes.apply_corrections bit for bit on the pass's own impulses,
  and the impulses equal the oracle's (cnt exact, dv / dx within 1e-9 relative);
* deterministic mode: the same bits from graph and eager launches, every narrow-phase grid,
  both Morton trees, a packed simulation and separate ones, over steps that advance the
  state;
* a candidate list that overflows grows and the step re-runs from the saved positions;
* caller-owned workspace, autotuning, and the fp32 path.
"""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def _scene(n, seed):
    from cccollisions_amd import scenes
    return scenes.folded_cloth(n, seed=seed)


def _sim(ccc, sc, dtype=np.float64, **kw):
    import torch
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    return ccc.Sim(t(sc.x0.astype(dtype)), t(sc.x1.astype(dtype)), t(sc.faces), t(sc.corner_masses().astype(dtype)), **kw)


def _np(x):
    return x.detach().cpu().numpy()


@pytest.mark.parametrize("det", [False, True])
@pytest.mark.parametrize("graph", [False, True])
def test_one_pass_matches_oracle_and_apply_corrections(ccc, oracle, det, graph):
    from cccollisions_amd import scenes
    sc = _scene(14, 7)
    sim = _sim(ccc, sc, iters=1, deterministic=det, graph=graph, max_candidates=512)
    sim.step()
    x0, x1 = (_np(a) for a in sim.positions())
    dv, dx, cnt = (_np(a) for a in sim.last_impulses())
    # the vertex update on the pass's own impulses: numpy's result, bit for bit
    np.testing.assert_array_equal(x1, scenes.apply_corrections(sc.faces, sc.x1, dx, cnt))
    np.testing.assert_array_equal(x0, sc.x0)
    # the impulses against the oracle's pass
    t0, t1, m = sc.x0[sc.faces], sc.x1[sc.faces], sc.corner_masses()
    bb, tree = oracle.bvh(t0, t1)[:2]
    odv, odx, ocnt, _, _, counts = oracle.compute_impulses(bb, tree, t0, t1, m)
    assert counts["collisions"] > 0
    assert sim.counts()[0] == [counts["candidates"], counts["subprims"], counts["roots"], counts["collisions"]]
    np.testing.assert_array_equal(cnt, ocnt)
    for a, b in ((dv, odv), (dx, odx)):
        fin = np.isfinite(b)
        np.testing.assert_array_equal(np.isfinite(a), fin)
        a, b = np.where(fin, a, 0.0), np.where(fin, b, 0.0)
        scale = np.abs(b).max()
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * scale)


def _run(ccc, sc, steps, **kw):
    sim = _sim(ccc, sc, **kw)
    for _ in range(steps):
        sim.step(advance=True)
    return [_np(a) for a in sim.positions()]


def test_deterministic_steps_give_the_same_bits_every_way(ccc):
    """Three advancing steps of three passes: graph or eager, narrow-phase grid 64 or 256
    blocks per CU, the reference's or a fitted Morton tree -- the same positions, bit for
    bit, and the same again on a second run."""
    sc = _scene(18, 3)
    ref = _run(ccc, sc, 3, iters=3, deterministic=True, graph=True, max_candidates=512)
    assert not np.array_equal(ref[1], sc.x1)
    for kw in (dict(graph=True), dict(graph=False), dict(ccd_blocks_per_cu=64), dict(ccd_blocks_per_cu=256),
               dict(morton=1), dict(morton=2)):
        got = _run(ccc, sc, 3, iters=3, deterministic=True, max_candidates=512, **kw)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b, err_msg=str(kw))


def test_packed_simulation_equals_separate_ones(ccc):
    """Three cloths of different sizes packed in one simulation against each alone:
    identical positions after two advancing steps (deterministic mode)."""
    import torch
    cl = [_scene(n, 20 + n) for n in (10, 16, 7)]
    alone = [_run(ccc, sc, 2, iters=2, deterministic=True, max_candidates=512) for sc in cl]
    voff = np.cumsum([0] + [sc.x0.shape[0] for sc in cl])
    foff = np.cumsum([0] + [sc.faces.shape[0] for sc in cl]).tolist()
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    sim = ccc.Sim(t(np.concatenate([sc.x0 for sc in cl])), t(np.concatenate([sc.x1 for sc in cl])),
                  t(np.concatenate([sc.faces + voff[i] for i, sc in enumerate(cl)])),
                  t(np.concatenate([sc.corner_masses() for sc in cl])), mesh_offsets=foff, iters=2,
                  deterministic=True, max_candidates=512)
    for _ in range(2):
        sim.step(advance=True)
    x0, x1 = (_np(a) for a in sim.positions())
    for i, (a0, a1) in enumerate(alone):
        np.testing.assert_array_equal(x0[voff[i]:voff[i + 1]], a0)
        np.testing.assert_array_equal(x1[voff[i]:voff[i + 1]], a1)


def test_candidate_overflow_regrows_and_reruns(ccc):
    """One candidate per triangle cannot hold the folded cloth's pairs: the step grows the
    list and runs again from the saved positions, giving the large list's result."""
    sc = _scene(16, 5)
    ref = _run(ccc, sc, 2, iters=2, deterministic=True, max_candidates=512)
    sim = _sim(ccc, sc, iters=2, deterministic=True, max_candidates=1)
    for _ in range(2):
        sim.step(advance=True)
    assert sim.config()["candidate_capacity"] > sc.faces.shape[0]
    for a, b in zip((_np(a) for a in sim.positions()), ref):
        np.testing.assert_array_equal(a, b)


def test_caller_workspace(ccc):
    import torch
    sc = _scene(16, 6)
    ref = _run(ccc, sc, 2, iters=2, deterministic=True, max_candidates=512)
    sim = _sim(ccc, sc, iters=2, deterministic=True, max_candidates=512)
    ws = torch.empty(sim.workspace_bytes(), dtype=torch.uint8, device="cuda:0")
    for _ in range(2):
        sim.step(advance=True, workspace=ws)
    for a, b in zip((_np(a) for a in sim.positions()), ref):
        np.testing.assert_array_equal(a, b)
    small = torch.empty(1024, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(RuntimeError, match="workspace"):
        sim.step(workspace=small)


def test_autotune_keeps_the_state_and_records_its_choice(ccc):
    sc = _scene(20, 8)
    sim = _sim(ccc, sc, iters=2, max_candidates=512)
    before = [_np(a) for a in sim.positions()]
    sim.autotune()
    after = [_np(a) for a in sim.positions()]
    for a, b in zip(after, before):
        np.testing.assert_array_equal(a, b)
    cfg = sim.config()
    assert cfg["ccd_blocks_per_cu"] in (64, 128, 256) and cfg["morton"] in (0, 1)
    assert len(sim.tuning_times()) == 6
    # a second simulation of the same size bucket takes the remembered choice without timing
    sim2 = _sim(ccc, _scene(20, 9), iters=2, max_candidates=512)
    sim2.autotune()
    assert sim2.tuning_times() == [] and sim2.config()["ccd_blocks_per_cu"] == cfg["ccd_blocks_per_cu"]
    sim.step()
    assert sim.counts()[0][3] > 0


def test_fp32_simulation(ccc, oracle):
    from cccollisions_amd import scenes
    sc = _scene(12, 4)
    sim = _sim(ccc, sc, dtype=np.float32, iters=1, max_candidates=512)
    sim.step()
    x1 = _np(sim.positions()[1])
    dv, dx, cnt = (_np(a) for a in sim.last_impulses())
    assert x1.dtype == np.float32
    np.testing.assert_array_equal(x1, scenes.apply_corrections(sc.faces, sc.x1.astype(np.float32), dx, cnt))
    t0, t1, m = (a.astype(np.float32) for a in (sc.x0[sc.faces], sc.x1[sc.faces], sc.corner_masses()))
    bb, tree = oracle.bvh(t0, t1)[:2]
    _, _, ocnt, _, _, counts = oracle.compute_impulses(bb, tree, t0, t1, m)
    np.testing.assert_array_equal(cnt, ocnt)


def test_bad_faces_raise(ccc):
    import torch
    sc = _scene(6, 1)
    f = sc.faces.copy()
    f[0, 0] = sc.x0.shape[0]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    with pytest.raises(RuntimeError, match="out of range"):
        ccc.Sim(t(sc.x0), t(sc.x1), t(f), t(sc.corner_masses()))
