"""The serving loop: admission logic on the CPU (FIFO, HBM budget, one request per
simulation at a time) and, on the GPU, that a request's result is the same whether its
simulation ran alone or packed with others."""
import threading

import numpy as np
import pytest

from cccollisions_amd import serving


def _sched(budget=100, fp=None):
    fp = fp or {0: 40, 1: 40, 2: 40, 3: 90}
    return serving.Scheduler(fp, budget)


def test_admission_fifo_with_budget():
    s = _sched()
    for sim, t in [(0, 0.0), (1, 0.1), (2, 0.2), (3, 0.3)]:
        s.submit(serving.Request(sim, t))
    b = s.take()
    assert [r.sim for r in b] == [0, 1]          # 40 + 40 <= 100, a third would exceed it
    b2 = s.take()
    assert [r.sim for r in b2] == [2]            # 3 (90) does not fit beside 2 (40)
    s.release(b)
    s.release(b2)
    assert [r.sim for r in s.take()] == [3]


def test_oversized_request_runs_alone():
    s = _sched(budget=50, fp={0: 500, 1: 10})
    s.submit(serving.Request(0, 0.0))
    s.submit(serving.Request(1, 0.1))
    assert [r.sim for r in s.take()] == [0]
    assert [r.sim for r in s.take()] == [1]


def test_one_request_per_simulation_in_flight():
    s = _sched()
    s.submit(serving.Request(0, 0.0))
    s.submit(serving.Request(0, 0.1))
    s.submit(serving.Request(1, 0.2))
    b = s.take()
    assert [r.sim for r in b] == [0, 1]          # the second request of sim 0 waits
    got = []
    t = threading.Thread(target=lambda: got.append(s.take()))
    t.start()
    t.join(0.2)
    assert t.is_alive()                          # blocked: sim 0 is busy
    s.release(b)
    t.join(2.0)
    assert [r.sim for r in got[0]] == [0] and got[0][0].arrival == 0.1


def test_close_drains_then_stops():
    s = _sched()
    s.submit(serving.Request(2, 0.0))
    s.close()
    assert [r.sim for r in s.take()] == [2]
    assert s.take() == []


def test_footprint_grows_with_mesh_and_capacity():
    a = serving.footprint_bytes(20000, 512)
    assert serving.footprint_bytes(40000, 512) > a > serving.footprint_bytes(20000, 64)


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix(ccc):
    """Twelve requests over five folded cloths of different sizes, two worker slots: every
    result equals the simulation's step computed alone (impulse sums reordered: 1e-9 of the
    correction)."""
    import torch
    from cccollisions_amd import scenes
    cl = [scenes.folded_cloth(n, seed=40 + n) for n in (6, 10, 14, 8, 12)]
    sims = [serving.Simulation(torch, sc.faces, sc.x0, sc.x1, sc.corner_masses()) for sc in cl]
    server = serving.CollisionServer(ccc, torch, sims, slots=2, iters=1, max_candidates=512)
    reqs = [server.submit(i % len(sims)) for i in range(12)]
    server.close()
    assert max(r.batch_size for r in reqs) > 1
    for r in reqs:
        alone = serving.impulse_step(ccc, torch, [sims[r.sim]], 1, 512)[0]
        corr = float((alone - sims[r.sim].x1).abs().max())
        assert corr > 0
        diff = float((r.result - alone).abs().max())
        assert diff <= 1e-9 * corr + 1e-15, (r.sim, diff, corr)
        assert r.done >= r.start >= r.arrival


@pytest.mark.gpu
def test_impulse_step_matches_the_oracle_step(ccc, oracle):
    """The caller step on the GPU (bvh_motion + compute_impulses + mean-correction vertex
    update, serving.impulse_step) against the same pass composed on the CPU from the oracle
    and scenes.apply_corrections, for three passes in a row, each from the GPU's previous
    positions: positions agree to 1e-9 of the pass's correction.  (Chained passes are not
    compared end to end: the atomic summation order leaves 1e-17 differences in the
    positions, and the next pass's borderline root and filter decisions amplify them.)"""
    import torch
    from cccollisions_amd import scenes
    sc = scenes.folded_cloth(10, seed=7)
    t0 = sc.x0[sc.faces]
    m = sc.corner_masses()
    x_prev = sc.x1.copy()
    for k in range(3):
        sim = serving.Simulation(torch, sc.faces, sc.x0, x_prev, m)
        got = serving.impulse_step(ccc, torch, [sim], 1, 512)[0].cpu().numpy()
        bb, tree = oracle.bvh(t0, x_prev[sc.faces])[:2]
        _, dx, cnt, _, _, counts = oracle.compute_impulses(bb, tree, t0, x_prev[sc.faces], m)
        assert counts["collisions"] > 0
        ref = scenes.apply_corrections(sc.faces, x_prev, dx, cnt)
        corr = float(np.abs(ref - x_prev).max())
        assert corr > 0
        diff = float(np.abs(got - ref).max())
        assert diff <= 1e-9 * corr + 1e-15, (k, diff, corr)
        x_prev = got
