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


def _cloths(sizes, seed0=40):
    from cccollisions_amd import scenes
    return [scenes.folded_cloth(n, seed=seed0 + n) for n in sizes]


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix(ccc):
    """Five resident cloths of different sizes, three requests each (fifteen, interleaved),
    two worker slots: every request's resulting state is, bit for bit, the state the same
    simulation reaches after the same number of steps run alone (deterministic sums)."""
    import torch
    cl = _cloths((6, 10, 14, 8, 12))
    mk = lambda sc: serving.Simulation(ccc, torch, sc.faces, sc.x0, sc.x1, sc.corner_masses(), iters=3)  # noqa: E731
    sims = [mk(sc) for sc in cl]
    server = serving.CollisionServer(ccc, torch, sims, slots=2, keep_results=True)
    reqs = [server.submit(i % len(sims)) for i in range(3 * len(sims))]
    server.close()
    assert all(s.steps == 3 for s in sims)
    assert max(r.batch_size for r in reqs) > 1
    for i, sc in enumerate(cl):
        alone = mk(sc)
        mine = [r for r in reqs if r.sim == i]
        assert len(mine) == 3
        for r in mine:
            alone.step()
            x0 = alone.positions()[0]
            assert not torch.equal(x0, torch.from_numpy(sc.x0).to(x0.device))
            assert torch.equal(r.result, x0), (i, float((r.result - x0).abs().max()))
            assert r.done >= r.start >= r.arrival


@pytest.mark.gpu
def test_server_steps_match_the_oracle_step(ccc, oracle):
    """Three advancing steps of one pass each: every step, from the GPU's state before it,
    against the same pass composed on the CPU from the oracle and scenes.apply_corrections
    (positions within 1e-9 of the pass's correction), and the advance rule x0 <- x1,
    x1 <- x1 + (x1 - x0) exact."""
    import torch
    from cccollisions_amd import scenes
    sc = _cloths((10,), seed0=7)[0]
    m = sc.corner_masses()
    sim = serving.Simulation(ccc, torch, sc.faces, sc.x0, sc.x1, m, iters=1)
    for k in range(3):
        x0, x1 = (a.cpu().numpy() for a in sim.positions())
        sim.step()
        g0, g1 = (a.cpu().numpy() for a in sim.positions())
        t0, t1 = x0[sc.faces], x1[sc.faces]
        bb, tree = oracle.bvh(t0, t1)[:2]
        _, dx, cnt, _, _, counts = oracle.compute_impulses(bb, tree, t0, t1, m)
        assert counts["collisions"] > 0
        ref = scenes.apply_corrections(sc.faces, x1, dx, cnt)
        corr = float(np.abs(ref - x1).max())
        assert corr > 0
        assert float(np.abs(g0 - ref).max()) <= 1e-9 * corr + 1e-15, k
        np.testing.assert_array_equal(g1, g0 + (g0 - x0))


@pytest.mark.gpu
def test_server_rejects_an_hbm_budget_it_cannot_hold(ccc):
    import torch
    sims = [serving.Simulation(ccc, torch, sc.faces, sc.x0, sc.x1, sc.corner_masses(), iters=1) for sc in _cloths((6, 8))]
    with pytest.raises(ValueError, match="HBM budget"):
        serving.CollisionServer(ccc, torch, sims, slots=2, hbm_budget=1 << 20)
