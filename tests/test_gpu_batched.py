"""The batched workload at full size: 64 independent folded cloths of 2k-200k triangles
(log-uniform sizes, 3.69 M triangles in all), first impulse pass.

* packed into one resident simulation (one BVH, one traversal and one narrow phase over all
  meshes) against the plain per-mesh path (bvh_motion + compute_impulses on each mesh
  alone): for every mesh cnt equal, dv / dx within 1e-9 relative, and the packed stage counts
  are the sums of the per-mesh ones (no pair crosses a mesh);
* the packed functional entry point (compute_impulses_packed) against the same per-mesh
  results;
* six meshes spread over the size range against the oracle's multithreaded pass.

The per-mesh loop is the reference's calling pattern, one mesh per call
(src/collisions_impulses.cu:1122, compute_impulses over a [B, N] batch).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _close(a, b, name):
    fin = np.isfinite(b)
    np.testing.assert_array_equal(np.isfinite(a), fin, err_msg=name)
    a, b = np.where(fin, a, 0.0), np.where(fin, b, 0.0)
    scale = max(float(np.abs(b).max()), 1e-300)
    assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * scale), name


@pytest.fixture(scope="module")
def batch():
    from cccollisions_amd import scenes
    return scenes.batched_cloths(64)


@pytest.fixture(scope="module")
def per_mesh(ccc, batch):
    """(dv, dx, cnt, counts) of each mesh alone, on the host."""
    import torch
    dev = torch.device("cuda:0")
    out = []
    for sc in batch:
        t0 = torch.from_numpy(sc.x0[sc.faces][None]).to(dev)
        t1 = torch.from_numpy(sc.x1[sc.faces][None]).to(dev)
        m = torch.from_numpy(sc.corner_masses()[None]).to(dev)
        bb, tree = ccc.bvh_motion(t0, t1)
        dv, dx, cnt = ccc.compute_impulses(bb, tree, t0, t1, m, 512, 16)
        out.append((dv[0].cpu().numpy(), dx[0].cpu().numpy(), cnt[0].cpu().numpy(), ccc.last_counts()[0]))
    return out


def _packed_inputs(batch):
    voff = np.cumsum([0] + [sc.x0.shape[0] for sc in batch])
    foff = np.cumsum([0] + [sc.faces.shape[0] for sc in batch]).tolist()
    x0 = np.concatenate([sc.x0 for sc in batch])
    x1 = np.concatenate([sc.x1 for sc in batch])
    f = np.concatenate([sc.faces + voff[i] for i, sc in enumerate(batch)])
    m = np.concatenate([sc.corner_masses() for sc in batch])
    return x0, x1, f, m, foff


def test_batch_is_the_full_workload(batch):
    sizes = [sc.num_triangles for sc in batch]
    assert len(batch) == 64 and min(sizes) < 3000 and max(sizes) > 180000
    assert sum(sizes) == 3687810


def test_packed_simulation_equals_per_mesh_path(ccc, batch, per_mesh):
    import torch
    dev = torch.device("cuda:0")
    x0, x1, f, m, foff = _packed_inputs(batch)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    sim = ccc.Sim(t(x0), t(x1), t(f), t(m), mesh_offsets=foff, iters=1, max_candidates=512)
    sim.step()
    dv, dx, cnt = (a.cpu().numpy() for a in sim.last_impulses())
    packed_counts = sim.counts()[0]
    del sim
    tot = np.zeros(4, np.int64)
    for i, (sdv, sdx, scnt, counts) in enumerate(per_mesh):
        lo, hi = foff[i], foff[i + 1]
        np.testing.assert_array_equal(cnt[lo:hi], scnt, err_msg=f"cnt mesh {i}")
        _close(dv[lo:hi], sdv, f"dv mesh {i}")
        _close(dx[lo:hi], sdx, f"dx mesh {i}")
        tot += counts
    assert tot[3] > 10**8
    assert list(packed_counts) == [int(v) for v in tot]


def test_packed_entry_point_equals_per_mesh_path(ccc, batch, per_mesh):
    import torch
    dev = torch.device("cuda:0")
    foff = np.cumsum([0] + [sc.faces.shape[0] for sc in batch]).tolist()
    t0 = torch.from_numpy(np.concatenate([sc.x0[sc.faces] for sc in batch])).to(dev)
    t1 = torch.from_numpy(np.concatenate([sc.x1[sc.faces] for sc in batch])).to(dev)
    m = torch.from_numpy(np.concatenate([sc.corner_masses() for sc in batch])).to(dev)
    dv, dx, cnt = ccc.compute_impulses_packed(t0, t1, m, foff, 512)
    K, M, V, R = ccc.last_counts()[0]
    dv, dx, cnt = dv.cpu().numpy(), dx.cpu().numpy(), cnt.cpu().numpy()
    tot = np.zeros(4, np.int64)
    for i, (sdv, sdx, scnt, counts) in enumerate(per_mesh):
        lo, hi = foff[i], foff[i + 1]
        np.testing.assert_array_equal(cnt[lo:hi], scnt, err_msg=f"cnt mesh {i}")
        _close(dv[lo:hi], sdv, f"dv mesh {i}")
        _close(dx[lo:hi], sdx, f"dx mesh {i}")
        tot += counts
    assert (K, M, V, R) == tuple(int(v) for v in tot)


def test_spread_of_meshes_against_the_oracle(oracle, batch, per_mesh):
    sizes = np.array([sc.num_triangles for sc in batch])
    order = np.argsort(sizes, kind="stable")
    picks = sorted({int(order[round(q * (len(order) - 1))]) for q in (0.0, 0.2, 0.4, 0.6, 0.8, 1.0)})
    assert len(picks) == 6
    for i in picks:
        sc = batch[i]
        t0, t1, m = sc.x0[sc.faces], sc.x1[sc.faces], sc.corner_masses()
        dv, dx, cnt, counts = oracle.impulse_pass(t0, t1, m, nthreads=16)
        sdv, sdx, scnt, scounts = per_mesh[i]
        print(f"mesh {i}: {sizes[i]} triangles, {counts['collisions']} collisions", flush=True)
        assert tuple(scounts) == (counts["candidates"], counts["subprims"], counts["roots"], counts["collisions"])
        np.testing.assert_array_equal(scnt, cnt)
        _close(sdv, dv, f"dv mesh {i}")
        _close(sdx, dx, f"dx mesh {i}")
