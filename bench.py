#!/usr/bin/env python3
"""Benchmark: CCD + impulse steps per second on the 130k-triangle folded cloth.

Workload (one "step"): a 256x256-vertex cloth (130,050 triangles, fp64) folded over itself
at x = 0.5 with a 1e-3 gap, end-of-step positions t1 = t0 + v/30 with v ~ N(0, 0.5)
(seeded).  A step runs 10 impulse passes; each pass gathers the end-of-step triangle soup,
builds the swept BVH (bvh_motion), runs compute_impulses (traversal -> 15-way VF/EE filter
-> cubic CCD -> impulse scatter) and moves every vertex of t1 by the mean of the finite
position corrections of its triangle corners.  The cloth is a resident simulation
(cccollisions_amd.Sim): inputs are in HBM before timing and each step is one replayed HIP
graph.

Prints one JSON line: throughput, per-step latency, the roofline of the dominant kernel
(its per-launch duration from HIP events on its stream, in one extra profiled step of the
same workload) and a CPU baseline (the oracle's OpenMP pass on this host's cores), with
sections for collect mode, the garment scene, the 64-mesh batch (packed + graph against the
plain per-mesh path), serving (256 resident simulations, Poisson arrivals) and Morton trees
on a face-shuffled mesh.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CCD+impulse steps/sec at 130k tris (batched vs plain); p99 request latency"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
FLOPS_PER_SUBPRIM = 162          # cubic-coefficient construction per VF/EE pair


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256, help="vertices per side of the cloth grid")
    ap.add_argument("--iters", type=int, default=10, help="impulse passes per step")
    ap.add_argument("--max-candidates", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed runs of each CPU baseline (median reported)")
    ap.add_argument("--cpu-single-stride", type=int, default=32,
                    help="the single-thread CPU baseline runs every n-th query leaf of a pass (scaled by n)")
    ap.add_argument("--batched-sims", type=int, default=64,
                    help="also time the packed multi-mesh path against the plain per-mesh path on this many "
                         "folded cloths of 2k-200k triangles (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=1)
    ap.add_argument("--no-collect", dest="collect", action="store_false",
                    help="skip timing find_collisions_continuous (collect mode) on the cloth")
    ap.add_argument("--collect-rows", type=int, default=288, help="collision rows per triangle for the collect timing")
    ap.add_argument("--no-garment", dest="garment", action="store_false",
                    help="skip the garment-on-body scene (static 24k-triangle capsule body + 100k-triangle "
                         "sleeve moving inward)")
    ap.add_argument("--serving-sims", type=int, default=256,
                    help="also run the serving loop over this many resident folded cloths (lognormal sizes, "
                         "median 20k triangles; 0: skip)")
    ap.add_argument("--serving-max-tris", type=int, default=500000)
    ap.add_argument("--serving-requests", type=int, default=256)
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--serving-load", type=float, default=0.5,
                    help="offered load of the open-loop run as a fraction of the saturation rate")
    ap.add_argument("--no-morton", dest="morton", action="store_false",
                    help="skip the Morton-tree comparison on a face-shuffled cloth")
    ap.add_argument("--no-proximity", dest="proximity", action="store_false",
                    help="skip the static proximity query on the 7,938-triangle cloth")
    ap.add_argument("--main-only", action="store_true",
                    help="time only the 130k-triangle cloth (no other section)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB once (a plain copy dispatch of "
                         "known size, used to scale the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.main_only:
        args.garment, args.collect, args.batched_sims, args.serving_sims, args.morton = False, False, 0, 0, False
        args.proximity = False
    return args


def tensors(torch, *arrays):
    dev = torch.device("cuda:0")
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrays]


def make_sim(ccc, torch, sc, iters, max_candidates, **kw):
    x0, x1, f, m = tensors(torch, sc.x0, sc.x1, sc.faces, sc.corner_masses())
    return ccc.Sim(x0, x1, f, m, iters=iters, max_candidates=max_candidates, **kw)


def time_sim(torch, sim, reset, steps, warmup):
    """Steps of a resident simulation, each from the same state (reset() restores it, outside
    the timed region of the step)."""
    for _ in range(warmup):
        reset()
        sim.step()
    torch.cuda.synchronize()
    lat = []
    for _ in range(steps):
        reset()
        torch.cuda.synchronize()
        t = time.perf_counter()
        sim.step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
    return lat


class PlainBatch:
    """The plain per-mesh GPU path (no packing, no graph): per mesh and pass, bvh_motion +
    compute_impulses and the vertex update as torch ops."""

    def __init__(self, ccc, torch, scenes_list, iters, max_candidates):
        self.ccc, self.torch, self.iters, self.maxc = ccc, torch, iters, max_candidates
        self.meshes = []
        for sc in scenes_list:
            x0, x1, f, m = tensors(torch, sc.x0, sc.x1, sc.faces, sc.corner_masses())
            self.meshes.append((f, f.reshape(-1), x0[f].unsqueeze(0).contiguous(), x1, m.unsqueeze(0).contiguous(), x0.shape[0]))

    def run(self):
        torch, ccc = self.torch, self.ccc
        for f, flat, t0, x1, m, nv in self.meshes:
            for _ in range(self.iters):
                t1 = x1[f].unsqueeze(0).contiguous()
                bb, tree = ccc.bvh_motion(t0, t1)
                _, dx, cnt = ccc.compute_impulses(bb, tree, t0, t1, m, self.maxc, 16)
                d = dx[0].reshape(-1, 3)
                ok = torch.isfinite(d).all(dim=1)
                d = torch.where(ok[:, None], d, torch.zeros_like(d))
                k = torch.where(ok, cnt[0].reshape(-1).to(x1.dtype), torch.zeros_like(ok, dtype=x1.dtype))
                s = torch.zeros((nv, 3), dtype=x1.dtype, device=x1.device).index_add_(0, flat, d)
                c = torch.zeros((nv,), dtype=x1.dtype, device=x1.device).index_add_(0, flat, k)
                x1 = x1 + torch.where(c[:, None] > 0, s / c.clamp(min=1)[:, None], torch.zeros_like(s))


def garment_section(ccc, torch, scenes, args):
    """Garment vs body: 24k-triangle static capsule body + 100k-triangle sleeve around the
    right arm at radius + 2e-3 with inward speeds |N(0, 0.3)|, one soup, fp64, as a resident
    simulation (steps of `iters` passes)."""
    sc = scenes.garment_on_body()
    sim = make_sim(ccc, torch, sc, args.iters, args.max_candidates)
    x0, x1 = tensors(torch, sc.x0, sc.x1)
    lat = time_sim(torch, sim, lambda: sim.set_positions(x0, x1), args.steps, 1)
    counts = sim.counts()
    return {"triangles": sc.num_triangles, "steps": args.steps, "steps_per_s": len(lat) / sum(lat),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "first_pass_counts": counts[0],
            "mean_collisions_per_pass": float(np.mean([c[3] for c in counts]))}


def collect_section(ccc, torch, sc, args):
    """find_collisions_continuous (collect mode: collision rows and roots instead of impulses)
    on the cloth's first-pass inputs, with room for every collision."""
    f, x0, x1 = tensors(torch, sc.faces, sc.x0, sc.x1)
    t0, t1 = x0[f].unsqueeze(0).contiguous(), x1[f].unsqueeze(0).contiguous()
    bb, tree = ccc.bvh_motion(t0, t1)

    def call():
        return ccc.find_collisions_continuous(bb, tree, t0, t1, args.max_candidates, args.collect_rows)

    call()
    torch.cuda.synchronize()
    ccc.set_profiling(True)
    ts = []
    for _ in range(3):
        torch.cuda.synchronize()
        a = time.perf_counter()
        call()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - a)
    kt = ccc.kernel_times()
    ccc.set_profiling(False)
    counts = ccc.last_counts()[0]
    return {"calls": 3, "ms_per_call": 1e3 * float(np.median(ts)), "ccd_collect_ms_per_call": kt["ccd_collect"][0] / 3,
            "rows_capacity": sc.num_triangles * args.collect_rows, "counts": list(map(int, counts))}


def proximity_section(ccc, torch, scenes):
    """Static proximity (find_proximity, threshold 2e-3) on a 64x64-vertex static cloth of
    7,938 triangles: bvh + query per call, inputs resident, median of 20 calls."""
    v, f = scenes.proximity_cloth(64, seed=1)
    tri = torch.from_numpy(np.ascontiguousarray(v[f][None])).cuda()

    def call():
        bb, tree = ccc.bvh(tri)
        return ccc.find_proximity(bb, tree, tri, 2e-3, 16)

    call()
    torch.cuda.synchronize()
    ts = []
    for _ in range(20):
        a = time.perf_counter()
        call()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - a)
    counts = ccc.last_counts()[0]
    return {"triangles": int(f.shape[0]), "calls": 20, "ms_per_call": 1e3 * float(np.median(ts)),
            "p99_ms": 1e3 * float(np.percentile(ts, 99)), "counts": list(map(int, counts))}


def batched_section(ccc, torch, scenes, args):
    """64 folded cloths of 2k-200k triangles (log-uniform sizes, seeded), one step each:
    packed into one resident simulation (one BVH / traversal / narrow phase over all meshes
    per pass, the step replayed as a HIP graph, launch parameters from the autotuner) against
    the plain per-mesh path (per mesh and pass: bvh_motion + compute_impulses + torch vertex
    update, no graph), and against one resident simulation per mesh run one after another."""
    sims = scenes.batched_cloths(args.batched_sims)
    tris = int(sum(sc.num_triangles for sc in sims))
    voff = np.cumsum([0] + [sc.x0.shape[0] for sc in sims])
    foff = np.cumsum([0] + [sc.faces.shape[0] for sc in sims]).tolist()
    X0, X1 = np.concatenate([sc.x0 for sc in sims]), np.concatenate([sc.x1 for sc in sims])
    x0, x1, f, m = tensors(torch, X0, X1, np.concatenate([sc.faces + voff[i] for i, sc in enumerate(sims)]),
                           np.concatenate([sc.corner_masses() for sc in sims]))
    packed = ccc.Sim(x0, x1, f, m, mesh_offsets=foff, iters=args.iters, max_candidates=args.max_candidates)
    packed.autotune()
    lat_p = time_sim(torch, packed, lambda: packed.set_positions(x0, x1), args.batched_steps, 1)
    counts = packed.counts()
    first = packed.positions()[1]
    cfg = packed.config()
    del packed
    torch.cuda.empty_cache()
    plain = PlainBatch(ccc, torch, sims, args.iters, args.max_candidates)
    plain.run()
    torch.cuda.synchronize()
    lat_q = []
    for _ in range(args.batched_steps):
        t = time.perf_counter()
        plain.run()
        torch.cuda.synchronize()
        lat_q.append(time.perf_counter() - t)
    del plain
    torch.cuda.empty_cache()
    tp, tq = sum(lat_p), sum(lat_q)
    return {
        "sims": len(sims), "triangles": tris, "sizes": [int(sc.num_triangles) for sc in sims],
        "steps": args.batched_steps, "impulse_iterations": args.iters,
        "packed_steps_per_s": len(lat_p) / tp, "plain_steps_per_s": len(lat_q) / tq, "packed_over_plain": tq / tp,
        "packed_p99_ms": 1e3 * float(np.percentile(lat_p, 99)), "plain_p99_ms": 1e3 * float(np.percentile(lat_q, 99)),
        "packed_config": cfg, "first_pass_counts": counts[0],
        "mean_collisions_per_pass": float(np.mean([c[3] for c in counts])),
        "positions_finite": bool(torch.isfinite(first).all()),
    }


def serving_section(ccc, torch, scenes, args):
    """Resident folded cloths with lognormal sizes (median 20k triangles, capped), one
    Simulation each (deterministic impulse sums: a request's result does not depend on what
    ran beside it); each request is one step of one simulation and advances it.  First the
    saturation throughput (all requests queued at t = 0), then an open loop with Poisson
    arrivals at args.serving_load of that rate: completed steps/s and p50 / p99 latency."""
    from cccollisions_amd import serving
    rng = np.random.default_rng(5)
    sizes = np.minimum(np.exp(rng.normal(np.log(20000), 0.8, args.serving_sims)), args.serving_max_tris)
    cl = [scenes.folded_cloth(scenes.grid_size_for_triangles(int(s)), seed=500 + i) for i, s in enumerate(sizes)]
    sims = [serving.Simulation(ccc, torch, sc.faces, sc.x0, sc.x1, sc.corner_masses(), iters=args.iters,
                               max_candidates=args.max_candidates) for sc in cl]
    del cl
    tris = [s.num_triangles for s in sims]
    t_tune = time.perf_counter()
    server = serving.CollisionServer(ccc, torch, sims, slots=args.serving_slots, autotune=True)
    t_tune = time.perf_counter() - t_tune

    def run(offsets, sim_ids):
        # request k is submitted offsets[k] seconds after the run starts
        start = server.now()
        reqs = []
        for a, s in zip(offsets, sim_ids):
            while True:
                left = start + a - server.now()
                if left <= 0:
                    break
                time.sleep(min(0.002, left))
            reqs.append(server.submit(int(s), arrival=server.now()))
        while any(r.done is None for r in reqs):
            if server.errors:
                raise server.errors[0]
            time.sleep(0.002)
        span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        lat = np.array([r.latency for r in reqs])
        return {"requests": len(reqs), "steps_per_s": len(reqs) / span,
                "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                "mean_batch": float(np.mean([r.batch_size for r in reqs]))}

    # warm-up: one request per simulation (captures every simulation's graph per slot it meets)
    run(np.zeros(len(sims)), np.arange(len(sims)))
    n = args.serving_requests
    sat = run(np.zeros(n), rng.integers(0, len(sims), n))
    rate = args.serving_load * sat["steps_per_s"]
    gaps = rng.exponential(1.0 / rate, n)
    open_loop = run(np.cumsum(gaps) - gaps[0], rng.integers(0, len(sims), n))
    server.close()
    return {"sims": len(sims), "triangles_median": int(np.median(tris)), "triangles_max": int(max(tris)),
            "triangles_total": int(sum(tris)), "slots": args.serving_slots, "impulse_iterations": args.iters,
            "deterministic": True, "resident_bytes": int(sum(s.footprint() for s in sims)),
            "setup_and_autotune_s": t_tune, "saturation": sat, "offered_steps_per_s": rate, "open_loop": open_loop,
            "offered_spec_steps_per_s": 2000.0}


def morton_section(ccc, torch, scenes, args):
    """A 128x128 folded cloth whose faces are stored in random order: one step as a resident
    simulation on the reference's tree (leaves in face order) and on fitted 30- / 63-bit
    Morton trees; step time, the traversal kernel's time and the (equal) pair counts."""
    sc = scenes.folded_cloth(128, seed=9)
    perm = np.random.default_rng(9).permutation(sc.faces.shape[0])
    sc.faces = sc.faces[perm]
    out = {"triangles": int(sc.faces.shape[0])}
    for mode in (0, 1, 2):
        sim = make_sim(ccc, torch, sc, args.iters, args.max_candidates, morton=mode)
        x0, x1 = tensors(torch, sc.x0, sc.x1)
        lat = time_sim(torch, sim, lambda: sim.set_positions(x0, x1), 2, 1)
        sim.set_positions(x0, x1)
        ccc.set_profiling(True)
        sim.step()
        kt = ccc.kernel_times()
        ccc.set_profiling(False)
        out[f"mode{mode}"] = {"step_ms": 1e3 * float(np.median(lat)), "traverse_ms_per_step": kt.get("traverse", (0, 0))[0],
                              "first_pass_counts": sim.counts()[0]}
        del sim
    return out


def roofline(kt, counts):
    """Roofline of the dominant kernel of one profiled step."""
    name, (ms, launches) = max(((k, v) for k, v in kt.items() if k.startswith("ccd")), key=lambda kv: kv[1][0])
    m_total = float(sum(c[1] for c in counts))
    avg_s = ms / 1e3 / max(launches, 1)
    work = FLOPS_PER_SUBPRIM * m_total / max(launches, 1)
    achieved = work / avg_s / 1e12
    traffic, src = measured_traffic("k_ccd")
    return {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic, "kernel": name, "avg_launch_ms": avg_s * 1e3,
            "launches": launches,
            "work_per_launch": f"{FLOPS_PER_SUBPRIM} fp64 flop x {m_total / max(launches, 1):.4g} VF/EE pairs",
            "peak_note": "MI355X fp64 vector peak (the kernel runs on the fp64 VALU; no matrix instructions)",
            "traffic_source": src}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (profiles/traffic_<kernel>_folded_cloth_130k.md), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_folded_cloth_130k.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return float(json.loads(block)["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)
    except (OSError, IndexError, KeyError, ValueError):
        return None, None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(scene, iters, runs, single_stride):
    """The oracle's impulse pass (bvh_motion + compute_impulses, OpenMP over query leaves) on
    this host: a full pass with the process's thread share (median of `runs`), and one thread
    on every single_stride-th query leaf (median of `runs`, scaled by single_stride)."""
    from oracle import oracle as O
    t0, t1 = scene.soups()
    m = scene.corner_masses()
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    threads = max(1, min(share, int(os.environ.get("OMP_NUM_THREADS", share))))
    O.impulse_pass(t0[:2000], t1[:2000], m[:2000], threads, 1)   # load / warm the library
    full, counts = [], None
    for _ in range(runs):
        t = time.perf_counter()
        _, _, _, counts = O.impulse_pass(t0, t1, m, threads, 1)
        full.append(time.perf_counter() - t)
    single = []
    for _ in range(runs):
        t = time.perf_counter()
        O.impulse_pass(t0, t1, m, 1, single_stride)
        single.append(time.perf_counter() - t)
    pass_s = float(np.median(full))
    single_pass_s = float(np.median(single)) * single_stride
    return {"value": 1.0 / (iters * pass_s), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": (f"full impulse passes (bvh_motion + compute_impulses) of the 130,050-triangle cloth, median of {runs} "
                       f"runs with {threads} threads: {pass_s:.2f} s per pass ({counts['candidates']} candidate pairs, "
                       f"{counts['collisions']} collisions), x{iters} passes per step"),
            "pass_s_runs": full, "single_thread_steps_per_s": 1.0 / (iters * single_pass_s),
            "single_thread_sample": (f"one thread, every {single_stride}-th query leaf of a pass, median of {runs} runs "
                                     f"({float(np.median(single)):.2f} s) scaled x{single_stride}"),
            "host_cpus": os.cpu_count(), "affinity_cpus": share, "cpu_model": cpu_model()}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              "(across GPUs the workload would shard by simulation, one process per GPU)", file=sys.stderr)
        sys.exit(2)
    import torch
    import cccollisions_amd as ccc
    from cccollisions_amd import scenes

    scene = scenes.folded_cloth(args.grid, seed=2)
    sim = make_sim(ccc, torch, scene, args.iters, args.max_candidates)
    x0, x1 = tensors(torch, scene.x0, scene.x1)
    for _ in range(args.warmup):
        sim.set_positions(x0, x1)
        sim.step()
    torch.cuda.synchronize()
    # timed region: K steps of the resident simulation, each from the cloth's initial state
    # (the reset is a device copy of the positions, inside the timed region)
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        sim.set_positions(x0, x1)
        sim.step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
    elapsed = time.perf_counter() - t_start
    counts = sim.counts()
    # one more step with per-kernel HIP events (eager launches), for the roofline and the
    # per-kernel times
    sim.set_positions(x0, x1)
    torch.cuda.synchronize()
    ccc.set_profiling(True)
    sim.step()
    kt = ccc.kernel_times()
    ccc.set_profiling(False)
    prof_counts = sim.counts()
    cfg = sim.config()
    del sim
    torch.cuda.empty_cache()

    if args.copy_calibration_mib:
        src = torch.empty(args.copy_calibration_mib << 20, dtype=torch.uint8, device="cuda:0").fill_(1)
        torch.cuda.synchronize()
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"folded cloth {args.grid}x{args.grid} vertices ({scene.num_triangles} triangles), fp64, "
                         f"{args.iters} impulse passes per step (bvh_motion + compute_impulses + vertex update), "
                         "resident simulation, each step one replayed HIP graph"),
            "triangles": scene.num_triangles,
            "impulse_iterations": args.iters,
            "max_candidates_per_triangle": args.max_candidates,
            "mean_candidates_per_pass": float(np.mean([c[0] for c in counts])),
            "mean_subprims_per_pass": float(np.mean([c[1] for c in counts])),
            "mean_roots_per_pass": float(np.mean([c[2] for c in counts])),
            "mean_collisions_per_pass": float(np.mean([c[3] for c in counts])),
            "sim_config": cfg,
            "parallelism": "single GPU",
        },
        "roofline": roofline(kt, prof_counts),
        "kernels_ms_per_step": {k: v[0] for k, v in sorted(kt.items())},
    }
    if args.collect:
        out["collect"] = collect_section(ccc, torch, scene, args)
    if args.garment:
        out["garment"] = garment_section(ccc, torch, scenes, args)
    if args.morton:
        out["morton_shuffled"] = morton_section(ccc, torch, scenes, args)
    if args.proximity:
        out["proximity"] = proximity_section(ccc, torch, scenes)
    if args.batched_sims:
        out["batched"] = batched_section(ccc, torch, scenes, args)
    if args.serving_sims:
        out["serving"] = serving_section(ccc, torch, scenes, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, args.iters, args.cpu_runs, args.cpu_single_stride)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
