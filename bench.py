#!/usr/bin/env python3
"""Benchmark: CCD + impulse steps per second on the 130k-triangle folded cloth.

Workload (one "step"): a 256x256-vertex cloth (130,050 triangles, fp64) folded over itself
at x = 0.5 with a 1e-3 gap, end-of-step positions t1 = t0 + v/30 with v ~ N(0, 0.5)
(seeded).  A step runs 10 impulse iterations; each iteration gathers the triangle soups,
builds the swept BVH (`bvh_motion`), runs `compute_impulses` (traversal -> 15-way VF/EE
filter -> cubic CCD -> impulse scatter) and moves every vertex of t1 by the mean of the
finite position corrections of its triangle corners (scenes.apply_corrections).  Inputs are resident in HBM before timing.

Prints one JSON line: throughput, per-step latency, the roofline of the dominant kernel
(its duration measured with HIP events on the stream it runs on) and a CPU baseline (the
oracle's OpenMP pass on a bounded sample of the same workload, on this host's cores).
"""
import argparse
import json
import os
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
    ap.add_argument("--iters", type=int, default=10, help="impulse iterations per step")
    ap.add_argument("--max-candidates", type=int, default=512)
    ap.add_argument("--cpu-stride", type=int, default=2,
                    help="CPU baseline runs every n-th query leaf of one pass")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    ap.add_argument("--serving-sims", type=int, default=32,
                    help="also run the serving loop over this many resident folded cloths (lognormal sizes, "
                         "median 20k triangles; 0: skip)")
    ap.add_argument("--serving-max-tris", type=int, default=200000)
    ap.add_argument("--serving-requests", type=int, default=128)
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--serving-load", type=float, default=0.5,
                    help="offered load of the open-loop run as a fraction of the saturation rate")
    ap.add_argument("--main-only", action="store_true",
                    help="time only the 130k-triangle cloth (no collect, garment, batched or serving section)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB once (a plain copy dispatch of "
                         "known size, used to scale the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.main_only:
        args.garment, args.collect, args.batched_sims, args.serving_sims = False, False, 0, 0
    return args


class Step:
    """One CCD + impulse step on the GPU (inputs already on the device)."""

    def __init__(self, ccc, torch, scene, iters, max_candidates):
        dev = torch.device("cuda:0")
        self.ccc, self.torch, self.iters, self.maxc = ccc, torch, iters, max_candidates
        self.faces = torch.from_numpy(scene.faces).to(dev)
        self.flat = self.faces.reshape(-1)
        self.x0 = torch.from_numpy(scene.x0).to(dev)
        self.x1 = torch.from_numpy(scene.x1).to(dev)
        self.t0 = self.x0[self.faces].unsqueeze(0).contiguous()
        self.masses = torch.from_numpy(scene.corner_masses()).to(dev).unsqueeze(0).contiguous()
        self.nv = scene.x0.shape[0]
        self.counts = []

    def run(self, iters=None):
        torch, ccc = self.torch, self.ccc
        x1 = self.x1
        for _ in range(self.iters if iters is None else iters):
            t1 = x1[self.faces].unsqueeze(0).contiguous()
            bb, tree = ccc.bvh_motion(self.t0, t1)
            dv, dx, cnt = ccc.compute_impulses(bb, tree, self.t0, t1, self.masses, self.maxc, 16)
            self.counts.append(ccc.last_counts()[0])
            d = dx[0].reshape(-1, 3)
            ok = torch.isfinite(d).all(dim=1)
            d = torch.where(ok[:, None], d, torch.zeros_like(d))
            k = torch.where(ok, cnt[0].reshape(-1).to(x1.dtype), torch.zeros_like(ok, dtype=x1.dtype))
            s = torch.zeros((self.nv, 3), dtype=x1.dtype, device=x1.device).index_add_(0, self.flat, d)
            c = torch.zeros((self.nv,), dtype=x1.dtype, device=x1.device).index_add_(0, self.flat, k)
            x1 = x1 + torch.where(c[:, None] > 0, s / c.clamp(min=1)[:, None], torch.zeros_like(s))
        return x1


class BatchStep:
    """One CCD + impulse step of many independent meshes of different sizes.  packed=True:
    each iteration is one compute_impulses_packed call over all meshes (one BVH, one
    traversal, one narrow phase); packed=False: the plain path, bvh_motion + compute_impulses
    per mesh.  The vertex update is the same packed index_add in both."""

    def __init__(self, ccc, torch, scenes_list, iters, max_candidates, packed):
        dev = torch.device("cuda:0")
        self.ccc, self.torch, self.iters, self.maxc, self.packed = ccc, torch, iters, max_candidates, packed
        voff = np.cumsum([0] + [sc.x0.shape[0] for sc in scenes_list])
        self.toff = np.cumsum([0] + [sc.faces.shape[0] for sc in scenes_list]).tolist()
        faces = np.concatenate([sc.faces + voff[i] for i, sc in enumerate(scenes_list)])
        self.faces = torch.from_numpy(faces).to(dev)
        self.flat = self.faces.reshape(-1)
        self.x0 = torch.from_numpy(np.concatenate([sc.x0 for sc in scenes_list])).to(dev)
        self.x1 = torch.from_numpy(np.concatenate([sc.x1 for sc in scenes_list])).to(dev)
        self.t0 = self.x0[self.faces].contiguous()
        self.masses = torch.from_numpy(np.concatenate([sc.corner_masses() for sc in scenes_list])).to(dev).contiguous()
        self.nv = self.x0.shape[0]
        self.counts = []
        self.first_pass = None      # positions after the very first impulse pass

    def run(self, iters=None):
        torch, ccc = self.torch, self.ccc
        x1 = self.x1
        for _ in range(self.iters if iters is None else iters):
            if self.first_pass is None and self.counts:
                self.first_pass = x1.clone()
            t1 = x1[self.faces].contiguous()
            if self.packed:
                dv, dx, cnt = ccc.compute_impulses_packed(self.t0, t1, self.masses, self.toff, self.maxc)
                self.counts.append(ccc.last_counts()[0])
            else:
                dxs, cnts, tot = [], [], np.zeros(4, np.int64)
                for a, b in zip(self.toff[:-1], self.toff[1:]):
                    t0b, t1b = self.t0[a:b].unsqueeze(0), t1[a:b].unsqueeze(0)
                    bb, tree = ccc.bvh_motion(t0b, t1b)
                    _, dxb, cb = ccc.compute_impulses(bb, tree, t0b, t1b, self.masses[a:b].unsqueeze(0), self.maxc, 16)
                    dxs.append(dxb[0])
                    cnts.append(cb[0])
                    tot += np.asarray(ccc.last_counts()[0])
                dx, cnt = torch.cat(dxs), torch.cat(cnts)
                self.counts.append(tot.tolist())
            d = dx.reshape(-1, 3)
            ok = torch.isfinite(d).all(dim=1)
            d = torch.where(ok[:, None], d, torch.zeros_like(d))
            k = torch.where(ok, cnt.reshape(-1).to(x1.dtype), torch.zeros_like(ok, dtype=x1.dtype))
            s = torch.zeros((self.nv, 3), dtype=x1.dtype, device=x1.device).index_add_(0, self.flat, d)
            c = torch.zeros((self.nv,), dtype=x1.dtype, device=x1.device).index_add_(0, self.flat, k)
            x1 = x1 + torch.where(c[:, None] > 0, s / c.clamp(min=1)[:, None], torch.zeros_like(s))
        return x1


def time_steps(torch, step, steps, warmup, warmup_iters=None):
    for _ in range(warmup):
        step.run(warmup_iters)
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        out = step.run()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
    return time.perf_counter() - t_start, lat, out


def garment_section(ccc, torch, scenes, args):
    """Garment vs body: 24k-triangle static capsule body + 100k-triangle sleeve around the
    right arm at radius + 2e-3 with inward speeds |N(0, 0.3)|, one soup, fp64; a step is
    `iters` impulse passes as for the cloth."""
    sc = scenes.garment_on_body()
    step = Step(ccc, torch, sc, args.iters, args.max_candidates)
    el, lat, _ = time_steps(torch, step, args.steps, 1)
    timed = step.counts[-args.steps * args.iters:]
    return {"triangles": sc.num_triangles, "steps": args.steps, "steps_per_s": args.steps / el,
            "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "first_pass_counts": list(map(int, step.counts[0])),
            "mean_collisions_per_iter": float(np.mean([c[3] for c in timed]))}


def collect_section(ccc, torch, step, args):
    """find_collisions_continuous (collect mode: collision rows and roots instead of impulses)
    on the cloth's first-pass inputs, with room for every collision."""
    t1 = step.x1[step.faces].unsqueeze(0).contiguous()
    bb, tree = ccc.bvh_motion(step.t0, t1)

    def call():
        return ccc.find_collisions_continuous(bb, tree, step.t0, t1, args.max_candidates, args.collect_rows)

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
            "rows_capacity": step.t0.shape[1] * args.collect_rows, "counts": list(map(int, counts))}


def batched_section(ccc, torch, scenes, args):
    """Folded cloths of 2k-200k triangles (log-uniform sizes, seeded): the packed path
    against the plain per-mesh path on the same simulation.  Reports how far apart the
    positions of the two paths are after their first pass next to the largest correction of
    that pass."""
    sims = scenes.batched_cloths(args.batched_sims)
    tris = int(sum(sc.num_triangles for sc in sims))
    packed = BatchStep(ccc, torch, sims, args.iters, args.max_candidates, True)
    plain = BatchStep(ccc, torch, sims, args.iters, args.max_candidates, False)
    # warm-up: two impulse passes each (allocator, code paths); the first one is also the
    # first pass of the comparison below
    tp, latp, _ = time_steps(torch, packed, args.batched_steps, 1, 2)
    tq, latq, _ = time_steps(torch, plain, args.batched_steps, 1, 2)
    # after one pass from the same input the two paths differ only by the order of the fp64
    # atomic sums; later passes amplify such differences through borderline collisions
    moved = float((plain.first_pass - plain.x1).abs().max())
    diff = float((packed.first_pass - plain.first_pass).abs().max())
    return {
        "sims": len(sims), "triangles": tris,
        "sizes": [int(sc.num_triangles) for sc in sims],
        "steps": args.batched_steps, "impulse_iterations": args.iters,
        "packed_steps_per_s": args.batched_steps / tp, "plain_steps_per_s": args.batched_steps / tq,
        "packed_over_plain": tq / tp,
        "packed_p99_ms": 1e3 * float(np.percentile(latp, 99)), "plain_p99_ms": 1e3 * float(np.percentile(latq, 99)),
        "mean_candidates_per_iter": float(np.mean([c[0] for c in packed.counts])),
        "mean_collisions_per_iter": float(np.mean([c[3] for c in packed.counts])),
        # the first pass of both runs sees the same input: equal K, M, V, R over the batch
        "first_pass_counts_equal": list(map(int, packed.counts[0])) == list(map(int, plain.counts[0])),
        "first_pass_max_position_difference": diff, "first_pass_max_correction": moved,
    }


def serving_section(ccc, torch, scenes, args):
    """Resident folded cloths with lognormal sizes (median 20k triangles), Poisson request
    arrivals; each request is one CCD + impulse step of one simulation.  First the
    saturation throughput (every request queued at t = 0), then an open-loop run at
    args.serving_load of that rate: completed steps/s and p50 / p99 latency."""
    from cccollisions_amd import serving
    rng = np.random.default_rng(5)
    sizes = np.minimum(np.exp(rng.normal(np.log(20000), 0.8, args.serving_sims)), args.serving_max_tris)
    cl = [scenes.folded_cloth(scenes.grid_size_for_triangles(int(s)), seed=500 + i) for i, s in enumerate(sizes)]
    sims = [serving.Simulation(torch, sc.faces, sc.x0, sc.x1, sc.corner_masses()) for sc in cl]
    tris = [s.num_triangles for s in sims]

    def run(arrivals, sim_ids):
        server = serving.CollisionServer(ccc, torch, sims, slots=args.serving_slots, iters=args.iters,
                                         max_candidates=args.max_candidates)
        t0 = time.perf_counter()
        reqs = []
        for a, s in zip(arrivals, sim_ids):
            while True:
                left = a - (time.perf_counter() - t0)
                if left <= 0:
                    break
                time.sleep(min(0.002, left))
            reqs.append(server.submit(int(s), arrival=server.now()))
        server.close()
        span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        lat = np.array([r.latency for r in reqs])
        return {"requests": len(reqs), "steps_per_s": len(reqs) / span,
                "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                "mean_batch": float(np.mean([r.batch_size for r in reqs]))}

    # warm-up (allocator, code paths) with one request per simulation
    run(np.zeros(len(sims)), np.arange(len(sims)))
    n_sat = args.serving_requests
    sat = run(np.zeros(n_sat), rng.integers(0, len(sims), n_sat))
    rate = args.serving_load * sat["steps_per_s"]
    gaps = rng.exponential(1.0 / rate, n_sat)
    open_loop = run(np.cumsum(gaps) - gaps[0], rng.integers(0, len(sims), n_sat))
    return {"sims": len(sims), "triangles_median": int(np.median(tris)), "triangles_max": int(max(tris)),
            "triangles_total": int(sum(tris)), "slots": args.serving_slots, "impulse_iterations": args.iters,
            "saturation": sat, "offered_steps_per_s": rate, "open_loop": open_loop}


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


def roofline(kt, counts, steps, iters):
    """Roofline of the dominant kernel over the timed region."""
    name, (ms, launches) = max(kt.items(), key=lambda kv: kv[1][0])
    timed = counts[-steps * iters:]
    m_total = float(sum(c[1] for c in timed))
    k_total = float(sum(c[0] for c in timed))
    avg_s = ms / 1e3 / max(launches, 1)
    if name.startswith("ccd"):
        work = FLOPS_PER_SUBPRIM * m_total / max(launches, 1)
        achieved = work / avg_s / 1e12
        traffic, src = measured_traffic("k_ccd")
        return name, {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                      "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                      "kernel": name, "avg_launch_ms": avg_s * 1e3,
                      "work_per_launch": f"{FLOPS_PER_SUBPRIM} fp64 flop x {m_total / max(launches, 1):.4g} VF/EE pairs",
                      "traffic_source": src}
    # memory-bound kernels: unique bytes per launch
    per = {"traverse": 64.0, "subprim": 8.0}
    b = per.get(name, 0.0) * (k_total / max(launches, 1))
    achieved = b / avg_s / 1e9
    return name, {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                  "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": name, "avg_launch_ms": avg_s * 1e3}


def cpu_baseline(scene, stride, iters):
    from oracle import oracle as O
    t0, t1 = scene.soups()
    m = scene.corner_masses()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)), 16))
    O.impulse_pass(t0[:2000], t1[:2000], m[:2000], cores, 1)   # load / warm the library
    t = time.perf_counter()
    _, _, _, counts = O.impulse_pass(t0, t1, m, cores, stride)
    dt = time.perf_counter() - t
    pass_s = dt * stride
    return {"value": 1.0 / (iters * pass_s), "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": (f"one impulse pass (bvh_motion + compute_impulses) over every {stride}-th query leaf "
                       f"of the 130,050-triangle cloth: {dt:.2f} s, {counts['candidates']} candidate pairs, "
                       f"{counts['collisions']} collisions; scaled x{stride} per pass and x{iters} passes per step")}


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
    step = Step(ccc, torch, scene, args.iters, args.max_candidates)
    for _ in range(args.warmup):
        step.run()
    torch.cuda.synchronize()
    ccc.set_profiling(True)
    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        step.run()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    kt = ccc.kernel_times()
    ccc.set_profiling(False)

    if args.copy_calibration_mib:
        src = torch.empty(args.copy_calibration_mib << 20, dtype=torch.uint8, device="cuda:0").fill_(1)
        torch.cuda.synchronize()
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    name, roof = roofline(kt, step.counts, args.steps, args.iters)
    timed = step.counts[-args.steps * args.iters:]
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
                         f"{args.iters} impulse iterations per step (bvh_motion + compute_impulses + vertex update)"),
            "triangles": scene.num_triangles,
            "impulse_iterations": args.iters,
            "max_candidates_per_triangle": args.max_candidates,
            "mean_candidates_per_iter": float(np.mean([c[0] for c in timed])),
            "mean_subprims_per_iter": float(np.mean([c[1] for c in timed])),
            "mean_roots_per_iter": float(np.mean([c[2] for c in timed])),
            "mean_collisions_per_iter": float(np.mean([c[3] for c in timed])),
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in sorted(kt.items())},
    }
    if args.collect:
        out["collect"] = collect_section(ccc, torch, step, args)
    if args.garment:
        out["garment"] = garment_section(ccc, torch, scenes, args)
    if args.batched_sims:
        out["batched"] = batched_section(ccc, torch, scenes, args)
    if args.serving_sims:
        out["serving"] = serving_section(ccc, torch, scenes, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_stride, args.iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
