"""Continuous-batching request loop for CCD + impulse steps of many resident simulations.

Requests name a resident simulation; each asks for one CCD + impulse step (`iters` impulse
passes: bvh_motion + compute_impulses + vertex update) on that simulation's current
positions.  Worker slots (one host thread and one HIP stream each) take batches of pending
requests as they free up: requests are admitted in arrival order while the batch's
estimated HBM footprint stays within the slot's budget and no other slot is working on the
same simulation.  A batch runs through the packed entry point (one BVH / traversal /
narrow phase over all its meshes), so requests that arrive while other batches are on the
GPU join the next free slot and their work overlaps on the device with the running ones.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np


@dataclasses.dataclass
class Request:
    sim: int
    arrival: float                  # seconds, on the server's clock
    start: Optional[float] = None
    done: Optional[float] = None
    result: object = None
    batch_size: int = 0

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def footprint_bytes(num_triangles: int, max_candidates: int, itemsize: int = 8) -> int:
    """HBM a one-mesh step needs beyond the resident state: the two soups, the BVH and its
    traversal nodes and leaf records, the candidate list at capacity, one chunk of
    sub-primitive items, the narrow phase's vertex records and the impulse outputs."""
    n = num_triangles
    soups = 2 * n * 9 * itemsize
    bvh = (2 * n - 1) * (6 * itemsize + 4 * 8) + (n - 1) * 64 + n * 32
    pairs = max_candidates * n * 8
    items = 15 * min(max_candidates * n, 1 << 25) * 8
    outs = n * (18 * itemsize + 12)
    records = n * 18 * itemsize        # per-corner [t0 | t1] vertex records of the narrow phase
    return soups + bvh + pairs + items + outs + records


class Scheduler:
    """Admission only (no GPU): FIFO with HBM budget and per-simulation exclusivity.
    Thread-safe; `take` blocks until a batch can be formed or the scheduler is closed."""

    def __init__(self, footprints: Dict[int, int], budget: int):
        self.footprints = footprints
        self.budget = budget
        self.pending: collections.deque = collections.deque()
        self.busy: set = set()
        self.cv = threading.Condition()
        self.closed = False

    def submit(self, req: Request) -> None:
        with self.cv:
            self.pending.append(req)
            self.cv.notify_all()

    def close(self) -> None:
        with self.cv:
            self.closed = True
            self.cv.notify_all()

    def _form(self) -> List[Request]:
        batch, used, seen = [], 0, set()
        for req in list(self.pending):
            if req.sim in self.busy or req.sim in seen:
                continue
            f = self.footprints[req.sim]
            if batch and used + f > self.budget:
                continue
            batch.append(req)
            seen.add(req.sim)
            used += f
        for req in batch:
            self.pending.remove(req)
            self.busy.add(req.sim)
        return batch

    def take(self) -> List[Request]:
        with self.cv:
            while True:
                batch = self._form()
                if batch:
                    return batch
                if self.closed and not self.pending:
                    return []
                self.cv.wait(0.05)

    def release(self, batch: Sequence[Request]) -> None:
        with self.cv:
            for req in batch:
                self.busy.discard(req.sim)
            self.cv.notify_all()


class Simulation:
    """Resident state of one simulation on the device: faces, positions at the start of
    the step (x0), predicted end-of-step positions (x1), per-corner masses, and t0 = x0
    gathered per triangle (fixed for the step)."""

    def __init__(self, torch, faces: np.ndarray, x0: np.ndarray, x1: np.ndarray, corner_masses: np.ndarray,
                 device="cuda:0"):
        dev = torch.device(device)
        self.faces = torch.from_numpy(np.ascontiguousarray(faces)).to(dev)
        self.x0 = torch.from_numpy(np.ascontiguousarray(x0)).to(dev)
        self.x1 = torch.from_numpy(np.ascontiguousarray(x1)).to(dev)
        self.masses = torch.from_numpy(np.ascontiguousarray(corner_masses)).to(dev)
        self.t0 = self.x0[self.faces].contiguous()
        self.num_triangles = int(faces.shape[0])
        self.num_vertices = int(x0.shape[0])


def impulse_step(ccc, torch, sims: Sequence[Simulation], iters: int, max_candidates: int):
    """One CCD + impulse step of every simulation in `sims`, packed into one pass of
    launches per impulse iteration.  Returns the corrected end-of-step positions per
    simulation (new tensors; the resident state is not modified)."""
    toff = np.cumsum([0] + [s.num_triangles for s in sims]).tolist()
    voff = np.cumsum([0] + [s.num_vertices for s in sims]).tolist()
    if len(sims) == 1:
        faces, t0, masses, x1 = sims[0].faces, sims[0].t0, sims[0].masses, sims[0].x1
    else:
        faces = torch.cat([s.faces + voff[i] for i, s in enumerate(sims)])
        t0 = torch.cat([s.t0 for s in sims])
        masses = torch.cat([s.masses for s in sims])
        x1 = torch.cat([s.x1 for s in sims])
    flat = faces.reshape(-1)
    nv = x1.shape[0]
    for _ in range(iters):
        t1 = x1[faces].contiguous()
        _, dx, cnt = ccc.compute_impulses_packed(t0, t1, masses, toff, max_candidates)
        d = dx.reshape(-1, 3)
        ok = torch.isfinite(d).all(dim=1)
        d = torch.where(ok[:, None], d, torch.zeros_like(d))
        k = torch.where(ok, cnt.reshape(-1).to(x1.dtype), torch.zeros_like(ok, dtype=x1.dtype))
        s = torch.zeros((nv, 3), dtype=x1.dtype, device=x1.device).index_add_(0, flat, d)
        c = torch.zeros((nv,), dtype=x1.dtype, device=x1.device).index_add_(0, flat, k)
        x1 = x1 + torch.where(c[:, None] > 0, s / c.clamp(min=1)[:, None], torch.zeros_like(s))
    return [x1[voff[i]:voff[i + 1]] for i in range(len(sims))]


class CollisionServer:
    """Worker slots over a Scheduler.  Each slot: one host thread, one HIP stream."""

    def __init__(self, ccc, torch, sims: Sequence[Simulation], slots: int = 4, hbm_budget: Optional[int] = None,
                 iters: int = 10, max_candidates: int = 512):
        self.ccc, self.torch, self.sims = ccc, torch, list(sims)
        self.iters, self.max_candidates = iters, max_candidates
        fp = {i: footprint_bytes(s.num_triangles, max_candidates) for i, s in enumerate(self.sims)}
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.6 * free)
        self.slot_budget = max(hbm_budget // slots, 1)
        self.scheduler = Scheduler(fp, self.slot_budget)
        self.clock0 = time.perf_counter()
        self.completed: List[Request] = []
        self.lock = threading.Lock()
        self.streams = [torch.cuda.Stream() for _ in range(slots)]
        self.threads = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(slots)]
        self.errors: List[BaseException] = []
        for t in self.threads:
            t.start()

    def now(self) -> float:
        return time.perf_counter() - self.clock0

    def submit(self, sim: int, arrival: Optional[float] = None) -> Request:
        req = Request(sim, self.now() if arrival is None else arrival)
        self.scheduler.submit(req)
        return req

    def _worker(self, slot: int) -> None:
        torch = self.torch
        stream = self.streams[slot]
        while True:
            batch = self.scheduler.take()
            if not batch:
                return
            t = self.now()
            for r in batch:
                r.start = t
                r.batch_size = len(batch)
            try:
                with torch.cuda.stream(stream):
                    out = impulse_step(self.ccc, torch, [self.sims[r.sim] for r in batch], self.iters,
                                       self.max_candidates)
                stream.synchronize()
            except BaseException as e:       # reported by close()
                self.errors.append(e)
                out = [None] * len(batch)
            t = self.now()
            for r, o in zip(batch, out):
                r.result = o
                r.done = t
            with self.lock:
                self.completed.extend(batch)
            self.scheduler.release(batch)

    def close(self) -> None:
        self.scheduler.close()
        for t in self.threads:
            t.join()
        if self.errors:
            raise self.errors[0]
