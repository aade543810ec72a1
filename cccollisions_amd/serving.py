"""Continuous-batching request loop for CCD + impulse steps of many resident simulations.

Every simulation lives on the device as a `cccollisions_amd.Sim`: its positions, faces,
masses, triangle soups, BVH and candidate-pair list stay resident between requests.  A
request asks for one step of one simulation -- `iters` impulse passes (bvh_motion +
compute_impulses + vertex update) replayed as one HIP graph -- after which the simulation
advances (x0 <- the corrected positions, x1 <- a constant-velocity prediction), so that the
next request continues from there.  Worker slots (one host thread, one HIP stream and one
device workspace each) take batches of pending requests as they free up: requests are
admitted in arrival order while the batch's working set (the steps' workspace bytes) stays
within the slot's workspace and no other slot holds a request of the same simulation.
Requests that arrive while other slots' steps run are admitted as soon as a slot frees up,
and the slots' work overlaps on the device.  The server checks at start that the resident
simulations plus the slots' workspaces fit its HBM budget.  With deterministic impulse sums
(`Sim(..., deterministic=True)`, the default here) a simulation's states are the same bits
whether its requests ran alone or beside others.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import Dict, List, Optional, Sequence

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
    """Estimated device bytes of one resident simulation of a folded-cloth-like mesh
    (V ~ N / 2 vertices) with `max_candidates` candidate pairs per triangle, plus the workspace
    of one of its steps (tree preparation, vertex records and one chunk of sub-primitive
    items); an estimate for planning before the simulations exist -- a created `Sim` reports
    its own `footprint()` and `workspace_bytes()`."""
    n = num_triangles
    resident = (3 * (n // 2 + 1) * 3 + 4 * 9 * n + 3 * n) * itemsize + (2 * n - 1) * (6 * itemsize + 32) + 36 * n
    pairs = max_candidates * n * 8
    workspace = n * (300 + 18 * itemsize) + 15 * min(max_candidates * n, 1 << 26) * 4
    return resident + pairs + workspace


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
    """One resident simulation: a `cccollisions_amd.Sim` over its faces, start-of-step
    positions x0, predicted end-of-step positions x1 and per-corner masses."""

    def __init__(self, ccc, torch, faces: np.ndarray, x0: np.ndarray, x1: np.ndarray, corner_masses: np.ndarray,
                 iters: int = 10, max_candidates: int = 512, deterministic: bool = True, device="cuda:0"):
        dev = torch.device(device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.sim = ccc.Sim(t(x0), t(x1), t(faces), t(corner_masses), iters=iters, max_candidates=max_candidates,
                           deterministic=deterministic)
        self.num_triangles = int(faces.shape[0])
        self.num_vertices = int(x0.shape[0])
        self.steps = 0

    def footprint(self) -> int:
        return int(self.sim.footprint())

    def workspace_bytes(self) -> int:
        return int(self.sim.workspace_bytes())

    def step(self, workspace=None):
        """One step on the current stream, advancing the state."""
        self.sim.step(advance=True, workspace=workspace)
        self.steps += 1

    def positions(self):
        return self.sim.positions()


class CollisionServer:
    """Worker slots over a Scheduler.  Each slot: one host thread, one HIP stream, one device
    workspace sized for the largest simulation's step (grown if a simulation's candidate list
    grows past it)."""

    def __init__(self, ccc, torch, sims: Sequence[Simulation], slots: int = 4, hbm_budget: Optional[int] = None,
                 keep_results: bool = False, autotune: bool = False):
        self.ccc, self.torch, self.sims = ccc, torch, list(sims)
        self.keep_results = keep_results
        ws = max(s.workspace_bytes() for s in self.sims)
        resident = sum(s.footprint() for s in self.sims)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.9 * free) + resident
        if resident + slots * ws > hbm_budget:
            raise ValueError(f"the simulations ({resident} bytes) and {slots} slot workspaces ({ws} bytes each) do not fit "
                             f"the HBM budget of {hbm_budget} bytes")
        self.workspaces = [torch.empty(ws, dtype=torch.uint8, device="cuda:0") for _ in range(slots)]
        self.streams = [torch.cuda.Stream() for _ in range(slots)]
        if autotune:
            # one timing per mesh-size bucket; the other simulations of a bucket reuse its choice
            with torch.cuda.stream(self.streams[0]):
                for s in self.sims:
                    s.sim.autotune(workspace=self.workspaces[0])
            self.streams[0].synchronize()
        self.scheduler = Scheduler({i: s.workspace_bytes() for i, s in enumerate(self.sims)}, ws)
        self.clock0 = time.perf_counter()
        self.completed: List[Request] = []
        self.lock = threading.Lock()
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
                    for r in batch:
                        sim = self.sims[r.sim]
                        need = sim.workspace_bytes()
                        if need > self.workspaces[slot].numel():
                            self.workspaces[slot] = torch.empty(need, dtype=torch.uint8, device="cuda:0")
                        sim.step(workspace=self.workspaces[slot])
                        if self.keep_results:
                            r.result = sim.positions()[0]
                        r.done = self.now()
                stream.synchronize()
            except BaseException as e:       # reported by close()
                self.errors.append(e)
                for r in batch:
                    r.done = self.now() if r.done is None else r.done
            with self.lock:
                self.completed.extend(batch)
            self.scheduler.release(batch)

    def close(self) -> None:
        self.scheduler.close()
        for t in self.threads:
            t.join()
        if self.errors:
            raise self.errors[0]
