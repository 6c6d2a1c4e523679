"""A serving loop over the batch arena: scene-step requests arrive over
time, are admitted while device memory and scene slots allow, stay
resident on the device for all of their frames, and are stepped together
with every other resident scene, one launch per pass (continuous
batching). Finished scenes are retired at once and their blocks reused by
the next admission.

Two servers share one interface so their throughput and latency compare
directly on the same request stream:

* `BatchServer` -- one batch arena; each tick steps every resident scene by
  one frame in a single submission (a hipGraph replay; a changed mix updates
  the captured graph in place).
* `PlainServer` -- one `JGS2GpuSolver` per request, each stepped by its own
  submission: the one-scene-per-launch path the reference offers.

A tick launches the frame of the resident scenes and then admits the
requests that have arrived while the device runs it: admission's host work
(validation, coloring, packing, the contact capacity scan) overlaps the
frame, and its uploads queue behind the frame on the device. A request
admitted during a tick joins from the next frame.

Latency is reported per tick (one frame of every resident scene, as a
client sees it: launch to completion) and per request: arrival to last
frame, split into queueing (arrival to admission) and residency (admission
to last frame). Percentiles use the nearest-rank rule of the reference's
metrics (src/performance/metrics.ts:42-46).
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import _native as N
from .batch import BatchCapacity, JGS2Batch, footprint
from .solver import GpuInput, JGS2GpuSolver


def nearest_rank(values: Sequence[float], q: float) -> float:
    """Nearest-rank percentile (src/performance/metrics.ts:42-46)."""
    s = sorted(values)
    if not s:
        return 0.0
    rank = max(1, int(math.ceil(q / 100.0 * len(s))))
    return float(s[min(rank, len(s)) - 1])


@dataclass
class Request:
    """One scene-step request: run `scene` for `frames` frames."""

    id: int
    scene: GpuInput
    frames: int
    arrival: float  # seconds from the start of the run (or ticks, with a tick clock)
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[np.ndarray] = None  # final positions when results are kept
    free_vertices: int = 0


def poisson_arrivals(rate: float, duration: float, seed: int) -> np.ndarray:
    """Arrival times of a Poisson process of `rate` per second over
    [0, duration)."""
    rng = np.random.default_rng(seed)
    out = []
    t = rng.exponential(1.0 / rate)
    while t < duration:
        out.append(t)
        t += rng.exponential(1.0 / rate)
    return np.asarray(out)


def capacity_for_budget(template: Sequence[GpuInput], budget_bytes: int, max_scenes: int = 4096) -> BatchCapacity:
    """The largest arena, in multiples of the template mix's footprint (the
    per-scene average), whose device bytes fit `budget_bytes`."""
    fp = [footprint(x) for x in template]
    mean = {k: sum(f[k] for f in fp) / len(fp) for k in fp[0]}
    k_max = max(x.cubature_k for x in template)

    def cap(n: int) -> BatchCapacity:
        return BatchCapacity(max_vertices=int(mean["vertices"] * n) + 256, max_tets=int(mean["tets"] * n) + 256,
                             max_scenes=n, cubature_k=k_max, max_triangles=int(mean["triangles"] * n),
                             max_hinges=int(mean["hinges"] * n), max_candidates=int(mean["candidates"] * n),
                             max_bodies=int(mean["bodies"] * n) + 1,
                             max_surface_triangles=int(mean["surface_triangles"] * n),
                             max_surface_edges=int(mean["surface_edges"] * n))

    def fits(n: int) -> bool:
        try:
            return cap(n).arena_bytes() <= budget_bytes
        except N.JGS2RangeError:  # beyond the arena's 32-bit indexing
            return False

    lo, hi = 0, max_scenes
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if fits(mid):
            lo = mid
        else:
            hi = mid - 1
    if lo < 1:
        raise ValueError(f"An HBM budget of {budget_bytes} bytes holds no scene of this mix.")
    return cap(lo)


class TickClock:
    """A deterministic clock for tests: time advances by one unit per served
    tick, and jumps ahead to the next arrival when the server is idle."""

    def __init__(self):
        self.t = 0.0

    def __call__(self) -> float:
        return self.t

    def step(self) -> None:
        self.t += 1.0

    def advance_to(self, t: float) -> None:
        self.t = max(self.t, t)


@dataclass
class ServeReport:
    ticks: int = 0
    wall_s: float = 0.0
    vertex_updates: int = 0
    tick_ms: List[float] = field(default_factory=list)
    completed: List[Request] = field(default_factory=list)
    rejected_admissions: int = 0
    admissions: int = 0
    # Wall time inside admission calls: host work that overlaps the frame in
    # flight, then the wait for that frame (an admission's uploads and
    # feasibility preflight queue behind it).
    admission_s: float = 0.0
    extra: dict = field(default_factory=dict)

    def summary(self) -> dict:
        lat = [r.finished - r.arrival for r in self.completed]
        wait = [r.admitted - r.arrival for r in self.completed]
        resident = [r.finished - r.admitted for r in self.completed]
        return {
            "ticks": self.ticks,
            "wall_s": self.wall_s,
            "vertex_updates": self.vertex_updates,
            "vertex_updates_per_s": self.vertex_updates / self.wall_s if self.wall_s > 0 else 0.0,
            "frame_ms": {"p50": nearest_rank(self.tick_ms, 50), "p99": nearest_rank(self.tick_ms, 99)},
            "requests_completed": len(self.completed),
            "request_latency_s": {"p50": nearest_rank(lat, 50), "p99": nearest_rank(lat, 99)},
            "queueing_s": {"p50": nearest_rank(wait, 50), "p99": nearest_rank(wait, 99)},
            "resident_s": {"p50": nearest_rank(resident, 50), "p99": nearest_rank(resident, 99)},
            "admissions": self.admissions,
            "admission_wall_s": self.admission_s,
            "admissions_deferred": self.rejected_admissions,
            **self.extra,
        }


class _Server:
    """Shared request loop. A tick admits what has arrived and fits, steps
    every resident scene by one frame, and retires finished scenes."""

    def __init__(self, settings: dict, iterations: int, keep_results: bool = False,
                 clock: Optional[Callable[[], float]] = None):
        self.settings = dict(settings)
        self.iterations = int(iterations)
        self.keep_results = keep_results
        self.clock = clock or time.perf_counter
        self.queue: deque = deque()
        self.active: list = []  # [request, handle, frames_left]
        self.report = ServeReport()
        self._t0 = 0.0

    def now(self) -> float:
        return self.clock() - self._t0

    # Subclasses: admit one request (None if it does not fit now), launch one
    # frame of every active handle, wait for it, release one handle. A
    # subclass whose steps complete synchronously implements `_step_all`.
    def _try_admit(self, request: Request):
        raise NotImplementedError

    def _step_all(self) -> None:
        raise NotImplementedError

    def _launch_all(self) -> None:
        self._step_all()

    def _wait_all(self) -> None:
        pass

    def _release(self, handle) -> None:
        raise NotImplementedError

    def _admit_arrivals(self) -> None:
        t = time.perf_counter()
        while self.queue:
            request = self.queue[0]
            handle = self._try_admit(request)
            if handle is None:
                self.report.rejected_admissions += 1
                break  # first come, first served: the head waits for room
            self.queue.popleft()
            request.admitted = self.now()
            request.free_vertices = handle.free_vertex_count
            self.active.append([request, handle, request.frames])
            self.report.admissions += 1
        self.report.admission_s += time.perf_counter() - t

    def tick(self) -> bool:
        if not self.active:  # an idle device: nothing to overlap admission with
            self._admit_arrivals()
            if not self.active:
                return False
        stepped = list(self.active)
        t = self.clock()
        self._launch_all()
        self._admit_arrivals()  # overlaps the frame in flight; joins from the next frame
        self._wait_all()
        if isinstance(self.clock, TickClock):
            self.clock.step()
        self.report.tick_ms.append((self.clock() - t) * 1e3)
        self.report.ticks += 1
        self.report.vertex_updates += self.iterations * sum(r.free_vertices for r, _, _ in stepped)
        joined = self.active[len(stepped):]
        still = []
        for entry in stepped:
            entry[2] -= 1
            request, handle, left = entry
            if left == 0:
                if self.keep_results:
                    request.result = handle.read_positions()
                self._release(handle)
                request.finished = self.now()
                self.report.completed.append(request)
            else:
                still.append(entry)
        self.active = still + joined
        return True

    def run(self, requests: Sequence[Request], duration: Optional[float] = None,
            progress: Optional[Callable[[ServeReport, float], None]] = None, every: float = 20.0) -> ServeReport:
        """Serve `requests` (sorted by arrival) until all are done or
        `duration` seconds pass; requests still queued then are dropped
        from the report. `progress(report, now)` is called about every
        `every` seconds of the run."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self._t0 = self.clock()
        next_report = every
        while pending or self.queue or self.active:
            now = self.now()
            if duration is not None and now >= duration:
                break
            if progress is not None and now >= next_report:
                progress(self.report, now)
                next_report = now + every
            while pending and pending[0].arrival <= now:
                self.queue.append(pending.popleft())
            if self.tick():
                continue
            if self.queue:  # nothing resident and the head still does not fit
                raise ValueError(f"Request {self.queue[0].id} does not fit an empty server.")
            if pending:  # idle until the next arrival
                if isinstance(self.clock, TickClock):
                    self.clock.advance_to(self._t0 + pending[0].arrival)
                else:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - self.now())))
        self.report.wall_s = self.now()
        return self.report

    def close(self) -> None:
        for _, handle, _ in self.active:
            self._release(handle)
        self.active = []


class BatchServer(_Server):
    def __init__(self, capacity: BatchCapacity, settings: dict, iterations: int, hbm_budget_bytes: Optional[int] = None,
                 use_graphs: bool = True, keep_results: bool = False, clock=None):
        super().__init__(settings, iterations, keep_results, clock)
        arena = capacity.arena_bytes()
        if hbm_budget_bytes is not None and arena > hbm_budget_bytes:
            raise ValueError(f"The arena needs {arena} bytes, over the HBM budget of {hbm_budget_bytes}.")
        self.arena_bytes = arena
        self.batch = JGS2Batch.create(capacity, self.settings)
        self.batch.use_graphs(use_graphs)

    def _try_admit(self, request: Request):
        try:
            return self.batch.admit(request.scene)
        except N.JGS2MemoryError:
            return None

    def _launch_all(self) -> None:
        self.batch.step_frames_exact_iterations(1, self.iterations, self.settings)

    def _wait_all(self) -> None:
        self.batch.sync()

    def _release(self, handle) -> None:
        self.batch.retire(handle)

    def run(self, requests, duration=None, progress=None, every=20.0) -> ServeReport:
        report = super().run(requests, duration, progress, every)
        report.extra["graphs"] = self.batch.graph_stats()
        return report

    def close(self) -> None:
        super().close()
        self.batch.destroy()


class PlainServer(_Server):
    """One solver per request, stepped one scene per submission; admission
    by a running total of the scenes' arena bytes against the budget."""

    def __init__(self, settings: dict, iterations: int, hbm_budget_bytes: int, max_scenes: int,
                 keep_results: bool = False, clock=None):
        super().__init__(settings, iterations, keep_results, clock)
        self.budget = int(hbm_budget_bytes)
        self.max_scenes = int(max_scenes)
        self.used = 0
        self._bytes = {}

    def _try_admit(self, request: Request):
        f = footprint(request.scene)
        need = BatchCapacity(max_vertices=f["vertices"], max_tets=f["tets"], max_scenes=1,
                             cubature_k=request.scene.cubature_k, max_triangles=f["triangles"],
                             max_hinges=f["hinges"], max_candidates=f["candidates"],
                             max_bodies=f["bodies"], max_surface_triangles=f["surface_triangles"],
                             max_surface_edges=f["surface_edges"]).arena_bytes()
        if len(self.active) >= self.max_scenes or self.used + need > self.budget:
            return None
        solver = JGS2GpuSolver.create(request.scene, self.settings)
        self.used += need
        self._bytes[id(solver)] = need
        return solver

    def _launch_all(self) -> None:
        self._stepped = [solver for _, solver, _ in self.active]
        for solver in self._stepped:
            solver.step_exact_iterations(self.iterations, self.settings)

    def _wait_all(self) -> None:
        for solver in self._stepped:
            solver.await_idle()

    def _release(self, handle) -> None:
        self.used -= self._bytes.pop(id(handle), 0)
        handle.destroy()
