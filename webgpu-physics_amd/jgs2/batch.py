"""Many scenes per launch: the batch arena of libjgs2hip.

The reference steps one scene per `JGS2GpuSolver` object. A `JGS2Batch`
holds many independent scenes of uneven size in one device arena; every
pass of a frame is one launch over all resident scenes, with reductions,
line-search control, history and the convergence latch kept per scene, so
each scene's results are bit-identical to stepping it alone. Scenes are
admitted and retired between submissions (continuous batching); an
admission that does not fit raises `JGS2MemoryError` and leaves the batch
unchanged.
"""

from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass
from typing import Mapping, Optional

from . import _native as N
from .solver import GpuInput, JGS2GpuSolver, settings_struct


@dataclass
class BatchCapacity:
    """Arena capacities; scene ranges are rounded up to 256 entries."""

    max_vertices: int
    max_tets: int
    max_scenes: int
    cubature_k: int
    max_triangles: int = 0
    max_hinges: int = 0
    max_candidates: int = 0
    max_bodies: int = 1
    corotated: bool = False
    max_surface_triangles: int = 0
    max_surface_edges: int = 0

    def struct(self) -> N.BatchConfig:
        c = N.BatchConfig()
        c.max_vertices = int(self.max_vertices)
        c.max_tets = int(self.max_tets)
        c.max_triangles = int(self.max_triangles)
        c.max_hinges = int(self.max_hinges)
        c.max_candidates = int(self.max_candidates)
        c.max_bodies = int(self.max_bodies)
        c.max_scenes = int(self.max_scenes)
        c.cubature_k = int(self.cubature_k)
        c.corotated = 1 if self.corotated else 0
        c.max_surface_triangles = int(self.max_surface_triangles)
        c.max_surface_edges = int(self.max_surface_edges)
        return c

    def arena_bytes(self) -> int:
        """Device bytes a batch of this capacity allocates."""
        out = C.c_uint64()
        cfg = self.struct()
        N.check(N.lib.jgs2_batch_arena_bytes(C.byref(cfg), C.byref(out)))
        return int(out.value)


def blocks(n: int) -> int:
    return (int(n) + 255) // 256


def footprint(input_: GpuInput) -> dict:
    """Arena units one scene occupies (vertex/tet/triangle/candidate
    ranges in 256-entry blocks, the rest in entries)."""
    V, T = input_.vertex_count, input_.tet_count
    NT = input_.cloth.triangle_count if input_.cloth is not None else 0
    NH = input_.cloth.hinge_count if input_.cloth is not None else 0
    NC = input_.contact.candidate_count if input_.contact is not None else 0
    NS, NE = input_.contact.canonical_counts() if hasattr(input_.contact, "canonical_counts") else (0, 0)
    bodies = int(input_.vertex_info.reshape(-1, 4)[:, 3].max()) + 1 if V else 1
    return {
        "vertices": 256 * blocks(V), "tets": 256 * blocks(T), "triangles": 256 * blocks(NT), "hinges": NH,
        "candidates": 256 * blocks(NC), "bodies": bodies, "surface_triangles": 256 * blocks(NS),
        "surface_edges": 256 * blocks(NE),
    }


class BatchScene(JGS2GpuSolver):
    """A scene resident in a batch: reads, objectives and diagnostics of
    `JGS2GpuSolver`; stepping goes through the batch."""

    def __init__(self, handle, input_, batch: "JGS2Batch"):
        super().__init__(handle, input_)
        self._batch = weakref.ref(batch)

    def _no_step(self, *args, **kwargs):
        raise N.JGS2StateError(N.JGS2_ERROR_STATE, "A batch-resident scene is stepped through its batch.")

    step = step_frames = step_exact_iterations = step_frames_exact_iterations = _no_step

    def destroy(self) -> None:
        """Retires the scene from its batch."""
        if self._handle:
            batch = self._batch()
            if batch is not None and batch._handle:
                batch.retire(self)
            self._handle = None


class JGS2Batch:
    def __init__(self, handle, owned: bool = True):
        self._handle = handle
        self._owned = owned
        self._scenes: "weakref.WeakSet[BatchScene]" = weakref.WeakSet()

    @classmethod
    def create(cls, capacity: BatchCapacity, settings: Optional[Mapping] = None) -> "JGS2Batch":
        s, mask = settings_struct(settings)
        cfg = capacity.struct()
        handle = C.c_void_p()
        N.check(N.lib.jgs2_batch_create(C.byref(cfg), C.byref(s), mask, C.byref(handle)))
        return cls(handle)

    @classmethod
    def of(cls, solver: JGS2GpuSolver) -> "JGS2Batch":
        """The private one-scene batch behind a `JGS2GpuSolver`, for the
        launch-shape, graph and autotune controls (not owned)."""
        handle = C.c_void_p()
        N.check(N.lib.jgs2_solver_batch(solver._h(), C.byref(handle)))
        return cls(handle, owned=False)

    def _h(self):
        if not self._handle:
            raise N.JGS2StateError(N.JGS2_ERROR_STATE, "The batch has been destroyed.")
        return self._handle

    def destroy(self) -> None:
        if self._handle and self._owned:
            for scene in list(self._scenes):
                scene._handle = None
            N.check(N.lib.jgs2_batch_destroy(self._handle))
        self._handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # -------------------------------------------------------------- residency
    def admit(self, input_: GpuInput) -> BatchScene:
        st = input_.struct()
        handle = C.c_void_p()
        N.check(N.lib.jgs2_batch_admit(self._h(), C.byref(st), C.byref(handle)))
        scene = BatchScene(handle, input_, self)
        self._scenes.add(scene)
        return scene

    def retire(self, scene: BatchScene) -> None:
        if scene._handle:
            N.check(N.lib.jgs2_batch_retire(self._h(), scene._handle))
            scene._handle = None
        self._scenes.discard(scene)

    # -------------------------------------------------------------- stepping
    def step_frames(self, frame_count: int, overrides: Optional[Mapping] = None) -> None:
        s, mask = settings_struct(overrides)
        N.check(N.lib.jgs2_batch_step_frames(self._h(), int(frame_count), C.byref(s), mask))

    def step_frames_exact_iterations(self, frame_count: int, iterations: int,
                                     overrides: Optional[Mapping] = None) -> None:
        s, mask = settings_struct(overrides)
        N.check(N.lib.jgs2_batch_step_frames_exact_iterations(self._h(), int(frame_count), int(iterations),
                                                              C.byref(s), mask))

    def sync(self) -> None:
        N.check(N.lib.jgs2_batch_sync(self._h()))

    # -------------------------------------------------------------- controls
    def info(self) -> dict:
        i = N.BatchInfo()
        N.check(N.lib.jgs2_batch_get_info(self._h(), C.byref(i)))
        return {name: getattr(i, name) for name, _ in N.BatchInfo._fields_}

    def use_graphs(self, enable: bool = True) -> None:
        N.check(N.lib.jgs2_batch_use_graphs(self._h(), 1 if enable else 0))

    def graph_stats(self) -> dict:
        """Frame graphs instantiated from scratch and captures applied to the
        existing exec in place."""
        inst, upd = C.c_uint64(), C.c_uint64()
        N.check(N.lib.jgs2_batch_graph_stats(self._h(), C.byref(inst), C.byref(upd)))
        return {"instantiations": int(inst.value), "updates": int(upd.value)}

    def set_launch(self, solve_block: int, xcd_remap: bool, solve_lanes: Optional[int] = None) -> None:
        lanes = self.info()["solve_lanes"] if solve_lanes is None else int(solve_lanes)
        N.check(N.lib.jgs2_batch_set_launch(self._h(), int(solve_block), 1 if xcd_remap else 0, lanes))

    def autotune(self, overrides: Optional[Mapping] = None, repeats: int = 20) -> float:
        """Keeps the fastest local-solve launch shape; returns its ms."""
        s, mask = settings_struct(overrides)
        best = C.c_double()
        N.check(N.lib.jgs2_batch_autotune(self._h(), C.byref(s), mask, int(repeats), C.byref(best)))
        return best.value

    def profile_begin(self) -> None:
        N.check(N.lib.jgs2_batch_profile_begin(self._h()))

    def profile_end(self):
        """(local-solve ms per launch, launches, total ms of the window)."""
        ms = C.c_double()
        n = C.c_uint64()
        total = C.c_double()
        N.check(N.lib.jgs2_batch_profile_end(self._h(), C.byref(ms), C.byref(n), C.byref(total)))
        return ms.value, n.value, total.value
