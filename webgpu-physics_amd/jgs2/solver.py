"""Python host mirror of the reference's `JGS2GpuSolver`
(src/simulation/gpu/jgs2-solver.ts:905-3051) over libjgs2hip.

Method names follow the reference in snake_case; arguments keep their
meaning and the same validation errors are raised (RangeError maps to
JGS2RangeError, a ValueError). Every computing method runs on the GPU
through the C ABI.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Mapping, Optional

import numpy as np

from . import _native as N

# DEFAULT_JGS2_STEP_SETTINGS (jgs2-solver.ts:164-187) in snake_case.
DEFAULT_STEP_SETTINGS = {
    "schedule": "jacobi",
    "timestep": 1.0 / 60.0,
    "gravity": (0.0, -9.81, 0.0),
    "iterations": 9,
    "floor_height": 0.0,
    "floor_stiffness": 20000.0,
    "velocity_damping": 0.999,
    "regularization": 1e-6,
    "rotation_epsilon": 1e-7,
    "max_step": 0.1,
    "residual_tolerance": 1e-3,
    "normalized_update_tolerance": 1e-3,
    "contact_tangential_damping": 8.0,
    "contact_margin": 0.01,
    "ipc_activation_distance": 0.05,
    "ipc_minimum_distance": 0.001,
    "ipc_barrier_stiffness": 0.0,
    "ipc_friction_coefficient": 0.0,
    "ipc_friction_velocity_epsilon": 0.01,
    "ipc_step_safety": 0.9,
    "horizontal_body_correction": True,
    "parity_mode": False,
}

SCHEDULES = {"jacobi": N.SCHEDULE_JACOBI, "graph-colored-gauss-seidel": N.SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL}


def settings_struct(overrides: Optional[Mapping] = None):
    """Pack a Partial<JGS2StepSettings>-like mapping into (Settings, mask)."""
    s = N.Settings()
    N.lib.jgs2_default_settings(C.byref(s))
    mask = 0
    for key, value in (overrides or {}).items():
        if key not in N.SETTINGS_BITS:
            raise TypeError(f"Unknown JGS2 step setting: {key}")
        mask |= N.SETTINGS_BITS[key]
        if key == "schedule":
            if value not in SCHEDULES:
                raise N.JGS2RangeError(N.JGS2_ERROR_RANGE, f"Unknown JGS2 schedule: {value}.")
            s.schedule = SCHEDULES[value]
        elif key == "gravity":
            for i in range(3):
                s.gravity[i] = float(value[i])
        elif key in ("horizontal_body_correction", "parity_mode"):
            setattr(s, key, 1 if value else 0)
        elif key == "iterations":
            s.iterations = int(value)
        else:
            setattr(s, key, float(value))
    return s, mask


@dataclass
class ClothData:
    """Arrays of `JGS2GpuClothInput` (cloth-layout.ts:89-104)."""

    triangle_indices: np.ndarray
    triangle_inverse_rest_bases: np.ndarray
    triangle_rest_areas: np.ndarray
    hinge_indices: np.ndarray
    hinge_rest_angles: np.ndarray
    hinge_rest_edge_lengths: np.ndarray
    young_modulus: float
    poisson_ratio: float
    thickness: float
    bending_stiffness: float

    def __post_init__(self):
        for name in ("triangle_indices", "hinge_indices"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=np.uint32).reshape(-1))
        for name in ("triangle_inverse_rest_bases", "triangle_rest_areas", "hinge_rest_angles",
                     "hinge_rest_edge_lengths"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=np.float32).reshape(-1))

    @property
    def triangle_count(self) -> int:
        return self.triangle_indices.size // 3

    @property
    def hinge_count(self) -> int:
        return self.hinge_indices.size // 4

    def struct(self) -> N.ClothInput:
        c = N.ClothInput()
        c.triangle_count = self.triangle_count
        c.triangle_indices = N.uptr(self.triangle_indices)
        c.triangle_inverse_rest_bases = N.fptr(self.triangle_inverse_rest_bases)
        c.triangle_rest_areas = N.fptr(self.triangle_rest_areas)
        c.hinge_count = self.hinge_count
        c.hinge_indices = N.uptr(self.hinge_indices)
        c.hinge_rest_angles = N.fptr(self.hinge_rest_angles)
        c.hinge_rest_edge_lengths = N.fptr(self.hinge_rest_edge_lengths)
        c.young_modulus = float(self.young_modulus)
        c.poisson_ratio = float(self.poisson_ratio)
        c.thickness = float(self.thickness)
        c.bending_stiffness = float(self.bending_stiffness)
        return c


@dataclass
class ContactData:
    """`StaticIpcContactCandidates` (src/simulation/cpu/ipc-contact.ts:20-38):
    VT records first, then EE, four vertex ids each."""

    packed_indices: np.ndarray
    vertex_triangle_count: int
    edge_edge_count: int

    def __post_init__(self):
        self.packed_indices = np.ascontiguousarray(self.packed_indices, dtype=np.uint32).reshape(-1)

    @property
    def candidate_count(self) -> int:
        return self.vertex_triangle_count + self.edge_edge_count

    def struct(self) -> N.ContactInput:
        c = N.ContactInput()
        c.vertex_triangle_count = int(self.vertex_triangle_count)
        c.edge_edge_count = int(self.edge_edge_count)
        c.packed_indices = N.uptr(self.packed_indices)
        return c


@dataclass
class BroadPhaseContact:
    """IPC contact whose candidates are the static superset's pairs within
    AABB reach, built and rebuilt on the device (no counterpart in the
    reference, which enumerates the whole superset on the host,
    src/simulation/cpu/ipc-contact.ts:600-665). Stepping gives the
    superset's results bit for bit; Jacobi schedule only.
    `candidate_capacity` 0 lets the library choose (twice the count at the
    initial pose plus 1024); `reach_margin` 0 is half the
    activation reach."""

    surface_triangles: np.ndarray
    surface_edges: np.ndarray
    candidate_capacity: int = 0
    reach_margin: float = 0.0
    vertex_triangle_count: int = 0
    edge_edge_count: int = 0

    def __post_init__(self):
        self.surface_triangles = np.ascontiguousarray(self.surface_triangles, dtype=np.uint32).reshape(-1)
        self.surface_edges = np.ascontiguousarray(self.surface_edges, dtype=np.uint32).reshape(-1)

    @property
    def candidate_count(self) -> int:
        """Candidate slots the scene occupies in an arena."""
        return int(self.candidate_capacity)

    @property
    def surface_triangle_count(self) -> int:
        return self.surface_triangles.size // 3

    @property
    def surface_edge_count(self) -> int:
        return self.surface_edges.size // 2

    def canonical_counts(self):
        """(unique triangles, unique edges including triangle edges): the
        surface ranges the scene occupies in an arena."""
        t = np.unique(np.sort(self.surface_triangles.reshape(-1, 3), axis=1), axis=0)
        e = np.concatenate([np.sort(self.surface_edges.reshape(-1, 2), axis=1), t[:, [0, 1]], t[:, [1, 2]], t[:, [0, 2]]])
        return len(t), len(np.unique(e, axis=0)) if len(e) else 0

    def struct(self) -> N.ContactInput:
        c = N.ContactInput()
        c.surface_triangle_count = self.surface_triangle_count
        c.surface_edge_count = self.surface_edge_count
        c.surface_triangles = N.uptr(self.surface_triangles)
        c.surface_edges = N.uptr(self.surface_edges)
        c.candidate_capacity = int(self.candidate_capacity)
        c.reach_margin = float(self.reach_margin)
        return c


@dataclass
class GpuInput:
    """Arrays of `JGS2GpuInput` (layout.ts:41-90) as contiguous numpy arrays."""

    vertex_count: int
    tet_count: int
    cubature_k: int
    positions: np.ndarray
    vertex_rest: np.ndarray
    vertex_info: np.ndarray
    tet_indices: np.ndarray
    tet_inverse_dm: np.ndarray
    tet_meta: np.ndarray
    adjacency: np.ndarray
    cubature_tet_ids: np.ndarray
    cubature_weights: np.ndarray
    cubature_basis: np.ndarray
    tet_rest_stiffness: Optional[np.ndarray] = None
    velocities: Optional[np.ndarray] = None
    external_forces: Optional[np.ndarray] = None
    target_positions: Optional[np.ndarray] = None
    target_stiffnesses: Optional[np.ndarray] = None
    cloth: Optional[ClothData] = None
    contact: Optional[object] = None  # ContactData or BroadPhaseContact
    extras: dict = field(default_factory=dict)

    def __post_init__(self):
        f32 = ("positions", "vertex_rest", "tet_inverse_dm", "tet_meta", "cubature_weights", "cubature_basis",
               "tet_rest_stiffness", "velocities", "external_forces", "target_positions", "target_stiffnesses")
        u32 = ("vertex_info", "tet_indices", "adjacency", "cubature_tet_ids")
        for name in f32:
            value = getattr(self, name)
            if value is not None:
                setattr(self, name, np.ascontiguousarray(value, dtype=np.float32).reshape(-1))
        for name in u32:
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=np.uint32).reshape(-1))

    def struct(self) -> N.SceneInput:
        s = N.SceneInput()
        s.vertex_count = self.vertex_count
        s.tet_count = self.tet_count
        s.cubature_k = self.cubature_k
        s.positions = N.fptr(self.positions)
        s.velocities = N.fptr(self.velocities)
        s.vertex_rest = N.fptr(self.vertex_rest)
        s.vertex_info = N.uptr(self.vertex_info)
        s.tet_indices = N.uptr(self.tet_indices)
        s.tet_inverse_dm = N.fptr(self.tet_inverse_dm)
        s.tet_meta = N.fptr(self.tet_meta)
        s.tet_rest_stiffness = N.fptr(self.tet_rest_stiffness)
        s.adjacency = N.uptr(self.adjacency)
        s.adjacency_length = int(self.adjacency.size)
        s.cubature_tet_ids = N.uptr(self.cubature_tet_ids)
        s.cubature_weights = N.fptr(self.cubature_weights)
        s.cubature_basis = N.fptr(self.cubature_basis)
        s.external_forces = N.fptr(self.external_forces)
        s.target_positions = N.fptr(self.target_positions)
        s.target_stiffnesses = N.fptr(self.target_stiffnesses)
        if self.cloth is not None:
            # The structure holds raw pointers; keep the cloth block alive
            # with the returned SceneInput.
            self._cloth_struct = self.cloth.struct()
            s.cloth = C.pointer(self._cloth_struct)
        if self.contact is not None:
            self._contact_struct = self.contact.struct()
            s.contact = C.pointer(self._contact_struct)
        return s

    def packed_objectives(self) -> np.ndarray:
        """packJGS2VertexObjectives (layout.ts:680-716): 8 floats per vertex."""
        V = self.vertex_count
        out = np.zeros((V, 8), dtype=np.float32)
        if self.external_forces is not None:
            out[:, 0:3] = self.external_forces.reshape(V, 3)
        if self.target_positions is not None:
            out[:, 4:7] = self.target_positions.reshape(V, 3)
            out[:, 7] = self.target_stiffnesses
        else:
            out[:, 4:7] = self.positions.reshape(V, 4)[:, :3]
        return out.reshape(-1)


class JGS2GpuSolver:
    """GPU JGS2 solver for one scene (the reference's JGS2GpuSolver)."""

    def __init__(self, handle: C.c_void_p, input_: GpuInput):
        self._handle = handle
        self._input = input_
        info = N.SolverInfo()
        N.check(N.lib.jgs2_get_info(self._handle, C.byref(info)))
        self.vertex_count = info.vertex_count
        self.tet_count = info.tet_count
        self.cubature_k = info.cubature_k
        self.body_count = info.body_count
        self.schedule_color_count = info.schedule_color_count
        self.globalization_enabled = bool(info.globalization_enabled)
        self.free_vertex_count = info.free_vertex_count
        self.scene_scale = info.scene_scale

    # -------------------------------------------------------------- lifetime
    @classmethod
    def create(cls, input_: GpuInput, settings: Optional[Mapping] = None) -> "JGS2GpuSolver":
        """JGS2GpuSolver.create (jgs2-solver.ts:971-1502)."""
        s, mask = settings_struct(settings)
        st = input_.struct()
        handle = C.c_void_p()
        N.check(N.lib.jgs2_create(C.byref(st), C.byref(s), mask, C.byref(handle)))
        return cls(handle, input_)

    def destroy(self) -> None:
        if self._handle:
            N.check(N.lib.jgs2_destroy(self._handle))
            self._handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _h(self):
        if not self._handle:
            raise N.JGS2StateError(N.JGS2_ERROR_STATE, "JGS2GpuSolver has been destroyed.")
        return self._handle

    # -------------------------------------------------------------- stepping
    def step(self, overrides: Optional[Mapping] = None) -> None:
        s, mask = settings_struct(overrides)
        N.check(N.lib.jgs2_step(self._h(), C.byref(s), mask))

    def step_frames(self, frame_count: int, overrides: Optional[Mapping] = None) -> None:
        s, mask = settings_struct(overrides)
        N.check(N.lib.jgs2_step_frames(self._h(), int(frame_count), C.byref(s), mask))

    def step_exact_iterations(self, iterations: int, overrides: Optional[Mapping] = None) -> None:
        s, mask = settings_struct(overrides)
        N.check(N.lib.jgs2_step_exact_iterations(self._h(), int(iterations), C.byref(s), mask))

    def step_frames_exact_iterations(self, frame_count: int, iterations: int,
                                     overrides: Optional[Mapping] = None) -> None:
        s, mask = settings_struct(overrides)
        N.check(N.lib.jgs2_step_frames_exact_iterations(self._h(), int(frame_count), int(iterations), C.byref(s),
                                                        mask))

    def await_idle(self) -> None:
        N.check(N.lib.jgs2_sync(self._h()))

    # -------------------------------------------------------------- contact list
    def read_contact_candidates(self):
        """The current IPC candidate list: (packed uint32 [n, 4] in caller
        vertex ids, VT count, reach of the build; +inf for a static
        superset)."""
        vt, total, reach = C.c_uint32(), C.c_uint64(), C.c_float()
        N.check(N.lib.jgs2_read_contact_candidates(self._h(), None, 0, C.byref(vt), C.byref(total), C.byref(reach)))
        out = np.zeros((max(int(total.value), 1), 4), np.uint32)
        N.check(N.lib.jgs2_read_contact_candidates(self._h(), N.uptr(out), int(total.value), C.byref(vt),
                                                   C.byref(total), C.byref(reach)))
        return out[: int(total.value)], int(vt.value), float(reach.value)

    def contact_info(self) -> dict:
        """Contact list facts: length, VT part, capacity, broad-phase builds
        so far and the last build's reach."""
        i = N.ContactInfo()
        N.check(N.lib.jgs2_get_contact_info(self._h(), C.byref(i)))
        return {name: getattr(i, name) for name, _ in N.ContactInfo._fields_}

    def rebuild_contact_candidates(self) -> None:
        """Rebuilds a broad-phase scene's list at its current positions."""
        N.check(N.lib.jgs2_rebuild_contact_candidates(self._h()))

    # -------------------------------------------------------------- objectives
    def set_external_force(self, vertex: int, force) -> None:
        f = np.ascontiguousarray(force, dtype=np.float32)
        N.check(N.lib.jgs2_set_external_force(self._h(), int(vertex), N.fptr(f)))

    def replace_external_forces(self, forces) -> None:
        f = np.ascontiguousarray(forces, dtype=np.float32).reshape(-1)
        N.check(N.lib.jgs2_replace_external_forces(self._h(), N.fptr(f), f.size))

    def clear_external_forces(self) -> None:
        N.check(N.lib.jgs2_clear_external_forces(self._h()))

    def set_quadratic_target(self, vertex: int, target, stiffness: float) -> None:
        t = np.ascontiguousarray(target, dtype=np.float32)
        N.check(N.lib.jgs2_set_quadratic_target(self._h(), int(vertex), N.fptr(t), float(stiffness)))

    def replace_quadratic_targets(self, positions, stiffnesses) -> None:
        p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        k = np.ascontiguousarray(stiffnesses, dtype=np.float32).reshape(-1)
        N.check(N.lib.jgs2_replace_quadratic_targets(self._h(), N.fptr(p), p.size, N.fptr(k), k.size))

    def release_quadratic_target(self, vertex: int) -> None:
        N.check(N.lib.jgs2_release_quadratic_target(self._h(), int(vertex)))

    def release_all_quadratic_targets(self) -> None:
        N.check(N.lib.jgs2_release_all_quadratic_targets(self._h()))

    # -------------------------------------------------------------- reads
    def _read(self, fn, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.float32)
        N.check(fn(self._h(), N.fptr(out), out.size))
        return out

    def read_positions(self) -> np.ndarray:
        return self._read(N.lib.jgs2_read_positions, 4 * self.vertex_count)

    def read_velocities(self) -> np.ndarray:
        return self._read(N.lib.jgs2_read_velocities, 4 * self.vertex_count)

    def read_predicted_positions(self) -> np.ndarray:
        return self._read(N.lib.jgs2_read_predicted_positions, 4 * self.vertex_count)

    def read_final_iteration_updates(self) -> np.ndarray:
        return self._read(N.lib.jgs2_read_final_iteration_updates, 4 * self.vertex_count)

    def read_vertex_rotations(self) -> np.ndarray:
        return self._read(N.lib.jgs2_read_vertex_rotations, 12 * self.vertex_count)

    def read_local_diagnostics(self) -> np.ndarray:
        return self._read(N.lib.jgs2_read_local_diagnostics, 16 * self.vertex_count)

    def read_globalization_control(self) -> np.ndarray:
        return self._read(N.lib.jgs2_read_globalization_control, 16)

    def read_globalization_history(self) -> np.ndarray:
        return self._read(N.lib.jgs2_read_globalization_history, 64 * 32)

    def read_schedule_colors(self) -> np.ndarray:
        out = np.empty(self.vertex_count, dtype=np.uint32)
        N.check(N.lib.jgs2_read_schedule_colors(self._h(), N.uptr(out), out.size))
        return out

    def debug_local_systems(self, overrides: Optional[Mapping] = None):
        """Per-vertex (g, H) of the local solve at the accepted pose."""
        s, mask = settings_struct(overrides)
        g = np.empty(3 * self.vertex_count, dtype=np.float32)
        h = np.empty(9 * self.vertex_count, dtype=np.float32)
        N.check(N.lib.jgs2_debug_local_systems(self._h(), C.byref(s), mask, N.fptr(g), g.size, N.fptr(h), h.size))
        return g.reshape(-1, 3), h.reshape(-1, 3, 3)

    def read_oracle_diagnostics(self, overrides: Optional[Mapping] = None, positions=None, predicted=None,
                                records: bool = False):
        """readOracleDiagnostics (gpu/jgs2-diagnostics.ts:61-111): the exact
        FP64 implicit-Euler energy components, gradient norms and metrics at
        the current pose (or `positions`, 4 floats per vertex) against the
        last predicted positions (or `predicted`). With `records`, also the
        per-vertex [V, 36] and per-tetrahedron [T, 4] records."""
        s, mask = settings_struct(overrides)
        out = N.Diagnostics()
        keep = []

        def pose(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float32).reshape(-1)
            if a.size != 4 * self.vertex_count:
                raise ValueError("poses need 4 floats per vertex")
            keep.append(a)
            return N.fptr(a)

        vrec = np.empty((self.vertex_count, 36)) if records else None
        trec = np.empty((self.tet_count, 4)) if records else None
        dp = C.POINTER(C.c_double)
        N.check(N.lib.jgs2_read_oracle_diagnostics(
            self._h(), C.byref(s), mask, pose(positions), pose(predicted), C.byref(out),
            vrec.ctypes.data_as(dp) if records else None, vrec.size if records else 0,
            trec.ctypes.data_as(dp) if records and trec.size else None, trec.size if records else 0))
        summary = {name: getattr(out, name) for name, _ in N.Diagnostics._fields_ if name != "reserved"}
        return (summary, vrec, trec) if records else summary

    def read_oracle_diagnostics_noalloc(self, overrides: Optional[Mapping] = None) -> dict:
        """The summary of read_oracle_diagnostics at the solver's own pose,
        evaluated in the library's fixed device buffer (no device
        allocation); scenes too large for that buffer raise JGS2RangeError."""
        s, mask = settings_struct(overrides)
        out = N.Diagnostics()
        N.check(N.lib.jgs2_read_oracle_diagnostics_noalloc(self._h(), C.byref(s), mask, C.byref(out)))
        return {name: getattr(out, name) for name, _ in N.Diagnostics._fields_ if name != "reserved"}

    # -------------------------------------------------------------- profiling
    def profile_begin(self) -> None:
        N.check(N.lib.jgs2_profile_begin(self._h()))

    def profile_end(self):
        """(local-solve ms per launch, launches, total ms of the window)."""
        ms = C.c_double()
        n = C.c_uint64()
        total = C.c_double()
        N.check(N.lib.jgs2_profile_end(self._h(), C.byref(ms), C.byref(n), C.byref(total)))
        return ms.value, n.value, total.value


def estimate_contact_capacity(input_: GpuInput, settings: Optional[Mapping] = None) -> int:
    """The candidate capacity the library chooses for a broad-phase scene
    given no capacity (0 without contact)."""
    s, mask = settings_struct(settings)
    st = input_.struct()
    out = C.c_uint64()
    N.check(N.lib.jgs2_estimate_contact_capacity(C.byref(st), C.byref(s), mask, C.byref(out)))
    return int(out.value)
