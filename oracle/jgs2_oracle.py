"""ctypes wrapper of the CPU oracle (oracle/jgs2_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg. The product (webgpu-physics_amd/) never loads it.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = ORACLE_DIR / "_build" / "libjgs2_oracle.so"

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)


class OracleInput(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32),
        ("tet_count", C.c_uint32),
        ("cubature_k", C.c_uint32),
        ("body_count", C.c_uint32),
        ("positions", f32p),
        ("velocities", f32p),
        ("vertex_rest", f32p),
        ("vertex_info", u32p),
        ("tet_indices", u32p),
        ("tet_inverse_dm", f32p),
        ("tet_meta", f32p),
        ("tet_rest_stiffness", f32p),
        ("adjacency", u32p),
        ("adjacency_length", C.c_uint64),
        ("cubature_tet_ids", u32p),
        ("cubature_weights", f32p),
        ("cubature_basis", f32p),
        ("objectives", f32p),
        ("colors", u32p),
        ("color_count", C.c_uint32),
        ("cloth_triangle_count", C.c_uint32),
        ("cloth_triangles", u32p),
        ("cloth_inverse_rest_bases", f32p),
        ("cloth_rest_areas", f32p),
        ("cloth_hinge_count", C.c_uint32),
        ("cloth_hinges", u32p),
        ("cloth_hinge_rest", f32p),
        ("cloth_young_modulus", C.c_double),
        ("cloth_poisson_ratio", C.c_double),
        ("cloth_thickness", C.c_double),
        ("cloth_bending_stiffness", C.c_double),
        ("contact_vt_count", C.c_uint32),
        ("contact_ee_count", C.c_uint32),
        ("contact_indices", u32p),
        ("ipc_minimum_distance", C.c_double),
        ("ipc_friction_coefficient", C.c_double),
        ("ipc_friction_velocity_epsilon", C.c_double),
        ("ipc_step_safety", C.c_double),
    ]


class OracleSettings(C.Structure):
    _fields_ = [
        ("schedule", C.c_int32),
        ("timestep", C.c_double),
        ("gravity", C.c_double * 3),
        ("floor_height", C.c_double),
        ("floor_stiffness", C.c_double),
        ("velocity_damping", C.c_double),
        ("regularization", C.c_double),
        ("rotation_epsilon", C.c_double),
        ("max_step", C.c_double),
        ("residual_tolerance", C.c_double),
        ("normalized_update_tolerance", C.c_double),
        ("contact_tangential_damping", C.c_double),
        ("contact_margin", C.c_double),
        ("ipc_activation_distance", C.c_double),
        ("ipc_barrier_stiffness", C.c_double),
        ("horizontal_body_correction", C.c_int32),
    ]


def build() -> Path:
    subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, stdout=subprocess.DEVNULL)
    return LIBRARY_PATH


# Mesh arguments of the precompute entry points (oracle/jgs2_oracle_precompute.cpp):
# V, positions (f64, 3 per vertex), T, tetrahedra, fixed flags, material ids,
# material count, materials (density, young, poisson), models (1 = stable
# Neo-Hookean, 0 = co-rotated linear), timestep.
_MESH = [C.c_uint32, f64p, C.c_uint32, u32p, u32p, u32p, C.c_uint32, f64p, u32p, C.c_double]


def _load() -> C.CDLL:
    if not LIBRARY_PATH.exists():
        build()
    lib = C.CDLL(str(LIBRARY_PATH))
    sig = {
        "jgs2o_create": (C.c_void_p, [C.POINTER(OracleInput), C.c_int]),
        "jgs2o_destroy": (None, [C.c_void_p]),
        "jgs2o_globalization_enabled": (C.c_int, [C.c_void_p]),
        "jgs2o_step_frames": (None, [C.c_void_p, C.POINTER(OracleSettings), C.c_uint32, C.c_uint32, C.c_int]),
        "jgs2o_set_objectives": (None, [C.c_void_p, f32p]),
        "jgs2o_read_positions": (None, [C.c_void_p, f64p]),
        "jgs2o_read_velocities": (None, [C.c_void_p, f64p]),
        "jgs2o_read_predicted": (None, [C.c_void_p, f64p]),
        "jgs2o_read_final_updates": (None, [C.c_void_p, f64p]),
        "jgs2o_read_local_diagnostics": (None, [C.c_void_p, f64p]),
        "jgs2o_read_control": (None, [C.c_void_p, f64p]),
        "jgs2o_read_history": (None, [C.c_void_p, f64p]),
        "jgs2o_read_vertex_rotations": (None, [C.c_void_p, f64p]),
        "jgs2o_local_systems": (None, [C.c_void_p, C.POINTER(OracleSettings), f64p, f64p]),
        "jgs2o_set_threads": (None, [C.c_int]),
        "jgs2o_assembled_candidate": (None, [C.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p]),
        "jgs2o_convergence": (None, [C.c_void_p, f64p, f64p, f64p, C.c_double, C.c_double, C.c_double, C.c_int,
                                     f64p, f64p]),
        "jgs2o_restricted_model": (None, [C.c_void_p, C.POINTER(OracleSettings), C.c_uint32, f64p, f64p]),
        "jgs2o_set_pose": (None, [C.c_void_p, f32p, f32p]),
        "jgs2o_snh": (None, [C.c_int, f64p, f64p, C.c_double, C.c_double, f64p, f64p, f64p]),
        "jgs2o_snh_energy_delta": (C.c_double, [C.c_int, f64p, f64p, C.c_double, C.c_double]),
        "jgs2o_globalization_solve": (C.c_int, [C.c_int, f64p, f64p, C.c_double, f64p]),
        "jgs2o_min_eigenvalue": (C.c_double, [C.c_int, f64p]),
        "jgs2o_greedy_coloring": (C.c_uint32, [C.c_uint32, u32p, C.c_uint32, u32p, u32p]),
        "jgs2o_cloth_probe": (None, [C.c_int, f64p, f64p, C.c_double, C.c_double, C.c_double, C.c_double, f64p]),
        "jgs2o_triangle_rest": (None, [f64p, f64p, f64p]),
        "jgs2o_ipc_probe": (None, [C.c_int, C.c_int, f64p, C.c_double, C.c_double, f64p]),
        "jgs2o_ipc_friction": (None, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, f64p]),
        "jgs2o_read_contact": (None, [C.c_void_p, f64p]),
        "jgs2o_contact_candidates_within": (C.c_int64, [C.c_uint32, f32p, u32p, C.c_uint32, u32p, C.c_uint32,
                                                        C.c_float, u32p, C.c_uint64, u32p]),
        "jgs2o_precompute_error": (C.c_char_p, []),
        "jgs2o_cubature_corpus": (C.c_int, _MESH + [C.c_uint32, f64p, f64p, f64p, u32p]),
        "jgs2o_low_frequency_modes": (C.c_int, _MESH + [C.c_uint32, C.c_uint32, f64p, f64p]),
        "jgs2o_exact_basis": (C.c_int, _MESH + [C.c_uint32, f64p]),
        "jgs2o_precompute_cubature": (C.c_int, _MESH + [C.c_uint32, C.c_double, u32p, f32p, f32p, f64p]),
        "jgs2o_select_columns": (C.c_int, [C.c_uint32, C.c_uint32, f64p, f64p, C.c_uint32, C.c_double, C.c_uint32,
                                           u32p, f64p, f64p]),
        "jgs2o_fit_columns": (C.c_int, [C.c_uint32, C.c_uint32, f64p, f64p, f64p, f64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# DEFAULT_JGS2_STEP_SETTINGS (src/simulation/gpu/jgs2-solver.ts:164-187).
DEFAULTS = {
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


def resolve(overrides=None) -> dict:
    """resolveJGS2StepSettings (jgs2-solver.ts:531-550) without validation."""
    s = dict(DEFAULTS)
    s.update(overrides or {})
    if s.get("parity_mode"):
        s["velocity_damping"] = 1.0
        s["contact_tangential_damping"] = 0.0
        s["horizontal_body_correction"] = False
    return s


def settings_struct(resolved: dict) -> OracleSettings:
    o = OracleSettings()
    o.schedule = 1 if resolved["schedule"] == "graph-colored-gauss-seidel" else 0
    for key in ("timestep", "floor_height", "floor_stiffness", "velocity_damping", "regularization",
                "rotation_epsilon", "max_step", "residual_tolerance", "normalized_update_tolerance",
                "contact_tangential_damping", "contact_margin", "ipc_activation_distance", "ipc_barrier_stiffness"):
        setattr(o, key, float(resolved[key]))
    for i in range(3):
        o.gravity[i] = float(resolved["gravity"][i])
    o.horizontal_body_correction = 1 if resolved["horizontal_body_correction"] else 0
    return o


def set_threads(threads: int) -> None:
    """Host threads of the oracle's per-element and per-vertex passes
    (results are identical for every count)."""
    lib.jgs2o_set_threads(int(threads))


def greedy_coloring(vertex_count: int, groups) -> np.ndarray:
    offsets = [0]
    members = []
    for g in groups:
        members.extend(int(v) for v in g)
        offsets.append(len(members))
    off = np.asarray(offsets, dtype=np.uint32)
    mem = np.asarray(members if members else [0], dtype=np.uint32)
    colors = np.empty(vertex_count, dtype=np.uint32)
    lib.jgs2o_greedy_coloring(vertex_count, off.ctypes.data_as(u32p), len(offsets) - 1, mem.ctypes.data_as(u32p),
                              colors.ctypes.data_as(u32p))
    return colors


def conflict_groups(scene) -> list:
    """buildJGS2InputVertexColoring conflict groups (jgs2-solver.ts:308-367),
    cloth triangles, hinges and IPC candidates included (jgs2-solver.ts:
    329-364)."""
    T, K, V = scene.tet_count, scene.cubature_k, scene.vertex_count
    tets = scene.tet_indices.reshape(T, 4)
    groups = [list(t) for t in tets]
    if K:
        ids = scene.cubature_tet_ids.reshape(V, K)
        w = scene.cubature_weights.reshape(V, K)
        for v in range(V):
            for k in range(K):
                t = int(ids[v, k])
                if t == 0xFFFFFFFF or not (w[v, k] > 0):
                    continue
                groups.append([v, *tets[t]])
    cloth = getattr(scene, "cloth", None)
    if cloth is not None:
        groups.extend(list(t) for t in cloth.triangle_indices.reshape(-1, 3))
        groups.extend(list(h) for h in cloth.hinge_indices.reshape(-1, 4))
    contact = getattr(scene, "contact", None)
    if contact is not None:
        groups.extend(list(c) for c in contact.packed_indices.reshape(-1, 4))
    return groups


class Oracle:
    """The reference step restated on the CPU, FP32 (precision=32) or FP64."""

    def __init__(self, scene, precision: int = 32, colors=None, objectives=None, settings=None):
        self.scene = scene
        self.V = scene.vertex_count
        self.precision = precision
        info = scene.vertex_info.reshape(-1, 4)
        self._keep = []
        inp = OracleInput()
        inp.vertex_count = scene.vertex_count
        inp.tet_count = scene.tet_count
        inp.cubature_k = scene.cubature_k
        inp.body_count = int(info[:, 3].max()) + 1

        def fp(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            self._keep.append(a)
            return a.ctypes.data_as(f32p)

        def up(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.uint32)
            self._keep.append(a)
            return a.ctypes.data_as(u32p)

        inp.positions = fp(scene.positions)
        inp.velocities = fp(scene.velocities)
        inp.vertex_rest = fp(scene.vertex_rest)
        inp.vertex_info = up(scene.vertex_info)
        inp.tet_indices = up(scene.tet_indices)
        inp.tet_inverse_dm = fp(scene.tet_inverse_dm)
        inp.tet_meta = fp(scene.tet_meta)
        inp.tet_rest_stiffness = fp(scene.tet_rest_stiffness)
        inp.adjacency = up(scene.adjacency)
        inp.adjacency_length = int(scene.adjacency.size)
        inp.cubature_tet_ids = up(scene.cubature_tet_ids)
        inp.cubature_weights = fp(scene.cubature_weights)
        inp.cubature_basis = fp(scene.cubature_basis)
        inp.objectives = fp(objectives if objectives is not None else scene.packed_objectives())
        if colors is not None:
            inp.colors = up(colors)
            inp.color_count = int(np.max(colors)) + 1
        cloth = getattr(scene, "cloth", None)
        if cloth is not None:
            inp.cloth_triangle_count = cloth.triangle_count
            inp.cloth_triangles = up(cloth.triangle_indices)
            inp.cloth_inverse_rest_bases = fp(cloth.triangle_inverse_rest_bases)
            inp.cloth_rest_areas = fp(cloth.triangle_rest_areas)
            inp.cloth_hinge_count = cloth.hinge_count
            inp.cloth_hinges = up(cloth.hinge_indices)
            rest = np.stack([cloth.hinge_rest_angles, cloth.hinge_rest_edge_lengths], axis=1).reshape(-1)
            inp.cloth_hinge_rest = fp(rest)
            inp.cloth_young_modulus = float(cloth.young_modulus)
            inp.cloth_poisson_ratio = float(cloth.poisson_ratio)
            inp.cloth_thickness = float(cloth.thickness)
            inp.cloth_bending_stiffness = float(cloth.bending_stiffness)
        contact = getattr(scene, "contact", None)
        created = resolve(settings)  # contact parameters are fixed at create (SV:1096-1105)
        inp.ipc_minimum_distance = float(created["ipc_minimum_distance"])
        inp.ipc_friction_coefficient = float(created["ipc_friction_coefficient"])
        inp.ipc_friction_velocity_epsilon = float(created["ipc_friction_velocity_epsilon"])
        inp.ipc_step_safety = float(created["ipc_step_safety"])
        if contact is not None:
            inp.contact_vt_count = contact.vertex_triangle_count
            inp.contact_ee_count = contact.edge_edge_count
            inp.contact_indices = up(contact.packed_indices)
        self._handle = lib.jgs2o_create(C.byref(inp), precision)

    def __del__(self):
        if getattr(self, "_handle", None):
            lib.jgs2o_destroy(self._handle)
            self._handle = None

    @property
    def globalization(self) -> bool:
        return bool(lib.jgs2o_globalization_enabled(self._handle))

    def step_frames(self, frames: int, iterations: int, settings=None, stop: bool = False) -> None:
        st = settings_struct(resolve(settings))
        lib.jgs2o_step_frames(self._handle, C.byref(st), int(frames), int(iterations), 1 if stop else 0)

    def step(self, settings=None) -> None:
        """step(): odd-normalized iterations with the stop latch (layout.ts:742-748)."""
        it = int(resolve(settings)["iterations"])
        it = max(1, it)
        it = it + 1 if it % 2 == 0 else it
        self.step_frames(1, it, settings, stop=True)

    def set_objectives(self, packed: np.ndarray) -> None:
        a = np.ascontiguousarray(packed, dtype=np.float32)
        lib.jgs2o_set_objectives(self._handle, a.ctypes.data_as(f32p))

    def set_pose(self, positions: np.ndarray, predicted: np.ndarray = None) -> None:
        p = np.ascontiguousarray(positions, dtype=np.float32)
        q = None if predicted is None else np.ascontiguousarray(predicted, dtype=np.float32)
        lib.jgs2o_set_pose(self._handle, p.ctypes.data_as(f32p), None if q is None else q.ctypes.data_as(f32p))

    def _read(self, fn, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.float64)
        fn(self._handle, out.ctypes.data_as(f64p))
        return out

    def positions(self) -> np.ndarray:
        return self._read(lib.jgs2o_read_positions, 3 * self.V).reshape(-1, 3)

    def velocities(self) -> np.ndarray:
        return self._read(lib.jgs2o_read_velocities, 3 * self.V).reshape(-1, 3)

    def predicted(self) -> np.ndarray:
        return self._read(lib.jgs2o_read_predicted, 3 * self.V).reshape(-1, 3)

    def final_updates(self) -> np.ndarray:
        return self._read(lib.jgs2o_read_final_updates, 4 * self.V).reshape(-1, 4)

    def local_diagnostics(self) -> np.ndarray:
        return self._read(lib.jgs2o_read_local_diagnostics, 16 * self.V).reshape(-1, 4, 4)

    def control(self) -> np.ndarray:
        return self._read(lib.jgs2o_read_control, 16).reshape(4, 4)

    def history(self) -> np.ndarray:
        return self._read(lib.jgs2o_read_history, 64 * 32).reshape(64, 8, 4)

    def contact_state(self):
        """Per candidate: lagged normal/force, lagged weights, cached source
        contact (normal, distance) and weights; and the last safe alpha."""
        contact = getattr(self.scene, "contact", None)
        n = 0 if contact is None else contact.candidate_count
        out = self._read(lib.jgs2o_read_contact, 16 * n + 1)
        return out[:-1].reshape(n, 4, 4), out[-1]

    def vertex_rotations(self) -> np.ndarray:
        return self._read(lib.jgs2o_read_vertex_rotations, 9 * self.V).reshape(-1, 3, 3)

    def local_systems(self, settings=None):
        st = settings_struct(resolve(settings))
        g = np.empty(3 * self.V)
        h = np.empty(9 * self.V)
        lib.jgs2o_local_systems(self._handle, C.byref(st), g.ctypes.data_as(f64p), h.ctypes.data_as(f64p))
        return g.reshape(-1, 3), h.reshape(-1, 3, 3)

    def assembled_candidate(self, source, candidate, vertex_energy=None):
        """One iteration's assembled accept/revert at caller poses (3 doubles
        per vertex): (control [4, 4], positions after apply [V, 3], final
        update records [V, 4])."""
        src = np.ascontiguousarray(source, np.float64).reshape(-1)
        cand = np.ascontiguousarray(candidate, np.float64).reshape(-1)
        ve = None if vertex_energy is None else np.ascontiguousarray(vertex_energy, np.float64).reshape(-1)
        ctl, pos, upd = np.empty(16), np.empty(3 * self.V), np.empty(4 * self.V)
        lib.jgs2o_assembled_candidate(self._handle, src.ctypes.data_as(f64p), cand.ctypes.data_as(f64p),
                                      None if ve is None else ve.ctypes.data_as(f64p), ctl.ctypes.data_as(f64p),
                                      pos.ctypes.data_as(f64p), upd.ctypes.data_as(f64p))
        return ctl.reshape(4, 4), pos.reshape(-1, 3), upd.reshape(-1, 4)

    def convergence(self, gradients, updates, control, scene_scale, residual_tolerance, update_tolerance,
                    stop=False):
        """The convergence reduction of caller records: gradients [V, 5, 3]
        (inertia, material, external force, target, contact), accepted
        update vectors [V, 3], control [4, 4] (control[1, 3] = history
        row). Returns (history row [8, 4], control after [4, 4])."""
        g = np.ascontiguousarray(gradients, np.float64).reshape(-1)
        u = np.ascontiguousarray(updates, np.float64).reshape(-1)
        c = np.ascontiguousarray(control, np.float64).reshape(-1)
        row, ctl = np.empty(32), np.empty(16)
        lib.jgs2o_convergence(self._handle, g.ctypes.data_as(f64p), u.ctypes.data_as(f64p), c.ctypes.data_as(f64p),
                              float(scene_scale), float(residual_tolerance), float(update_tolerance),
                              1 if stop else 0, row.ctypes.data_as(f64p), ctl.ctypes.data_as(f64p))
        return row.reshape(8, 4), ctl.reshape(4, 4)

    def restricted_model(self, vertex: int, d, settings=None):
        st = settings_struct(resolve(settings))
        dd = np.ascontiguousarray(d, dtype=np.float64)
        out = np.empty(3)
        lib.jgs2o_restricted_model(self._handle, C.byref(st), int(vertex), dd.ctypes.data_as(f64p),
                                   out.ctypes.data_as(f64p))
        return out[0], bool(out[1]), out[2]


def snh(F, H, lam, mu, precision=64):
    f = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    h = np.ascontiguousarray(H, dtype=np.float64).reshape(9)
    e = C.c_double()
    p = np.empty(9)
    t = np.empty(9)
    lib.jgs2o_snh(precision, f.ctypes.data_as(f64p), h.ctypes.data_as(f64p), lam, mu, C.byref(e),
                  p.ctypes.data_as(f64p), t.ctypes.data_as(f64p))
    return e.value, p.reshape(3, 3), t.reshape(3, 3)


def snh_energy_delta(F0, F1, lam, mu, precision=64) -> float:
    a = np.ascontiguousarray(F0, dtype=np.float64).reshape(9)
    b = np.ascontiguousarray(F1, dtype=np.float64).reshape(9)
    return lib.jgs2o_snh_energy_delta(precision, a.ctypes.data_as(f64p), b.ctypes.data_as(f64p), lam, mu)


def globalization_solve(H, g, inertia_scale, precision=64):
    h = np.ascontiguousarray(H, dtype=np.float64).reshape(9)
    gg = np.ascontiguousarray(g, dtype=np.float64).reshape(3)
    out = np.empty(9)
    status = lib.jgs2o_globalization_solve(precision, h.ctypes.data_as(f64p), gg.ctypes.data_as(f64p),
                                           float(inertia_scale), out.ctypes.data_as(f64p))
    return status, {
        "direction": out[0:3].copy(), "gradient_dot_direction": out[3], "scale": out[4],
        "minimum_eigenvalue_normalized": out[5], "normalized_shift": out[6], "maximum_relative_asymmetry": out[7],
        "shifted_relative_residual": out[8],
    }


def min_eigenvalue(H, precision=64) -> float:
    h = np.ascontiguousarray(H, dtype=np.float64).reshape(9)
    return lib.jgs2o_min_eigenvalue(precision, h.ctypes.data_as(f64p))


def cloth_probe(rest12, pos12, young, poisson, thickness, bending, precision=64) -> dict:
    """StVK membrane of triangle (0,1,2) and dihedral bending of hinge
    (0,1,2,3) of a four-vertex sheet with rest pose `rest12` evaluated at
    `pos12` (src/simulation/cpu/cloth.ts restated)."""
    r = np.ascontiguousarray(rest12, dtype=np.float64).reshape(12)
    x = np.ascontiguousarray(pos12, dtype=np.float64).reshape(12)
    out = np.empty(36)
    lib.jgs2o_cloth_probe(precision, r.ctypes.data_as(f64p), x.ctypes.data_as(f64p), float(young), float(poisson),
                          float(thickness), float(bending), out.ctypes.data_as(f64p))
    return {
        "membrane_energy": out[0], "membrane_gradient": out[1:10].reshape(3, 3).copy(),
        "bending_energy": out[10], "bending_gradient": out[11:23].reshape(4, 3).copy(), "angle": out[23],
        "rest_angle": out[24], "rest_edge_length": out[25], "membrane_hessian_vertex0": out[26:35].reshape(3, 3).copy(),
        "hinge_valid": bool(out[35]),
    }


def triangle_rest(x9):
    x = np.ascontiguousarray(x9, dtype=np.float64).reshape(9)
    inv = np.empty(4)
    area = C.c_double()
    lib.jgs2o_triangle_rest(x.ctypes.data_as(f64p), inv.ctypes.data_as(f64p), C.byref(area))
    return inv.reshape(2, 2), area.value


def ipc_probe(kind, p12, dhat, kappa, precision=64) -> dict:
    """One IPC candidate (kind "vt" or "ee") at 4 positions: contact distance,
    normal, weights, validity, barrier value and derivatives and the
    per-vertex barrier gradient."""
    x = np.ascontiguousarray(p12, dtype=np.float64).reshape(12)
    out = np.empty(24)
    lib.jgs2o_ipc_probe(precision, 0 if kind == "vt" else 1, x.ctypes.data_as(f64p), float(dhat), float(kappa),
                        out.ctypes.data_as(f64p))
    return {"distance": out[0], "normal": out[1:4].copy(), "weights": out[4:8].copy(), "valid": bool(out[8]),
            "energy": out[9], "d1": out[10], "d2": out[11], "gradient": out[12:24].reshape(4, 3).copy()}


def ipc_friction(slip, eps, mu, force, precision=64):
    out = np.empty(2)
    lib.jgs2o_ipc_friction(precision, float(slip), float(eps), float(mu), float(force), out.ctypes.data_as(f64p))
    return out[0], out[1]


def contact_candidates_within(vertex_count: int, positions, triangles, edges, reach: float):
    """The reference's static VT/EE superset restricted to pairs whose f32
    AABB distance lower bound at `positions` (4 floats per vertex) is not
    above `reach` (float('inf') = the whole superset). Returns
    (packed uint32 [n, 4], vertex-triangle count)."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    tri = np.ascontiguousarray(triangles, np.uint32).reshape(-1)
    edg = np.ascontiguousarray(edges, np.uint32).reshape(-1)
    vt = C.c_uint32()
    args = (int(vertex_count), pos.ctypes.data_as(f32p), tri.ctypes.data_as(u32p), tri.size // 3,
            edg.ctypes.data_as(u32p), edg.size // 2, float(reach))
    n = lib.jgs2o_contact_candidates_within(*args, None, 0, C.byref(vt))
    if n < 0:
        raise ValueError("invalid surface topology")
    out = np.zeros((max(n, 1), 4), np.uint32)
    lib.jgs2o_contact_candidates_within(*args, out.ctypes.data_as(u32p), n, C.byref(vt))
    return out[:n], vt.value


class PrecomputeError(RuntimeError):
    pass


def _mesh(positions, tetrahedra, fixed, material_ids, materials, timestep):
    x = np.ascontiguousarray(positions, np.float64).reshape(-1)
    t = np.ascontiguousarray(tetrahedra, np.uint32).reshape(-1)
    f = np.ascontiguousarray(fixed, np.uint32).reshape(-1)
    m = np.ascontiguousarray(material_ids, np.uint32).reshape(-1)
    mats = np.ascontiguousarray([[q[0], q[1], q[2]] for q in materials], np.float64).reshape(-1)
    models = np.ascontiguousarray([1 if (len(q) > 3 and q[3]) else 0 for q in materials], np.uint32)
    keep = (x, t, f, m, mats, models)
    args = [x.size // 3, x.ctypes.data_as(f64p), t.size // 4, t.ctypes.data_as(u32p), f.ctypes.data_as(u32p),
            m.ctypes.data_as(u32p), len(materials), mats.ctypes.data_as(f64p), models.ctypes.data_as(u32p),
            float(timestep)]
    return args, keep


def _check(rc):
    if rc < 0:
        raise PrecomputeError(lib.jgs2o_precompute_error().decode())
    return rc


def cubature_corpus(positions, tetrahedra, fixed, material_ids, materials, timestep):
    """Training and validation poses of the nonlinear cubature corpus:
    (training, validation), each a list of (positions, predicted, min J)."""
    args, keep = _mesh(positions, tetrahedra, fixed, material_ids, materials, timestep)
    V = args[0]
    cap = 64
    pos = np.zeros((cap, V * 3)); pred = np.zeros((cap, V * 3)); mj = np.zeros(cap)
    nv = C.c_uint32()
    nt = _check(lib.jgs2o_cubature_corpus(*args, cap, pos.ctypes.data_as(f64p), pred.ctypes.data_as(f64p),
                                          mj.ctypes.data_as(f64p), C.byref(nv)))
    poses = [(pos[i], pred[i], float(mj[i])) for i in range(nt + nv.value)]
    return poses[:nt], poses[nt:]


def low_frequency_modes(positions, tetrahedra, fixed, material_ids, materials, timestep, count, iterations):
    """(modes [count, 3V], dense rest Hessian over free coordinates)."""
    args, keep = _mesh(positions, tetrahedra, fixed, material_ids, materials, timestep)
    V = args[0]
    n = 3 * int((np.asarray(fixed).reshape(-1) == 0).sum())
    out = np.zeros((count, 3 * V)); hess = np.zeros((n, n))
    k = _check(lib.jgs2o_low_frequency_modes(*args, count, iterations, out.ctypes.data_as(f64p), hess.ctypes.data_as(f64p)))
    return out[:k], hess


def exact_basis(positions, tetrahedra, fixed, material_ids, materials, timestep, vertex):
    """Exact rest basis of one free vertex: [V, 3, 3] (row = coordinate)."""
    args, keep = _mesh(positions, tetrahedra, fixed, material_ids, materials, timestep)
    out = np.zeros(args[0] * 9)
    _check(lib.jgs2o_exact_basis(*args, int(vertex), out.ctypes.data_as(f64p)))
    return out.reshape(-1, 3, 3)


def precompute_cubature(positions, tetrahedra, fixed, material_ids, materials, timestep, samples, bound=0.01):
    """Packed cubature records (tetrahedra [V, K], f32 weights [V, K], f32
    basis [V, K, 36]) and per-vertex statistics [V, 8]: packed residual, f64
    residual, valid and trivial training poses, nonzero candidates, column
    rank, packed nonzero candidates, packed column rank (NaN when fixed)."""
    args, keep = _mesh(positions, tetrahedra, fixed, material_ids, materials, timestep)
    V, K = args[0], int(samples)
    tets = np.zeros((V, K), np.uint32); w = np.zeros((V, K), np.float32)
    basis = np.zeros((V, K, 36), np.float32); stats = np.zeros((V, 8))
    _check(lib.jgs2o_precompute_cubature(*args, K, float(bound), tets.ctypes.data_as(u32p), w.ctypes.data_as(f32p),
                                         basis.ctypes.data_as(f32p), stats.ctypes.data_as(f64p)))
    return tets, w, basis, stats


def select_columns(columns, target, maximum_columns, tolerance=0.01, refinement_passes=4):
    """Greedy nonnegative column selection: (indices, weights, target norm,
    residual norm, normalized residual)."""
    t = np.ascontiguousarray(target, np.float64).reshape(-1)
    cols = np.ascontiguousarray(np.asarray(columns, np.float64).reshape(-1, t.size))
    sel = np.zeros(max(1, len(cols)), np.uint32); w = np.zeros(max(1, len(cols))); norms = np.zeros(3)
    k = _check(lib.jgs2o_select_columns(len(cols), t.size, cols.ctypes.data_as(f64p), t.ctypes.data_as(f64p),
                                        int(maximum_columns), float(tolerance), int(refinement_passes),
                                        sel.ctypes.data_as(u32p), w.ctypes.data_as(f64p), norms.ctypes.data_as(f64p)))
    return [int(i) for i in sel[:k]], w[:k].copy(), norms[0], norms[1], norms[2]


def fit_columns(columns, target):
    """Exact nonnegative fit of the given columns: (weights, residual norm,
    normalized residual)."""
    t = np.ascontiguousarray(target, np.float64).reshape(-1)
    cols = np.ascontiguousarray(np.asarray(columns, np.float64).reshape(-1, t.size))
    w = np.zeros(max(1, len(cols))); norms = np.zeros(2)
    _check(lib.jgs2o_fit_columns(len(cols), t.size, cols.ctypes.data_as(f64p), t.ctypes.data_as(f64p),
                                 w.ctypes.data_as(f64p), norms.ctypes.data_as(f64p)))
    return w[:len(cols)].copy(), norms[0], norms[1]
