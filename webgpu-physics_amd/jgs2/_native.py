"""ctypes bindings of libjgs2hip (include/jgs2hip.h, include/jgs2hip_scene.h).

The library is required: importing this module raises if it has not been
built. There is no CPU fallback for any entry point that computes the step.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PACKAGE_ROOT = Path(__file__).resolve().parent.parent
LIBRARY_PATH = Path(os.environ.get("JGS2HIP_LIBRARY", PACKAGE_ROOT / "lib" / "libjgs2hip.so"))

JGS2_OK = 0
JGS2_ERROR_RANGE = 1
JGS2_ERROR_TYPE = 2
JGS2_ERROR_STATE = 3
JGS2_ERROR_DEVICE = 4
JGS2_ERROR_INFEASIBLE = 5
JGS2_ERROR_MEMORY = 6

SCHEDULE_JACOBI = 0
SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL = 1
MATERIAL_COROTATED_LINEAR = 0
MATERIAL_STABLE_NEO_HOOKEAN = 1

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)


class ClothInput(C.Structure):
    _fields_ = [
        ("triangle_count", C.c_uint32),
        ("triangle_indices", u32p),
        ("triangle_inverse_rest_bases", f32p),
        ("triangle_rest_areas", f32p),
        ("hinge_count", C.c_uint32),
        ("hinge_indices", u32p),
        ("hinge_rest_angles", f32p),
        ("hinge_rest_edge_lengths", f32p),
        ("young_modulus", C.c_double),
        ("poisson_ratio", C.c_double),
        ("thickness", C.c_double),
        ("bending_stiffness", C.c_double),
    ]


class ContactInput(C.Structure):
    _fields_ = [
        ("vertex_triangle_count", C.c_uint32),
        ("edge_edge_count", C.c_uint32),
        ("packed_indices", u32p),
        ("surface_triangle_count", C.c_uint32),
        ("surface_edge_count", C.c_uint32),
        ("surface_triangles", u32p),
        ("surface_edges", u32p),
        ("candidate_capacity", C.c_uint64),
        ("reach_margin", C.c_double),
    ]


class SceneInput(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32),
        ("tet_count", C.c_uint32),
        ("cubature_k", C.c_uint32),
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
        ("external_forces", f32p),
        ("target_positions", f32p),
        ("target_stiffnesses", f32p),
        ("cloth", C.POINTER(ClothInput)),
        ("contact", C.POINTER(ContactInput)),
    ]


SETTINGS_FIELDS = [
    ("schedule", C.c_int32),
    ("timestep", C.c_double),
    ("gravity", C.c_double * 3),
    ("iterations", C.c_int32),
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
    ("ipc_minimum_distance", C.c_double),
    ("ipc_barrier_stiffness", C.c_double),
    ("ipc_friction_coefficient", C.c_double),
    ("ipc_friction_velocity_epsilon", C.c_double),
    ("ipc_step_safety", C.c_double),
    ("horizontal_body_correction", C.c_int32),
    ("parity_mode", C.c_int32),
]


class Settings(C.Structure):
    _fields_ = SETTINGS_FIELDS


# Override mask bit of each settings field (include/jgs2hip.h JGS2_SET_*).
SETTINGS_BITS = {name: 1 << i for i, (name, _) in enumerate(SETTINGS_FIELDS)}


class SolverInfo(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32),
        ("tet_count", C.c_uint32),
        ("cubature_k", C.c_uint32),
        ("body_count", C.c_uint32),
        ("schedule_color_count", C.c_uint32),
        ("globalization_enabled", C.c_uint32),
        ("free_vertex_count", C.c_uint32),
        ("scene_scale", C.c_float),
    ]


class CuboidDesc(C.Structure):
    _fields_ = [
        ("cells", C.c_uint32 * 3),
        ("origin", C.c_double * 3),
        ("size", C.c_double * 3),
        ("jitter_fraction", C.c_double),
        ("seed", C.c_uint32),
        ("pin_axis", C.c_uint32),
        ("material_model", C.c_int32),
        ("density", C.c_double),
        ("poisson_ratio", C.c_double),
        ("young_modulus", C.c_double * 2),
        ("layer_axis", C.c_uint32),
        ("layer_cells", C.c_uint32),
        ("cubature_k", C.c_uint32),
        ("pack_rest_stiffness", C.c_int32),
    ]


class ClothDesc(C.Structure):
    _fields_ = [
        ("cells", C.c_uint32),
        ("origin", C.c_double * 3),
        ("size", C.c_double * 2),
        ("fold_column", C.c_uint32),
        ("fold_lift", C.c_double),
        ("collider_origin", C.c_double * 3),
        ("collider_size", C.c_double * 3),
        ("density", C.c_double),
        ("young_modulus", C.c_double),
        ("poisson_ratio", C.c_double),
        ("thickness", C.c_double),
        ("bending_stiffness", C.c_double),
        ("cubature_k", C.c_uint32),
    ]


class TroughDesc(C.Structure):
    _fields_ = [
        ("box_count", C.c_uint32),
        ("box_cells", C.c_uint32),
        ("seed", C.c_uint32),
        ("cubature_k", C.c_uint32),
        ("build_candidates", C.c_int32),
    ]


# Every exported entry point with its ctypes signature. Tests compare this
# table with the declarations in include/*.h.
class ContactInfo(C.Structure):
    _fields_ = [
        ("candidates", C.c_uint64),
        ("vertex_triangle", C.c_uint64),
        ("capacity", C.c_uint64),
        ("builds", C.c_uint64),
        ("reach", C.c_float),
        ("broad_phase", C.c_uint32),
    ]


class BatchConfig(C.Structure):
    _fields_ = [
        ("max_vertices", C.c_uint64),
        ("max_tets", C.c_uint64),
        ("max_triangles", C.c_uint64),
        ("max_hinges", C.c_uint64),
        ("max_candidates", C.c_uint64),
        ("max_bodies", C.c_uint64),
        ("max_scenes", C.c_uint32),
        ("cubature_k", C.c_uint32),
        ("corotated", C.c_uint32),
        ("reserved", C.c_uint32),
        ("max_surface_triangles", C.c_uint64),
        ("max_surface_edges", C.c_uint64),
    ]


class BatchInfo(C.Structure):
    _fields_ = [
        ("scene_count", C.c_uint32),
        ("max_scenes", C.c_uint32),
        ("solve_block", C.c_uint32),
        ("xcd_remap", C.c_uint32),
        ("use_graphs", C.c_uint32),
        ("solve_lanes", C.c_uint32),
        ("vertex_capacity", C.c_uint64),
        ("vertex_blocks_used", C.c_uint64),
        ("vertex_extent", C.c_uint64),
        ("tet_extent", C.c_uint64),
        ("arena_bytes", C.c_uint64),
    ]


class Diagnostics(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "energy", "inertia", "elasticity", "external_force", "quadratic_target", "floor_contact", "contact_barrier",
        "gradient_norm_inertia", "gradient_norm_material", "gradient_norm_external_force", "gradient_norm_target",
        "gradient_norm_floor", "gradient_norm", "residual_denominator", "relative_residual", "maximum_update",
        "minimum_deformation_determinant", "minimum_floor_distance", "minimum_contact_distance")] + [
        (n, C.c_uint32) for n in ("active_floor_contacts", "active_vertices", "active_contact_pairs", "finite",
                                  "maximum_update_valid", "minimum_determinant_valid", "floor_metrics_valid",
                                  "reserved")]


class PrecomputeMaterial(C.Structure):
    _fields_ = [
        ("density", C.c_double),
        ("young_modulus", C.c_double),
        ("poisson_ratio", C.c_double),
        ("model", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32),
        ("tet_count", C.c_uint32),
        ("positions", C.POINTER(C.c_double)),
        ("tetrahedra", C.POINTER(C.c_uint32)),
        ("material_ids", C.POINTER(C.c_uint32)),
        ("fixed", C.POINTER(C.c_uint8)),
        ("body_ids", C.POINTER(C.c_uint32)),
        ("material_count", C.c_uint32),
        ("cubature_k", C.c_uint32),
        ("materials", C.POINTER(PrecomputeMaterial)),
        ("cubature_length", C.c_double),
        ("pack_rest_stiffness", C.c_int32),
        ("reserved", C.c_int32),
    ]


class PrecomputeInput(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32),
        ("tetrahedron_count", C.c_uint32),
        ("material_count", C.c_uint32),
        ("cubature_samples", C.c_uint32),
        ("positions", C.POINTER(C.c_double)),
        ("tetrahedra", u32p),
        ("fixed", C.POINTER(C.c_uint8)),
        ("material_ids", u32p),
        ("materials", C.POINTER(PrecomputeMaterial)),
        ("timestep", C.c_double),
    ]


class PrecomputeOptions(C.Structure):
    _fields_ = [
        ("basis_rings", C.c_uint32),
        ("threads", C.c_uint32),
        ("mode_solver", C.c_uint32),
        ("training", C.c_uint32),
        ("cg_tolerance", C.c_double),
        ("residual_bound", C.c_double),
        ("exact_all_elements", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class PrecomputeOutput(C.Structure):
    _fields_ = [
        ("cubature_tet_ids", u32p),
        ("cubature_weights", f32p),
        ("cubature_basis", f32p),
        ("vertex_residuals", C.POINTER(C.c_double)),
        ("vertex_ranks", u32p),
        ("timings", C.POINTER(C.c_double)),
    ]


SIGNATURES = {
    "jgs2_default_settings": (None, [C.POINTER(Settings)]),
    "jgs2_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "jgs2_device_count": (C.c_int, []),
    "jgs2_create": (C.c_int, [C.POINTER(SceneInput), C.POINTER(Settings), C.c_uint64, C.POINTER(C.c_void_p)]),
    "jgs2_destroy": (C.c_int, [C.c_void_p]),
    "jgs2_step": (C.c_int, [C.c_void_p, C.POINTER(Settings), C.c_uint64]),
    "jgs2_step_frames": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Settings), C.c_uint64]),
    "jgs2_step_exact_iterations": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Settings), C.c_uint64]),
    "jgs2_step_frames_exact_iterations": (
        C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(Settings), C.c_uint64]),
    "jgs2_set_external_force": (C.c_int, [C.c_void_p, C.c_uint32, f32p]),
    "jgs2_replace_external_forces": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_clear_external_forces": (C.c_int, [C.c_void_p]),
    "jgs2_set_quadratic_target": (C.c_int, [C.c_void_p, C.c_uint32, f32p, C.c_float]),
    "jgs2_replace_quadratic_targets": (C.c_int, [C.c_void_p, f32p, C.c_size_t, f32p, C.c_size_t]),
    "jgs2_release_quadratic_target": (C.c_int, [C.c_void_p, C.c_uint32]),
    "jgs2_release_all_quadratic_targets": (C.c_int, [C.c_void_p]),
    "jgs2_read_positions": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_read_velocities": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_read_predicted_positions": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_read_final_iteration_updates": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_read_vertex_rotations": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_read_local_diagnostics": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_read_globalization_control": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_read_globalization_history": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "jgs2_sync": (C.c_int, [C.c_void_p]),
    "jgs2_read_contact_candidates": (C.c_int, [C.c_void_p, u32p, C.c_uint64, C.POINTER(C.c_uint32),
                                               C.POINTER(C.c_uint64), C.POINTER(C.c_float)]),
    "jgs2_rebuild_contact_candidates": (C.c_int, [C.c_void_p]),
    "jgs2_get_contact_info": (C.c_int, [C.c_void_p, C.POINTER(ContactInfo)]),
    "jgs2_estimate_contact_capacity": (C.c_int, [C.POINTER(SceneInput), C.POINTER(Settings), C.c_uint64,
                                                 C.POINTER(C.c_uint64)]),
    "jgs2_get_info": (C.c_int, [C.c_void_p, C.POINTER(SolverInfo)]),
    "jgs2_read_schedule_colors": (C.c_int, [C.c_void_p, u32p, C.c_size_t]),
    "jgs2_debug_local_systems": (
        C.c_int, [C.c_void_p, C.POINTER(Settings), C.c_uint64, f32p, C.c_size_t, f32p, C.c_size_t]),
    "jgs2_profile_begin": (C.c_int, [C.c_void_p]),
    "jgs2_profile_end": (
        C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "jgs2_batch_arena_bytes": (C.c_int, [C.POINTER(BatchConfig), C.POINTER(C.c_uint64)]),
    "jgs2_batch_create": (C.c_int, [C.POINTER(BatchConfig), C.POINTER(Settings), C.c_uint64, C.POINTER(C.c_void_p)]),
    "jgs2_batch_destroy": (C.c_int, [C.c_void_p]),
    "jgs2_batch_admit": (C.c_int, [C.c_void_p, C.POINTER(SceneInput), C.POINTER(C.c_void_p)]),
    "jgs2_batch_retire": (C.c_int, [C.c_void_p, C.c_void_p]),
    "jgs2_batch_step_frames": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Settings), C.c_uint64]),
    "jgs2_batch_step_frames_exact_iterations": (
        C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(Settings), C.c_uint64]),
    "jgs2_batch_sync": (C.c_int, [C.c_void_p]),
    "jgs2_batch_get_info": (C.c_int, [C.c_void_p, C.POINTER(BatchInfo)]),
    "jgs2_batch_use_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "jgs2_batch_graph_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "jgs2_batch_set_launch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    "jgs2_batch_autotune": (
        C.c_int, [C.c_void_p, C.POINTER(Settings), C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    "jgs2_batch_profile_begin": (C.c_int, [C.c_void_p]),
    "jgs2_batch_profile_end": (
        C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "jgs2_solver_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "jgs2_scene_build_cuboid": (C.c_int, [C.POINTER(CuboidDesc), C.POINTER(C.c_void_p)]),
    "jgs2_scene_build_mesh": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    "jgs2_scene_build_cloth_drape": (C.c_int, [C.POINTER(ClothDesc), C.POINTER(C.c_void_p)]),
    "jgs2_scene_build_trough": (C.c_int, [C.POINTER(TroughDesc), C.POINTER(C.c_void_p)]),
    "jgs2_scene_get_surface": (
        C.c_int, [C.c_void_p, C.POINTER(u32p), C.POINTER(C.c_uint32), C.POINTER(u32p), C.POINTER(C.c_uint32)]),
    "jgs2_build_static_contact_candidates": (
        C.c_int, [C.c_uint32, u32p, C.c_uint32, u32p, C.c_uint32, u32p, C.c_uint64, C.POINTER(C.c_uint32),
                  C.POINTER(C.c_uint32)]),
    "jgs2_scene_get_input": (C.c_int, [C.c_void_p, C.POINTER(SceneInput)]),
    "jgs2_scene_set_objectives": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "jgs2_scene_destroy": (None, [C.c_void_p]),
    "jgs2_build_adjacency": (C.c_int, [C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p]),
    "jgs2_greedy_vertex_coloring": (C.c_int, [C.c_uint32, u32p, C.c_uint32, u32p, u32p, u32p]),
    "jgs2_build_element_incidence": (C.c_int, [C.c_uint32, u32p, C.c_uint32, C.c_uint32, u32p, u32p]),
    "jgs2_read_oracle_diagnostics": (C.c_int, [C.c_void_p, C.POINTER(Settings), C.c_uint64, f32p, f32p,
                                               C.POINTER(Diagnostics), C.POINTER(C.c_double), C.c_size_t,
                                               C.POINTER(C.c_double), C.c_size_t]),
    "jgs2_read_oracle_diagnostics_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Settings), C.c_uint64,
                                                       C.POINTER(Diagnostics)]),
    "jgs2_precompute_cubature": (C.c_int, [C.POINTER(PrecomputeInput), C.POINTER(PrecomputeOptions),
                                           C.POINTER(PrecomputeOutput)]),
}


class JGS2Error(Exception):
    """Base class; `status` is the jgs2_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class JGS2RangeError(JGS2Error, ValueError):
    pass


class JGS2TypeError(JGS2Error, TypeError):
    pass


class JGS2StateError(JGS2Error, RuntimeError):
    pass


class JGS2DeviceError(JGS2Error, RuntimeError):
    pass


class JGS2InfeasibleError(JGS2Error, ValueError):
    pass


class JGS2MemoryError(JGS2Error, MemoryError):
    pass


_ERRORS = {
    JGS2_ERROR_RANGE: JGS2RangeError,
    JGS2_ERROR_TYPE: JGS2TypeError,
    JGS2_ERROR_STATE: JGS2StateError,
    JGS2_ERROR_DEVICE: JGS2DeviceError,
    JGS2_ERROR_INFEASIBLE: JGS2InfeasibleError,
    JGS2_ERROR_MEMORY: JGS2MemoryError,
}


def _load() -> C.CDLL:
    if not LIBRARY_PATH.exists():
        raise ImportError(
            f"libjgs2hip is not built ({LIBRARY_PATH}); run __graft_entry__.build() first. "
            "There is no CPU fallback.")
    lib = C.CDLL(str(LIBRARY_PATH))
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


lib = _load()


def last_error() -> str:
    buf = C.create_string_buffer(2048)
    lib.jgs2_last_error(buf, len(buf))
    return buf.value.decode("utf-8", "replace")


def check(status: int) -> None:
    if status != JGS2_OK:
        raise _ERRORS.get(status, JGS2Error)(status, last_error())


def fptr(array):
    return None if array is None else array.ctypes.data_as(f32p)


def uptr(array):
    return None if array is None else array.ctypes.data_as(u32p)
