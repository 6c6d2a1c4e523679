"""Cubature precompute (libjgs2hip `jgs2_precompute_cubature`).

The host step the reference runs once per scene, buildPrecomputedScene
(src/simulation/cpu/precompute.ts:218-380): per free vertex, up to K cubature
records (sample tetrahedron, nonnegative f32 weight, four f32 rest-basis
blocks) trained on the nonlinear pose corpus for stable Neo-Hookean scenes
(src/simulation/cpu/nonlinear-cubature-training.ts:731-843) or on
displacement modes for co-rotated linear scenes (src/simulation/cpu/
cubature.ts:714-826). `rings=0` solves each vertex's exact basis over the
whole mesh like the reference; `rings=k` over the free vertices within k
edges of the source, which keeps the cost per vertex independent of the mesh.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .solver import GpuInput

MODELS = {"corotated-linear": N.MATERIAL_COROTATED_LINEAR, "stable-neo-hookean": N.MATERIAL_STABLE_NEO_HOOKEAN}
MODE_SOLVERS = {"auto": 0, "dense": 1, "cg": 2, "gpu-cg": 3}
TRAINING = {"auto": 0, "host": 1, "gpu": 2}


@dataclass
class CubatureRecords:
    tet_ids: np.ndarray     # [V, K] uint32, 0xffffffff where unused
    weights: np.ndarray     # [V, K] float32
    basis: np.ndarray       # [V, K, 36] float32: four row-major 3x3 rest-basis blocks
    residuals: np.ndarray   # [V] packed training residual (NaN for pinned vertices)
    ranks: np.ndarray       # [V, 3] nonzero candidates, column rank, packed column rank
    timings: dict           # seconds per phase; GPU CG milliseconds and iterations


def precompute_cubature(positions, tetrahedra, fixed, material_ids, materials, timestep, samples=6, *,
                        rings=0, threads=0, mode_solver="auto", cg_tolerance=0.0,
                        residual_bound=0.0, training="auto", exact_all_elements=False) -> CubatureRecords:
    """`materials`: sequence of (density, young_modulus, poisson_ratio, model)
    with model "stable-neo-hookean" / "corotated-linear" (or the MATERIAL_*
    integer). Raises the library's error classes (JGS2RangeError for invalid
    input, JGS2InfeasibleError when a vertex's packed residual exceeds
    `residual_bound`, 0 meaning the reference's 0.01). `exact_all_elements`:
    the reference's exact all-element packing (every tetrahedron with weight
    1 and the whole-mesh exact basis; `samples` = tetrahedron count)."""
    x = np.ascontiguousarray(positions, np.float64).reshape(-1)
    t = np.ascontiguousarray(tetrahedra, np.uint32).reshape(-1)
    f = np.ascontiguousarray(np.asarray(fixed).reshape(-1) != 0, np.uint8)
    m = np.ascontiguousarray(material_ids, np.uint32).reshape(-1)
    V, T, K = x.size // 3, t.size // 4, int(samples)
    mats = (N.PrecomputeMaterial * max(1, len(materials)))()
    for i, q in enumerate(materials):
        model = MODELS[q[3]] if isinstance(q[3], str) else int(q[3])
        mats[i] = N.PrecomputeMaterial(float(q[0]), float(q[1]), float(q[2]), model, 0)
    inp = N.PrecomputeInput(V, T, len(materials), K, x.ctypes.data_as(C.POINTER(C.c_double)),
                            t.ctypes.data_as(N.u32p), f.ctypes.data_as(C.POINTER(C.c_uint8)),
                            m.ctypes.data_as(N.u32p), mats, float(timestep))
    opt = N.PrecomputeOptions(int(rings), int(threads), MODE_SOLVERS[mode_solver], TRAINING[training],
                              float(cg_tolerance), float(residual_bound), 1 if exact_all_elements else 0, 0)
    out = CubatureRecords(np.zeros((V, K), np.uint32), np.zeros((V, K), np.float32),
                          np.zeros((V, K, 36), np.float32), np.zeros(V), np.zeros((V, 3), np.uint32), {})
    timings = np.zeros(8)
    o = N.PrecomputeOutput(out.tet_ids.ctypes.data_as(N.u32p), out.weights.ctypes.data_as(N.f32p),
                           out.basis.ctypes.data_as(N.f32p), out.residuals.ctypes.data_as(C.POINTER(C.c_double)),
                           out.ranks.ctypes.data_as(N.u32p), timings.ctypes.data_as(C.POINTER(C.c_double)))
    N.check(N.lib.jgs2_precompute_cubature(C.byref(inp), C.byref(opt), C.byref(o)))
    out.timings = {"rest_s": timings[0], "modes_s": timings[1], "training_s": timings[2],
                   "gpu_cg_ms": timings[3], "gpu_cg_iterations": int(timings[4]), "gpu": bool(timings[5]),
                   "gpu_select_ms": timings[6], "gpu_training": bool(timings[7])}
    return out


def train_scene(scene: GpuInput, materials, material_ids, timestep, *, rings=0, threads=0,
                residual_bound=0.0) -> CubatureRecords:
    """Train cubature records for a packed scene from its own rest positions,
    tetrahedra and pins, and install them in place of its records."""
    rest = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    fixed = scene.vertex_info.reshape(-1, 4)[:, 2] != 0
    rec = precompute_cubature(rest, scene.tet_indices.reshape(-1, 4), fixed, material_ids, materials, timestep,
                              scene.cubature_k, rings=rings, threads=threads, residual_bound=residual_bound)
    scene.cubature_tet_ids = np.ascontiguousarray(rec.tet_ids.reshape(-1))
    scene.cubature_weights = np.ascontiguousarray(rec.weights.reshape(-1))
    scene.cubature_basis = np.ascontiguousarray(rec.basis.reshape(-1))
    scene.extras["cubature"] = f"trained (rings={rings})"
    return rec
