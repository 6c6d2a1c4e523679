"""The GPU step with trained cubature records (jgs2_precompute_cubature)
against the FP64 oracle stepping the same records.

The records replace the closed-form stand-in: sparse ring-restricted
bases with exact zero blocks, weights far from uniform and samples that are
not incident to their source vertex, so the cubature correction of the local
solve runs on the data it is meant for. Tolerances as in test_gpu_parity.py.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
from jgs2.precompute import train_scene
import jgs2_oracle as O
from gpu_compare import check_local_systems

pytestmark = pytest.mark.gpu


def xyz(p):
    return np.asarray(p, np.float64).reshape(-1, 4)[:, :3]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


# The whole-mesh basis (rings = 0) on the reference's fixture size; ring-
# restricted bases on a longer beam, where whole-mesh bases make parallel
# Jacobi iterations revert (in the oracle as on the GPU).
@pytest.mark.parametrize("cells,size,rings", [((2, 1, 1), (1.0, 1.0, 1.0), 0), ((8, 2, 2), (2.0, 0.5, 0.5), 1),
                                              ((8, 2, 2), (2.0, 0.5, 0.5), 2)])
def test_trained_cantilever_matches_fp64_oracle(cells, size, rings):
    print(f"case: trained cubature cantilever {cells}, rings {rings}", flush=True)
    scene = scenes.cantilever_beam(cells=cells, size=size)
    settings = dict(scenes.CANTILEVER_SETTINGS)
    rec = train_scene(scene, [(1000.0, 1e5, 0.4, "stable-neo-hookean")], np.zeros(scene.tet_count), 1.0 / 60.0,
                      rings=rings, residual_bound=1.0)
    assert np.nanmax(rec.residuals) < 1.0
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    solver.step_frames_exact_iterations(3, 8, settings)
    g_gpu, h_gpu = solver.debug_local_systems(settings)
    pos, pred = solver.read_positions(), solver.read_predicted_positions()
    systems = {}
    for precision in (32, 64):
        ref = O.Oracle(scene, precision)
        ref.set_pose(pos, pred)
        systems[precision] = ref.local_systems(settings)
    g_ref, h_ref = systems[64]
    free = scene.vertex_info.reshape(-1, 4)[:, 2] == 0
    h_err = np.abs(h_gpu[free] - h_ref[free]).max(axis=(1, 2)) / np.abs(h_ref[free]).max(axis=(1, 2))
    den = np.maximum(np.linalg.norm(g_ref[free], axis=1), 1e-6 * np.abs(h_ref[free]).max(axis=(1, 2)))
    g_err = np.linalg.norm(g_gpu[free] - g_ref[free], axis=1) / den
    g_err32 = np.linalg.norm(systems[32][0][free] - g_ref[free], axis=1) / den
    print(f"local systems: max H rel {h_err.max():.3e}, max g rel {g_err.max():.3e} "
          f"(fp32 oracle {g_err32.max():.3e})", flush=True)
    # Near equilibrium the trained weights cancel large terms: the gradient is
    # held to the f32 rounding the FP32 oracle itself shows at the same pose.
    assert h_err.max() < 1e-5
    assert g_err.max() <= 3 * g_err32.max() + 1e-5
    check_local_systems(g_gpu, h_gpu, systems, free, 1e-5, "trained cubature")
    solver.destroy()
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    frames = O.Oracle(scene, 64)
    solver.step_frames_exact_iterations(10, 8, settings)
    frames.step_frames(10, 8, settings)
    x0 = xyz(scene.positions)
    x = xyz(solver.read_positions())
    err, derr = rel(x, frames.positions()), rel(x - x0, frames.positions() - x0)
    print(f"10 frames: rel L2 {err:.3e}, displacement rel {derr:.3e}", flush=True)
    assert np.isfinite(x).all() and err < 1e-5 and derr < 1e-3
    hist = solver.read_globalization_history().reshape(64, 8, 4)
    assert np.array_equal(hist[:8, 1, 0], frames.history()[:8, 1, 0])  # same accept decisions
    solver.destroy()


def test_device_mode_solver_agrees_with_the_dense_factor():
    from jgs2.precompute import precompute_cubature
    from test_precompute_oracle import cuboid
    print("case: precompute modes by device CG vs dense factor, 4x2x2 and 8x4x4", flush=True)
    pos, tets, fixed = cuboid((4, 2, 2), (0, 0, 0), (2, 1, 1), lambda p, c: c[0] == 0)
    args = (pos, tets, fixed, np.zeros(len(tets)), [(1000, 1e5, 0.4, 1)], 1 / 60)
    a = precompute_cubature(*args, 6, rings=2, residual_bound=1.0, mode_solver="dense")
    b = precompute_cubature(*args, 6, rings=2, residual_bound=1.0, mode_solver="gpu-cg", cg_tolerance=1e-13)
    assert b.timings["gpu"] and b.timings["gpu_cg_iterations"] > 0
    same = np.mean([np.array_equal(a.tet_ids[v], b.tet_ids[v]) for v in range(len(pos))])
    free = fixed == 0
    print(f"same selections {same:.3f}, GPU CG {b.timings['gpu_cg_iterations']} iterations "
          f"{b.timings['gpu_cg_ms']:.1f} ms", flush=True)
    assert same >= 0.9
    np.testing.assert_allclose(a.residuals[free], b.residuals[free], rtol=0.05, atol=1e-3)
    # a larger system: device CG against host CG
    pos, tets, fixed = cuboid((16, 8, 8), (0, 0, 0), (2, 1, 1), lambda p, c: c[0] == 0)
    args = (pos, tets, fixed, np.zeros(len(tets)), [(1000, 1e5, 0.4, 1)], 1 / 60)
    h = precompute_cubature(*args, 6, rings=1, residual_bound=1.0, mode_solver="cg", cg_tolerance=1e-12)
    g = precompute_cubature(*args, 6, rings=1, residual_bound=1.0, mode_solver="gpu-cg", cg_tolerance=1e-12)
    assert g.timings["gpu"] and not h.timings["gpu"]
    same = np.mean([np.array_equal(h.tet_ids[v], g.tet_ids[v]) for v in range(len(pos))])
    free = fixed == 0
    print(f"V={len(pos)}: same selections {same:.3f}; host modes {h.timings['modes_s']:.2f} s, "
          f"device modes {g.timings['modes_s']:.2f} s ({g.timings['gpu_cg_iterations']} CG iterations, "
          f"{g.timings['gpu_cg_ms']:.0f} ms)", flush=True)
    assert same >= 0.9
    np.testing.assert_allclose(h.residuals[free], g.residuals[free], rtol=0.05, atol=1e-3)


@pytest.mark.parametrize("rings,two_materials", [(1, False), (2, True)])
def test_device_selection_equals_the_host_selection(rings, two_materials):
    from jgs2.precompute import precompute_cubature
    from test_precompute_oracle import cuboid
    print(f"case: device greedy selection vs host, 10x6x6, rings {rings}, two materials {two_materials}", flush=True)
    pos, tets, fixed = cuboid((10, 6, 6), (0, 0, 0), (2.5, 1.5, 1.5), lambda p, c: c[0] == 0)
    mid = ((np.arange(len(tets)) // 48) % 2) if two_materials else np.zeros(len(tets))
    mats = [(1000, 1e5, 0.4, 1), (1000, 1e6, 0.45, 1)]
    args = (pos, tets, fixed, mid, mats, 1 / 60)
    host = precompute_cubature(*args, 6, rings=rings, residual_bound=1.0, training="host")
    dev = precompute_cubature(*args, 6, rings=rings, residual_bound=1.0, training="gpu")
    assert dev.timings["gpu_training"] and not host.timings["gpu_training"]
    same = np.mean([np.array_equal(host.tet_ids[v], dev.tet_ids[v]) for v in range(len(pos))])
    print(f"V={len(pos)}: identical selections {same:.4f}; host training {host.timings['training_s']:.2f} s, "
          f"device {dev.timings['training_s']:.2f} s (selection kernel {dev.timings['gpu_select_ms']:.0f} ms)",
          flush=True)
    assert np.array_equal(host.tet_ids, dev.tet_ids)
    assert np.array_equal(host.weights, dev.weights)
    assert np.array_equal(host.basis, dev.basis)
    assert np.array_equal(host.residuals, dev.residuals, equal_nan=True)
