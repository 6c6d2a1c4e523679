"""The legacy co-rotated linear step on the GPU against the oracle (whose
co-rotated path is pinned to closed forms in tests/test_oracle_corotated.py):
tetPolarRotation (jgs2-shader.ts:2932-2957), the frozen-rotation gradient
and R K_ss R^T blocks (:1022-1043, :1163-1198), the regularized 3x3 solve
(:2479-2512) and the maxStep clamp (:3475-3486), with the cubature
correction on the 425-vertex cantilever.

Bar: per-vertex H within 1e-5 of both oracles and g within three times the
FP32 oracle's own distance from FP64 (the rest-frame offset difference in
the co-rotated force costs FP32 about four digits); positions after one
frame within 1e-5 relative L2 of both oracles; after 5 Jacobi frames within
1e-5 of the FP32 oracle; after 30 graph-colored Gauss-Seidel frames within
1e-3 of the FP64 oracle. Jacobi on this legacy path grows an unstable mode
after about five frames in the FP64 oracle as well (f32 and f64 oracles
differ by 1e-2 at frame 10), which is why long runs use Gauss-Seidel, where
the two oracles stay within 3e-6.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import _native as N
from jgs2 import scenes
import jgs2_oracle as O
from gpu_compare import check_local_systems, rel_l2, xyz

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def beam():
    return scenes.cantilever_beam(model=N.MATERIAL_COROTATED_LINEAR)


def test_corotated_local_systems_match_oracle(beam):
    settings = dict(scenes.CANTILEVER_SETTINGS)
    print(f"case: co-rotated local systems, cantilever V={beam.vertex_count}", flush=True)
    solver = jgs2.JGS2GpuSolver.create(beam, settings)
    assert not solver.globalization_enabled
    solver.step_frames_exact_iterations(3, 8, settings)
    g_gpu, h_gpu = solver.debug_local_systems(settings)
    pos, pred = solver.read_positions(), solver.read_predicted_positions()
    free = beam.vertex_info.reshape(-1, 4)[:, 2] == 0
    systems = {}
    for precision in (32, 64):
        oracle = O.Oracle(beam, precision)
        oracle.set_pose(pos, pred)
        systems[precision] = oracle.local_systems(settings)
    # The co-rotated force R K (R^T (x - c) - (X - C)) subtracts two nearly
    # equal rest-frame offsets, so FP32 resolves it to ~1e-4 of its largest
    # value (the FP32 oracle included): g is held to three times the FP32
    # oracle's distance from FP64 there, H to 1e-5.
    check_local_systems(g_gpu, h_gpu, systems, free, TOL, "co-rotated")
    solver.destroy()


def test_corotated_jacobi_frames_match_oracle(beam):
    settings = dict(scenes.CANTILEVER_SETTINGS)
    print("case: co-rotated Jacobi, 1 and 5 frames", flush=True)
    solver = jgs2.JGS2GpuSolver.create(beam, settings)
    ref32, ref64 = O.Oracle(beam, 32), O.Oracle(beam, 64)
    x0 = xyz(beam.positions)
    for frame in range(5):
        solver.step_exact_iterations(8, settings)
        ref32.step_frames(1, 8, settings)
        ref64.step_frames(1, 8, settings)
        x = xyz(solver.read_positions())
        e32, e64 = rel_l2(x, ref32.positions()), rel_l2(x, ref64.positions())
        d32 = rel_l2(x - x0, ref32.positions() - x0)
        print(f"frame {frame}: rel L2 vs f32 {e32:.2e}, vs f64 {e64:.2e}, displacement vs f32 {d32:.2e}", flush=True)
        assert e32 < TOL
        if frame == 0:
            assert e64 < TOL
    upd = solver.read_final_iteration_updates().reshape(-1, 4)
    assert np.array_equal(upd[:, 3], ref32.final_updates()[:, 3])
    assert np.abs(upd[:, 0]).max() <= settings.get("max_step", 0.1) * (1 + 1e-6)  # the maxStep clamp
    solver.destroy()


def test_corotated_gauss_seidel_thirty_frames_match_fp64_oracle(beam):
    settings = dict(scenes.CANTILEVER_SETTINGS, schedule="graph-colored-gauss-seidel")
    print("case: co-rotated GS, 30 frames", flush=True)
    solver = jgs2.JGS2GpuSolver.create(beam, settings)
    colors = solver.read_schedule_colors()
    assert np.array_equal(colors, O.greedy_coloring(beam.vertex_count, O.conflict_groups(beam)))
    ref = O.Oracle(beam, 64, colors=colors)
    solver.step_frames_exact_iterations(30, 8, settings)
    ref.step_frames(30, 8, settings)
    x0 = xyz(beam.positions)
    x = xyz(solver.read_positions())
    err = rel_l2(x, ref.positions())
    derr = rel_l2(x - x0, ref.positions() - x0)
    print(f"30-frame rel L2 {err:.2e}, displacement rel {derr:.2e}", flush=True)
    assert derr < 1e-3
    assert np.isfinite(x).all()
    solver.destroy()
