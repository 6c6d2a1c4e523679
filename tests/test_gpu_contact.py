"""GPU parity of the IPC contact path against the oracle on the reference's
trough (src/scenes/index.ts:551-631, ten boxes, 31,801 static candidates)
with the reference app's contact settings (src/App.tsx:529-574).

Bars: positions after one frame within 1e-5 relative L2 of the FP64 oracle,
after 10 frames within 1e-3 relative to the displacement; schedule colors
bit-exact; no candidate closer than the minimum distance after any of 60
GPU frames (checked with an independent distance); every accepted local
solve an Armijo decrease; local Hessians within 1e-5, local gradients
within twice the FP32 oracle's own distance from FP64.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
import jgs2_oracle as O
from gpu_compare import check_local_systems
from ipc_geometry import candidate_distances

pytestmark = pytest.mark.gpu

IT = scenes.TROUGH_ITERATIONS


def xyz(p):
    return np.asarray(p, np.float64).reshape(-1, 4)[:, :3]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def trough():
    return scenes.trough()


def test_trough_one_frame_matches_oracle(trough):
    settings = dict(scenes.TROUGH_SETTINGS)
    print("case: trough one frame", flush=True)
    solver = jgs2.JGS2GpuSolver.create(trough, settings)
    solver.step_exact_iterations(IT, settings)
    ref = O.Oracle(trough, 64, settings=settings)
    ref.step_frames(1, IT, settings)
    x, x0 = xyz(solver.read_positions()), xyz(trough.positions)
    e, d = rel(x, ref.positions()), rel(x - x0, ref.positions() - x0)
    print(f"rel L2 {e:.3e}, displacement rel {d:.3e}", flush=True)
    assert e < 1e-5
    assert d < 1e-4
    solver.destroy()


@pytest.mark.parametrize("schedule", ["jacobi", "graph-colored-gauss-seidel"])
def test_trough_frames_match_fp64_oracle(trough, schedule):
    settings = dict(scenes.TROUGH_SETTINGS, schedule=schedule)
    print(f"case: trough 10 frames, {schedule}", flush=True)
    solver = jgs2.JGS2GpuSolver.create(trough, settings)
    colors = None
    if schedule != "jacobi":
        colors = solver.read_schedule_colors()
        assert np.array_equal(colors, O.greedy_coloring(trough.vertex_count, O.conflict_groups(trough)))
    ref = O.Oracle(trough, 64, colors=colors, settings=settings)
    # Every IPC candidate is a conflict group, so GS needs one color per
    # trough vertex here (92 colors x 3 sweeps per frame); frames go one per
    # submission under the 2048-iteration submission limit.
    for _ in range(10):
        solver.step_exact_iterations(IT, settings)
    ref.step_frames(10, IT, settings)
    x, xr, x0 = xyz(solver.read_positions()), ref.positions(), xyz(trough.positions)
    d = rel(x - x0, xr - x0)
    print(f"10-frame displacement rel {d:.3e}", flush=True)
    assert d < 1e-3
    hist = solver.read_globalization_history().reshape(64, 8, 4)
    assert np.array_equal(hist[:IT, 1, 0], ref.history()[:IT, 1, 0])
    solver.destroy()


def test_trough_sixty_frames_keep_clearance(trough):
    settings = dict(scenes.TROUGH_SETTINGS)
    print("case: trough 60 frames, clearance", flush=True)
    solver = jgs2.JGS2GpuSolver.create(trough, settings)
    dmin = settings["ipc_minimum_distance"]
    closest = np.inf
    for frame in range(60):
        solver.step_exact_iterations(IT, settings)
        d = candidate_distances(xyz(solver.read_positions()), trough.contact)
        closest = min(closest, d.min())
        assert d.min() > dmin, frame
        diag = solver.read_local_diagnostics().reshape(-1, 4, 4)
        accepted = diag[:, 3, 3] == 0
        assert np.all(diag[accepted, 2, 0] <= diag[accepted, 2, 1])
    print(f"closest candidate distance over 60 frames {closest:.4e}", flush=True)
    assert closest < 0.08
    assert np.isfinite(xyz(solver.read_positions())).all()
    solver.destroy()


def test_trough_local_systems_in_contact_match_oracle(trough):
    settings = dict(scenes.TROUGH_SETTINGS)
    print("case: trough local systems in contact", flush=True)
    solver = jgs2.JGS2GpuSolver.create(trough, settings)
    solver.step_frames_exact_iterations(40, IT, settings)
    # Both sides evaluate the systems of a frame start at the GPU pose: the
    # contact caches, lagged forces and active rows are rebuilt from it.
    g_gpu, h_gpu = solver.debug_local_systems(settings)
    pose = (solver.read_positions(), solver.read_predicted_positions())
    d = candidate_distances(xyz(pose[0]), trough.contact)
    assert d.min() < settings["ipc_activation_distance"]  # barrier terms are active
    systems = {}
    for precision in (32, 64):
        o = O.Oracle(trough, precision, settings=settings)
        o.set_pose(*pose)
        systems[precision] = o.local_systems(settings)
    g_ref, h_ref = systems[64]
    free = trough.vertex_info.reshape(-1, 4)[:, 2] == 0
    scale = np.abs(h_ref[free]).max(axis=(1, 2))
    den = np.maximum(np.linalg.norm(g_ref[free], axis=1), 1e-6 * scale)
    g_err = np.linalg.norm(g_gpu[free] - g_ref[free], axis=1) / den
    g_err32 = np.linalg.norm(systems[32][0][free] - g_ref[free], axis=1) / den
    h_err = np.abs(h_gpu[free] - h_ref[free]).max(axis=(1, 2)) / scale
    print("max g rel", g_err.max(), "fp32 oracle", g_err32.max(), "max H rel", h_err.max(), flush=True)
    assert h_err.max() < 1e-4
    assert g_err.max() <= max(2.0 * g_err32.max(), 1e-5)
    check_local_systems(g_gpu, h_gpu, systems, free, 1e-5, "trough in contact")
    solver.destroy()
