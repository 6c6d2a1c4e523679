"""The floor penalty, grounded tangential damping and the horizontal body
correction inside the step (jgs2-shader.ts:1944-1956, :4499-4589), under the
reference app's settings for scenes without IPC (src/App.tsx:552-563: floor
stiffness 250,000, tangential damping 12 within a 0.01 m margin, horizontal
body correction on, velocity damping 0.997), on two free stable Neo-Hookean
boxes (two bodies) thrown onto the floor.

Bars: one parallel Jacobi frame within 1e-5 relative L2 of both oracles
(later Jacobi frames are not comparable: the FP32 and FP64 oracles part by
1e-1 of the displacement by frame 5 as the penalty's stiffness drives the
parallel iteration); 60 graph-colored Gauss-Seidel frames within 1e-3 of
the FP64 oracle's displacement and velocities (FP32 oracle 1.4e-4); and each
of the body correction and the tangential damping changes the GPU result,
so neither is a switched-off branch.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
import jgs2_oracle as O
from gpu_compare import rel_l2, xyz

pytestmark = pytest.mark.gpu

APP = dict(timestep=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), floor_height=0.0, floor_stiffness=250000.0,
           velocity_damping=0.997, contact_tangential_damping=12.0, contact_margin=0.01,
           horizontal_body_correction=True, parity_mode=False, max_step=0.1, schedule="jacobi")


def two_boxes():
    a = scenes.build_cuboid(scenes._desc((3, 3, 3), (0, 0.05, 0), (0.6, 0.6, 0.6), pin_axis=3, young=(2e5, 2e5),
                                         poisson=0.35))
    b = scenes.build_cuboid(scenes._desc((2, 2, 2), (1.0, 0.3, 0.2), (0.4, 0.4, 0.4), pin_axis=3, young=(1e5, 1e5),
                                         poisson=0.3))
    xa, xb = (q.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64) for q in (a, b))
    x = np.vstack([xa, xb])
    tets = np.vstack([a.tet_indices.reshape(-1, 4), b.tet_indices.reshape(-1, 4) + len(xa)])
    body = np.r_[np.zeros(len(xa)), np.ones(len(xb))].astype(np.uint32)
    mid = np.r_[np.zeros(a.tet_count), np.ones(b.tet_count)].astype(np.uint32)
    s = scenes.from_mesh(x, tets, [(1000, 2e5, 0.35), (1000, 1e5, 0.3)], material_ids=mid, body_ids=body, cubature_k=6)
    v = np.zeros((len(x), 4), np.float32)
    v[:len(xa), 0], v[:len(xa), 2] = 1.5, 0.5
    v[len(xa):, 0], v[len(xa):, 1] = -1.0, -1.0
    s.velocities = v.reshape(-1)
    return s


@pytest.fixture(scope="module")
def scene():
    return two_boxes()


def test_one_jacobi_frame_with_floor_and_body_correction(scene):
    print("case: two boxes, one Jacobi frame, app floor settings", flush=True)
    solver = jgs2.JGS2GpuSolver.create(scene, APP)
    solver.step_exact_iterations(8, APP)
    x = xyz(solver.read_positions())
    for p in (32, 64):
        o = O.Oracle(scene, p)
        o.step_frames(1, 8, APP)
        e = rel_l2(x, o.positions())
        ev = rel_l2(xyz(solver.read_velocities()), o.velocities())
        print(f"vs FP{p}: positions {e:.2e}, velocities {ev:.2e}", flush=True)
        assert e < 1e-5 and ev < 1e-4
    solver.destroy()


def run_gs(scene, settings, frames):
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    for _ in range(frames):  # a submission holds at most 2048 colored sweeps
        solver.step_exact_iterations(8, settings)
    out = xyz(solver.read_positions()), xyz(solver.read_velocities()), solver.read_schedule_colors()
    solver.destroy()
    return out


def test_sixty_gauss_seidel_frames_with_floor_damping_and_body_correction(scene):
    settings = dict(APP, schedule="graph-colored-gauss-seidel")
    print("case: two boxes, 60 GS frames on the floor", flush=True)
    x, v, colors = run_gs(scene, settings, 60)
    o = O.Oracle(scene, 64, colors=colors)
    o.step_frames(60, 8, settings)
    x0 = xyz(scene.positions)
    e = rel_l2(x - x0, o.positions() - x0)
    ev = rel_l2(v, o.velocities())
    print(f"60 frames: displacement vs FP64 {e:.2e}, velocities {ev:.2e}, lowest point {x[:, 1].min():.4f} m",
          flush=True)
    assert e < 1e-3 and ev < 1e-3
    assert x[:, 1].min() < 0.01  # resting on the floor's penalty band
    # Each branch changes the result.
    xb, _, _ = run_gs(scene, dict(settings, horizontal_body_correction=False), 60)
    xd, _, _ = run_gs(scene, dict(settings, contact_tangential_damping=0.0), 60)
    db, dd = rel_l2(xb - x0, x - x0), rel_l2(xd - x0, x - x0)
    print(f"without body correction: {db:.2e}; without tangential damping: {dd:.2e}", flush=True)
    assert db > 1e-3 and dd > 1e-3
