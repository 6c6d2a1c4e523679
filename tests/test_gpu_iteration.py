"""Assembled accept/revert, convergence, history and the stop latch on the
GPU against the FP32 oracle (whose reductions are pinned to the reference's
nonlinear-iteration.test.ts cases in tests/test_oracle_iteration.py).

* Forced whole-pose revert: the reference's one-tetrahedron case
  (nonlinear-iteration.test.ts:53-100) inside a real step. Two vertices are
  launched toward moves that are each feasible alone (J = 1) but invert the
  tetrahedron together (J ~ -0.21), with inertia dominating the local solves,
  so every Jacobi iteration of the frame reverts the complete pose.
* step(): odd-normalized iteration counts and the convergence stop latch
  (layout.ts:742-748, jgs2-shader.ts:4421-4486) over three frames: all
  reverts, a one-row latch, and a three-row latch after a gravity override.
* Every history row, the control records and the final updates of a
  Jacobi cantilever through stepFramesExactIterations and stepFrames.

Bar: integer and flag fields (accept, revert, failures, validity bits,
converged, finite, row index, backtracks, trial counts) equal; float fields
(minimum determinants, shift) within 1e-5 of the FP32 oracle relative to
the field's magnitude; the assembled energies, the gradient component
norms, the summed-gradient norm, the relative residual and the update norms
(f32 sums or differences of nearly cancelling terms) within 3e-4 of the
scale of their terms (the frame's largest energy, its largest residual
denominator, 1, its largest update; updates at least to eight f32 ulps of
the largest coordinate) against both oracles; positions within 1e-5 relative L2, and bit-identical to
the source pose after a reverted frame.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
import jgs2_oracle as O
from gpu_compare import (CANCELLING_TOL, CONTROL_CANCELLING_FIELDS, CONTROL_INT_FIELDS, HISTORY_INT_FIELDS,
                         F32_ULP, cancelling_scales, compare_records, rel_l2, xyz)

pytestmark = pytest.mark.gpu

TOL = 1e-5
L, DT = 0.05, 1.0 / 60.0
REVERT_SETTINGS = dict(timestep=DT, gravity=(0.0, 0.0, 0.0), floor_stiffness=0.0, parity_mode=True,
                       schedule="jacobi", iterations=8)


def colliding_tet():
    X = np.array([[0, 0, 0], [L, 0, 0], [0, L, 0], [0, 0, L]], np.float64)
    scene = scenes.from_mesh(X, [[0, 1, 2, 3]], [(1e6, 1e4, 0.3)])
    vel = np.zeros((4, 4), np.float32)
    vel[1, 1] = 1.1 * L / DT  # alone: (L, 1.1 L, 0), J = 1
    vel[2, 0] = 1.1 * L / DT  # alone: (1.1 L, L, 0), J = 1; together J = 1 - 1.1^2
    scene.velocities = vel.reshape(-1)
    return scene


def compare_frame(solver, oracle, oracle64, label):
    ref_ctl = oracle.control()
    n = int(ref_ctl[1, 3])
    assert n >= 1 and int(oracle64.control()[1, 3]) == n
    max_coordinate = np.abs(oracle64.positions()).max()
    scales = cancelling_scales(oracle64.history()[:n], max_coordinate, solver.scene_scale)
    ctl = solver.read_globalization_control().reshape(4, 4)
    compare_records(ctl, ref_ctl, CONTROL_INT_FIELDS, TOL, f"{label} control", oracle64.control(),
                    {f: scales[f] for f in CONTROL_CANCELLING_FIELDS})
    hist = solver.read_globalization_history().reshape(64, 8, 4)
    worst = compare_records(hist[:n], oracle.history()[:n], HISTORY_INT_FIELDS, TOL, f"{label} history",
                            oracle64.history()[:n], scales)
    err = rel_l2(xyz(solver.read_positions()), oracle.positions())
    upd = solver.read_final_iteration_updates().reshape(-1, 4)
    ref_upd, ref_upd64 = oracle.final_updates(), oracle64.final_updates()
    assert np.array_equal(upd[:, 3], ref_upd[:, 3])
    allowed = max(CANCELLING_TOL * np.abs(ref_upd64[:, 0]).max(), 8 * F32_ULP * max_coordinate)
    for ref in (ref_upd, ref_upd64):
        assert np.abs(upd[:, 0] - ref[:, 0]).max() <= allowed
    print(f"{label}: {n} history rows, worst float field {worst:.2e}, positions rel L2 {err:.2e}", flush=True)
    assert err < TOL
    return ctl, hist[:n]


def test_forced_whole_pose_revert_and_stop_latch_match_oracle():
    scene = colliding_tet()
    print("case: one-tetrahedron forced revert, step() x3", flush=True)
    solver = jgs2.JGS2GpuSolver.create(scene, REVERT_SETTINGS)
    oracle, oracle64 = O.Oracle(scene, 32), O.Oracle(scene, 64)
    # Frame 1: step() normalizes 8 iterations to 9; every one reverts.
    for o in (solver, oracle, oracle64):
        o.step(REVERT_SETTINGS)
    ctl, hist = compare_frame(solver, oracle, oracle64, "frame 1")
    assert len(hist) == 9
    assert (hist[:, 1, 0] == 0).all() and (hist[:, 1, 1] == 1).all()
    assert (hist[:, 0, 1] < -0.2).all() and (hist[:, 0, 2] == 1.0).all()
    assert (hist[:, 6, 0] == 0).all()  # no convergence through a revert
    assert np.array_equal(solver.read_positions().reshape(-1, 4)[:, :3], scene.positions.reshape(-1, 4)[:, :3])
    # Frame 2: at rest, the first iteration converges and latches.
    for o in (solver, oracle, oracle64):
        o.step(REVERT_SETTINGS)
    ctl, hist = compare_frame(solver, oracle, oracle64, "frame 2")
    assert len(hist) == 1 and hist[0, 6, 0] == 1 and int(ctl[0, 3]) & 16
    # Frame 3: gravity switched on; converges after three rows.
    g = dict(REVERT_SETTINGS, gravity=(0.0, -9.81, 0.0))
    for o in (solver, oracle, oracle64):
        o.step(g)
    ctl, hist = compare_frame(solver, oracle, oracle64, "frame 3")
    assert len(hist) == 3 and (hist[:, 6, 0] == [0, 0, 1]).all() and int(ctl[0, 3]) & 16
    solver.destroy()


@pytest.mark.parametrize("api", ["step_frames_exact_iterations", "step_frames"])
def test_history_rows_match_oracle_through_frames(api):
    # The 425-vertex cantilever (E = 1e5 Pa), where parallel Jacobi is well
    # conditioned over these frames (FP32 and FP64 oracles within 1e-7).
    scene = scenes.cantilever_beam()
    settings = dict(scenes.CANTILEVER_SETTINGS, iterations=8)
    print(f"case: cantilever 16x4x4 history rows through {api}", flush=True)
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    oracle, oracle64 = O.Oracle(scene, 32), O.Oracle(scene, 64)
    for frame in range(3):
        if api == "step_frames":
            solver.step_frames(1, settings)
            oracle.step(settings)
            oracle64.step(settings)
        else:
            solver.step_frames_exact_iterations(1, 8, settings)
            oracle.step_frames(1, 8, settings)
            oracle64.step_frames(1, 8, settings)
        compare_frame(solver, oracle, oracle64, f"frame {frame}")
    solver.destroy()
