"""Cloth with IPC contact on the GPU: the reference's cloth scene is an IPC
scene (isIpcSceneId, src/scenes/index.ts:113-115) whose candidates are the
static VT/EE superset of the collider's boundary plus the cloth triangles
and edges (src/scenes/index.ts:1098-1101, surfaceWithCloth
src/simulation/cpu/precompute.ts:98-133), stepped with the app's cloth
settings (src/App.tsx:521-572: activation distance and maxStep 0.01,
friction 0.3, barrier stiffness 1e5, minimum distance 3 mm, dt = 1/120,
5 Jacobi iterations). This exercises the cloth-and-contact variant of every
pass (membrane, bending and barrier rows together in one local solve).

Bars: local g and H against both oracles (the local-system bar of
tests/gpu_compare.py); positions after one frame within 1e-5 relative L2 of
both oracles; 10 Jacobi frames within 1e-3 of the FP64 oracle's
displacement with the same accept decisions; the canonical cloth fixture's
metrics at frame 48 (pinned error <= 2e-5, smallest area ratio > 0.1,
nearest candidate pair > 3 mm, from independent distances) with no pair
ever at or below the minimum distance; the device broad phase over the same
surface steps bit-identically to the superset.
"""

import json
from pathlib import Path

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
import jgs2_oracle as O
from gpu_compare import check_local_systems, rel_l2, xyz
from ipc_geometry import candidate_distances

pytestmark = pytest.mark.gpu

SETTINGS = scenes.CLOTH_IPC_SETTINGS
IT = scenes.CLOTH_ITERATIONS
MANIFEST = json.loads((Path(__file__).parent / "golden" / "canonical-scenes.v1.json").read_text())


@pytest.fixture(scope="module")
def drape():
    return scenes.cloth_drape(contact="superset")


def test_cloth_contact_local_systems_match_oracle(drape):
    print(f"case: cloth + IPC local systems, {drape.contact.candidate_count} candidates", flush=True)
    solver = jgs2.JGS2GpuSolver.create(drape, SETTINGS)
    solver.step_frames_exact_iterations(20, IT, SETTINGS)
    g_gpu, h_gpu = solver.debug_local_systems(SETTINGS)
    pos, pred = solver.read_positions(), solver.read_predicted_positions()
    d = candidate_distances(xyz(pos), drape.contact)
    print(f"nearest pair {d.min():.4f} m", flush=True)
    assert d.min() < SETTINGS["ipc_activation_distance"]  # barrier rows are active
    systems = {}
    for precision in (32, 64):
        o = O.Oracle(drape, precision, settings=SETTINGS)
        o.set_pose(pos, pred)
        systems[precision] = o.local_systems(SETTINGS)
    free = drape.vertex_info.reshape(-1, 4)[:, 2] == 0
    check_local_systems(g_gpu, h_gpu, systems, free, 1e-5, "cloth + IPC")
    solver.destroy()


def test_cloth_contact_frames_match_oracle(drape):
    print("case: cloth + IPC, 1 and 10 Jacobi frames", flush=True)
    solver = jgs2.JGS2GpuSolver.create(drape, SETTINGS)
    ref32, ref64 = O.Oracle(drape, 32, settings=SETTINGS), O.Oracle(drape, 64, settings=SETTINGS)
    x0 = xyz(drape.positions)
    for frame in range(10):
        solver.step_exact_iterations(IT, SETTINGS)
        ref32.step_frames(1, IT, SETTINGS)
        ref64.step_frames(1, IT, SETTINGS)
        x = xyz(solver.read_positions())
        hist = solver.read_globalization_history().reshape(64, 8, 4)
        assert np.array_equal(hist[:IT, 1, 0], ref64.history()[:IT, 1, 0]), frame
        if frame == 0:
            e32, e64 = rel_l2(x, ref32.positions()), rel_l2(x, ref64.positions())
            print(f"one frame: rel L2 vs f32 {e32:.2e}, vs f64 {e64:.2e}", flush=True)
            assert e32 < 1e-5 and e64 < 1e-5
    d = rel_l2(x - x0, ref64.positions() - x0)
    print(f"10 frames: displacement rel L2 vs f64 {d:.2e}", flush=True)
    assert d < 1e-3
    solver.destroy()


def test_cloth_fixture_with_contact_at_frame_48(drape):
    f = next(q for q in MANIFEST["fixtures"] if q["id"] == "baseline.cloth")
    bound = {m["id"]: m["value"] for m in f["expectedMetrics"]}
    frames = f["sampledFrames"][-1]["frame"]
    print(f"case: canonical cloth fixture with IPC, {frames} frames", flush=True)
    solver = jgs2.JGS2GpuSolver.create(drape, SETTINGS)
    nearest = np.inf
    for _ in range(frames):
        solver.step_exact_iterations(f["solver"]["iterations"], SETTINGS)
        p = xyz(solver.read_positions())
        nearest = min(nearest, float(candidate_distances(p, drape.contact).min()))
        assert nearest > SETTINGS["ipc_minimum_distance"]
    rest = xyz(drape.vertex_rest)
    pinned = drape.vertex_info.reshape(-1, 4)[:, 2] != 0
    tri = drape.cloth.triangle_indices.reshape(-1, 3).astype(np.int64)
    area = 0.5 * np.linalg.norm(np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]]), axis=1)
    ratio = float((area / drape.cloth.triangle_rest_areas.astype(np.float64)).min())
    pin_err = float(np.abs(p[pinned] - rest[pinned]).max())
    final = float(candidate_distances(p, drape.contact).min())
    moved = float(np.abs(p - xyz(drape.positions)).max())
    print(f"pinned error {pin_err:.2e}, area ratio {ratio:.3f}, nearest pair {final:.4f} m "
          f"(over all frames {nearest:.4f}), largest motion {moved:.3f} m", flush=True)
    assert np.isfinite(p).all() and moved > 0.01
    assert pin_err <= bound["pinned-max-error"]
    assert ratio > bound["minimum-area-ratio"]
    assert final > bound["minimum-ipc-distance"]
    solver.destroy()


def test_cloth_broad_phase_steps_like_the_superset(drape):
    print("case: cloth + IPC broad phase vs superset, 24 frames", flush=True)
    bp = scenes.cloth_drape(contact="broad_phase", settings=SETTINGS)
    a = jgs2.JGS2GpuSolver.create(drape, SETTINGS)
    b = jgs2.JGS2GpuSolver.create(bp, SETTINGS)
    for frame in range(24):
        a.step_exact_iterations(IT, SETTINGS)
        b.step_exact_iterations(IT, SETTINGS)
        assert np.array_equal(a.read_positions(), b.read_positions()), frame
    assert np.array_equal(a.read_velocities(), b.read_velocities())
    assert np.array_equal(a.read_local_diagnostics(), b.read_local_diagnostics())
    print(f"broad phase: {b.contact_info()}", flush=True)
    a.destroy()
    b.destroy()
