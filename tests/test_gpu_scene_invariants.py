"""Scene invariants of the reference's canonical fixtures on the GPU.

The fixtures and their expected metrics are the reference's manifest
(manifests/canonical-scenes.v1.json, kept as tests/golden/canonical-scenes.v1.json);
the metric definitions follow the reference's e2e spec
(tests/e2e/jgs2.spec.ts:480-575 for the momentum errors):

* phase0.force-free-cuboid: a 2x2x2-cell cuboid rotated by [0.17, -0.23,
  0.31] rad with linear velocity [0.37, -0.11, 0.23] and angular velocity
  [0.41, 0.29, -0.19], no gravity, floor, damping or pins, dt = 1/120, 12
  Jacobi iterations, 1200 frames: |P_end - P_0| / max(|P_0|, M s) and
  |L_end - L_0| / max(|L_0|, M s^2) <= 5e-3 (s = initial bounds diagonal),
  with the fixture's exact-all-elements cubature (every vertex's exact
  whole-mesh basis over all 48 tetrahedra, weight 1); the GPU's drift also
  equals the FP64 oracle's at 100 frames.
* baseline.trough at frame 90: pinned error <= 2e-5, minimum IPC distance
  > 3e-3 (independent distances over the static superset), every box's
  lowest point above the trough apex.
* baseline.cloth at frame 48: pinned error <= 2e-5, smallest triangle area
  over its rest area > 0.1 (the drape here has no contact candidates, so
  the fixture's IPC distance metric does not apply).
The trough and the drape use this build's stand-in cubature records (K per
vertex), not the reference's trained selection.
"""

import json
from pathlib import Path

import numpy as np
import pytest

import jgs2
from jgs2 import manifests, scenes
from ipc_geometry import candidate_distances
import jgs2_oracle as O

pytestmark = pytest.mark.gpu

MANIFEST = json.loads((Path(__file__).parent / "golden" / "canonical-scenes.v1.json").read_text())


def fixture(fid):
    return next(f for f in MANIFEST["fixtures"] if f["id"] == fid)


def metric(f, mid):
    return next(m["value"] for m in f["expectedMetrics"] if m["id"] == mid)


def xyz(p):
    return np.asarray(p, np.float64).reshape(-1, 4)[:, :3]


def momenta(x, v, m):
    c = (m[:, None] * x).sum(0) / m.sum()
    return (m[:, None] * v).sum(0), (m[:, None] * np.cross(x - c, v)).sum(0)


def step(solver, frames, iterations, settings):
    per = max(1, 2048 // iterations)  # frames per submission within the globalized-iteration limit
    done = 0
    while done < frames:
        n = min(per, 240, frames - done)
        solver.step_frames_exact_iterations(n, iterations, settings)
        done += n
    solver.await_idle()


def test_force_free_cuboid_momentum():
    f = fixture("phase0.force-free-cuboid")
    print("case: force-free cuboid, 1200 frames", flush=True)
    x, settings, it, _ = manifests.force_free_cuboid(MANIFEST)
    p0 = xyz(x.positions)
    vel = xyz(x.velocities)
    solver = jgs2.JGS2GpuSolver.create(x, settings)
    mass = x.vertex_rest.reshape(-1, 4)[:, 3].astype(np.float64)
    P0, L0 = momenta(p0, vel, mass)
    scale = float(np.linalg.norm(p0.max(0) - p0.min(0)))
    M = mass.sum()

    def errors(xe, ve):
        P1, L1 = momenta(xe, ve, mass)
        return (np.linalg.norm(P1 - P0) / max(np.linalg.norm(P0), M * scale),
                np.linalg.norm(L1 - L0) / max(np.linalg.norm(L0), M * scale * scale))

    step(solver, 100, it, settings)
    gpu100 = errors(xyz(solver.read_positions()), xyz(solver.read_velocities()))
    ref = O.Oracle(x, 64, settings=settings)
    ref.step_frames(100, it, settings)
    ref100 = errors(np.asarray(ref.positions()).reshape(-1, 3), np.asarray(ref.velocities()).reshape(-1, 3))
    print(f"100 frames: GPU linear/angular {gpu100[0]:.4e}/{gpu100[1]:.4e}, FP64 oracle {ref100[0]:.4e}/{ref100[1]:.4e}",
          flush=True)
    assert abs(gpu100[0] - ref100[0]) <= 0.02 * ref100[0] + 1e-6
    assert abs(gpu100[1] - ref100[1]) <= 0.02 * ref100[1] + 1e-6
    frames = f["sampledFrames"][-1]["frame"]
    step(solver, frames - 100, it, settings)
    xe, ve = xyz(solver.read_positions()), xyz(solver.read_velocities())
    assert np.isfinite(xe).all() and np.isfinite(ve).all()
    lin, ang = errors(xe, ve)
    travel = np.linalg.norm(xe.mean(0) - p0.mean(0))
    print(f"{frames} frames: linear momentum error {lin:.3e}, angular {ang:.3e}, centre travel {travel:.3f} m", flush=True)
    assert travel > 1.0  # the body moved (0.45 m/s for 10 s)
    assert lin <= metric(f, "linear-momentum-error")
    assert ang <= metric(f, "angular-momentum-error")
    solver.destroy()


def test_trough_invariants_at_frame_90():
    f = fixture("baseline.trough")
    print("case: canonical trough, 90 frames", flush=True)
    x = scenes.trough()
    settings = dict(scenes.TROUGH_SETTINGS)
    it = f["solver"]["iterations"]
    frame = f["sampledFrames"][-1]["frame"]
    solver = jgs2.JGS2GpuSolver.create(x, settings)
    step(solver, frame, it, settings)
    p = xyz(solver.read_positions())
    info = x.vertex_info.reshape(-1, 4)
    rest = xyz(x.vertex_rest)
    pinned = info[:, 2] != 0
    assert np.isfinite(p).all()
    pin_err = float(np.abs(p[pinned] - rest[pinned]).max())
    dist = float(candidate_distances(p, x.contact).min())
    apex = 0.05  # the trough's top surface at x = 0 (warp y = 0.05 + 0.65|x| - 0.18 (1 - source y))
    lows = [float(p[info[:, 3] == b, 1].min()) for b in range(1, int(info[:, 3].max()) + 1)]
    print(f"pinned error {pin_err:.2e}, nearest pair {dist:.4f} m, lowest box point {min(lows):.4f} m", flush=True)
    assert pin_err <= metric(f, "pinned-max-error")
    assert dist > metric(f, "minimum-ipc-distance")
    assert min(lows) > apex
    solver.destroy()


def test_cloth_invariants_at_frame_48():
    f = fixture("baseline.cloth")
    print("case: canonical cloth drape, 48 frames", flush=True)
    x = scenes.cloth_drape()
    sim = f["simulation"]
    settings = dict(scenes.CLOTH_SETTINGS, timestep=sim["timestep"], parity_mode=sim["parityMode"])
    it = f["solver"]["iterations"]
    frame = f["sampledFrames"][-1]["frame"]
    solver = jgs2.JGS2GpuSolver.create(x, settings)
    step(solver, frame, it, settings)
    p = xyz(solver.read_positions())
    info = x.vertex_info.reshape(-1, 4)
    rest = xyz(x.vertex_rest)
    pinned = info[:, 2] != 0
    tri = x.cloth.triangle_indices.reshape(-1, 3).astype(np.int64)
    area = 0.5 * np.linalg.norm(np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]]), axis=1)
    ratio = float((area / x.cloth.triangle_rest_areas.astype(np.float64)).min())
    pin_err = float(np.abs(p[pinned] - rest[pinned]).max())
    moved = float(np.abs(p - xyz(x.positions)).max())
    print(f"pinned error {pin_err:.2e}, smallest area ratio {ratio:.3f}, largest motion {moved:.3f} m", flush=True)
    assert np.isfinite(p).all() and moved > 0.01
    assert pin_err <= metric(f, "pinned-max-error")
    assert ratio > metric(f, "minimum-area-ratio")
    solver.destroy()
