"""One GPU test per benchmark configuration.

The configurations (the 425-vertex cantilever, the layered box, the
contact boxes, the batched uneven stream and the served stream) are run as
configured, or at a reduced size where the FP64 oracle must follow them,
with the bar each one admits. What the oracle itself shows is part of each
bar, because FP32 arithmetic (the reference's) cannot follow FP64 past the
point where the reference's iteration amplifies rounding:

* cantilever 16x4x4, parallel Jacobi, 8 exact iterations, 60 frames, the
  scripted tip target at the reference's 3e4 N/m: the FP32 and FP64 oracles
  agree to ~1e-5 of the displacement for five frames and part by ~1e-1 by
  frame 11 (an unstable mode of parallel Jacobi on this beam, gravity
  alone). Bar: FP64 parity to 1e-4 through frame 5; over all 60 frames
  finite positions and an accepted minimum determinant above 1e-4 in every
  iteration (the assembled accept/revert gate).
* layered box (alternating 1e4 / 1e6 Pa slabs, jittered, z = 0 pinned) at
  16x16x32 cells: graph-colored Gauss-Seidel within 1e-4 of the FP64
  oracle over 3 frames (FP32 oracle 8e-6); parallel Jacobi, which reverts
  most iterations at this stiffness ratio (14 of 24 in the oracle), with
  the FP32 oracle's accept decisions and within three times its distance
  from FP64. At the full 64x64x128 cells (545,025 vertices): finite, the
  accept gate in every iteration, Gauss-Seidel accepting every iteration.
* contact boxes (d_hat 1 mm, kappa 1e5, friction 0.3, 12 Jacobi
  iterations, device broad phase): 4 boxes of 2x2x2 cells within 1e-4 of
  the FP64 oracle (static superset) through the fall (24 frames); then
  through contact, and for the full 20 boxes of 8x8x8 cells over 240
  frames, no candidate pair at or below the minimum distance (independent
  distances) and an Armijo decrease in every accepted local solve.
* batched uneven stream: a sample of the generator (its first small
  tetrahedral scenes and folded self-contact sheets) in one batch, each
  scene within max(1e-3, 3 x the FP32 oracle's distance) of the FP64
  oracle's displacement over 12 frames, sheets interpenetration-free with
  Armijo decreases, and a scene bit-identical alone and in the batch.
* served stream: Poisson arrivals of contact-box requests through the
  continuous-batching server, every result bit-identical to the same scene
  stepped alone for its frames.
"""

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
from jgs2.batch import JGS2Batch
from jgs2.solver import ContactData
import jgs2_oracle as O
from gpu_compare import rel_l2, xyz
from ipc_geometry import candidate_distances
from test_gpu_batch import capacity_for

pytestmark = pytest.mark.gpu

DETERMINANT_FLOOR = 1e-4


def accept_gate_holds(solver, iterations):
    h = solver.read_globalization_history().reshape(64, 8, 4)[:iterations]
    return bool(np.all(h[:, 0, 2] > DETERMINANT_FLOOR)), int(h[:, 1, 1].sum())


def armijo_holds(solver):
    diag = solver.read_local_diagnostics().reshape(-1, 4, 4)
    ok = diag[:, 3, 3] == 0
    return bool(np.all(diag[ok, 2, 0] <= diag[ok, 2, 1]))


def set_target(solver, oracles, scene, objectives, frame, stiffness):
    v = scene.extras["target_vertex"]
    active, target, _ = scenes.scripted_target_at_frame(scene, frame)
    k = stiffness if active else 0.0
    if objectives[v, 7] != k or (active and np.any(objectives[v, 4:7] != np.float32(target))):
        objectives[v, 4:7] = np.float32(target)
        objectives[v, 7] = k
        solver.set_quadratic_target(v, target, k)
        for o in oracles:
            o.set_objectives(objectives.reshape(-1))


def test_c1_cantilever_as_configured():
    beam = scenes.cantilever_beam()
    settings = dict(scenes.CANTILEVER_SETTINGS)
    it = scenes.CANTILEVER_ITERATIONS
    print("case: C1 cantilever 16x4x4, Jacobi, 8 iterations, 60 frames, target 3e4", flush=True)
    solver = jgs2.JGS2GpuSolver.create(beam, settings)
    o64, o32 = O.Oracle(beam, 64), O.Oracle(beam, 32)
    obj = beam.packed_objectives().reshape(-1, 8).copy()
    x0 = xyz(beam.positions)
    reverts = 0
    for frame in range(scenes.CANTILEVER_FRAMES):
        set_target(solver, (o64, o32) if frame < 12 else (), beam, obj, frame, scenes.TARGET_STIFFNESS)
        solver.step_exact_iterations(it, settings)
        gate, rev = accept_gate_holds(solver, it)
        assert gate, frame
        reverts += rev
        if frame < 12:
            o64.step_frames(1, it, settings)
            o32.step_frames(1, it, settings)
            x = xyz(solver.read_positions())
            e = rel_l2(x - x0, o64.positions() - x0)
            e32 = rel_l2(o32.positions() - x0, o64.positions() - x0)
            if frame in (5, 11):
                print(f"frame {frame}: displacement vs FP64 {e:.2e} (FP32 oracle {e32:.2e})", flush=True)
            if frame <= 5:
                assert e < 1e-4
    x = xyz(solver.read_positions())
    print(f"60 frames: {reverts} reverted iterations, tip at {x[beam.extras['target_vertex']]}", flush=True)
    assert np.isfinite(x).all()
    solver.destroy()


def oracle_positions(scene, precision, frames, iterations, settings, colors=None):
    o = O.Oracle(scene, precision, colors=colors, settings=settings)
    o.step_frames(frames, iterations, settings)
    return o


def test_c2_layered_box_reduced_against_fp64_and_full_size_invariants():
    small = scenes.layered_box(cells=(16, 16, 32))
    it = scenes.LAYERED_BOX_ITERATIONS
    x0 = xyz(small.positions)
    print(f"case: C2 layered box 16x16x32 (V={small.vertex_count}), GS and Jacobi, 3 frames", flush=True)
    gs = dict(scenes.LAYERED_BOX_SETTINGS, schedule="graph-colored-gauss-seidel")
    solver = jgs2.JGS2GpuSolver.create(small, gs)
    colors = solver.read_schedule_colors()
    solver.step_frames_exact_iterations(3, it, gs)
    ref = oracle_positions(small, 64, 3, it, gs, colors)
    e = rel_l2(xyz(solver.read_positions()) - x0, ref.positions() - x0)
    print(f"GS: displacement vs FP64 {e:.2e}", flush=True)
    assert e < 1e-4
    assert np.all(solver.read_globalization_history().reshape(64, 8, 4)[:it, 1, 0] == 1)
    solver.destroy()
    jac = dict(scenes.LAYERED_BOX_SETTINGS)
    solver = jgs2.JGS2GpuSolver.create(small, jac)
    o32, o64 = O.Oracle(small, 32), O.Oracle(small, 64)
    for frame in range(3):
        solver.step_exact_iterations(it, jac)
        o32.step_frames(1, it, jac)
        o64.step_frames(1, it, jac)
        h = solver.read_globalization_history().reshape(64, 8, 4)[:it]
        assert np.array_equal(h[:, 1, 0], o32.history()[:it, 1, 0]), frame
        assert np.all(h[:, 0, 2] > DETERMINANT_FLOOR)
    e = rel_l2(xyz(solver.read_positions()) - x0, o64.positions() - x0)
    e32 = rel_l2(o32.positions() - x0, o64.positions() - x0)
    print(f"Jacobi: displacement vs FP64 {e:.2e} (FP32 oracle {e32:.2e}), reverts "
          f"{int(o32.history()[:it, 1, 1].sum())} in the last frame", flush=True)
    assert e <= max(1e-4, 3.0 * e32)
    solver.destroy()
    full = scenes.layered_box()
    print(f"case: C2 layered box 64x64x128 (V={full.vertex_count}, T={full.tet_count}), 3 frames each", flush=True)
    for settings in (jac, gs):
        solver = jgs2.JGS2GpuSolver.create(full, settings)
        accepted = 0
        for _ in range(3):
            solver.step_exact_iterations(it, settings)
            gate, rev = accept_gate_holds(solver, it)
            assert gate
            accepted += it - rev
        x = xyz(solver.read_positions())
        print(f"{settings['schedule']}: {accepted} of {3 * it} iterations accepted, largest motion "
              f"{np.abs(x - xyz(full.positions)).max():.4f} m", flush=True)
        assert np.isfinite(x).all()
        if settings is gs:
            assert accepted == 3 * it
        solver.destroy()


def nearest_listed_pair(solver):
    packed, vt, _ = solver.read_contact_candidates()
    if len(packed) == 0:
        return np.inf
    d = candidate_distances(xyz(solver.read_positions()), ContactData(packed.reshape(-1), vt, len(packed) - vt))
    return float(d.min())


def test_c3_contact_boxes_reduced_and_full():
    settings = dict(scenes.CONTACT_BOXES_SETTINGS)
    it = scenes.CONTACT_BOXES_ITERATIONS
    dmin = settings["ipc_minimum_distance"]
    small = scenes.trough(box_count=4, box_cells=2, seed=7, broad_phase=True, settings=settings)
    superset = scenes.trough(box_count=4, box_cells=2, seed=7)
    print(f"case: C3 reduced, 4 boxes of 2x2x2 cells, {superset.contact.candidate_count} superset pairs", flush=True)
    solver = jgs2.JGS2GpuSolver.create(small, settings)
    o64 = O.Oracle(superset, 64, settings=settings)
    x0 = xyz(small.positions)
    for frame in range(60):
        solver.step_exact_iterations(it, settings)
        if frame < 24:
            o64.step_frames(1, it, settings)
        if frame == 23:
            e = rel_l2(xyz(solver.read_positions()) - x0, o64.positions() - x0)
            print(f"24 frames (falling): displacement vs FP64 {e:.2e}", flush=True)
            assert e < 1e-4
        d = candidate_distances(xyz(solver.read_positions()), superset.contact).min()
        assert d > dmin, frame
        assert armijo_holds(solver), frame
    print(f"60 frames: nearest pair {d:.3e} m (minimum {dmin:g})", flush=True)
    assert d < 2e-3  # the boxes are in contact
    solver.destroy()

    full = scenes.contact_boxes()
    print(f"case: C3 full, 20 boxes of 8x8x8 cells (V={full.vertex_count}), 240 frames", flush=True)
    solver = jgs2.JGS2GpuSolver.create(full, settings)
    JGS2Batch.of(solver).use_graphs(True)
    nearest, nearest_device = np.inf, np.inf
    for frame in range(240):
        solver.step_exact_iterations(it, settings)
        d = solver.read_oracle_diagnostics(settings)
        nearest_device = min(nearest_device, d["minimum_contact_distance"])
        assert d["minimum_contact_distance"] > dmin and d["finite"], frame
        assert armijo_holds(solver), frame
        if frame % 60 == 59:
            n = nearest_listed_pair(solver)
            nearest = min(nearest, n)
            print(f"frame {frame}: nearest listed pair {n:.3e} m (FP64 device {d['minimum_contact_distance']:.3e}), "
                  f"{solver.contact_info()['builds']} builds", flush=True)
            assert n > dmin
    assert nearest_device < 5e-3  # the boxes reached the trough
    solver.destroy()


def test_c4_batched_stream_sample_against_fp64():
    settings = dict(scenes.MIXED_SETTINGS)
    it, frames = scenes.MIXED_ITERATIONS, 12
    tet_specs = [s for s in scenes.mixed_scene_specs(count=512)
                 if s["kind"] == "tet" and 6 * int(np.prod(s["cells"])) <= 3000][:6]
    specs = tet_specs + [{"kind": "cloth", "cells": n} for n in (6, 8)]
    gpu_inputs = [scenes.build_mixed_scene(s, "broad_phase") for s in specs]
    ref_inputs = [scenes.build_mixed_scene(s, "superset") for s in specs]
    print(f"case: C4 sample, {len(specs)} scenes in one batch, {frames} frames", flush=True)
    batch = JGS2Batch.create(capacity_for(gpu_inputs), settings)
    members = [batch.admit(x) for x in gpu_inputs]
    for _ in range(frames):
        batch.step_frames_exact_iterations(1, it, settings)
        for s, m in zip(specs, members):
            if s["kind"] == "cloth":
                assert nearest_listed_pair(m) > settings["ipc_minimum_distance"]
                assert armijo_holds(m)
    gpu = [xyz(m.read_positions()) for m in members]
    cloth_member = members[-1].read_positions()
    batch.destroy()

    def run(args):
        x, p = args
        return oracle_positions(x, p, frames, it, settings).positions()

    with ThreadPoolExecutor(8) as pool:
        r64 = list(pool.map(run, [(x, 64) for x in ref_inputs]))
        r32 = list(pool.map(run, [(x, 32) for x in ref_inputs]))
    for s, x, g, a, b in zip(specs, gpu_inputs, gpu, r64, r32):
        x0 = xyz(x.positions)
        e, e32 = rel_l2(g - x0, a - x0), rel_l2(b - x0, a - x0)
        print(f"{s['kind']} {s.get('cells')}: displacement vs FP64 {e:.2e} (FP32 oracle {e32:.2e})", flush=True)
        assert e <= max(1e-3, 3.0 * e32)
    alone = jgs2.JGS2GpuSolver.create(gpu_inputs[-1], settings)
    alone.step_frames_exact_iterations(frames, it, settings)
    assert np.array_equal(alone.read_positions(), cloth_member)
    alone.destroy()


def test_c5_poisson_stream_results_equal_each_scene_alone():
    from jgs2.serving import BatchServer, Request, TickClock, poisson_arrivals

    settings = dict(scenes.CONTACT_BOXES_SETTINGS)
    it = scenes.CONTACT_BOXES_ITERATIONS
    pool = [scenes.trough(box_count=3, box_cells=2, seed=s, broad_phase=True, settings=settings) for s in (1, 2, 3, 4)]
    arrivals = poisson_arrivals(2000.0, 0.012, 5)  # ~24 requests, in ticks of the test clock below
    rng = np.random.default_rng(5)
    reqs = [Request(id=i, scene=pool[int(rng.integers(0, len(pool)))], frames=int(rng.integers(1, 9)),
                    arrival=float(t * 1000.0)) for i, t in enumerate(arrivals)]
    cap = capacity_for(pool[:2] * 2)
    print(f"case: C5 served stream, {len(reqs)} Poisson requests, arena for {cap.max_scenes} scenes", flush=True)
    server = BatchServer(cap, settings, it, use_graphs=True, keep_results=True, clock=TickClock())
    rep = server.run(reqs)
    server.close()
    assert len(rep.completed) == len(reqs)
    assert rep.rejected_admissions > 0
    cache = {}
    for r in rep.completed:
        key = (id(r.scene), r.frames)
        if key not in cache:
            solver = jgs2.JGS2GpuSolver.create(r.scene, settings)
            solver.step_frames_exact_iterations(r.frames, it, settings)
            cache[key] = solver.read_positions()
            solver.destroy()
        assert np.array_equal(r.result, cache[key]), r.id
    s = rep.summary()
    print(f"served: {s['requests_completed']} requests, {s['ticks']} ticks, queueing p50 {s['queueing_s']['p50']} "
          f"p99 {s['queueing_s']['p99']} ticks", flush=True)
