"""Batched scenes on the GPU: every scene of a mixed batch (tetrahedral
scenes of different sizes and materials, the cloth drape, a contact
trough) must produce exactly the results it produces alone, through
admission and retirement between submissions, under hipGraph replay and
under every local-solve launch shape.

Bar: bit-exact (np.array_equal) positions, velocities, vertex frames,
local diagnostics, line-search control and history. Each scene is
independent and its reductions run over its own blocks in the same order
as alone, so no tolerance applies.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
from jgs2.batch import BatchCapacity, JGS2Batch, footprint

pytestmark = pytest.mark.gpu

# One settings set for the whole mix; IPC parameters are fixed at create.
SETTINGS = dict(scenes.TROUGH_SETTINGS)
IT = 3


def mix():
    return [
        scenes.cantilever_beam(cells=(6, 2, 2), size=(1.5, 0.5, 0.5)),
        scenes.trough(box_count=3),
        scenes.layered_box(cells=(5, 6, 9), size=(0.5, 0.6, 0.9), seed=11, layer_cells=3),
        scenes.cloth_drape(),
        scenes.cantilever_beam(cells=(3, 3, 7), size=(0.6, 0.6, 1.4), young=3e5, poisson=0.3),
    ]


def capacity_for(inputs, slack_scenes=0):
    fp = [footprint(x) for x in inputs]
    total = {k: sum(f[k] for f in fp) for k in fp[0]}
    return BatchCapacity(max_vertices=total["vertices"], max_tets=total["tets"], max_scenes=len(inputs) + slack_scenes,
                         cubature_k=max(x.cubature_k for x in inputs), max_triangles=total["triangles"],
                         max_hinges=total["hinges"], max_candidates=total["candidates"], max_bodies=total["bodies"],
                         max_surface_triangles=total["surface_triangles"], max_surface_edges=total["surface_edges"])


def state(solver):
    return {
        "positions": solver.read_positions(),
        "velocities": solver.read_velocities(),
        "rotations": solver.read_vertex_rotations(),
        "diagnostics": solver.read_local_diagnostics(),
        "control": solver.read_globalization_control(),
        "history": solver.read_globalization_history(),
    }


def assert_same(a, b, label):
    for key in a:
        bad = np.flatnonzero(a[key] != b[key])
        assert bad.size == 0, (f"{label}: {key} differs at {bad.size} entries, first {bad[:8]} "
                               f"(max {np.abs(a[key] - b[key]).max()})")


def run_alone(scene, frames, settings, exact):
    solver = jgs2.JGS2GpuSolver.create(scene, SETTINGS)
    for _ in range(frames):
        if exact:
            solver.step_exact_iterations(IT, settings)
        else:
            solver.step(settings)
    out = state(solver)
    solver.destroy()
    return out


def step_batch(batch, frames, settings, exact):
    for _ in range(frames):
        if exact:
            batch.step_frames_exact_iterations(1, IT, settings)
        else:
            batch.step_frames(1, settings)


@pytest.mark.parametrize("schedule,exact", [("jacobi", True), ("jacobi", False),
                                            ("graph-colored-gauss-seidel", True)])
def test_mixed_batch_matches_each_scene_alone(schedule, exact):
    inputs = mix()
    settings = dict(SETTINGS, schedule=schedule, iterations=IT)
    frames = 4
    print(f"case: mixed batch of {len(inputs)} scenes, {schedule}, exact={exact}", flush=True)
    batch = JGS2Batch.create(capacity_for(inputs), SETTINGS)
    members = [batch.admit(x) for x in inputs]
    info = batch.info()
    assert info["scene_count"] == len(inputs)
    step_batch(batch, frames, settings, exact)
    for i, (x, m) in enumerate(zip(inputs, members)):
        assert_same(state(m), run_alone(x, frames, settings, exact), f"scene {i}")
    batch.destroy()


def test_admission_and_retirement_between_submissions():
    inputs = mix()
    late = scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5), young=2e5)
    settings = dict(SETTINGS, iterations=IT)
    print("case: retire scene 2 after 2 frames, admit a new scene into its blocks", flush=True)
    batch = JGS2Batch.create(capacity_for(inputs), SETTINGS)
    members = [batch.admit(x) for x in inputs]
    step_batch(batch, 2, settings, True)
    extent = batch.info()["vertex_extent"]
    batch.retire(members[2])
    newcomer = batch.admit(late)
    assert batch.info()["vertex_extent"] == extent  # first fit reused the freed blocks
    step_batch(batch, 3, settings, True)
    for i in (0, 1, 3, 4):
        assert_same(state(members[i]), run_alone(inputs[i], 5, settings, True), f"scene {i}")
    assert_same(state(newcomer), run_alone(late, 3, settings, True), "newcomer")
    with pytest.raises(jgs2.JGS2StateError):
        members[2].read_positions()
    batch.destroy()


def test_full_arena_rejects_admission_and_stays_usable():
    inputs = mix()[:2]
    settings = dict(SETTINGS, iterations=IT)
    print("case: admission into a full arena", flush=True)
    batch = JGS2Batch.create(capacity_for(inputs), SETTINGS)
    members = [batch.admit(x) for x in inputs]
    before = batch.info()
    with pytest.raises(jgs2.JGS2MemoryError):
        batch.admit(scenes.cantilever_beam(cells=(3, 2, 2), size=(0.75, 0.5, 0.5)))
    assert batch.info() == before
    step_batch(batch, 2, settings, True)
    assert_same(state(members[0]), run_alone(inputs[0], 2, settings, True), "scene 0")
    batch.retire(members[0])
    members[0] = batch.admit(scenes.cantilever_beam(cells=(3, 2, 2), size=(0.75, 0.5, 0.5)))
    assert batch.info()["scene_count"] == 2
    batch.destroy()


def test_graph_replays_match_eager_frames():
    inputs = mix()
    settings = dict(SETTINGS, iterations=IT)
    print("case: hipGraph frame replays vs eager frames", flush=True)
    eager = JGS2Batch.create(capacity_for(inputs), SETTINGS)
    graphed = JGS2Batch.create(capacity_for(inputs), SETTINGS)
    em = [eager.admit(x) for x in inputs]
    gm = [graphed.admit(x) for x in inputs]
    graphed.use_graphs(True)
    tip = inputs[0].extras["target_vertex"]
    for frame in range(6):
        if frame == 3:  # an objective change rebuilds the graph
            for m in (em[0], gm[0]):
                m.set_quadratic_target(tip, (1.6, 0.3, 0.1), 3e4)
        eager.step_frames_exact_iterations(1, IT, settings)
        graphed.step_frames_exact_iterations(1, IT, settings)
        eager.sync()
        graphed.sync()
        for i in range(len(inputs)):
            assert np.array_equal(em[i].read_positions(), gm[i].read_positions()), (frame, i)
    for i in range(len(inputs)):
        assert_same(state(gm[i]), state(em[i]), f"scene {i}")
    eager.destroy()
    graphed.destroy()


def test_recapture_between_queued_replays_matches_eager():
    # Two settings alternate between submissions with no synchronization in
    # between, so every submission recaptures the frame graph while the
    # previous one's replays may still be queued on the stream.
    inputs = mix()[:3]
    a = dict(SETTINGS, iterations=IT)
    b = dict(SETTINGS, iterations=IT, gravity=(0.0, -4.0, 1.0))
    print("case: alternating overrides, async graph recaptures", flush=True)
    eager = JGS2Batch.create(capacity_for(inputs), SETTINGS)
    graphed = JGS2Batch.create(capacity_for(inputs), SETTINGS)
    em = [eager.admit(x) for x in inputs]
    gm = [graphed.admit(x) for x in inputs]
    graphed.use_graphs(True)
    for k in range(8):
        st = a if k % 2 == 0 else b
        frames = 1 + k % 3
        eager.step_frames_exact_iterations(frames, IT, st)
        graphed.step_frames_exact_iterations(frames, IT, st)
    for i in range(len(inputs)):
        assert_same(state(gm[i]), state(em[i]), f"scene {i}")
    eager.destroy()
    graphed.destroy()


def test_launch_shapes_and_autotune_do_not_change_results():
    inputs = mix()
    settings = dict(SETTINGS, iterations=IT)
    results = []
    for shape in [(256, True), (64, False), (128, True), None]:
        print(f"case: launch shape {shape}", flush=True)
        batch = JGS2Batch.create(capacity_for(inputs), SETTINGS)
        members = [batch.admit(x) for x in inputs]
        if shape is None:
            with pytest.raises(jgs2.JGS2StateError, match="at least one frame"):
                batch.autotune(settings, repeats=5)
            step_batch(batch, 1, settings, True)
            ms = batch.autotune(settings, repeats=5)
            info = batch.info()
            print(f"autotune: block {info['solve_block']} remap {info['xcd_remap']} {ms:.4f} ms", flush=True)
            assert info["solve_block"] in (64, 128, 256) and ms > 0
        else:
            batch.set_launch(*shape)
            step_batch(batch, 1, settings, True)
        step_batch(batch, 2, settings, True)
        results.append([m.read_positions() for m in members])
        batch.destroy()
    for r in results[1:]:
        for a, b in zip(r, results[0]):
            assert np.array_equal(a, b)


def test_batch_rejects_mismatched_scenes():
    print("case: batch compatibility errors", flush=True)
    small = scenes.cantilever_beam(cells=(2, 1, 1), size=(0.5, 0.25, 0.25))
    inputs = [mix()[0], small]
    batch = JGS2Batch.create(capacity_for(inputs, slack_scenes=1), SETTINGS)
    first = batch.admit(inputs[0])  # a dropped handle retires its scene
    with pytest.raises(jgs2.JGS2RangeError, match="co-rotated"):
        batch.admit(scenes.cantilever_beam(cells=(2, 1, 1), size=(0.5, 0.25, 0.25), model=0))
    assert batch.info()["scene_count"] == 1
    member = batch.admit(small)
    assert first.vertex_count == inputs[0].vertex_count
    with pytest.raises(jgs2.JGS2StateError):
        member.step()
    batch.destroy()


def test_served_requests_match_each_scene_alone():
    from jgs2.serving import BatchServer, Request, TickClock

    pool = mix()
    rng = np.random.default_rng(5)
    reqs = [Request(id=i, scene=pool[int(rng.integers(0, len(pool)))], frames=int(rng.integers(1, 5)),
                    arrival=float(rng.integers(0, 6))) for i in range(12)]
    # Room for three of the largest scenes at once: later arrivals wait for
    # retirements and reuse their blocks.
    big = max(pool, key=lambda x: x.tet_count)
    cap = capacity_for([big, big, big] + [p for p in pool if p.contact is not None or p.cloth is not None])
    print(f"case: served stream of {len(reqs)} requests through a {cap.max_scenes}-slot arena", flush=True)
    server = BatchServer(cap, SETTINGS, IT, use_graphs=True, keep_results=True, clock=TickClock())
    rep = server.run(reqs)
    server.close()
    assert len(rep.completed) == len(reqs)
    assert rep.rejected_admissions > 0  # some requests had to wait for room
    settings = dict(SETTINGS, iterations=IT)
    for r in rep.completed:
        alone = run_alone(r.scene, r.frames, settings, True)["positions"]
        assert np.array_equal(r.result, alone), r.id
