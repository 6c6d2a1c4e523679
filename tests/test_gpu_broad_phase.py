"""GPU broad phase: the device-built candidate list against the oracle, and
stepping with it against stepping the reference's static superset.

Bars: the list a broad-phase scene holds equals, pair for pair and in the
superset's canonical order, the oracle's restriction of the superset
(src/simulation/cpu/ipc-contact.ts:600-665) to pairs whose f32 AABB lower
bound (src/simulation/gpu/jgs2-shader.ts:525-550) at the build pose is within
the build's reach, on 1- to 20-box troughs, at admission and after frames;
positions, velocities and local diagnostics of a broad-phase trough equal
those of the same trough with the static superset bit for bit, frame by
frame, with the default reach margin and with a margin small enough to
rebuild inside frames (lagged friction and the contact cache carried over);
a broad-phase scene in a batch steps exactly as alone, with and without
graph replay; a capacity too small is an error, not a silent truncation.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
from jgs2.batch import JGS2Batch
import jgs2_oracle as O

from ipc_geometry import candidate_distances
from test_gpu_batch import capacity_for

pytestmark = pytest.mark.gpu

IT = scenes.TROUGH_ITERATIONS


def surface(x):
    s = x.extras["surface"]
    return s["triangles"], s["edges"]


def assert_list_matches_oracle(solver, x, positions):
    packed, vt, reach = solver.read_contact_candidates()
    tri, edg = surface(x)
    ref, ref_vt = O.contact_candidates_within(x.vertex_count, positions, tri, edg, reach)
    print(f"list {len(packed)} (VT {vt}) at reach {reach:.6f}; oracle {len(ref)} (VT {ref_vt})", flush=True)
    assert vt == ref_vt
    np.testing.assert_array_equal(packed, ref)
    return len(packed)


@pytest.mark.parametrize("boxes,cells,seed", [(1, 1, 0), (10, 1, 0), (20, 1, 3), (4, 2, 5), (20, 2, 1)])
def test_admission_list_equals_oracle(boxes, cells, seed):
    print(f"case: broad-phase list at admission, {boxes} boxes x {cells} cells, seed {seed}", flush=True)
    x = scenes.trough(box_count=boxes, box_cells=cells, seed=seed, broad_phase=True)
    solver = jgs2.JGS2GpuSolver.create(x, scenes.TROUGH_SETTINGS)
    n = assert_list_matches_oracle(solver, x, x.positions)
    assert n > 0
    solver.destroy()


def test_list_after_frames_equals_oracle():
    print("case: broad-phase list after frames, 20 boxes x 2 cells", flush=True)
    settings = dict(scenes.TROUGH_SETTINGS)
    x = scenes.trough(box_count=20, box_cells=2, seed=1, broad_phase=True)
    solver = jgs2.JGS2GpuSolver.create(x, settings)
    for _ in range(3):
        solver.step_frames_exact_iterations(10, IT, settings)
        solver.rebuild_contact_candidates()
        assert_list_matches_oracle(solver, x, solver.read_positions())
    solver.destroy()


def xyz(p):
    return np.asarray(p, np.float64).reshape(-1, 4)[:, :3]


def state(solver):
    return {
        "positions": solver.read_positions(),
        "velocities": solver.read_velocities(),
        "diagnostics": solver.read_local_diagnostics(),
        "history": solver.read_globalization_history(),
    }


def assert_same(a, b, what):
    for k in a:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), f"{what}: {k} differs"


# The static superset of a trough with more or finer boxes needs over 256
# schedule colors, which the reference rejects at creation.
@pytest.mark.parametrize("boxes,cells,margin", [(10, 1, 0.0), (20, 1, 0.0), (8, 2, 0.0), (8, 2, 0.002), (2, 4, 0.002)])
def test_broad_phase_steps_like_the_static_superset(boxes, cells, margin):
    print(f"case: broad phase vs superset, {boxes} boxes x {cells} cells, margin {margin}", flush=True)
    settings = dict(scenes.TROUGH_SETTINGS)
    sup = scenes.trough(box_count=boxes, box_cells=cells, seed=2)
    bp = scenes.trough(box_count=boxes, box_cells=cells, seed=2, broad_phase=True, reach_margin=margin)
    a = jgs2.JGS2GpuSolver.create(sup, settings)
    b = jgs2.JGS2GpuSolver.create(bp, settings)
    nearest = np.inf
    for frame in range(60):  # the boxes land in the trough after about 35 frames
        a.step_frames_exact_iterations(1, IT, settings)
        b.step_frames_exact_iterations(1, IT, settings)
        assert_same(state(a), state(b), f"frame {frame}")
        if frame % 10 == 9:
            d = candidate_distances(xyz(a.read_positions()), sup.contact)
            nearest = min(nearest, float(d.min()))
    info = b.contact_info()
    print(f"nearest pair {nearest:.4f} m; {info['builds']} builds, last list {info['candidates']}", flush=True)
    assert nearest < settings["ipc_activation_distance"]  # inside the barrier: contact forces acted
    assert info["builds"] > 1
    a.destroy()
    b.destroy()


def test_broad_phase_scene_in_a_batch_steps_as_alone():
    print("case: broad-phase trough in a mixed batch, eager and graphs", flush=True)
    settings = dict(scenes.TROUGH_SETTINGS)
    bp = scenes.trough(box_count=6, box_cells=2, seed=4, broad_phase=True, reach_margin=0.004)
    beam = scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5))
    sup = scenes.trough(box_count=3)
    alone = jgs2.JGS2GpuSolver.create(bp, settings)
    for graphs in (False, True):
        batch = JGS2Batch.create(capacity_for([bp, beam, sup]), settings)
        batch.use_graphs(graphs)
        m = [batch.admit(x) for x in (beam, bp, sup)]
        ref = jgs2.JGS2GpuSolver.create(bp, settings)
        for frame in range(12):
            batch.step_frames_exact_iterations(1, IT, settings)
            ref.step_frames_exact_iterations(1, IT, settings)
            assert_same(state(ref), state(m[1]), f"graphs {graphs}, frame {frame}")
        ref.destroy()
        batch.destroy()
    alone.destroy()


def test_broad_phase_rejects_gauss_seidel_and_small_capacity():
    print("case: broad-phase errors", flush=True)
    settings = dict(scenes.TROUGH_SETTINGS)
    x = scenes.trough(box_count=4, box_cells=1, broad_phase=True)
    solver = jgs2.JGS2GpuSolver.create(x, settings)
    with pytest.raises(jgs2.JGS2RangeError, match="Jacobi"):
        solver.step_exact_iterations(IT, dict(settings, schedule="graph-colored-gauss-seidel"))
    solver.destroy()
    small = scenes.trough(box_count=4, box_cells=1, broad_phase=True, candidate_capacity=8)
    with pytest.raises(jgs2.JGS2MemoryError, match="capacity"):
        jgs2.JGS2GpuSolver.create(small, settings)
