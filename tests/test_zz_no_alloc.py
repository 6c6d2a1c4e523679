"""jgs2_read_oracle_diagnostics_noalloc: the allocation-free variant of the
exact-diagnostics entry point, called through the C ABI.

Bar: its summary equals the allocating entry point's summary bit for bit
(same kernels, same fixed-order reductions) on two calls in a row, the
second on a smaller scene than the first, so the shared device buffer is
reused with stale contents from the larger call; and a scene whose records
exceed the fixed buffer is refused with a range error before any GPU work.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import scenes

pytestmark = pytest.mark.gpu


def _same(a, b, label):
    for key in a:
        va, vb = a[key], b[key]
        same = va == vb or (isinstance(va, float) and np.isnan(va) and np.isnan(vb))
        assert same, f"{label}: {key} {va!r} != {vb!r}"


def test_noalloc_summary_equals_allocating_entry_point_on_shrinking_calls():
    settings = dict(scenes.CANTILEVER_SETTINGS)
    large = scenes.cantilever_beam(cells=(12, 3, 3), size=(3.0, 0.75, 0.75))
    small = scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5))
    solvers = []
    for scene in (large, small):
        solver = jgs2.JGS2GpuSolver.create(scene, settings)
        solver.step_frames_exact_iterations(3, 8, settings)
        solvers.append(solver)
    # The allocation-free calls come first, larger scene then smaller; the
    # allocating entry point is evaluated afterwards as the reference value.
    print("case: noalloc diagnostics, 12x3x3 then 4x2x2 cantilever", flush=True)
    noalloc = [s.read_oracle_diagnostics_noalloc(settings) for s in solvers]
    for label, s, got in zip(("large", "small"), solvers, noalloc):
        want = s.read_oracle_diagnostics(settings)
        print(f"{label}: energy {got['energy']:.9e} vs {want['energy']:.9e}", flush=True)
        assert got["energy"] > 0 and got["finite"] == 1
        _same(got, want, label)
    # Repeat on the small scene after the large one's evaluation filled the
    # buffer again: same summary.
    solvers[0].read_oracle_diagnostics_noalloc(settings)
    _same(solvers[1].read_oracle_diagnostics_noalloc(settings), noalloc[1], "small again")
    for s in solvers:
        s.destroy()


def test_noalloc_refuses_scenes_beyond_the_fixed_buffer():
    # 36 doubles per vertex + 4 per tetrahedron: a 90x40x40 box needs
    # about 9.0 M doubles, over the 8,388,608 the buffer holds.
    print("case: noalloc diagnostics size limit", flush=True)
    scene = scenes.cantilever_beam(cells=(90, 40, 40), size=(9.0, 4.0, 4.0))
    need = 36 * scene.vertex_count + 4 * scene.tet_count + 32
    assert need > 8 * 2 ** 20
    solver = jgs2.JGS2GpuSolver.create(scene, scenes.CANTILEVER_SETTINGS)
    with pytest.raises(jgs2.JGS2RangeError, match="at most 8388608 doubles"):
        solver.read_oracle_diagnostics_noalloc(scenes.CANTILEVER_SETTINGS)
    solver.destroy()
