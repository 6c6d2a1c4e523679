"""Host logic of batching and serving (no GPU): the seeded mixed-scene
generator, HBM-budget sizing of the arena, Poisson arrivals, and the
request loop's admission, retirement and latency accounting driven through
a stand-in server whose "device" is a Python counter."""

import numpy as np
import pytest

from jgs2 import scenes
from jgs2.batch import BatchCapacity, footprint
from jgs2.serving import Request, TickClock, _Server, capacity_for_budget, nearest_rank, poisson_arrivals


def test_mixed_scene_specs_are_seeded_and_in_range():
    a = scenes.mixed_scene_specs(count=512, seed=7)
    b = scenes.mixed_scene_specs(count=512, seed=7)
    assert a == b
    assert a != scenes.mixed_scene_specs(count=512, seed=8)
    kinds = [s["kind"] for s in a]
    assert 0.25 < kinds.count("cloth") / len(a) < 0.35
    for s in a:
        if s["kind"] == "cloth":
            assert 32 <= s["cells"] <= 256
        else:
            assert 1e4 <= s["young"] <= 1e7 and 0.3 <= s["poisson"] <= 0.45
            tets = 6 * int(np.prod(s["cells"]))
            assert 6 <= tets <= 2.6e5, s


def test_mixed_scene_sizes_follow_the_specs():
    specs = scenes.mixed_scene_specs(count=6, seed=3, tet_range=(200, 2000), cloth_cells=(6, 12))
    for s in specs:
        x = scenes.build_mixed_scene(s)
        if s["kind"] == "cloth":
            n = s["cells"]
            assert x.cloth.triangle_count == 2 * n * n
            assert x.vertex_count == (n + 1) ** 2 + 8  # the sheet and its one-cell fixed block
        else:
            cx, cy, cz = s["cells"]
            assert x.tet_count == 6 * cx * cy * cz
            assert x.vertex_count == (cx + 1) * (cy + 1) * (cz + 1)


def test_footprint_rounds_ranges_to_blocks():
    x = scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5))
    f = footprint(x)
    assert f["vertices"] == 256 and f["tets"] == 256 and f["candidates"] == 0 and f["bodies"] == 1
    t = scenes.trough(box_count=3)
    assert footprint(t)["candidates"] == 256 * ((t.contact.candidate_count + 255) // 256)


def test_arena_bytes_grow_with_capacity_and_budget_sizing_fits():
    small = BatchCapacity(max_vertices=1000, max_tets=5000, max_scenes=4, cubature_k=6)
    big = BatchCapacity(max_vertices=2000, max_tets=5000, max_scenes=4, cubature_k=6)
    assert big.arena_bytes() > small.arena_bytes() > 0
    # Cubature records dominate per vertex: 10 uint4 per record.
    assert big.arena_bytes() - small.arena_bytes() >= 1024 * 6 * 160
    pool = [scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5)), scenes.trough(box_count=2)]
    for budget in (64 << 20, 256 << 20):
        cap = capacity_for_budget(pool, budget)
        assert cap.arena_bytes() <= budget
        bigger = BatchCapacity(**{**cap.__dict__, "max_scenes": cap.max_scenes + 1})
        grown = capacity_for_budget(pool, budget * 2)
        assert grown.max_scenes > cap.max_scenes
        assert bigger.max_scenes == cap.max_scenes + 1
    with pytest.raises(ValueError):
        capacity_for_budget(pool, 1000)


def test_poisson_arrivals_rate_and_order():
    t = poisson_arrivals(2000.0, 5.0, seed=1)
    assert np.all(np.diff(t) > 0) and t[-1] < 5.0
    assert abs(t.size / 5.0 - 2000.0) < 4 * np.sqrt(2000.0 * 5.0) / 5.0


def test_nearest_rank_percentiles():
    v = list(range(1, 101))
    assert nearest_rank(v, 50) == 50 and nearest_rank(v, 99) == 99 and nearest_rank(v, 100) == 100
    assert nearest_rank([3.0], 99) == 3.0 and nearest_rank([], 50) == 0.0


class _Handle:
    def __init__(self, size):
        self.free_vertex_count = size

    def read_positions(self):
        return np.zeros(4)


class CounterServer(_Server):
    """A server whose device holds `slots` scenes; records every call."""

    def __init__(self, slots, iterations=2):
        super().__init__({}, iterations, keep_results=True, clock=TickClock())
        self.slots = slots
        self.log = []

    def _try_admit(self, request):
        if len(self.active) >= self.slots:
            return None
        self.log.append(("admit", request.id, self.now()))
        return _Handle(request.scene)

    def _step_all(self):
        self.log.append(("step", [r.id for r, _, _ in self.active]))

    def _release(self, handle):
        self.log.append(("release",))


def test_continuous_batching_admits_fifo_as_slots_free_up():
    # Scenes are stand-ins: `scene` carries the free-vertex count.
    reqs = [Request(id=i, scene=10 * (i + 1), frames=f, arrival=a)
            for i, (f, a) in enumerate([(3, 0.0), (1, 0.0), (2, 0.5), (1, 4.0)])]
    server = CounterServer(slots=2)
    rep = server.run(reqs)
    steps = [e[1] for e in server.log if e[0] == "step"]
    # Tick 0: an idle server admits requests 0 and 1, then steps them.
    # Request 2 (arrived at 0.5) is admitted during tick 1, once request 1
    # has finished, while request 0's frame is in flight, and joins from the
    # next frame; request 3 arrives at 4 into an idle server.
    assert steps == [[0, 1], [0], [0, 2], [2], [3]]
    assert [r.id for r in rep.completed] == [1, 0, 2, 3]
    done = {r.id: r for r in rep.completed}
    assert done[2].admitted == 1.0 and done[2].finished == 4.0
    assert done[3].admitted == 4.0 and done[3].finished == 5.0
    assert rep.vertex_updates == 2 * (10 + 20 + 10 + 10 + 30 + 30 + 40)
    s = rep.summary()
    assert s["requests_completed"] == 4 and s["ticks"] == 5 and s["admissions"] == 4
    assert s["queueing_s"]["p99"] == 0.5
    assert s["resident_s"]["p50"] == 1.0 and s["resident_s"]["p99"] == 3.0
    assert all(r.result is not None for r in rep.completed)


def test_request_larger_than_an_empty_server_is_an_error():
    server = CounterServer(slots=0)
    with pytest.raises(ValueError, match="does not fit"):
        server.run([Request(id=0, scene=1, frames=1, arrival=0.0)])


def test_duration_stops_the_loop():
    reqs = [Request(id=i, scene=1, frames=100, arrival=0.0) for i in range(3)]
    server = CounterServer(slots=3)
    rep = server.run(reqs, duration=10.0)
    assert rep.ticks == 10 and not rep.completed
