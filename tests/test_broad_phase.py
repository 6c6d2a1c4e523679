"""The candidate set within AABB reach (no GPU): the oracle's restatement of
the reference's static VT/EE superset (src/simulation/cpu/ipc-contact.ts:
600-665) restricted to pairs whose AABB distance lower bound
(ipcAabbDistanceLowerBound, src/simulation/gpu/jgs2-shader.ts:525-550) is
within a reach.

Pins: at infinite reach the oracle reproduces the reference's golden
superset counts for the default trough (src/scenes/scenes.test.ts:520-524)
and the library's superset builder pair for pair (same canonical order) for
1 to 20 boxes; at finite reach it equals the superset filtered by a numpy
float32 evaluation of the lower bound; hand-built cases fix the boundary
(a pair exactly at the reach is kept), the one-ring exclusions and the
inclusion of non-finite points."""

import numpy as np
import pytest

from jgs2 import scenes

import jgs2_oracle as O


def _surface(scene):
    s = scene.extras["surface"]
    return s["triangles"], s["edges"]


def _lower_bounds(packed, vt, positions):
    """Float32 AABB lower bound of every candidate, one rounding per op."""
    p = np.asarray(positions, np.float32).reshape(-1, 4)[:, :3]
    q = p[packed.astype(np.int64)]  # [n, 4, 3]
    lo = np.empty((len(packed), 3), np.float32)
    hi = np.empty_like(lo)
    lo2 = np.empty_like(lo)
    hi2 = np.empty_like(lo)
    lo[:vt] = hi[:vt] = q[:vt, 0]
    lo2[:vt] = np.minimum(np.minimum(q[:vt, 1], q[:vt, 2]), q[:vt, 3])
    hi2[:vt] = np.maximum(np.maximum(q[:vt, 1], q[:vt, 2]), q[:vt, 3])
    lo[vt:] = np.minimum(q[vt:, 0], q[vt:, 1])
    hi[vt:] = np.maximum(q[vt:, 0], q[vt:, 1])
    lo2[vt:] = np.minimum(q[vt:, 2], q[vt:, 3])
    hi2[vt:] = np.maximum(q[vt:, 2], q[vt:, 3])
    g = np.maximum(np.maximum(lo2 - hi, lo - hi2), np.float32(0))
    s = (g[:, 0] * g[:, 0] + g[:, 1] * g[:, 1]) + g[:, 2] * g[:, 2]
    return np.sqrt(s.astype(np.float32))


def test_infinite_reach_is_the_reference_superset_of_the_default_trough():
    t = scenes.trough(box_count=10, candidates=False)
    tri, edg = _surface(t)
    packed, vt = O.contact_candidates_within(t.vertex_count, t.positions, tri, edg, float("inf"))
    assert vt == 11_728 and len(packed) - vt == 20_073 and len(packed) == 31_801


@pytest.mark.parametrize("boxes", [1, 2, 5, 13, 20])
def test_infinite_reach_matches_the_library_superset_pair_for_pair(boxes):
    t = scenes.trough(box_count=boxes, candidates=False)
    tri, edg = _surface(t)
    packed, vt = O.contact_candidates_within(t.vertex_count, t.positions, tri, edg, float("inf"))
    ref = scenes.static_contact_candidates(t.vertex_count, tri, edg)
    assert vt == ref.vertex_triangle_count
    np.testing.assert_array_equal(packed.reshape(-1), ref.packed_indices)


@pytest.mark.parametrize("boxes,cells,reach", [(3, 1, 0.1), (10, 1, 0.151), (4, 2, 0.2), (20, 1, 0.05)])
def test_finite_reach_is_the_superset_filtered_by_the_lower_bound(boxes, cells, reach):
    t = scenes.trough(box_count=boxes, box_cells=cells, seed=3, candidates=False)
    tri, edg = _surface(t)
    ref = scenes.static_contact_candidates(t.vertex_count, tri, edg)
    full = ref.packed_indices.reshape(-1, 4)
    lb = _lower_bounds(full, ref.vertex_triangle_count, t.positions)
    keep = ~(lb > np.float32(reach))
    packed, vt = O.contact_candidates_within(t.vertex_count, t.positions, tri, edg, reach)
    np.testing.assert_array_equal(packed, full[keep])
    assert vt == int(keep[: ref.vertex_triangle_count].sum())
    assert 0 < len(packed) < len(full)


def _two_triangles(gap):
    # Triangle (0,1,2) in z = 0 and triangle (3,4,5) in z = gap, no shared edges.
    pos = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1],
                    [0, 0, gap, 1], [1, 0, gap, 1], [0, 1, gap, 1]], np.float32)
    return pos, [0, 1, 2, 3, 4, 5], []


def test_reach_boundary_and_one_ring_exclusion():
    pos, tri, edg = _two_triangles(0.25)
    everything, vt_all = O.contact_candidates_within(6, pos, tri, edg, float("inf"))
    # Each vertex against the other triangle (6 VT), and each of the 3x3
    # edge pairs across the triangles (9 EE); pairs within a triangle share
    # its one-ring and are excluded.
    assert vt_all == 6 and len(everything) == 15
    gap = np.float32(0.25)
    at, vt_at = O.contact_candidates_within(6, pos, tri, edg, float(gap))
    assert vt_at == 6 and len(at) == 15  # a bound equal to the reach is kept
    below, _ = O.contact_candidates_within(6, pos, tri, edg, float(np.nextafter(gap, np.float32(0))))
    assert len(below) == 0


def test_non_finite_points_are_always_candidates():
    pos, tri, edg = _two_triangles(5.0)
    assert len(O.contact_candidates_within(6, pos, tri, edg, 1.0)[0]) == 0
    pos[4, 1] = np.nan
    packed, vt = O.contact_candidates_within(6, pos, tri, edg, 1.0)
    # Every pair that touches vertex 4: VT (0|1|2 vs triangle 3,4,5, and 4
    # vs triangle 0,1,2) and EE (each edge of the first triangle vs (3,4)
    # and (4,5)).
    assert vt == 4 and len(packed) == 4 + 6
    assert all(4 in row for row in packed.tolist())


@pytest.mark.parametrize("boxes,cells", [(10, 1), (20, 2)])
def test_library_capacity_estimate_counts_the_oracle_list(boxes, cells):
    # The library sizes a broad-phase scene at twice the number of pairs
    # within its first build's reach (activation reach plus half) plus
    # 1024; its host count must be the oracle's count at that reach.
    st = scenes.TROUGH_SETTINGS
    x = scenes.trough(box_count=boxes, box_cells=cells, seed=1, broad_phase=True)
    p = x.positions.reshape(-1, 4)[:, :3].astype(np.float64)
    free = x.vertex_info.reshape(-1, 4)[:, 2] == 0
    scale = float(np.float32(np.linalg.norm(p[free].max(0) - p[free].min(0))))
    rr = max(scale, st["ipc_activation_distance"], 1e-12) * 2.0**-24 * 8.0
    act = st["ipc_activation_distance"] + 2.0 * (st["max_step"] + rr)
    tri, edg = _surface(x)
    packed, _ = O.contact_candidates_within(x.vertex_count, x.positions, tri, edg, float(np.float32(1.5 * act)))
    assert x.contact.candidate_capacity == 2 * len(packed) + 1024
