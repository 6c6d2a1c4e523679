"""The library's cubature precompute (jgs2_precompute_cubature) against the
reference's golden selections and the FP64 oracle. Host only: these run
without a GPU.

* rings = 0 (the reference's whole-mesh basis): every packed selection of the
  reference's 12-tetrahedron beam (manifests/phase1-cubature.v1.json) bit for
  bit, and the oracle's records bit for bit on larger stable Neo-Hookean and
  co-rotated cuboids;
* rings = k: blocks outside the k-ring are zero, the source block is the
  identity, residuals stay within the bound, and rings past the mesh diameter
  reproduce rings = 0;
* error classes of the reference's failures.
"""

import json
from pathlib import Path

import numpy as np
import pytest

import jgs2
from jgs2.precompute import precompute_cubature
import jgs2_oracle as O
from test_precompute_oracle import beam, cuboid

MANIFEST = json.loads((Path(__file__).parent / "golden" / "phase1-cubature.v1.json").read_text())


def test_golden_selections_bit_for_bit():
    args = beam()
    r = precompute_cubature(*args, 6, rings=0, threads=4)
    for e in MANIFEST["expectedPackedSelections"]:
        v, n = e["vertex"], len(e["tetrahedra"])
        assert list(r.tet_ids[v, :n]) == e["tetrahedra"], v
        assert np.array_equal(r.weights[v, :n], np.array(e["weights"], np.float32)), v
        assert np.all(r.tet_ids[v, n:] == 0xFFFFFFFF)
        assert list(r.ranks[v]) == [e["nonzeroCandidateCount"], e["trainingColumnRank"], e["packedTrainingColumnRank"]]
        assert r.residuals[v] == pytest.approx(e["packedResidual"], rel=1e-9)
    assert np.isnan(r.residuals[[0, 6]]).all()
    ot, ow, ob, _ = O.precompute_cubature(*args, 6)
    assert np.array_equal(r.basis, ob)


@pytest.mark.parametrize("case", ["snh-4x2x2", "corotated-6x2x2"])
def test_matches_the_oracle_bit_for_bit(case):
    if case == "snh-4x2x2":
        pos, tets, fixed = cuboid((4, 2, 2), (0, 0, 0), (2, 1, 1), lambda p, c: c[0] == 0)
        mats = [(1000, 1e5, 0.4, 1)]
    else:
        pos, tets, fixed = cuboid((6, 2, 2), (0, 0, 0), (1.5, 0.5, 0.5), lambda p, c: c[0] == 0)
        mats = [(950, 280000, 0.32, 0)]
    args = (pos, tets, fixed, np.zeros(len(tets)), mats, 1 / 60)
    ot, ow, ob, os_ = O.precompute_cubature(*args, 6, bound=1.0)
    r = precompute_cubature(*args, 6, rings=0, residual_bound=1.0)
    assert np.array_equal(r.tet_ids, ot)
    assert np.array_equal(r.weights, ow)
    assert np.array_equal(r.basis, ob)
    free = fixed == 0
    np.testing.assert_allclose(r.residuals[free], os_[free, 0], rtol=1e-9)


def _graph_distance(tets, V, src):
    nb = [set() for _ in range(V)]
    for t in tets:
        for a in t:
            nb[a].update(int(b) for b in t if b != a)
    dist = np.full(V, 1 << 30)
    dist[src] = 0
    frontier = [src]
    while frontier:
        nxt = []
        for v in frontier:
            for u in nb[v]:
                if dist[u] > dist[v] + 1:
                    dist[u] = dist[v] + 1
                    nxt.append(u)
        frontier = nxt
    return dist


@pytest.mark.parametrize("rings", [1, 2])
def test_ring_restricted_basis(rings):
    pos, tets, fixed = cuboid((6, 3, 3), (0, 0, 0), (1.5, 0.75, 0.75), lambda p, c: c[0] == 0)
    args = (pos, tets, fixed, np.zeros(len(tets)), [(1000, 1e5, 0.4, 1)], 1 / 60)
    r = precompute_cubature(*args, 6, rings=rings, residual_bound=1.0)
    free = np.flatnonzero(fixed == 0)
    assert np.all(r.residuals[free] <= 1.0) and np.all(np.isfinite(r.residuals[free]))
    for v in free[::7]:
        # distance over the whole mesh; the ring walk does not pass through
        # pinned vertices, so its distances are never shorter than these
        dist = _graph_distance(tets, len(pos), v)
        used = r.tet_ids[v] != 0xFFFFFFFF
        assert used.any() and np.all(r.weights[v][used] > 0)
        for k in np.flatnonzero(used):
            blocks = r.basis[v, k].reshape(4, 3, 3)
            for l, u in enumerate(tets[r.tet_ids[v, k]]):
                if fixed[u]:
                    assert not blocks[l].any()
                elif u == v:
                    np.testing.assert_allclose(blocks[l], np.eye(3), atol=1e-6)
                elif dist[u] > rings:
                    assert not blocks[l].any()


def test_rings_past_the_diameter_equal_the_whole_mesh_basis():
    pos, tets, fixed = cuboid((4, 2, 2), (0, 0, 0), (2, 1, 1), lambda p, c: c[0] == 0)
    args = (pos, tets, fixed, np.zeros(len(tets)), [(1000, 1e5, 0.4, 1)], 1 / 60)
    a = precompute_cubature(*args, 6, rings=0, residual_bound=1.0)
    b = precompute_cubature(*args, 6, rings=50, residual_bound=1.0, threads=3)
    assert np.array_equal(a.tet_ids, b.tet_ids) and np.array_equal(a.weights, b.weights)
    assert np.array_equal(a.basis, b.basis)


def test_conjugate_gradient_modes_agree_with_the_dense_factor():
    pos, tets, fixed = cuboid((4, 2, 2), (0, 0, 0), (2, 1, 1), lambda p, c: c[0] == 0)
    args = (pos, tets, fixed, np.zeros(len(tets)), [(1000, 1e5, 0.4, 1)], 1 / 60)
    a = precompute_cubature(*args, 6, rings=2, residual_bound=1.0, mode_solver="dense")
    b = precompute_cubature(*args, 6, rings=2, residual_bound=1.0, mode_solver="cg", cg_tolerance=1e-13)
    same = np.mean([np.array_equal(a.tet_ids[v], b.tet_ids[v]) for v in range(len(pos))])
    assert same >= 0.9
    free = fixed == 0
    np.testing.assert_allclose(a.residuals[free], b.residuals[free], rtol=0.05, atol=1e-3)


def test_errors():
    args = beam()
    pos, tets, fixed = cuboid((4, 2, 2), (0, 0, 0), (2, 1, 1), lambda p, c: c[0] == 0)
    with pytest.raises(jgs2.JGS2InfeasibleError, match="packed training residual"):
        precompute_cubature(pos, tets, fixed, np.zeros(len(tets)), [(1000, 1e5, 0.4, 1)], 1 / 60, 6)
    mid = np.arange(len(tets)) % 2
    with pytest.raises(jgs2.JGS2RangeError, match="Mixed stable Neo-Hookean"):
        precompute_cubature(pos, tets, fixed, mid, [(1000, 1e5, 0.3, 1), (1000, 1e5, 0.3, 0)], 1 / 60, 6)
    with pytest.raises(jgs2.JGS2RangeError, match="movable vertex"):
        precompute_cubature(pos, tets, np.ones(len(pos)), np.zeros(len(tets)), [(1000, 1e5, 0.3, 1)], 1 / 60, 6)
    with pytest.raises(jgs2.JGS2RangeError, match="Cubature samples"):
        precompute_cubature(*args, 0)
    with pytest.raises(jgs2.JGS2RangeError, match="timestep"):
        precompute_cubature(*args[:5], 0.0, 6)
    with pytest.raises(jgs2.JGS2RangeError, match="Poisson"):
        precompute_cubature(*args[:4], [(1000, 5000, 0.5, 1)], args[5], 6)
    bad = args[1].copy()
    bad[0] = 99
    with pytest.raises(jgs2.JGS2RangeError, match="out of range"):
        precompute_cubature(args[0], bad, *args[2:], 6)


def test_exact_all_element_packing_of_the_force_free_fixture():
    # scenes.test.ts:127-200: cubatureK = tetCount = 48, every vertex packs
    # tetrahedra 0..47 in order with weight 1 and its exact basis at each
    # tetrahedron's four vertices (to 5 decimals); pinned here against the
    # oracle's exact basis (itself pinned to the reference's beam golden).
    import json
    from pathlib import Path

    import jgs2_oracle as O
    from jgs2 import manifests

    m = json.loads((Path(__file__).parent / "golden" / "canonical-scenes.v1.json").read_text())
    x, settings, _, _ = manifests.force_free_cuboid(m)
    V, T = x.vertex_count, x.tet_count
    assert (V, T, x.cubature_k) == (27, 48, 48)
    ids = x.cubature_tet_ids.reshape(V, T)
    assert np.array_equal(ids, np.tile(np.arange(T, dtype=np.uint32), (V, 1)))
    assert np.array_equal(x.cubature_weights.reshape(V, T), np.ones((V, T), np.float32))
    rest = x.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    tets = x.tet_indices.reshape(-1, 4)
    basis = x.cubature_basis.reshape(V, T, 4, 9)
    for v in (0, 5, 13, 26):
        U = O.exact_basis(rest, tets, np.zeros(V), np.zeros(T), [(1000.0, 8e4, 0.3, 1)], settings["timestep"], v)
        for t in range(T):
            for l in range(4):
                assert np.abs(basis[v, t, l] - U[tets[t, l]].reshape(9)).max() < 5e-6
