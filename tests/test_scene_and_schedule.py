"""Host scene setup, canonical CSR and the greedy coloring (CPU).

Pins: the reference's coloring golden vector
(src/simulation/cpu/vertex-coloring.test.ts:17-28); the canonical CSR rule
(src/simulation/gpu/layout.ts:295-401) against a brute-force scan; mesh
counts of the Kuhn split (src/simulation/cpu/mesh.ts:101-200); lumped mass
= density x volume; the stable Neo-Hookean parameter map
(src/simulation/cpu/stable-neo-hookean.ts:68-109).
"""

import ctypes as C

import numpy as np
import pytest

import jgs2_oracle as O
from jgs2 import _native as N
from jgs2 import scenes

GOLDEN_GROUPS = [[0, 1, 2, 3], [0, 4, 5], [1, 4, 6, 7], [2, 5, 6, 8]]
GOLDEN_COLORS = [0, 1, 2, 3, 2, 1, 0, 3, 3]


def product_coloring(vertex_count, groups):
    offsets = [0]
    members = []
    for g in groups:
        members.extend(int(v) for v in g)
        offsets.append(len(members))
    off = np.asarray(offsets, np.uint32)
    mem = np.asarray(members or [0], np.uint32)
    colors = np.empty(vertex_count, np.uint32)
    count = C.c_uint32()
    N.check(N.lib.jgs2_greedy_vertex_coloring(vertex_count, N.uptr(off), len(groups), N.uptr(mem), N.uptr(colors),
                                              C.byref(count)))
    return colors, count.value


def test_coloring_golden_vector_product_and_oracle():
    colors, count = product_coloring(9, GOLDEN_GROUPS)
    assert colors.tolist() == GOLDEN_COLORS
    assert count == 4
    assert O.greedy_coloring(9, GOLDEN_GROUPS).tolist() == GOLDEN_COLORS
    reversed_colors, _ = product_coloring(9, GOLDEN_GROUPS[::-1])
    assert reversed_colors.tolist() == GOLDEN_COLORS
    for g in GOLDEN_GROUPS:
        assert len({int(colors[v]) for v in g}) == len(g)


def test_coloring_of_beam_conflicts_is_bit_exact_and_proper():
    beam = scenes.cantilever_beam()
    groups = O.conflict_groups(beam)
    colors, count = product_coloring(beam.vertex_count, groups)
    assert np.array_equal(colors, O.greedy_coloring(beam.vertex_count, groups))
    for g in groups:
        assert len({int(colors[v]) for v in g}) == len(set(int(v) for v in g))
    assert count == int(colors.max()) + 1


def brute_force_csr(V, tets):
    rows = [[] for _ in range(V)]
    for t, tet in enumerate(tets):
        for v in dict.fromkeys(int(x) for x in tet):
            rows[v].append(t)
    return rows


@pytest.mark.parametrize("cells,jitter", [((16, 4, 4), 0.0), ((5, 3, 4), 0.1)])
def test_cuboid_mesh_and_csr(cells, jitter):
    scene = scenes.build_cuboid(scenes._desc(cells, (0, 0, 0), (2.0, 1.0, 1.5), jitter=jitter, seed=11,
                                             cubature_k=6))
    nx, ny, nz = cells
    assert scene.vertex_count == (nx + 1) * (ny + 1) * (nz + 1)
    assert scene.tet_count == 6 * nx * ny * nz
    tets = scene.tet_indices.reshape(-1, 4)
    x = scene.positions.reshape(-1, 4)[:, :3].astype(np.float64)
    d = np.linalg.det(np.stack([x[tets[:, 1]] - x[tets[:, 0]], x[tets[:, 2]] - x[tets[:, 0]],
                                x[tets[:, 3]] - x[tets[:, 0]]], axis=2))
    assert (d > 0).all()  # orientTetrahedra
    vol = scene.tet_meta.reshape(-1, 4)[:, 0]
    assert np.allclose(vol, d / 6, rtol=1e-6)
    assert vol.sum() == pytest.approx(2.0 * 1.0 * 1.5, rel=1e-6 if jitter == 0 else 0.05)
    mass = scene.vertex_rest.reshape(-1, 4)[:, 3]
    assert mass.sum() == pytest.approx(1000.0 * vol.astype(np.float64).sum(), rel=1e-5)
    info = scene.vertex_info.reshape(-1, 4)
    rows = brute_force_csr(scene.vertex_count, tets)
    start = 0
    for v, row in enumerate(rows):
        assert info[v, 0] == start and info[v, 1] == len(row)
        assert scene.adjacency[start:start + len(row)].tolist() == row
        start += len(row)
    assert start == scene.adjacency.size
    # Inverse rest matrices: Dm^-1 Dm = I for every tetrahedron.
    inv = scene.tet_inverse_dm.reshape(-1, 3, 4)[:, :, :3].transpose(0, 2, 1)
    rest = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    dm = np.stack([rest[tets[:, 1]] - rest[tets[:, 0]], rest[tets[:, 2]] - rest[tets[:, 0]],
                   rest[tets[:, 3]] - rest[tets[:, 0]]], axis=2)
    assert np.allclose(np.einsum("tij,tjk->tik", inv, dm), np.eye(3), atol=1e-5)


def test_stable_neo_hookean_parameters():
    beam = scenes.cantilever_beam()
    E, nu = 1e5, 0.4
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    meta = beam.tet_meta.reshape(-1, 4)
    assert np.allclose(meta[:, 1], lam + 5.0 / 6.0 * mu, rtol=1e-7)
    assert np.allclose(meta[:, 2], 4.0 / 3.0 * mu, rtol=1e-7)
    assert (meta[:, 3] == 1).all()


def test_cubature_records_have_reference_form():
    beam = scenes.cantilever_beam()
    V, K = beam.vertex_count, beam.cubature_k
    ids = beam.cubature_tet_ids.reshape(V, K)
    w = beam.cubature_weights.reshape(V, K)
    basis = beam.cubature_basis.reshape(V, K, 4, 3, 3)
    tets = beam.tet_indices.reshape(-1, 4)
    info = beam.vertex_info.reshape(-1, 4)
    for v in range(V):
        if info[v, 2]:
            assert (ids[v] == 0xFFFFFFFF).all() and (w[v] == 0).all()
            continue
        for k in range(K):
            if w[v, k] <= 0:
                continue
            tet = tets[ids[v, k]]
            for l in range(4):
                if tet[l] == v:
                    assert np.allclose(basis[v, k, l], np.eye(3))
                if info[tet[l], 2]:
                    assert not basis[v, k, l].any()
        assert (w[v] > 0).sum() >= 1
    # Incident samples carry weight 1, second-ring samples 1/4.
    assert set(np.unique(w[w > 0]).tolist()) <= {1.0, 0.25}


def test_large_box_counts():
    box = scenes.layered_box(cells=(8, 8, 16))
    assert box.vertex_count == 9 * 9 * 17
    assert box.tet_count == 6 * 8 * 8 * 16
    E = box.tet_meta.reshape(-1, 4)[:, 2]
    assert len(np.unique(E)) == 2  # alternating soft/stiff layers
    assert (box.vertex_info.reshape(-1, 4)[:, 2] == 1).sum() == 81  # z = 0 face pinned
