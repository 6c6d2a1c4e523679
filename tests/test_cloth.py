"""Cloth layer: StVK membrane triangles and quadratic dihedral hinges (CPU).

Pins of the oracle's cloth terms (oracle/jgs2_oracle.cpp, restating
src/simulation/gpu/jgs2-shader.ts:1225-1845) and of the product's cloth
host code:
* zero energy and gradient at rest, and finite-difference gradients of the
  reference's own deformed membrane triangle and bent hinge
  (src/simulation/cpu/cloth.test.ts:11-87), FP64 within 2e-8 relative;
* the membrane corner Hessian block against a finite difference of the
  analytic gradient;
* the plane-stress material map (lambda = 320, mu = 480 for E = 1200,
  nu = 0.25: src/simulation/gpu/cloth-layout.test.ts:40-67) through the
  closed-form energies of a uniaxial stretch and a simple shear;
* triangle rest data golden values (inverse basis [1, -0.25; 0, 1.25], area
  0.4: cloth-layout.test.ts:66-75);
* the dihedral sign and energy on a hinge folded by a known angle;
* the vertex-to-element CSR golden vectors (cloth-layout.test.ts:125-148);
* the drape scene (src/scenes/index.ts:632-705): counts, pins, hinge
  topology and order, rest data, masses, coloring;
* the oracle's assembled local gradient of a cloth vertex against a finite
  difference of its restricted energy.
"""

import ctypes as C

import numpy as np
import pytest

import jgs2
import jgs2_oracle as O
from jgs2 import _native as N
from jgs2 import scenes

REST_TRIANGLE = np.array([0, 0, 0, 1, 0, 0, 0.2, 0.8, 0], np.float64)
REST_HINGE = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, -1, 0], np.float64)
MEMBRANE = dict(young=8000.0, poisson=0.3, thickness=0.02)


def central_difference(fun, x, step=1e-6):
    """centralDifferenceGradient (src/simulation/cpu/finite-difference.ts:27-49)."""
    x = np.asarray(x, np.float64)
    out = np.empty_like(x)
    for i in range(x.size):
        h = step * max(1.0, abs(x[i]))
        p, m = x.copy(), x.copy()
        p[i] += h
        m[i] -= h
        out[i] = (fun(p) - fun(m)) / (2 * h)
    return out


def relative_error(a, b):
    """relativeError (finite-difference.ts:94-118)."""
    a, b = np.ravel(a), np.ravel(b)
    return float(np.linalg.norm(a - b) / max(1.0, np.linalg.norm(a), np.linalg.norm(b)))


def membrane(rest9, pos9, young, poisson, thickness):
    rest12 = np.concatenate([rest9, [1.0, -1.0, 0.0]])
    pos12 = np.concatenate([pos9, [1.0, -1.0, 0.0]])
    return O.cloth_probe(rest12, pos12, young, poisson, thickness, 0.0)


def hinge(pos12, k):
    return O.cloth_probe(REST_HINGE, pos12, 1000.0, 0.3, 0.01, k)


def test_zero_energy_and_gradient_at_rest():
    m = membrane(REST_TRIANGLE, REST_TRIANGLE, **MEMBRANE)
    assert m["membrane_energy"] < 1e-24
    assert np.linalg.norm(m["membrane_gradient"]) < 1e-12
    b = hinge(REST_HINGE, 2.5)
    assert b["hinge_valid"]
    assert b["bending_energy"] == 0.0
    assert b["angle"] == b["rest_angle"] == 0.0
    assert np.all(b["bending_gradient"] == 0.0)


def test_membrane_gradient_matches_finite_differences():
    x = np.array([-0.03, 0.02, 0.04, 1.08, -0.04, 0.12, 0.16, 0.91, -0.08])
    analytic = membrane(REST_TRIANGLE, x, **MEMBRANE)["membrane_gradient"]
    fd = central_difference(lambda y: membrane(REST_TRIANGLE, y, **MEMBRANE)["membrane_energy"], x)
    assert relative_error(analytic, fd) < 2e-8


def test_bending_gradient_matches_finite_differences():
    x = np.array([-0.02, 0.01, 0.03, 1.04, -0.03, -0.02, 0.08, 0.94, 0.22, 0.91, -1.07, -0.14])
    b = hinge(x, 3.25)
    assert b["bending_energy"] > 0
    fd = central_difference(lambda y: hinge(y, 3.25)["bending_energy"], x)
    assert relative_error(b["bending_gradient"], fd) < 2e-8


def test_membrane_corner_hessian_matches_gradient_difference():
    x = np.array([-0.03, 0.02, 0.04, 1.08, -0.04, 0.12, 0.16, 0.91, -0.08])
    H = membrane(REST_TRIANGLE, x, **MEMBRANE)["membrane_hessian_vertex0"]
    cols = []
    for a in range(3):
        h = 1e-6
        p, m = x.copy(), x.copy()
        p[a] += h
        m[a] -= h
        gp = membrane(REST_TRIANGLE, p, **MEMBRANE)["membrane_gradient"][0]
        gm = membrane(REST_TRIANGLE, m, **MEMBRANE)["membrane_gradient"][0]
        cols.append((gp - gm) / (2 * h))
    fd = np.stack(cols, axis=1)
    fd = 0.5 * (fd + fd.T)
    assert np.allclose(H, H.T)
    assert np.abs(H - fd).max() < 1e-6 * np.abs(fd).max()


def test_plane_stress_material_closed_form():
    # E = 1200, nu = 0.25: plane-stress lambda = 320, mu = 480.
    lam, mu, thickness, area = 320.0, 480.0, 0.125, 0.4
    s = 0.05
    x = REST_TRIANGLE.reshape(3, 3).copy()
    x[:, 0] *= 1 + s  # uniaxial stretch along the first rest axis
    e00 = 0.5 * ((1 + s) ** 2 - 1)
    expected = area * thickness * (mu * e00 ** 2 + 0.5 * lam * e00 ** 2)
    got = membrane(REST_TRIANGLE, x.reshape(-1), 1200.0, 0.25, thickness)["membrane_energy"]
    assert got == pytest.approx(expected, rel=1e-12)
    g = 0.07
    x = REST_TRIANGLE.reshape(3, 3).copy()
    x[:, 0] += g * x[:, 1]  # simple shear: f1 = (g, 1, 0)
    e01, e11 = 0.5 * g, 0.5 * g * g
    expected = area * thickness * (mu * (2 * e01 ** 2 + e11 ** 2) + 0.5 * lam * e11 ** 2)
    got = membrane(REST_TRIANGLE, x.reshape(-1), 1200.0, 0.25, thickness)["membrane_energy"]
    assert got == pytest.approx(expected, rel=1e-12)


def test_triangle_rest_golden_values():
    inv, area = O.triangle_rest(REST_TRIANGLE)
    assert np.allclose(inv, [[1.0, -0.25], [0.0, 1.25]], atol=1e-15)
    assert area == pytest.approx(0.4, abs=1e-15)


@pytest.mark.parametrize("theta", [0.3, 1.0, 2.5])
def test_dihedral_sign_and_energy_of_a_known_fold(theta):
    x = REST_HINGE.copy()
    x[9:12] = [1.0, -np.cos(theta), np.sin(theta)]  # vertex 3 turned about the edge (x axis)
    b = hinge(x, 2.0)
    assert b["angle"] == pytest.approx(-theta, abs=1e-14)
    assert b["rest_edge_length"] == 1.0
    assert b["bending_energy"] == pytest.approx(0.5 * 2.0 * 1.0 * theta ** 2, rel=1e-13)


def test_dihedral_gradient_entries_are_finite_above_the_edge_floor():
    """The device evaluates a hinge's validity without forming its twelve
    gradient entries when no corner gradient is wanted and the shared edge
    is at least 1e-20 long (they are then finite, so validity is decided by
    the gate and the angle alone). Pinned to the FP32 oracle, which forms
    all twelve as the reference does (S:1700-1810): over seeded hinges at
    scales 1e-19 to 1e15, down to nearly degenerate wings, every hinge that
    passes the gate with a margin is valid."""
    rng = np.random.default_rng(7)
    f = np.float32
    checked = 0
    for scale in (1e-19, 1e-15, 1e-9, 1e-3, 1.0, 1e3, 1e9, 1e15):
        for flat in (1.0, 1e-3, 1e-5):
            for _ in range(40):
                x = rng.normal(size=(4, 3))
                x[1] = x[0] + rng.normal(size=3)
                for w in (2, 3):  # wings pressed toward the edge line
                    t = rng.uniform(-0.5, 1.5)
                    x[w] = x[0] + t * (x[1] - x[0]) + flat * (x[w] - x[0] - t * (x[1] - x[0]))
                x = (x * scale).astype(f)
                e = x[1] - x[0]
                el = np.sqrt(np.dot(e, e))
                if not el >= f(1e-20):
                    continue
                rn0, rn1 = np.cross(e, x[2] - x[0]), np.cross(-e, x[3] - x[1])
                nl0, nl1 = np.sqrt(np.dot(rn0, rn0)), np.sqrt(np.dot(rn1, rn1))
                es0 = max(el, np.linalg.norm(x[2] - x[0]), np.linalg.norm(x[2] - x[1]))
                es1 = max(el, np.linalg.norm(x[3] - x[1]), np.linalg.norm(x[3] - x[0]))
                margin = f(2e-12)
                if not (np.isfinite([nl0, nl1, es0, es1]).all() and nl0 > margin * es0 * es0 and nl1 > margin * es1 * es1):
                    continue
                b = O.cloth_probe(REST_HINGE, x.astype(np.float64).ravel(), 1000.0, 0.3, 0.01, 1.0, precision=32)
                assert b["hinge_valid"], (scale, flat, x.tolist())
                assert np.isfinite(b["angle"])
                checked += 1
    assert checked > 500


def test_dihedral_corner_gradients_closed_form():
    """The device forms a hinge's four corner gradients in closed form
    (a_k = N_k / |N_k|^2 with N_0 = e x (x2 - x0), N_1 = (-e) x (x3 - x1):
    grad_2 = -|e| a_0, grad_3 = -|e| a_1, grad_0 = ((x1 - x2).e a_0 +
    (x1 - x3).e a_1) / |e|, grad_1 = ((x2 - x0).e a_0 + (x3 - x0).e a_1) / |e|)
    instead of the reference's twelve directional derivatives; the identity
    holds against the FP64 oracle, which forms the twelve (S:1700-1810), on
    seeded hinges over seven decades of scale."""
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(600):
        scale = 10.0 ** rng.uniform(-4, 3)
        x = (REST_HINGE + rng.normal(scale=rng.uniform(0.01, 1.0), size=12)) * scale
        b = O.cloth_probe(REST_HINGE * scale, x, 1000.0, 0.3, 0.01, 1.0)
        if not b["hinge_valid"]:
            continue
        ad = np.arctan2(np.sin(b["angle"] - b["rest_angle"]), np.cos(b["angle"] - b["rest_angle"]))
        if abs(ad) < 1e-3:
            continue
        grad = b["bending_gradient"] / (b["rest_edge_length"] * ad)
        x0, x1, x2, x3 = x.reshape(4, 3)
        e = x1 - x0
        el = np.linalg.norm(e)
        n0, n1 = np.cross(e, x2 - x0), np.cross(-e, x3 - x1)
        a0, a1 = n0 / n0.dot(n0), n1 / n1.dot(n1)
        closed = np.array([((x1 - x2).dot(e) * a0 + (x1 - x3).dot(e) * a1) / el,
                           ((x2 - x0).dot(e) * a0 + (x3 - x0).dot(e) * a1) / el, -el * a0, -el * a1])
        assert np.abs(closed - grad).max() <= 1e-12 * np.abs(grad).max()
        checked += 1
    assert checked > 400


def incidence(V, idx, width):
    idx = np.asarray(idx, np.uint32)
    count = idx.size // width
    rows = np.empty(V + 1, np.uint32)
    elems = np.empty(max(idx.size, 1), np.uint32)
    N.check(N.lib.jgs2_build_element_incidence(V, N.uptr(idx), width, count, N.uptr(rows), N.uptr(elems)))
    return rows.tolist(), elems[: idx.size].tolist()


def test_incidence_csr_golden_vectors():
    rows, elems = incidence(5, [0, 1, 2, 1, 0, 3], 3)
    assert rows == [0, 2, 4, 5, 6, 6]
    assert elems == [0, 1, 0, 1, 0, 1]
    rows, elems = incidence(5, [0, 1, 2, 3], 4)
    assert rows == [0, 1, 2, 3, 4, 4]
    assert elems == [0, 0, 0, 0]
    rows, elems = incidence(3, [], 3)
    assert rows == [0, 0, 0, 0] and elems == []


@pytest.fixture(scope="module")
def drape():
    return scenes.cloth_drape()


def triangle_areas(x, tris):
    a, b, c = x[tris[:, 0]], x[tris[:, 1]], x[tris[:, 2]]
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)


def test_drape_counts_pins_and_masses(drape):
    cloth = drape.cloth
    assert drape.vertex_count == 8 + 25 and drape.tet_count == 6
    assert cloth.triangle_count == 32 and cloth.hinge_count == 40
    info = drape.vertex_info.reshape(-1, 4)
    off = drape.extras["cloth_offset"]
    assert off == 8
    assert np.all(info[:off, 2] == 1)
    pinned = sorted(int(v) - off for v in np.nonzero(info[off:, 2])[0] + off)
    assert pinned == [0, 2]  # grid vertices (0, 0) and (2, 0)
    assert np.all(info[off:, 3] == 1) and np.all(info[:off, 3] == 0)
    x = drape.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    tris = cloth.triangle_indices.reshape(-1, 3).astype(np.int64)
    areas = triangle_areas(x, tris)
    assert np.allclose(cloth.triangle_rest_areas, areas, rtol=1e-6)
    mass = drape.vertex_rest.reshape(-1, 4)[off:, 3].astype(np.float64)
    assert mass.sum() == pytest.approx(480.0 * 0.012 * areas.sum(), rel=1e-6)
    # The pre-folded columns x >= 3 sit 0.079 above the sheet (index.ts:651-661).
    y = x[off:, 1].reshape(5, 5)
    assert np.allclose(y[:, :3], 1.15) and np.allclose(y[:, 3:], 1.229)


def test_drape_inverse_rest_bases_map_rest_edges_to_identity(drape):
    cloth = drape.cloth
    x = drape.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    for t, (a, b, c) in enumerate(cloth.triangle_indices.reshape(-1, 3)):
        e01, e02 = x[b] - x[a], x[c] - x[a]
        tangent = e01 / np.linalg.norm(e01)
        normal = np.cross(e01, e02)
        normal /= np.linalg.norm(normal)
        bitangent = np.cross(normal, tangent)
        Dm = np.array([[e01 @ tangent, e02 @ tangent], [e01 @ bitangent, e02 @ bitangent]])
        inv = cloth.triangle_inverse_rest_bases.reshape(-1, 2, 2)[t]
        assert np.allclose(Dm @ inv, np.eye(2), atol=1e-5)


def test_drape_hinges_cover_interior_edges_in_closing_order(drape):
    cloth = drape.cloth
    tris = cloth.triangle_indices.reshape(-1, 3).tolist()
    x = drape.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)

    def find(a, b, c):
        for t, tri in enumerate(tris):
            for r in range(3):
                if tri[r:] + tri[:r] == [a, b, c]:
                    return t, r
        raise AssertionError((a, b, c))

    edge_count = {}
    for tri in tris:
        for r in range(3):
            e = tuple(sorted((tri[r], tri[(r + 1) % 3])))
            edge_count[e] = edge_count.get(e, 0) + 1
    interior = {e for e, n in edge_count.items() if n == 2}
    keys, seen = [], set()
    for h, (a, b, c, d) in enumerate(cloth.hinge_indices.reshape(-1, 4).tolist()):
        t0, _ = find(a, b, c)
        t1, r1 = find(b, a, d)
        assert t0 < t1
        e = tuple(sorted((a, b)))
        assert e in interior and e not in seen
        seen.add(e)
        keys.append((t1, r1))
        assert cloth.hinge_rest_edge_lengths[h] == pytest.approx(np.linalg.norm(x[b] - x[a]), rel=1e-6)
    assert seen == interior
    assert keys == sorted(keys)  # each hinge appears when its second triangle closes the edge
    flat = np.abs(cloth.hinge_rest_angles) < 1e-6
    assert flat.sum() >= 30 and np.all(np.abs(cloth.hinge_rest_angles) <= np.pi)


def test_drape_coloring_is_bit_exact_and_proper(drape):
    groups = O.conflict_groups(drape)
    offsets = np.cumsum([0] + [len(g) for g in groups]).astype(np.uint32)
    members = np.asarray([v for g in groups for v in g], np.uint32)
    colors = np.empty(drape.vertex_count, np.uint32)
    count = C.c_uint32()
    N.check(N.lib.jgs2_greedy_vertex_coloring(drape.vertex_count, N.uptr(offsets), len(groups), N.uptr(members),
                                              N.uptr(colors), C.byref(count)))
    assert np.array_equal(colors, O.greedy_coloring(drape.vertex_count, groups))
    for g in groups:
        assert len({int(colors[v]) for v in g}) == len(g)


def test_oracle_cloth_gradient_matches_restricted_energy(drape):
    settings = dict(scenes.CLOTH_SETTINGS)
    ref = O.Oracle(drape, 64)
    ref.step_frames(6, scenes.CLOTH_ITERATIONS, settings)
    g, H = ref.local_systems(settings)
    off = drape.extras["cloth_offset"]
    free = np.nonzero(drape.vertex_info.reshape(-1, 4)[:, 2] == 0)[0]
    assert np.all(free >= off)
    for v in free[::4]:
        fd = np.empty(3)
        for a in range(3):
            h = 1e-6
            d = np.zeros(3)
            d[a] = h
            ep = ref.restricted_model(int(v), d, settings)[2]
            d[a] = -h
            em = ref.restricted_model(int(v), d, settings)[2]
            fd[a] = (ep - em) / (2 * h)
        assert relative_error(g[v], fd) < 1e-6, v
        assert np.allclose(H[v], H[v].T, atol=1e-9 * np.abs(H[v]).max())


def test_oracle_cloth_fp32_tracks_fp64(drape):
    settings = dict(scenes.CLOTH_SETTINGS)
    a, b = O.Oracle(drape, 32), O.Oracle(drape, 64)
    a.step_frames(20, scenes.CLOTH_ITERATIONS, settings)
    b.step_frames(20, scenes.CLOTH_ITERATIONS, settings)
    xa, xb = a.positions(), b.positions()
    x0 = drape.positions.reshape(-1, 4)[:, :3]
    assert np.isfinite(xb).all()
    moved = np.linalg.norm(xb - x0)
    assert moved > 1e-3  # the sheet drapes
    assert np.linalg.norm(xa - xb) / np.linalg.norm(xb - x0) < 1e-3
    off = drape.extras["cloth_offset"]
    pinned = drape.vertex_info.reshape(-1, 4)[:, 2] == 1
    assert np.array_equal(xb[pinned], x0[pinned].astype(np.float64))
    assert np.all(xb[off:, 1][~pinned[off:]] < x0[off:, 1][~pinned[off:]] + 1e-6)


def cloth_scene_with(**changes):
    scene = scenes.cloth_drape()
    c = scene.cloth
    for key, value in changes.items():
        setattr(c, key, value)
    return scene


@pytest.mark.parametrize("changes,match", [
    (dict(triangle_indices=np.array([8, 8, 9] + [0] * 93, np.uint32)), "must reference 3 distinct vertices"),
    (dict(thickness=0.0), "thickness"),
    (dict(bending_stiffness=-1.0), "bending stiffness"),
    (dict(poisson_ratio=0.5), "Poisson ratio"),
    (dict(triangle_rest_areas=np.zeros(32, np.float32)), r"triangleRestAreas\[0\] must be positive"),
    (dict(hinge_rest_edge_lengths=np.zeros(40, np.float32)), r"hingeRestEdgeLengths\[0\] must be positive"),
])
def test_cloth_validation(changes, match):
    scene = cloth_scene_with(**changes)
    with pytest.raises(jgs2.JGS2RangeError, match=match):
        jgs2.JGS2GpuSolver.create(scene, scenes.CLOTH_SETTINGS)


def test_cloth_index_out_of_range_rejected():
    scene = scenes.cloth_drape()
    tri = scene.cloth.triangle_indices.copy()
    tri[0] = 99
    scene.cloth.triangle_indices = tri
    with pytest.raises(jgs2.JGS2RangeError, match="Triangle 0 index 99 is outside vertexCount 33"):
        jgs2.JGS2GpuSolver.create(scene, scenes.CLOTH_SETTINGS)


def test_drape_contact_surface_and_superset_match_the_reference_scene():
    # scenes.test.ts:531-620 and surfaceWithCloth (precompute.ts:98-133):
    # the surface ends with the cloth triangles in their packed order, holds
    # every cloth edge, and the static superset starts with its nearest pair
    # inside (0.003, 0.08).
    from ipc_geometry import candidate_distances

    scene = scenes.cloth_drape(contact="superset")
    surf = scene.extras["surface"]
    tris = surf["triangles"].reshape(-1, 3)
    cloth_tris = scene.cloth.triangle_indices.reshape(-1, 3)
    assert len(cloth_tris) == 32 and scene.cloth.hinge_count == 40
    assert np.array_equal(tris[-len(cloth_tris):], cloth_tris)
    assert len(tris) - len(cloth_tris) == 12  # the collider cube's boundary faces
    edges = {tuple(e) for e in surf["edges"].reshape(-1, 2)}
    assert len(edges) == len(surf["edges"]) // 2
    for t in cloth_tris:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            assert (min(a, b), max(a, b)) in edges
    d = candidate_distances(scene.positions.reshape(-1, 4)[:, :3], scene.contact)
    assert 0.003 < d.min() < 0.08
    # The superset is the one the library builds from that surface.
    again = scenes.static_contact_candidates(scene.vertex_count, surf["triangles"], surf["edges"])
    assert np.array_equal(again.packed_indices, scene.contact.packed_indices)
