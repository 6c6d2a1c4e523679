"""Pins of the cubature precompute oracle (oracle/jgs2_oracle_precompute.cpp).

* The reference's 12-tetrahedron beam fixture (manifests/phase1-cubature.v1.json,
  kept as tests/golden/phase1-cubature.v1.json; checked by the reference in
  src/reproducibility/phase1-cubature-manifest.test.ts:44-171): the first
  training and last validation pose anchors, and every packed selection
  (tetrahedra in slot order, f32 weights bit for bit, candidate counts, column
  ranks, packed residual).
* The reference's selector known answers (src/simulation/cpu/cubature.test.ts:8-160).
* The exact basis identities and the mode properties the reference tests
  (src/simulation/cpu/precompute.test.ts:334-436), plus the lowest rest-Hessian
  eigenvalues from a direct eigensolver and NNLS optima from scipy.
* The co-rotated linear selection: the reference's bounds
  (src/simulation/cpu/precompute.test.ts:580-615); it has no golden vector.
"""

import json
from pathlib import Path

import numpy as np
import pytest
from scipy.optimize import nnls

import jgs2_oracle as O

MANIFEST = json.loads((Path(__file__).parent / "golden" / "phase1-cubature.v1.json").read_text())


def beam():
    g = MANIFEST["fixture"]["generator"]["parameters"]
    m = MANIFEST["fixture"]["material"]
    return (np.array(g["positions"], np.float64), np.array(g["tetrahedra"]), np.array(g["fixed"]),
            np.array(g["materialIds"]), [(m["density"], m["youngModulus"], m["poissonRatio"], 1)],
            MANIFEST["fixture"]["simulation"]["timestep"])


def cuboid(cells, origin, size, fixed_fn):
    """generateRegularCuboidMesh (src/simulation/cpu/mesh.ts:101-200): vertices
    x-fastest, six tetrahedra per cell around the v000-v111 diagonal."""
    nx, ny, nz = cells
    idx = lambda x, y, z: x + y * (nx + 1) + z * (nx + 1) * (ny + 1)
    pos, fixed = [], []
    for z in range(nz + 1):
        for y in range(ny + 1):
            for x in range(nx + 1):
                p = [origin[0] + size[0] * x / nx, origin[1] + size[1] * y / ny, origin[2] + size[2] * z / nz]
                pos.append(p)
                fixed.append(1 if fixed_fn(p, (x, y, z)) else 0)
    tets = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                v000, v100, v010, v110 = idx(x, y, z), idx(x + 1, y, z), idx(x, y + 1, z), idx(x + 1, y + 1, z)
                v001, v101, v011, v111 = idx(x, y, z + 1), idx(x + 1, y, z + 1), idx(x, y + 1, z + 1), idx(x + 1, y + 1, z + 1)
                tets += [[v000, v100, v110, v111], [v000, v110, v010, v111], [v000, v010, v011, v111],
                         [v000, v011, v001, v111], [v000, v001, v101, v111], [v000, v101, v100, v111]]
    return np.array(pos, np.float64), np.array(tets), np.array(fixed)


def test_cuboid_generator_matches_the_fixture():
    pos, tets, fixed = cuboid((2, 1, 1), (0, 0, 0), (1, 1, 1), lambda p, c: c[0] == 0 and c[1] == 0)
    x, t, f, *_ = beam()
    assert np.array_equal(pos.reshape(-1), x) and np.array_equal(tets.reshape(-1), t) and np.array_equal(fixed, f)


def test_pose_corpus_anchors():
    training, validation = O.cubature_corpus(*beam())
    t = MANIFEST["training"]
    v = MANIFEST["validation"]
    assert len(training) == t["count"] == t["validCount"]
    assert len(validation) == v["count"]
    for pose, anchor in ((training[0], MANIFEST["anchors"]["firstTraining"]),
                         (validation[-1], MANIFEST["anchors"]["lastValidation"])):
        np.testing.assert_allclose(pose[0], anchor["positions"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(pose[1], anchor["predictedPositions"], rtol=0, atol=1e-12)
        assert abs(pose[2] - anchor["minimumDeformationDeterminant"]) <= 1e-12
    assert min(p[2] for p in training + validation) >= v["minimumDeformationDeterminant"]


def test_packed_selections_match_the_reference():
    args = beam()
    K = MANIFEST["cubature"]["maximumSamples"]
    tets, w, basis, stats = O.precompute_cubature(*args, K)
    fixed = args[2]
    expected = {e["vertex"]: e for e in MANIFEST["expectedPackedSelections"]}
    assert sorted(expected) == [v for v in range(len(fixed)) if not fixed[v]]
    for v, e in expected.items():
        n = len(e["tetrahedra"])
        assert list(tets[v, :n]) == e["tetrahedra"], v
        assert np.all(tets[v, n:] == 0xFFFFFFFF)
        assert np.array_equal(w[v, :n], np.array(e["weights"], np.float32)), v  # f32 bit for bit
        assert stats[v, 4] == e["nonzeroCandidateCount"]
        assert stats[v, 5] == e["trainingColumnRank"]
        assert stats[v, 7] == e["packedTrainingColumnRank"]
        assert stats[v, 0] == pytest.approx(e["packedResidual"], rel=1e-9)
        assert stats[v, 0] <= MANIFEST["cubature"]["normalizedTrainingResidual"]
        assert stats[v, 2] == 16 and stats[v, 3] == 0
    # the packed basis blocks are the f32 rest basis of each sample's vertices
    x, t4, f, mid, mats, h = args
    for v in (1, 11):
        U = O.exact_basis(*args, v).astype(np.float32)
        for k in range(K):
            tet = t4.reshape(-1, 4)[tets[v, k]]
            assert np.array_equal(basis[v, k].reshape(4, 9), U[tet].reshape(4, 9))


@pytest.mark.parametrize("which", ["beam", "minimal", "layered"])
def test_exact_basis_identities(which):
    if which == "beam":
        args = beam()
    elif which == "minimal":  # src/scenes/index.ts:330-336 with the soft rubber (:124-136)
        pos, tets, fixed = cuboid((1, 1, 1), (-1, 1.35, -0.4), (2, 0.8, 0.8), lambda p, c: c[0] == 0)
        args = (pos, tets, fixed, np.zeros(len(tets)), [(850, 35000, 0.35, 1)], 1 / 30)
    else:
        pos, tets, fixed = cuboid((4, 2, 2), (0, 0, 0), (2, 1, 1), lambda p, c: c[0] == 0)
        mid = (np.arange(len(tets)) // 24) % 2
        args = (pos, tets, fixed, mid, [(1000, 1e5, 0.4, 1), (1000, 1e7, 0.3, 1)], 1 / 60)
    fixed = np.asarray(args[2])
    free = np.flatnonzero(fixed == 0)
    _, H = O.low_frequency_modes(*args, 1, 1)
    for v in free[:: max(1, len(free) // 5)]:
        U = O.exact_basis(*args, v)
        np.testing.assert_allclose(U[v], np.eye(3), atol=1e-9)
        assert np.all(U[fixed != 0] == 0)
        Ua = U[free].reshape(-1, 3)
        E = H @ Ua
        base = 3 * int(np.searchsorted(free, v))
        schur = E[base:base + 3]
        rest = np.delete(E, np.s_[base:base + 3], axis=0)
        assert np.abs(rest).max() / np.abs(schur).max() < 1e-9  # complementary equilibrium
        np.testing.assert_allclose(schur, schur.T, rtol=1e-9, atol=1e-9 * np.abs(schur).max())
        # S^T H^-1 S is the inverse of that block
        Hinv = np.linalg.inv(H)
        np.testing.assert_allclose(np.linalg.inv(schur), Hinv[base:base + 3, base:base + 3], rtol=1e-7)


def test_low_frequency_modes_are_the_lowest_eigenvectors():
    pos, tets, fixed = cuboid((4, 2, 2), (0, 0, 0), (2, 1, 1), lambda p, c: c[0] == 0)
    args = (pos, tets, fixed, np.zeros(len(tets)), [(950, 280000, 0.32, 0)], 1 / 60)
    modes, H = O.low_frequency_modes(*args, 8, 24)
    free = np.flatnonzero(fixed == 0)
    A = np.stack([m.reshape(-1, 3)[free].reshape(-1) for m in modes])
    np.testing.assert_allclose(A @ A.T, np.eye(8), atol=1e-8)
    q = np.einsum("ij,jk,ik->i", A, H, A)
    assert np.all(q[1:] >= q[:-1] * (1 - 1e-9))
    lam = np.linalg.eigvalsh(H)[:8]
    # Rayleigh-Ritz values bound the eigenvalues from above; 24 block inverse
    # iterations converge the lowest five to 1e-7 on this mesh.
    assert np.all(q >= lam * (1 - 1e-12))
    np.testing.assert_allclose(q[:5], lam[:5], rtol=1e-7)
    for m in modes:  # sign convention: the largest-magnitude entry is positive
        assert m[np.argmax(np.abs(m))] > 0
    assert np.all(np.stack(modes).reshape(8, -1, 3)[:, fixed != 0] == 0)


def test_selector_known_answers():
    e = np.eye(3)
    sel, w, tn, rn, nr = O.select_columns([e[0], e[1], e[2]], [2, 3, 0], 2)
    assert sel == [1, 0] and list(w) == [3, 2] and rn == 0 and nr == 0 and tn == pytest.approx(np.sqrt(13), abs=1e-14)
    sel, w, *_ = O.select_columns([[0, 0], [1, 0], [1, 0]], [1, 0], 1)
    assert sel == [1] and list(w) == [1]
    sel, w, tn, rn, nr = O.select_columns([[1, 0], [0, 1]], [1, 1], 1)
    assert sel == [0] and list(w) == [1] and rn == 1 and nr == pytest.approx(1 / np.sqrt(2), abs=1e-14)
    sel, w, tn, rn, nr = O.select_columns([[1, 0], [0, 1]], [1, 1], 2, tolerance=0.8, refinement_passes=0)
    assert sel == [0] and nr == pytest.approx(1 / np.sqrt(2), abs=1e-14)
    sel, w, tn, rn, nr = O.select_columns([[1, 0], [0, 1]], [1, 1], 2, tolerance=0.0, refinement_passes=0)
    assert sel == [0, 1] and nr == 0
    sel, w, tn, rn, nr = O.select_columns([[-1, 0], [0, 1]], [1, 1], 2)
    assert sel == [1] and list(w) == [1] and nr == pytest.approx(1 / np.sqrt(2), abs=1e-14)
    sel, w, tn, rn, nr = O.select_columns([[1, 0]], [0, 0], 1)
    assert sel == [] and nr == 0
    sel, w, tn, rn, nr = O.select_columns(np.zeros((0, 2)), [2, -1], 4)
    assert sel == [] and rn == pytest.approx(np.sqrt(5), abs=1e-14) and nr == 1
    w, rn, nr = O.fit_columns([[1, 0], [1, 1]], [2, 1])
    assert list(w) == [1, 1] and rn == 0 and nr == 0
    w, rn, nr = O.fit_columns([[1, 0], [1, 1]], [0, 1])
    assert list(w) == [0, 0.5] and nr == pytest.approx(np.sqrt(0.5), abs=1e-14)


@pytest.mark.parametrize("seed", range(6))
def test_exhaustive_nnls_is_the_nnls_optimum(seed):
    rng = np.random.default_rng(seed)
    cols = rng.standard_normal((5, 9))
    target = rng.standard_normal(9)
    w, rn, nr = O.fit_columns(cols, target)
    ref, rref = nnls(cols.T, target)
    np.testing.assert_allclose(w, ref, atol=1e-10)
    assert rn == pytest.approx(rref, rel=1e-10)


def test_linear_corotated_selection_bounds():
    # the reference's bounds on a co-rotated scene (src/simulation/cpu/precompute.test.ts:580-615)
    pos, tets, fixed = cuboid((3, 2, 2), (0, 0, 0), (1.5, 1, 1), lambda p, c: c[0] == 0)
    args = (pos, tets, fixed, np.zeros(len(tets)), [(950, 280000, 0.32, 0)], 1 / 60)
    K = 6
    t, w, basis, stats = O.precompute_cubature(*args, K)
    incident = False
    for v in np.flatnonzero(fixed == 0):
        used = t[v] != 0xFFFFFFFF
        assert used.sum() <= K and np.all(w[v][used] > 0) and stats[v, 0] < 0.3
        incident |= any(v in tets[k] for k in t[v][used])
    assert incident


def test_precompute_rejects_mixed_models_and_reports_errors():
    pos, tets, fixed = cuboid((2, 1, 1), (0, 0, 0), (1, 1, 1), lambda p, c: c[0] == 0)
    mid = np.arange(len(tets)) % 2
    with pytest.raises(O.PrecomputeError, match="Mixed stable Neo-Hookean"):
        O.precompute_cubature(pos, tets, fixed, mid, [(1000, 1e5, 0.3, 1), (1000, 1e5, 0.3, 0)], 1 / 60, 6)
    with pytest.raises(O.PrecomputeError, match="at least one movable vertex"):
        O.precompute_cubature(pos, tets, np.ones(len(pos)), np.zeros(len(tets)), [(1000, 1e5, 0.3, 1)], 1 / 60, 6)
