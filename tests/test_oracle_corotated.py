"""The co-rotated linear path of the oracle, pinned without the reference's
code:

* the packed 12x12 rest stiffness of every tetrahedron equals the textbook
  closed form of isotropic linear elasticity, K_ab = vol (lambda g_a g_b^T +
  mu g_b g_a^T + mu (g_a . g_b) I) with g the shape gradients, a different
  route from the reference's B^T C B (src/simulation/cpu/fem.ts:89-165); it is
  symmetric, positive semidefinite of rank 6 and annihilates the rigid
  translations and infinitesimal rotations;
* the oracle's per-vertex co-rotated gradient and Hessian
  (tetGradient / stiffnessBlock, jgs2-shader.ts:1022-1043, :1163-1198,
  :3083-3111) equal a numpy evaluation that takes each tetrahedron's rotation
  from an SVD polar decomposition: g = m/h^2 (x - x~) + sum R K_{s,:}
  (R^T (x - c) - (X - C)), H = m/h^2 I + sum R K_ss R^T;
* under a rigid rotation of pose and prediction the elastic force vanishes
  and (g, H) rotate as Q g, Q H Q^T.
"""

import numpy as np
import pytest

from jgs2 import _native as N
from jgs2 import scenes
import jgs2_oracle as O

YOUNG, POISSON, DENSITY = 2.0e5, 0.3, 1000.0
SETTINGS = dict(scenes.CANTILEVER_SETTINGS, schedule="jacobi")


def mesh():
    rng = np.random.default_rng(3)
    scene = scenes.cantilever_beam(cells=(3, 2, 2), size=(0.9, 0.6, 0.6), cubature_k=0,
                                   model=N.MATERIAL_COROTATED_LINEAR, young=YOUNG, poisson=POISSON)
    # A jittered copy through the generic mesh packer, so the elements are
    # not all congruent.
    x = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    x = x + rng.uniform(-0.03, 0.03, x.shape)
    tets = scene.tet_indices.reshape(-1, 4)
    fixed = scene.vertex_info.reshape(-1, 4)[:, 2]
    return scenes.from_mesh(x, tets, [(DENSITY, YOUNG, POISSON, N.MATERIAL_COROTATED_LINEAR)], fixed=fixed,
                            rest_stiffness=True)


def lame():
    lam = YOUNG * POISSON / ((1 + POISSON) * (1 - 2 * POISSON))
    mu = YOUNG / (2 * (1 + POISSON))
    return lam, mu


def shape_gradients(X4):
    Dm = np.stack([X4[1] - X4[0], X4[2] - X4[0], X4[3] - X4[0]], axis=1)
    inv = np.linalg.inv(Dm)
    g123 = inv  # rows are the gradients of N1..N3
    g0 = -g123.sum(0)
    return np.vstack([g0, g123]), abs(np.linalg.det(Dm)) / 6.0


def closed_form_stiffness(X4):
    lam, mu = lame()
    g, vol = shape_gradients(X4)
    K = np.zeros((12, 12))
    for a in range(4):
        for b in range(4):
            K[3 * a:3 * a + 3, 3 * b:3 * b + 3] = vol * (lam * np.outer(g[a], g[b]) + mu * np.outer(g[b], g[a])
                                                         + mu * np.dot(g[a], g[b]) * np.eye(3))
    return K


def test_packed_rest_stiffness_is_the_closed_form_and_has_the_rigid_null_space():
    scene = mesh()
    X = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    tets = scene.tet_indices.reshape(-1, 4)
    packed = scene.tet_rest_stiffness.reshape(-1, 12, 12).astype(np.float64)
    for t, ix in enumerate(tets):
        X4 = X[ix]
        K = closed_form_stiffness(X4)
        assert np.abs(packed[t] - K).max() <= 2e-6 * np.abs(K).max()  # f32 packing of the f64 value
        assert np.allclose(K, K.T, atol=1e-9 * np.abs(K).max())
        rigid = []
        for a in range(3):
            e = np.zeros(3); e[a] = 1
            rigid.append(np.tile(e, 4))
            rigid.append(np.cross(e, X4 - X4.mean(0)).reshape(-1))
        for m in rigid:
            assert np.abs(K @ m).max() <= 1e-9 * np.abs(K).max() * np.abs(m).max()
        w = np.linalg.eigvalsh(K)
        assert (w > -1e-9 * w.max()).all() and (w > 1e-8 * w.max()).sum() == 6


def polar(F):
    U, _, Vt = np.linalg.svd(F)
    R = U @ Vt
    if np.linalg.det(R) < 0:
        U[:, -1] *= -1
        R = U @ Vt
    return R


def numpy_local_systems(scene, x, xp, dt):
    X = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float32).astype(np.float64)
    m = scene.vertex_rest.reshape(-1, 4)[:, 3].astype(np.float64)
    tets = scene.tet_indices.reshape(-1, 4)
    K = scene.tet_rest_stiffness.reshape(-1, 12, 12).astype(np.float64)
    inv_dm = scene.tet_inverse_dm.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)  # columns of Dm^-1
    V = len(X)
    g = (m / dt ** 2)[:, None] * (x - xp)
    H = np.repeat((np.eye(3))[None], V, 0) * (m / dt ** 2)[:, None, None]
    for t, ix in enumerate(tets):
        Ds = np.stack([x[ix[1]] - x[ix[0]], x[ix[2]] - x[ix[0]], x[ix[3]] - x[ix[0]]], axis=1)
        R = polar(Ds @ inv_dm[t].T)
        u = (R.T @ (x[ix] - x[ix].mean(0)).T).T - (X[ix] - X[ix].mean(0))
        f = (K[t] @ u.reshape(-1)).reshape(4, 3)
        for s, v in enumerate(ix):
            g[v] += R @ f[s]
            H[v] += R @ K[t][3 * s:3 * s + 3, 3 * s:3 * s + 3] @ R.T
    return g, H


def deformed_pose(scene, rotation=np.eye(3), seed=9, amplitude=0.02):
    rng = np.random.default_rng(seed)
    X = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    x = X + amplitude * rng.standard_normal(X.shape)
    xp = x + 0.01 * rng.standard_normal(X.shape)
    return x @ rotation.T, xp @ rotation.T


def pose4(x):
    return np.concatenate([x, np.ones((len(x), 1))], 1).astype(np.float32).reshape(-1)


def oracle_systems(scene, x, xp):
    oracle = O.Oracle(scene, 64)
    assert not oracle.globalization  # the legacy co-rotated path
    oracle.set_pose(pose4(x), pose4(xp))
    x32 = x.astype(np.float32).astype(np.float64)
    xp32 = xp.astype(np.float32).astype(np.float64)
    g, H = oracle.local_systems(SETTINGS)
    return g, H, x32, xp32


def test_oracle_corotated_local_systems_match_an_svd_polar_evaluation():
    scene = mesh()
    free = scene.vertex_info.reshape(-1, 4)[:, 2] == 0
    x, xp = deformed_pose(scene)
    g, H, x32, xp32 = oracle_systems(scene, x, xp)
    g_ref, H_ref = numpy_local_systems(scene, x32, xp32, SETTINGS["timestep"])
    scale_g = np.abs(g_ref[free]).max()
    scale_h = np.abs(H_ref[free]).max()
    assert np.abs(g[free] - g_ref[free]).max() <= 1e-9 * scale_g
    assert np.abs(H[free] - H_ref[free]).max() <= 1e-9 * scale_h


def test_rigid_rotation_has_no_elastic_force_and_rotates_the_local_systems():
    scene = mesh()
    free = scene.vertex_info.reshape(-1, 4)[:, 2] == 0
    th = 0.7
    Q = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]]) @ \
        np.array([[1, 0, 0], [0, np.cos(0.4), -np.sin(0.4)], [0, np.sin(0.4), np.cos(0.4)]])
    X = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    m = scene.vertex_rest.reshape(-1, 4)[:, 3].astype(np.float64)
    # Rigid pose, prediction equal to the pose: the elastic force is zero.
    x = X @ Q.T
    g, H, x32, _ = oracle_systems(scene, x, x)
    assert np.abs(g[free]).max() <= 1e-6 * np.abs(H[free]).max() * 1e-2
    # Deformed pose and its rotated copy: g -> Q g, H -> Q H Q^T.
    x0, xp0 = deformed_pose(scene)
    g0, H0, _, _ = oracle_systems(scene, x0, xp0)
    g1, H1, _, _ = oracle_systems(scene, x0 @ Q.T, xp0 @ Q.T)
    assert np.abs(g1[free] - g0[free] @ Q.T).max() <= 1e-5 * np.abs(g0[free]).max()
    HQ = np.einsum("ij,vjk,lk->vil", Q, H0[free], Q)
    assert np.abs(H1[free] - HQ).max() <= 1e-6 * np.abs(H0[free]).max()
    assert m.min() > 0
