"""Pins of the oracle's frame step (CPU).

* The local gradient and Hessian the step solves with are the first and
  second derivatives of the restricted trial model the line search
  evaluates (src/simulation/gpu/jgs2-shader.ts:2992-3263 vs :1980-2337);
  checked by central finite differences in FP64, with cubature.
* Without cubature, the restricted energy is the total implicit-Euler
  energy restricted to one vertex: checked against a direct numpy sum of
  SNH densities (closed form) and inertia.
* A scene at rest without gravity stays bit-identical (zero gradient).
* Rigid free fall of a soft free body: the implicit-Euler minimizer is the
  predicted position (elastic energy is zero for a rigid translation), so
  Jacobi sweeps converge to it when inertia dominates the local systems.
* FP32 and FP64 runs of the oracle agree after one frame.
"""

import math

import numpy as np
import pytest

import jgs2_oracle as O
from jgs2 import scenes

SETTINGS = dict(scenes.CANTILEVER_SETTINGS)


def small_beam(k=6):
    return scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5), cubature_k=k)


def deformed(scene, frames=3):
    o = O.Oracle(scene, 64)
    o.step_frames(frames, 8, SETTINGS)
    return o


@pytest.mark.parametrize("k", [0, 6])
def test_local_system_is_derivative_of_restricted_model(k):
    scene = small_beam(k)
    o = deformed(scene)
    pos = np.zeros((scene.vertex_count, 4), np.float32)
    pos[:, :3] = o.positions()
    pred = np.zeros_like(pos)
    pred[:, :3] = o.predicted()
    o.set_pose(pos.reshape(-1), pred.reshape(-1))
    g, H = o.local_systems(SETTINGS)
    free = np.nonzero(scene.vertex_info.reshape(-1, 4)[:, 2] == 0)[0]
    h = 1e-5
    for v in free[::7]:
        fd_g = np.zeros(3)
        for a in range(3):
            d = np.zeros(3)
            d[a] = h
            fd_g[a] = (o.restricted_model(v, d, SETTINGS)[2] - o.restricted_model(v, -d, SETTINGS)[2]) / (2 * h)
        assert np.allclose(fd_g, g[v], rtol=1e-5, atol=1e-6 * np.abs(H[v]).max() * h + 1e-9), (v, fd_g, g[v])
        fd_H = np.zeros((3, 3))
        hh = 1e-4
        for a in range(3):
            for b in range(3):
                da = np.zeros(3)
                db = np.zeros(3)
                da[a] = hh
                db[b] = hh
                e = lambda d: o.restricted_model(v, d, SETTINGS)[2]
                fd_H[a, b] = (e(da + db) - e(da - db) - e(db - da) + e(-da - db)) / (4 * hh * hh)
        assert np.abs(fd_H - 0.5 * (H[v] + H[v].T)).max() < 1e-4 * np.abs(H[v]).max(), v


def snh_energy(F, lam, mu):
    i1 = np.sum(F * F, axis=(1, 2))
    j = np.linalg.det(F)
    return 0.5 * mu * (i1 - 3) - 0.75 * mu * (j - 1) + 0.5 * lam * (j - 1) ** 2 - 0.5 * mu * np.log((i1 + 1) / 4)


def total_energy(scene, x, pred, dt):
    tets = scene.tet_indices.reshape(-1, 4)
    inv = scene.tet_inverse_dm.reshape(-1, 3, 4)[:, :, :3].transpose(0, 2, 1).astype(np.float64)
    meta = scene.tet_meta.reshape(-1, 4).astype(np.float64)
    ds = np.stack([x[tets[:, 1]] - x[tets[:, 0]], x[tets[:, 2]] - x[tets[:, 0]], x[tets[:, 3]] - x[tets[:, 0]]], 2)
    F = ds @ inv
    e = float(np.sum(meta[:, 0] * snh_energy(F, meta[:, 1], meta[:, 2])))
    info = scene.vertex_info.reshape(-1, 4)
    m = scene.vertex_rest.reshape(-1, 4)[:, 3].astype(np.float64)
    free = info[:, 2] == 0
    r = x - pred
    return e + float(np.sum(0.5 * m[free] / dt ** 2 * np.sum(r[free] ** 2, axis=1)))


def test_restricted_energy_without_cubature_is_total_energy_difference():
    scene = small_beam(0)
    o = deformed(scene)
    x = o.positions()
    pred = o.predicted()
    pos = np.zeros((scene.vertex_count, 4), np.float32)
    pos[:, :3] = x
    pr = np.zeros_like(pos)
    pr[:, :3] = pred
    o.set_pose(pos.reshape(-1), pr.reshape(-1))
    xs = pos[:, :3].astype(np.float64)
    ps = pr[:, :3].astype(np.float64)
    dt = np.float64(np.float32(1 / 60))
    rng = np.random.default_rng(5)
    free = np.nonzero(scene.vertex_info.reshape(-1, 4)[:, 2] == 0)[0]
    base = total_energy(scene, xs, ps, dt)
    for v in free[::5]:
        d = 1e-3 * rng.normal(size=3)
        x1 = xs.copy()
        x1[v] += d
        direct = total_energy(scene, x1, ps, dt) - base
        _, ok, model = o.restricted_model(v, d, SETTINGS)
        assert ok
        assert model == pytest.approx(direct, rel=1e-6, abs=1e-9 * abs(base))


def test_rest_without_gravity_is_stationary():
    scene = small_beam()
    o = O.Oracle(scene, 32)
    o.step_frames(2, 4, dict(SETTINGS, gravity=(0.0, 0.0, 0.0)))
    x0 = scene.positions.reshape(-1, 4)[:, :3].astype(np.float64)
    assert np.array_equal(o.positions(), x0)
    diag = o.local_diagnostics()
    free = scene.vertex_info.reshape(-1, 4)[:, 2] == 0
    assert (diag[free, 3, 3] == 2).all()  # zero-gradient status
    assert (diag[~free, 3, 3] == 1).all()  # pinned status


def test_rigid_free_fall_converges_to_prediction():
    scene = scenes.cantilever_beam(cells=(3, 2, 2), size=(0.6, 0.4, 0.4), cubature_k=0, young=100.0)
    info = scene.vertex_info.reshape(-1, 4).copy()
    info[:, 2] = 0  # nothing pinned: a free body
    scene.vertex_info = info.reshape(-1)
    o = O.Oracle(scene, 64)
    o.step_frames(1, 31, SETTINGS)
    x0 = scene.positions.reshape(-1, 4)[:, :3].astype(np.float64)
    g = np.float64(np.float32(-9.81))
    dt = np.float64(np.float32(1 / 60))
    expected = x0 + np.array([0.0, dt * dt * g, 0.0])
    assert np.abs(o.positions() - expected).max() < 1e-4 * abs(dt * dt * g)


def test_fp32_and_fp64_agree_after_one_frame():
    scene = scenes.cantilever_beam()
    a = O.Oracle(scene, 32)
    b = O.Oracle(scene, 64)
    a.step_frames(1, 8, SETTINGS)
    b.step_frames(1, 8, SETTINGS)
    x0 = scene.positions.reshape(-1, 4)[:, :3].astype(np.float64)
    err = np.linalg.norm(a.positions() - b.positions()) / np.linalg.norm(b.positions() - x0)
    assert err < 1e-4


def test_accept_and_history_records():
    scene = small_beam()
    o = O.Oracle(scene, 64)
    o.step_frames(1, 3, SETTINGS)
    ctrl = o.control()
    assert ctrl[1, 0] == 1.0 and ctrl[1, 1] == 0.0  # accepted, not reverted
    assert ctrl[0, 2] > 1e-4  # accepted minimum determinant
    hist = o.history()
    assert [hist[i, 6, 2] for i in range(3)] == [0.0, 1.0, 2.0]
    assert ctrl[1, 3] == 3.0
    assert math.isfinite(hist[2, 4, 3])
