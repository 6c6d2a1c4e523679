"""Exact FP64 diagnostics (jgs2_read_oracle_diagnostics) on the GPU.

The reference's opt-in evaluator (gpu/jgs2-diagnostics.ts:712-1110) has no
golden vectors; these pins fix it instead:

* every energy component against an independent FP64 assembly from the same
  f32 state (per-tetrahedron density from the oracle's stable Neo-Hookean
  routine, itself pinned to closed forms and finite differences);
* the gradient against central differences of the evaluator's own energy,
  and the diagonal Hessian blocks against central differences of its
  gradient (stable Neo-Hookean; the co-rotated gradient is the reference's
  frozen-rotation R K u, which is not the energy's derivative, and is
  checked against the same formula evaluated independently);
* a rest pose with x = x~ has zero energy and gradient up to the f32
  rounding of the rest data;
* in the contact trough the exact incremental potential (with the barrier
  term) decreases over every graph-colored Gauss-Seidel frame: the
  per-frame incremental-potential decrease the step is built to deliver.
"""

import numpy as np
import pytest

import jgs2
from jgs2 import scenes
import jgs2_oracle as O

pytestmark = pytest.mark.gpu


def pos4(x3):
    return np.concatenate([x3, np.ones((len(x3), 1))], 1).astype(np.float32).reshape(-1)


def hand_energy(scene, x4, xp4, settings, forces=(), targets=()):
    """Independent FP64 assembly of the implicit-Euler energy components."""
    x = np.asarray(x4, np.float32).reshape(-1, 4)[:, :3].astype(np.float64)
    xp = np.asarray(xp4, np.float32).reshape(-1, 4)[:, :3].astype(np.float64)
    rest = scene.vertex_rest.reshape(-1, 4).astype(np.float64)
    free = scene.vertex_info.reshape(-1, 4)[:, 2] == 0
    ih2 = 1.0 / settings["timestep"] ** 2
    inertia = 0.5 * ih2 * float((rest[free, 3] * ((x[free] - xp[free]) ** 2).sum(1)).sum())
    tets = scene.tet_indices.reshape(-1, 4)
    dminv = scene.tet_inverse_dm.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)  # 3 columns of Dm^-1
    meta = scene.tet_meta.reshape(-1, 4).astype(np.float64)
    elastic = 0.0
    for t, (a, b, c, d) in enumerate(tets):
        Ds = np.stack([x[b] - x[a], x[c] - x[a], x[d] - x[a]], axis=1)
        F = Ds @ dminv[t].T
        elastic += meta[t, 0] * O.snh(F, np.zeros((3, 3)), meta[t, 1], meta[t, 2])[0]
    ext = -sum(float(np.dot(np.float32(f).astype(np.float64), x[v])) for v, f in forces if free[v])
    tgt = sum(0.5 * k * float(((x[v] - np.float32(p).astype(np.float64)) ** 2).sum()) for v, p, k in targets if free[v])
    return {"inertia": inertia, "elasticity": elastic, "external_force": ext, "quadratic_target": tgt}


def test_components_match_an_independent_fp64_assembly():
    print("case: diagnostics components, cantilever 8x2x2 with force and target", flush=True)
    scene = scenes.cantilever_beam(cells=(8, 2, 2), size=(2.0, 0.5, 0.5))
    settings = dict(scenes.CANTILEVER_SETTINGS)
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    rest = scene.vertex_rest.reshape(-1, 4)[:, :3]
    forces = [(8, (0.0, 5.0, 0.0))]
    targets = [(17, tuple(rest[17] + np.array([0.0, 0.05, 0.02])), 3e3)]
    for v, f in forces:
        solver.set_external_force(v, f)
    for v, p, k in targets:
        solver.set_quadratic_target(v, p, k)
    solver.step_frames_exact_iterations(4, 8, settings)
    d = solver.read_oracle_diagnostics(settings)
    x, xp = solver.read_positions(), solver.read_predicted_positions()
    ref = hand_energy(scene, x, xp, settings, forces, targets)
    print({k: (d[k], ref[k]) for k in ref}, flush=True)
    for k, v in ref.items():
        assert d[k] == pytest.approx(v, rel=1e-9, abs=1e-9 * max(1.0, abs(v)))
    assert d["floor_contact"] == 0.0 and d["contact_barrier"] == 0.0
    total = sum(d[k] for k in ("inertia", "elasticity", "external_force", "quadratic_target", "floor_contact",
                               "contact_barrier"))
    assert d["energy"] == pytest.approx(total, rel=1e-12)
    assert d["finite"] == 1 and d["active_vertices"] == int((scene.vertex_info.reshape(-1, 4)[:, 2] == 0).sum())
    assert d["minimum_determinant_valid"] == 1
    assert 0 < d["minimum_deformation_determinant"] < 2
    assert d["relative_residual"] == pytest.approx(d["gradient_norm"] / max(1.0, d["residual_denominator"]))
    solver.destroy()


def test_gradient_and_hessian_are_derivatives_of_the_energy():
    print("case: diagnostics derivatives, stable Neo-Hookean with the floor active", flush=True)
    scene = scenes.cantilever_beam(cells=(6, 2, 2), size=(1.5, 0.5, 0.5))
    settings = dict(scenes.CANTILEVER_SETTINGS, floor_height=0.3, floor_stiffness=2e4)
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    solver.step_frames_exact_iterations(6, 8, settings)
    x = solver.read_positions().reshape(-1, 4).astype(np.float64)
    xp = solver.read_predicted_positions()
    d, vrec, _ = solver.read_oracle_diagnostics(settings, positions=x.astype(np.float32), predicted=xp, records=True)
    assert d["active_floor_contacts"] > 0  # the floor term is exercised
    free = scene.vertex_info.reshape(-1, 4)[:, 2] == 0
    rng = np.random.default_rng(7)
    delta = rng.normal(size=(scene.vertex_count, 3)) * free[:, None]
    eps = 1e-4  # small enough that the floor penalty's jump in curvature stays negligible
    xa, xb = x.copy(), x.copy()
    xa[:, :3] += eps * delta
    xb[:, :3] -= eps * delta
    xa32, xb32 = xa.astype(np.float32), xb.astype(np.float32)
    ea = solver.read_oracle_diagnostics(settings, positions=xa32, predicted=xp)["energy"]
    eb = solver.read_oracle_diagnostics(settings, positions=xb32, predicted=xp)["energy"]
    step = (xa32.astype(np.float64) - xb32.astype(np.float64))[:, :3]
    g = vrec[:, :3]
    fd, an = ea - eb, float((g * step).sum())
    print(f"directional derivative: FD {fd:.9e} analytic {an:.9e}", flush=True)
    assert fd == pytest.approx(an, rel=2e-4)
    # Hessian blocks of a few free vertices by central differences of the gradient
    away = free & (np.abs(vrec[:, 23]) > 1e-3)  # Hessian of the floor term jumps at contact
    for v in np.flatnonzero(away)[::9][:5]:
        Hfd = np.zeros((3, 3))
        for c in range(3):
            xa, xb = x.copy(), x.copy()
            xa[v, c] += 1e-4
            xb[v, c] -= 1e-4
            xa32, xb32 = xa.astype(np.float32), xb.astype(np.float32)
            ga = solver.read_oracle_diagnostics(settings, positions=xa32, predicted=xp, records=True)[1][v, :3]
            gb = solver.read_oracle_diagnostics(settings, positions=xb32, predicted=xp, records=True)[1][v, :3]
            Hfd[:, c] = (ga - gb) / (float(xa32[v, c]) - float(xb32[v, c]))
        H = vrec[v, 4:13].reshape(3, 3)
        err = np.abs(Hfd - H).max() / np.abs(H).max()
        print(f"vertex {v}: Hessian block FD rel {err:.2e}", flush=True)
        assert err < 2e-3
    solver.destroy()


def test_rest_pose_has_zero_energy():
    print("case: diagnostics at rest", flush=True)
    scene = scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5))
    settings = dict(scenes.CANTILEVER_SETTINGS, gravity=(0.0, 0.0, 0.0))
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    x = solver.read_positions()
    d = solver.read_oracle_diagnostics(settings, positions=x, predicted=x)
    mu = float(scene.tet_meta.reshape(-1, 4)[:, 2].max())
    vol = float(scene.tet_meta.reshape(-1, 4)[:, 0].sum())
    print(d, flush=True)
    assert d["inertia"] == 0.0
    assert abs(d["elasticity"]) < 1e-10 * mu * vol
    assert d["gradient_norm"] < 1e-5 * mu * vol ** (2 / 3)
    assert d["minimum_deformation_determinant"] == pytest.approx(1.0, abs=1e-6)
    solver.destroy()


def test_trough_incremental_potential_decreases_every_frame():
    print("case: trough incremental potential per frame, graph-colored Gauss-Seidel", flush=True)
    scene = scenes.trough()
    settings = dict(scenes.TROUGH_SETTINGS, schedule="graph-colored-gauss-seidel")
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    worst, rises = 0.0, 0
    for frame in range(40):
        x0 = solver.read_positions()
        solver.step_frames_exact_iterations(1, 12, settings)
        xp = solver.read_predicted_positions()
        before = solver.read_oracle_diagnostics(settings, positions=x0, predicted=xp)
        after = solver.read_oracle_diagnostics(settings)
        assert before["finite"] == 1 and after["finite"] == 1
        change = after["energy"] - before["energy"]
        scale = max(abs(before["energy"]), 1e-9)
        worst = max(worst, change / scale)
        rises += change > 1e-9 * scale
    print(f"40 frames: largest relative rise {worst:.3e}, frames with a rise {rises}; barrier pairs "
          f"{after['active_contact_pairs']}, nearest pair {after['minimum_contact_distance']:.4f}", flush=True)
    assert after["active_contact_pairs"] > 0  # the barrier term is in play
    assert rises == 0
    solver.destroy()


def test_corotated_gradient_is_the_frozen_rotation_force():
    print("case: diagnostics co-rotated gradient R K u", flush=True)
    from jgs2 import _native as N
    scene = scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5), model=N.MATERIAL_COROTATED_LINEAR)
    settings = dict(scenes.CANTILEVER_SETTINGS)
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    solver.step_frames_exact_iterations(4, 8, settings)
    d, vrec, trec = solver.read_oracle_diagnostics(settings, records=True)
    x = solver.read_positions().reshape(-1, 4)[:, :3].astype(np.float64)
    xp = solver.read_predicted_positions().reshape(-1, 4)[:, :3].astype(np.float64)
    rest = scene.vertex_rest.reshape(-1, 4).astype(np.float64)
    tets = scene.tet_indices.reshape(-1, 4)
    dminv = scene.tet_inverse_dm.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    K = scene.tet_rest_stiffness.reshape(-1, 12, 12).astype(np.float64)
    free = scene.vertex_info.reshape(-1, 4)[:, 2] == 0
    grad = np.zeros_like(x)
    energy = 0.0
    for t, ix in enumerate(tets):
        F = np.stack([x[ix[1]] - x[ix[0]], x[ix[2]] - x[ix[0]], x[ix[3]] - x[ix[0]]], axis=1) @ dminv[t].T
        U, _, Vt = np.linalg.svd(F)
        R = U @ Vt  # polar rotation (det F > 0 here)
        u = ((x[ix] - x[ix].mean(0)) @ R) - (rest[ix, :3] - rest[ix, :3].mean(0))
        f = (K[t] @ u.reshape(12)).reshape(4, 3)
        energy += 0.5 * float(u.reshape(12) @ K[t] @ u.reshape(12))
        grad[ix] += f @ R.T
    ih2 = 1.0 / settings["timestep"] ** 2
    grad[free] += ih2 * rest[free, 3:4] * (x[free] - xp[free])
    grad[~free] = 0
    err = np.abs(vrec[:, :3] - grad).max() / np.abs(grad).max()
    print(f"co-rotated: gradient rel {err:.2e}, energy {d['elasticity']:.9e} vs {energy:.9e}", flush=True)
    assert err < 1e-5  # the library's frame takes 7 polar steps, not the exact SVD
    assert d["elasticity"] == pytest.approx(energy, rel=1e-5)
    solver.destroy()
