"""The oracle's assembled accept/revert (candidateTetrahedron +
reduceCandidate + applyCandidate, jgs2-shader.ts:3499-3983) and convergence
reduction (reduceConvergence, jgs2-shader.ts:4106-4486), pinned to the
reference's own known answers for the CPU twins of those passes,
acceptOrRevertFeasibleJacobiIteration and evaluateJGS2Convergence
(src/simulation/cpu/nonlinear-iteration.test.ts:52-388) and the assembled
revert gate of manifests/phase1-globalization.v1.json (referenceGates
assembledRevertDeterminant = -0.2100000000000002).

The GPU passes keep the reference's f32 finiteness rule (|v| <= 3e38), so
the reference's Number.MAX_VALUE inputs become FLT_MAX here and its
representable 1e200 norm becomes 1e30; its dimension-validation case
(:224-264) has no counterpart in the fixed-layout device records.
"""

import numpy as np
import pytest

from jgs2 import _native as N
from jgs2 import scenes
import jgs2_oracle as O

FLT_MAX = float(np.finfo(np.float32).max)
MATERIAL = [(1000.0, 10000.0, 0.3, N.MATERIAL_STABLE_NEO_HOOKEAN)]  # nonlinear-iteration.test.ts:12-19
MESH = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)  # :21-32
MULTI = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                  [3, 0, 0], [4, 0, 0], [3, 1, 0], [3, 0, 1]], np.float64)  # :34-50


def one_tet():
    return scenes.from_mesh(MESH, [[0, 1, 2, 3]], MATERIAL)


def two_tets():
    return scenes.from_mesh(MULTI, [[0, 1, 2, 3], [4, 5, 6, 7]], MATERIAL, body_ids=[0, 0, 0, 0, 1, 1, 1, 1])


@pytest.fixture(params=[32, 64], ids=["f32", "f64"])
def precision(request):
    return request.param


def accept(ctl):
    return ctl[1, 0] == 1.0, ctl[1, 1] == 1.0


def test_individually_feasible_updates_that_invert_together_revert_the_whole_pose(precision):
    # :53-100. Each single move keeps J = 1; both together give
    # J = 1 - 1.1^2 = -0.21, so the assembled pose reverts.
    oracle = O.Oracle(one_tet(), precision)
    move_b = MESH.copy(); move_b[1, 1] += 1.1
    move_c = MESH.copy(); move_c[2, 0] += 1.1
    for move in (move_b, move_c):
        ctl, pos, _ = oracle.assembled_candidate(MESH, move)
        assert accept(ctl) == (True, False)
        assert np.array_equal(pos, np.float32(move) if precision == 32 else move)
    both = MESH.copy(); both[1, 1] += 1.1; both[2, 0] += 1.1
    ctl, pos, upd = oracle.assembled_candidate(MESH, both)
    assert accept(ctl) == (False, True)
    assert ctl[0, 1] < 0  # candidate minimum determinant
    assert ctl[0, 2] == 1.0  # accepted minimum = the source's
    if precision == 64:
        assert ctl[0, 1] == -0.2100000000000002  # the manifest's assembledRevertDeterminant
    else:
        assert abs(ctl[0, 1] + 0.21) < 1e-6
    assert np.array_equal(pos, MESH)  # every vertex back at the source, not a subset
    assert np.array_equal(upd, np.tile([0.0, 0.0, 0.0, 1.0], (4, 1)))  # reverted updates are zero
    assert ctl[0, 3] == 15  # every validity bit kept: the revert is a decision, not a failure


def test_overflowing_determinant_is_an_invalid_candidate_with_finite_sentinels(precision):
    # :102-127: the candidate's determinant overflows; the candidate minimum
    # is a zero sentinel flagged invalid, the accepted minimum stays finite.
    oracle = O.Oracle(one_tet(), precision)
    cand = np.array([[0, 0, 0], [FLT_MAX, 0, 0], [0, FLT_MAX, 0], [0, 0, FLT_MAX]])
    ctl, pos, _ = oracle.assembled_candidate(MESH, cand)
    assert accept(ctl) == (False, True)
    bits = int(ctl[0, 3])
    assert bits & 1 and not bits & 2  # source geometry valid, candidate geometry invalid
    assert ctl[0, 1] == 0.0
    assert np.isfinite(ctl[0, 2]) and ctl[0, 2] == 1.0
    assert ctl[2, 3] == 0.0  # energy pair not valid
    assert ctl[2, 1] == 0.0 and ctl[2, 2] == ctl[2, 0] == 0.0  # zero sentinels, accepted = source (rest) energy
    assert np.array_equal(pos, MESH)


def test_positive_overflow_after_a_finite_tetrahedron_rejects_the_pose(precision):
    # :129-158: the second tetrahedron's determinant is +inf; a plain
    # minimum over tetrahedra would keep the first one's J = 1.
    oracle = O.Oracle(two_tets(), precision)
    cand = MULTI.copy()
    cand[4:] = [[3, 0, 0], [FLT_MAX, 0, 0], [3, FLT_MAX, 0], [3, 0, FLT_MAX]]
    ctl, pos, _ = oracle.assembled_candidate(MULTI, cand)
    assert accept(ctl) == (False, True)
    assert not int(ctl[0, 3]) & 2
    assert ctl[0, 1] == 0.0
    assert np.array_equal(pos, MULTI)


def test_assembled_energy_is_recorded_but_not_a_gate(precision):
    # :160-185: a stretched candidate with higher energy is accepted.
    oracle = O.Oracle(one_tet(), precision)
    cand = MESH.copy(); cand[3, 2] = 1.1
    ctl, pos, _ = oracle.assembled_candidate(MESH, cand)
    assert accept(ctl) == (True, False)
    assert ctl[2, 3] == 1.0  # energy pair valid
    assert ctl[2, 1] > ctl[2, 0] == 0.0  # candidate energy above the (zero) rest energy
    assert ctl[2, 2] == ctl[2, 1]  # accepted energy = candidate energy
    # Vertex energies are summed into the pair without affecting the decision.
    ve = np.tile([0.5, 0.75, 1.0, 1.0], (4, 1))  # sums 2 (source) and 3 (candidate)
    ctl2, _, _ = oracle.assembled_candidate(MESH, cand, ve)
    assert accept(ctl2) == (True, False)
    assert abs(ctl2[2, 0] - 2.0) < 1e-6 and abs(ctl2[2, 1] - (ctl[2, 1] + 3.0)) < 1e-5
    # A non-finite vertex energy invalidates the pair (the CPU twin throws).
    ve[0, 1] = np.nan
    ctl3, _, _ = oracle.assembled_candidate(MESH, cand, ve)
    assert ctl3[2, 3] == 0.0 and accept(ctl3) == (True, False)


# ---------------------------------------------------------------- convergence

FEASIBLE = np.array([[1, 1, 1, 15], [1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 0]], np.float64)


def conv(precision, inertia, material, updates=(0, 0, 0), scene_scale=1.0, res_tol=1e-5, upd_tol=1e-5,
         control=FEASIBLE, external=(0, 0, 0)):
    """One-vertex case of the reference's tests on the 4-vertex mesh: the
    other vertices carry zero gradients and updates, which change no norm,
    maximum or sum."""
    oracle = O.Oracle(one_tet(), precision)
    g = np.zeros((4, 5, 3))
    g[0, 0], g[0, 1], g[0, 2] = inertia, material, external
    u = np.zeros((4, 3))
    u[0] = updates
    row, ctl = oracle.convergence(g, u, control, scene_scale, res_tol, upd_tol)
    return {
        "norms": row[3].tolist() + [row[4, 0]], "gradient_norm": row[4, 1], "denominator": row[4, 2],
        "relative": row[4, 3], "max_update": row[5, 0], "normalized_update": row[5, 1],
        "residual_ok": row[5, 2] == 1, "update_ok": row[5, 3] == 1, "converged": row[6, 0] == 1,
        "finite": row[6, 1] == 1, "reverted": row[1, 3] == 1, "row": row, "control": ctl,
    }


def test_both_gates_are_required(precision):
    # :187-231 (scene scale 2, residual tolerance 1e-5, update tolerance 1e-4)
    base = dict(scene_scale=2.0, res_tol=1e-5, upd_tol=1e-4)
    ok = conv(precision, (1, 0, 0), (-1, 0, 0), (1e-5, 0, 0), **base)
    assert ok["gradient_norm"] == 0.0 and ok["denominator"] == 2.0
    assert ok["converged"] and ok["finite"]
    high_update = conv(precision, (1, 0, 0), (-1, 0, 0), (1e-3, 0, 0), **base)
    assert high_update["residual_ok"] and not high_update["update_ok"] and not high_update["converged"]
    high_res = conv(precision, (1, 0, 0), (-0.5, 0, 0), (0, 0, 0), **base)
    assert not high_res["residual_ok"] and high_res["update_ok"] and not high_res["converged"]
    assert abs(high_res["relative"] - 0.5 / 1.5) < 1e-7


def test_invalid_diagnostics_stay_finite_and_cannot_converge(precision):
    # :233-252: a NaN gradient with an infeasible, failed iteration.
    infeasible = np.array([[1, 0.5, 1, 15], [0, 1, 1, 0], [0, 0, 0, 1], [0, 0, 0, 0]], np.float64)
    r = conv(precision, (np.nan, 0, 0), (0, 0, 0), control=infeasible)
    assert not r["finite"] and not r["converged"]
    assert np.isfinite(r["relative"]) and np.isfinite(r["normalized_update"])
    assert np.isfinite(r["row"]).all()


def test_updates_are_normalized_by_max_scene_scale_and_1e_12(precision):
    # :266-282: scene scale 0, update 5e-16 -> 5e-4.
    r = conv(precision, (0, 0, 0), (0, 0, 0), (5e-16, 0, 0), scene_scale=0.0, res_tol=1e-5, upd_tol=1e-3)
    tol = 1e-14 if precision == 64 else 1e-10
    assert abs(r["normalized_update"] - 5e-4) < tol
    assert r["converged"]


def test_overflow_and_revert_cannot_converge(precision):
    # :284-357
    over = conv(precision, (FLT_MAX,) * 3, (FLT_MAX,) * 3)
    assert not over["finite"] and not over["converged"]
    assert np.isfinite(over["row"]).all()
    upd = conv(precision, (0, 0, 0), (0, 0, 0), (FLT_MAX,) * 3)
    assert not upd["finite"] and not upd["converged"]
    assert np.isfinite(upd["max_update"]) and np.isfinite(upd["normalized_update"])
    reverted = FEASIBLE.copy()
    reverted[1, 1] = 1.0  # feasible, but the assembled candidate reverted
    rv = conv(precision, (0, 0, 0), (0, 0, 0), control=reverted)
    assert rv["residual_ok"] and rv["update_ok"] and rv["reverted"] and not rv["converged"]


def test_large_representable_norms_stay_valid(precision):
    # :359-388 with the f32 range: a 1e30 gradient is kept by the scaled
    # norms although its square overflows.
    r = conv(precision, (1e30, 0, 0), (0, 0, 0))
    assert r["finite"]
    assert r["gradient_norm"] == pytest.approx(1e30, rel=1e-7)
    assert r["norms"][0] == pytest.approx(1e30, rel=1e-7)
    assert r["denominator"] == pytest.approx(1e30, rel=1e-7)
    assert r["relative"] == 1.0
    assert not r["converged"]


def test_local_failures_block_convergence_and_the_stop_latch_needs_convergence(precision):
    # converged requires no failed local solve (nonlinear-iteration.ts:322-327);
    # the stop latch is set only by a converged iteration with stop enabled.
    failed = FEASIBLE.copy()
    failed[1, 2] = 1.0
    r = conv(precision, (0, 0, 0), (0, 0, 0), control=failed)
    assert r["residual_ok"] and r["update_ok"] and not r["converged"]
    oracle = O.Oracle(one_tet(), precision)
    g = np.zeros((4, 5, 3))
    u = np.zeros((4, 3))
    _, ctl = oracle.convergence(g, u, FEASIBLE, 1.0, 1e-5, 1e-5, stop=True)
    assert int(ctl[0, 3]) & 16 and ctl[1, 3] == 1  # latched, history index advanced
    _, ctl = oracle.convergence(g, u, failed, 1.0, 1e-5, 1e-5, stop=True)
    assert not int(ctl[0, 3]) & 16
