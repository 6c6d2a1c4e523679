"""Pins of the oracle's stable Neo-Hookean material (SN = stable-neo-hookean-wgsl.ts).

Checks that a mistake in the restatement would fail: the closed-form density
at chosen F, zero energy and stress at every rotation (the corrected density
is normalized so, stable-neo-hookean-wgsl.ts:42-43), central finite
differences of energy -> stress -> tangent, and the cancellation-safe
energy delta (jgs2-globalization-wgsl.ts:529-559) against a direct
difference.
"""

import math

import numpy as np
import pytest

import jgs2_oracle as O

LAM, MU = 3.0e5, 4.0e4


def closed_form_energy(F, lam, mu):
    i1 = float(np.sum(F * F))
    j = float(np.linalg.det(F))
    return 0.5 * mu * (i1 - 3) - 0.75 * mu * (j - 1) + 0.5 * lam * (j - 1) ** 2 - 0.5 * mu * math.log((i1 + 1) / 4)


def random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


@pytest.mark.parametrize("seed", range(4))
def test_energy_matches_closed_form(seed):
    rng = np.random.default_rng(seed)
    F = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
    e, _, _ = O.snh(F, np.zeros((3, 3)), LAM, MU)
    assert e == pytest.approx(closed_form_energy(F, LAM, MU), rel=1e-12, abs=1e-9)


@pytest.mark.parametrize("seed", range(4))
def test_rest_and_rotations_are_stress_free(seed):
    R = random_rotation(np.random.default_rng(seed))
    for F in (np.eye(3), R):
        e, P, _ = O.snh(F, np.zeros((3, 3)), LAM, MU)
        assert abs(e) < 1e-10 * MU
        assert np.abs(P).max() < 1e-10 * MU


@pytest.mark.parametrize("seed", range(4))
def test_stress_and_tangent_are_derivatives(seed):
    rng = np.random.default_rng(100 + seed)
    F = np.eye(3) + 0.25 * rng.normal(size=(3, 3))
    H = rng.normal(size=(3, 3))
    e, P, DP = O.snh(F, H, LAM, MU)
    h = 1e-6
    fd_P = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            d = np.zeros((3, 3))
            d[i, j] = h
            fd_P[i, j] = (O.snh(F + d, H, LAM, MU)[0] - O.snh(F - d, H, LAM, MU)[0]) / (2 * h)
    assert np.abs(fd_P - P).max() < 1e-5 * np.abs(P).max()
    fd_DP = (O.snh(F + h * H, H, LAM, MU)[1] - O.snh(F - h * H, H, LAM, MU)[1]) / (2 * h)
    assert np.abs(fd_DP - DP).max() < 1e-4 * np.abs(DP).max()


@pytest.mark.parametrize("seed", range(4))
def test_energy_delta_equals_difference(seed):
    rng = np.random.default_rng(200 + seed)
    F0 = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
    for scale in (1e-1, 1e-3, 1e-6):
        F1 = F0 + scale * rng.normal(size=(3, 3))
        direct = closed_form_energy(F1, LAM, MU) - closed_form_energy(F0, LAM, MU)
        delta = O.snh_energy_delta(F0, F1, LAM, MU)
        assert delta == pytest.approx(direct, rel=1e-6, abs=1e-9 * MU * scale)


def test_energy_delta_keeps_small_changes_in_f32():
    # The f32 direct difference of two nearby energies cancels; the factored
    # delta keeps the first-order change (the reason for GL:510-559).
    rng = np.random.default_rng(7)
    F0 = (np.eye(3) + 0.2 * rng.normal(size=(3, 3))).astype(np.float32).astype(np.float64)
    F1 = (F0 + 1e-6 * rng.normal(size=(3, 3))).astype(np.float32).astype(np.float64)
    exact = O.snh_energy_delta(F0, F1, LAM, MU, precision=64)
    f32 = O.snh_energy_delta(F0, F1, LAM, MU, precision=32)
    assert f32 == pytest.approx(exact, rel=1e-2)
