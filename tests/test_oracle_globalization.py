"""Pins of the oracle's globalized 3x3 Newton direction
(jgs2-globalization-wgsl.ts:328-492) against the reference's own unit cases
(src/simulation/cpu/nonlinear-globalization.test.ts:39-140) and closed forms.
"""

import numpy as np
import pytest

import jgs2_oracle as O

ACCEPTED, ZERO_GRADIENT, SHIFT_LIMIT, NON_DESCENT, NONFINITE, CHOLESKY_FAILED = 0, 2, 3, 4, 6, 7
FLOOR = 2.0 ** -16
MAX_SHIFT = 1e-3


def test_minimum_eigenvalue_diagonal_and_coupled():
    # nonlinear-globalization.test.ts:40-56
    assert O.min_eigenvalue(np.diag([3.0, -2.0, 5.0])) == -2.0
    lo, hi = -5e-4, 2.0
    mean, off = 0.5 * (lo + hi), 0.5 * (lo - hi)
    coupled = np.array([[mean, off, 0], [off, mean, 0], [0, 0, 4.0]])
    # The reference scales before the closed form; feed the scaled matrix.
    scale = 4.0
    assert O.min_eigenvalue(coupled / scale) * scale == pytest.approx(lo, abs=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_minimum_eigenvalue_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(3, 3))
    s = 0.5 * (a + a.T)
    s /= np.abs(s).max()
    assert O.min_eigenvalue(s) == pytest.approx(np.linalg.eigvalsh(s)[0], abs=1e-12)


def test_asymmetry_reported():
    # nonlinear-globalization.test.ts:57-67
    status, r = O.globalization_solve(np.array([[2, 0.2, 0], [0.1, 3, 0], [0, 0, 4.0]]), [1, 0, 0], 1.0)
    assert r["maximum_relative_asymmetry"] == pytest.approx(0.025, abs=1e-12)


def test_capped_shift_and_descent():
    # nonlinear-globalization.test.ts:70-92
    H = np.diag([-5e-4, 1.0, 0.75])
    g = np.array([1.0, -2.0, 0.5])
    status, r = O.globalization_solve(H, g, 1.0)
    assert status == ACCEPTED
    assert r["minimum_eigenvalue_normalized"] * r["scale"] == pytest.approx(-5e-4, abs=1e-12)
    assert r["normalized_shift"] * r["scale"] == pytest.approx(5e-4 + FLOOR, abs=1e-12)
    assert r["normalized_shift"] <= MAX_SHIFT
    assert r["gradient_dot_direction"] < 0
    assert r["shifted_relative_residual"] < 1e-10
    # The direction solves the shifted system exactly: -(H + shift I)^-1 g / scale * scale.
    shifted = H / r["scale"] + r["normalized_shift"] * np.eye(3)
    expected = np.linalg.solve(shifted, -g / r["scale"])
    assert np.allclose(r["direction"], expected, rtol=1e-12)


def test_rejects_rather_than_raising_shift_cap():
    # nonlinear-globalization.test.ts:94-125
    status, r = O.globalization_solve(np.diag([-2e-3, 1.0, 1.0]), [1, 0, 0], 1.0)
    assert status == SHIFT_LIMIT
    assert r["normalized_shift"] > MAX_SHIFT
    assert np.array_equal(r["direction"], [0, 0, 0])
    status, r = O.globalization_solve(np.diag([FLOOR - MAX_SHIFT, 1.0, 1.0]), [1, 0, 0], 1.0)
    assert status == ACCEPTED
    assert r["normalized_shift"] == pytest.approx(MAX_SHIFT, abs=1e-14)


def test_scale_invariance():
    # nonlinear-globalization.test.ts:127-150
    base = np.diag([-5e-4, 1.0, 0.75])
    g = np.array([1.0, -2.0, 0.5])
    ref = None
    for scale in (1e-12, 1.0, 1e12):
        status, r = O.globalization_solve(base * scale, g * scale, scale)
        assert status == ACCEPTED
        if ref is None:
            ref = r
        assert np.allclose(r["direction"], ref["direction"], rtol=1e-9)
        assert r["normalized_shift"] == pytest.approx(ref["normalized_shift"], rel=1e-9)


def test_status_codes_for_degenerate_inputs():
    assert O.globalization_solve(np.eye(3), [0, 0, 0], 1.0)[0] == ZERO_GRADIENT
    assert O.globalization_solve(np.full((3, 3), np.nan), [1, 0, 0], 1.0)[0] == NONFINITE
    assert O.globalization_solve(np.eye(3), [1, 0, 0], 0.0)[0] == NONFINITE
    # Positive definite SPD system: Newton direction is -H^-1 g.
    rng = np.random.default_rng(3)
    a = rng.normal(size=(3, 3))
    H = a @ a.T + 3 * np.eye(3)
    g = rng.normal(size=3)
    status, r = O.globalization_solve(H, g, 1.0)
    assert status == ACCEPTED
    assert r["normalized_shift"] == 0.0
    assert np.allclose(r["direction"], np.linalg.solve(H, -g), rtol=1e-10)
