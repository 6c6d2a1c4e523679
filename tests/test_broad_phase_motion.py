"""The broad phase's rebuild rule on the span of the motion (CPU).

A listed pair set stays complete while no unlisted pair's AABB lower bound
(ipcAabbDistanceLowerBound, jgs2-shader.ts:525-550) can have fallen to the
activation reach. The rule in k_bp_decide takes, per axis, the smallest
and largest motion lo_a, hi_a of any vertex since the build and bounds the
fall of every pair's bound by |hi - lo|: AABB gaps are invariant under a
common translation, and a box face moves by at most the span of its
vertices' motions relative to any other face. The old rule's 2|m|
(m the per-axis largest |motion|) is never smaller. These tests check both
claims on seeded primitives and motions, in float64, including motions
that are a pure translation (span 0: no bound may fall at all)."""

import numpy as np
import pytest


def aabb_lower_bound(a, b):
    """Euclidean norm of the per-axis gaps of the AABBs of point sets a, b."""
    lo_a, hi_a, lo_b, hi_b = a.min(0), a.max(0), b.min(0), b.max(0)
    gap = np.maximum(np.maximum(lo_b - hi_a, lo_a - hi_b), 0.0)
    return float(np.sqrt((gap * gap).sum()))


def primitive_pairs(rng, n):
    """Random point-triangle and edge-edge pairs (vertex index tuples) over n vertices."""
    pairs = []
    for _ in range(200):
        if rng.random() < 0.5:
            pairs.append((rng.choice(n, 1, replace=False), rng.choice(n, 3, replace=False)))
        else:
            pairs.append((rng.choice(n, 2, replace=False), rng.choice(n, 2, replace=False)))
    return pairs


@pytest.mark.parametrize("seed", range(6))
def test_bound_falls_by_at_most_the_motion_span(seed):
    rng = np.random.default_rng(seed)
    n = 64
    x = rng.uniform(-1.0, 1.0, (n, 3))
    base = rng.normal(0.0, 0.5, 3)  # a common translation
    scale = [0.0, 1e-3, 0.05, 0.3][seed % 4]
    d = base + rng.uniform(-1.0, 1.0, (n, 3)) * scale * rng.uniform(0.1, 1.0, 3)
    if seed % 3 == 2:
        d[rng.choice(n, 4, replace=False)] = 0.0  # pinned vertices
    lo, hi = d.min(0), d.max(0)
    span = float(np.linalg.norm(hi - lo))
    old = 2.0 * float(np.linalg.norm(np.abs(d).max(0)))
    assert span <= old + 1e-15
    y = x + d
    worst = 0.0
    for p, q in primitive_pairs(rng, n):
        fall = aabb_lower_bound(x[p], x[q]) - aabb_lower_bound(y[p], y[q])
        worst = max(worst, fall)
        assert fall <= span + 1e-12
    if scale == 0.0 and seed % 3 != 2:  # a pure translation moves no bound
        assert worst <= 1e-12


def test_span_rule_is_tight_for_two_vertices():
    """Two points moving apart along x by s: the bound falls by exactly the
    span when they move toward each other."""
    a, b = np.array([[0.0, 0.0, 0.0]]), np.array([[1.0, 0.0, 0.0]])
    s = 0.25
    d = np.array([[s / 2, 0.0, 0.0], [-s / 2, 0.0, 0.0]])
    before = aabb_lower_bound(a, b)
    after = aabb_lower_bound(a + d[0], b + d[1])
    span = float(np.linalg.norm(d.max(0) - d.min(0)))
    assert before - after == pytest.approx(span)
