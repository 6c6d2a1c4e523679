"""Independent closest-feature distances for checking IPC results.

Test infrastructure. The textbook Voronoi-region point-triangle test and the
clamped segment-segment solve of Ericson, Real-Time Collision Detection
(2005), sections 5.1.5 and 5.1.9, vectorized in numpy. They share no code or
formulation with the oracle's or the product's contact routines, which
follow the reference's barycentric test with an edge fallback.
"""

import numpy as np


def _dot(a, b):
    return np.einsum("ij,ij->i", a, b)


def point_triangle_distance(p, a, b, c):
    p, a, b, c = (np.asarray(v, np.float64).reshape(-1, 3) for v in (p, a, b, c))
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = _dot(ab, ap), _dot(ac, ap)
    bp = p - b
    d3, d4 = _dot(ab, bp), _dot(ac, bp)
    cp = p - c
    d5, d6 = _dot(ab, cp), _dot(ac, cp)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    closest = np.empty_like(p)
    done = np.zeros(len(p), bool)

    def take(mask, value):
        nonlocal done
        m = mask & ~done
        closest[m] = value[m]
        done |= m

    with np.errstate(divide="ignore", invalid="ignore"):
        take((d1 <= 0) & (d2 <= 0), a)
        take((d3 >= 0) & (d4 <= d3), b)
        take((vc <= 0) & (d1 >= 0) & (d3 <= 0), a + (d1 / (d1 - d3))[:, None] * ab)
        take((d6 >= 0) & (d5 <= d6), c)
        take((vb <= 0) & (d2 >= 0) & (d6 <= 0), a + (d2 / (d2 - d6))[:, None] * ac)
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        take((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0), b + w[:, None] * (c - b))
        den = 1.0 / (va + vb + vc)
        take(np.ones(len(p), bool), a + (vb * den)[:, None] * ab + (vc * den)[:, None] * ac)
    return np.linalg.norm(p - closest, axis=1)


def segment_segment_distance(p1, q1, p2, q2, eps=1e-20):
    p1, q1, p2, q2 = (np.asarray(v, np.float64).reshape(-1, 3) for v in (p1, q1, p2, q2))
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = _dot(d1, d1), _dot(d2, d2), _dot(d2, r)
    c, b = _dot(d1, r), _dot(d1, d2)
    s = np.zeros(len(p1))
    t = np.zeros(len(p1))
    with np.errstate(divide="ignore", invalid="ignore"):
        a0, e0 = a <= eps, e <= eps
        m = a0 & ~e0
        t[m] = np.clip(f[m] / e[m], 0, 1)
        m = ~a0 & e0
        s[m] = np.clip(-c[m] / a[m], 0, 1)
        g = ~a0 & ~e0
        den = a * e - b * b
        s[g] = np.where(den[g] != 0, np.clip((b[g] * f[g] - c[g] * e[g]) / den[g], 0, 1), 0.0)
        t[g] = (b[g] * s[g] + f[g]) / e[g]
        lo, hi = g & (t < 0), g & (t > 1)
        t[lo] = 0
        s[lo] = np.clip(-c[lo] / a[lo], 0, 1)
        t[hi] = 1
        s[hi] = np.clip((b[hi] - c[hi]) / a[hi], 0, 1)
    return np.linalg.norm((p1 + s[:, None] * d1) - (p2 + t[:, None] * d2), axis=1)


def candidate_distances(positions, contact):
    """Distances of every VT then EE candidate at `positions` (V x 3)."""
    x = np.asarray(positions, np.float64).reshape(-1, 3)
    ix = contact.packed_indices.reshape(-1, 4).astype(np.int64)
    nvt = contact.vertex_triangle_count
    vt, ee = ix[:nvt], ix[nvt:]
    out = np.empty(len(ix))
    out[:nvt] = point_triangle_distance(x[vt[:, 0]], x[vt[:, 1]], x[vt[:, 2]], x[vt[:, 3]])
    out[nvt:] = segment_segment_distance(x[ee[:, 0]], x[ee[:, 1]], x[ee[:, 2]], x[ee[:, 3]])
    return out
