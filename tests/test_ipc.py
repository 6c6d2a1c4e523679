"""IPC contact: candidate superset, contact geometry, barrier and friction,
and the contact step of the oracle (CPU).

Pins:
* the static VT/EE superset against the reference's enumeration example
  and order invariance (src/simulation/cpu/ipc-contact.test.ts:101-146) and
  the trough's golden counts 11,728 / 20,073 / 31,801 with 12 + 8n vertices
  and 12 + 6n tetrahedra (src/scenes/scenes.test.ts:478-528), bit-exact;
* the initial trough clearance > 0.08 (scenes.test.ts:526-528) with an
  independent textbook distance (tests/ipc_geometry.py);
* point-triangle and edge-edge distances of the reference's cases
  (ipc-contact.test.ts:148-187) and, on random configurations, against the
  independent distance; the weight identity sum_l w_l x_l = d n;
* the barrier closed form and finite-difference derivatives
  (ipc-contact.test.ts:14-63), the barrier gradient chain against a finite
  difference of kappa B(d(x)), and the friction primitives
  (ipc-contact.test.ts:65-98);
* the oracle's trough frames: no candidate closer than the minimum
  distance after any frame, every accepted local solve an Armijo decrease,
  and FP32 tracking FP64 over a short horizon.
"""

import ctypes as C

import numpy as np
import pytest

import jgs2
import jgs2_oracle as O
from jgs2 import _native as N
from jgs2 import scenes
from ipc_geometry import candidate_distances, point_triangle_distance, segment_segment_distance


def test_candidate_enumeration_matches_reference_example():
    forward = scenes.static_contact_candidates(6, [0, 1, 2, 3, 4, 5], [0, 1, 1, 2, 2, 0, 3, 4, 4, 5, 5, 3])
    assert forward.vertex_triangle_count == 6 and forward.edge_edge_count == 9
    ix = forward.packed_indices.reshape(-1, 4).tolist()
    assert ix[:6] == [[0, 3, 4, 5], [1, 3, 4, 5], [2, 3, 4, 5], [3, 0, 1, 2], [4, 0, 1, 2], [5, 0, 1, 2]]
    assert ix[6] == [0, 1, 3, 4] and ix[-1] == [1, 2, 4, 5]
    reordered = scenes.static_contact_candidates(6, [5, 3, 4, 2, 1, 0, 4, 5, 3],
                                                 [5, 4, 4, 3, 3, 5, 2, 1, 1, 0, 0, 2, 1, 2])
    assert np.array_equal(reordered.packed_indices, forward.packed_indices)
    patch = scenes.static_contact_candidates(4, [0, 1, 2, 2, 1, 3], [])
    assert patch.candidate_count == 0


@pytest.mark.parametrize("boxes", [1, 10, 20])
def test_trough_counts(boxes):
    t = scenes.trough(box_count=boxes, candidates=boxes == 10)
    assert t.vertex_count == 12 + 8 * boxes and t.tet_count == 12 + 6 * boxes
    info = t.vertex_info.reshape(-1, 4)
    assert np.array_equal(np.nonzero(info[:, 2])[0], np.arange(12))
    assert sorted(set(info[:, 3].tolist())) == list(range(boxes + 1))
    assert np.all((info[:, 3] == 0) == (info[:, 2] == 1))
    if boxes == 10:
        assert t.contact.vertex_triangle_count == 11728
        assert t.contact.edge_edge_count == 20073
        assert t.contact.packed_indices.size == 31801 * 4
        d = candidate_distances(t.positions.reshape(-1, 4)[:, :3], t.contact)
        assert d.min() > 0.08


def probe(kind, pts, dhat=1.2, kappa=1.0):
    return O.ipc_probe(kind, np.asarray(pts, np.float64).reshape(12), dhat, kappa)


def test_reference_point_triangle_cases():
    assert probe("vt", [0.25, 1, 0.25, 0, 0, 0, 1, 0, 0, 0, 0, 1])["distance"] == 1.0
    assert probe("vt", [0.25, 1, 0.25, 0, 0, 0, 2, 0, 0, 1, 0, 0])["distance"] == pytest.approx(np.sqrt(1.0625), abs=1e-14)
    assert probe("vt", [0.25, 1, 0.25, 0, 0, 0, 0, 0, 0, 0, 0, 0])["distance"] == pytest.approx(np.sqrt(1.125), abs=1e-14)


def test_reference_edge_edge_cases():
    assert probe("ee", [-1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0])["distance"] == 0.0
    assert probe("ee", [-1, 0, 0, 1, 0, 0, -1, 2, 0, 1, 2, 0])["distance"] == 2.0
    assert probe("ee", [0, 0, 0, 0, 0, 0, 0, 3, 0, 0, 4, 0])["distance"] == 3.0
    assert probe("ee", [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0])["distance"] == 0.0


@pytest.mark.parametrize("kind", ["vt", "ee"])
def test_contact_distance_weights_and_gradient(kind):
    rng = np.random.default_rng(7 if kind == "vt" else 8)
    for _ in range(200):
        x = rng.normal(size=(4, 3))
        r = probe(kind, x, dhat=10.0, kappa=3.0)
        ref = (point_triangle_distance(x[0], x[1], x[2], x[3]) if kind == "vt"
               else segment_segment_distance(x[0], x[1], x[2], x[3]))[0]
        assert r["valid"]
        assert r["distance"] == pytest.approx(ref, rel=1e-10, abs=1e-12)
        gap = (r["weights"][:, None] * x).sum(axis=0)
        assert np.allclose(gap, r["distance"] * r["normal"], atol=1e-10)
        assert np.isclose(r["weights"][:2].sum() if kind == "ee" else r["weights"][0], 1.0)
    # Gradient chain: kappa B'(d) w_l n equals d(kappa B(d(x)))/dx_l.
    x = np.array([[0.2, 0.3, 0.25], [0, 0, 0], [1, 0, 0], [0, 0, 1]]) if kind == "vt" else \
        np.array([[-1, 0.2, 0.1], [1, 0.3, -0.1], [0.1, -1, 0.4], [-0.2, 1, 0.35]])
    r = probe(kind, x, dhat=1.0, kappa=2.5)
    h = 1e-7
    fd = np.zeros(12)
    for i in range(12):
        xp, xm = x.reshape(-1).copy(), x.reshape(-1).copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (probe(kind, xp, 1.0, 2.5)["energy"] - probe(kind, xm, 1.0, 2.5)["energy"]) / (2 * h)
    assert np.allclose(r["gradient"].reshape(-1), fd, rtol=1e-6, atol=1e-8)


def test_barrier_closed_form_and_derivatives():
    dhat, d = 1.2, 0.37
    pts = [0.25, d, 0.25, 0, 0, 0, 1, 0, 0, 0, 0, 1]
    r = probe("vt", pts, dhat=dhat, kappa=1.0)
    assert r["distance"] == pytest.approx(d, abs=1e-15)
    assert r["energy"] == pytest.approx(-((d - dhat) ** 2) * np.log(d / dhat), abs=1e-14)
    step = 1e-5

    def value(y):
        return probe("vt", [0.25, y, 0.25, 0, 0, 0, 1, 0, 0, 0, 0, 1], dhat=dhat)["energy"]

    assert r["d1"] == pytest.approx((value(d + step) - value(d - step)) / (2 * step), abs=1e-8)
    assert r["d2"] == pytest.approx((value(d + step) - 2 * value(d) + value(d - step)) / step ** 2, abs=1e-4)
    outside = probe("vt", [0.25, dhat, 0.25, 0, 0, 0, 1, 0, 0, 0, 0, 1], dhat=dhat)
    assert outside["energy"] == 0.0 and outside["d1"] == 0.0 and outside["d2"] == 0.0


def test_friction_primitives():
    f1, _ = O.ipc_friction(0.1, 0.2, 1.0, 1.0)
    assert f1 == pytest.approx(0.75, abs=1e-14)
    assert O.ipc_friction(0.0, 0.2, 1.0, 1.0)[0] == 0.0
    assert O.ipc_friction(0.2, 0.2, 1.0, 1.0)[0] == 1.0
    assert O.ipc_friction(0.4, 0.2, 1.0, 1.0)[0] == 1.0
    # The lagged dissipation is mu lambda (f0(s) - f0(0)); its slip
    # derivative is mu lambda f1(s), and f0(0) = eps / 3.
    mu, lam, eps = 0.45, 3.0, 0.2
    for s in (0.05, 0.1, 0.19, 0.35):
        h = 1e-6
        fd = (O.ipc_friction(s + h, eps, mu, lam)[1] - O.ipc_friction(s - h, eps, mu, lam)[1]) / (2 * h)
        assert fd == pytest.approx(mu * lam * O.ipc_friction(s, eps, mu, lam)[0], rel=1e-8)
    assert O.ipc_friction(0.0, eps, mu, lam)[1] == 0.0
    assert O.ipc_friction(0.5, eps, mu, lam)[1] == pytest.approx(mu * lam * (0.5 - eps / 3), rel=1e-14)


@pytest.fixture(scope="module")
def trough():
    return scenes.trough()


def test_oracle_trough_frames_keep_clearance_and_armijo(trough):
    settings = dict(scenes.TROUGH_SETTINGS)
    ref = O.Oracle(trough, 64, settings=settings)
    dmin = settings["ipc_minimum_distance"]
    closest = np.inf
    for frame in range(60):
        ref.step_frames(1, scenes.TROUGH_ITERATIONS, settings)
        d = candidate_distances(ref.positions(), trough.contact)
        closest = min(closest, d.min())
        assert d.min() > dmin, frame
        diag = ref.local_diagnostics()
        accepted = diag[:, 3, 3] == 0  # local status accepted
        assert np.all(diag[accepted, 2, 0] <= diag[accepted, 2, 1])  # Armijo: dE <= c1 g.d
        assert np.all(diag[accepted, 2, 1] < 0)
    assert closest < 0.08  # the boxes reached the barrier band
    x = ref.positions()
    assert x[12:, 1].max() < 2.0  # every box fell into the trough


def test_oracle_trough_fp32_tracks_fp64(trough):
    settings = dict(scenes.TROUGH_SETTINGS)
    a, b = O.Oracle(trough, 32, settings=settings), O.Oracle(trough, 64, settings=settings)
    a.step_frames(10, scenes.TROUGH_ITERATIONS, settings)
    b.step_frames(10, scenes.TROUGH_ITERATIONS, settings)
    x0 = trough.positions.reshape(-1, 4)[:, :3]
    xa, xb = a.positions(), b.positions()
    assert np.linalg.norm(xa - xb) / np.linalg.norm(xb - x0) < 1e-4
    lag, alpha = b.contact_state()
    assert lag.shape == (31801, 4, 4) and 0.0 <= alpha <= 1.0


def contact_scene(**changes):
    scene = scenes.trough()
    for key, value in changes.items():
        setattr(scene.contact, key, value)
    return scene


def test_contact_validation():
    idx = scenes.trough().contact.packed_indices.copy()
    idx[1] = idx[0]
    with pytest.raises(jgs2.JGS2RangeError, match="VT candidate 0 must reference four distinct vertices"):
        jgs2.JGS2GpuSolver.create(contact_scene(packed_indices=idx), scenes.TROUGH_SETTINGS)
    idx = scenes.trough().contact.packed_indices.copy()
    idx[5] = 999
    with pytest.raises(jgs2.JGS2RangeError, match="references vertex 999"):
        jgs2.JGS2GpuSolver.create(contact_scene(packed_indices=idx), scenes.TROUGH_SETTINGS)
    scene = scenes.trough()
    pos = scene.positions.reshape(-1, 4).copy()
    top = int(np.argmax(pos[:12, 1]))  # a trough corner on its upper rim
    pos[12:20, :3] += pos[top, :3] + np.array([0.0, 1e-4, 0.0]) - pos[12, :3]  # first box touching it
    scene.positions = pos.reshape(-1)
    with pytest.raises(jgs2.JGS2RangeError, match="initially feasible source pose with distance"):
        jgs2.JGS2GpuSolver.create(scene, scenes.TROUGH_SETTINGS)


def test_oracle_contact_gradient_matches_restricted_energy(trough):
    # At a pose with active barriers, the assembled local gradient (barrier
    # and lagged friction included) is the derivative of the restricted
    # energy model the line search uses.
    settings = dict(scenes.TROUGH_SETTINGS)
    ref = O.Oracle(trough, 64, settings=settings)
    ref.step_frames(40, scenes.TROUGH_ITERATIONS, settings)
    x = ref.positions()
    d = candidate_distances(x, trough.contact)
    assert d.min() < settings["ipc_activation_distance"]
    g, _ = ref.local_systems(settings)
    ix = trough.contact.packed_indices.reshape(-1, 4)
    near = np.unique(ix[d < settings["ipc_activation_distance"]].reshape(-1))
    free = [int(v) for v in near if trough.vertex_info[4 * v + 2] == 0]
    assert free
    for v in free[:12]:
        fd = np.empty(3)
        for a in range(3):
            h = 1e-7
            e = np.zeros(3)
            e[a] = h
            ep = ref.restricted_model(v, e, settings)[2]
            e[a] = -h
            em = ref.restricted_model(v, e, settings)[2]
            fd[a] = (ep - em) / (2 * h)
        assert np.linalg.norm(g[v] - fd) <= 1e-5 * max(1.0, np.linalg.norm(g[v])), v
