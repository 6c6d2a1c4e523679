"""Loader of the reference's versioned manifests (manifests/*.json).

Validates the schema identity, version and the fields a caller builds from
(src/reproducibility/manifests.ts, phase1-cubature-manifest.ts), and turns
fixtures into this package's inputs:

* `org.jgs2.phase1-cubature` v1: the 12-tetrahedron beam as cubature
  precompute input (`cubature_fixture`) and its expected packed selections;
* `org.jgs2.canonical-scenes` v1: fixtures by id, their expected metrics and
  sampled frames; `force_free_cuboid` builds the phase-0 force-free cuboid
  (rotated, with its linear and angular velocity) as a packed scene.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

CUBATURE_SCHEMA = "org.jgs2.phase1-cubature"
CANONICAL_SCHEMA = "org.jgs2.canonical-scenes"


class ManifestError(ValueError):
    pass


def _require(cond, label):
    if not cond:
        raise ManifestError(f"Manifest field {label} is missing or malformed.")


def validate(manifest: dict) -> dict:
    _require(isinstance(manifest, dict), "root")
    schema, version = manifest.get("schema"), manifest.get("schemaVersion")
    if schema == CUBATURE_SCHEMA:
        _require(version == 1, "schemaVersion")
        fx = manifest.get("fixture", {})
        g = fx.get("generator", {}).get("parameters", {})
        pos, tets = g.get("positions"), g.get("tetrahedra")
        _require(isinstance(pos, list) and len(pos) % 3 == 0 and len(pos) > 0, "fixture.generator.parameters.positions")
        V = len(pos) // 3
        _require(isinstance(tets, list) and len(tets) % 4 == 0 and all(0 <= t < V for t in tets),
                 "fixture.generator.parameters.tetrahedra")
        _require(isinstance(g.get("fixed"), list) and len(g["fixed"]) == V, "fixture.generator.parameters.fixed")
        _require(isinstance(g.get("materialIds"), list) and len(g["materialIds"]) == len(tets) // 4,
                 "fixture.generator.parameters.materialIds")
        m = fx.get("material", {})
        for k in ("density", "youngModulus", "poissonRatio"):
            _require(isinstance(m.get(k), (int, float)) and m[k] > (-1 if k == "poissonRatio" else 0),
                     f"fixture.material.{k}")
        _require(m.get("model") == "stable-neo-hookean", "fixture.material.model")
        _require(isinstance(fx.get("simulation", {}).get("timestep"), (int, float)), "fixture.simulation.timestep")
        cub = manifest.get("cubature", {})
        _require(isinstance(cub.get("maximumSamples"), int) and cub["maximumSamples"] > 0, "cubature.maximumSamples")
        _require(isinstance(cub.get("normalizedTrainingResidual"), (int, float)) and
                 0 <= cub["normalizedTrainingResidual"] <= 1, "cubature.normalizedTrainingResidual")
        free = [v for v in range(V) if not g["fixed"][v]]
        sel = manifest.get("expectedPackedSelections")
        _require(isinstance(sel, list) and [e.get("vertex") for e in sel] == free, "expectedPackedSelections")
        for i, e in enumerate(sel):
            _require(isinstance(e.get("tetrahedra"), list) and isinstance(e.get("weights"), list) and
                     len(e["tetrahedra"]) == len(e["weights"]) <= cub["maximumSamples"] and
                     all(w > 0 for w in e["weights"]), f"expectedPackedSelections[{i}]")
            _require(isinstance(e.get("packedTrainingColumnRank"), int) and
                     e["packedTrainingColumnRank"] > cub["maximumSamples"],
                     f"expectedPackedSelections[{i}].packedTrainingColumnRank")
        return manifest
    if schema == CANONICAL_SCHEMA:
        _require(version == 1, "schemaVersion")
        fixtures = manifest.get("fixtures")
        _require(isinstance(fixtures, list) and fixtures, "fixtures")
        ids = [f.get("id") for f in fixtures]
        _require(len(set(ids)) == len(ids) and all(isinstance(i, str) for i in ids), "fixtures[].id")
        for f in fixtures:
            for k in ("generator", "materials", "simulation", "solver", "sampledFrames"):
                _require(k in f, f"fixtures[{f['id']}].{k}")
        return manifest
    raise ManifestError(f"Unsupported manifest schema {schema!r}.")


def load(path) -> dict:
    return validate(json.loads(Path(path).read_text()))


def cubature_fixture(manifest: dict):
    """(positions, tetrahedra, fixed, material_ids, materials, timestep,
    samples) of the phase-1 cubature beam for jgs2.precompute."""
    validate(manifest)
    fx = manifest["fixture"]
    g = fx["generator"]["parameters"]
    m = fx["material"]
    return (np.array(g["positions"], np.float64).reshape(-1, 3), np.array(g["tetrahedra"], np.uint32).reshape(-1, 4),
            np.array(g["fixed"], np.uint8), np.array(g["materialIds"], np.uint32),
            [(float(m["density"]), float(m["youngModulus"]), float(m["poissonRatio"]), m["model"])],
            float(fx["simulation"]["timestep"]), int(manifest["cubature"]["maximumSamples"]))


def fixture(manifest: dict, fixture_id: str) -> dict:
    validate(manifest)
    for f in manifest["fixtures"]:
        if f["id"] == fixture_id:
            return f
    raise ManifestError(f"Unknown fixture {fixture_id!r}.")


def expected_metric(fx: dict, metric_id: str) -> float:
    for m in fx.get("expectedMetrics", []):
        if m["id"] == metric_id:
            return m["value"]
    raise ManifestError(f"Fixture {fx['id']} has no metric {metric_id!r}.")


def _rotate_xyz(p, r):  # rotateEulerXyz (src/simulation/cpu/mesh.ts:210-235): x, then y, then z
    cx, sx, cy, sy, cz, sz = np.cos(r[0]), np.sin(r[0]), np.cos(r[1]), np.sin(r[1]), np.cos(r[2]), np.sin(r[2])
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return p @ (rz @ ry @ rx).T


def force_free_cuboid(manifest: dict):
    """The phase-0 force-free cuboid as (scene, settings, iterations,
    frames): 2x2x2 cells rotated by [0.17, -0.23, 0.31] rad with linear
    velocity [0.37, -0.11, 0.23] and angular velocity [0.41, 0.29, -0.19]
    about the vertex mean, no gravity, floor, damping or pins, and the
    manifest's exact-all-elements cubature: every vertex keeps its exact
    whole-mesh basis and packs all 48 tetrahedra with weight 1
    (buildPhase0ForceFreeScene, src/scenes/canonical-phase0.ts:144-205)."""
    from . import scenes
    from .precompute import precompute_cubature

    f = fixture(manifest, "phase0.force-free-cuboid")
    g = f["generator"]["parameters"]
    mat = f["materials"][0]
    _require(f["solver"]["cubature"]["mode"] == "exact-all-elements", "solver.cubature.mode")
    x = scenes.build_cuboid(scenes._desc(g["cells"], g["origin"], g["size"], pin_axis=3, density=mat["density"],
                                         poisson=mat["poissonRatio"], young=(mat["youngModulus"],) * 2))
    rest = x.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    rec = precompute_cubature(rest, x.tet_indices.reshape(-1, 4), np.zeros(x.vertex_count), np.zeros(x.tet_count),
                              [(mat["density"], mat["youngModulus"], mat["poissonRatio"], "stable-neo-hookean")],
                              f["simulation"]["timestep"], x.tet_count, exact_all_elements=True)
    x.cubature_k = x.tet_count
    x.cubature_tet_ids = np.ascontiguousarray(rec.tet_ids.reshape(-1))
    x.cubature_weights = np.ascontiguousarray(rec.weights.reshape(-1))
    x.cubature_basis = np.ascontiguousarray(rec.basis.reshape(-1))
    x.extras["cubature"] = "exact-all-elements"
    p0 = _rotate_xyz(x.positions.reshape(-1, 4)[:, :3].astype(np.float64), np.array([0.17, -0.23, 0.31]))
    v_lin, omega = np.array([0.37, -0.11, 0.23]), np.array([0.41, 0.29, -0.19])
    vel = v_lin + np.cross(omega, p0 - p0.mean(0))
    x.positions = np.concatenate([p0, np.ones((len(p0), 1))], 1).astype(np.float32).reshape(-1)
    x.velocities = np.concatenate([vel, np.zeros((len(vel), 1))], 1).astype(np.float32).reshape(-1)
    sim = f["simulation"]
    settings = {"timestep": sim["timestep"], "gravity": tuple(sim["gravity"]), "floor_stiffness": 0.0,
                "parity_mode": sim["parityMode"], "velocity_damping": 1.0, "contact_tangential_damping": 0.0,
                "horizontal_body_correction": False, "schedule": "jacobi"}
    return x, settings, f["solver"]["iterations"], f["sampledFrames"][-1]["frame"]
