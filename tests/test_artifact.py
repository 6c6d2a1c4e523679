"""Precomputation artifacts and manifest loading (jgs2.artifact, jgs2.manifests).

The artifact behaviors follow the reference's own tests
(src/reproducibility/precomputation-artifact.test.ts:48-134): deterministic
fingerprints independent of key order, sensitivity to content and typed-array
kind, rejection of non-finite and cyclic values, field-by-field header
validation, and one error naming every stale component. Number formatting
follows ECMAScript Number::toString, checked against the spec's cases. The
reference holds no fingerprint values, so the hashes themselves are not
compared with the reference's (parity unpinned for the exact digests).
"""

import copy
import json
from pathlib import Path

import numpy as np
import pytest

from jgs2 import artifact as A
from jgs2 import manifests as M
from jgs2.precompute import precompute_cubature

GOLDEN = Path(__file__).parent / "golden"


def fp(c):
    return "sha256:" + c * 64


def fingerprints():
    return {"topology": fp("1"), "materials": fp("2"), "timestep": fp("3"), "basis": fp("4"),
            "cubature": fp("5"), "solverSchedule": fp("6")}


def header():
    return {"schema": A.SCHEMA, "schemaVersion": A.SCHEMA_VERSION,
            "createdBy": {"name": "fixture-preprocessor", "version": "0.1.0"}, "fingerprints": fingerprints(),
            "payload": {"format": A.PAYLOAD_FORMAT, "formatVersion": A.PAYLOAD_FORMAT_VERSION, "byteLength": 4096,
                        "fingerprint": fp("a")}}


@pytest.mark.parametrize("value,text", [
    (0.0, "0"), (-0.0, "0"), (1.0, "1"), (123.0, "123"), (0.1, "0.1"), (1e21, "1e+21"), (1e20, "100000000000000000000"),
    (1e-7, "1e-7"), (1e-6, "0.000001"), (1.5e-5, "0.000015"), (-2.5e-10, "-2.5e-10"), (5e-324, "5e-324"),
    (1.7976931348623157e308, "1.7976931348623157e+308"), (1 / 3, "0.3333333333333333"), (123.456, "123.456"),
    (float(np.float32(0.1)), "0.10000000149011612")])
def test_numbers_format_like_ecmascript(value, text):
    assert A.js_number(value) == text


def test_fingerprints_are_deterministic_and_order_free():
    left = A.fingerprint({"timestep": 1 / 60, "topology": np.array([0, 1, 2, 3], np.uint32),
                          "nested": {"schedule": "jacobi", "iterations": 12}})
    reordered = A.fingerprint({"nested": {"iterations": 12, "schedule": "jacobi"},
                               "topology": np.array([0, 1, 2, 3], np.uint32), "timestep": 1 / 60})
    changed = A.fingerprint({"timestep": 1 / 60, "topology": np.array([0, 1, 3, 2], np.uint32),
                             "nested": {"schedule": "jacobi", "iterations": 12}})
    assert re_match(left)
    assert reordered == left and changed != left


def re_match(s):
    return s.startswith("sha256:") and len(s) == 71 and all(c in "0123456789abcdef" for c in s[7:])


def test_typed_kinds_differ_and_ambiguous_values_are_rejected():
    assert A.fingerprint(np.array([1, 2, 3], np.uint32)) != A.fingerprint(np.array([1, 2, 3], np.float32))
    assert A.canonical_string(np.array([1, 2, 3], np.uint32)) == "typed:Uint32Array:3:[number:1,number:2,number:3]"
    assert A.canonical_string({"b": [True, None], "a": "x"}) == 'object:2:{"a"=string:"x","b"=array:2:[boolean:true,null]}'
    assert A.canonical_string(b"\x01\xff") == "bytes:ArrayBuffer:2:01ff"
    with pytest.raises(A.ArtifactError, match="non-finite"):
        A.fingerprint(float("nan"))
    cyclic = {}
    cyclic["self"] = cyclic
    with pytest.raises(A.ArtifactError, match="cyclic"):
        A.fingerprint(cyclic)


def test_header_validation():
    assert A.validate_header(header()) == header()
    bad = dict(header(), schemaVersion=2)
    with pytest.raises(A.ArtifactError, match="schema version 2"):
        A.validate_header(bad)
    bad = header()
    bad["payload"]["byteLength"] = -1
    with pytest.raises(A.ArtifactError, match="byteLength"):
        A.validate_header(bad)
    bad = header()
    bad["fingerprints"]["topology"] = "sha256:ABC"
    with pytest.raises(A.ArtifactError, match="topology"):
        A.validate_header(bad)
    assert A.FINGERPRINT_COMPONENTS == ("topology", "materials", "timestep", "basis", "cubature", "solverSchedule")


def test_compatibility_names_every_stale_component():
    assert A.assert_compatible(header(), fingerprints()) == header()
    stale = dict(fingerprints(), topology=fp("b"), timestep=fp("c"), solverSchedule=fp("d"))
    with pytest.raises(A.ArtifactError, match="topology, timestep, solverSchedule"):
        A.assert_compatible(header(), stale)


def test_cubature_records_round_trip_and_stale_inputs_are_rejected(tmp_path):
    manifest = M.load(GOLDEN / "phase1-cubature.v1.json")
    pos, tets, fixed, mid, mats, h, K = M.cubature_fixture(manifest)
    rec = precompute_cubature(pos, tets, fixed, mid, mats, h, K)
    fps = A.precompute_fingerprints(pos, tets, fixed, mid, mats, h, K, rings=0)
    A.save_cubature_artifact(tmp_path / "beam", rec, fps)
    t, w, b, r = A.load_cubature_artifact(tmp_path / "beam", fps)
    assert np.array_equal(t, rec.tet_ids) and np.array_equal(w, rec.weights) and np.array_equal(b, rec.basis)
    assert np.array_equal(r, rec.residuals, equal_nan=True)
    # every selection still the reference's
    for e in manifest["expectedPackedSelections"]:
        n = len(e["tetrahedra"])
        assert list(t[e["vertex"], :n]) == e["tetrahedra"]
        assert np.array_equal(w[e["vertex"], :n], np.array(e["weights"], np.float32))
    moved = pos.copy()
    moved[5, 1] += 1e-9
    stale = A.precompute_fingerprints(moved, tets, fixed, mid, [(1000.0, 5001.0, 0.3, "stable-neo-hookean")], h, K,
                                      rings=2)
    with pytest.raises(A.ArtifactError, match="topology, materials, basis"):
        A.load_cubature_artifact(tmp_path / "beam", stale)
    data = bytearray((tmp_path / "beam.bin").read_bytes())
    data[40] ^= 1
    (tmp_path / "beam.bin").write_bytes(bytes(data))
    with pytest.raises(A.ArtifactError, match="payload fingerprint"):
        A.load_cubature_artifact(tmp_path / "beam", fps)


def test_manifests_load_and_reject_drift():
    canonical = M.load(GOLDEN / "canonical-scenes.v1.json")
    assert M.expected_metric(M.fixture(canonical, "baseline.trough"), "minimum-ipc-distance") > 0
    scene, settings, iterations, frames = M.force_free_cuboid(canonical)
    assert scene.vertex_count == 27 and iterations == 12 and frames == 1200
    v = scene.velocities.reshape(-1, 4)[:, :3]
    assert np.abs(v.mean(0) - [0.37, -0.11, 0.23]).max() < 1e-6
    cub = json.loads((GOLDEN / "phase1-cubature.v1.json").read_text())
    M.validate(cub)
    for mutate, field in [
        (lambda m: m.update(schemaVersion=2), "schemaVersion"),
        (lambda m: m["fixture"]["generator"]["parameters"]["positions"].pop(), "positions"),
        (lambda m: m["expectedPackedSelections"].pop(), "expectedPackedSelections"),
        (lambda m: m["expectedPackedSelections"][0].update(packedTrainingColumnRank=6), "packedTrainingColumnRank"),
        (lambda m: m["cubature"].update(normalizedTrainingResidual=-0.1), "normalizedTrainingResidual"),
    ]:
        bad = copy.deepcopy(cub)
        mutate(bad)
        with pytest.raises(M.ManifestError, match=field):
            M.validate(bad)
    with pytest.raises(M.ManifestError, match="Unsupported manifest schema"):
        M.validate({"schema": "org.jgs2.unknown"})
