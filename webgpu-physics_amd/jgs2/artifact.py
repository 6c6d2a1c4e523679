"""Versioned precomputation artifacts (src/reproducibility/precomputation-artifact.ts).

A header of metadata stored next to an offline precomputation payload, with
SHA-256 fingerprints of six input components (topology, materials, timestep,
basis, cubature, solver schedule) and of the payload. Fingerprints detect
stale or incompatible artifacts; they are not a trust mechanism.

Fingerprints hash a canonical string of JSON-like values, typed arrays and
byte buffers, built by the reference's rules (precomputation-artifact.ts:
163-258): sorted object keys, typed-array kinds included, numbers in the
ECMAScript Number-to-String form, non-finite numbers and cycles rejected.
numpy arrays stand for typed arrays (uint8 -> Uint8Array, float32 ->
Float32Array, ...), `bytes` for an ArrayBuffer.

`save_cubature_artifact` / `load_cubature_artifact` store the records of
`jgs2_precompute_cubature` with their input fingerprints, so a caller can
reuse them across runs and have a changed scene, material or setting
rejected by name.
"""

from __future__ import annotations

import hashlib
import json
import re
import struct
from decimal import Decimal
from pathlib import Path

import numpy as np

SCHEMA = "org.jgs2.precomputation-artifact"
SCHEMA_VERSION = 1
PAYLOAD_FORMAT = "jgs2-precomputation-binary-le"
PAYLOAD_FORMAT_VERSION = 1
FINGERPRINT_COMPONENTS = ("topology", "materials", "timestep", "basis", "cubature", "solverSchedule")
_SHA256 = re.compile(r"^sha256:[0-9a-f]{64}$")

_TYPED = {
    np.dtype(np.int8): "Int8Array", np.dtype(np.uint8): "Uint8Array",
    np.dtype(np.int16): "Int16Array", np.dtype(np.uint16): "Uint16Array",
    np.dtype(np.int32): "Int32Array", np.dtype(np.uint32): "Uint32Array",
    np.dtype(np.float32): "Float32Array", np.dtype(np.float64): "Float64Array",
}


class ArtifactError(ValueError):
    pass


def js_number(value: float) -> str:
    """Number::toString(10) (ECMAScript 2020, 7.1.12.1) of a finite double:
    the shortest round-trip digits, positional for exponents in (-7, 21]."""
    x = float(value)
    if not np.isfinite(x):
        raise ArtifactError("Cannot fingerprint a non-finite number.")
    if x == 0.0:
        return "0"
    sign = "-" if x < 0 else ""
    d = Decimal(repr(abs(x))).normalize()
    _, digits, exponent = d.as_tuple()
    s = "".join(str(c) for c in digits)
    k = len(s)
    n = exponent + k  # value = s * 10^(n - k)
    if k <= n <= 21:
        body = s + "0" * (n - k)
    elif 0 < n <= 21:
        body = s[:n] + "." + s[n:]
    elif -6 < n <= 0:
        body = "0." + "0" * (-n) + s
    else:
        e = n - 1
        mant = s if k == 1 else s[0] + "." + s[1:]
        body = f"{mant}e{'+' if e >= 0 else '-'}{abs(e)}"
    return sign + body


def canonical_string(value, _active=None) -> str:
    """The reference's canonical form of a JSON-like value."""
    active = set() if _active is None else _active
    if value is None:
        return "null"
    if isinstance(value, (bool, np.bool_)):
        return "boolean:true" if value else "boolean:false"
    if isinstance(value, (int, float, np.integer, np.floating)):
        return "number:" + js_number(value)
    if isinstance(value, str):
        return "string:" + json.dumps(value, ensure_ascii=False)
    if isinstance(value, (bytes, bytearray, memoryview)):
        b = bytes(value)
        return f"bytes:ArrayBuffer:{len(b)}:{b.hex()}"
    if id(value) in active:
        raise ArtifactError("Cannot fingerprint a cyclic value.")
    active.add(id(value))
    try:
        if isinstance(value, np.ndarray):
            kind = _TYPED.get(value.dtype)
            if kind is None:
                raise ArtifactError(f"Cannot fingerprint an array of {value.dtype}.")
            flat = value.reshape(-1)
            entries = ",".join("number:" + js_number(float(v)) for v in flat)
            return f"typed:{kind}:{flat.size}:[{entries}]"
        if isinstance(value, (list, tuple)):
            return f"array:{len(value)}:[" + ",".join(canonical_string(v, active) for v in value) + "]"
        if isinstance(value, dict):
            if not all(isinstance(k, str) for k in value):
                raise ArtifactError("Object keys must be strings.")
            keys = sorted(value, key=lambda k: k.encode("utf-16-be"))  # UTF-16 code-unit order
            return f"object:{len(keys)}:{{" + ",".join(
                f"{json.dumps(k, ensure_ascii=False)}={canonical_string(value[k], active)}" for k in keys) + "}"
        raise ArtifactError(f"Cannot fingerprint non-plain object {type(value).__name__}.")
    finally:
        active.discard(id(value))


def fingerprint(value) -> str:
    """`sha256:<hex>` of the UTF-8 canonical string (fingerprintCanonicalValue)."""
    return "sha256:" + hashlib.sha256(canonical_string(value).encode("utf-8")).hexdigest()


def _record(value, label):
    if not isinstance(value, dict):
        raise ArtifactError(f"{label} must be an object.")
    return value


def _fingerprint_field(value, label):
    if not isinstance(value, str) or not _SHA256.match(value):
        raise ArtifactError(f"{label} must be a lowercase SHA-256 fingerprint.")
    return value


def validate_fingerprints(value) -> dict:
    rec = _record(value, "fingerprints")
    return {c: _fingerprint_field(rec.get(c), f"fingerprints.{c}") for c in FINGERPRINT_COMPONENTS}


def validate_header(value) -> dict:
    """validatePrecomputationArtifactHeader (precomputation-artifact.ts:75-131)."""
    root = _record(value, "precomputation artifact")
    if root.get("schema") != SCHEMA:
        raise ArtifactError(f"Unsupported precomputation schema {root.get('schema')!r}; expected {SCHEMA}.")
    if root.get("schemaVersion") != SCHEMA_VERSION:
        raise ArtifactError(f"Unsupported precomputation schema version {root.get('schemaVersion')}; "
                            f"expected {SCHEMA_VERSION}.")
    created = _record(root.get("createdBy"), "createdBy")
    for key in ("name", "version"):
        if not isinstance(created.get(key), str) or not created[key].strip():
            raise ArtifactError(f"createdBy.{key} must be a nonempty string.")
    fps = validate_fingerprints(root.get("fingerprints"))
    payload = _record(root.get("payload"), "payload")
    if payload.get("format") != PAYLOAD_FORMAT:
        raise ArtifactError(f"Unsupported payload format {payload.get('format')!r}; expected {PAYLOAD_FORMAT}.")
    if payload.get("formatVersion") != PAYLOAD_FORMAT_VERSION:
        raise ArtifactError(f"Unsupported payload format version {payload.get('formatVersion')}; "
                            f"expected {PAYLOAD_FORMAT_VERSION}.")
    n = payload.get("byteLength")
    if isinstance(n, bool) or not isinstance(n, int) or n < 0 or n > 2 ** 53 - 1:
        raise ArtifactError("payload.byteLength must be a nonnegative safe integer.")
    return {"schema": SCHEMA, "schemaVersion": SCHEMA_VERSION,
            "createdBy": {"name": created["name"], "version": created["version"]}, "fingerprints": fps,
            "payload": {"format": PAYLOAD_FORMAT, "formatVersion": PAYLOAD_FORMAT_VERSION, "byteLength": n,
                        "fingerprint": _fingerprint_field(payload.get("fingerprint"), "payload.fingerprint")}}


def assert_compatible(artifact, expected) -> dict:
    """assertPrecomputationArtifactCompatible (:136-152): every stale component
    named in one error."""
    header = validate_header(artifact)
    want = validate_fingerprints(expected)
    stale = [c for c in FINGERPRINT_COMPONENTS if header["fingerprints"][c] != want[c]]
    if stale:
        raise ArtifactError(f"Precomputation artifact is incompatible with: {', '.join(stale)}.")
    return header


# ---------------------------------------------------------------- cubature records
MODEL_NAMES = {0: "corotated-linear", 1: "stable-neo-hookean"}


def precompute_fingerprints(positions, tetrahedra, fixed, material_ids, materials, timestep, samples, rings,
                            residual_bound=0.0) -> dict:
    """Input fingerprints of one jgs2_precompute_cubature call."""
    mats = []
    for q in materials:
        model = q[3] if isinstance(q[3], str) else MODEL_NAMES[int(q[3])]
        mats.append({"density": float(q[0]), "youngModulus": float(q[1]), "poissonRatio": float(q[2]),
                     "model": model})
    return {
        "topology": fingerprint({
            "positions": np.ascontiguousarray(positions, np.float64).reshape(-1),
            "tetrahedra": np.ascontiguousarray(tetrahedra, np.uint32).reshape(-1),
            "fixed": np.ascontiguousarray(np.asarray(fixed).reshape(-1) != 0, np.uint8),
            "materialIds": np.ascontiguousarray(material_ids, np.uint32).reshape(-1)}),
        "materials": fingerprint(mats),
        "timestep": fingerprint(float(timestep)),
        "basis": fingerprint({"coRotation": "R_v * Ubar_vi * transpose(R_i)", "rings": int(rings),
                              "vertexFrame": "polar(rest-volume-weighted-average-F)"}),
        "cubature": fingerprint({"samples": int(samples), "residualBound": float(residual_bound or 0.01),
                                 "corpusVersion": "phase1-nonlinear-cubature-v1",
                                 "selection": "deterministic-greedy-nnls-with-small-fixture-subset-audit"}),
        "solverSchedule": fingerprint({"modeCount": 8, "validationOnlyModeCount": 4, "inverseIterations": 24,
                                       "refinementPasses": 12}),
    }


def _payload(records) -> bytes:
    V, K = records.tet_ids.shape
    head = struct.pack("<4sII", b"JGC1", V, K)
    return (head + np.ascontiguousarray(records.tet_ids, "<u4").tobytes() +
            np.ascontiguousarray(records.weights, "<f4").tobytes() +
            np.ascontiguousarray(records.basis, "<f4").tobytes() +
            np.ascontiguousarray(records.residuals, "<f8").tobytes())


def save_cubature_artifact(path, records, fingerprints: dict, created_by=("libjgs2hip", "1")) -> dict:
    """Writes `<path>.json` (header) and `<path>.bin` (payload)."""
    payload = _payload(records)
    header = {"schema": SCHEMA, "schemaVersion": SCHEMA_VERSION,
              "createdBy": {"name": created_by[0], "version": created_by[1]},
              "fingerprints": validate_fingerprints(fingerprints),
              "payload": {"format": PAYLOAD_FORMAT, "formatVersion": PAYLOAD_FORMAT_VERSION,
                          "byteLength": len(payload), "fingerprint": fingerprint(payload)}}
    base = Path(path)
    base.with_suffix(".bin").write_bytes(payload)
    base.with_suffix(".json").write_text(json.dumps(header, indent=1, sort_keys=True))
    return header


def load_cubature_artifact(path, expected_fingerprints: dict):
    """Reads and verifies an artifact: header schema, every input fingerprint
    against `expected_fingerprints`, payload length and fingerprint. Returns
    (tet_ids [V, K], weights [V, K], basis [V, K, 36], residuals [V])."""
    base = Path(path)
    header = assert_compatible(json.loads(base.with_suffix(".json").read_text()), expected_fingerprints)
    payload = base.with_suffix(".bin").read_bytes()
    if len(payload) != header["payload"]["byteLength"]:
        raise ArtifactError("Precomputation payload length does not match its header.")
    if fingerprint(payload) != header["payload"]["fingerprint"]:
        raise ArtifactError("Precomputation payload fingerprint does not match its header.")
    magic, V, K = struct.unpack_from("<4sII", payload, 0)
    if magic != b"JGC1" or len(payload) != 12 + V * K * (4 + 4 + 144) + 8 * V:
        raise ArtifactError("Malformed cubature payload.")
    o = 12
    tet = np.frombuffer(payload, "<u4", V * K, o).reshape(V, K).copy()
    o += 4 * V * K
    w = np.frombuffer(payload, "<f4", V * K, o).reshape(V, K).copy()
    o += 4 * V * K
    b = np.frombuffer(payload, "<f4", 36 * V * K, o).reshape(V, K, 36).copy()
    o += 144 * V * K
    res = np.frombuffer(payload, "<f8", V, o).copy()
    return tet, w, b, res
