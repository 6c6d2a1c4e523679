"""The C-ABI library: it loads, exports every function declared in
include/*.h, and rejects invalid input synchronously before any GPU work
(the reference's validation, src/simulation/gpu/layout.ts:223-416)."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import jgs2
from jgs2 import _native as N
from jgs2 import scenes

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_functions():
    names = set()
    for header in INCLUDE.glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(jgs2_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 30
    lib = C.CDLL(str(N.LIBRARY_PATH))
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, name


def test_default_settings_match_reference_defaults():
    s = N.Settings()
    N.lib.jgs2_default_settings(C.byref(s))
    d = jgs2.DEFAULT_STEP_SETTINGS
    assert s.schedule == 0 and s.iterations == d["iterations"]
    for key in ("timestep", "floor_stiffness", "velocity_damping", "regularization", "rotation_epsilon", "max_step",
                "residual_tolerance", "normalized_update_tolerance", "contact_tangential_damping", "contact_margin",
                "ipc_activation_distance", "ipc_minimum_distance", "ipc_step_safety"):
        assert getattr(s, key) == pytest.approx(d[key]), key
    assert list(s.gravity) == [0.0, -9.81, 0.0]


def small_scene():
    return scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5))


def expect_range(scene, match, settings=None):
    with pytest.raises(jgs2.JGS2RangeError, match=match):
        jgs2.JGS2GpuSolver.create(scene, settings)


def test_validation_rejects_noncanonical_adjacency():
    scene = small_scene()
    scene.adjacency = scene.adjacency.copy()
    scene.adjacency[[0, 1]] = scene.adjacency[[1, 0]]
    expect_range(scene, "must equal incident tetrahedron")


def test_validation_rejects_bad_counts_and_values():
    scene = small_scene()
    scene.tet_meta = scene.tet_meta.copy()
    scene.tet_meta[0] = -1.0
    expect_range(scene, "positive volume and mu")
    scene = small_scene()
    scene.positions = scene.positions.copy()
    scene.positions[5] = np.nan
    expect_range(scene, r"positions\[5\] must be finite")
    scene = small_scene()
    scene.tet_meta = scene.tet_meta.copy()
    scene.tet_meta[3] = 0.0  # mixed materials
    expect_range(scene, "cannot mix")


def test_validation_rejects_inverted_source_and_pinned_objectives():
    scene = small_scene()
    pos = scene.positions.reshape(-1, 4).copy()
    pos[:, 0] *= -1.0  # mirror: every tetrahedron inverted
    scene.positions = pos.reshape(-1)
    expect_range(scene, "initially feasible pose")
    scene = small_scene()
    forces = np.zeros(3 * scene.vertex_count, np.float32)
    forces[0] = 1.0  # vertex 0 is on the pinned x = 0 face
    scene.external_forces = forces
    expect_range(scene, "Pinned vertex 0 cannot have")


def test_settings_validation():
    scene = small_scene()
    expect_range(scene, "timestep must be positive", {"timestep": 0.0})
    expect_range(scene, "velocityDamping", {"velocity_damping": 2.0})
    expect_range(scene, "residualTolerance", {"residual_tolerance": 0.5})
    with pytest.raises(jgs2.JGS2RangeError, match="Unknown JGS2 schedule"):
        jgs2.JGS2GpuSolver.create(scene, {"schedule": "gauss"})


def test_create_without_gpu_fails_loudly():
    if jgs2.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(jgs2.JGS2DeviceError, match="no CPU fallback"):
        jgs2.JGS2GpuSolver.create(small_scene())
