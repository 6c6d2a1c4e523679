"""Test configuration: markers, import paths and library builds."""

import os
import subprocess
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "webgpu-physics_amd"))
sys.path.insert(0, str(REPO / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: longer CPU cases")


def _ensure_built():
    lib = REPO / "webgpu-physics_amd" / "lib" / "libjgs2hip.so"
    if not lib.exists():
        subprocess.run([sys.executable, str(REPO / "webgpu-physics_amd" / "jgs2" / "build.py")], check=True)
    oracle = REPO / "oracle" / "_build" / "libjgs2_oracle.so"
    if not oracle.exists():
        subprocess.run(["make", "-C", str(REPO / "oracle")], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def has_gpu():
    import jgs2

    return jgs2.device_count() > 0


@pytest.fixture(autouse=True)
def _gpu_guard(request):
    if request.node.get_closest_marker("gpu"):
        import jgs2

        if jgs2.device_count() < 1:
            pytest.skip("no HIP device")
    yield
