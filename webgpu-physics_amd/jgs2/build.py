"""Builds libjgs2hip.so in-tree with hipcc for gfx950.

The library is the product: HIP kernels (csrc/jgs2_kernels.hip) plus the
C-ABI host code (csrc/jgs2_host.cpp, csrc/jgs2_scene*.cpp). It is written to
webgpu-physics_amd/lib/libjgs2hip.so so that it travels with the tree; the
Node binding (napi/) is linked against it as lib/jgs2hip.node.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PACKAGE_ROOT = Path(__file__).resolve().parent.parent
REPO_ROOT = PACKAGE_ROOT.parent
CSRC = PACKAGE_ROOT / "csrc"
LIB_DIR = PACKAGE_ROOT / "lib"
LIBRARY = LIB_DIR / "libjgs2hip.so"
BUILD_DIR = PACKAGE_ROOT / "_build"
NAPI_SOURCE = PACKAGE_ROOT / "napi" / "jgs2_napi.cpp"
NAPI_ADDON = LIB_DIR / "jgs2hip.node"
NODE_INCLUDE = Path(os.environ.get("NODE_INCLUDE", "/usr/include/node"))

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", f"{ROCM}/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
OFFLOAD_ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code; the compiler
# otherwise forms v_pk_*_f32 from float2-shaped code. Floating-point
# contraction is limited to single source expressions (-ffp-contract=on):
# every kernel variant (plain, cloth, contact) then rounds a shared
# expression the same way, so a scene's results do not depend on which
# other scenes share its batch.
DEVICE_FLAGS = [
    f"--offload-arch={OFFLOAD_ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=on",
]
COMMON_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", f"-I{REPO_ROOT / 'include'}", f"-I{CSRC}"]

SOURCES = [
    ("jgs2_kernels.hip", True),
    ("jgs2_host.cpp", False),
    ("jgs2_scene.cpp", False),
    ("jgs2_scene_contact.cpp", False),
    ("jgs2_precompute.cpp", False),
    ("jgs2_precompute_kernels.hip", True),
]


def _run(cmd: list[str]) -> None:
    print(" ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build(verbose: bool = True) -> Path:
    """Compile every source and link libjgs2hip.so; always rebuilds."""
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    objects = []
    for name, device in SOURCES:
        src = CSRC / name
        obj = BUILD_DIR / (name + ".o")
        if device:
            cmd = [HIPCC, *DEVICE_FLAGS, *COMMON_FLAGS, "-c", "-x", "hip", src, "-o", obj]
        else:
            # Host translation units: plain C++ against the HIP runtime API.
            cmd = [CXX, *COMMON_FLAGS, "-pthread", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-c", src, "-o", obj]
        _run(cmd)
        objects.append(obj)
    tmp = LIBRARY.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", *objects, "-o", tmp,
          f"-L{ROCM}/lib", "-lamdhip64", "-pthread", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIBRARY)
    build_napi()
    return LIBRARY


def build_napi() -> Path | None:
    """The Node binding (napi/jgs2_napi.cpp) as lib/jgs2hip.node, linked to
    libjgs2hip.so beside it. Skipped where Node's headers are absent."""
    if not (NODE_INCLUDE / "node_api.h").exists():
        print(f"N-API headers not found under {NODE_INCLUDE}; Node binding not built", flush=True)
        return None
    tmp = NAPI_ADDON.with_suffix(".node.tmp")
    _run([CXX, "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Wextra", f"-I{NODE_INCLUDE}",
          f"-I{REPO_ROOT / 'include'}", NAPI_SOURCE, "-o", tmp, f"-L{LIB_DIR}", "-ljgs2hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, NAPI_ADDON)
    return NAPI_ADDON


if __name__ == "__main__":
    build()
    sys.exit(0)
