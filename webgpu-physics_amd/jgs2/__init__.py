"""jgs2: MI355X implementation of the JGS2 per-frame step (libjgs2hip).

Importing this package loads webgpu-physics_amd/lib/libjgs2hip.so and fails
if it is missing; there is no CPU fallback for the step.
"""

from ._native import (  # noqa: F401
    JGS2DeviceError,
    JGS2Error,
    JGS2InfeasibleError,
    JGS2MemoryError,
    JGS2RangeError,
    JGS2StateError,
    JGS2TypeError,
    LIBRARY_PATH,
)
from .solver import DEFAULT_STEP_SETTINGS, ClothData, GpuInput, JGS2GpuSolver  # noqa: F401
from .batch import BatchCapacity, JGS2Batch  # noqa: F401


def device_count() -> int:
    from ._native import lib

    return int(lib.jgs2_device_count())
