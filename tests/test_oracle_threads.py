"""The oracle's threaded passes (jgs2o_set_threads) give the serial
results bit for bit: every parallel loop writes only its own element's
records and the reductions stay serial. The multi-core CPU baseline
relies on this."""

import numpy as np
import pytest

from jgs2 import scenes
import jgs2_oracle as O


def run(scene, settings, frames, iterations, precision, threads):
    O.set_threads(threads)
    try:
        o = O.Oracle(scene, precision, settings=settings)
        o.step_frames(frames, iterations, settings)
        return o.positions(), o.velocities(), o.local_diagnostics(), o.history(), o.control()
    finally:
        O.set_threads(1)


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("case", ["cantilever", "trough", "cloth-ipc"])
def test_threaded_oracle_is_bit_identical_to_serial(case, precision):
    if case == "cantilever":
        scene, settings, it = scenes.cantilever_beam(), scenes.CANTILEVER_SETTINGS, 8
    elif case == "trough":
        scene, settings, it = scenes.trough(box_count=2), scenes.TROUGH_SETTINGS, 3
    else:
        scene, settings, it = scenes.cloth_drape(contact="superset"), scenes.CLOTH_IPC_SETTINGS, 5
    serial = run(scene, settings, 3, it, precision, 1)
    threaded = run(scene, settings, 3, it, precision, 4)
    for a, b in zip(serial, threaded):
        assert np.array_equal(a, b)
