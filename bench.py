"""Benchmark of the JGS2 step on one MI355X.

Workload (default): the single large mesh -- a 1 m x 1 m x 2 m box on a
64x64x128 hex grid (545,025 vertices, 3,145,728 tetrahedra), seeded vertex
jitter, alternating soft/stiff layers, stable Neo-Hookean, K = 6 cubature
records per vertex, 8 Jacobi sweeps per frame. One "step" is one frame.

Prints ONE JSON line. `value` is free-vertex updates per second (one local
solve of one free vertex in one sweep), all inputs resident in HBM.
`roofline` reports the local-solve kernel: reference-ABI algorithmic bytes
per launch over its mean launch time measured with HIP events on the
solver's stream. `cpu_baseline` times the oracle (a single-threaded CPU
restatement of the same step) on a bounded sample. `accuracy` is the
trajectory error of the GPU against the FP64 oracle on a reduced grid of the
same generator.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "webgpu-physics_amd"))

METRIC = "JGS2 vertex-updates/s batched vs plain on 1 MI355X at tol; p50/p99 frame ms"
HBM_PEAK_GBS = 8000.0


def nearest_rank(values, q):
    """Nearest-rank percentile (src/performance/metrics.ts:42-46)."""
    s = sorted(values)
    if not s:
        return 0.0
    rank = max(1, int(np.ceil(q / 100.0 * len(s))))
    return float(s[min(rank, len(s)) - 1])


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--schedule", default="jacobi", choices=["jacobi", "graph-colored-gauss-seidel"])
    ap.add_argument("--cells", default="64,64,128")
    ap.add_argument("--iterations", type=int, default=8)
    ap.add_argument("--latency-frames", type=int, default=10)
    ap.add_argument("--other-schedule-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the step is single-GPU here).",
              file=sys.stderr)
        return 2

    import jgs2
    from jgs2 import scenes

    if jgs2.device_count() < 1:
        print("No HIP device visible; bench.py needs an MI355X.", file=sys.stderr)
        return 1

    cells = tuple(int(c) for c in args.cells.split(","))
    t0 = time.time()
    scene = scenes.layered_box(cells=cells)
    settings = dict(scenes.LAYERED_BOX_SETTINGS, schedule=args.schedule)
    log(f"scene {cells}: V={scene.vertex_count} T={scene.tet_count} built in {time.time() - t0:.1f}s")
    t0 = time.time()
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    log(f"solver created in {time.time() - t0:.1f}s; colors={solver.schedule_color_count}")
    it = args.iterations
    free = solver.free_vertex_count

    for w in range(args.warmup):
        solver.step_exact_iterations(it, settings)
        solver.await_idle()
        log(f"warmup frame {w + 1}/{args.warmup} done")

    solver.await_idle()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.step_exact_iterations(it, settings)
    solver.await_idle()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} frames in {elapsed:.3f}s")
    updates = args.steps * it * free
    value = updates / elapsed

    # Per-frame latency as a caller sees it (enqueue + completion).
    lat = []
    for _ in range(args.latency_frames):
        a = time.perf_counter()
        solver.step_exact_iterations(it, settings)
        solver.await_idle()
        lat.append((time.perf_counter() - a) * 1e3)
    log(f"latency frames done: p50 {nearest_rank(lat, 50):.2f} ms")

    # Local-solve kernel time per launch, HIP events on the solver's stream.
    solver.profile_begin()
    solver.step_exact_iterations(it, settings)
    solve_ms, launches, window_ms = solver.profile_end()
    bytes_per_sweep = scenes.algorithmic_bytes_per_sweep(scene.vertex_count, scene.tet_count, scene.cubature_k)
    achieved = bytes_per_sweep / (solve_ms * 1e-3) / 1e9 if solve_ms > 0 else 0.0
    x = solver.read_positions().reshape(-1, 4)[:, :3]
    finite = bool(np.isfinite(x).all())
    hist = solver.read_globalization_history().reshape(64, 8, 4)
    dynamics = {"reverted_iterations_last_frame": int(hist[:it, 1, 3].sum()),
                "max_backtracks_last_frame": int(hist[:it, 7, 0].max()),
                "relative_residual_last_iteration": float(hist[it - 1, 4, 3])}
    solver.destroy()

    other = None
    if args.other_schedule_steps > 0:
        other_name = "graph-colored-gauss-seidel" if args.schedule == "jacobi" else "jacobi"
        other = measure_schedule(scene, dict(settings, schedule=other_name), it, args.other_schedule_steps)

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"layered-box-{cells[0]}x{cells[1]}x{cells[2]}",
            "vertices": scene.vertex_count,
            "free_vertices": free,
            "tetrahedra": scene.tet_count,
            "cubature_k": scene.cubature_k,
            "schedule": args.schedule,
            "iterations_per_frame": it,
            "mode": "plain (one scene per launch)",
        },
        "frame_ms": {"p50": nearest_rank(lat, 50), "p99": nearest_rank(lat, 99), "frames": len(lat)},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_solve (local solve)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": bytes_per_sweep,
            "ms_per_launch": solve_ms,
            "launches_profiled": launches,
        },
        "positions_finite": finite,
        "dynamics": dynamics,
    }
    if other is not None:
        result["other_schedule"] = other

    if not args.no_accuracy:
        result["accuracy"] = accuracy_check(settings, it)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(settings, it)
    print(json.dumps(result), flush=True)
    return 0


def measure_schedule(scene, settings, iterations, steps):
    """Throughput of the second schedule on the same scene (fresh solver)."""
    import jgs2

    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    solver.step_exact_iterations(iterations, settings)
    solver.await_idle()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.step_exact_iterations(iterations, settings)
    solver.await_idle()
    dt = time.perf_counter() - t0
    hist = solver.read_globalization_history().reshape(64, 8, 4)
    out = {"schedule": settings["schedule"], "colors": solver.schedule_color_count, "steps": steps,
           "ms_per_step": dt / steps * 1e3,
           "vertex_updates_per_s": steps * iterations * solver.free_vertex_count / dt,
           "reverted_iterations_last_frame": int(hist[:iterations, 1, 3].sum())}
    solver.destroy()
    log(f"{settings['schedule']}: {out['ms_per_step']:.2f} ms/frame")
    return out


def _small_scene():
    from jgs2 import scenes

    return scenes.layered_box(cells=(12, 12, 24))


def accuracy_check(settings, iterations, frames=2):
    """GPU vs FP64 oracle on a 12x12x24-cell grid of the same generator."""
    import jgs2

    sys.path.insert(0, str(REPO / "oracle"))
    import jgs2_oracle as O

    scene = _small_scene()
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    solver.step_frames_exact_iterations(frames, iterations, settings)
    x_gpu = solver.read_positions().reshape(-1, 4)[:, :3].astype(np.float64)
    solver.destroy()
    ref = O.Oracle(scene, 64)
    ref.step_frames(frames, iterations, settings)
    x_ref = ref.positions()
    x0 = scene.positions.reshape(-1, 4)[:, :3].astype(np.float64)
    err = float(np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref))
    derr = float(np.linalg.norm(x_gpu - x_ref) / max(np.linalg.norm(x_ref - x0), 1e-30))
    log(f"accuracy: rel L2 {err:.3e}, displacement-relative {derr:.3e}")
    return {"reference": "fp64 oracle", "grid": "12x12x24", "frames": frames, "rel_l2": err,
            "rel_l2_of_displacement": derr, "tolerance": 1e-3, "within_tol": err <= 1e-3}


def cpu_baseline(settings, iterations, frames=2):
    """The oracle's FP32 step (single thread) on the 12x12x24-cell sample."""
    sys.path.insert(0, str(REPO / "oracle"))
    import jgs2_oracle as O

    scene = _small_scene()
    ref = O.Oracle(scene, 32)
    free = int((scene.vertex_info.reshape(-1, 4)[:, 2] == 0).sum())
    t0 = time.perf_counter()
    ref.step_frames(frames, iterations, settings)
    dt = time.perf_counter() - t0
    log(f"cpu baseline: {frames} frames in {dt:.2f}s")
    return {"value": frames * iterations * free / dt, "unit": "vertex-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP32 step, layered box 12x12x24 cells ({scene.vertex_count} vertices), "
                      f"{frames} frames x {iterations} Jacobi sweeps, {dt:.2f}s"}


if __name__ == "__main__":
    sys.exit(main())
