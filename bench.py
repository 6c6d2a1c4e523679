"""Benchmark of the JGS2 step on one MI355X.

Headline workload (`value`): the batched uneven-scene stream -- 512 seeded
independent scenes (70 % stable Neo-Hookean beams and boxes of 200 to
200,000 tetrahedra, log-uniform, E log-uniform in [1e4, 1e7] Pa; 30 % StVK
cloth sheets of 32x32 to 256x256 cells with dihedral bending), K = 6
cubature records, 8 exact Jacobi sweeps per frame, all resident in one
batch arena and stepped one launch per pass, frames replayed from a
hipGraph. One "step" is one frame of every scene. `value` is free-vertex
updates per second (one local solve of one free vertex in one sweep).
The same scenes stepped one solver per scene (the plain path) give
`plain`; `batched_over_plain` is the ratio.

Further legs, each in its own object:
* `serving`: Poisson arrivals of scene-step requests (contact troughs, 1 to
  120 frames each) served by continuous batching under an HBM budget, and
  by one solver per request; throughput, p50/p99 frame latency, request
  latency and queueing.
* `single_large_mesh`: the 64x64x128 layered box (545,025 vertices,
  3,145,728 tetrahedra) under Jacobi and graph-colored Gauss-Seidel.

`roofline` reports the local-solve kernel of the headline batch:
reference-ABI algorithmic bytes per launch over its mean launch time,
measured with HIP events on the batch's stream. `accuracy` compares GPU
trajectories of a reduced mix with the FP64 oracle. `cpu_baseline` times
the oracle's FP32 step on a bounded sample of the same generator.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def scene_bytes_per_sweep(scene, candidates=0) -> int:
    """Reference-ABI compulsory bytes of one sweep over one scene, counting
    the records the scene holds: V * 240 (vertex state and records) +
    T * 96 (tetrahedron record and adjacency ids) + 152 per cubature record
    in use (tet id set, weight > 0; cloth vertices hold none) + cloth records
    (triangle 4 vec4 + 3 incidence ids, hinge 2 vec4 + 4 incidence ids) +
    IPC candidates (144-byte record + 4 incidence ids)."""
    V, T = scene.vertex_count, scene.tet_count
    used = int(((scene.cubature_tet_ids != 0xFFFFFFFF) & (scene.cubature_weights > 0)).sum()) if scene.cubature_k else 0
    b = V * 240 + T * 96 + used * 152
    if scene.cloth is not None:
        b += scene.cloth.triangle_count * (64 + 12) + scene.cloth.hinge_count * (32 + 16)
    return b + int(candidates) * (144 + 16)


def build_scenes(specs, contact="broad_phase"):
    """Builds the scenes of `specs` on the host threads (the scene setup and
    the broad-phase capacity scans run in the library, outside the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    from jgs2 import scenes

    with ThreadPoolExecutor(_host_threads()) as pool:
        return list(pool.map(lambda s: scenes.build_mixed_scene(s, contact), specs))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)  # at least 1: the autotuner runs after the first
    ap.add_argument("--scenes", type=int, default=512)
    ap.add_argument("--seed", type=int, default=20251203)
    ap.add_argument("--iterations", type=int, default=8)
    ap.add_argument("--latency-frames", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=12.0)
    ap.add_argument("--contact-frames", type=int, default=10)
    ap.add_argument("--large-steps", type=int, default=5)
    ap.add_argument("--large-cells", default="64,64,128")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a cold 1 GiB device buffer (the counter calibration kernel) and exit")
    ap.add_argument("--profile-contact", type=int, default=0, metavar="FRAMES",
                    help="profiler driver: only the contact-box workload, FRAMES frames after 3 warm ones")
    ap.add_argument("--profile-large", default=None, metavar="LANES",
                    help="only step the layered box two frames per listed lane count (e.g. 1,4,8,16) for a profiler, "
                         "and exit")
    ap.add_argument("--precompute-cells", default="16,4,4",
                    help="cantilever cells of the cubature precompute leg ('' skips it)")
    ap.add_argument("--profile-precompute", default=None, metavar="CELLS:RINGS[:FRAMES]",
                    help="only the cubature precompute of the layered box, e.g. 32,32,64:1; with FRAMES, "
                         "then Jacobi frames on the trained records against the stand-in records")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving leg (for --serve-seconds long enough to complete many requests), "
                         "printed as one JSON object, and exit")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    args = ap.parse_args(argv)

    args.warmup = max(1, args.warmup)
    if args.copy_calibration:
        return copy_calibration()
    if args.profile_contact:
        return profile_contact(args.profile_contact)
    if args.profile_large:
        return profile_large([int(x) for x in args.profile_large.split(",")], args.large_cells, args.iterations)
    if args.serve_only:
        print(json.dumps({"serving": measure_serving(args.serve_seconds, args.seed)}), flush=True)
        return 0
    if args.profile_precompute:
        parts = args.profile_precompute.split(":")
        frames = int(parts[2]) if len(parts) > 2 else 0
        print(json.dumps(precompute_layered_box(tuple(int(c) for c in parts[0].split(",")), int(parts[1]), frames,
                                                args.iterations)), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (every leg runs on one GPU).",
              file=sys.stderr)
        return 2

    import jgs2
    from jgs2 import scenes
    from jgs2.batch import JGS2Batch
    from jgs2.serving import nearest_rank

    if jgs2.device_count() < 1:
        print("No HIP device visible; bench.py needs an MI355X.", file=sys.stderr)
        return 1

    it = args.iterations
    settings = dict(scenes.MIXED_SETTINGS)
    t0 = time.time()
    specs = scenes.mixed_scene_specs(count=args.scenes, seed=args.seed)
    inputs = build_scenes(specs)
    V = sum(x.vertex_count for x in inputs)
    T = sum(x.tet_count for x in inputs)
    NT = sum(x.cloth.triangle_count for x in inputs if x.cloth is not None)
    log(f"{len(inputs)} scenes: V={V} T={T} triangles={NT} built in {time.time() - t0:.1f}s")

    # ---------------------------------------------------------- batched
    t0 = time.time()
    batch = JGS2Batch.create(_capacity(inputs), settings)
    members = [batch.admit(x) for x in inputs]
    free = sum(m.free_vertex_count for m in members)
    batch.use_graphs(not args.no_graphs)
    info = batch.info()
    log(f"batch: {info['scene_count']} scenes admitted in {time.time() - t0:.1f}s, arena {info['arena_bytes'] / 1e9:.2f} GB")
    for w in range(args.warmup):
        batch.step_frames_exact_iterations(1, it, settings)
        batch.sync()
        log(f"warmup frame {w + 1}/{args.warmup} done")
        if w == 0:
            autotune_ms = batch.autotune(dict(settings), repeats=5)
            info = batch.info()
            log(f"autotune: block {info['solve_block']} xcd_remap {info['xcd_remap']} ({autotune_ms:.3f} ms/launch)")
    info = batch.info()
    batch.sync()

    def listed():
        return [m.contact_info()["candidates"] if x.contact is not None else 0 for x, m in zip(inputs, members)]

    lists_before = listed()
    # Timed region: the frames are launched pass by pass with HIP events
    # around every local-solve launch on the batch's stream, so `value` and
    # the roofline come from the same frames.
    batch.profile_begin()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.step_frames_exact_iterations(1, it, settings)
    batch.sync()
    elapsed = time.perf_counter() - t0
    solve_ms, launches, _ = batch.profile_end()
    value = args.steps * it * free / elapsed
    log(f"batched: {args.steps} frames in {elapsed:.3f}s, {value:.4g} vertex-updates/s, local solve {solve_ms:.3f} ms "
        f"per launch over {launches} launches")
    lists_after = listed()
    lat = []
    for _ in range(args.latency_frames):
        a = time.perf_counter()
        batch.step_frames_exact_iterations(1, it, settings)
        batch.sync()
        lat.append((time.perf_counter() - a) * 1e3)
    if lat:
        log(f"batched frame latency (graph replay) p50 {nearest_rank(lat, 50):.2f} ms")
    # Candidate lists change at rebuilds (folded sheets separate): the
    # algorithmic bytes of a sweep count the mean of each scene's list at
    # the start and the end of the timed frames.
    lists = [(a + b) / 2 for a, b in zip(lists_before, lists_after)]
    bytes_per_sweep = sum(scene_bytes_per_sweep(x, c) for x, c in zip(inputs, lists))
    achieved = bytes_per_sweep / (solve_ms * 1e-3) / 1e9 if solve_ms > 0 else 0.0
    traffic = measured_traffic(f"batched-uneven-scenes-{len(inputs)}", info["solve_block"])
    # Host readback of every scene's positions (PCIe), once per frame if a
    # caller wanted them all: the PCIe-inclusive rate next to `value`.
    batch.sync()
    t0 = time.perf_counter()
    finite = all(bool(np.isfinite(m.read_positions()).all()) for m in members)
    readback_s = time.perf_counter() - t0
    # Accepted iterations of the last frame, per scene, weighted by its free
    # vertices: the share of the vertex updates that the assembled
    # accept/revert kept.
    reverted, accepted_updates, builds, contact_scenes = 0, 0, 0, 0
    for x, m in zip(inputs, members):
        h = m.read_globalization_history().reshape(64, 8, 4)[:it]
        reverted += int(h[:, 1, 3].sum())
        accepted_updates += int(h[:, 1, 0].sum()) * m.free_vertex_count
        if x.contact is not None:
            contact_scenes += 1
            builds += int(m.contact_info()["builds"])
    accepted_fraction = accepted_updates / (it * free)
    batch.destroy()

    # ---------------------------------------------------------- plain
    plain = None if args.no_plain else measure_plain(inputs, settings, it, args.plain_steps)

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
            "workload": f"batched-uneven-scenes-{len(inputs)}",
            "scenes": len(inputs),
            "tet_scenes": sum(1 for s in specs if s["kind"] == "tet"),
            "cloth_scenes": sum(1 for s in specs if s["kind"] == "cloth"),
            "vertices": V,
            "free_vertices": free,
            "tetrahedra": T,
            "cloth_triangles": NT,
            "cubature_k": max(x.cubature_k for x in inputs),
            "schedule": "jacobi",
            "iterations_per_frame": it,
            "seed": args.seed,
            "mode": "batched (one launch per pass over all scenes); timed frames launched pass by pass with HIP "
                    "events around each local solve" + ("" if args.no_graphs else ", latency frames replayed from a "
                                                        "hipGraph"),
            "solve_block": info["solve_block"],
            "xcd_remap": bool(info["xcd_remap"]),
        },
        "frame_ms": {"p50": nearest_rank(lat, 50), "p99": nearest_rank(lat, 99), "frames": len(lat)} if lat else None,
        "plain": plain,
        "batched_over_plain": value / plain["vertex_updates_per_s"] if plain and plain["vertex_updates_per_s"] > 0 else None,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_solve_lanes (local solve, every launch of a sweep)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_per_sweep,
            "ms_per_launch": solve_ms,
            "launches_profiled": launches,
        },
        "positions_finite": finite,
        "readback_ms_per_frame": readback_s * 1e3,
        "pcie_inclusive_vertex_updates_per_s": it * free / (elapsed / args.steps + readback_s),
        "reverted_iterations_last_frame_all_scenes": reverted,
        "accepted_fraction_last_frame": accepted_fraction,
        "accepted_vertex_updates_per_s": value * accepted_fraction,
        "self_contact": {"cloth_scenes_with_self_ipc": contact_scenes,
                         "candidates_listed_timed_start": int(sum(lists_before)),
                         "candidates_listed_timed_end": int(sum(lists_after)),
                         "broad_phase_builds_total": builds},
    }
    del inputs, members

    if args.contact_frames > 0:
        result["contact_boxes"] = measure_contact_boxes(args.contact_frames)
    if args.serve_seconds > 0:
        result["serving"] = measure_serving(args.serve_seconds, args.seed)
    if args.large_steps > 0:
        cells = tuple(int(c) for c in args.large_cells.split(","))
        result["single_large_mesh"] = measure_large(cells, it, args.large_steps)
    if args.precompute_cells:
        result["precompute"] = measure_precompute(tuple(int(c) for c in args.precompute_cells.split(",")))
        if not args.no_cpu_baseline:
            result["precompute"]["cpu_baseline"] = precompute_cpu_baseline()
    if not args.no_accuracy:
        result["accuracy"] = accuracy_check(it, args.seed)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(it, args.seed)
    print(json.dumps(result), flush=True)
    return 0


def copy_calibration():
    """A plain copy of a cold 1 GiB device buffer into another (twice, so
    the buffers exceed every cache): the known-size kernel that scales the
    FETCH_SIZE and WRITE_SIZE counters to bytes."""
    import torch

    n = 1 << 28  # 2^28 floats = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    c.copy_(b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 2}), flush=True)
    return 0


def profile_contact(frames):
    """The contact-box workload alone, eager launches (profiler driver)."""
    import jgs2
    from jgs2 import scenes

    x = scenes.contact_boxes()
    settings = dict(scenes.CONTACT_BOXES_SETTINGS)
    it = scenes.CONTACT_BOXES_ITERATIONS
    solver = jgs2.JGS2GpuSolver.create(x, settings)
    solver.step_frames_exact_iterations(3, it, settings)
    solver.await_idle()
    b0 = solver.contact_info()["builds"]
    t0 = time.perf_counter()
    solver.step_frames_exact_iterations(frames, it, settings)
    solver.await_idle()
    dt = time.perf_counter() - t0
    info = solver.contact_info()
    log(f"contact boxes: {frames} frames, {dt / frames * 1e3:.2f} ms/frame, "
        f"{(info['builds'] - b0) / frames:.2f} builds/frame, list {info['candidates']}")
    solver.destroy()
    return 0


def profile_large(lanes_list, cells_arg, iterations):
    """Two Jacobi frames of the layered box per lane count (profiler driver)."""
    import jgs2
    from jgs2 import scenes
    from jgs2.batch import JGS2Batch

    cells = tuple(int(c) for c in cells_arg.split(","))
    scene = scenes.layered_box(cells=cells)
    settings = dict(scenes.LAYERED_BOX_SETTINGS)
    solver = jgs2.JGS2GpuSolver.create(scene, settings)
    batch = JGS2Batch.of(solver)
    for lanes in lanes_list:
        batch.set_launch(256, True, lanes)
        for f in range(2):
            solver.step_exact_iterations(iterations, settings)
            solver.await_idle()
        log(f"lanes {lanes}: 2 frames done")
    solver.destroy()
    return 0


def measured_traffic(workload, block):
    """HBM bytes per local-solve launch of this workload and block size from
    the rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarized in
    profiles/traffic.json (null when not profiled)."""
    path = REPO / "profiles" / "traffic.json"
    if not path.exists():
        return None
    entry = json.loads(path.read_text()).get(workload, {}).get(str(block))
    return None if entry is None else entry["hbm_bytes_per_launch"]


def _capacity(inputs, slack_scenes=0):
    from jgs2.batch import BatchCapacity, footprint

    fp = [footprint(x) for x in inputs]
    total = {k: sum(f[k] for f in fp) for k in fp[0]}
    return BatchCapacity(max_vertices=total["vertices"], max_tets=total["tets"], max_scenes=len(inputs) + slack_scenes,
                         cubature_k=max(x.cubature_k for x in inputs), max_triangles=total["triangles"],
                         max_hinges=total["hinges"], max_candidates=total["candidates"], max_bodies=total["bodies"],
                         max_surface_triangles=total["surface_triangles"], max_surface_edges=total["surface_edges"])


def measure_plain(inputs, settings, iterations, steps):
    """The same scenes, one solver each, every solver stepped by its own
    submission per frame."""
    import jgs2

    t0 = time.time()
    solvers = [jgs2.JGS2GpuSolver.create(x, settings) for x in inputs]
    free = sum(s.free_vertex_count for s in solvers)
    log(f"plain: {len(solvers)} solvers created in {time.time() - t0:.1f}s")
    for s in solvers:
        s.step_exact_iterations(iterations, settings)
    for s in solvers:
        s.await_idle()
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in solvers:
            s.step_exact_iterations(iterations, settings)
    for s in solvers:
        s.await_idle()
    dt = time.perf_counter() - t0
    for s in solvers:
        s.destroy()
    out = {"mode": "plain (one solver and one submission per scene)", "steps": steps, "ms_per_step": dt / steps * 1e3,
           "vertex_updates_per_s": steps * iterations * free / dt}
    log(f"plain: {out['ms_per_step']:.1f} ms/frame, {out['vertex_updates_per_s']:.4g} vertex-updates/s")
    return out


def measure_contact_boxes(frames):
    """The large contact workload (twenty 8x8x8-cell boxes dropped into the
    trough, broad-phase candidates) as one plain scene with graph frames:
    throughput, frame latency, broad-phase builds and the nearest pair."""
    import jgs2
    from jgs2 import scenes
    from jgs2.batch import JGS2Batch
    from jgs2.serving import nearest_rank
    from jgs2.solver import ContactData

    t0 = time.time()
    x = scenes.contact_boxes()
    settings = dict(scenes.CONTACT_BOXES_SETTINGS)
    it = scenes.CONTACT_BOXES_ITERATIONS
    solver = jgs2.JGS2GpuSolver.create(x, settings)
    JGS2Batch.of(solver).use_graphs(True)
    log(f"contact boxes: V={x.vertex_count} T={x.tet_count} capacity {x.contact.candidate_capacity}, "
        f"ready in {time.time() - t0:.1f}s")
    solver.step_frames_exact_iterations(1, it, settings)
    solver.await_idle()
    b0 = solver.contact_info()["builds"]
    lat = []
    for _ in range(frames):
        a = time.perf_counter()
        solver.step_frames_exact_iterations(1, it, settings)
        solver.await_idle()
        lat.append((time.perf_counter() - a) * 1e3)
    info = solver.contact_info()
    packed, vt, _ = solver.read_contact_candidates()
    p = np.asarray(solver.read_positions(), np.float64).reshape(-1, 4)[:, :3]
    sys.path.insert(0, str(REPO / "tests"))
    from ipc_geometry import candidate_distances

    nearest = float(candidate_distances(p, ContactData(packed.reshape(-1), vt, len(packed) - vt)).min())
    solver.destroy()
    ms = float(np.mean(lat))
    out = {
        "workload": "contact-boxes: 20 boxes of 8x8x8 cells (3,072 tetrahedra each) dropped at seeded offsets and "
                    "yaws into the pinned V-trough; dhat 1 mm, kappa 1e5, friction 0.3, 12 Jacobi iterations",
        "vertices": x.vertex_count, "tetrahedra": x.tet_count, "free_vertices": solver.free_vertex_count,
        "frames": frames, "ms_per_frame": ms, "vertex_updates_per_s": solver.free_vertex_count * it / (ms * 1e-3),
        "frame_ms": {"p50": nearest_rank(lat, 50), "p99": nearest_rank(lat, 99)},
        "broad_phase": {"candidates": int(info["candidates"]), "capacity": int(info["capacity"]),
                        "builds_per_frame": (int(info["builds"]) - b0) / frames, "reach_m": float(info["reach"])},
        "nearest_pair_m": nearest, "minimum_distance_m": settings["ipc_minimum_distance"],
        "interpenetration_free": nearest > settings["ipc_minimum_distance"],
    }
    log(f"contact boxes: {ms:.1f} ms/frame, {out['vertex_updates_per_s']:.4g} vertex-updates/s, "
        f"{out['broad_phase']['builds_per_frame']:.1f} builds/frame, nearest pair {nearest:.4g} m")
    return out


def measure_serving(seconds, seed):
    """Poisson arrivals (2000 requests/s) of contact-box requests (the large
    contact workload's generator, seeded placements) asking for 1 to 120
    frames, served for `seconds` by continuous batching under an HBM budget
    and by one solver per request under the same budget."""
    from jgs2 import scenes
    from jgs2.serving import BatchServer, PlainServer, Request, capacity_for_budget, poisson_arrivals

    rate, budget, max_scenes = 2000.0, 64 << 30, 512
    t0 = time.time()
    pool, _ = scenes.contact_boxes_pool(16)
    log(f"serving pool: {len(pool)} contact-box scenes built in {time.time() - t0:.1f}s")
    settings = dict(scenes.CONTACT_BOXES_SETTINGS)
    it = scenes.CONTACT_BOXES_ITERATIONS
    rng = np.random.default_rng(seed)
    arrivals = poisson_arrivals(rate, seconds, seed)
    frames = rng.integers(1, 121, size=arrivals.size)
    picks = rng.integers(0, len(pool), size=arrivals.size)

    def stream():
        return [Request(id=i, scene=pool[picks[i]], frames=int(frames[i]), arrival=float(arrivals[i]))
                for i in range(arrivals.size)]

    cap = capacity_for_budget(pool, budget, max_scenes=max_scenes)
    out = {"workload": f"Poisson {rate:g} requests/s of contact-box scenes (20 boxes of 8x8x8 cells in the "
                       f"trough, 16 seeded placements), 1-120 frames each, {it} exact Jacobi sweeps per frame, "
                       f"{seconds:g} s",
           "offered_requests": int(arrivals.size), "hbm_budget_bytes": budget, "max_resident_scenes": cap.max_scenes}
    server = BatchServer(cap, settings, it, hbm_budget_bytes=budget, use_graphs=True)
    out["arena_bytes"] = server.arena_bytes
    def progress(label):
        return lambda r, now: log(f"serving {label}: {now:.0f} s, {r.ticks} ticks, {len(r.completed)} completed")

    rep = server.run(stream(), duration=seconds, progress=progress("batched"))
    server.close()
    out["batched"] = rep.summary()
    log(f"serving batched: {out['batched']['vertex_updates_per_s']:.4g} vertex-updates/s, "
        f"frame p50 {out['batched']['frame_ms']['p50']:.2f} ms")
    plain = PlainServer(settings, it, hbm_budget_bytes=budget, max_scenes=cap.max_scenes)
    rep = plain.run(stream(), duration=seconds, progress=progress("plain"))
    plain.close()
    out["plain"] = rep.summary()
    log(f"serving plain: {out['plain']['vertex_updates_per_s']:.4g} vertex-updates/s, "
        f"frame p50 {out['plain']['frame_ms']['p50']:.2f} ms")
    p = out["plain"]["vertex_updates_per_s"]
    out["batched_over_plain"] = out["batched"]["vertex_updates_per_s"] / p if p > 0 else None
    return out


def measure_large(cells, iterations, steps):
    """The layered box as one plain scene: throughput, frame latency and the
    local-solve roofline under Jacobi, and Gauss-Seidel throughput."""
    import jgs2
    from jgs2 import scenes
    from jgs2.serving import nearest_rank

    scene = scenes.layered_box(cells=cells)
    out = {"workload": f"layered-box-{cells[0]}x{cells[1]}x{cells[2]}", "vertices": scene.vertex_count,
           "tetrahedra": scene.tet_count}
    for schedule in ("jacobi", "graph-colored-gauss-seidel"):
        settings = dict(scenes.LAYERED_BOX_SETTINGS, schedule=schedule)
        solver = jgs2.JGS2GpuSolver.create(scene, settings)
        free = solver.free_vertex_count
        solver.step_exact_iterations(iterations, settings)
        solver.await_idle()
        t0 = time.perf_counter()
        for _ in range(steps):
            solver.step_exact_iterations(iterations, settings)
        solver.await_idle()
        dt = time.perf_counter() - t0
        leg = {"ms_per_step": dt / steps * 1e3, "vertex_updates_per_s": steps * iterations * free / dt,
               "colors": solver.schedule_color_count}
        if schedule == "jacobi":
            lat = []
            for _ in range(5):
                a = time.perf_counter()
                solver.step_exact_iterations(iterations, settings)
                solver.await_idle()
                lat.append((time.perf_counter() - a) * 1e3)
            leg["frame_ms"] = {"p50": nearest_rank(lat, 50), "p99": nearest_rank(lat, 99)}
            solver.profile_begin()
            solver.step_exact_iterations(iterations, settings)
            solve_ms, launches, _ = solver.profile_end()
            b = scenes.algorithmic_bytes_per_sweep(scene.vertex_count, scene.tet_count, scene.cubature_k)
            achieved = b / (solve_ms * 1e-3) / 1e9
            leg["roofline"] = {"bound": "hbm", "kernel": "k_solve_lanes (local solve, every launch of a sweep)", "achieved": achieved,
                               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                               "algorithmic_bytes_per_launch": b, "ms_per_launch": solve_ms,
                               "launches_profiled": launches}
            hist = solver.read_globalization_history().reshape(64, 8, 4)
            leg["reverted_iterations_last_frame"] = int(hist[:iterations, 1, 3].sum())
            # exact FP64 diagnostics of the frame's result (one evaluation, host sync included)
            solver.read_oracle_diagnostics(settings)
            a = time.perf_counter()
            d = solver.read_oracle_diagnostics(settings)
            leg["diagnostics"] = {"ms": (time.perf_counter() - a) * 1e3, "energy": d["energy"],
                                  "elasticity": d["elasticity"], "inertia": d["inertia"],
                                  "relative_residual": d["relative_residual"],
                                  "minimum_deformation_determinant": d["minimum_deformation_determinant"],
                                  "finite": bool(d["finite"])}
        solver.destroy()
        out[schedule] = leg
        log(f"large mesh {schedule}: {leg['ms_per_step']:.2f} ms/frame")
    return out


def _host_threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def _train(scene, materials, material_ids, rings, records=None):
    from jgs2.precompute import precompute_cubature

    rest = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    fixed = scene.vertex_info.reshape(-1, 4)[:, 2] != 0
    t0 = time.perf_counter()
    r = precompute_cubature(rest, scene.tet_indices.reshape(-1, 4), fixed, material_ids, materials, 1.0 / 60.0,
                            scene.cubature_k, rings=rings, threads=_host_threads(), residual_bound=1.0)
    seconds = time.perf_counter() - t0
    free = ~fixed
    if records is not None:
        records.append(r)
    return {"rings": rings, "seconds": seconds, "vertices_per_s": float(free.sum()) / seconds,
            "phases_s": {k: round(float(v), 4) for k, v in r.timings.items() if k.endswith("_s")},
            "modes_on_gpu": bool(r.timings["gpu"]), "gpu_cg_ms": float(r.timings["gpu_cg_ms"]),
            "selection_on_gpu": bool(r.timings["gpu_training"]), "gpu_select_ms": float(r.timings["gpu_select_ms"]),
            "gpu_cg_iterations": r.timings["gpu_cg_iterations"],
            "packed_residual_max": float(np.nanmax(r.residuals)), "packed_residual_mean": float(np.nanmean(r.residuals)),
            "records": int((r.tet_ids != 0xFFFFFFFF).sum()), "host_threads": _host_threads()}


def measure_precompute(cells):
    """Cubature precompute (jgs2_precompute_cubature) of the stable Neo-Hookean
    cantilever: the reference's whole-mesh basis (rings 0) and the 2-ring
    basis, K = 6, dt = 1/60; host training, modes on the dense factor."""
    from jgs2 import scenes

    scene = scenes.cantilever_beam(cells=cells, size=(cells[0] * 0.25, cells[1] * 0.25, cells[2] * 0.25))
    out = {"workload": f"cantilever-{cells[0]}x{cells[1]}x{cells[2]}", "vertices": scene.vertex_count,
           "tetrahedra": scene.tet_count, "samples": scene.cubature_k}
    mats = [(1000.0, 1e5, 0.4, "stable-neo-hookean")]
    for rings in (0, 2):
        out[f"rings_{rings}"] = _train(scene, mats, np.zeros(scene.tet_count), rings)
        log(f"precompute {out['workload']} rings {rings}: {out[f'rings_{rings}']['seconds']:.2f} s")
    return out


def precompute_cpu_baseline():
    """The FP64 oracle's precompute (the reference's dense algorithm, one
    thread) beside the library's on the same 4x2x2 stable Neo-Hookean
    cantilever (45 vertices, 96 tetrahedra, whole-mesh basis): the oracle
    evaluates every tetrahedron as a candidate, as the reference does."""
    from jgs2 import scenes
    from jgs2.precompute import precompute_cubature

    sys.path.insert(0, str(REPO / "oracle"))
    import jgs2_oracle as O

    scene = scenes.cantilever_beam(cells=(4, 2, 2), size=(1.0, 0.5, 0.5))
    rest = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
    fixed = scene.vertex_info.reshape(-1, 4)[:, 2] != 0
    tets = scene.tet_indices.reshape(-1, 4)
    mats = [(1000.0, 1e5, 0.4, 1)]
    mid = np.zeros(scene.tet_count)
    t0 = time.perf_counter()
    ot, ow, _, _ = O.precompute_cubature(rest, tets, fixed, mid, mats, 1.0 / 60.0, 6, bound=1.0)
    oracle_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = precompute_cubature(rest, tets, fixed, mid, mats, 1.0 / 60.0, 6, rings=0, threads=_host_threads(),
                            residual_bound=1.0)
    lib_s = time.perf_counter() - t0
    return {"value": scene.vertex_count / oracle_s, "unit": "vertices/s", "cores": 1, "kind": "port",
            "sample": "FP64 oracle precompute of the 4x2x2 cantilever (45 vertices, 96 tetrahedra, K = 6)",
            "oracle_s": oracle_s, "library_s": lib_s, "library_threads": _host_threads(),
            "same_records": bool(np.array_equal(ot, r.tet_ids) and np.array_equal(ow, r.weights))}


def precompute_layered_box(cells, rings, frames=0, iterations=8):
    """Cubature precompute of the layered box (two stable Neo-Hookean layer
    materials, E = 1e4 / 1e6 Pa): modes and selection on the GPU, bases and
    columns on the host threads. With `frames`, Jacobi frames of the box on
    the trained records and on the stand-in records: frame time and the
    iterations the assembled accept/revert rejected."""
    import jgs2
    from jgs2 import scenes

    scene = scenes.layered_box(cells=cells)
    mu = scene.tet_meta.reshape(-1, 4)[:, 2]
    mid = (mu > np.sqrt(mu.min() * mu.max())).astype(np.uint32)
    mats = [(1000.0, 1e4, 0.45, "stable-neo-hookean"), (1000.0, 1e6, 0.45, "stable-neo-hookean")]
    out = {"workload": f"layered-box-{cells[0]}x{cells[1]}x{cells[2]}", "vertices": scene.vertex_count,
           "tetrahedra": scene.tet_count, "samples": scene.cubature_k}
    log(f"precompute {out['workload']} rings {rings}: start")
    trained = []
    out.update(_train(scene, mats, mid, rings, trained))
    log(f"precompute {out['workload']} rings {rings}: {out['seconds']:.1f} s")
    if frames > 0:
        rest = scene.vertex_rest.reshape(-1, 4)[:, :3].astype(np.float64)
        settings = dict(scenes.LAYERED_BOX_SETTINGS)
        for name in ("stand-in", "trained"):
            if name == "trained":
                r = trained[0]
                scene.cubature_tet_ids = np.ascontiguousarray(r.tet_ids.reshape(-1))
                scene.cubature_weights = np.ascontiguousarray(r.weights.reshape(-1))
                scene.cubature_basis = np.ascontiguousarray(r.basis.reshape(-1))
            solver = jgs2.JGS2GpuSolver.create(scene, settings)
            solver.step_exact_iterations(iterations, settings)
            solver.await_idle()
            reverted, t0 = 0, time.perf_counter()
            for _ in range(frames):
                solver.step_exact_iterations(iterations, settings)
                solver.await_idle()
                hist = solver.read_globalization_history().reshape(64, 8, 4)
                reverted += int(hist[:iterations, 1, 3].sum())
            ms = (time.perf_counter() - t0) / frames * 1e3
            x = solver.read_positions().reshape(-1, 4)[:, :3]
            out[f"frames_{name}"] = {"frames": frames, "ms_per_frame": ms, "reverted_iterations": reverted,
                                     "iterations": frames * iterations, "finite": bool(np.isfinite(x).all()),
                                     "largest_motion_m": float(np.abs(x - rest).max())}
            log(f"layered box on {name} records: {ms:.2f} ms/frame, {reverted} of {frames * iterations} "
                f"iterations reverted")
            solver.destroy()
    return out


def accuracy_check(iterations, seed, frames=36):
    """GPU trajectories of a sample of the mixed stream against the FP64
    oracle, the scenes stepped together as one batch exactly as the
    headline steps them (cloth self-contact through the device broad phase;
    the oracle steps the same sheets with the static superset, which the
    broad phase reproduces bit for bit). The sample: the generator's first
    six tetrahedral scenes of at most 3,000 tetrahedra, and folded sheets of
    6, 8 and 10 cells built like the generator's (its own sheets, 32 cells
    and up, have superset lists too long for the oracle).

    Bars: non-contact (tetrahedral) scenes within 1e-3 relative L2 of the
    FP64 oracle's displacement; contact scenes interpenetration-free at
    every frame (independent distances over the device's candidate list)
    with an Armijo decrease in every accepted local solve. The FP32
    oracle's own distance from FP64 is reported beside each scene: the
    reference computes in FP32, and where that distance exceeds 1e-3 the
    gate is beyond FP32 arithmetic on this algorithm."""
    from concurrent.futures import ThreadPoolExecutor

    from jgs2 import scenes
    from jgs2.batch import JGS2Batch
    from jgs2.solver import ContactData

    sys.path.insert(0, str(REPO / "oracle"))
    sys.path.insert(0, str(REPO / "tests"))
    import jgs2_oracle as O
    from ipc_geometry import candidate_distances

    settings = dict(scenes.MIXED_SETTINGS)
    tet_specs = [s for s in scenes.mixed_scene_specs(count=512, seed=seed)
                 if s["kind"] == "tet" and 6 * int(np.prod(s["cells"])) <= 3000][:6]
    specs = tet_specs + [{"kind": "cloth", "cells": n} for n in (6, 8, 10)]
    gpu_inputs = [scenes.build_mixed_scene(s, "broad_phase") for s in specs]
    ref_inputs = [scenes.build_mixed_scene(s, "superset") for s in specs]
    batch = JGS2Batch.create(_capacity(gpu_inputs), settings)
    members = [batch.admit(x) for x in gpu_inputs]
    dmin = settings["ipc_minimum_distance"]
    nearest = {i: np.inf for i, s in enumerate(specs) if s["kind"] == "cloth"}
    armijo_ok = True
    for _ in range(frames):
        batch.step_frames_exact_iterations(1, iterations, settings)
        for i in nearest:
            m = members[i]
            packed, vt, _ = m.read_contact_candidates()
            p = m.read_positions().reshape(-1, 4)[:, :3].astype(np.float64)
            if len(packed):
                d = candidate_distances(p, ContactData(packed.reshape(-1), vt, len(packed) - vt))
                nearest[i] = min(nearest[i], float(d.min()))
            diag = m.read_local_diagnostics().reshape(-1, 4, 4)
            ok = diag[:, 3, 3] == 0
            armijo_ok = armijo_ok and bool(np.all(diag[ok, 2, 0] <= diag[ok, 2, 1]))
    gpu = [m.read_positions().reshape(-1, 4)[:, :3].astype(np.float64) for m in members]
    # Relative residual of the last iteration (history row 4, w): about 1
    # means the frame's sweeps made no progress on the implicit system.
    residuals = [float(m.read_globalization_history().reshape(64, 8, 4)[iterations - 1, 4, 3]) for m in members]
    batch.destroy()

    def oracle_run(args):
        x, precision = args
        ref = O.Oracle(x, precision, settings=settings)
        ref.step_frames(frames, iterations, settings)
        return ref.positions()

    with ThreadPoolExecutor(_host_threads()) as pool:
        ref64 = list(pool.map(oracle_run, [(x, 64) for x in ref_inputs]))
        ref32 = list(pool.map(oracle_run, [(x, 32) for x in ref_inputs]))
    rows = []
    for s, x, g, a, b in zip(specs, gpu_inputs, gpu, ref64, ref32):
        x0 = x.positions.reshape(-1, 4)[:, :3].astype(np.float64)
        disp = max(float(np.linalg.norm(a - x0)), 1e-30)
        row = {"kind": s["kind"], "cells": list(s["cells"]) if s["kind"] == "tet" else s["cells"],
               "young": s.get("young"), "vertices": x.vertex_count,
               "last_frame_relative_residual": residuals[len(rows)],
               "gpu_vs_fp64_displacement": float(np.linalg.norm(g - a)) / disp,
               "fp32_oracle_vs_fp64_displacement": float(np.linalg.norm(b - a)) / disp,
               "gpu_vs_fp64_positions": float(np.linalg.norm(g - a) / np.linalg.norm(a))}
        rows.append(row)
    tet = [r for r in rows if r["kind"] == "tet"]
    worst_tet = max(r["gpu_vs_fp64_displacement"] for r in tet)
    interpenetration_free = all(v > dmin for v in nearest.values())
    within_floor = all(r["gpu_vs_fp64_displacement"] <= max(1e-3, 3.0 * r["fp32_oracle_vs_fp64_displacement"])
                       for r in rows)
    out = {"reference": "fp64 oracle", "frames": frames, "tolerance": 1e-3,
           "sample": f"{len(tet)} tetrahedral scenes of the generator (<= 3,000 tetrahedra) and folded sheets of "
                     "6, 8, 10 cells with self-contact, one batch",
           "scenes": rows,
           "max_rel_l2_of_displacement_tet": worst_tet,
           "max_rel_l2_of_displacement_all": max(r["gpu_vs_fp64_displacement"] for r in rows),
           "max_rel_l2_positions": max(r["gpu_vs_fp64_positions"] for r in rows),
           "contact": {"nearest_pair_m": min(nearest.values()), "minimum_distance_m": dmin,
                       "interpenetration_free": interpenetration_free, "armijo_every_accepted_solve": armijo_ok},
           "within_tol": bool(worst_tet <= 1e-3 and interpenetration_free and armijo_ok),
           "gpu_within_fp32_oracle_floor": within_floor}
    log(f"accuracy: tet displacement error {worst_tet:.3e} (FP32 oracle "
        f"{max(r['fp32_oracle_vs_fp64_displacement'] for r in tet):.3e}), nearest contact pair "
        f"{out['contact']['nearest_pair_m']:.4g} m, within_tol {out['within_tol']}")
    return out


def _first_build_list(x, settings):
    """The candidate list of the device's first broad-phase build for a
    cloth scene (the superset restricted to the build's reach at the initial
    pose, in canonical order; equal pair for pair to the oracle's
    restriction, tests/test_gpu_broad_phase.py), as static candidates."""
    import jgs2
    from jgs2.solver import ContactData

    solver = jgs2.JGS2GpuSolver.create(x, settings)
    packed, vt, _ = solver.read_contact_candidates()
    solver.destroy()
    return ContactData(packed_indices=np.ascontiguousarray(packed.reshape(-1)), vertex_triangle_count=vt,
                       edge_edge_count=len(packed) - vt)


def cpu_baseline(iterations, seed, frames=1, budget_s=15.0):
    """The oracle's FP32 step on the host threads over a bounded sample of
    the headline workload: the generator's scenes in order (tetrahedral
    scenes up to 20,000 tetrahedra, sheets up to 64 cells, each sheet with
    the candidate list of the device's first broad-phase build, so its
    self-contact rows are part of the work), one frame each, until about
    `budget_s` seconds; plus the cantilever workload in full (60 frames) and
    one frame of the 64x64x128 layered box."""
    from jgs2 import scenes

    sys.path.insert(0, str(REPO / "oracle"))
    import jgs2_oracle as O

    threads = _host_threads()
    O.set_threads(threads)
    settings = dict(scenes.MIXED_SETTINGS)
    specs = [s for s in scenes.mixed_scene_specs(count=512, seed=seed)
             if (s["kind"] == "tet" and 6 * int(np.prod(s["cells"])) <= 20000) or
             (s["kind"] == "cloth" and s["cells"] <= 64)]
    updates, dt, used, verts, rows = 0, 0.0, 0, 0, 0
    for spec in specs:
        if dt >= budget_s:
            break
        x = scenes.build_mixed_scene(spec)
        if spec["kind"] == "cloth":
            x.contact = _first_build_list(x, settings)
            rows += x.contact.candidate_count
        ref = O.Oracle(x, 32, settings=settings)
        free = int((x.vertex_info.reshape(-1, 4)[:, 2] == 0).sum())
        t0 = time.perf_counter()
        ref.step_frames(frames, iterations, settings)
        dt += time.perf_counter() - t0
        updates += frames * iterations * free
        used += 1
        verts += x.vertex_count
    out = {"value": updates / dt, "unit": "vertex-updates/s", "cores": threads, "kind": "port",
           "sample": f"oracle FP32 step on {threads} host threads, the first {used} scenes of the mixed generator "
                     f"up to 20,000 tetrahedra / 64x64 cloth cells ({verts} vertices, {rows} self-contact "
                     f"candidates), {frames} frame x {iterations} Jacobi sweeps each, {dt:.1f} s"}
    log(f"cpu baseline: {updates} vertex-updates in {dt:.2f}s on {threads} threads")
    legs = {}
    beam = scenes.cantilever_beam()
    st = dict(scenes.CANTILEVER_SETTINGS)
    ref = O.Oracle(beam, 32)
    free = int((beam.vertex_info.reshape(-1, 4)[:, 2] == 0).sum())
    t0 = time.perf_counter()
    ref.step_frames(scenes.CANTILEVER_FRAMES, scenes.CANTILEVER_ITERATIONS, st)
    t = time.perf_counter() - t0
    legs["cantilever-16x4x4"] = {"frames": scenes.CANTILEVER_FRAMES, "seconds": t,
                                 "vertex_updates_per_s": scenes.CANTILEVER_FRAMES * scenes.CANTILEVER_ITERATIONS * free / t}
    box = scenes.layered_box()
    st = dict(scenes.LAYERED_BOX_SETTINGS)
    ref = O.Oracle(box, 32)
    free = int((box.vertex_info.reshape(-1, 4)[:, 2] == 0).sum())
    t0 = time.perf_counter()
    ref.step_frames(1, iterations, st)
    t = time.perf_counter() - t0
    legs["layered-box-64x64x128"] = {"frames": 1, "seconds": t, "vertex_updates_per_s": iterations * free / t}
    del ref, box
    out["workloads"] = legs
    O.set_threads(1)
    log(f"cpu baseline legs: cantilever {legs['cantilever-16x4x4']['vertex_updates_per_s']:.4g}, layered box "
        f"{legs['layered-box-64x64x128']['vertex_updates_per_s']:.4g} vertex-updates/s")
    return out


if __name__ == "__main__":
    sys.exit(main())
