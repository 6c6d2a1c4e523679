"""Synthetic workloads of the JGS2 step, built by the host scene setup in
libjgs2hip (include/jgs2hip_scene.h).

* `cantilever_beam`: a 4x4x16-hex beam (1 m x 1 m x 4 m, 425 vertices, 1536
  tetrahedra), stable Neo-Hookean E = 1e5 Pa, nu = 0.4, rho = 1000, the x = 0
  face pinned, gravity and a scripted soft tip target following the
  reference's minimal-scene keyframes (src/scenes/index.ts:252-328).
* `layered_box`: a 1 m x 1 m x 2 m box on a 64x64x128 hex grid (545,025
  vertices, 3,145,728 tetrahedra) with seeded U(-0.1h, 0.1h) vertex jitter,
  alternating 8-cell z-slabs of E = 1e4 / 1e6 Pa, nu = 0.45, the z = 0 face
  pinned. `scale` shrinks the grid for tests.
* `cloth_drape`: the reference's pinned cloth drape, a fixed collider block
  under a pre-folded StVK triangle sheet with dihedral bending.
* `mixed_scene_specs` / `build_mixed_scene`: the batched uneven-scene
  workload, seeded: 70 % tetrahedral beams and boxes with log-uniform
  tetrahedron counts and Young's moduli, 30 % cloth sheets of n x n cells.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .solver import BroadPhaseContact, ClothData, ContactData, GpuInput, estimate_contact_capacity

# Minimal-scene scripted target (src/scenes/index.ts:252-262).
TARGET_PULL_START_FRAME = 48
TARGET_PULL_END_FRAME = 90
TARGET_RELEASE_FRAME = 108
TARGET_KEYFRAME_INTERVAL = 3
TARGET_STIFFNESS = 30000.0
TARGET_PULL_OFFSET = (0.45, 0.65, 0.25)


def _desc(cells, origin, size, *, jitter=0.0, seed=0, pin_axis=0, model=N.MATERIAL_STABLE_NEO_HOOKEAN,
          density=1000.0, poisson=0.4, young=(1e5, 1e5), layer_axis=2, layer_cells=0, cubature_k=6,
          rest_stiffness=False) -> N.CuboidDesc:
    d = N.CuboidDesc()
    for i in range(3):
        d.cells[i] = int(cells[i])
        d.origin[i] = float(origin[i])
        d.size[i] = float(size[i])
    d.jitter_fraction = float(jitter)
    d.seed = int(seed) & 0xFFFFFFFF
    d.pin_axis = int(pin_axis)
    d.material_model = int(model)
    d.density = float(density)
    d.poisson_ratio = float(poisson)
    d.young_modulus[0] = float(young[0])
    d.young_modulus[1] = float(young[1])
    d.layer_axis = int(layer_axis)
    d.layer_cells = int(layer_cells)
    d.cubature_k = int(cubature_k)
    d.pack_rest_stiffness = 1 if rest_stiffness else 0
    return d


def _from_handle(handle) -> GpuInput:
    s = N.SceneInput()
    N.check(N.lib.jgs2_scene_get_input(handle, C.byref(s)))
    V, T, K = s.vertex_count, s.tet_count, s.cubature_k

    def f(ptr, n, optional=False):
        if n == 0 or not ptr:
            return None if optional else np.zeros(0, dtype=np.float32)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    cloth = None
    if s.cloth:
        c = s.cloth.contents
        NT, NH = c.triangle_count, c.hinge_count
        cloth = ClothData(
            triangle_indices=f(c.triangle_indices, 3 * NT),
            triangle_inverse_rest_bases=f(c.triangle_inverse_rest_bases, 4 * NT),
            triangle_rest_areas=f(c.triangle_rest_areas, NT),
            hinge_indices=f(c.hinge_indices, 4 * NH),
            hinge_rest_angles=f(c.hinge_rest_angles, NH),
            hinge_rest_edge_lengths=f(c.hinge_rest_edge_lengths, NH),
            young_modulus=c.young_modulus, poisson_ratio=c.poisson_ratio, thickness=c.thickness,
            bending_stiffness=c.bending_stiffness)
    contact = None
    if s.contact:
        c = s.contact.contents
        contact = ContactData(packed_indices=f(c.packed_indices, 4 * (c.vertex_triangle_count + c.edge_edge_count)),
                              vertex_triangle_count=c.vertex_triangle_count, edge_edge_count=c.edge_edge_count)
    tri_p, tri_n, edge_p, edge_n = N.u32p(), C.c_uint32(), N.u32p(), C.c_uint32()
    N.check(N.lib.jgs2_scene_get_surface(handle, C.byref(tri_p), C.byref(tri_n), C.byref(edge_p), C.byref(edge_n)))
    surface = {"triangles": f(tri_p, 3 * tri_n.value), "edges": f(edge_p, 2 * edge_n.value)}
    return GpuInput(
        vertex_count=V, tet_count=T, cubature_k=K,
        positions=f(s.positions, 4 * V),
        vertex_rest=f(s.vertex_rest, 4 * V),
        vertex_info=f(s.vertex_info, 4 * V),
        tet_indices=f(s.tet_indices, 4 * T),
        tet_inverse_dm=f(s.tet_inverse_dm, 12 * T),
        tet_meta=f(s.tet_meta, 4 * T),
        adjacency=f(s.adjacency, int(s.adjacency_length)),
        cubature_tet_ids=f(s.cubature_tet_ids, V * K),
        cubature_weights=f(s.cubature_weights, V * K),
        cubature_basis=f(s.cubature_basis, V * K * 36),
        tet_rest_stiffness=f(s.tet_rest_stiffness, 144 * T, optional=True),
        cloth=cloth,
        contact=contact,
        extras={"surface": surface},
    )


def build_cuboid(desc: N.CuboidDesc) -> GpuInput:
    """Run the native scene builder and copy its arrays into a GpuInput."""
    handle = C.c_void_p()
    N.check(N.lib.jgs2_scene_build_cuboid(C.byref(desc), C.byref(handle)))
    try:
        return _from_handle(handle)
    finally:
        N.lib.jgs2_scene_destroy(handle)


def from_mesh(positions, tetrahedra, materials, material_ids=None, fixed=None, body_ids=None, cubature_k=0,
              cubature_length=0.0, rest_stiffness=False) -> GpuInput:
    """Pack a caller tetrahedral mesh like packJGS2GpuInput
    (src/scenes/index.ts:975-1109). `materials` is a list of (density,
    young, poisson, model) with model N.MATERIAL_*."""
    x = np.ascontiguousarray(positions, np.float64).reshape(-1)
    t = np.ascontiguousarray(tetrahedra, np.uint32).reshape(-1)
    V, T = x.size // 3, t.size // 4
    mats = (N.PrecomputeMaterial * len(materials))()
    for i, q in enumerate(materials):
        mats[i].density, mats[i].young_modulus, mats[i].poisson_ratio = float(q[0]), float(q[1]), float(q[2])
        mats[i].model = int(q[3]) if len(q) > 3 else N.MATERIAL_STABLE_NEO_HOOKEAN
    keep = [x, t]
    d = N.MeshDesc()
    d.vertex_count, d.tet_count = V, T
    d.positions = x.ctypes.data_as(C.POINTER(C.c_double))
    d.tetrahedra = t.ctypes.data_as(C.POINTER(C.c_uint32))
    if material_ids is not None:
        m = np.ascontiguousarray(material_ids, np.uint32)
        keep.append(m)
        d.material_ids = m.ctypes.data_as(C.POINTER(C.c_uint32))
    if fixed is not None:
        f = np.ascontiguousarray(fixed, np.uint8)
        keep.append(f)
        d.fixed = f.ctypes.data_as(C.POINTER(C.c_uint8))
    if body_ids is not None:
        b = np.ascontiguousarray(body_ids, np.uint32)
        keep.append(b)
        d.body_ids = b.ctypes.data_as(C.POINTER(C.c_uint32))
    d.material_count = len(materials)
    d.materials = mats
    d.cubature_k = int(cubature_k)
    d.cubature_length = float(cubature_length)
    d.pack_rest_stiffness = 1 if rest_stiffness else 0
    handle = C.c_void_p()
    N.check(N.lib.jgs2_scene_build_mesh(C.byref(d), C.byref(handle)))
    try:
        return _from_handle(handle)
    finally:
        N.lib.jgs2_scene_destroy(handle)


def trough(box_count=10, box_cells=1, seed=0, cubature_k=6, candidates=True, broad_phase=False,
           settings=None, reach_margin=0.0, candidate_capacity=0) -> GpuInput:
    """The reference's IPC trough (troughDefinition, src/scenes/index.ts:
    551-631) with its static VT/EE candidate superset; box_cells > 1 and a
    nonzero seed give the finer boxes and seeded placements of the large
    contact workload. `broad_phase` replaces the superset by the device
    broad phase over the boundary surface, its capacity sized for
    `settings` (the trough's by default) unless given."""
    d = N.TroughDesc()
    d.box_count, d.box_cells, d.seed, d.cubature_k = int(box_count), int(box_cells), int(seed), int(cubature_k)
    d.build_candidates = 1 if (candidates and not broad_phase) else 0
    handle = C.c_void_p()
    N.check(N.lib.jgs2_scene_build_trough(C.byref(d), C.byref(handle)))
    try:
        x = _from_handle(handle)
    finally:
        N.lib.jgs2_scene_destroy(handle)
    if broad_phase:
        surf = x.extras["surface"]
        x.contact = BroadPhaseContact(surf["triangles"], surf["edges"], int(candidate_capacity), float(reach_margin))
        if not candidate_capacity:
            x.contact.candidate_capacity = estimate_contact_capacity(x, settings or TROUGH_SETTINGS)
    return x


def static_contact_candidates(vertex_count, triangles, edges) -> ContactData:
    """buildStaticIpcContactCandidates (src/simulation/cpu/ipc-contact.ts:600-665)."""
    tri = np.ascontiguousarray(triangles, np.uint32).reshape(-1)
    edg = np.ascontiguousarray(edges, np.uint32).reshape(-1)
    vt, ee = C.c_uint32(), C.c_uint32()
    N.check(N.lib.jgs2_build_static_contact_candidates(vertex_count, N.uptr(tri), tri.size // 3, N.uptr(edg),
                                                        edg.size // 2, None, 0, C.byref(vt), C.byref(ee)))
    packed = np.empty(4 * (vt.value + ee.value), np.uint32)
    N.check(N.lib.jgs2_build_static_contact_candidates(vertex_count, N.uptr(tri), tri.size // 3, N.uptr(edg),
                                                        edg.size // 2, N.uptr(packed), vt.value + ee.value,
                                                        C.byref(vt), C.byref(ee)))
    return ContactData(packed_indices=packed, vertex_triangle_count=vt.value, edge_edge_count=ee.value)


# The reference app's IPC settings for its contact scenes (src/App.tsx:
# 529-574), in exact-iteration form.
TROUGH_SETTINGS = {
    "timestep": 1.0 / 60.0,
    "gravity": (0.0, -9.81, 0.0),
    "floor_height": -0.18,
    "floor_stiffness": 0.0,
    "velocity_damping": 0.997,
    "contact_tangential_damping": 0.0,
    "horizontal_body_correction": False,
    "ipc_activation_distance": 0.08,
    "ipc_minimum_distance": 0.003,
    "ipc_barrier_stiffness": 100000.0,
    "ipc_friction_coefficient": 0.45,
    "ipc_friction_velocity_epsilon": 0.05,
    "ipc_step_safety": 0.9,
    "max_step": 0.075,
    "schedule": "jacobi",
}
TROUGH_ITERATIONS = 3

# The large contact workload: twenty 8x8x8-cell boxes (3,072 tetrahedra
# each) dropped at seeded offsets and yaw angles into the pinned V-trough;
# barrier activation 1 mm (minimum distance 0.1 mm, below it as the
# reference requires), stiffness 1e5, lagged friction 0.3, 12 Jacobi
# iterations per frame, the trough's step bound; candidates from the device
# broad phase (the static superset of this scene has about 5e8 pairs).
CONTACT_BOXES = {"box_count": 20, "box_cells": 8, "seed": 20251203}
CONTACT_BOXES_SETTINGS = dict(TROUGH_SETTINGS, ipc_activation_distance=1e-3, ipc_minimum_distance=1e-4,
                              ipc_barrier_stiffness=1e5, ipc_friction_coefficient=0.3)
CONTACT_BOXES_ITERATIONS = 12


def contact_boxes(seed=CONTACT_BOXES["seed"], box_count=CONTACT_BOXES["box_count"],
                  box_cells=CONTACT_BOXES["box_cells"], candidate_capacity=0, reach_margin=0.0) -> GpuInput:
    """The large contact workload (see CONTACT_BOXES), with broad-phase contact.
    Raises JGS2RangeError for a seed whose placement leaves two boxes closer
    than the minimum distance (rotated boxes in neighboring slots can)."""
    return trough(box_count=box_count, box_cells=box_cells, seed=seed, broad_phase=True,
                  settings=CONTACT_BOXES_SETTINGS, candidate_capacity=candidate_capacity, reach_margin=reach_margin)


def contact_boxes_pool(count, first_seed=1, **kw):
    """`count` feasible contact-box scenes from consecutive seeds, skipping
    seeds whose initial pose is infeasible. Returns (scenes, seeds)."""
    out, seeds, seed = [], [], first_seed
    while len(out) < count:
        try:
            out.append(contact_boxes(seed=seed, **kw))
            seeds.append(seed)
        except N.JGS2RangeError:
            pass
        seed += 1
    return out, seeds


def cloth_drape(cells=4, origin=(-0.8, 1.15, -0.8), size=(1.6, 1.6), fold_column=3, fold_lift=0.079,
                collider_origin=(-0.65, 0.0, -0.45), collider_size=(0.9, 0.5, 0.9), density=480.0,
                young=9000.0, poisson=0.3, thickness=0.012, bending=0.045, cubature_k=4, contact=None,
                settings=None, candidate_capacity=0, reach_margin=0.0) -> GpuInput:
    """The pinned cloth drape (clothDefinition, src/scenes/index.ts:632-705);
    the defaults are the reference scene's 4x4 pre-folded sheet. Its contact
    surface is the collider's boundary plus the cloth triangles and edges
    (surfaceWithCloth, src/simulation/cpu/precompute.ts:98-133).
    `contact`: None (no IPC candidates), "superset" (the reference's static
    VT/EE superset of that surface, src/scenes/index.ts:1098-1101) or
    "broad_phase" (the device broad phase over it, capacity sized for
    `settings`, CLOTH_IPC_SETTINGS by default)."""
    d = N.ClothDesc()
    d.cells = int(cells)
    for i in range(3):
        d.origin[i] = float(origin[i])
        d.collider_origin[i] = float(collider_origin[i])
        d.collider_size[i] = float(collider_size[i])
    d.size[0], d.size[1] = float(size[0]), float(size[1])
    d.fold_column = int(fold_column)
    d.fold_lift = float(fold_lift)
    d.density = float(density)
    d.young_modulus = float(young)
    d.poisson_ratio = float(poisson)
    d.thickness = float(thickness)
    d.bending_stiffness = float(bending)
    d.cubature_k = int(cubature_k)
    handle = C.c_void_p()
    N.check(N.lib.jgs2_scene_build_cloth_drape(C.byref(d), C.byref(handle)))
    try:
        scene = _from_handle(handle)
    finally:
        N.lib.jgs2_scene_destroy(handle)
    surf = scene.extras["surface"]
    if contact == "superset":
        scene.contact = static_contact_candidates(scene.vertex_count, surf["triangles"], surf["edges"])
    elif contact == "broad_phase":
        scene.contact = BroadPhaseContact(surf["triangles"], surf["edges"], int(candidate_capacity),
                                          float(reach_margin))
        if not candidate_capacity:
            # Four times the pairs within the first build's reach (a sheet
            # folds and drapes onto itself), capped by the superset's size
            # bound V F + E (E - 1) / 2, which no list can exceed.
            first = (estimate_contact_capacity(scene, settings or CLOTH_IPC_SETTINGS) - 1024) // 2
            nf, ne = len(surf["triangles"]) // 3, len(surf["edges"]) // 2
            bound = scene.vertex_count * nf + ne * (ne - 1) // 2
            scene.contact.candidate_capacity = int(min(bound, 4 * first + 4096))
    elif contact is not None:
        raise ValueError("contact must be None, 'superset' or 'broad_phase'")
    side = int(cells) + 1
    scene.extras["cloth_offset"] = scene.vertex_count - side * side
    scene.extras["landmark_vertex"] = scene.extras["cloth_offset"] + 2 + 2 * side  # cloth centre of the 4x4 sheet
    return scene


def cantilever_beam(cells=(16, 4, 4), size=(4.0, 1.0, 1.0), cubature_k=6, model=N.MATERIAL_STABLE_NEO_HOOKEAN,
                    young=1e5, poisson=0.4) -> GpuInput:
    scene = build_cuboid(_desc(cells, (0.0, 0.0, 0.0), size, pin_axis=0, model=model, density=1000.0,
                               poisson=poisson, young=(young, young), cubature_k=cubature_k,
                               rest_stiffness=model == N.MATERIAL_COROTATED_LINEAR))
    scene.extras["target_vertex"] = int(cells[0])  # grid vertex (nx, 0, 0): the free tip corner
    return scene


def scripted_target_at_frame(scene: GpuInput, frame: int):
    """minimalScriptedTargetAtFrame (src/scenes/index.ts:269-328) around the
    tip's rest position: (active, position, stiffness)."""
    if frame < 0:
        raise ValueError("Minimal scripted-target frame must be nonnegative.")
    vertex = scene.extras["target_vertex"]
    rest = scene.vertex_rest.reshape(-1, 4)[vertex, :3].astype(np.float64)
    if frame < TARGET_PULL_START_FRAME:
        return False, rest, 0.0
    pull_frame = min(frame, TARGET_PULL_END_FRAME)
    keyframe = min(TARGET_PULL_END_FRAME,
                   TARGET_PULL_START_FRAME + ((pull_frame - TARGET_PULL_START_FRAME) // TARGET_KEYFRAME_INTERVAL)
                   * TARGET_KEYFRAME_INTERVAL)
    linear = (keyframe - TARGET_PULL_START_FRAME) / (TARGET_PULL_END_FRAME - TARGET_PULL_START_FRAME)
    progress = linear * linear * (3 - 2 * linear)
    position = rest + np.array(TARGET_PULL_OFFSET) * progress
    if frame >= TARGET_RELEASE_FRAME:
        return False, position, 0.0
    return True, position, TARGET_STIFFNESS


def layered_box(cells=(64, 64, 128), size=(1.0, 1.0, 2.0), seed=20251203, cubature_k=6,
                layer_cells=8) -> GpuInput:
    return build_cuboid(_desc(cells, (0.0, 0.0, 0.0), size, jitter=0.1, seed=seed, pin_axis=2,
                              density=1000.0, poisson=0.45, young=(1e4, 1e6), layer_axis=2,
                              layer_cells=layer_cells, cubature_k=cubature_k))


# Settings of the two single-scene workloads (exact iteration counts).
CANTILEVER_SETTINGS = {
    "timestep": 1.0 / 60.0,
    "gravity": (0.0, -9.81, 0.0),
    "floor_stiffness": 0.0,
    "parity_mode": True,
    "schedule": "jacobi",
}
CANTILEVER_FRAMES = 60
CANTILEVER_ITERATIONS = 8

LAYERED_BOX_SETTINGS = {
    "timestep": 1.0 / 60.0,
    "gravity": (0.0, -9.81, 0.0),
    "floor_stiffness": 0.0,
    "parity_mode": True,
    "schedule": "jacobi",
}
LAYERED_BOX_ITERATIONS = 8

# The reference cloth scene's settings (src/scenes/index.ts:698-704).
CLOTH_SETTINGS = {
    "timestep": 1.0 / 120.0,
    "gravity": (0.0, -9.81, 0.0),
    "floor_height": 0.0,
    "parity_mode": True,
    "schedule": "jacobi",
}
CLOTH_ITERATIONS = 5
# The reference app's settings for its cloth scene, an IPC scene
# (src/App.tsx:521-572): activation distance and maxStep 0.01, friction 0.3,
# barrier stiffness 1e5, minimum distance 0.003, damping 0.997, no floor,
# no tangential damping or body correction.
CLOTH_IPC_SETTINGS = dict(CLOTH_SETTINGS, parity_mode=False, velocity_damping=0.997, floor_stiffness=0.0,
                          contact_tangential_damping=0.0, contact_margin=0.01, horizontal_body_correction=False,
                          ipc_activation_distance=0.01, ipc_minimum_distance=0.003, ipc_barrier_stiffness=100000.0,
                          ipc_friction_coefficient=0.3, ipc_friction_velocity_epsilon=0.05, ipc_step_safety=0.9,
                          max_step=0.01, regularization=1e-6, rotation_epsilon=1e-7)


def algorithmic_bytes_per_sweep(V: int, T: int, K: int, objectives: bool = False) -> int:
    """Reference-ABI compulsory bytes of one sweep of the local solve:
    V*(16 x_S + 16 x~ + 48 VertexStatic + 48 R_v + 16 x_T + 16 update
    + 64 local diagnostics + 16 vertex energy) + T*(80 TetStatic + 16
    adjacency ids) + V*K*152 (+ 32 V with objectives)."""
    b = V * 240 + T * 96 + V * K * 152
    if objectives:
        b += 32 * V
    return b


# ---------------------------------------------------------------- mixed scenes
# One settings set for the whole batch: the reference's cloth-scene step
# (dt = 1/120 and the app's IPC parameters, src/scenes/index.ts:698-704,
# src/App.tsx:521-572), under which its StVK sheets step without reverts;
# tetrahedral scenes have no candidates, so the contact terms leave them
# unchanged.
MIXED_SETTINGS = dict(CLOTH_IPC_SETTINGS)
MIXED_ITERATIONS = 8
# Metres per hex (the cantilever's 0.25 m cell) and per cloth cell.
MIXED_HEX = 0.25
MIXED_CLOTH_CELL = 0.05
# Each sheet's last quarter of columns is folded back over the sheet at this
# gap (inside the 0.01 m activation distance, above the 0.003 m minimum), so
# same-sheet barrier rows are active from the first frame, the way the
# reference pre-folds its drape to exercise nonlocal same-sheet candidates
# (src/scenes/index.ts:653-666).
MIXED_FOLD_GAP = 0.008
# Broad-phase reach past the activation reach for the sheets (0 = the
# library default, half the activation reach).
MIXED_REACH_MARGIN = 0.0


def mixed_scene_specs(count=512, seed=20251203, cloth_fraction=0.3, tet_range=(200, 200_000),
                      cloth_cells=(32, 256), young_range=(1e4, 1e7)):
    """Seeded descriptors of the batched uneven-scene stream: tetrahedral
    scenes with log-uniform tetrahedron counts in `tet_range` (beams 4:1:1
    or cubes, six tetrahedra per hex, E log-uniform in `young_range`,
    nu uniform in [0.3, 0.45], the x = 0 face pinned) and cloth sheets of
    n x n cells with n log-uniform in `cloth_cells` (pinned at two corners
    over a one-cell fixed block far below, the last quarter folded back over
    the sheet, self-contact through the device broad phase)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        if rng.random() < cloth_fraction:
            lo, hi = np.log(cloth_cells[0]), np.log(cloth_cells[1])
            specs.append({"kind": "cloth", "cells": int(round(np.exp(rng.uniform(lo, hi))))})
            continue
        tets = float(np.exp(rng.uniform(np.log(tet_range[0]), np.log(tet_range[1]))))
        hexes = max(1.0, tets / 6.0)
        if rng.random() < 0.5:
            a = max(1, int(round((hexes / 4.0) ** (1.0 / 3.0))))
            cells = (4 * a, a, a)
        else:
            a = max(1, int(round(hexes ** (1.0 / 3.0))))
            cells = (a, a, a)
        specs.append({
            "kind": "tet", "cells": cells,
            "young": float(np.exp(rng.uniform(np.log(young_range[0]), np.log(young_range[1])))),
            "poisson": float(rng.uniform(0.3, 0.45)), "seed": int(rng.integers(1, 2**31)),
        })
    return specs


def build_mixed_scene(spec, contact="broad_phase") -> GpuInput:
    """One scene of the mixed stream. Cloth sheets carry self-contact over
    their own surface: `contact` "broad_phase" (the stream's), "superset"
    (the static superset, for the oracle) or None."""
    if spec["kind"] == "cloth":
        n = int(spec["cells"])
        side = n * MIXED_CLOTH_CELL
        return cloth_drape(cells=n, origin=(-0.5 * side, 2.0, -0.5 * side), size=(side, side),
                           fold_column=max(2, n - n // 4), fold_lift=MIXED_FOLD_GAP,
                           collider_origin=(-0.05, -50.0, -0.05), collider_size=(0.1, 0.1, 0.1), contact=contact,
                           settings=MIXED_SETTINGS, reach_margin=MIXED_REACH_MARGIN)
    cells = spec["cells"]
    size = tuple(c * MIXED_HEX for c in cells)
    return build_cuboid(_desc(cells, (0.0, 0.0, 0.0), size, jitter=0.1, seed=spec["seed"], pin_axis=0,
                              density=1000.0, poisson=spec["poisson"], young=(spec["young"], spec["young"]),
                              cubature_k=6))
