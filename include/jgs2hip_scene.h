/*
 * libjgs2hip host scene setup: the reference's CPU scene path
 * (src/scenes/index.ts buildScene/toJGS2GpuInput over src/simulation/cpu
 * mesh.ts and fem.ts) for the synthetic workloads the step is measured on.
 * Pure host code; usable without a GPU.
 */
#ifndef JGS2HIP_SCENE_H
#define JGS2HIP_SCENE_H

#include "jgs2hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jgs2_scene jgs2_scene;

/*
 * A regular cuboid split into six Kuhn tetrahedra per hexahedron
 * (generateRegularCuboidMesh, src/simulation/cpu/mesh.ts:101-200), with
 * optional seeded vertex jitter, alternating material layers and one pinned
 * face, packed like packJGS2GpuInput (src/scenes/index.ts:975-1109).
 */
typedef struct {
  uint32_t cells[3];
  double origin[3];
  double size[3];
  /* Each coordinate moves by U(-f*h_axis, f*h_axis) drawn from `seed`
   * (0 = no jitter). */
  double jitter_fraction;
  uint32_t seed;
  /* Vertices with grid index 0 along this axis are pinned (3 = none). */
  uint32_t pin_axis;
  int32_t material_model; /* JGS2_MATERIAL_* */
  double density;
  double poisson_ratio;
  /* Young's modulus of even and odd layers; layers are `layer_cells` cells
   * thick along `layer_axis` (layer_cells = 0: one material). */
  double young_modulus[2];
  uint32_t layer_axis;
  uint32_t layer_cells;
  /* Cubature records per vertex (0 = none). */
  uint32_t cubature_k;
  /* Pack the 12x12 co-rotated rest stiffness (needed by the co-rotated
   * path only; 576 bytes per tetrahedron). */
  int32_t pack_rest_stiffness;
} jgs2_cuboid_desc;

jgs2_status jgs2_scene_build_cuboid(const jgs2_cuboid_desc* desc, jgs2_scene** out);

/*
 * A caller tetrahedral mesh packed like packJGS2GpuInput
 * (src/scenes/index.ts:975-1109): rest matrices, volumes and lumped masses
 * (computeRestTetraData / lumped masses, src/simulation/cpu/fem.ts:167-238),
 * adjusted stable Neo-Hookean parameters (stable-neo-hookean.ts:68-109), the
 * canonical CSR and vertex records. Tetrahedra keep the caller's vertex
 * order. cubature_k > 0 adds the closed-form stand-in records of
 * jgs2_scene_build_cuboid with decay length `cubature_length` (0: the edge
 * of a cube of six times the mean tetrahedron volume).
 */
typedef struct {
  uint32_t vertex_count;
  uint32_t tet_count;
  const double* positions;          /* 3 per vertex */
  const uint32_t* tetrahedra;       /* 4 per tetrahedron */
  const uint32_t* material_ids;     /* 1 per tetrahedron; NULL = all 0 */
  const uint8_t* fixed;             /* 1 per vertex, nonzero = pinned; NULL = none */
  const uint32_t* body_ids;         /* 1 per vertex; NULL = all 0 */
  uint32_t material_count;
  uint32_t cubature_k;
  const jgs2_precompute_material* materials;  /* density, E, nu, model */
  double cubature_length;
  int32_t pack_rest_stiffness;
  int32_t reserved;
} jgs2_mesh_desc;

jgs2_status jgs2_scene_build_mesh(const jgs2_mesh_desc* desc, jgs2_scene** out);

/*
 * A pinned cloth drape in the shape of the reference's cloth scene
 * (clothDefinition, src/scenes/index.ts:632-705): a fixed one-cell stable
 * Neo-Hookean collider block (rho 1000, E 1.1e5, nu 0.34) and an n x n quad
 * sheet split into triangles (i00, i01, i10), (i10, i01, i11), pinned at
 * grid vertices (0, 0) and (n/2, 0), with StVK membrane and quadratic
 * dihedral bending. Columns x >= fold_column are mirrored back over column
 * fold_column - 1 and lifted by fold_lift (fold_column = 0: flat sheet).
 * Rest data follow precomputeStVKTriangleRest and the signed dihedral
 * (src/simulation/cpu/cloth.ts:115-140, :258-350), masses
 * density * thickness * area / 3 per corner (precompute.ts:46-93), hinges in
 * the half-edge order of buildJGS2GpuClothInput (src/scenes/index.ts:882-973).
 */
typedef struct {
  uint32_t cells;             /* quads per side */
  double origin[3];           /* sheet corner (x, y, z); the sheet spans x and z */
  double size[2];             /* sheet extent along x and z */
  uint32_t fold_column;
  double fold_lift;
  double collider_origin[3];
  double collider_size[3];
  double density;             /* kg / m^3 */
  double young_modulus;
  double poisson_ratio;
  double thickness;
  double bending_stiffness;
  uint32_t cubature_k;        /* records per vertex (collider vertices are pinned) */
} jgs2_cloth_desc;

jgs2_status jgs2_scene_build_cloth_drape(const jgs2_cloth_desc* desc, jgs2_scene** out);

/*
 * The reference's IPC trough (troughDefinition, src/scenes/index.ts:551-631):
 * a pinned 2x1x1 cuboid warped into a V (E 1.2e5, nu 0.34, rho 1000) and
 * `box_count` free stable Neo-Hookean boxes of 0.46 m (E 7e4, nu 0.35,
 * rho 850) in layers of four slots. box_cells = 1 and seed = 0 give the
 * reference scene exactly; box_cells = n splits each box into n^3 hexes and
 * a nonzero seed draws each box's offset within its slot and its yaw. With
 * `build_candidates` the static VT/EE superset of the boundary surface is
 * attached (buildStaticIpcContactCandidates).
 */
typedef struct {
  uint32_t box_count;        /* 1..20 */
  uint32_t box_cells;        /* hexes per box edge */
  uint32_t seed;
  uint32_t cubature_k;
  int32_t build_candidates;
} jgs2_trough_desc;

jgs2_status jgs2_scene_build_trough(const jgs2_trough_desc* desc, jgs2_scene** out);

/* The boundary surface of the scene's tetrahedra (extractBoundarySurface,
 * src/simulation/cpu/mesh.ts:366-440): outward triangles in ascending
 * sorted-key order and their unique sorted edges. Pointers stay valid until
 * jgs2_scene_destroy(). */
jgs2_status jgs2_scene_get_surface(const jgs2_scene* scene, const uint32_t** triangles, uint32_t* triangle_count,
                                   const uint32_t** edges, uint32_t* edge_count);

/* The static IPC candidate superset (buildStaticIpcContactCandidates,
 * src/simulation/cpu/ipc-contact.ts:600-665): vertex-triangle pairs over all
 * vertices and canonical triangles, then edge-edge pairs, omitting pairs in
 * a closed one-ring. Counts are always returned; `packed` (4 indices per
 * candidate, `capacity` candidates) may be NULL to query them. */
jgs2_status jgs2_build_static_contact_candidates(uint32_t vertex_count, const uint32_t* triangles,
                                                 uint32_t triangle_count, const uint32_t* edges, uint32_t edge_count,
                                                 uint32_t* packed, uint64_t capacity, uint32_t* vt_count,
                                                 uint32_t* ee_count);

/* Pointers into the scene's arrays, valid until jgs2_scene_destroy(). */
jgs2_status jgs2_scene_get_input(const jgs2_scene* scene, jgs2_scene_input* out);

/* Attach per-vertex objectives (3 floats per vertex forces, 3 per vertex
 * targets, 1 per vertex stiffness; any may be NULL). */
jgs2_status jgs2_scene_set_objectives(jgs2_scene* scene, const float* forces, const float* targets,
                                      const float* stiffnesses);

void jgs2_scene_destroy(jgs2_scene* scene);

/* Canonical vertex->tetrahedron CSR (buildAdjacency, src/scenes/index.ts:
 * 811-837). `starts`/`counts` receive V entries, `tets` 4T entries. */
jgs2_status jgs2_build_adjacency(uint32_t vertex_count, uint32_t tet_count, const uint32_t* tet_indices,
                                 uint32_t* starts, uint32_t* counts, uint32_t* tets);

/* Vertex -> element incidence CSR of triangles (width 3) or hinges (width
 * 4) (buildIncidentElementCsr, src/simulation/gpu/cloth-layout.ts:183-207):
 * `rows` receives vertex_count + 1 offsets, `elements` width * element_count
 * element ids, ascending within each row. */
jgs2_status jgs2_build_element_incidence(uint32_t vertex_count, const uint32_t* element_indices, uint32_t width,
                                         uint32_t element_count, uint32_t* rows, uint32_t* elements);

/* Deterministic greedy coloring over CSR conflict groups
 * (buildGreedyVertexColoring, src/simulation/cpu/vertex-coloring.ts:20-110).
 * Returns the color count through `color_count`. */
jgs2_status jgs2_greedy_vertex_coloring(uint32_t vertex_count, const uint32_t* group_offsets, uint32_t group_count,
                                        const uint32_t* group_members, uint32_t* colors, uint32_t* color_count);

#ifdef __cplusplus
}
#endif

#endif /* JGS2HIP_SCENE_H */
