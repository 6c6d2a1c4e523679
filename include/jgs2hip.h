/*
 * libjgs2hip -- MI355X (gfx950) implementation of the JGS2 per-frame step.
 *
 * This C ABI is the drop-in boundary for the reference's WebGPU solver class
 * `JGS2GpuSolver` (src/simulation/gpu/jgs2-solver.ts:905-3051). Each entry
 * point names the reference member it replaces. Plain pointers and sizes
 * only; every function returns a jgs2_status and never throws across the
 * boundary. The calling thread's last error message is available from
 * jgs2_last_error().
 *
 * Ownership: jgs2_create() deep-copies every input array into device memory
 * the handle owns (jgs2-solver.ts:389-408). Read functions copy into
 * caller-owned host buffers (jgs2-solver.ts:2996-3027). jgs2_destroy() frees
 * everything; any call on a destroyed handle is an error
 * (jgs2-solver.ts:3046-3050).
 *
 * Threading: a handle owns one HIP stream. Step and objective calls only
 * enqueue work; reads and jgs2_sync() wait for it. Calls on one handle must
 * not overlap from several threads.
 */
#ifndef JGS2HIP_H
#define JGS2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference throws RangeError/TypeError/Error
 * synchronously before any GPU work (layout.ts:223-416,
 * jgs2-solver.ts:429-529) and rejects asynchronously on an infeasible
 * source pose (jgs2-solver.ts:873-889); the codes mirror those classes. */
typedef enum {
  JGS2_OK = 0,
  JGS2_ERROR_RANGE = 1,      /* RangeError: invalid sizes, values, limits */
  JGS2_ERROR_TYPE = 2,       /* TypeError: malformed arguments, NULL */
  JGS2_ERROR_STATE = 3,      /* use after destroy, unsupported mode */
  JGS2_ERROR_DEVICE = 4,     /* HIP runtime failure */
  JGS2_ERROR_INFEASIBLE = 5, /* initial pose fails the f32 feasibility gate */
  JGS2_ERROR_MEMORY = 6      /* device or host allocation failed */
} jgs2_status;

/* Material tags stored in tet_meta[4*t+3] (layout.ts:28-29). */
#define JGS2_MATERIAL_COROTATED_LINEAR 0
#define JGS2_MATERIAL_STABLE_NEO_HOOKEAN 1

/* Record widths of the reference ABI (layout.ts:16-27). */
#define JGS2_CUBATURE_BASIS_FLOATS 36
#define JGS2_REST_STIFFNESS_FLOATS 144
#define JGS2_GLOBALIZATION_HISTORY_CAPACITY 64
#define JGS2_MAX_BATCH_FRAMES 240
#define JGS2_MAX_GLOBALIZED_ITERATIONS_PER_SUBMISSION 2048

/*
 * Optional triangle cloth layer sharing the vertex state: `JGS2GpuClothInput`
 * (src/simulation/gpu/cloth-layout.ts:89-104). Hinges list the oriented
 * triangles (0,1,2) and (1,0,3) of each interior edge.
 */
typedef struct {
  uint32_t triangle_count;
  const uint32_t* triangle_indices;          /* 3 per triangle */
  const float* triangle_inverse_rest_bases;  /* 4 per triangle, row-major */
  const float* triangle_rest_areas;          /* 1 per triangle */
  uint32_t hinge_count;
  const uint32_t* hinge_indices;             /* 4 per hinge */
  const float* hinge_rest_angles;            /* 1 per hinge */
  const float* hinge_rest_edge_lengths;      /* 1 per hinge */
  double young_modulus;
  double poisson_ratio;
  double thickness;
  double bending_stiffness;
} jgs2_cloth_input;

/*
 * Optional static IPC contact candidates: `StaticIpcContactCandidates`
 * (src/simulation/cpu/ipc-contact.ts:20-38) in the order packed by
 * packIpcContactBuffer (src/simulation/gpu/ipc-contact-layout.ts:271-342):
 * vertex-triangle records (point, t0, t1, t2) first, then edge-edge records
 * (a0, a1, b0, b1), four distinct vertices each.
 */
typedef struct {
  uint32_t vertex_triangle_count;
  uint32_t edge_edge_count;
  const uint32_t* packed_indices;  /* 4 per candidate */
  /*
   * Broad-phase contact (no counterpart in the reference, which enumerates
   * the whole superset on the host, ipc-contact.ts:600-665): with no packed
   * candidates (counts 0, packed_indices NULL) and a surface given, the
   * candidates are the superset's pairs (same canonical order and one-ring
   * exclusions) within AABB reach, built on the device at admission and
   * rebuilt at the frame start or before an iteration when motion since the
   * last build could bring an unlisted pair within the activation reach.
   * Stepping gives the superset's results bit for bit. Jacobi schedule only.
   */
  uint32_t surface_triangle_count;
  uint32_t surface_edge_count;
  const uint32_t* surface_triangles;  /* 3 per triangle (canonicalized like canonicalTriangles, :500-527) */
  const uint32_t* surface_edges;      /* 2 per edge; triangle edges are added (canonicalEdges, :546-569) */
  uint64_t candidate_capacity;        /* resident candidates at most; 0 = twice the count at the initial pose + 1024 */
  double reach_margin;                /* reach past the activation reach in m; 0 = half of it */
} jgs2_contact_input;

/*
 * Scene input: the arrays of `JGS2GpuInput` (layout.ts:41-90), same element
 * order and meaning. Optional arrays may be NULL.
 */
typedef struct {
  uint32_t vertex_count;
  uint32_t tet_count;
  uint32_t cubature_k;
  const float* positions;          /* 4 per vertex, xyz position, w = 1 */
  const float* velocities;         /* optional, 4 per vertex */
  const float* vertex_rest;        /* 4 per vertex: rest xyz, mass */
  const uint32_t* vertex_info;     /* 4 per vertex: adjStart, adjCount, pinned, body */
  const uint32_t* tet_indices;     /* 4 per tet */
  const float* tet_inverse_dm;     /* 12 per tet: three padded columns of Dm^-1 */
  const float* tet_meta;           /* 4 per tet: volume, lambda, mu, material tag */
  const float* tet_rest_stiffness; /* optional for Neo-Hookean scenes; 144 per tet row-major */
  const uint32_t* adjacency;       /* canonical vertex->tet CSR payload */
  uint64_t adjacency_length;
  const uint32_t* cubature_tet_ids; /* vertex_count * cubature_k, 0xffffffff = empty */
  const float* cubature_weights;    /* vertex_count * cubature_k */
  const float* cubature_basis;      /* vertex_count * cubature_k * 36, row-major Ubar blocks */
  const float* external_forces;     /* optional, 3 per vertex */
  const float* target_positions;    /* optional, 3 per vertex (with target_stiffnesses) */
  const float* target_stiffnesses;  /* optional, 1 per vertex */
  const jgs2_cloth_input* cloth;    /* optional triangle cloth layer */
  const jgs2_contact_input* contact; /* optional IPC candidates (contactCandidates) */
} jgs2_scene_input;

/* Schedules (jgs2-solver.ts:62). */
#define JGS2_SCHEDULE_JACOBI 0
#define JGS2_SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL 1

/*
 * `JGS2StepSettings` (jgs2-solver.ts:64-100). Values are doubles on the
 * boundary and rounded to f32 when packed, as the reference's uniform
 * packer does (jgs2-solver.ts:604-730).
 */
typedef struct {
  int32_t schedule;
  double timestep;
  double gravity[3];
  int32_t iterations;
  double floor_height;
  double floor_stiffness;
  double velocity_damping;
  double regularization;
  double rotation_epsilon;
  double max_step;
  double residual_tolerance;
  double normalized_update_tolerance;
  double contact_tangential_damping;
  double contact_margin;
  double ipc_activation_distance;
  double ipc_minimum_distance;
  double ipc_barrier_stiffness;
  double ipc_friction_coefficient;
  double ipc_friction_velocity_epsilon;
  double ipc_step_safety;
  int32_t horizontal_body_correction;
  int32_t parity_mode;
} jgs2_settings;

/* Override mask bits: a set bit means "use this field of the passed struct",
 * the equivalent of a key present in the reference's Partial<> overrides. */
enum {
  JGS2_SET_SCHEDULE = 1u << 0,
  JGS2_SET_TIMESTEP = 1u << 1,
  JGS2_SET_GRAVITY = 1u << 2,
  JGS2_SET_ITERATIONS = 1u << 3,
  JGS2_SET_FLOOR_HEIGHT = 1u << 4,
  JGS2_SET_FLOOR_STIFFNESS = 1u << 5,
  JGS2_SET_VELOCITY_DAMPING = 1u << 6,
  JGS2_SET_REGULARIZATION = 1u << 7,
  JGS2_SET_ROTATION_EPSILON = 1u << 8,
  JGS2_SET_MAX_STEP = 1u << 9,
  JGS2_SET_RESIDUAL_TOLERANCE = 1u << 10,
  JGS2_SET_NORMALIZED_UPDATE_TOLERANCE = 1u << 11,
  JGS2_SET_CONTACT_TANGENTIAL_DAMPING = 1u << 12,
  JGS2_SET_CONTACT_MARGIN = 1u << 13,
  JGS2_SET_IPC_ACTIVATION_DISTANCE = 1u << 14,
  JGS2_SET_IPC_MINIMUM_DISTANCE = 1u << 15,
  JGS2_SET_IPC_BARRIER_STIFFNESS = 1u << 16,
  JGS2_SET_IPC_FRICTION_COEFFICIENT = 1u << 17,
  JGS2_SET_IPC_FRICTION_VELOCITY_EPSILON = 1u << 18,
  JGS2_SET_IPC_STEP_SAFETY = 1u << 19,
  JGS2_SET_HORIZONTAL_BODY_CORRECTION = 1u << 20,
  JGS2_SET_PARITY_MODE = 1u << 21
};

typedef struct jgs2_solver jgs2_solver;

/* DEFAULT_JGS2_STEP_SETTINGS (jgs2-solver.ts:164-187). */
void jgs2_default_settings(jgs2_settings* out);

/* Message of the calling thread's last failed call ("" after success). */
size_t jgs2_last_error(char* buffer, size_t capacity);

/* Number of visible HIP devices (0 without a GPU). */
int jgs2_device_count(void);

/* JGS2GpuSolver.create (jgs2-solver.ts:971-1502): validate, color, pack,
 * upload and run the f32 source-feasibility preflight. `overrides` may be
 * NULL (mask ignored). */
jgs2_status jgs2_create(const jgs2_scene_input* input, const jgs2_settings* overrides, uint64_t override_mask,
                        jgs2_solver** out);

/* JGS2GpuSolver.destroy (jgs2-solver.ts:3029-3044). NULL is a no-op. */
jgs2_status jgs2_destroy(jgs2_solver* solver);

/* step / stepFrames (jgs2-solver.ts:1767-1812): odd-normalized iteration
 * count and the convergence stop latch. */
jgs2_status jgs2_step(jgs2_solver* solver, const jgs2_settings* overrides, uint64_t override_mask);
jgs2_status jgs2_step_frames(jgs2_solver* solver, uint32_t frame_count, const jgs2_settings* overrides,
                             uint64_t override_mask);

/* stepExactIterations / stepFramesExactIterations (jgs2-solver.ts:1814-1862):
 * exactly `iterations` sweeps per frame, even counts kept, no stop latch. */
jgs2_status jgs2_step_exact_iterations(jgs2_solver* solver, uint32_t iterations, const jgs2_settings* overrides,
                                       uint64_t override_mask);
jgs2_status jgs2_step_frames_exact_iterations(jgs2_solver* solver, uint32_t frame_count, uint32_t iterations,
                                              const jgs2_settings* overrides, uint64_t override_mask);

/* Objective mutation API (jgs2-solver.ts:1576-1760). */
jgs2_status jgs2_set_external_force(jgs2_solver* solver, uint32_t vertex, const float force[3]);
jgs2_status jgs2_replace_external_forces(jgs2_solver* solver, const float* forces, size_t count);
jgs2_status jgs2_clear_external_forces(jgs2_solver* solver);
jgs2_status jgs2_set_quadratic_target(jgs2_solver* solver, uint32_t vertex, const float target[3],
                                      float stiffness);
jgs2_status jgs2_replace_quadratic_targets(jgs2_solver* solver, const float* positions, size_t position_count,
                                           const float* stiffnesses, size_t stiffness_count);
jgs2_status jgs2_release_quadratic_target(jgs2_solver* solver, uint32_t vertex);
jgs2_status jgs2_release_all_quadratic_targets(jgs2_solver* solver);

/* Reads (jgs2-solver.ts:2501-2604). `count` is the float capacity of `dst`;
 * it must equal the exact size stated. All wait for queued work. */
jgs2_status jgs2_read_positions(jgs2_solver* solver, float* dst, size_t count);        /* 4 * V */
jgs2_status jgs2_read_velocities(jgs2_solver* solver, float* dst, size_t count);       /* 4 * V */
jgs2_status jgs2_read_predicted_positions(jgs2_solver* solver, float* dst, size_t count); /* 4 * V */
jgs2_status jgs2_read_final_iteration_updates(jgs2_solver* solver, float* dst, size_t count); /* 4 * V */
jgs2_status jgs2_read_vertex_rotations(jgs2_solver* solver, float* dst, size_t count); /* 12 * V */
jgs2_status jgs2_read_local_diagnostics(jgs2_solver* solver, float* dst, size_t count); /* 16 * V */
jgs2_status jgs2_read_globalization_control(jgs2_solver* solver, float* dst, size_t count); /* 16 */
jgs2_status jgs2_read_globalization_history(jgs2_solver* solver, float* dst, size_t count); /* 64 * 32 */

/* awaitIdle (jgs2-solver.ts:2496-2499). Also reports JGS2_ERROR_MEMORY
 * when a broad-phase build needed more candidates (or vertex row entries)
 * than the scene's capacity; results from that build on are invalid. */
jgs2_status jgs2_sync(jgs2_solver* solver);

/*
 * The current IPC candidate list (no counterpart in the reference): the
 * packed records in caller vertex ids and canonical order, VT first. For a
 * broad-phase scene it is the list of the last build, whose reach is
 * returned; for a static superset the reach is +inf. `capacity` is in
 * candidates; `total` receives the count even when it exceeds capacity.
 */
jgs2_status jgs2_read_contact_candidates(jgs2_solver* solver, uint32_t* packed, uint64_t capacity,
                                         uint32_t* vertex_triangle_count, uint64_t* total, float* reach);
/* Contact list facts of a scene (no counterpart in the reference). */
typedef struct {
  uint64_t candidates;          /* current list length (superset: its size) */
  uint64_t vertex_triangle;     /* of which vertex-triangle */
  uint64_t capacity;            /* candidate slots the scene holds */
  uint64_t builds;              /* broad-phase builds so far, admission included */
  float reach;                  /* reach of the last build (+inf for a static superset) */
  uint32_t broad_phase;         /* 1 for a broad-phase scene */
} jgs2_contact_info;
jgs2_status jgs2_get_contact_info(jgs2_solver* solver, jgs2_contact_info* out);
/* Rebuilds a broad-phase scene's list at its current positions now (no
 * counterpart in the reference; stepping results do not change). */
jgs2_status jgs2_rebuild_contact_candidates(jgs2_solver* solver);
/* The candidate capacity jgs2_create would choose for a broad-phase scene
 * whose candidate_capacity is 0 (no counterpart in the reference; for
 * sizing batch arenas). Validates the input like jgs2_create; 0 when the
 * scene has no contact. Host only. */
jgs2_status jgs2_estimate_contact_capacity(const jgs2_scene_input* input, const jgs2_settings* overrides,
                                           uint64_t override_mask, uint64_t* capacity);

/* Scene facts the reference exposes as readonly members
 * (jgs2-solver.ts:906-916). */
typedef struct {
  uint32_t vertex_count;
  uint32_t tet_count;
  uint32_t cubature_k;
  uint32_t body_count;
  uint32_t schedule_color_count;
  uint32_t globalization_enabled;
  uint32_t free_vertex_count;
  float scene_scale;
} jgs2_solver_info;
jgs2_status jgs2_get_info(jgs2_solver* solver, jgs2_solver_info* out);

/* Greedy coloring of the schedule's conflict graph
 * (buildJGS2InputVertexColoring, jgs2-solver.ts:308-367, over
 * buildGreedyVertexColoring, src/simulation/cpu/vertex-coloring.ts:20-110).
 * `colors` receives V entries. */
jgs2_status jgs2_read_schedule_colors(jgs2_solver* solver, uint32_t* colors, size_t count);

/*
 * Test hook in the spirit of readOracleDiagnostics (jgs2-solver.ts:2882-2907):
 * evaluates the local solve's gradient (3 per vertex) and Hessian (9 per
 * vertex, row-major) at the current accepted pose with vertex frames
 * refreshed there, using the given settings. Pinned vertices report zeros.
 */
jgs2_status jgs2_debug_local_systems(jgs2_solver* solver, const jgs2_settings* overrides, uint64_t override_mask,
                                     float* gradients, size_t gradient_count, float* hessians,
                                     size_t hessian_count);

/* Average device time per launch of each kernel over the last
 * `jgs2_profile_*` window, measured with HIP events on the solver's stream.
 * Used by the benchmark to report the local-solve kernel's bandwidth. */
jgs2_status jgs2_profile_begin(jgs2_solver* solver);
jgs2_status jgs2_profile_end(jgs2_solver* solver, double* local_solve_ms_per_launch, uint64_t* local_solve_launches,
                             double* frame_ms_total);

/*
 * Batches. No counterpart in the reference, whose JGS2GpuSolver steps one
 * scene per object; these entry points serve callers that step many
 * independent scenes of uneven size at once (continuous-batching servers).
 *
 * A batch is one device arena created at fixed capacities. Scenes are
 * admitted into free 256-entry blocks of its vertex, tetrahedron, triangle
 * and candidate ranges and retired at any time between submissions. Each
 * pass of a frame is one kernel launch over every resident scene;
 * reductions, the line-search control, history and the convergence latch
 * stay per scene, so every scene's results are bit-identical to stepping it
 * alone. Admitted scenes are jgs2_solver handles: every read, objective,
 * info and debug call above applies to them; jgs2_step* does not (they are
 * stepped by jgs2_batch_step_frames*), and jgs2_destroy() retires them.
 * jgs2_batch_destroy() frees the arena and invalidates every handle of it.
 *
 * Scenes of one batch share: the material path (stable Neo-Hookean or
 * co-rotated), the settings given at create (timestep, contact
 * parameters), per-step overrides, and one cloth material.
 */
typedef struct jgs2_batch jgs2_batch;

typedef struct {
  uint64_t max_vertices;    /* summed over resident scenes, each rounded up to 256 */
  uint64_t max_tets;        /* likewise */
  uint64_t max_triangles;   /* likewise */
  uint64_t max_hinges;
  uint64_t max_candidates;  /* IPC candidates, each scene rounded up to 256 */
  uint64_t max_bodies;
  uint32_t max_scenes;
  uint32_t cubature_k;      /* records per vertex; scenes may use fewer */
  uint32_t corotated;       /* 1 = legacy co-rotated material path */
  uint32_t reserved;
  uint64_t max_surface_triangles;  /* broad-phase surfaces, each scene rounded up to 256 */
  uint64_t max_surface_edges;      /* likewise */
} jgs2_batch_config;

typedef struct {
  uint32_t scene_count;
  uint32_t max_scenes;
  uint32_t solve_block;  /* threads per block of the local solve */
  uint32_t xcd_remap;    /* XCD-aware block order of per-vertex kernels */
  uint32_t use_graphs;
  uint32_t solve_lanes;  /* lanes of a wavefront sharing one vertex's local solve */
  uint64_t vertex_capacity;
  uint64_t vertex_blocks_used;
  uint64_t vertex_extent;  /* vertices every launch covers (highest used block) */
  uint64_t tet_extent;
  uint64_t arena_bytes;    /* device bytes the arena holds */
} jgs2_batch_info;

/* Device bytes a batch of this configuration allocates (for admission
 * against an HBM budget before creating it). */
jgs2_status jgs2_batch_arena_bytes(const jgs2_batch_config* config, uint64_t* bytes);
jgs2_status jgs2_batch_create(const jgs2_batch_config* config, const jgs2_settings* overrides,
                              uint64_t override_mask, jgs2_batch** out);
jgs2_status jgs2_batch_destroy(jgs2_batch* batch);
/* Validates and uploads a scene (the checks of jgs2_create). Returns
 * JGS2_ERROR_MEMORY, with the batch unchanged, when the arena has no room
 * for it now. */
jgs2_status jgs2_batch_admit(jgs2_batch* batch, const jgs2_scene_input* input, jgs2_solver** scene);
jgs2_status jgs2_batch_retire(jgs2_batch* batch, jgs2_solver* scene);
/* stepFrames / stepFramesExactIterations over every resident scene. */
jgs2_status jgs2_batch_step_frames(jgs2_batch* batch, uint32_t frame_count, const jgs2_settings* overrides,
                                   uint64_t override_mask);
jgs2_status jgs2_batch_step_frames_exact_iterations(jgs2_batch* batch, uint32_t frame_count, uint32_t iterations,
                                                    const jgs2_settings* overrides, uint64_t override_mask);
jgs2_status jgs2_batch_sync(jgs2_batch* batch);
jgs2_status jgs2_batch_get_info(jgs2_batch* batch, jgs2_batch_info* out);
/* Records each distinct frame once as a hipGraph and replays it. After an
 * admission, retirement, objective change or different step settings the
 * frame is captured again; when only extents and arguments changed the
 * capture updates the existing exec in place (hipGraphExecUpdate), and
 * otherwise the exec is rebuilt. */
jgs2_status jgs2_batch_use_graphs(jgs2_batch* batch, int enable);
/* How many frame graphs were instantiated from scratch and how many
 * captures updated the existing exec in place (no counterpart in the
 * reference). */
jgs2_status jgs2_batch_graph_stats(jgs2_batch* batch, uint64_t* instantiations, uint64_t* updates);
/* Launch shape of the local solve: 64, 128 or 256 threads per block, the
 * XCD-aware block order on or off, and 1, 4, 8 or 16 lanes per vertex
 * (default 4). Results do not depend on the block size or block order; the
 * lane count sets the summation order of each vertex's gathered terms, so
 * results agree across lane counts to FP32 rounding, not bit for bit. */
jgs2_status jgs2_batch_set_launch(jgs2_batch* batch, uint32_t solve_block, uint32_t xcd_remap,
                                  uint32_t solve_lanes);
/* Times the local solve over the resident scenes under every block size and
 * block order (`repeats` launches each, lanes per vertex unchanged) and
 * keeps the fastest; results are unchanged. Needs a frame stepped since the
 * last admission or retirement (JGS2_ERROR_STATE otherwise). */
jgs2_status jgs2_batch_autotune(jgs2_batch* batch, const jgs2_settings* overrides, uint64_t override_mask,
                                uint32_t repeats, double* best_ms_per_launch);
jgs2_status jgs2_batch_profile_begin(jgs2_batch* batch);
jgs2_status jgs2_batch_profile_end(jgs2_batch* batch, double* local_solve_ms_per_launch,
                                   uint64_t* local_solve_launches, double* frame_ms_total);
/* The batch a handle lives in (jgs2_create() makes a private one-scene
 * batch), for the launch, graph and autotune controls above. */
jgs2_status jgs2_solver_batch(jgs2_solver* solver, jgs2_batch** out);

/* Exact oracle diagnostics, readOracleDiagnostics / the opt-in evaluator
 * (gpu/jgs2-diagnostics.ts:61-111, evaluator :712-1110; the reference reads
 * them through jgs2-solver.ts:2620-2907), evaluated on the device in FP64
 * at the current pose against the last predicted positions, or at caller
 * poses (4 floats per vertex; NULL = the solver's own). Energy components
 * of the implicit-Euler incremental potential: inertia 1/2 m/h^2 |x - x~|^2,
 * elasticity (stable Neo-Hookean vol psi(F), or the co-rotated quadratic),
 * external force -f.x, quadratic target, analytic floor; plus, with contact
 * candidates, the barrier sum kappa b(d) (no counterpart in the reference's
 * evaluator; cloth energies are in neither). */
typedef struct jgs2_diagnostics {
  double energy;                     /* sum of the six components */
  double inertia, elasticity, external_force, quadratic_target, floor_contact, contact_barrier;
  double gradient_norm_inertia, gradient_norm_material, gradient_norm_external_force, gradient_norm_target,
      gradient_norm_floor;
  double gradient_norm;              /* |g| over free vertices */
  double residual_denominator;       /* sum of the five component norms */
  double relative_residual;          /* |g| / max(1, denominator) */
  double maximum_update;             /* largest valid final-iteration |delta| */
  double minimum_deformation_determinant;
  double minimum_floor_distance;     /* min signed y - floor over free vertices */
  double minimum_contact_distance;   /* +inf without candidates */
  uint32_t active_floor_contacts, active_vertices, active_contact_pairs;
  uint32_t finite, maximum_update_valid, minimum_determinant_valid, floor_metrics_valid, reserved;
} jgs2_diagnostics;
/* vertex_records (optional, 36 doubles per vertex, caller order): gradient
 * xyz, free flag, row-major 3x3 Hessian block, inertia energy, floor
 * energy, final update, |g|^2, finite, inertia weight, update valid,
 * |g_inertia|^2, |g_material|^2, |g_floor|^2, signed floor distance, floor
 * active, external-force gradient xyz and energy, target gradient xyz and
 * energy, |g_force|^2, |g_target|^2, 0. tet_records (optional, 4 per
 * tetrahedron): elastic energy, det F, finite, 0. */
jgs2_status jgs2_read_oracle_diagnostics(jgs2_solver* solver, const jgs2_settings* overrides, uint64_t override_mask,
                                         const float* positions, const float* predicted, jgs2_diagnostics* out,
                                         double* vertex_records, size_t vertex_record_count, double* tet_records,
                                         size_t tet_record_count);

/* No counterpart in the reference. jgs2_read_oracle_diagnostics for callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc. Same kernels and summary,
 * evaluated at the solver's own pose against its own last predicted
 * positions (no caller poses, no per-vertex or per-tetrahedron records).
 * The records of one evaluation live in a fixed 64 MiB device array of the
 * library (8,388,608 doubles: 36 per vertex, 4 per tetrahedron, 9 more per
 * co-rotated tetrahedron, 2 per candidate slot, 32 of summary); a larger
 * scene is a JGS2_ERROR_RANGE. Calls no hipMalloc or hipFree. Every call on
 * one GPU shares that array, so calls that overlap on one GPU (several
 * threads, or several solvers from several threads) race; calls from one
 * thread are serialized, since each waits for its result. */
jgs2_status jgs2_read_oracle_diagnostics_noalloc(jgs2_solver* solver, const jgs2_settings* overrides,
                                                 uint64_t override_mask, jgs2_diagnostics* out);

/* ---------------------------------------------------------------------
 * Cubature precompute: buildPrecomputedScene (src/simulation/cpu/
 * precompute.ts:218-380) with its exact vertex basis (:135-200), nonlinear
 * pose corpus (nonlinear-cubature.ts:462-620) and greedy NNLS training
 * (nonlinear-cubature-training.ts:731-843; co-rotated linear scenes:
 * cubature.ts:714-826). Produces the cubature_* arrays of jgs2_scene_input
 * in the reference's packing (src/scenes/index.ts:1054-1068). Host only.
 * ------------------------------------------------------------------- */
typedef struct jgs2_precompute_material {
  double density;
  double young_modulus;
  double poisson_ratio;
  uint32_t model;                  /* JGS2_MATERIAL_* */
  uint32_t reserved;
} jgs2_precompute_material;

typedef struct jgs2_precompute_input {
  uint32_t vertex_count;
  uint32_t tetrahedron_count;
  uint32_t material_count;
  uint32_t cubature_samples;       /* K records per vertex (settings.cubatureSamples), 1..12 */
  const double* positions;         /* 3 per vertex, rest pose (SceneDefinition.mesh.positions) */
  const uint32_t* tetrahedra;      /* 4 per tetrahedron */
  const uint8_t* fixed;            /* 1 per vertex, nonzero = pinned */
  const uint32_t* material_ids;    /* 1 per tetrahedron */
  const jgs2_precompute_material* materials;
  double timestep;
} jgs2_precompute_input;

typedef struct jgs2_precompute_options {
  /* 0: each vertex's basis over the whole mesh (the reference's definition;
   * one dense factorization, for meshes up to a few thousand vertices).
   * k > 0: over the free vertices within k edges of the source, zero outside
   * (a local solve per vertex; the cost no longer grows with the mesh). */
  uint32_t basis_rings;
  uint32_t threads;                /* 0: every hardware thread */
  /* 0: automatic (dense up to 1536 free coordinates, else CG on the GPU when
   * one is present, else on the host), 1: dense Cholesky, 2: host CG,
   * 3: GPU CG (JGS2_ERROR_DEVICE without a GPU) */
  uint32_t mode_solver;
  /* Greedy selection of the training: 0 automatic (on the GPU for stable
   * Neo-Hookean scenes of 256+ free vertices with ring bases, more than 12
   * tetrahedra and K <= 8; else on the host threads), 1 host, 2 GPU. Same
   * records either way. */
  uint32_t training;
  double cg_tolerance;             /* relative residual of the CG mode solves; 0: 1e-12 */
  /* Largest packed training residual accepted per vertex; 0: the
   * reference's 0.01 (NONLINEAR_CUBATURE_TRAINING_RESIDUAL_TOLERANCE).
   * Meshes beyond a few cells need more than K = 6 samples to meet 0.01. */
  double residual_bound;
  /* Nonzero: the reference's exact-all-element fixture packing
   * (buildPhase0ForceFreeScene, src/scenes/canonical-phase0.ts:144-205)
   * instead of training: each free vertex keeps its exact equilibrium basis
   * over the whole mesh (basis_rings must be 0) and packs every tetrahedron
   * once with weight 1, so cubature_samples must equal tetrahedron_count;
   * residuals are 0. */
  uint32_t exact_all_elements;
  uint32_t reserved;
} jgs2_precompute_options;

typedef struct jgs2_precompute_output {
  uint32_t* cubature_tet_ids;      /* V*K; 0xffffffff marks unused records */
  float* cubature_weights;         /* V*K */
  float* cubature_basis;           /* V*K*36: four row-major 3x3 rest-basis blocks per record */
  double* vertex_residuals;        /* optional, V: packed training residual (NaN when pinned) */
  uint32_t* vertex_ranks;          /* optional, 3 per vertex: nonzero candidates, column rank, packed rank */
  /* optional, 8 values: seconds for the rest system, the modes and poses,
   * the per-vertex training; GPU CG milliseconds and iterations; 1 when the
   * GPU solved the modes; GPU selection milliseconds; 1 when the GPU
   * selected */
  double* timings;
} jgs2_precompute_output;

/* Errors: JGS2_ERROR_TYPE for NULL arrays; JGS2_ERROR_RANGE for the
 * reference's input RangeErrors (bad materials, degenerate tetrahedra, mixed
 * stable Neo-Hookean / co-rotated scenes, no free vertex); JGS2_ERROR_INFEASIBLE
 * when a vertex's packed training residual exceeds the bound (the
 * reference's Error at precompute.ts:313-323). */
jgs2_status jgs2_precompute_cubature(const jgs2_precompute_input* input, const jgs2_precompute_options* options,
                                     const jgs2_precompute_output* output);

#ifdef __cplusplus
}
#endif

#endif /* JGS2HIP_H */
