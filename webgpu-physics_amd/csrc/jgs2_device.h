// Device-side state and launch interface of libjgs2hip (host view).
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>
#include <cstdint>

namespace jgs2 {

// Per-pass uniform block, the role of the reference's SimParams
// (src/simulation/gpu/jgs2-shader.ts:141-164), passed by value to kernels.
// Counts are arena extents: kernels run over every block below them and
// skip blocks that belong to no scene.
struct PassParams {
  uint32_t vertex_count, tet_count, cubature_k, body_count;
  uint32_t src, tgt;          // 0 = posA, 1 = posB
  uint32_t colored;           // in-place single-color pass
  uint32_t color_begin, color_end;  // range of the color-sorted vertex list
  uint32_t objective_flags;
  uint32_t globalization;     // stable Neo-Hookean production path
  uint32_t stop_enabled;      // convergence stop latch honored
  uint32_t corotated;         // legacy co-rotated material path
  uint32_t fused_rot;         // solve reads frames from vrot_next and adopts them
  uint32_t compute_next_rot;  // convergence pass also computes the next frames
  uint32_t tri_count, hinge_count;  // cloth layer sizes (0 = no cloth)
  float dt, inv_dt, inv_dt2, max_step;
  float gx, gy, gz, floor_y;
  float floor_k, regularization, polar_eps, damping;
  float tangential_damping, contact_margin, ipc_dhat, ipc_kappa;
  float residual_tol, update_tol;
  uint32_t solve_block;         // threads per block of the local solve (autotuned)
  uint32_t energy_after_sweep;  // vertex-energy pass that replaces the in-solve one
  uint32_t contact;             // IPC enabled for this step (ipcEnabled)
  uint32_t scene_count;         // scene slots in use
  uint32_t xcd_remap;           // XCD-aware block order for vertex kernels (autotuned)
  uint32_t color_index;         // colored pass: the color being swept
  uint32_t solve_lanes;         // lanes per vertex of the local solve for scenes that use lanes (4, 8, 16)
  uint32_t solve_classes;       // bit 0: scenes solved with lanes resident; bit 1: scenes solved one lane per vertex
  uint32_t cloth_base;          // cloth element records hold this pass's source pose (k_cloth_base ran)
  uint32_t target_records;      // energy pass: the records hold the target pose, the source energies in spare words
  uint32_t ipc_local;           // the solve's contact gather terms are in ipc_local (k_ipc_local_lanes ran)
  uint32_t cub_local;           // the lane sweep's cubature gather terms are in cub_local (k_cub_local_lanes ran)
  uint32_t kind_filter;         // lane sweep over: 0 every scene; 1-4 the scenes of that kind_blocks list
  uint32_t kind_blocks1, kind_blocks2, kind_blocks3, kind_blocks4;  // lengths of the four kind_blocks lists
  uint32_t fric_records;        // Jacobi sweeps: k_candidate_keep writes the friction records (fric)
};

// Device buffers of one solver. Layout (16-byte records throughout):
//   vtx_rest  float4[V]  rest xyz, lumped mass
//   vtx_info  uint4[V]   adjacency start/count, pinned, body
//   tet_rec   uint4[4T]  one 64-byte record per tetrahedron: vertex ids,
//                        then the columns of Dm^-1 as float bits with
//                        w = volume, lambda, mu
//   tet_k     float[144T] co-rotated rest stiffness (legacy path only)
//   adjacency uint32[4T]
//   cub       uint4[10 * V * K] record = {tet, weight bits, 0, 0} + 9 float4 Ubar
//   objectives float4[2V]
//   tri_ix    uint4[NT]  cloth triangle {v0, v1, v2, owner}; owner = the
//                        vertex with the smallest caller id, which sums the
//                        element into its vertex energy
//   tri_inv   float4[NT] inverse rest basis (row-major 2x2)
//   tri_area  float[NT]  rest area
//   hinge_ix  uint4[NH]  hinge {v0, v1, v2, v3}: triangles (0,1,2), (1,0,3)
//   hinge_rest float4[NH] {rest angle, rest edge length, owner bits, 0}
//   tri_row/hinge_row uint2[V] {begin, count}, *_elems: vertex -> element
//                        CSR, each row ascending in element id
//   IPC candidates (NC of them, the first NVT vertex-triangle):
//   cand      uint4[NC]  vertex ids (device order)
//   lag_n, lag_w, src_c, src_w, tgt_c, tgt_w  float4[NC]: lagged normal +
//                        force and weights, cached source/target contact
//                        (normal, distance) + weights
//   contact_ctl float4[1] w = the step's safe alpha
//   cand_row uint2[V] {begin, count}, cand_ids uint32[4 NC]: vertex ->
//   candidate CSR; active_count uint32[V], active_ids uint32[4 NC]: the
//   active rows; cand_meta uint32[NC]: bit 0 set for edge-edge records,
//   bits 1-2 the record slot of the vertex owning the candidate's energy
// Scenes: every scene owns whole 256-entry blocks of the vertex,
// tetrahedron, triangle and candidate ranges (block tables *_scene map a
// block to its scene, kNoScene for free blocks); per-scene control (4),
// history (64 x 8), contact alpha, scene scale and convergence scales.
struct DeviceState {
  uint32_t V = 0, T = 0, K = 0, bodies = 1;
  float4* vtx_rest = nullptr;
  uint4* vtx_info = nullptr;
  uint4* tet_rec = nullptr;
  float* tet_k = nullptr;
  uint32_t* adjacency = nullptr;
  uint4* cub = nullptr;
  float4* objectives = nullptr;
  uint32_t* color_order = nullptr;  // vertices sorted by color (ascending id per color)
  uint32_t* colors = nullptr;
  float4* pos[2] = {nullptr, nullptr};
  float4* predicted = nullptr;
  float4* velocity = nullptr;
  float4* old = nullptr;
  float4* vrot = nullptr;       // 3 per vertex (columns)
  float4* vrot_next = nullptr;  // frames for the next sweep, 3 per vertex
  // Per-vertex record the cubature terms gather, one 64-byte line per
  // vertex, paired with vrot ([0]) and vrot_next ([1]): {R.c0, x.x},
  // {R.c1, x.y}, {R.c2, x.z}, {x - x_pred, inertia weight m/(h^2 n), or -1
  // for a pinned vertex}, x being the pose the frame was computed at.
  float4* vcub[2] = {nullptr, nullptr};
  float4* trot = nullptr;       // 3 per tet (legacy path)
  float4* body_corr = nullptr;
  float4* final_update = nullptr;
  float4* local_diag = nullptr;  // 4 per vertex
  float4* tet_glob = nullptr;    // 2 per tet
  float4* vertex_energy = nullptr;
  float2* ipc_energy = nullptr;  // per vertex: owned contact energy at the source and the target pose
  float4* ipc_grad = nullptr;    // per vertex: contact gradient at the candidate pose (convergence)
  float4* ipc_local = nullptr;   // per vertex, 3 float4: contact g and H at a Jacobi sweep's source pose
  float4* cub_local = nullptr;   // per vertex, 3 float4: cubature g, symmetric H and record count at a sweep's source pose
  float4* ipc_erec = nullptr;    // per candidate: owned contact energies at the source and target pose, flags, owner
  float4* conv = nullptr;        // 5 per vertex
  float4* control = nullptr;     // 4
  float4* history = nullptr;     // 64 * 8
  void* partials = nullptr;      // block partial reductions
  uint32_t partial_capacity = 0;
  float* debug_g = nullptr;      // debug local systems
  float* debug_h = nullptr;
  // Exact diagnostics scratch: caller poses (arena order) and FP64 records.
  float4* diag_x = nullptr;
  float4* diag_xp = nullptr;
  double* diag_buf = nullptr;
  size_t diag_doubles = 0;
  // Cloth layer.
  uint32_t NT = 0, NH = 0;
  float cloth_lam = 0, cloth_mu = 0, cloth_thick = 0, cloth_bend = 0;
  uint4* tri_ix = nullptr;
  float4* tri_inv = nullptr;
  float* tri_area = nullptr;
  uint4* hinge_ix = nullptr;
  float4* hinge_rest = nullptr;
  uint2* tri_row = nullptr;
  uint32_t* tri_elems = nullptr;
  uint2* hinge_row = nullptr;
  uint32_t* hinge_elems = nullptr;
  float4* tri_glob = nullptr;   // 2 per triangle: candidate area stretch
  float4* tri_base = nullptr;   // 3 per triangle: membrane state at a sweep's source pose
  float4* hinge_base = nullptr; // 4 per hinge: corner gradients, angle, validity, energy at that pose
  // IPC contact candidates.
  uint32_t NC = 0, NVT = 0;
  float ipc_dmin = 0, ipc_mu = 0, ipc_eps = 0, ipc_safety = 0;
  uint4* cand = nullptr;
  float4 *lag_n = nullptr, *lag_w = nullptr;
  // Per candidate, 4 float4, written by k_candidate_keep before each Jacobi
  // sweep for the kept candidates with lagged friction: {x_l - old_l} of its
  // four vertices at the sweep's source pose, and in the first w the row's
  // energy at that pose (the trials' source term).
  float4* fric = nullptr;
  float4 *src_c = nullptr, *src_w = nullptr, *tgt_c = nullptr, *tgt_w = nullptr;
  float2* src_d = nullptr;  // per candidate: cached source distance (-1 invalid), lagged force
  float2* tgt_d = nullptr;  // per candidate: cached target distance (-1 invalid), lagged force
  float4* contact_ctl = nullptr;
  uint2* cand_row = nullptr;
  uint32_t *cand_ids = nullptr, *active_count = nullptr, *active_ids = nullptr, *cand_meta = nullptr;
  uint8_t* cand_keep = nullptr;   // per candidate: kept in its vertices' active rows this iteration
  // Broad-phase contact: scenes whose candidates are the static superset's
  // pairs within AABB reach, rebuilt on the device when motion since the
  // last build could bring an unlisted pair into reach.
  //   bp_tri    uint4[NS]  canonical surface triangles (ascending caller ids)
  //                        as device vertex ids; w = 0
  //   bp_edge   uint2[NE]  canonical surface edges, device ids
  //   bp_tbox, bp_ebox float4[2 NS], [2 NE]: AABB lo (w = 1 if a point is
  //                        non-finite or an unbounded pin) and hi at the build pose
  //   bp_tchunk, bp_echunk float4[2 NS/64], [2 NE/64]: AABBs of 64 consecutive
  //                        primitives (w as above, or-ed)
  //   bp_tsuper, bp_esuper same shape: AABBs of 64 consecutive chunks of a
  //                        scene, at the index of the superchunk's first chunk
  //   ring_row  uint2[V], ring_ids uint32[]: surface one-ring (device ids)
  //   bp_qv     uint32[V]  canonical (caller-order) vertex q of a scene at
  //                        vertex-range index VB + q -> its device vertex
  //   bp_vcnt   uint32[V]  VT matches of query VB + q, then their offsets
  //   bp_ecnt   uint32[NE] EE matches of edge query, then offsets
  //   bp_x      float4[V]  positions at the last build
  //   bp_cpose  float4[V]  the pose the source contact cache was last computed
  //                        at (the frame start, or the validated target of the
  //                        last promoted iteration); bp_tval the validated
  //                        target of the current iteration
  //   cand_key  uint2[NC]  {query (| 0x80000000 for EE), other} canonical key
  //   prev_key, prev_lag_n, prev_lag_w, prev_src_c, prev_src_w: the list
  //                        before a rebuild inside a frame
  //   bp_rowcnt, bp_rowfill uint32[V]: candidate CSR construction
  //   bp_ctl    uint4[S]  {flags, VT count, total, needed total}
  //   bp_reach  float4[S] {reach of the last build, margin, builds, 0}
  //   bp_motion uint4[3S] per-axis max of x - x_build (order keys), max |x - x_frame|
  //                        (float bits), min of x - x_build (order keys)
  //   bp_size   uint4[S]  {surface triangles, surface edges, candidate capacity,
  //                        candidates before the current build}
  //   bp_blk_v, bp_blk_e, bp_blk_r uint32[blocks]: block sums, then block
  //                        offsets, of the VT, EE and row-count scans
  //   bp_list   uint32[]   arena blocks of broad-phase scenes: bp_nv vertex
  //                        blocks, bp_ns surface-triangle blocks, bp_ne edge
  //                        blocks, bp_nc candidate blocks
  //   bp_work   uint32[]   per bp_list segment, the positions of the blocks
  //                        whose scene rebuilds at the current check point
  //   bp_work_n uint32[8]  their counts, then the count of bp_long
  //   bp_long   uint32[V]  vertices whose candidate row is longer than a
  //                        wavefront sorts in registers
  uint32_t NS = 0, NE = 0;
  uint32_t bp_nv = 0, bp_ns = 0, bp_ne = 0, bp_nc = 0;
  uint32_t* bp_list = nullptr;
  uint32_t *bp_work = nullptr, *bp_work_n = nullptr, *bp_long = nullptr;
  uint4* bp_size = nullptr;
  uint32_t* bp_rows_at = nullptr;  // per scene: first entry of its cand_ids / active_ids rows
  uint32_t *bp_blk_v = nullptr, *bp_blk_e = nullptr, *bp_blk_r = nullptr;
  uint4* bp_tri = nullptr;
  uint2* bp_edge = nullptr;
  float4 *bp_tbox = nullptr, *bp_ebox = nullptr, *bp_tchunk = nullptr, *bp_echunk = nullptr;
  float4 *bp_tsuper = nullptr, *bp_esuper = nullptr;
  uint2* ring_row = nullptr;
  uint32_t* ring_ids = nullptr;
  uint32_t *bp_qv = nullptr, *bp_vcnt = nullptr, *bp_ecnt = nullptr;
  uint32_t *bp_vprev = nullptr, *bp_eprev = nullptr;  // the query offsets of the last build
  float4 *bp_x = nullptr, *bp_cpose = nullptr, *bp_tval = nullptr;
  uint2 *cand_key = nullptr, *prev_key = nullptr;
  float4 *prev_lag_n = nullptr, *prev_lag_w = nullptr, *prev_src_c = nullptr, *prev_src_w = nullptr;
  uint32_t *bp_rowcnt = nullptr, *bp_rowfill = nullptr;
  uint4* bp_ctl = nullptr;
  float4* bp_reach = nullptr;
  uint4* bp_motion = nullptr;
  // Body-relative motion: a pair's AABB bound moves by at most its bodies'
  // deviations from their reference vertices plus the difference of those
  // references' displacements, and no pair of two bodies is closer than the
  // bodies' AABB gap at the build.
  //   bp_bodies uint4[S]  {first body, body count, 1 if the body rule applies, 0}
  //   bp_bref   uint32[B] reference (device) vertex of each body
  //   bp_bdev   uint4[B]  per-axis max deviation from the reference's displacement (float bits)
  //   bp_bbox   uint4[2B] body AABB at the build (order-preserving integer keys)
  uint4* bp_bodies = nullptr;
  uint32_t* bp_bref = nullptr;
  uint4 *bp_bdev = nullptr, *bp_bbox = nullptr;
  uint32_t *sblock_scene = nullptr, *eblock_scene = nullptr;
  uint4* scene_blk2 = nullptr;  // {sb0, sb1, eb0, eb1}: surface triangle and edge blocks
  // Scene tables.
  uint32_t S = 0;
  uint32_t *vblock_scene = nullptr, *tblock_scene = nullptr, *rblock_scene = nullptr, *cblock_scene = nullptr;
  uint4* scene_blk0 = nullptr;  // {vb0, vb1, tb0, tb1}
  uint4* scene_blk1 = nullptr;  // {rb0, rb1, cb0, cb1}
  uint4* scene_cnt = nullptr;   // {V, T, schedule colors (0 = free slot), bodies}
  uint4* scene_end = nullptr;   // absolute ends {vertex, tet, triangle, candidate}
  float* scene_scale = nullptr;
  uint32_t* scene_lanes = nullptr;  // lanes per vertex of the scene's local solve (1 = one thread per vertex)
  uint32_t* scene_kind = nullptr;   // per scene: bit 0 cloth elements, bit 1 contact candidates, bit 2 cubature,
                                    // bit 3 tetrahedral terms at an unpinned vertex (kKind*)
  uint32_t* kind_blocks = nullptr;  // vertex blocks of lane-swept scenes: kind 0 (no cloth or contact), the others
                                    // with tetrahedral terms, those of them with cubature records, then the
                                    // others without (counts in PassParams)
  void* scene_conv = nullptr;   // per-scene convergence scale records
  uint4* body_rng = nullptr;    // per body {scene vertex begin, end, 0, 0}
};

constexpr uint32_t kNoScene = 0xffffffffu;
constexpr uint32_t kSceneBlock = 256;  // scene ranges are multiples of this

constexpr uint32_t kHistoryCapacity = 64;
// Scene kinds (scene_kind): element terms beyond tetrahedra that a scene holds.
constexpr uint32_t kKindCloth = 1u;
constexpr uint32_t kKindContact = 2u;
constexpr uint32_t kKindCubature = 4u;  // a cubature record in use
constexpr uint32_t kKindTets = 8u;      // an unpinned vertex with incident tetrahedra or cubature records

// Launchers (jgs2_kernels.hip). All enqueue on `stream`.
void launch_predict(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_tet_polar(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_vertex_polar(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_solve(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_vertex_energy(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_reduce_candidate(const DeviceState& s, const PassParams& p, hipStream_t stream);
// The same reduction over one scene's blocks only ({vb0, vb1, tb0, tb1},
// {rb0, rb1, cb0, cb1}), for its admission preflight.
void launch_reduce_candidate_scene(const DeviceState& s, const PassParams& p, uint32_t scene, uint4 blk0, uint4 blk1,
                                   hipStream_t stream);
void launch_apply_candidate(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_convergence(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_copy_position(const DeviceState& s, const PassParams& p, hipStream_t stream);
// Cloth element records at positions `which` (0 = posA, 1 = posB).
// mode 0: element records at positions `which`; 1: the same, first keeping
// each element's energy (at the pose the records held) in a spare word;
// 2: only the elements of scenes whose candidate was rejected.
void launch_cloth_base(const DeviceState& s, const PassParams& p, uint32_t which, hipStream_t stream, uint32_t mode = 0);
void launch_body_correction(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_finalize(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_clear_frame(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_debug_local_systems(const DeviceState& s, const PassParams& p, hipStream_t stream);

// Exact FP64 diagnostics of one scene (arena ranges [vb, vb + V), [tb, tb + T),
// candidate slots [cb, cb + C)); records and summary in device doubles.
struct DiagParams {
  const float4* x;   // evaluated positions
  const float4* xp;  // predicted positions
  uint32_t vb, V, tb, T, cb, C, scene, objective_flags, corotated, pad;
  double ih2, floor_y, floor_k, kappa, dhat, polar_eps;
  double* vrec;      // 36 per vertex
  double* trec;      // 4 per tetrahedron
  double* trot;      // 9 per tetrahedron (co-rotated)
  double* crec;      // 2 per candidate slot
  double* summary;   // 32
};
void launch_diagnostics(const DeviceState& s, const DiagParams& q, hipStream_t stream);
// Fixed device buffer of the allocation-free diagnostics, in doubles (64 MiB).
constexpr size_t kDiagScratchDoubles = size_t(8) << 20;
// Device address of `device`'s copy of that buffer.
hipError_t diagnostics_scratch(int device, double** ptr);
// IPC passes (encodeFrame's contact dispatches, jgs2-solver.ts:2231-2400).
void launch_lag_contact(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_build_active_rows(const DeviceState& s, const PassParams& p, hipStream_t stream);
void launch_contact_step(const DeviceState& s, const PassParams& p, hipStream_t stream);
// Call with the source and target cache arrays swapped (see k_promote_contact).
void launch_promote_contact(const DeviceState& s, const PassParams& p, hipStream_t stream);
// Source cache copied into the target array before the host swaps them back.
void launch_contact_cache_home(const DeviceState& s, hipStream_t stream);
uint32_t partial_blocks_needed(uint32_t vertices, uint32_t tets, uint32_t triangles, uint32_t candidates);
size_t scene_conv_bytes(uint32_t scenes);
// Broad-phase contact (jgs2_broad_phase.hip). `check` decides per scene
// whether the list must be rebuilt before this point of the frame (always
// for scenes whose flags ask for it); `carry` keeps lagged friction data
// and recomputes the source contact cache, for a rebuild inside a frame.
constexpr uint32_t kBpScene = 1u;      // the scene's candidates come from the broad phase
constexpr uint32_t kBpNeed = 2u;       // rebuild at the next check point
constexpr uint32_t kBpOverflow = 4u;   // a build needed more candidates or row entries than fit (sticky)
constexpr uint32_t kBpForce = 8u;      // rebuild at the next check point regardless of motion
void launch_broad_phase(const DeviceState& s, const PassParams& p, bool check, bool carry, hipStream_t stream);
// Local-solve timing of one launch configuration (autotuner).
void launch_solve_config(const DeviceState& s, const PassParams& p, hipStream_t stream);
// Reads `records` uint4 of the cubature array to evict the caches.
void launch_cache_flush(const DeviceState& s, uint32_t records, hipStream_t stream);
size_t partial_bytes(uint32_t blocks);

}  // namespace jgs2
