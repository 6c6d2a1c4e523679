// libjgs2hip host side: the device half of JGS2GpuSolver
// (src/simulation/gpu/jgs2-solver.ts) behind a C ABI. Validation, coloring,
// packing and frame orchestration run here; every pass of the frame is a
// HIP kernel in jgs2_kernels.hip enqueued on the batch's stream.
//
// Every scene lives in a batch: a device arena whose vertex, tetrahedron,
// triangle and candidate ranges are handed out in 256-entry blocks. A frame
// is one launch per pass over all resident scenes; reductions, line-search
// control, history and the convergence latch are per scene. jgs2_create()
// builds a private one-scene batch sized exactly for its input.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "jgs2_device.h"
#include "jgs2_internal.h"
#include "jgs2hip.h"
#include "jgs2hip_scene.h"

namespace jgs2 {

namespace {
thread_local std::string g_last_error;
}

jgs2_status fail(jgs2_status status, const char* format, ...) {
  char buffer[1024];
  va_list args;
  va_start(args, format);
  vsnprintf(buffer, sizeof(buffer), format, args);
  va_end(args);
  g_last_error = buffer;
  return status;
}
void clear_error() { g_last_error.clear(); }

namespace {

constexpr double kDeterminantFloor = 1e-4;
constexpr uint32_t kMaxColor = 255;  // jgs2-solver.ts:255
constexpr double kMaxRuntimeTolerance = 1e-3;  // jgs2-globalization.ts:24

#define HIP_CHECK(expr)                                                                              \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) return fail(JGS2_ERROR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define TRYS(expr)                      \
  do {                                  \
    jgs2_status t_ = (expr);            \
    if (t_ != JGS2_OK) return t_;       \
  } while (0)

float round_f32(double v) { return float(v); }

}  // namespace
}  // namespace jgs2

using namespace jgs2;

namespace {

// First-fit allocator over [0, capacity) units with coalescing free list.
struct RangeAlloc {
  uint64_t capacity = 0;
  std::map<uint64_t, uint64_t> free;  // begin -> length
  void init(uint64_t cap) {
    capacity = cap;
    free.clear();
    if (cap) free[0] = cap;
  }
  bool alloc(uint64_t n, uint64_t& at) {
    if (n == 0) {
      at = 0;
      return true;
    }
    for (auto it = free.begin(); it != free.end(); ++it) {
      if (it->second < n) continue;
      at = it->first;
      const uint64_t rest = it->second - n;
      free.erase(it);
      if (rest) free[at + n] = rest;
      return true;
    }
    return false;
  }
  void release(uint64_t at, uint64_t n) {
    if (n == 0) return;
    auto it = free.emplace(at, n).first;
    auto next = std::next(it);
    if (next != free.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free.erase(next);
    }
    if (it != free.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free.erase(it);
      }
    }
  }
  // One past the highest allocated unit.
  uint64_t high() const {
    if (free.empty()) return capacity;
    auto last = std::prev(free.end());
    return last->first + last->second == capacity ? last->first : capacity;
  }
};

struct Span {
  uint64_t at = 0, n = 0;
};

}  // namespace

// One scene resident in a batch arena. Handles returned by jgs2_create
// own a private one-scene batch; handles returned by jgs2_batch_admit
// belong to the batch that stepped them.
struct jgs2_solver {
  bool destroyed = false;
  jgs2_batch* batch = nullptr;
  bool owns_batch = false;
  uint32_t slot = 0;
  jgs2_settings defaults{};
  double preprocessing_timestep = 0;
  float scene_scale = 0;
  bool globalization = true;
  bool corotated = false;
  uint32_t V = 0, T = 0, K = 0, NT = 0, NH = 0, NC = 0, NVT = 0, bodies = 1, color_count = 1, free_vertices = 0;
  uint32_t cub_used = 0;  // cubature records in use (tetrahedron set, weight > 0)
  bool free_tets = false;  // an unpinned vertex has an incident tetrahedron or a cubature record in use
  // Arena spans: vertex/tet/triangle/candidate spans in 256-entry blocks;
  // the others in entries.
  Span vblk, tblk, rblk, cblk, adj, hin, tel, hel, cel, body;
  // Broad-phase contact: canonical surface, its spans (triangle and edge
  // blocks, one-ring entries), the candidate capacity and the reach margin.
  bool bp = false;
  uint32_t NS = 0, NE = 0;
  Surface surface;
  float bp_margin = 0;
  Span sblk, eblk, rng;
  std::vector<uint32_t> colors;         // by caller vertex id
  std::vector<uint32_t> color_offsets;  // color_count + 1
  std::vector<uint32_t> color_order;    // arena vertex ids by color
  // Device order within the scene: vertices may be stored along a Morton
  // curve of their rest positions; the boundary maps ids both ways.
  std::vector<uint32_t> vnew, vold, tnew, told;
  std::vector<float> objectives;        // 8 per vertex (host mirror)
  std::vector<uint8_t> pinned;
  uint32_t objective_flags = 0;
  uint32_t active_forces = 0, active_targets = 0;
  // packJGS2GpuClothArena material words (cloth-layout.ts:335-343): lambda,
  // mu, thickness, bending stiffness; f64 expressions rounded to f32.
  bool has_cloth = false;
  float cloth_mat[4] = {0, 0, 0, 0};
  uint64_t vbase() const { return vblk.at * kSceneBlock; }
  uint64_t tbase() const { return tblk.at * kSceneBlock; }
  uint64_t rbase() const { return rblk.at * kSceneBlock; }
  uint64_t cbase() const { return cblk.at * kSceneBlock; }
};

// A device arena holding many independent scenes. Every pass of a frame is
// one launch over all resident scenes; reductions are per scene.
struct jgs2_batch {
  DeviceState dev;
  hipStream_t stream = nullptr;
  int device = 0;  // the device current at creation, which owns the stream
  jgs2_settings defaults{};
  jgs2_batch_config cfg{};
  bool corotated = false;
  RangeAlloc va, ta, ra, ca, adja, hina, tela, hela, cela, bodya, slota, sa, ea, rga;
  std::vector<jgs2_solver*> slots;
  uint64_t version = 0;      // bumps on every admission or retirement
  uint64_t stepped_version = ~0ull;  // the version the last submitted frame ran at
  uint64_t synced_version = ~0ull;
  uint32_t color_count = 1;
  std::vector<uint32_t> color_offsets;
  uint32_t kind_blocks_n[4] = {0, 0, 0, 0};  // lengths of the dev.kind_blocks lists
  // Launch shape of the local solve. Block size and XCD remap leave every
  // result unchanged (autotuned); lanes per vertex set the summation order
  // of the gathered terms and stay fixed unless a caller sets them.
  uint32_t solve_block = 256, xcd_remap = 1, solve_lanes = 4;
  bool solve_serial_only = false;  // every scene one thread per vertex
  uint64_t arena_bytes = 0;
  // Cloth material shared by every resident cloth scene (one set of
  // material words per arena).
  uint32_t cloth_scenes = 0;
  // hipGraph frames
  bool use_graphs = false;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_key = 0;
  bool graph_valid = false;
  uint64_t graph_instantiations = 0;  // execs built from scratch
  uint64_t graph_updates = 0;         // captures applied to the existing exec in place
  // Local-solve profiling
  bool profiling = false;
  std::vector<hipEvent_t> events;
  size_t event_used = 0;
  hipEvent_t frame_begin = nullptr, frame_end = nullptr;
  uint32_t active_scenes() const {
    uint32_t n = 0;
    for (auto* s : slots) n += s ? 1u : 0u;
    return n;
  }
};
namespace {

void free_device(DeviceState& d) {
  void* ptrs[] = {d.vtx_rest, d.vtx_info, d.tet_rec, d.vcub[0], d.vcub[1], d.tet_k, d.adjacency, d.cub, d.objectives,
                  d.color_order, d.colors, d.pos[0], d.pos[1], d.predicted, d.velocity, d.old, d.vrot, d.vrot_next, d.trot,
                  d.body_corr, d.final_update, d.local_diag, d.tet_glob, d.vertex_energy, d.ipc_energy, d.ipc_grad, d.ipc_local, d.cub_local, d.ipc_erec, d.conv, d.control,
                  d.history, d.partials, d.debug_g, d.debug_h, d.diag_x, d.diag_xp, d.diag_buf, d.tri_ix, d.tri_inv, d.tri_area, d.hinge_ix,
                  d.hinge_rest, d.tri_row, d.tri_elems, d.hinge_row, d.hinge_elems, d.tri_glob, d.tri_base, d.hinge_base, d.cand, d.lag_n, d.lag_w, d.fric,
                  d.src_c, d.src_w, d.src_d, d.tgt_d, d.tgt_c, d.tgt_w, d.contact_ctl, d.cand_row, d.cand_ids,
                  d.active_count, d.active_ids, d.cand_meta, d.cand_keep, d.vblock_scene, d.tblock_scene, d.rblock_scene,
                  d.cblock_scene, d.scene_blk0, d.scene_blk1, d.scene_cnt, d.scene_end, d.scene_scale, d.scene_lanes, d.scene_kind, d.kind_blocks, d.scene_conv,
                  d.body_rng, d.bp_list, d.bp_work, d.bp_work_n, d.bp_long, d.bp_size, d.bp_rows_at, d.bp_blk_v, d.bp_blk_e, d.bp_blk_r, d.bp_tri, d.bp_edge, d.bp_tbox,
                  d.bp_ebox, d.bp_tchunk, d.bp_echunk, d.bp_tsuper, d.bp_esuper, d.ring_row, d.ring_ids, d.bp_qv, d.bp_vcnt, d.bp_ecnt, d.bp_vprev, d.bp_eprev, d.bp_x,
                  d.bp_cpose, d.bp_tval, d.cand_key, d.prev_key, d.prev_lag_n, d.prev_lag_w, d.prev_src_c, d.prev_src_w,
                  d.bp_rowcnt, d.bp_rowfill, d.bp_ctl, d.bp_reach, d.bp_motion, d.bp_bodies, d.bp_bref, d.bp_bdev, d.bp_bbox, d.sblock_scene, d.eblock_scene,
                  d.scene_blk2};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  d = DeviceState{};
}

// Allocates `count` elements (at least one); with `dry` only counts bytes.
template <class T> jgs2_status dalloc(T** ptr, size_t count, uint64_t* bytes = nullptr, bool dry = false) {
  *ptr = nullptr;
  if (count == 0) count = 1;
  if (dry) {
    if (bytes) *bytes += count * sizeof(T);
    return JGS2_OK;
  }
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T));
  if (e != hipSuccess)
    return fail(JGS2_ERROR_MEMORY, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
  if (bytes) *bytes += count * sizeof(T);
  return JGS2_OK;
}


// ---------------------------------------------------------------- settings
// resolveJGS2StepSettings + validateStepSettings (jgs2-solver.ts:429-550).
jgs2_status merge_settings(const jgs2_settings& defaults, const jgs2_settings* ov, uint64_t mask, jgs2_settings& out) {
  out = defaults;
  if (ov) {
#define TAKE(bit, field) \
  if (mask & (bit)) out.field = ov->field;
    TAKE(JGS2_SET_SCHEDULE, schedule)
    TAKE(JGS2_SET_TIMESTEP, timestep)
    if (mask & JGS2_SET_GRAVITY) std::memcpy(out.gravity, ov->gravity, sizeof(out.gravity));
    TAKE(JGS2_SET_ITERATIONS, iterations)
    TAKE(JGS2_SET_FLOOR_HEIGHT, floor_height)
    TAKE(JGS2_SET_FLOOR_STIFFNESS, floor_stiffness)
    TAKE(JGS2_SET_VELOCITY_DAMPING, velocity_damping)
    TAKE(JGS2_SET_REGULARIZATION, regularization)
    TAKE(JGS2_SET_ROTATION_EPSILON, rotation_epsilon)
    TAKE(JGS2_SET_MAX_STEP, max_step)
    TAKE(JGS2_SET_RESIDUAL_TOLERANCE, residual_tolerance)
    TAKE(JGS2_SET_NORMALIZED_UPDATE_TOLERANCE, normalized_update_tolerance)
    TAKE(JGS2_SET_CONTACT_TANGENTIAL_DAMPING, contact_tangential_damping)
    TAKE(JGS2_SET_CONTACT_MARGIN, contact_margin)
    TAKE(JGS2_SET_IPC_ACTIVATION_DISTANCE, ipc_activation_distance)
    TAKE(JGS2_SET_IPC_MINIMUM_DISTANCE, ipc_minimum_distance)
    TAKE(JGS2_SET_IPC_BARRIER_STIFFNESS, ipc_barrier_stiffness)
    TAKE(JGS2_SET_IPC_FRICTION_COEFFICIENT, ipc_friction_coefficient)
    TAKE(JGS2_SET_IPC_FRICTION_VELOCITY_EPSILON, ipc_friction_velocity_epsilon)
    TAKE(JGS2_SET_IPC_STEP_SAFETY, ipc_step_safety)
    TAKE(JGS2_SET_HORIZONTAL_BODY_CORRECTION, horizontal_body_correction)
    TAKE(JGS2_SET_PARITY_MODE, parity_mode)
#undef TAKE
  }
  if (out.parity_mode) {
    out.velocity_damping = 1;
    out.contact_tangential_damping = 0;
    out.horizontal_body_correction = 0;
  }
  const jgs2_settings& s = out;
  if (s.schedule != JGS2_SCHEDULE_JACOBI && s.schedule != JGS2_SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL)
    return fail(JGS2_ERROR_RANGE, "Unknown JGS2 schedule: %d.", s.schedule);
  struct {
    const char* name;
    double value;
  } finite[] = {{"timestep", s.timestep},
                {"floorHeight", s.floor_height},
                {"floorStiffness", s.floor_stiffness},
                {"velocityDamping", s.velocity_damping},
                {"regularization", s.regularization},
                {"rotationEpsilon", s.rotation_epsilon},
                {"maxStep", s.max_step},
                {"residualTolerance", s.residual_tolerance},
                {"normalizedUpdateTolerance", s.normalized_update_tolerance},
                {"contactTangentialDamping", s.contact_tangential_damping},
                {"contactMargin", s.contact_margin},
                {"ipcActivationDistance", s.ipc_activation_distance},
                {"ipcMinimumDistance", s.ipc_minimum_distance},
                {"ipcBarrierStiffness", s.ipc_barrier_stiffness},
                {"ipcFrictionCoefficient", s.ipc_friction_coefficient},
                {"ipcFrictionVelocityEpsilon", s.ipc_friction_velocity_epsilon},
                {"ipcStepSafety", s.ipc_step_safety},
                {"gravity.x", s.gravity[0]},
                {"gravity.y", s.gravity[1]},
                {"gravity.z", s.gravity[2]}};
  for (auto& f : finite)
    if (!std::isfinite(f.value)) return fail(JGS2_ERROR_RANGE, "%s must be finite; got %g.", f.name, f.value);
  if (!(s.timestep > 0)) return fail(JGS2_ERROR_RANGE, "timestep must be positive.");
  if (s.floor_stiffness < 0) return fail(JGS2_ERROR_RANGE, "floorStiffness must be nonnegative.");
  if (s.velocity_damping < 0 || s.velocity_damping > 1)
    return fail(JGS2_ERROR_RANGE, "velocityDamping must be between zero and one.");
  if (!(s.regularization > 0)) return fail(JGS2_ERROR_RANGE, "regularization must be positive.");
  if (!(s.rotation_epsilon > 0)) return fail(JGS2_ERROR_RANGE, "rotationEpsilon must be positive.");
  if (s.max_step < 0) return fail(JGS2_ERROR_RANGE, "maxStep must be nonnegative.");
  if (!(s.residual_tolerance > 0) || s.residual_tolerance > kMaxRuntimeTolerance)
    return fail(JGS2_ERROR_RANGE, "residualTolerance must be positive and no greater than 0.001.");
  if (!(s.normalized_update_tolerance > 0) || s.normalized_update_tolerance > kMaxRuntimeTolerance)
    return fail(JGS2_ERROR_RANGE, "normalizedUpdateTolerance must be positive and no greater than 0.001.");
  if (s.contact_tangential_damping < 0)
    return fail(JGS2_ERROR_RANGE, "contactTangentialDamping must be finite and nonnegative.");
  if (s.contact_margin < 0) return fail(JGS2_ERROR_RANGE, "contactMargin must be finite and nonnegative.");
  if (!(s.ipc_activation_distance > 0)) return fail(JGS2_ERROR_RANGE, "ipcActivationDistance must be positive.");
  if (s.ipc_minimum_distance < 0 || s.ipc_minimum_distance >= s.ipc_activation_distance)
    return fail(JGS2_ERROR_RANGE, "ipcMinimumDistance must be nonnegative and below ipcActivationDistance.");
  if (s.ipc_barrier_stiffness < 0) return fail(JGS2_ERROR_RANGE, "ipcBarrierStiffness must be nonnegative.");
  if (s.ipc_friction_coefficient < 0) return fail(JGS2_ERROR_RANGE, "ipcFrictionCoefficient must be nonnegative.");
  if (!(s.ipc_friction_velocity_epsilon > 0))
    return fail(JGS2_ERROR_RANGE, "ipcFrictionVelocityEpsilon must be positive.");
  if (!(s.ipc_step_safety > 0) || !(s.ipc_step_safety < 1))
    return fail(JGS2_ERROR_RANGE, "ipcStepSafety must be strictly between zero and one.");
  return JGS2_OK;
}

// ---------------------------------------------------------------- input validation
bool finite_array(const float* a, size_t n, size_t* bad) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) {
      *bad = i;
      return false;
    }
  return true;
}

// validateJGS2GpuInput (layout.ts:223-416).
jgs2_status validate_input(const jgs2_scene_input& in) {
  const uint32_t V = in.vertex_count, T = in.tet_count, K = in.cubature_k;
  if (V < 1) return fail(JGS2_ERROR_RANGE, "vertexCount must be an integer >= 1; got %u.", V);
  if (T < 1) return fail(JGS2_ERROR_RANGE, "tetCount must be an integer >= 1; got %u.", T);
  if (!in.positions || !in.vertex_rest || !in.vertex_info || !in.tet_indices || !in.tet_inverse_dm || !in.tet_meta ||
      !in.adjacency || (K > 0 && (!in.cubature_tet_ids || !in.cubature_weights || !in.cubature_basis)))
    return fail(JGS2_ERROR_TYPE, "A required JGS2 input array is NULL.");
  if ((in.target_positions == nullptr) != (in.target_stiffnesses == nullptr))
    return fail(JGS2_ERROR_RANGE,
                "objectives.targetPositions and objectives.targetStiffnesses must be provided together.");
  size_t bad = 0;
  if (in.external_forces && !finite_array(in.external_forces, 3 * size_t(V), &bad))
    return fail(JGS2_ERROR_RANGE, "objectives.externalForces[%zu] must be finite.", bad);
  if (in.target_positions) {
    if (!finite_array(in.target_positions, 3 * size_t(V), &bad))
      return fail(JGS2_ERROR_RANGE, "objectives.targetPositions[%zu] must be finite.", bad);
    if (!finite_array(in.target_stiffnesses, V, &bad))
      return fail(JGS2_ERROR_RANGE, "objectives.targetStiffnesses[%zu] must be finite.", bad);
    for (uint32_t v = 0; v < V; ++v)
      if (in.target_stiffnesses[v] < 0)
        return fail(JGS2_ERROR_RANGE, "objectives.targetStiffnesses[%u] must be nonnegative.", v);
  }
  for (uint32_t v = 0; v < V; ++v) {
    if (in.vertex_info[4 * size_t(v) + 2] == 0) continue;
    bool force = in.external_forces && (in.external_forces[3 * size_t(v)] != 0 ||
                                        in.external_forces[3 * size_t(v) + 1] != 0 ||
                                        in.external_forces[3 * size_t(v) + 2] != 0);
    bool target = in.target_stiffnesses && in.target_stiffnesses[v] != 0;
    if (force || target)
      return fail(JGS2_ERROR_RANGE,
                  "Pinned vertex %u cannot have a nonzero external force or quadratic-target stiffness.", v);
  }
  uint32_t max_body = 0;
  for (uint32_t v = 0; v < V; ++v) max_body = std::max(max_body, in.vertex_info[4 * size_t(v) + 3]);
  if (uint64_t(max_body) + 1 > V)
    return fail(JGS2_ERROR_RANGE,
                "vertexInfo body ids must be compact nonnegative values whose maximum is smaller than vertexCount (%u); got %u.",
                V, max_body);
  if (!finite_array(in.positions, 4 * size_t(V), &bad)) return fail(JGS2_ERROR_RANGE, "positions[%zu] must be finite.", bad);
  if (in.velocities && !finite_array(in.velocities, 4 * size_t(V), &bad))
    return fail(JGS2_ERROR_RANGE, "velocities[%zu] must be finite.", bad);
  if (!finite_array(in.vertex_rest, 4 * size_t(V), &bad)) return fail(JGS2_ERROR_RANGE, "vertexRest[%zu] must be finite.", bad);
  if (!finite_array(in.tet_inverse_dm, 12 * size_t(T), &bad))
    return fail(JGS2_ERROR_RANGE, "tetInverseDm[%zu] must be finite.", bad);
  if (!finite_array(in.tet_meta, 4 * size_t(T), &bad)) return fail(JGS2_ERROR_RANGE, "tetMeta[%zu] must be finite.", bad);
  if (in.tet_rest_stiffness && !finite_array(in.tet_rest_stiffness, 144 * size_t(T), &bad))
    return fail(JGS2_ERROR_RANGE, "tetRestStiffness[%zu] must be finite.", bad);
  const size_t R = size_t(V) * K;
  if (K > 0) {
    if (!finite_array(in.cubature_weights, R, &bad)) return fail(JGS2_ERROR_RANGE, "cubatureWeights[%zu] must be finite.", bad);
    if (!finite_array(in.cubature_basis, R * 36, &bad)) return fail(JGS2_ERROR_RANGE, "cubatureBasis[%zu] must be finite.", bad);
  }
  for (size_t i = 0; i < 4 * size_t(T); ++i)
    if (in.tet_indices[i] >= V)
      return fail(JGS2_ERROR_RANGE, "tetIndices[%zu] references vertex %u, but vertexCount is %u.", i,
                  in.tet_indices[i], V);
  uint32_t snh = 0;
  for (uint32_t t = 0; t < T; ++t) {
    const float* m = &in.tet_meta[4 * size_t(t)];
    if (!(m[0] > 0) || !(m[2] > 0))
      return fail(JGS2_ERROR_RANGE, "tetMeta for tetrahedron %u requires positive volume and mu.", t);
    if (m[3] != JGS2_MATERIAL_COROTATED_LINEAR && m[3] != JGS2_MATERIAL_STABLE_NEO_HOOKEAN)
      return fail(JGS2_ERROR_RANGE, "tetMeta for tetrahedron %u has unknown material model %g.", t, double(m[3]));
    if (m[3] == JGS2_MATERIAL_STABLE_NEO_HOOKEAN && !(m[1] > 0))
      return fail(JGS2_ERROR_RANGE, "Stable Neo-Hookean tetMeta for tetrahedron %u requires positive lambda.", t);
    if (m[3] == JGS2_MATERIAL_COROTATED_LINEAR && !(3 * m[1] + 2 * m[2] > 0))
      return fail(JGS2_ERROR_RANGE, "Co-rotated tetMeta for tetrahedron %u requires positive bulk modulus.", t);
    snh += m[3] == JGS2_MATERIAL_STABLE_NEO_HOOKEAN ? 1u : 0u;
  }
  if (snh != 0 && snh != T)  // inferJGS2MaterialMode (layout.ts:510-526)
    return fail(JGS2_ERROR_RANGE,
                "A production JGS2 solve cannot mix stable Neo-Hookean and co-rotated regression tetrahedra in one precomputation.");
  if (snh == 0 && !in.tet_rest_stiffness)
    return fail(JGS2_ERROR_TYPE, "The co-rotated material requires tetRestStiffness.");
  for (uint64_t i = 0; i < in.adjacency_length; ++i)
    if (in.adjacency[i] >= T)
      return fail(JGS2_ERROR_RANGE, "adjacency[%llu] references tetrahedron %u, but tetCount is %u.",
                  (unsigned long long)i, in.adjacency[i], T);
  // Canonical CSR: rows in vertex order, each listing its incident
  // tetrahedra once, ascending (layout.ts:291-401).
  std::vector<uint32_t> expected_count(V, 0);
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t* ix = &in.tet_indices[4 * size_t(t)];
    for (int l = 0; l < 4; ++l) {
      bool dup = false;
      for (int q = 0; q < l; ++q) dup |= ix[q] == ix[l];
      if (!dup) expected_count[ix[l]] += 1;
    }
  }
  std::vector<uint32_t> cursor(V);
  uint64_t offset = 0;
  for (uint32_t v = 0; v < V; ++v) {
    const uint32_t start = in.vertex_info[4 * size_t(v)], count = in.vertex_info[4 * size_t(v) + 1];
    if (uint64_t(start) + count > in.adjacency_length)
      return fail(JGS2_ERROR_RANGE, "vertexInfo for vertex %u addresses adjacency range [%u, %llu), beyond %llu.", v,
                  start, (unsigned long long)(uint64_t(start) + count), (unsigned long long)in.adjacency_length);
    if (start != offset)
      return fail(JGS2_ERROR_RANGE, "vertexInfo for vertex %u must start at canonical adjacency offset %llu; got %u.",
                  v, (unsigned long long)offset, start);
    if (count != expected_count[v])
      return fail(JGS2_ERROR_RANGE, "vertexInfo for vertex %u must list exactly %u unique incident tetrahedra; got %u.",
                  v, expected_count[v], count);
    cursor[v] = start;
    offset += count;
    const float mass = in.vertex_rest[4 * size_t(v) + 3];
    if (in.vertex_info[4 * size_t(v) + 2] == 0 && !(mass > 0))
      return fail(JGS2_ERROR_RANGE, "Unpinned vertex %u must have positive mass.", v);
  }
  if (offset != in.adjacency_length)
    return fail(JGS2_ERROR_RANGE, "adjacency contains %llu unused entries.",
                (unsigned long long)(in.adjacency_length - offset));
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t* ix = &in.tet_indices[4 * size_t(t)];
    for (int l = 0; l < 4; ++l) {
      bool dup = false;
      for (int q = 0; q < l; ++q) dup |= ix[q] == ix[l];
      if (dup) continue;
      const uint32_t v = ix[l];
      const uint32_t at = cursor[v]++;
      if (in.adjacency[at] != t)
        return fail(JGS2_ERROR_RANGE, "adjacency[%u] for vertex %u must equal incident tetrahedron %u; got %u.", at, v,
                    t, in.adjacency[at]);
    }
  }
  for (size_t r = 0; r < R; ++r) {
    if (in.cubature_weights[r] < 0) return fail(JGS2_ERROR_RANGE, "cubatureWeights[%zu] must be nonnegative.", r);
    if (in.cubature_weights[r] > 0 && in.cubature_tet_ids[r] >= T)
      return fail(JGS2_ERROR_RANGE, "Positive cubature record %zu references tetrahedron %u, but tetCount is %u.", r,
                  in.cubature_tet_ids[r], T);
  }
  return JGS2_OK;
}

// validateJGS2GpuClothInput (cloth-layout.ts:209-297).
jgs2_status validate_cloth(const jgs2_cloth_input& c, uint32_t V) {
  const uint32_t NT = c.triangle_count, NH = c.hinge_count;
  if ((NT && (!c.triangle_indices || !c.triangle_inverse_rest_bases || !c.triangle_rest_areas)) ||
      (NH && (!c.hinge_indices || !c.hinge_rest_angles || !c.hinge_rest_edge_lengths)))
    return fail(JGS2_ERROR_TYPE, "A required cloth array is NULL.");
  auto distinct = [&](const uint32_t* ix, uint32_t width, uint32_t element, const char* label) -> jgs2_status {
    for (uint32_t l = 0; l < width; ++l) {
      if (ix[l] >= V)
        return fail(JGS2_ERROR_RANGE, "%s %u index %u is outside vertexCount %u.", label, element, ix[l], V);
      for (uint32_t q = 0; q < l; ++q)
        if (ix[q] == ix[l]) return fail(JGS2_ERROR_RANGE, "%s %u must reference %u distinct vertices.", label, element, width);
    }
    return JGS2_OK;
  };
  for (uint32_t t = 0; t < NT; ++t) {
    jgs2_status st = distinct(c.triangle_indices + 3 * size_t(t), 3, t, "Triangle");
    if (st != JGS2_OK) return st;
  }
  for (uint32_t h = 0; h < NH; ++h) {
    jgs2_status st = distinct(c.hinge_indices + 4 * size_t(h), 4, h, "Hinge");
    if (st != JGS2_OK) return st;
  }
  size_t bad = 0;
  if (NT && !finite_array(c.triangle_inverse_rest_bases, 4 * size_t(NT), &bad))
    return fail(JGS2_ERROR_RANGE, "triangleInverseRestBases[%zu] must be finite.", bad);
  if (NT && !finite_array(c.triangle_rest_areas, NT, &bad))
    return fail(JGS2_ERROR_RANGE, "triangleRestAreas[%zu] must be finite.", bad);
  if (NH && !finite_array(c.hinge_rest_angles, NH, &bad))
    return fail(JGS2_ERROR_RANGE, "hingeRestAngles[%zu] must be finite.", bad);
  if (NH && !finite_array(c.hinge_rest_edge_lengths, NH, &bad))
    return fail(JGS2_ERROR_RANGE, "hingeRestEdgeLengths[%zu] must be finite.", bad);
  for (uint32_t t = 0; t < NT; ++t) {
    if (!(c.triangle_rest_areas[t] > 0)) return fail(JGS2_ERROR_RANGE, "triangleRestAreas[%u] must be positive.", t);
    const float* b = c.triangle_inverse_rest_bases + 4 * size_t(t);
    const float det = b[0] * b[3] - b[1] * b[2];
    if (det == 0 || !std::isfinite(det))
      return fail(JGS2_ERROR_RANGE, "triangleInverseRestBases for triangle %u must be invertible.", t);
  }
  for (uint32_t h = 0; h < NH; ++h)
    if (!(c.hinge_rest_edge_lengths[h] > 0)) return fail(JGS2_ERROR_RANGE, "hingeRestEdgeLengths[%u] must be positive.", h);
  if (!(c.young_modulus > 0) || !std::isfinite(c.young_modulus))
    return fail(JGS2_ERROR_RANGE, "Cloth Young's modulus must be finite and positive.");
  if (!(c.poisson_ratio >= 0 && c.poisson_ratio < 0.5))
    return fail(JGS2_ERROR_RANGE, "Cloth Poisson ratio must be finite and in [0, 0.5).");
  if (!(c.thickness > 0) || !std::isfinite(c.thickness))
    return fail(JGS2_ERROR_RANGE, "Cloth thickness must be finite and positive.");
  if (!(c.bending_stiffness >= 0) || !std::isfinite(c.bending_stiffness))
    return fail(JGS2_ERROR_RANGE, "Cloth bending stiffness must be finite and nonnegative.");
  return JGS2_OK;
}

// validateIpcContactCandidatesForGpu and the incidence packer's range check
// (ipc-contact-layout.ts:117-178, :209-218).
jgs2_status validate_contact(const jgs2_contact_input& c, uint32_t V) {
  const uint64_t n = uint64_t(c.vertex_triangle_count) + c.edge_edge_count;
  if (n > 0x3fffffffull) return fail(JGS2_ERROR_RANGE, "Total IPC candidate count must be an unsigned 32-bit integer.");
  if (n && !c.packed_indices) return fail(JGS2_ERROR_TYPE, "Packed IPC indices must not be NULL.");
  for (uint64_t k = 0; k < n; ++k) {
    const uint32_t* ix = &c.packed_indices[4 * k];
    const char* type = k < c.vertex_triangle_count ? "VT" : "EE";
    for (int l = 0; l < 4; ++l) {
      if (ix[l] >= V)
        return fail(JGS2_ERROR_RANGE, "IPC candidate references vertex %u, but vertexCount is %u.", ix[l], V);
      for (int q = 0; q < l; ++q)
        if (ix[q] == ix[l])
          return fail(JGS2_ERROR_RANGE, "%s candidate %llu must reference four distinct vertices.", type,
                      (unsigned long long)k);
    }
  }
  return JGS2_OK;
}

// minimumJGS2InputDeformationDeterminant (layout.ts:549-588), f64.
double min_input_determinant(const jgs2_scene_input& in) {
  auto det_cols = [](double c0x, double c0y, double c0z, double c1x, double c1y, double c1z, double c2x, double c2y,
                     double c2z) {
    return c0x * (c1y * c2z - c1z * c2y) - c1x * (c0y * c2z - c0z * c2y) + c2x * (c0y * c1z - c0z * c1y);
  };
  double minimum = INFINITY;
  for (uint32_t t = 0; t < in.tet_count; ++t) {
    const uint32_t* ix = &in.tet_indices[4 * size_t(t)];
    const float* p = in.positions;
    size_t i0 = 4 * size_t(ix[0]), i1 = 4 * size_t(ix[1]), i2 = 4 * size_t(ix[2]), i3 = 4 * size_t(ix[3]);
    double ds = det_cols(double(p[i1]) - p[i0], double(p[i1 + 1]) - p[i0 + 1], double(p[i1 + 2]) - p[i0 + 2],
                         double(p[i2]) - p[i0], double(p[i2 + 1]) - p[i0 + 1], double(p[i2 + 2]) - p[i0 + 2],
                         double(p[i3]) - p[i0], double(p[i3 + 1]) - p[i0 + 1], double(p[i3 + 2]) - p[i0 + 2]);
    const float* m = &in.tet_inverse_dm[12 * size_t(t)];
    double id = det_cols(m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]);
    double d = ds * id;
    if (!std::isfinite(d)) return d;
    minimum = std::min(minimum, d);
  }
  return minimum;
}

// buildJGS2InputVertexColoring (jgs2-solver.ts:308-367): conflict groups of
// every tetrahedron, every positive cubature record's {vertex} + tet, every
// cloth triangle and hinge, and every IPC candidate.
jgs2_status color_vertices(const jgs2_scene_input& in, std::vector<uint32_t>& colors, uint32_t& count) {
  const uint32_t V = in.vertex_count, T = in.tet_count, K = in.cubature_k;
  std::vector<uint32_t> offsets, members;
  offsets.reserve(size_t(T) + size_t(V) * K + 1);
  members.reserve(4 * size_t(T) + 5 * size_t(V) * K);
  offsets.push_back(0);
  for (uint32_t t = 0; t < T; ++t) {
    for (int l = 0; l < 4; ++l) members.push_back(in.tet_indices[4 * size_t(t) + l]);
    offsets.push_back(uint32_t(members.size()));
  }
  for (uint32_t v = 0; v < V; ++v)
    for (uint32_t k = 0; k < K; ++k) {
      size_t r = size_t(v) * K + k;
      uint32_t t = in.cubature_tet_ids[r];
      if (t == 0xffffffffu || in.cubature_weights[r] <= 0) continue;
      members.push_back(v);
      for (int l = 0; l < 4; ++l) members.push_back(in.tet_indices[4 * size_t(t) + l]);
      offsets.push_back(uint32_t(members.size()));
    }
  if (in.cloth) {  // every cloth triangle and hinge is a conflict group
    for (uint32_t t = 0; t < in.cloth->triangle_count; ++t) {
      for (int l = 0; l < 3; ++l) members.push_back(in.cloth->triangle_indices[3 * size_t(t) + l]);
      offsets.push_back(uint32_t(members.size()));
    }
    for (uint32_t h = 0; h < in.cloth->hinge_count; ++h) {
      for (int l = 0; l < 4; ++l) members.push_back(in.cloth->hinge_indices[4 * size_t(h) + l]);
      offsets.push_back(uint32_t(members.size()));
    }
  }
  if (in.contact) {  // every IPC candidate (jgs2-solver.ts:352-364)
    const uint32_t n = in.contact->vertex_triangle_count + in.contact->edge_edge_count;
    for (uint32_t c = 0; c < n; ++c) {
      for (int l = 0; l < 4; ++l) members.push_back(in.contact->packed_indices[4 * size_t(c) + l]);
      offsets.push_back(uint32_t(members.size()));
    }
  }
  colors.resize(V);
  return jgs2_greedy_vertex_coloring(V, offsets.data(), uint32_t(offsets.size() - 1), members.data(), colors.data(),
                                     &count);
}

// ---------------------------------------------------------------- device order
uint64_t spread_bits(uint64_t x) {  // 21 bits -> every third of 63
  x &= 0x1fffffull;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

// Device order of vertices and tetrahedra: caller order, or (on request)
// Morton order of rest positions and rest centroids with ties in caller
// order. Returns both directions of each map.
void morton_orders(const jgs2_scene_input& in, std::vector<uint32_t>& vnew, std::vector<uint32_t>& vold,
                   std::vector<uint32_t>& tnew, std::vector<uint32_t>& told) {
  const uint32_t V = in.vertex_count, T = in.tet_count;
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < V; ++v)
    for (int a = 0; a < 3; ++a) {
      mn[a] = std::min(mn[a], double(in.vertex_rest[4 * size_t(v) + a]));
      mx[a] = std::max(mx[a], double(in.vertex_rest[4 * size_t(v) + a]));
    }
  const double extent = std::max({mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2], 1e-30});
  auto code = [&](double x, double y, double z) {
    const double q = double((1u << 21) - 1) / extent;
    uint64_t ix = uint64_t(std::max(0.0, (x - mn[0]) * q));
    uint64_t iy = uint64_t(std::max(0.0, (y - mn[1]) * q));
    uint64_t iz = uint64_t(std::max(0.0, (z - mn[2]) * q));
    return spread_bits(ix) | spread_bits(iy) << 1 | spread_bits(iz) << 2;
  };
  std::vector<std::pair<uint64_t, uint32_t>> keys(V);
  for (uint32_t v = 0; v < V; ++v) {
    const float* r = &in.vertex_rest[4 * size_t(v)];
    keys[v] = {code(r[0], r[1], r[2]), v};
  }
  // Caller order by default (builders already emit grid-coherent ids);
  // JGS2HIP_DEVICE_ORDER=morton reorders unstructured inputs.
  const char* order = std::getenv("JGS2HIP_DEVICE_ORDER");
  const bool identity = !(order && std::strcmp(order, "morton") == 0);
  if (identity)
    for (uint32_t v = 0; v < V; ++v) keys[v].first = 0;
  std::sort(keys.begin(), keys.end());
  vold.resize(V);
  vnew.resize(V);
  for (uint32_t i = 0; i < V; ++i) {
    vold[i] = keys[i].second;
    vnew[keys[i].second] = i;
  }
  std::vector<std::pair<uint64_t, uint32_t>> tkeys(T);
  for (uint32_t t = 0; t < T; ++t) {
    double c[3] = {0, 0, 0};
    for (int l = 0; l < 4; ++l)
      for (int a = 0; a < 3; ++a) c[a] += 0.25 * in.vertex_rest[4 * size_t(in.tet_indices[4 * size_t(t) + l]) + a];
    tkeys[t] = {code(c[0], c[1], c[2]), t};
  }
  if (identity)
    for (uint32_t t = 0; t < T; ++t) tkeys[t].first = 0;
  std::sort(tkeys.begin(), tkeys.end());
  told.resize(T);
  tnew.resize(T);
  for (uint32_t i = 0; i < T; ++i) {
    told[i] = tkeys[i].second;
    tnew[tkeys[i].second] = i;
  }
}


// ---------------------------------------------------------------- batch arena
uint64_t blocks_of(uint64_t n) { return (n + kSceneBlock - 1) / kSceneBlock; }

// Lanes per vertex of a scene's local solve: the batch's lane count, or one
// thread per vertex on the co-rotated path or when the caller asks for it.
// Kept per scene so a rule that depends on the scene alone (never on the
// rest of the batch) can choose it; its results then do not depend on the
// batch it is in.
uint32_t scene_lanes(const jgs2_batch* b, const jgs2_solver* s) {
  (void)s;
  return (b->corotated || b->solve_serial_only) ? 1u : b->solve_lanes;
}


// Element terms that a scene holds (scene_kind): the Jacobi lane sweep of a
// batch with cloth or contact solves scenes without either with the
// tetrahedral kernel variant, the others with the batch's full variant, or
// with its variant without tetrahedral terms when no unpinned vertex has
// any (a sheet over a pinned tetrahedral collider).
uint32_t scene_kind(const jgs2_solver* s) {
  return ((s->NT > 0 || s->NH > 0) ? kKindCloth : 0u) | ((s->NC > 0 || s->bp) ? kKindContact : 0u) |
         (s->cub_used > 0 ? kKindCubature : 0u) | (s->free_tets ? kKindTets : 0u);
}

// Allocates every arena array for the configured capacities.
jgs2_status batch_allocate(jgs2_batch* b, bool dry = false) {
  const jgs2_batch_config& c = b->cfg;
  DeviceState& d = b->dev;
  uint64_t* by = &b->arena_bytes;
  const uint64_t Vc = std::max<uint64_t>(1, blocks_of(c.max_vertices)) * kSceneBlock;
  const uint64_t Tc = std::max<uint64_t>(1, blocks_of(c.max_tets)) * kSceneBlock;
  const uint64_t Rc = blocks_of(c.max_triangles) * kSceneBlock;
  const uint64_t Cc = blocks_of(c.max_candidates) * kSceneBlock;
  const uint64_t Hc = c.max_hinges, Bc = std::max<uint64_t>(1, c.max_bodies), S = std::max<uint32_t>(1, c.max_scenes);
  const uint64_t K = c.cubature_k;
  if (Vc > 0xffffff00ull || Tc * 4 > 0xffffff00ull || Cc * 4 > 0xffffff00ull || Rc * 3 > 0xffffff00ull ||
      Hc * 4 > 0xffffff00ull)
    return fail(JGS2_ERROR_RANGE, "Batch capacities exceed 32-bit indexing.");
  d.K = uint32_t(K);
  d.S = uint32_t(S);
  TRYS(dalloc(&d.vtx_rest, Vc, by, dry));
  TRYS(dalloc(&d.vtx_info, Vc, by, dry));
  TRYS(dalloc(&d.objectives, 2 * Vc, by, dry));
  TRYS(dalloc(&d.color_order, Vc, by, dry));
  TRYS(dalloc(&d.colors, Vc, by, dry));
  TRYS(dalloc(&d.pos[0], Vc, by, dry));
  TRYS(dalloc(&d.pos[1], Vc, by, dry));
  TRYS(dalloc(&d.predicted, Vc, by, dry));
  TRYS(dalloc(&d.velocity, Vc, by, dry));
  TRYS(dalloc(&d.old, Vc, by, dry));
  TRYS(dalloc(&d.vrot, 3 * Vc, by, dry));
  TRYS(dalloc(&d.vrot_next, 3 * Vc, by, dry));
  TRYS(dalloc(&d.vcub[0], 4 * Vc, by, dry));
  TRYS(dalloc(&d.vcub[1], 4 * Vc, by, dry));
  TRYS(dalloc(&d.final_update, Vc, by, dry));
  TRYS(dalloc(&d.local_diag, 4 * Vc, by, dry));
  TRYS(dalloc(&d.vertex_energy, Vc, by, dry));
  TRYS(dalloc(&d.ipc_energy, Vc, by, dry));
  if (Cc) TRYS(dalloc(&d.ipc_grad, Vc, by, dry));
  if (Cc) TRYS(dalloc(&d.ipc_local, 3 * Vc, by, dry));
  if (Cc) TRYS(dalloc(&d.ipc_erec, Cc, by, dry));
  if (K) TRYS(dalloc(&d.cub_local, 3 * Vc, by, dry));
  TRYS(dalloc(&d.conv, 5 * Vc, by, dry));
  TRYS(dalloc(&d.tri_row, Vc, by, dry));
  TRYS(dalloc(&d.hinge_row, Vc, by, dry));
  TRYS(dalloc(&d.cand_row, Vc, by, dry));
  TRYS(dalloc(&d.active_count, Vc, by, dry));
  TRYS(dalloc(&d.cub, 10 * Vc * std::max<uint64_t>(K, 1), by, dry));
  TRYS(dalloc(&d.tet_rec, 4 * Tc, by, dry));
  if (b->corotated) TRYS(dalloc(&d.tet_k, 144 * Tc, by, dry));
  TRYS(dalloc(&d.trot, 3 * Tc, by, dry));
  TRYS(dalloc(&d.tet_glob, 2 * Tc, by, dry));
  TRYS(dalloc(&d.adjacency, 4 * Tc, by, dry));
  TRYS(dalloc(&d.tri_ix, Rc, by, dry));
  TRYS(dalloc(&d.tri_inv, Rc, by, dry));
  TRYS(dalloc(&d.tri_area, Rc, by, dry));
  TRYS(dalloc(&d.tri_glob, 2 * Rc, by, dry));
  TRYS(dalloc(&d.tri_base, 3 * Rc, by, dry));
  TRYS(dalloc(&d.hinge_base, 4 * Hc, by, dry));
  TRYS(dalloc(&d.tri_elems, 3 * Rc, by, dry));
  TRYS(dalloc(&d.hinge_ix, Hc, by, dry));
  TRYS(dalloc(&d.hinge_rest, Hc, by, dry));
  TRYS(dalloc(&d.hinge_elems, 4 * Hc, by, dry));
  TRYS(dalloc(&d.cand, Cc, by, dry));
  TRYS(dalloc(&d.cand_meta, Cc, by, dry));
  TRYS(dalloc(&d.cand_keep, Cc, by, dry));
  for (float4** a : {&d.lag_n, &d.lag_w, &d.src_c, &d.src_w, &d.tgt_c, &d.tgt_w}) TRYS(dalloc(a, Cc, by, dry));
  if (Cc) TRYS(dalloc(&d.fric, 4 * Cc, by, dry));
  TRYS(dalloc(&d.src_d, Cc, by, dry));
  TRYS(dalloc(&d.tgt_d, Cc, by, dry));
  TRYS(dalloc(&d.cand_ids, 4 * Cc, by, dry));
  TRYS(dalloc(&d.active_ids, 4 * Cc, by, dry));
  TRYS(dalloc(&d.body_corr, Bc, by, dry));
  TRYS(dalloc(&d.body_rng, Bc, by, dry));
  TRYS(dalloc(&d.control, 4 * S, by, dry));
  TRYS(dalloc(&d.history, size_t(kHistoryCapacity) * 8 * S, by, dry));
  TRYS(dalloc(&d.contact_ctl, S, by, dry));
  TRYS(dalloc(&d.scene_scale, S, by, dry));
  TRYS(dalloc(&d.scene_lanes, S, by, dry));
  TRYS(dalloc(&d.scene_kind, S, by, dry));
  TRYS(dalloc(&d.kind_blocks, 2 * (Vc / kSceneBlock), by, dry));
  TRYS(dalloc(reinterpret_cast<char**>(&d.scene_conv), scene_conv_bytes(uint32_t(S)), by, dry));
  TRYS(dalloc(&d.scene_blk0, S, by, dry));
  TRYS(dalloc(&d.scene_blk1, S, by, dry));
  TRYS(dalloc(&d.scene_cnt, S, by, dry));
  TRYS(dalloc(&d.scene_end, S, by, dry));
  TRYS(dalloc(&d.vblock_scene, Vc / kSceneBlock, by, dry));
  TRYS(dalloc(&d.tblock_scene, Tc / kSceneBlock, by, dry));
  TRYS(dalloc(&d.rblock_scene, std::max<uint64_t>(1, Rc / kSceneBlock), by, dry));
  TRYS(dalloc(&d.cblock_scene, std::max<uint64_t>(1, Cc / kSceneBlock), by, dry));
  d.partial_capacity = partial_blocks_needed(uint32_t(Vc), uint32_t(Tc), uint32_t(Rc), uint32_t(Cc));
  TRYS(dalloc(reinterpret_cast<char**>(&d.partials), partial_bytes(d.partial_capacity), by, dry));
  // Broad-phase contact arrays, only for batches that may hold such scenes.
  const uint64_t Sc = blocks_of(c.max_surface_triangles) * kSceneBlock, Ec = blocks_of(c.max_surface_edges) * kSceneBlock;
  if (Sc * 3 > 0xffffff00ull || Ec * 2 > 0xffffff00ull) return fail(JGS2_ERROR_RANGE, "Batch capacities exceed 32-bit indexing.");
  const bool bp = Sc + Ec > 0;
  if (bp) {
    TRYS(dalloc(&d.bp_tri, Sc, by, dry));
    TRYS(dalloc(&d.bp_tbox, 2 * Sc, by, dry));
    TRYS(dalloc(&d.bp_tchunk, 2 * (Sc / 64), by, dry));
    TRYS(dalloc(&d.bp_tsuper, 2 * (Sc / 64), by, dry));
    TRYS(dalloc(&d.bp_edge, Ec, by, dry));
    TRYS(dalloc(&d.bp_ebox, 2 * Ec, by, dry));
    TRYS(dalloc(&d.bp_echunk, 2 * (Ec / 64), by, dry));
    TRYS(dalloc(&d.bp_esuper, 2 * (Ec / 64), by, dry));
    TRYS(dalloc(&d.bp_ecnt, Ec, by, dry));
    TRYS(dalloc(&d.bp_eprev, Ec, by, dry));
    TRYS(dalloc(&d.bp_blk_e, Ec / kSceneBlock, by, dry));
    TRYS(dalloc(&d.ring_row, Vc, by, dry));
    TRYS(dalloc(&d.ring_ids, 2 * Ec, by, dry));
    for (uint32_t** a : {&d.bp_qv, &d.bp_vcnt, &d.bp_vprev, &d.bp_rowcnt, &d.bp_rowfill}) TRYS(dalloc(a, Vc, by, dry));
    for (float4** a : {&d.bp_x, &d.bp_cpose, &d.bp_tval}) TRYS(dalloc(a, Vc, by, dry));
    TRYS(dalloc(&d.bp_blk_v, Vc / kSceneBlock, by, dry));
    TRYS(dalloc(&d.bp_blk_r, Vc / kSceneBlock, by, dry));
    TRYS(dalloc(&d.cand_key, Cc, by, dry));
    TRYS(dalloc(&d.prev_key, Cc, by, dry));
    for (float4** a : {&d.prev_lag_n, &d.prev_lag_w, &d.prev_src_c, &d.prev_src_w}) TRYS(dalloc(a, Cc, by, dry));
    TRYS(dalloc(&d.bp_ctl, S, by, dry));
    TRYS(dalloc(&d.bp_reach, S, by, dry));
    TRYS(dalloc(&d.bp_size, S, by, dry));
    TRYS(dalloc(&d.bp_rows_at, S, by, dry));
    TRYS(dalloc(&d.bp_motion, 3 * S, by, dry));
    TRYS(dalloc(&d.bp_bodies, S, by, dry));
    TRYS(dalloc(&d.bp_bref, Bc, by, dry));
    TRYS(dalloc(&d.bp_bdev, Bc, by, dry));
    TRYS(dalloc(&d.bp_bbox, 2 * Bc, by, dry));
    TRYS(dalloc(&d.scene_blk2, S, by, dry));
    TRYS(dalloc(&d.sblock_scene, Sc / kSceneBlock, by, dry));
    TRYS(dalloc(&d.eblock_scene, Ec / kSceneBlock, by, dry));
    TRYS(dalloc(&d.bp_list, (Vc + Sc + Ec + Cc) / kSceneBlock, by, dry));
    TRYS(dalloc(&d.bp_work, (Vc + Sc + Ec + Cc) / kSceneBlock, by, dry));
    TRYS(dalloc(&d.bp_work_n, 8, by, dry));
    TRYS(dalloc(&d.bp_long, Vc, by, dry));
  }
  if (dry) return JGS2_OK;
  // Free blocks belong to no scene; scene slots start inactive.
  hipStream_t q = b->stream;
  std::vector<uint32_t> none(std::max({Vc, Tc, Rc, Cc}) / kSceneBlock + 1, kNoScene);
  HIP_CHECK(hipMemcpyAsync(d.vblock_scene, none.data(), 4 * (Vc / kSceneBlock), hipMemcpyHostToDevice, q));
  HIP_CHECK(hipMemcpyAsync(d.tblock_scene, none.data(), 4 * (Tc / kSceneBlock), hipMemcpyHostToDevice, q));
  if (Rc) HIP_CHECK(hipMemcpyAsync(d.rblock_scene, none.data(), 4 * (Rc / kSceneBlock), hipMemcpyHostToDevice, q));
  if (Cc) HIP_CHECK(hipMemcpyAsync(d.cblock_scene, none.data(), 4 * (Cc / kSceneBlock), hipMemcpyHostToDevice, q));
  std::vector<uint4> zero_cnt(S, make_uint4(0, 0, 0, 0));
  HIP_CHECK(hipMemcpyAsync(d.scene_cnt, zero_cnt.data(), S * sizeof(uint4), hipMemcpyHostToDevice, q));
  // Hinge entries no scene owns reference no vertex (the element pass skips them).
  HIP_CHECK(hipMemsetAsync(d.hinge_ix, 0xff, std::max<uint64_t>(Hc, 1) * sizeof(uint4), q));
  std::vector<uint4> no_body(Bc, make_uint4(0, 0, 0, kNoScene));
  HIP_CHECK(hipMemcpyAsync(d.body_rng, no_body.data(), Bc * sizeof(uint4), hipMemcpyHostToDevice, q));
  if (bp) {
    std::vector<uint32_t> no_scene(std::max(Sc, Ec) / kSceneBlock + 1, kNoScene);
    if (Sc) HIP_CHECK(hipMemcpyAsync(d.sblock_scene, no_scene.data(), 4 * (Sc / kSceneBlock), hipMemcpyHostToDevice, q));
    if (Ec) HIP_CHECK(hipMemcpyAsync(d.eblock_scene, no_scene.data(), 4 * (Ec / kSceneBlock), hipMemcpyHostToDevice, q));
    HIP_CHECK(hipMemsetAsync(d.bp_ctl, 0, S * sizeof(uint4), q));
    HIP_CHECK(hipMemsetAsync(d.bp_motion, 0, 3 * S * sizeof(uint4), q));
  }
  HIP_CHECK(hipStreamSynchronize(q));
  b->va.init(Vc / kSceneBlock);
  b->ta.init(Tc / kSceneBlock);
  b->ra.init(Rc / kSceneBlock);
  b->ca.init(Cc / kSceneBlock);
  b->adja.init(4 * Tc);
  b->hina.init(Hc);
  b->tela.init(3 * Rc);
  b->hela.init(4 * Hc);
  b->cela.init(4 * Cc);
  b->bodya.init(Bc);
  b->slota.init(S);
  b->sa.init(Sc / kSceneBlock);
  b->ea.init(Ec / kSceneBlock);
  b->rga.init(2 * Ec);
  b->slots.assign(S, nullptr);
  return JGS2_OK;
}

void release_spans(jgs2_batch* b, jgs2_solver* s) {
  b->va.release(s->vblk.at, s->vblk.n);
  b->ta.release(s->tblk.at, s->tblk.n);
  b->ra.release(s->rblk.at, s->rblk.n);
  b->ca.release(s->cblk.at, s->cblk.n);
  b->adja.release(s->adj.at, s->adj.n);
  b->hina.release(s->hin.at, s->hin.n);
  b->tela.release(s->tel.at, s->tel.n);
  b->hela.release(s->hel.at, s->hel.n);
  b->cela.release(s->cel.at, s->cel.n);
  b->bodya.release(s->body.at, s->body.n);
  b->slota.release(s->slot, 1);
  b->sa.release(s->sblk.at, s->sblk.n);
  b->ea.release(s->eblk.at, s->eblk.n);
  b->rga.release(s->rng.at, s->rng.n);
}

template <class T> hipError_t upload(T* base, uint64_t at, const std::vector<T>& v, hipStream_t q) {
  if (v.empty()) return hipSuccess;
  return hipMemcpyAsync(base + at, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, q);
}

// Packs one validated scene into freshly allocated spans of the arena
// (packJGS2GpuInput, src/scenes/index.ts:975-1109, with the cloth arena and
// IPC buffers of cloth-layout.ts and ipc-contact-layout.ts), in device order
// and with arena-absolute indices.
jgs2_status pack_into_arena(jgs2_batch* b, jgs2_solver* s, const jgs2_scene_input& in) {
  DeviceState& d = b->dev;
  hipStream_t q = b->stream;
  const uint32_t V = s->V, T = s->T, K = b->cfg.cubature_k, NT = s->NT, NH = s->NH, NC = s->NC;
  const uint64_t VB = s->vbase(), TB = s->tbase(), RB = s->rbase(), CB = s->cbase();
  const uint64_t Vp = s->vblk.n * kSceneBlock;  // padded vertex count
  const std::vector<uint32_t>& vnew = s->vnew;
  const std::vector<uint32_t>& vold = s->vold;
  const std::vector<uint32_t>& tnew = s->tnew;
  const std::vector<uint32_t>& told = s->told;
  // Vertex records; padding vertices are pinned, massless and unattached.
  std::vector<float4> vtx_rest(Vp, make_float4(0, 0, 0, 0)), positions(Vp, make_float4(0, 0, 0, 1)),
      velocities(Vp, make_float4(0, 0, 0, 0)), rot_v(3 * Vp), zeros4(Vp, make_float4(0, 0, 0, 0));
  std::vector<uint4> vtx_info(Vp, make_uint4(uint32_t(s->adj.at), 0, 1, kNoScene));
  std::vector<uint32_t> adjacency(s->adj.n);
  {
    uint32_t cursor = 0;
    for (uint32_t nv = 0; nv < V; ++nv) {
      const uint32_t ov = vold[nv];
      const float* r = &in.vertex_rest[4 * size_t(ov)];
      const uint32_t* i = &in.vertex_info[4 * size_t(ov)];
      vtx_rest[nv] = make_float4(r[0], r[1], r[2], r[3]);
      vtx_info[nv] = make_uint4(uint32_t(s->adj.at) + cursor, i[1], i[2], uint32_t(s->body.at) + i[3]);
      // Rows stay ascending in caller tet ids (the reference's summation
      // order); entries are arena tet ids.
      for (uint32_t k = 0; k < i[1]; ++k) adjacency[cursor + k] = uint32_t(TB) + tnew[in.adjacency[i[0] + k]];
      cursor += i[1];
      if (i[2] == 0u && i[1] > 0u) s->free_tets = true;
      const float* x = &in.positions[4 * size_t(ov)];
      positions[nv] = make_float4(x[0], x[1], x[2], 1.0f);
      if (in.velocities)
        velocities[nv] = make_float4(in.velocities[4 * size_t(ov)], in.velocities[4 * size_t(ov) + 1],
                                     in.velocities[4 * size_t(ov) + 2], 0.0f);
    }
  }
  for (uint64_t v = 0; v < Vp; ++v) {
    rot_v[3 * v] = make_float4(1, 0, 0, 0);
    rot_v[3 * v + 1] = make_float4(0, 1, 0, 0);
    rot_v[3 * v + 2] = make_float4(0, 0, 1, 0);
  }
  std::vector<float4> dev_objectives(2 * Vp, make_float4(0, 0, 0, 0));
  for (uint32_t nv = 0; nv < V; ++nv) {
    const float* o = &s->objectives[8 * size_t(vold[nv])];
    dev_objectives[2 * size_t(nv)] = make_float4(o[0], o[1], o[2], o[3]);
    dev_objectives[2 * size_t(nv) + 1] = make_float4(o[4], o[5], o[6], o[7]);
  }
  // Tetrahedra.
  std::vector<uint4> tet_rec(4 * size_t(T));
  std::vector<float4> rot_t(3 * size_t(T));
  auto bits4 = [](float a, float b, float c, float d) {
    uint4 u;
    std::memcpy(&u.x, &a, 4);
    std::memcpy(&u.y, &b, 4);
    std::memcpy(&u.z, &c, 4);
    std::memcpy(&u.w, &d, 4);
    return u;
  };
  std::vector<float> tet_k;
  if (s->corotated) tet_k.resize(144 * size_t(T));
  for (uint32_t nt = 0; nt < T; ++nt) {
    const uint32_t ot = told[nt];
    const uint32_t* i = &in.tet_indices[4 * size_t(ot)];
    const float* m = &in.tet_inverse_dm[12 * size_t(ot)];
    const float* a = &in.tet_meta[4 * size_t(ot)];
    tet_rec[4 * size_t(nt)] = make_uint4(uint32_t(VB) + vnew[i[0]], uint32_t(VB) + vnew[i[1]], uint32_t(VB) + vnew[i[2]],
                             uint32_t(VB) + vnew[i[3]]);
    tet_rec[4 * size_t(nt) + 1] = bits4(m[0], m[1], m[2], a[0]);
    tet_rec[4 * size_t(nt) + 2] = bits4(m[4], m[5], m[6], a[1]);
    tet_rec[4 * size_t(nt) + 3] = bits4(m[8], m[9], m[10], a[2]);
    if (s->corotated) std::memcpy(&tet_k[144 * size_t(nt)], &in.tet_rest_stiffness[144 * size_t(ot)], 144 * sizeof(float));
    rot_t[3 * size_t(nt)] = make_float4(1, 0, 0, 0);
    rot_t[3 * size_t(nt) + 1] = make_float4(0, 1, 0, 0);
    rot_t[3 * size_t(nt) + 2] = make_float4(0, 0, 1, 0);
  }
  // Cubature records, K per vertex; unused records are empty.
  const uint32_t Kin = in.cubature_k;
  std::vector<uint4> cub(10 * Vp * std::max<uint32_t>(K, 1), make_uint4(0, 0, 0, 0));
  for (uint64_t r = 0; r < Vp * K; ++r) cub[10 * r] = make_uint4(0xffffffffu, 0, 0, 0);
  for (uint32_t nv = 0; nv < V; ++nv)
    for (uint32_t k = 0; k < Kin; ++k) {
      const size_t ro = size_t(vold[nv]) * Kin + k, rn = size_t(nv) * K + k;
      uint32_t wbits;
      std::memcpy(&wbits, &in.cubature_weights[ro], 4);
      const uint32_t t = in.cubature_tet_ids[ro];
      cub[10 * rn] = make_uint4(t < T ? uint32_t(TB) + tnew[t] : 0xffffffffu, wbits, 0, 0);
      if (t < T && in.cubature_weights[ro] > 0) {
        ++s->cub_used;
        if (in.vertex_info[4 * size_t(vold[nv]) + 2] == 0u) s->free_tets = true;
      }
      std::memcpy(&cub[10 * rn + 1], &in.cubature_basis[36 * ro], 36 * sizeof(float));
    }
  // Cloth records, vertex -> element CSR rows {begin, count}.
  std::vector<uint2> tri_row(Vp, make_uint2(uint32_t(s->tel.at), 0)), hinge_row(Vp, make_uint2(uint32_t(s->hel.at), 0)),
      cand_row(Vp, make_uint2(uint32_t(s->cel.at), 0));
  std::vector<uint4> tri_ix(NT), hinge_ix(NH), cand(NC);
  std::vector<float4> tri_inv(NT), hinge_rest(NH);
  std::vector<float> tri_area(NT);
  std::vector<uint32_t> tri_elems, hinge_elems, cand_ids, cand_meta(NC, 0);
  auto rows_of = [&](const uint32_t* idx, uint32_t width, uint32_t count, uint64_t el_base, uint64_t id_base,
                     std::vector<uint2>& rows, std::vector<uint32_t>& elems) -> jgs2_status {
    std::vector<uint32_t> crow(size_t(V) + 1), celem(size_t(count) * width);
    TRYS(jgs2_build_element_incidence(V, idx, width, count, crow.data(), celem.data()));
    elems.clear();
    for (uint32_t nv = 0; nv < V; ++nv) {
      const uint32_t ov = vold[nv];
      rows[nv] = make_uint2(uint32_t(el_base + elems.size()), crow[ov + 1] - crow[ov]);
      for (uint32_t k = crow[ov]; k < crow[ov + 1]; ++k) elems.push_back(uint32_t(id_base) + celem[k]);
    }
    return JGS2_OK;
  };
  if (in.cloth) {
    const jgs2_cloth_input& c = *in.cloth;
    for (uint32_t t = 0; t < NT; ++t) {
      const uint32_t* ix = &c.triangle_indices[3 * size_t(t)];
      const uint32_t owner = std::min(ix[0], std::min(ix[1], ix[2]));  // S:1838-1840
      tri_ix[t] = make_uint4(uint32_t(VB) + vnew[ix[0]], uint32_t(VB) + vnew[ix[1]], uint32_t(VB) + vnew[ix[2]],
                             uint32_t(VB) + vnew[owner]);
      const float* bb = &c.triangle_inverse_rest_bases[4 * size_t(t)];
      tri_inv[t] = make_float4(bb[0], bb[1], bb[2], bb[3]);
      tri_area[t] = c.triangle_rest_areas[t];
    }
    for (uint32_t h = 0; h < NH; ++h) {
      const uint32_t* ix = &c.hinge_indices[4 * size_t(h)];
      const uint32_t owner = std::min(std::min(ix[0], ix[1]), std::min(ix[2], ix[3]));  // S:1842-1844
      hinge_ix[h] = make_uint4(uint32_t(VB) + vnew[ix[0]], uint32_t(VB) + vnew[ix[1]], uint32_t(VB) + vnew[ix[2]],
                               uint32_t(VB) + vnew[ix[3]]);
      const uint32_t ob = uint32_t(VB) + vnew[owner];
      float of;
      std::memcpy(&of, &ob, 4);
      hinge_rest[h] = make_float4(c.hinge_rest_angles[h], c.hinge_rest_edge_lengths[h], of, 0.0f);
    }
    TRYS(rows_of(c.triangle_indices, 3, NT, s->tel.at, RB, tri_row, tri_elems));
    TRYS(rows_of(c.hinge_indices, 4, NH, s->hel.at, s->hin.at, hinge_row, hinge_elems));
  }
  if (NC > 0 && !s->bp) {
    for (uint32_t c = 0; c < NC; ++c) {
      const uint32_t* ix = &in.contact->packed_indices[4 * size_t(c)];
      cand[c] = make_uint4(uint32_t(VB) + vnew[ix[0]], uint32_t(VB) + vnew[ix[1]], uint32_t(VB) + vnew[ix[2]],
                           uint32_t(VB) + vnew[ix[3]]);
      // Bit 0: edge-edge; bits 1-2: the record slot of the energy owner,
      // the first unpinned vertex, else the first (S:742-748); bit 3: a
      // vertex is pinned.
      uint32_t owner = 0;
      while (owner < 4 && in.vertex_info[4 * size_t(ix[owner]) + 2] != 0u) ++owner;
      bool pinned = false;
      for (int l = 0; l < 4; ++l) pinned = pinned || in.vertex_info[4 * size_t(ix[l]) + 2] != 0u;
      cand_meta[c] = (c < s->NVT ? 0u : 1u) | ((owner < 4 ? owner : 0u) << 1) | (pinned ? 8u : 0u);
    }
    TRYS(rows_of(in.contact->packed_indices, 4, NC, s->cel.at, CB, cand_row, cand_ids));
  }
#define UP(base, at, vec) HIP_CHECK(upload((base), (at), (vec), q))
  UP(d.vtx_rest, VB, vtx_rest);
  UP(d.vtx_info, VB, vtx_info);
  UP(d.adjacency, s->adj.at, adjacency);
  UP(d.objectives, 2 * VB, dev_objectives);
  for (float4* region : {d.pos[0], d.pos[1], d.predicted, d.old}) UP(region, VB, positions);
  UP(d.velocity, VB, velocities);
  UP(d.vrot, 3 * VB, rot_v);
  UP(d.vrot_next, 3 * VB, rot_v);
  {  // Cubature vertex records at the initial pose (identity frames).
    std::vector<float4> vc(4 * Vp);
    for (uint64_t v = 0; v < Vp; ++v) {
      const float4 x = positions[v];
      const uint4 inf = vtx_info[v];
      const float n = float(inf.y > 1u ? inf.y : 1u);
      const float w = inf.z != 0u ? -1.0f : vtx_rest[v].w * float(1.0 / (s->defaults.timestep * s->defaults.timestep)) / n;
      // {x, w}, {R.c0, diff.x}, {R.c1, diff.y}, {R.c2, diff.z}; diff is
      // rewritten by the first polar pass.
      vc[4 * v] = make_float4(x.x, x.y, x.z, w);
      vc[4 * v + 1] = make_float4(1, 0, 0, 0);
      vc[4 * v + 2] = make_float4(0, 1, 0, 0);
      vc[4 * v + 3] = make_float4(0, 0, 1, 0);
    }
    UP(d.vcub[0], 4 * VB, vc);
    UP(d.vcub[1], 4 * VB, vc);
  }
  for (float4* region : {d.final_update, d.vertex_energy}) UP(region, VB, zeros4);
  {
    std::vector<float4> z(5 * Vp, make_float4(0, 0, 0, 0));
    HIP_CHECK(hipMemcpyAsync(d.local_diag + 4 * VB, z.data(), 4 * Vp * sizeof(float4), hipMemcpyHostToDevice, q));
    HIP_CHECK(hipMemcpyAsync(d.conv + 5 * VB, z.data(), 5 * Vp * sizeof(float4), hipMemcpyHostToDevice, q));
  }
  UP(d.cub, 10 * VB * std::max<uint32_t>(K, 1), cub);
  UP(d.tet_rec, 4 * TB, tet_rec);
  if (s->corotated) UP(d.tet_k, 144 * TB, tet_k);
  UP(d.trot, 3 * TB, rot_t);
  UP(d.tri_row, VB, tri_row);
  UP(d.hinge_row, VB, hinge_row);
  UP(d.cand_row, VB, cand_row);
  {
    std::vector<uint32_t> zc(Vp, 0);
    UP(d.active_count, VB, zc);
  }
  UP(d.tri_ix, RB, tri_ix);
  UP(d.tri_inv, RB, tri_inv);
  UP(d.tri_area, RB, tri_area);
  UP(d.tri_elems, s->tel.at, tri_elems);
  UP(d.hinge_ix, s->hin.at, hinge_ix);
  UP(d.hinge_rest, s->hin.at, hinge_rest);
  UP(d.hinge_elems, s->hel.at, hinge_elems);
  if (NC > 0 && !s->bp) {  // packIpcContactBuffer initial records (ipc-contact-layout.ts:302-341)
    std::vector<float4> z(NC, make_float4(0, 0, 0, 0)),
        none(NC, make_float4(0, 0, 0, -1));
    UP(d.cand, CB, cand);
    UP(d.cand_meta, CB, cand_meta);
    for (float4* a : {d.lag_n, d.lag_w, d.src_w, d.tgt_w}) UP(a, CB, z);
    UP(d.src_c, CB, none);
    std::vector<float2> summary(NC, make_float2(-1, 0));
    UP(d.src_d, CB, summary);
    UP(d.tgt_c, CB, none);
    UP(d.cand_ids, s->cel.at, cand_ids);
    UP(d.active_ids, s->cel.at, cand_ids);
  }
  // Bodies: every body of the scene scans the scene's vertex range.
  {
    std::vector<uint4> rng(s->body.n, make_uint4(uint32_t(VB), uint32_t(VB + V), s->objective_flags, s->slot));
    std::vector<float4> z(s->body.n, make_float4(0, 0, 0, 0));
    UP(d.body_rng, s->body.at, rng);
    UP(d.body_corr, s->body.at, z);
  }
  // Block tables and scene records.
  const uint32_t sl = s->slot;
  {
    std::vector<uint32_t> tab(std::max({s->vblk.n, s->tblk.n, s->rblk.n, s->cblk.n}), sl);
    HIP_CHECK(hipMemcpyAsync(d.vblock_scene + s->vblk.at, tab.data(), 4 * s->vblk.n, hipMemcpyHostToDevice, q));
    HIP_CHECK(hipMemcpyAsync(d.tblock_scene + s->tblk.at, tab.data(), 4 * s->tblk.n, hipMemcpyHostToDevice, q));
    if (s->rblk.n) HIP_CHECK(hipMemcpyAsync(d.rblock_scene + s->rblk.at, tab.data(), 4 * s->rblk.n, hipMemcpyHostToDevice, q));
    if (s->cblk.n) HIP_CHECK(hipMemcpyAsync(d.cblock_scene + s->cblk.at, tab.data(), 4 * s->cblk.n, hipMemcpyHostToDevice, q));
  }
  const uint4 blk0 = make_uint4(uint32_t(s->vblk.at), uint32_t(s->vblk.at + s->vblk.n), uint32_t(s->tblk.at),
                                uint32_t(s->tblk.at + s->tblk.n));
  const uint4 blk1 = make_uint4(uint32_t(s->rblk.at), uint32_t(s->rblk.at + s->rblk.n), uint32_t(s->cblk.at),
                                uint32_t(s->cblk.at + s->cblk.n));
  const uint4 cnt = make_uint4(V, T, s->color_count, s->bodies);
  const uint4 end = make_uint4(uint32_t(VB + V), uint32_t(TB + T), uint32_t(RB + NT), uint32_t(CB + (s->bp ? 0u : NC)));
  const float4 ctl = make_float4(0, 0, 0, 1);
  std::vector<float4> zero_ctl(4, make_float4(0, 0, 0, 0)), zero_hist(size_t(kHistoryCapacity) * 8, make_float4(0, 0, 0, 0));
  HIP_CHECK(hipMemcpyAsync(d.scene_blk0 + sl, &blk0, sizeof(uint4), hipMemcpyHostToDevice, q));
  HIP_CHECK(hipMemcpyAsync(d.scene_blk1 + sl, &blk1, sizeof(uint4), hipMemcpyHostToDevice, q));
  HIP_CHECK(hipMemcpyAsync(d.scene_end + sl, &end, sizeof(uint4), hipMemcpyHostToDevice, q));
  HIP_CHECK(hipMemcpyAsync(d.scene_scale + sl, &s->scene_scale, sizeof(float), hipMemcpyHostToDevice, q));
  const uint32_t lanes = scene_lanes(b, s);
  HIP_CHECK(hipMemcpyAsync(d.scene_lanes + sl, &lanes, sizeof(uint32_t), hipMemcpyHostToDevice, q));
  const uint32_t kind = scene_kind(s);
  HIP_CHECK(hipMemcpyAsync(d.scene_kind + sl, &kind, sizeof(uint32_t), hipMemcpyHostToDevice, q));
  HIP_CHECK(hipMemcpyAsync(d.contact_ctl + sl, &ctl, sizeof(float4), hipMemcpyHostToDevice, q));
  HIP_CHECK(hipMemcpyAsync(d.control + 4 * size_t(sl), zero_ctl.data(), 4 * sizeof(float4), hipMemcpyHostToDevice, q));
  HIP_CHECK(hipMemcpyAsync(d.history + size_t(sl) * kHistoryCapacity * 8, zero_hist.data(),
                           zero_hist.size() * sizeof(float4), hipMemcpyHostToDevice, q));
  HIP_CHECK(hipMemcpyAsync(d.scene_cnt + sl, &cnt, sizeof(uint4), hipMemcpyHostToDevice, q));
  if (d.bp_ctl) {  // broad-phase records (every slot of a batch that may hold such scenes)
    const uint4 ctl_bp = make_uint4(s->bp ? (kBpScene | kBpForce) : 0u, 0, 0, 0);
    const float4 reach = make_float4(0, s->bp_margin, 0, 0);
    const uint4 size = make_uint4(s->NS, s->NE, s->bp ? NC : 0u, 0);
    const uint4 blk2 = make_uint4(uint32_t(s->sblk.at), uint32_t(s->sblk.at + s->sblk.n), uint32_t(s->eblk.at),
                                  uint32_t(s->eblk.at + s->eblk.n));
    // Motion accumulators: maxima start at key 0, minima at the largest key.
    const uint4 zero2[3] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0),
                            make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu)};
    HIP_CHECK(hipMemcpyAsync(d.bp_ctl + sl, &ctl_bp, sizeof(uint4), hipMemcpyHostToDevice, q));
    HIP_CHECK(hipMemcpyAsync(d.bp_reach + sl, &reach, sizeof(float4), hipMemcpyHostToDevice, q));
    HIP_CHECK(hipMemcpyAsync(d.bp_size + sl, &size, sizeof(uint4), hipMemcpyHostToDevice, q));
    const uint32_t rows_at = uint32_t(s->cel.at);
    HIP_CHECK(hipMemcpyAsync(d.bp_rows_at + sl, &rows_at, sizeof(uint32_t), hipMemcpyHostToDevice, q));
    HIP_CHECK(hipMemcpyAsync(d.bp_motion + 3 * size_t(sl), zero2, sizeof(zero2), hipMemcpyHostToDevice, q));
    HIP_CHECK(hipMemcpyAsync(d.scene_blk2 + sl, &blk2, sizeof(uint4), hipMemcpyHostToDevice, q));
    // The body rule: up to 64 bodies, every surface primitive within one body.
    std::vector<uint32_t> bref(s->body.n, 0xffffffffu);  // no vertex
    bool bodies_ok = s->bp && s->body.n <= 64;
    for (uint32_t nv = 0; nv < V && bodies_ok; ++nv) {
      const uint32_t body = in.vertex_info[4 * size_t(vold[nv]) + 3];
      if (bref[body] == 0xffffffffu) bref[body] = uint32_t(VB) + nv;
    }
    if (bodies_ok) {
      auto body_of = [&](uint32_t v) { return in.vertex_info[4 * size_t(v) + 3]; };
      for (const auto& f : s->surface.tris)
        bodies_ok = bodies_ok && body_of(f[0]) == body_of(f[1]) && body_of(f[1]) == body_of(f[2]);
      for (const auto& e : s->surface.edges) bodies_ok = bodies_ok && body_of(e[0]) == body_of(e[1]);
    }
    const uint4 bodies = make_uint4(uint32_t(s->body.at), uint32_t(s->body.n), bodies_ok ? 1u : 0u, 0u);
    HIP_CHECK(hipMemcpyAsync(d.bp_bodies + sl, &bodies, sizeof(uint4), hipMemcpyHostToDevice, q));
    if (bodies_ok) {
      std::vector<uint4> zb(s->body.n, make_uint4(0, 0, 0, 0));
      UP(d.bp_bref, s->body.at, bref);
      UP(d.bp_bdev, s->body.at, zb);
    }
  }
  if (s->bp) {  // canonical surface in device ids, one-ring rows, query order
    const Surface& sf = s->surface;
    const uint64_t SB = s->sblk.at * kSceneBlock, EB = s->eblk.at * kSceneBlock;
    std::vector<uint4> tri(sf.tris.size());
    for (size_t t = 0; t < sf.tris.size(); ++t)
      tri[t] = make_uint4(uint32_t(VB) + vnew[sf.tris[t][0]], uint32_t(VB) + vnew[sf.tris[t][1]],
                          uint32_t(VB) + vnew[sf.tris[t][2]], 0u);
    std::vector<uint2> edge(sf.edges.size());
    for (size_t e = 0; e < sf.edges.size(); ++e)
      edge[e] = make_uint2(uint32_t(VB) + vnew[sf.edges[e][0]], uint32_t(VB) + vnew[sf.edges[e][1]]);
    std::vector<uint2> ring_row(Vp, make_uint2(uint32_t(s->rng.at), 0));
    std::vector<uint32_t> ring_ids(sf.ring.size()), qv(Vp, 0);
    for (uint32_t v = 0; v < V; ++v) {
      ring_row[vnew[v]] = make_uint2(uint32_t(s->rng.at) + sf.ring_start[v], sf.ring_start[v + 1] - sf.ring_start[v]);
      for (uint32_t k = sf.ring_start[v]; k < sf.ring_start[v + 1]; ++k) ring_ids[k] = uint32_t(VB) + vnew[sf.ring[k]];
      qv[v] = uint32_t(VB) + vnew[v];
    }
    UP(d.bp_tri, SB, tri);
    UP(d.bp_edge, EB, edge);
    UP(d.ring_row, VB, ring_row);
    UP(d.ring_ids, s->rng.at, ring_ids);
    UP(d.bp_qv, VB, qv);
    std::vector<uint32_t> tab(std::max<uint64_t>({s->sblk.n, s->eblk.n, 1}), sl);
    if (s->sblk.n) HIP_CHECK(hipMemcpyAsync(d.sblock_scene + s->sblk.at, tab.data(), 4 * s->sblk.n, hipMemcpyHostToDevice, q));
    if (s->eblk.n) HIP_CHECK(hipMemcpyAsync(d.eblock_scene + s->eblk.at, tab.data(), 4 * s->eblk.n, hipMemcpyHostToDevice, q));
  }
#undef UP
  HIP_CHECK(hipStreamSynchronize(q));
  return JGS2_OK;
}

// ---------------------------------------------------------------- pass parameters
PassParams make_params(jgs2_batch* b, const jgs2_settings& st, bool stop) {
  PassParams p{};
  b->dev.NC = uint32_t(b->ca.high() * kSceneBlock);
  uint32_t flags = 0;
  for (auto* s : b->slots)
    if (s) flags |= s->objective_flags;
  p.vertex_count = uint32_t(b->va.high() * kSceneBlock);
  p.tet_count = uint32_t(b->ta.high() * kSceneBlock);
  p.tri_count = uint32_t(b->ra.high() * kSceneBlock);
  p.hinge_count = uint32_t(b->hina.high());
  p.body_count = uint32_t(b->bodya.high());
  p.scene_count = uint32_t(b->slota.high());
  p.cubature_k = b->cfg.cubature_k;
  p.objective_flags = flags;
  p.globalization = b->corotated ? 0u : 1u;
  p.stop_enabled = (!b->corotated && stop) ? 1u : 0u;
  p.corotated = b->corotated ? 1u : 0u;
  // contactEnabled (jgs2-solver.ts:2213-2214).
  p.contact = (b->dev.NC > 0 && st.ipc_barrier_stiffness > 0) ? 1u : 0u;
  p.solve_block = b->solve_block;
  p.solve_lanes = b->solve_lanes;
  for (auto* s : b->slots)
    if (s) {
      p.solve_classes |= scene_lanes(b, s) > 1u ? 1u : 2u;
      if (scene_lanes(b, s) > 1u) p.solve_classes |= (scene_kind(s) & (kKindCloth | kKindContact)) == 0u ? 4u : 8u;
    }
  p.kind_blocks1 = b->kind_blocks_n[0];
  p.kind_blocks2 = b->kind_blocks_n[1];
  p.kind_blocks3 = b->kind_blocks_n[2];
  p.kind_blocks4 = b->kind_blocks_n[3];
  p.xcd_remap = b->xcd_remap;
  // packUniforms (jgs2-solver.ts:604-730): f64 expressions rounded to f32.
  const double inv = 1.0 / st.timestep;
  p.dt = round_f32(st.timestep);
  p.inv_dt = round_f32(inv);
  p.inv_dt2 = round_f32(inv * inv);
  p.max_step = round_f32(st.max_step);
  p.gx = round_f32(st.gravity[0]);
  p.gy = round_f32(st.gravity[1]);
  p.gz = round_f32(st.gravity[2]);
  p.floor_y = round_f32(st.floor_height);
  p.floor_k = round_f32(st.floor_stiffness);
  p.regularization = round_f32(st.regularization);
  p.polar_eps = round_f32(st.rotation_epsilon);
  p.damping = round_f32(st.velocity_damping);
  p.tangential_damping = round_f32(st.contact_tangential_damping);
  p.contact_margin = round_f32(st.contact_margin);
  p.ipc_dhat = round_f32(st.ipc_activation_distance);
  p.ipc_kappa = round_f32(st.ipc_barrier_stiffness);
  p.residual_tol = round_f32(st.residual_tolerance);
  p.update_tol = round_f32(st.normalized_update_tolerance);
  return p;
}

// The batch's color-sorted vertex list: color c holds every resident
// scene's color-c vertices, ascending in arena id (vertex-coloring.ts:92-107
// per scene).
jgs2_status sync_colors(jgs2_batch* b) {
  if (b->synced_version == b->version) return JGS2_OK;
  std::vector<jgs2_solver*> order;
  uint32_t cc = 1;
  for (auto* s : b->slots)
    if (s) {
      order.push_back(s);
      cc = std::max(cc, s->color_count);
    }
  std::sort(order.begin(), order.end(), [](jgs2_solver* a, jgs2_solver* c) { return a->vblk.at < c->vblk.at; });
  b->color_count = cc;
  b->color_offsets.assign(cc + 1, 0);
  std::vector<uint32_t> list;
  for (uint32_t c = 0; c < cc; ++c) {
    for (auto* s : order)
      if (c < s->color_count)
        list.insert(list.end(), s->color_order.begin() + s->color_offsets[c], s->color_order.begin() + s->color_offsets[c + 1]);
    b->color_offsets[c + 1] = uint32_t(list.size());
  }
  if (!list.empty()) {
    HIP_CHECK(hipMemcpyAsync(b->dev.color_order, list.data(), 4 * list.size(), hipMemcpyHostToDevice, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  }
  // Vertex blocks of the lane-swept scenes by kind, in arena order: the lane
  // sweep of a mixed batch launches over each list instead of every block.
  std::vector<uint32_t> kinds[4];
  for (auto* s : order) {
    if (scene_lanes(b, s) <= 1u) continue;
    const uint32_t k = scene_kind(s);
    const bool plain = (k & (kKindCloth | kKindContact)) == 0u;
    for (uint64_t i = 0; i < s->vblk.n; ++i) {
      const uint32_t blk = uint32_t(s->vblk.at + i);
      kinds[plain ? 0 : (k & kKindTets) ? 1 : 3].push_back(blk);
      if (!plain && (k & kKindCubature)) kinds[2].push_back(blk);
    }
  }
  std::vector<uint32_t> all;
  for (int i = 0; i < 4; ++i) {
    b->kind_blocks_n[i] = uint32_t(kinds[i].size());
    all.insert(all.end(), kinds[i].begin(), kinds[i].end());
  }
  if (!all.empty()) {
    HIP_CHECK(hipMemcpyAsync(b->dev.kind_blocks, all.data(), 4 * all.size(), hipMemcpyHostToDevice, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  }
  b->synced_version = b->version;
  b->graph_valid = false;
  return JGS2_OK;
}

void record_solve(jgs2_batch* b, const PassParams& p) {
  if (b->profiling && b->event_used + 2 <= b->events.size()) {
    (void)hipEventRecord(b->events[b->event_used], b->stream);
    launch_solve(b->dev, p, b->stream);
    (void)hipEventRecord(b->events[b->event_used + 1], b->stream);
    b->event_used += 2;
    return;
  }
  launch_solve(b->dev, p, b->stream);
}

// The source and target contact cache arrays and summaries trade roles at a
// promotion.
void swap_contact_caches(DeviceState& d) {
  std::swap(d.src_c, d.tgt_c);
  std::swap(d.src_w, d.tgt_w);
  std::swap(d.src_d, d.tgt_d);
}

// encodeFrame (jgs2-solver.ts:2207-2494), one launch per pass over every
// resident scene.
void enqueue_frame(jgs2_batch* b, const jgs2_settings& st, uint32_t iterations, bool stop) {
  const DeviceState& d = b->dev;
  hipStream_t q = b->stream;
  PassParams base = make_params(b, st, stop);
  const bool globalization = !b->corotated;
  const bool gs = st.schedule == JGS2_SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL;
  const bool cloth = base.tri_count != 0u || base.hinge_count != 0u;
  const bool contact = base.contact != 0u;
  base.src = 0;
  base.tgt = 0;
  // Jacobi sweeps read friction rows from per-candidate records written by
  // the active-row pass at the sweep's source pose.
  base.fric_records = (contact && !gs) ? 1u : 0u;
  uint32_t cache_swaps = 0;
  if (globalization) launch_clear_frame(d, base, q);
  launch_predict(d, base, q);
  if (contact) launch_broad_phase(d, base, true, false, q);  // lists complete at the frame start
  if (contact) launch_lag_contact(d, base, q);
  if (globalization) launch_reduce_candidate(d, base, q);  // source feasibility preflight
  for (uint32_t it = 0; it < iterations; ++it) {
    PassParams p = base;
    p.src = (it % 2 == 0) ? 1u : 0u;
    p.tgt = (it % 2 == 0) ? 0u : 1u;
    if (contact && it > 0) launch_broad_phase(d, p, true, true, q);  // and at each later sweep's source
    if (contact) launch_build_active_rows(d, p, q);
    if (gs) {
      launch_copy_position(d, p, q);
      for (uint32_t c = 0; c < b->color_count; ++c) {
        PassParams pc = p;
        pc.src = p.tgt;
        pc.colored = 1;
        pc.color_index = c;
        pc.color_begin = b->color_offsets[c];
        pc.color_end = b->color_offsets[c + 1];
        if (b->corotated) launch_tet_polar(d, pc, q);
        launch_vertex_polar(d, pc, q);
        record_solve(b, pc);
      }
      if (globalization && !contact) launch_vertex_energy(d, p, q);
    } else {
      // Sweeps after the first read the frames (and cloth element records)
      // the previous convergence stage computed at this sweep's source pose.
      const bool fused = globalization && it > 0;
      if (b->corotated) launch_tet_polar(d, p, q);
      if (!fused) launch_vertex_polar(d, p, q);
      if (!fused && globalization && cloth) launch_cloth_base(d, p, p.src, q);
      PassParams ps = p;
      ps.fused_rot = fused ? 1u : 0u;
      ps.cloth_base = (globalization && cloth) ? 1u : 0u;
      ps.ipc_local = contact ? 1u : 0u;
      record_solve(b, ps);
      if (globalization && cloth && !contact) {
        // Element records move to the candidate (keeping the source
        // energies), so the energy pair reads both from them.
        launch_cloth_base(d, p, p.tgt, q, 1u);
        PassParams pe = p;
        pe.energy_after_sweep = 1;
        pe.cloth_base = 1;
        pe.target_records = 1;
        launch_vertex_energy(d, pe, q);
      }
    }
    if (contact) {  // safe step, validation, then the energy pair (jgs2-solver.ts:2347-2400)
      launch_contact_step(d, p, q);
      if (globalization) {
        PassParams pe = p;
        if (!gs && cloth) {  // as above, at the candidate after the contact step
          launch_cloth_base(d, p, p.tgt, q, 1u);
          pe.cloth_base = 1;
          pe.target_records = 1;
        }
        launch_vertex_energy(d, pe, q);
      }
    }
    if (globalization) {
      launch_reduce_candidate(d, p, q);
      launch_apply_candidate(d, p, q);
      if (contact) {
        // Promotion by renaming: the validated target caches become the
        // source arrays for every later pass of the frame.
        swap_contact_caches(b->dev);
        ++cache_swaps;
        launch_promote_contact(d, p, q);
      }
      PassParams pc = p;
      pc.compute_next_rot = (!gs && it + 1 < iterations) ? 1u : 0u;
      if (!gs && cloth) {
        // The records hold the candidate; scenes that rejected it are back
        // at the source pose and recompute theirs. This pass and the next
        // sweep read them.
        launch_cloth_base(d, p, p.tgt, q, 2u);
        pc.cloth_base = 1;
      }
      launch_convergence(d, pc, q);
    }
  }
  if (cache_swaps % 2 == 1) {  // the frame ends with the cache arrays in their roles at its start
    launch_contact_cache_home(d, q);
    swap_contact_caches(b->dev);
  }
  if (iterations % 2 == 0) {
    PassParams p = base;
    p.src = 1;
    p.tgt = 0;
    launch_copy_position(d, p, q);
  }
  if (st.horizontal_body_correction) launch_body_correction(d, base, q);
  launch_finalize(d, base, q);
}

uint64_t fnv(uint64_t h, const void* data, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  return h;
}

jgs2_status check_batch(jgs2_batch* b) {
  if (!b) return fail(JGS2_ERROR_TYPE, "NULL batch.");
  return JGS2_OK;
}

// resolveStepSettings (jgs2-solver.ts:1982-2009) for the batch.
jgs2_status resolve_batch_step(jgs2_batch* b, const jgs2_settings* ov, uint64_t mask, const char* api,
                               jgs2_settings& out) {
  if (ov && (mask & JGS2_SET_TIMESTEP)) {
    const double pre = b->defaults.timestep, run = ov->timestep;
    if (!(std::isfinite(run) && run > 0 && float(pre) == float(run)))
      return fail(JGS2_ERROR_RANGE,
                  "JGS2 was precomputed for timestep %.17g, but %s() requested %.17g. Regenerate the precomputation "
                  "before changing timestep.",
                  pre, api, run);
  }
  bool any_contact = false;
  for (auto* s : b->slots)
    if (s) any_contact |= s->NC > 0;
  if (any_contact && ov && (mask & JGS2_SET_IPC_MINIMUM_DISTANCE) &&
      float(ov->ipc_minimum_distance) != float(b->defaults.ipc_minimum_distance))
    return fail(JGS2_ERROR_RANGE,
                "ipcMinimumDistance is fixed when an IPC solver is created; recreate the solver before changing it.");
  jgs2_status st = merge_settings(b->defaults, ov, mask, out);
  if (st != JGS2_OK) return st;
  // Objectives and mesh contact keep their solved centre-of-mass motion
  // (jgs2-solver.ts:1994-2008); the body-correction kernels apply that rule
  // per scene.
  return JGS2_OK;
}

uint32_t normalize_odd(int32_t iterations) {  // normalizeOddIterationCount (layout.ts:742-748)
  int32_t i = std::max(1, iterations);
  return uint32_t(i % 2 == 0 ? i + 1 : i);
}

// submitFrames (jgs2-solver.ts:2026-2072): limits, then the frames, replayed
// from a captured hipGraph when graphs are enabled.
jgs2_status submit(jgs2_batch* b, const jgs2_settings& st, uint32_t iterations, uint32_t frames, bool stop) {
  TRYS(sync_colors(b));
  if (!b->corotated) {
    if (iterations > kHistoryCapacity)
      return fail(JGS2_ERROR_RANGE, "Stable JGS2 records at most %u nonlinear iterations per frame; got %u.",
                  kHistoryCapacity, iterations);
    const uint32_t sub = st.schedule == JGS2_SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL ? b->color_count : 1u;
    if (uint64_t(iterations) * sub > uint64_t(JGS2_MAX_GLOBALIZED_ITERATIONS_PER_SUBMISSION / frames))
      return fail(JGS2_ERROR_RANGE,
                  "Stable JGS2 may encode at most %u total nonlinear iterations per submission; got %u frames * %u "
                  "iterations.",
                  JGS2_MAX_GLOBALIZED_ITERATIONS_PER_SUBMISSION, frames, iterations * sub);
  }
  if (b->active_scenes() == 0) return fail(JGS2_ERROR_STATE, "The batch holds no scene.");
  if (st.schedule == JGS2_SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL && b->dev.bp_nv)
    return fail(JGS2_ERROR_RANGE, "Broad-phase contact scenes step with the Jacobi schedule only.");
  if (b->use_graphs && !b->profiling) {
    uint64_t key = 1469598103934665603ull;
    key = fnv(key, &b->version, sizeof(b->version));
    key = fnv(key, &st, sizeof(st));
    key = fnv(key, &iterations, sizeof(iterations));
    const uint32_t flags[4] = {stop ? 1u : 0u, b->solve_block, b->xcd_remap, b->solve_lanes};
    key = fnv(key, flags, sizeof(flags));
    for (auto* s : b->slots) {
      const uint32_t f = s ? s->objective_flags : 0u;
      key = fnv(key, &f, sizeof(f));
    }
    if (!b->graph_valid || key != b->graph_key) {
      // Earlier launches of the exec may still be queued on the stream: it
      // is changed or destroyed only once they have completed.
      if (b->graph_exec) HIP_CHECK(hipStreamSynchronize(b->stream));
      b->graph_valid = false;
      hipGraph_t graph = nullptr;
      HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      enqueue_frame(b, st, iterations, stop);
      HIP_CHECK(hipStreamEndCapture(b->stream, &graph));
      // A mix that changed only in extents and arguments (an admission or
      // retirement of scenes like the resident ones) keeps the pass
      // sequence: the new capture is applied to the existing exec in place.
      // Any other change (a kernel variant, an iteration count, a color
      // count) fails the update, and the exec is rebuilt.
      bool updated = false;
      if (b->graph_exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult why = hipGraphExecUpdateSuccess;
        updated = hipGraphExecUpdate(b->graph_exec, graph, &bad, &why) == hipSuccess && why == hipGraphExecUpdateSuccess;
        (void)hipGetLastError();
        if (!updated) {
          (void)hipGraphExecDestroy(b->graph_exec);
          b->graph_exec = nullptr;
        }
      }
      hipError_t e = hipSuccess;
      if (!updated) e = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        b->graph_exec = nullptr;
        return fail(JGS2_ERROR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
      }
      (updated ? b->graph_updates : b->graph_instantiations) += 1;
      b->graph_key = key;
      b->graph_valid = true;
    }
    for (uint32_t f = 0; f < frames; ++f) HIP_CHECK(hipGraphLaunch(b->graph_exec, b->stream));
  } else {
    for (uint32_t f = 0; f < frames; ++f) enqueue_frame(b, st, iterations, stop);
  }
  HIP_CHECK(hipGetLastError());
  b->stepped_version = b->version;
  clear_error();
  return JGS2_OK;
}

jgs2_status validate_frames(uint32_t frames) {
  if (frames < 1 || frames > JGS2_MAX_BATCH_FRAMES)
    return fail(JGS2_ERROR_RANGE, "frameCount must be a positive safe integer no greater than %d; got %u.",
                JGS2_MAX_BATCH_FRAMES, frames);
  return JGS2_OK;
}

void recount_objectives(jgs2_solver* s) {  // computePackedObjectiveActivity (jgs2-solver.ts:233-262)
  uint32_t forces = 0, targets = 0;
  for (uint32_t v = 0; v < s->V; ++v) {
    const float* o = &s->objectives[8 * size_t(v)];
    if (o[0] != 0 || o[1] != 0 || o[2] != 0) ++forces;
    if (o[7] > 0) ++targets;
  }
  s->active_forces = forces;
  s->active_targets = targets;
  s->objective_flags = (forces ? 1u : 0u) | (targets ? 2u : 0u);
}

// Broad-phase contact of a scene: canonical surface, reach margin, the
// initially feasible pose (the check of jgs2-solver.ts:1041-1066 over the
// pairs that can be that close) and the candidate capacity.
jgs2_status prepare_broad_phase(const jgs2_scene_input& in, const jgs2_settings& st, jgs2_solver* s) {
  const jgs2_contact_input& c = *in.contact;
  const uint32_t V = in.vertex_count;
  TRYS(canonical_surface(V, c.surface_triangles, c.surface_triangle_count, c.surface_edges, c.surface_edge_count,
                         s->surface));
  for (uint32_t v = 0; v < V; ++v) {
    const uint32_t n = s->surface.ring_start[v + 1] - s->surface.ring_start[v];
    if (n > 64)
      return fail(JGS2_ERROR_RANGE,
                  "Broad-phase contact supports surface vertices with at most 64 neighbors; vertex %u has %u.", v, n);
  }
  if (!(std::isfinite(c.reach_margin) && c.reach_margin >= 0))
    return fail(JGS2_ERROR_RANGE, "reach_margin must be finite and nonnegative.");
  if (!superset_nonempty(s->surface, V)) return JGS2_OK;  // no pair can ever touch: no contact
  const double rr = std::max({double(s->scene_scale), st.ipc_activation_distance, 1e-12}) * 0x1p-24 * 8.0;
  const double act = st.max_step > 0 ? st.ipc_activation_distance + 2.0 * (st.max_step + rr) : INFINITY;
  const double margin = c.reach_margin > 0 ? c.reach_margin : 0.5 * act;
  const double near = st.ipc_minimum_distance + 8.0 * 0x1p-24 * std::max(1.0, st.ipc_activation_distance);
  const float distance_reach = float(2.0 * near + 1e-6);
  // Count the pairs within the first build's reach only when sizing the
  // capacity; the feasibility check needs the nearest pairs alone.
  const SurfaceScan scan = scan_surface_pairs(s->surface, V, in.positions,
                                              c.candidate_capacity ? distance_reach : float(act + margin), distance_reach);
  if (std::isfinite(scan.min_distance)) {
    const double m = 8.0 * 0x1p-24 * std::max({1.0, st.ipc_activation_distance, std::fabs(scan.min_distance)});
    const double required = st.ipc_minimum_distance + m;
    if (!(scan.min_distance > required))
      return fail(JGS2_ERROR_RANGE,
                  "IPC candidates require an initially feasible source pose with distance greater than %.17g "
                  "(minimum plus an f32 safety margin); got %.17g.",
                  required, scan.min_distance);
  }
  const uint64_t estimate = 2 * (scan.vertex_triangle + scan.edge_edge) + 1024;
  const uint64_t capacity = c.candidate_capacity ? c.candidate_capacity : estimate;
  if (capacity > 0x3fffffffull)
    return fail(JGS2_ERROR_RANGE, "Broad-phase candidate capacity %llu exceeds 32-bit indexing.",
                (unsigned long long)capacity);
  s->bp = true;
  s->NC = uint32_t(capacity);
  s->NS = uint32_t(s->surface.tris.size());
  s->NE = uint32_t(s->surface.edges.size());
  s->bp_margin = float(margin);
  return JGS2_OK;
}

// Validates a scene input and fills the host-side scene record (everything
// but the arena spans) under the batch's resolved settings. Runs entirely
// on the host, before any device is touched (layout.ts:223-416,
// jgs2-solver.ts:940-1066).
jgs2_status prepare_scene(const jgs2_scene_input& in, const jgs2_settings& resolved, jgs2_solver* s) {
  TRYS(validate_input(in));
  const uint32_t V = in.vertex_count, T = in.tet_count, K = in.cubature_k;
  const bool corotated = in.tet_meta[3] == JGS2_MATERIAL_COROTATED_LINEAR;
  const bool globalization = !corotated;
  const uint32_t NT = in.cloth ? in.cloth->triangle_count : 0u, NH = in.cloth ? in.cloth->hinge_count : 0u;
  if (in.cloth) {
    TRYS(validate_cloth(*in.cloth, V));
    if (NT > 0 && !globalization)  // jgs2-solver.ts:998-1001
      return fail(JGS2_ERROR_RANGE, "Triangle cloth requires the stable Neo-Hookean production path.");
  }
  const uint32_t NC = in.contact ? in.contact->vertex_triangle_count + in.contact->edge_edge_count : 0u;
  const bool bp_mode = in.contact && NC == 0 &&
                       in.contact->surface_triangle_count + in.contact->surface_edge_count > 0;
  if (in.contact && NC > 0 && in.contact->surface_triangle_count + in.contact->surface_edge_count > 0)
    return fail(JGS2_ERROR_RANGE, "IPC contact takes either packed candidates or a broad-phase surface, not both.");
  if (in.contact) TRYS(validate_contact(*in.contact, V));
  if (NC > 0) {  // initially feasible contact distances (jgs2-solver.ts:1041-1066)
    const double minimum = min_static_contact_distance(in.positions, *in.contact);
    const double margin = 8.0 * 0x1p-24 * std::max({1.0, resolved.ipc_activation_distance, std::fabs(minimum)});
    const double required = resolved.ipc_minimum_distance + margin;
    if (!std::isfinite(minimum) || !(minimum > required))
      return fail(JGS2_ERROR_RANGE,
                  "IPC candidates require an initially feasible source pose with distance greater than %.17g "
                  "(minimum plus an f32 safety margin); got %.17g.",
                  required, minimum);
  }
  uint32_t color_count = 0;
  TRYS(color_vertices(in, s->colors, color_count));
  if (color_count > kMaxColor + 1)
    return fail(JGS2_ERROR_RANGE,
                "Colored Gauss-Seidel supports at most %u colors; the small-scene conflict graph needs %u.",
                kMaxColor + 1, color_count);
  s->V = V;
  s->objectives.assign(8 * size_t(V), 0.0f);
  for (uint32_t v = 0; v < V; ++v) {  // packJGS2VertexObjectives (layout.ts:680-716)
    float* o = &s->objectives[8 * size_t(v)];
    if (in.external_forces) {
      o[0] = in.external_forces[3 * size_t(v)];
      o[1] = in.external_forces[3 * size_t(v) + 1];
      o[2] = in.external_forces[3 * size_t(v) + 2];
    }
    o[4] = in.target_positions ? in.target_positions[3 * size_t(v)] : in.positions[4 * size_t(v)];
    o[5] = in.target_positions ? in.target_positions[3 * size_t(v) + 1] : in.positions[4 * size_t(v) + 1];
    o[6] = in.target_positions ? in.target_positions[3 * size_t(v) + 2] : in.positions[4 * size_t(v) + 2];
    o[7] = in.target_stiffnesses ? in.target_stiffnesses[v] : 0.0f;
  }
  recount_objectives(s);
  if (!globalization && s->objective_flags != 0)
    return fail(JGS2_ERROR_RANGE,
                "Active external forces and quadratic targets require the stable Neo-Hookean production path.");
  if (globalization) {
    const double minimum = min_input_determinant(in);
    if (!std::isfinite(minimum) || !(minimum > kDeterminantFloor))
      return fail(JGS2_ERROR_RANGE,
                  "A stable Neo-Hookean production solve requires an initially feasible pose with minimum J greater "
                  "than 0.0001; got %.17g.",
                  minimum);
  }
  s->defaults = resolved;
  s->preprocessing_timestep = resolved.timestep;
  s->globalization = globalization;
  s->corotated = corotated;
  s->T = T;
  s->K = K;
  s->NT = NT;
  s->NH = NH;
  s->NC = NC;
  s->NVT = in.contact ? in.contact->vertex_triangle_count : 0u;
  s->has_cloth = NT + NH > 0;
  if (s->has_cloth) {
    const double E = in.cloth->young_modulus, nu = in.cloth->poisson_ratio;
    s->cloth_mat[0] = float(E * nu / (1 - nu * nu));
    s->cloth_mat[1] = float(E / (2 * (1 + nu)));
    s->cloth_mat[2] = float(in.cloth->thickness);
    s->cloth_mat[3] = float(in.cloth->bending_stiffness);
  }
  uint32_t max_body = 0, free_count = 0;
  s->pinned.resize(V);
  for (uint32_t v = 0; v < V; ++v) {
    max_body = std::max(max_body, in.vertex_info[4 * size_t(v) + 3]);
    s->pinned[v] = in.vertex_info[4 * size_t(v) + 2] != 0;
    free_count += s->pinned[v] ? 0u : 1u;
  }
  s->bodies = max_body + 1;
  s->free_vertices = free_count;
  s->color_count = color_count;
  {  // Initial dynamic scene scale (jgs2-solver.ts:732-766).
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t dyn = 0;
    for (uint32_t v = 0; v < V; ++v) {
      if (s->pinned[v]) continue;
      ++dyn;
      for (int a = 0; a < 3; ++a) {
        double c = in.positions[4 * size_t(v) + a];
        mn[a] = std::min(mn[a], c);
        mx[a] = std::max(mx[a], c);
      }
    }
    double diag = dyn ? std::hypot(std::hypot(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]) : 0.0;
    s->scene_scale = float((dyn > 0 && std::isfinite(diag) && diag > 0) ? diag : 1e-12);
  }
  if (bp_mode) TRYS(prepare_broad_phase(in, resolved, s));
  morton_orders(in, s->vnew, s->vold, s->tnew, s->told);
  return JGS2_OK;
}

// The arena configuration that holds exactly one prepared scene.
jgs2_batch_config exact_config(const jgs2_solver* s) {
  jgs2_batch_config c{};
  c.max_vertices = s->V;
  c.max_tets = s->T;
  c.max_triangles = s->NT;
  c.max_hinges = s->NH;
  c.max_candidates = s->NC;
  c.max_bodies = s->bodies;
  c.max_scenes = 1;
  c.cubature_k = s->K;
  c.corotated = s->corotated ? 1u : 0u;
  c.max_surface_triangles = s->NS;
  c.max_surface_edges = s->NE;
  return c;
}

// Facts of a prepared scene the batch must share.
jgs2_status check_fits(const jgs2_batch* b, const jgs2_solver* s) {
  if (s->corotated != b->corotated)
    return fail(JGS2_ERROR_RANGE, "A batch holds either stable Neo-Hookean or co-rotated scenes, not both.");
  if (s->K > b->cfg.cubature_k)
    return fail(JGS2_ERROR_RANGE, "The scene has %u cubature records per vertex; the batch holds at most %u.", s->K,
                b->cfg.cubature_k);
  if (s->has_cloth && b->cloth_scenes > 0) {
    const float* m = s->cloth_mat;
    const DeviceState& d = b->dev;
    if (m[0] != d.cloth_lam || m[1] != d.cloth_mu || m[2] != d.cloth_thick || m[3] != d.cloth_bend)
      return fail(JGS2_ERROR_RANGE, "Every resident cloth scene of a batch shares one cloth material.");
  }
  if (float(s->defaults.ipc_minimum_distance) != b->dev.ipc_dmin && s->NC > 0)
    return fail(JGS2_ERROR_RANGE, "The scene's ipcMinimumDistance differs from the batch's.");
  return JGS2_OK;
}

// Marks a scene's blocks and slot free on the device.
jgs2_status clear_scene_tables(jgs2_batch* b, const jgs2_solver* s) {
  const uint4 zero = make_uint4(0, 0, 0, 0);
  HIP_CHECK(hipMemcpy(b->dev.scene_cnt + s->slot, &zero, sizeof(uint4), hipMemcpyHostToDevice));
  std::vector<uint32_t> none(std::max({s->vblk.n, s->tblk.n, s->rblk.n, s->cblk.n, uint64_t(1)}), kNoScene);
  HIP_CHECK(hipMemcpy(b->dev.vblock_scene + s->vblk.at, none.data(), 4 * s->vblk.n, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(b->dev.tblock_scene + s->tblk.at, none.data(), 4 * s->tblk.n, hipMemcpyHostToDevice));
  if (s->rblk.n) HIP_CHECK(hipMemcpy(b->dev.rblock_scene + s->rblk.at, none.data(), 4 * s->rblk.n, hipMemcpyHostToDevice));
  if (s->cblk.n) HIP_CHECK(hipMemcpy(b->dev.cblock_scene + s->cblk.at, none.data(), 4 * s->cblk.n, hipMemcpyHostToDevice));
  if (s->body.n) {
    std::vector<uint4> rng(s->body.n, make_uint4(0, 0, 0, kNoScene));
    HIP_CHECK(hipMemcpy(b->dev.body_rng + s->body.at, rng.data(), rng.size() * sizeof(uint4), hipMemcpyHostToDevice));
  }
  if (b->dev.bp_ctl) {
    HIP_CHECK(hipMemcpy(b->dev.bp_ctl + s->slot, &zero, sizeof(uint4), hipMemcpyHostToDevice));
    std::vector<uint32_t> no(std::max<uint64_t>({s->sblk.n, s->eblk.n, 1}), kNoScene);
    if (s->sblk.n) HIP_CHECK(hipMemcpy(b->dev.sblock_scene + s->sblk.at, no.data(), 4 * s->sblk.n, hipMemcpyHostToDevice));
    if (s->eblk.n) HIP_CHECK(hipMemcpy(b->dev.eblock_scene + s->eblk.at, no.data(), 4 * s->eblk.n, hipMemcpyHostToDevice));
  }
  return JGS2_OK;
}

// The arena blocks of every resident broad-phase scene, the grid of every
// broad-phase launch: vertex blocks, surface triangle blocks, surface edge
// blocks, candidate blocks.
jgs2_status update_bp_list(jgs2_batch* b) {
  DeviceState& d = b->dev;
  if (!d.bp_list) return JGS2_OK;
  std::vector<uint32_t> lists[4];
  for (auto* s : b->slots) {
    if (!s || !s->bp) continue;
    const Span* sp[4] = {&s->vblk, &s->sblk, &s->eblk, &s->cblk};
    for (int k = 0; k < 4; ++k)
      for (uint64_t i = 0; i < sp[k]->n; ++i) lists[k].push_back(uint32_t(sp[k]->at + i));
  }
  std::vector<uint32_t> all;
  for (auto& l : lists) all.insert(all.end(), l.begin(), l.end());
  if (!all.empty()) {
    HIP_CHECK(hipMemcpyAsync(d.bp_list, all.data(), 4 * all.size(), hipMemcpyHostToDevice, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  }
  d.bp_nv = uint32_t(lists[0].size());
  d.bp_ns = uint32_t(lists[1].size());
  d.bp_ne = uint32_t(lists[2].size());
  d.bp_nc = uint32_t(lists[3].size());
  return JGS2_OK;
}

// A broad-phase build that needed more room than the scene has makes every
// later result invalid; reported by the next synchronizing call.
jgs2_status check_bp_overflow(jgs2_batch* b) {
  DeviceState& d = b->dev;
  if (!d.bp_ctl || d.bp_nv == 0) return JGS2_OK;
  std::vector<uint4> ctl(b->slots.size());
  HIP_CHECK(hipMemcpy(ctl.data(), d.bp_ctl, ctl.size() * sizeof(uint4), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < ctl.size(); ++i) {
    const jgs2_solver* s = b->slots[i];
    if (s && s->bp && (ctl[i].x & kBpOverflow))
      return fail(JGS2_ERROR_MEMORY,
                  "A broad-phase contact build needed %u candidates (capacity %u) or a vertex row above 8192 entries; "
                  "results are invalid. Create the scene with a larger candidate_capacity.",
                  ctl[i].w, s->NC);
  }
  return JGS2_OK;
}

void forget_scene(jgs2_batch* b, jgs2_solver* s) {
  b->slots[s->slot] = nullptr;
  if (s->has_cloth) b->cloth_scenes -= 1;
  release_spans(b, s);
  if (s->bp) (void)update_bp_list(b);
  b->version++;
  b->graph_valid = false;
}

// Places a prepared scene: spans for every arena range, packing, and the
// GPU preflight of the uploaded f32 source (validateGpuInitialStableSource,
// jgs2-solver.ts:820-893). Takes ownership of `s` (deleted on failure).
// JGS2_ERROR_MEMORY means the arena has no room for the scene now.
jgs2_status place_scene(jgs2_batch* b, jgs2_solver* s, const jgs2_scene_input& in, jgs2_solver** out) {
  jgs2_status st = check_fits(b, s);
  if (st != JGS2_OK) {
    delete s;
    return st;
  }
  s->batch = b;
  uint64_t slot = 0;
  if (!b->slota.alloc(1, slot)) {
    delete s;
    return fail(JGS2_ERROR_MEMORY, "The batch holds its maximum of %u scenes.", b->cfg.max_scenes);
  }
  s->slot = uint32_t(slot);
  bool ok = true;
  auto take = [&](RangeAlloc& a, uint64_t n, Span& sp) {
    if (!ok) return;
    ok = a.alloc(n, sp.at);
    sp.n = ok ? n : 0;
  };
  take(b->va, blocks_of(s->V), s->vblk);
  take(b->ta, blocks_of(s->T), s->tblk);
  take(b->ra, blocks_of(s->NT), s->rblk);
  take(b->ca, blocks_of(s->NC), s->cblk);
  take(b->adja, in.adjacency_length, s->adj);
  take(b->hina, s->NH, s->hin);
  take(b->tela, 3 * uint64_t(s->NT), s->tel);
  take(b->hela, 4 * uint64_t(s->NH), s->hel);
  take(b->cela, 4 * uint64_t(s->NC), s->cel);
  take(b->bodya, s->bodies, s->body);
  take(b->sa, blocks_of(s->NS), s->sblk);
  take(b->ea, blocks_of(s->NE), s->eblk);
  take(b->rga, 2 * uint64_t(s->NE), s->rng);
  if (!ok) {
    release_spans(b, s);
    delete s;
    return fail(JGS2_ERROR_MEMORY, "The batch arena has no room for this scene now.");
  }
  // Color-sorted vertex list of the scene in arena ids.
  s->color_offsets.assign(s->color_count + 1, 0);
  for (uint32_t v = 0; v < s->V; ++v) s->color_offsets[s->colors[v] + 1] += 1;
  for (uint32_t c = 0; c < s->color_count; ++c) s->color_offsets[c + 1] += s->color_offsets[c];
  s->color_order.assign(s->V, 0);
  std::vector<uint32_t> dev_colors(size_t(s->vblk.n) * kSceneBlock, 0);
  {
    std::vector<uint32_t> cursor(s->color_offsets.begin(), s->color_offsets.end() - 1);
    for (uint32_t nv = 0; nv < s->V; ++nv) {
      const uint32_t c = s->colors[s->vold[nv]];
      s->color_order[cursor[c]++] = uint32_t(s->vbase()) + nv;
      dev_colors[nv] = c;
    }
  }
  if (s->has_cloth && b->cloth_scenes == 0) {
    b->dev.cloth_lam = s->cloth_mat[0];
    b->dev.cloth_mu = s->cloth_mat[1];
    b->dev.cloth_thick = s->cloth_mat[2];
    b->dev.cloth_bend = s->cloth_mat[3];
  }
  st = pack_into_arena(b, s, in);
  if (st == JGS2_OK && hipMemcpy(b->dev.colors + s->vbase(), dev_colors.data(), 4 * dev_colors.size(),
                                 hipMemcpyHostToDevice) != hipSuccess)
    st = fail(JGS2_ERROR_DEVICE, "Uploading schedule colors failed.");
  b->slots[s->slot] = s;
  if (s->has_cloth) b->cloth_scenes += 1;
  b->version++;
  b->graph_valid = false;
  if (st == JGS2_OK && s->bp) {  // the candidate list at the uploaded pose
    st = update_bp_list(b);
    if (st == JGS2_OK) {
      PassParams p = make_params(b, b->defaults, false);
      p.src = 0;
      p.tgt = 0;
      launch_broad_phase(b->dev, p, false, false, b->stream);
      uint4 ctl = make_uint4(0, 0, 0, 0);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(&ctl, b->dev.bp_ctl + s->slot, sizeof(uint4), hipMemcpyDeviceToHost, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (e != hipSuccess)
        st = fail(JGS2_ERROR_DEVICE, "The broad-phase contact build failed: %s", hipGetErrorString(e));
      else if (ctl.x & kBpOverflow)
        st = fail(JGS2_ERROR_MEMORY,
                  "The initial broad-phase contact list needs %u candidates (or a vertex row above 8192 entries); "
                  "the scene's capacity is %u.",
                  ctl.w, s->NC);
    }
  }
  if (st == JGS2_OK && !b->corotated) {
    PassParams p = make_params(b, b->defaults, false);
    p.src = 0;
    p.tgt = 0;
    const uint4 blk0 = make_uint4(uint32_t(s->vblk.at), uint32_t(s->vblk.at + s->vblk.n), uint32_t(s->tblk.at),
                                  uint32_t(s->tblk.at + s->tblk.n));
    const uint4 blk1 = make_uint4(uint32_t(s->rblk.at), uint32_t(s->rblk.at + s->rblk.n), uint32_t(s->cblk.at),
                                  uint32_t(s->cblk.at + s->cblk.n));
    launch_reduce_candidate_scene(b->dev, p, s->slot, blk0, blk1, b->stream);
    float control[4] = {0, 0, 0, 0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(control, b->dev.control + 4 * size_t(s->slot), sizeof(control), hipMemcpyDeviceToHost,
                         b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    const uint32_t required = 1u | 2u | 4u;
    const float bits = control[3];
    const bool feasible = std::isfinite(control[0]) && std::isfinite(control[1]) && std::isfinite(control[2]) &&
                          control[0] > kDeterminantFloor && control[1] > kDeterminantFloor &&
                          control[2] > kDeterminantFloor && bits == std::floor(bits) &&
                          (uint32_t(bits) & required) == required;
    if (e != hipSuccess)
      st = fail(JGS2_ERROR_DEVICE, "The source-feasibility preflight failed: %s", hipGetErrorString(e));
    else if (!feasible)
      st = fail(JGS2_ERROR_INFEASIBLE,
                "The uploaded stable Neo-Hookean source pose is not feasible under the production GPU f32 "
                "determinant calculation; source J=%g, candidate J=%g, accepted J=%g.",
                double(control[0]), double(control[1]), double(control[2]));
  }
  if (st != JGS2_OK) {
    const std::string message = g_last_error;
    (void)clear_scene_tables(b, s);
    forget_scene(b, s);
    delete s;
    g_last_error = message;
    return st;
  }
  *out = s;
  clear_error();
  return JGS2_OK;
}

jgs2_status retire(jgs2_batch* b, jgs2_solver* s) {
  HIP_CHECK(hipStreamSynchronize(b->stream));
  TRYS(clear_scene_tables(b, s));
  forget_scene(b, s);
  s->destroyed = true;
  return JGS2_OK;
}

jgs2_status batch_create(const jgs2_batch_config& cfg, const jgs2_settings& resolved, jgs2_batch** out) {
  int device_count = 0;
  if (hipGetDeviceCount(&device_count) != hipSuccess || device_count < 1)
    return fail(JGS2_ERROR_DEVICE, "No HIP device is available; libjgs2hip has no CPU fallback.");
  jgs2_batch* b = new (std::nothrow) jgs2_batch();
  if (!b) return fail(JGS2_ERROR_MEMORY, "Out of host memory.");
  b->cfg = cfg;
  b->defaults = resolved;
  b->corotated = cfg.corotated != 0;
  jgs2_status st = JGS2_OK;
  if (hipGetDevice(&b->device) != hipSuccess) st = fail(JGS2_ERROR_DEVICE, "hipGetDevice failed.");
  if (st == JGS2_OK && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    st = fail(JGS2_ERROR_DEVICE, "hipStreamCreateWithFlags failed.");
  if (st == JGS2_OK) st = batch_allocate(b);
  // Contact parameters are fixed when the arena is created
  // (jgs2-solver.ts:1096-1105).
  b->dev.ipc_dmin = float(resolved.ipc_minimum_distance);
  b->dev.ipc_mu = float(resolved.ipc_friction_coefficient);
  b->dev.ipc_eps = float(resolved.ipc_friction_velocity_epsilon);
  b->dev.ipc_safety = float(resolved.ipc_step_safety);
  if (st != JGS2_OK) {
    free_device(b->dev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return st;
  }
  *out = b;
  return JGS2_OK;
}

void batch_destroy(jgs2_batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto*& s : b->slots)
    if (s) {
      s->destroyed = true;
      delete s;
      s = nullptr;
    }
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  free_device(b->dev);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  if (b->frame_begin) (void)hipEventDestroy(b->frame_begin);
  if (b->frame_end) (void)hipEventDestroy(b->frame_end);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// ---------------------------------------------------------------- per-scene access
jgs2_status check_usable(jgs2_solver* s) {
  if (!s) return fail(JGS2_ERROR_TYPE, "NULL solver.");
  if (s->destroyed) return fail(JGS2_ERROR_STATE, "JGS2GpuSolver has been destroyed.");
  return JGS2_OK;
}

jgs2_status check_owner(jgs2_solver* s) {
  TRYS(check_usable(s));
  if (!s->owns_batch) return fail(JGS2_ERROR_STATE, "A batch-resident scene is stepped through jgs2_batch_step_frames.");
  return JGS2_OK;
}

jgs2_status upload_objectives(jgs2_solver* s, size_t first_vertex, size_t vertex_count) {
  // The host mirror is in caller vertex ids; device records are in arena order.
  jgs2_batch* b = s->batch;
  float4* base = b->dev.objectives + 2 * s->vbase();
  if (vertex_count == s->V) {
    std::vector<float> staged(8 * vertex_count);
    for (uint32_t nv = 0; nv < s->V; ++nv)
      std::memcpy(&staged[8 * size_t(nv)], &s->objectives[8 * size_t(s->vold[nv])], 32);
    HIP_CHECK(hipMemcpyAsync(base, staged.data(), 32 * size_t(s->V), hipMemcpyHostToDevice, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
  } else {
    for (size_t i = 0; i < vertex_count; ++i) {
      const size_t ov = first_vertex + i;
      HIP_CHECK(hipMemcpyAsync(base + 2 * size_t(s->vnew[ov]), &s->objectives[8 * ov], 32, hipMemcpyHostToDevice,
                               b->stream));
    }
    HIP_CHECK(hipStreamSynchronize(b->stream));
  }
  recount_objectives(s);
  // The scene's bodies carry its objective flags for the body correction.
  std::vector<uint4> rng(s->body.n, make_uint4(uint32_t(s->vbase()), uint32_t(s->vbase() + s->V), s->objective_flags,
                                               s->slot));
  HIP_CHECK(hipMemcpy(b->dev.body_rng + s->body.at, rng.data(), rng.size() * sizeof(uint4), hipMemcpyHostToDevice));
  b->graph_valid = false;
  return JGS2_OK;
}

jgs2_status objective_supported(jgs2_solver* s) {
  jgs2_status st = check_usable(s);
  if (st != JGS2_OK) return st;
  if (!s->globalization)
    return fail(JGS2_ERROR_STATE, "Vertex objectives require the stable Neo-Hookean production path.");
  return JGS2_OK;
}

// Copies a per-vertex arena array (`stride` floats per vertex) of one scene
// to the caller in caller vertex ids; `stride` 0 copies `expected` floats
// verbatim from `src`.
jgs2_status read_buffer(jgs2_solver* s, const float* src, float* dst, size_t count, size_t expected, const char* what,
                        size_t stride = 0) {
  jgs2_status st = check_usable(s);
  if (st != JGS2_OK) return st;
  if (!dst) return fail(JGS2_ERROR_TYPE, "NULL destination for %s.", what);
  if (count != expected) return fail(JGS2_ERROR_RANGE, "%s needs exactly %zu floats; got %zu.", what, expected, count);
  HIP_CHECK(hipStreamSynchronize(s->batch->stream));
  if (stride == 0) {
    HIP_CHECK(hipMemcpy(dst, src, expected * sizeof(float), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> staged(expected);
    HIP_CHECK(hipMemcpy(staged.data(), src + stride * s->vbase(), expected * sizeof(float), hipMemcpyDeviceToHost));
    for (uint32_t ov = 0; ov < s->V; ++ov)
      std::memcpy(dst + stride * ov, &staged[stride * size_t(s->vnew[ov])], stride * sizeof(float));
  }
  clear_error();
  return JGS2_OK;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

void jgs2_default_settings(jgs2_settings* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->schedule = JGS2_SCHEDULE_JACOBI;
  o->timestep = 1.0 / 60.0;
  o->gravity[0] = 0;
  o->gravity[1] = -9.81;
  o->gravity[2] = 0;
  o->iterations = 9;
  o->floor_height = 0;
  o->floor_stiffness = 20000;
  o->velocity_damping = 0.999;
  o->regularization = 1e-6;
  o->rotation_epsilon = 1e-7;
  o->max_step = 0.1;
  o->residual_tolerance = 1e-3;
  o->normalized_update_tolerance = 1e-3;
  o->contact_tangential_damping = 8;
  o->contact_margin = 0.01;
  o->ipc_activation_distance = 0.05;
  o->ipc_minimum_distance = 0.001;
  o->ipc_barrier_stiffness = 0;
  o->ipc_friction_coefficient = 0;
  o->ipc_friction_velocity_epsilon = 0.01;
  o->ipc_step_safety = 0.9;
  o->horizontal_body_correction = 1;
  o->parity_mode = 0;
}

size_t jgs2_last_error(char* buffer, size_t capacity) {
  const std::string& e = jgs2::g_last_error;
  if (buffer && capacity > 0) {
    size_t n = std::min(capacity - 1, e.size());
    std::memcpy(buffer, e.data(), n);
    buffer[n] = '\0';
  }
  return e.size();
}

int jgs2_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

jgs2_status jgs2_create(const jgs2_scene_input* input, const jgs2_settings* overrides, uint64_t mask,
                        jgs2_solver** out) {
  if (!input || !out) return fail(JGS2_ERROR_TYPE, "jgs2_create requires an input and an output handle.");
  *out = nullptr;
  jgs2_settings defaults;
  jgs2_default_settings(&defaults);
  jgs2_settings resolved;
  TRYS(merge_settings(defaults, overrides, mask, resolved));
  jgs2_solver* s = new (std::nothrow) jgs2_solver();
  if (!s) return fail(JGS2_ERROR_MEMORY, "Out of host memory.");
  jgs2_status st = prepare_scene(*input, resolved, s);
  if (st != JGS2_OK) {
    delete s;
    return st;
  }
  jgs2_batch* b = nullptr;
  st = batch_create(exact_config(s), resolved, &b);
  if (st != JGS2_OK) {
    delete s;
    return st;
  }
  st = place_scene(b, s, *input, out);
  if (st != JGS2_OK) {
    const std::string message = g_last_error;
    batch_destroy(b);
    g_last_error = message;
    return st;
  }
  (*out)->owns_batch = true;
  return JGS2_OK;
}

jgs2_status jgs2_destroy(jgs2_solver* s) {
  if (!s) return JGS2_OK;
  if (s->destroyed) return fail(JGS2_ERROR_STATE, "JGS2GpuSolver has been destroyed.");
  if (s->owns_batch) {
    batch_destroy(s->batch);  // deletes s
  } else {
    TRYS(retire(s->batch, s));
    delete s;
  }
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_step(jgs2_solver* s, const jgs2_settings* ov, uint64_t mask) {
  TRYS(check_owner(s));
  jgs2_settings r;
  TRYS(resolve_batch_step(s->batch, ov, mask, "step", r));
  return submit(s->batch, r, normalize_odd(r.iterations), 1, true);
}

jgs2_status jgs2_step_frames(jgs2_solver* s, uint32_t frames, const jgs2_settings* ov, uint64_t mask) {
  TRYS(check_owner(s));
  TRYS(validate_frames(frames));
  jgs2_settings r;
  TRYS(resolve_batch_step(s->batch, ov, mask, "stepFrames", r));
  return submit(s->batch, r, normalize_odd(r.iterations), frames, true);
}

jgs2_status jgs2_step_exact_iterations(jgs2_solver* s, uint32_t iterations, const jgs2_settings* ov, uint64_t mask) {
  TRYS(check_owner(s));
  if (iterations < 1) return fail(JGS2_ERROR_RANGE, "iterations must be a positive safe integer; got %u.", iterations);
  jgs2_settings r;
  TRYS(resolve_batch_step(s->batch, ov, mask, "stepExactIterations", r));
  return submit(s->batch, r, iterations, 1, false);
}

jgs2_status jgs2_step_frames_exact_iterations(jgs2_solver* s, uint32_t frames, uint32_t iterations,
                                              const jgs2_settings* ov, uint64_t mask) {
  TRYS(check_owner(s));
  TRYS(validate_frames(frames));
  if (iterations < 1) return fail(JGS2_ERROR_RANGE, "iterations must be a positive safe integer; got %u.", iterations);
  jgs2_settings r;
  TRYS(resolve_batch_step(s->batch, ov, mask, "stepFramesExactIterations", r));
  return submit(s->batch, r, iterations, frames, false);
}

jgs2_status jgs2_set_external_force(jgs2_solver* s, uint32_t vertex, const float force[3]) {
  TRYS(objective_supported(s));
  if (vertex >= s->V) return fail(JGS2_ERROR_RANGE, "Objective vertex %u is outside vertexCount %u.", vertex, s->V);
  if (!force) return fail(JGS2_ERROR_TYPE, "NULL force.");
  for (int c = 0; c < 3; ++c)
    if (!std::isfinite(force[c]))
      return fail(JGS2_ERROR_RANGE, "External force %d for vertex %u must be finite after f32 conversion.", c, vertex);
  if (s->pinned[vertex] && (force[0] != 0 || force[1] != 0 || force[2] != 0))
    return fail(JGS2_ERROR_RANGE, "Pinned vertex %u cannot have a nonzero external force.", vertex);
  float* o = &s->objectives[8 * size_t(vertex)];
  o[0] = force[0];
  o[1] = force[1];
  o[2] = force[2];
  o[3] = 0;
  return upload_objectives(s, vertex, 1);
}

jgs2_status jgs2_replace_external_forces(jgs2_solver* s, const float* forces, size_t count) {
  TRYS(objective_supported(s));
  if (!forces) return fail(JGS2_ERROR_TYPE, "NULL forces.");
  if (count != 3 * size_t(s->V))
    return fail(JGS2_ERROR_RANGE, "External forces must contain %zu values; got %zu.", 3 * size_t(s->V), count);
  for (uint32_t v = 0; v < s->V; ++v) {
    bool nonzero = false;
    for (int c = 0; c < 3; ++c) {
      float f = forces[3 * size_t(v) + c];
      if (!std::isfinite(f))
        return fail(JGS2_ERROR_RANGE, "External forces[%zu] must be finite after f32 conversion.", 3 * size_t(v) + c);
      nonzero |= f != 0;
    }
    if (s->pinned[v] && nonzero) return fail(JGS2_ERROR_RANGE, "Pinned vertex %u cannot have a nonzero external force.", v);
  }
  for (uint32_t v = 0; v < s->V; ++v) {
    float* o = &s->objectives[8 * size_t(v)];
    for (int c = 0; c < 3; ++c) o[c] = forces[3 * size_t(v) + c];
    o[3] = 0;
  }
  return upload_objectives(s, 0, s->V);
}

jgs2_status jgs2_clear_external_forces(jgs2_solver* s) {
  TRYS(objective_supported(s));
  std::vector<float> zeros(3 * size_t(s->V), 0.0f);
  return jgs2_replace_external_forces(s, zeros.data(), zeros.size());
}

jgs2_status jgs2_set_quadratic_target(jgs2_solver* s, uint32_t vertex, const float target[3], float stiffness) {
  TRYS(objective_supported(s));
  if (vertex >= s->V) return fail(JGS2_ERROR_RANGE, "Objective vertex %u is outside vertexCount %u.", vertex, s->V);
  if (!target) return fail(JGS2_ERROR_TYPE, "NULL target.");
  for (int c = 0; c < 3; ++c)
    if (!std::isfinite(target[c]))
      return fail(JGS2_ERROR_RANGE, "Quadratic target %d for vertex %u must be finite after f32 conversion.", c, vertex);
  if (!std::isfinite(stiffness))
    return fail(JGS2_ERROR_RANGE, "Quadratic-target stiffness for vertex %u must be finite after f32 conversion.", vertex);
  if (stiffness < 0) return fail(JGS2_ERROR_RANGE, "Quadratic-target stiffness for vertex %u must be nonnegative.", vertex);
  if (s->pinned[vertex] && stiffness != 0)
    return fail(JGS2_ERROR_RANGE, "Pinned vertex %u cannot have a nonzero quadratic-target stiffness.", vertex);
  float* o = &s->objectives[8 * size_t(vertex)];
  o[4] = target[0];
  o[5] = target[1];
  o[6] = target[2];
  o[7] = stiffness;
  return upload_objectives(s, vertex, 1);
}

jgs2_status jgs2_replace_quadratic_targets(jgs2_solver* s, const float* positions, size_t position_count,
                                           const float* stiffnesses, size_t stiffness_count) {
  TRYS(objective_supported(s));
  if (!positions || !stiffnesses) return fail(JGS2_ERROR_TYPE, "NULL targets.");
  if (position_count != 3 * size_t(s->V))
    return fail(JGS2_ERROR_RANGE, "Quadratic target positions must contain %zu values; got %zu.", 3 * size_t(s->V),
                position_count);
  if (stiffness_count != s->V)
    return fail(JGS2_ERROR_RANGE, "Quadratic target stiffnesses must contain %u values; got %zu.", s->V, stiffness_count);
  for (uint32_t v = 0; v < s->V; ++v) {
    for (int c = 0; c < 3; ++c)
      if (!std::isfinite(positions[3 * size_t(v) + c]))
        return fail(JGS2_ERROR_RANGE, "Quadratic target positions[%zu] must be finite after f32 conversion.",
                    3 * size_t(v) + c);
    float k = stiffnesses[v];
    if (!std::isfinite(k))
      return fail(JGS2_ERROR_RANGE, "Quadratic target stiffnesses[%u] must be finite after f32 conversion.", v);
    if (k < 0) return fail(JGS2_ERROR_RANGE, "Quadratic target stiffnesses[%u] must be nonnegative.", v);
    if (s->pinned[v] && k != 0)
      return fail(JGS2_ERROR_RANGE, "Pinned vertex %u cannot have a nonzero quadratic-target stiffness.", v);
  }
  for (uint32_t v = 0; v < s->V; ++v) {
    float* o = &s->objectives[8 * size_t(v)];
    for (int c = 0; c < 3; ++c) o[4 + c] = positions[3 * size_t(v) + c];
    o[7] = stiffnesses[v];
  }
  return upload_objectives(s, 0, s->V);
}

jgs2_status jgs2_release_quadratic_target(jgs2_solver* s, uint32_t vertex) {
  TRYS(objective_supported(s));
  if (vertex >= s->V) return fail(JGS2_ERROR_RANGE, "Objective vertex %u is outside vertexCount %u.", vertex, s->V);
  s->objectives[8 * size_t(vertex) + 7] = 0;
  return upload_objectives(s, vertex, 1);
}

jgs2_status jgs2_release_all_quadratic_targets(jgs2_solver* s) {
  TRYS(objective_supported(s));
  for (uint32_t v = 0; v < s->V; ++v) s->objectives[8 * size_t(v) + 7] = 0;
  return upload_objectives(s, 0, s->V);
}

jgs2_status jgs2_read_positions(jgs2_solver* s, float* dst, size_t count) {
  TRYS(check_usable(s));
  return read_buffer(s, &s->batch->dev.pos[0]->x, dst, count, 4 * size_t(s->V), "readPositions", 4);
}
jgs2_status jgs2_read_velocities(jgs2_solver* s, float* dst, size_t count) {
  TRYS(check_usable(s));
  return read_buffer(s, &s->batch->dev.velocity->x, dst, count, 4 * size_t(s->V), "readVelocities", 4);
}
jgs2_status jgs2_read_predicted_positions(jgs2_solver* s, float* dst, size_t count) {
  TRYS(check_usable(s));
  return read_buffer(s, &s->batch->dev.predicted->x, dst, count, 4 * size_t(s->V), "readPredictedPositions", 4);
}
jgs2_status jgs2_read_final_iteration_updates(jgs2_solver* s, float* dst, size_t count) {
  TRYS(check_usable(s));
  return read_buffer(s, &s->batch->dev.final_update->x, dst, count, 4 * size_t(s->V), "readFinalIterationUpdates", 4);
}
jgs2_status jgs2_read_vertex_rotations(jgs2_solver* s, float* dst, size_t count) {
  TRYS(check_usable(s));
  return read_buffer(s, &s->batch->dev.vrot->x, dst, count, 12 * size_t(s->V), "readVertexRotations", 12);
}
jgs2_status jgs2_read_local_diagnostics(jgs2_solver* s, float* dst, size_t count) {
  TRYS(check_usable(s));
  if (!s->globalization) return fail(JGS2_ERROR_STATE, "Globalization diagnostics require the stable material path.");
  return read_buffer(s, &s->batch->dev.local_diag->x, dst, count, 16 * size_t(s->V), "readLocalDiagnostics", 16);
}
jgs2_status jgs2_read_globalization_control(jgs2_solver* s, float* dst, size_t count) {
  TRYS(check_usable(s));
  return read_buffer(s, &s->batch->dev.control[4 * size_t(s->slot)].x, dst, count, 16, "readGlobalizationControl");
}
jgs2_status jgs2_read_globalization_history(jgs2_solver* s, float* dst, size_t count) {
  TRYS(check_usable(s));
  return read_buffer(s, &s->batch->dev.history[size_t(s->slot) * kHistoryCapacity * 8].x, dst, count,
                     size_t(kHistoryCapacity) * 32, "readGlobalizationHistory");
}

jgs2_status jgs2_sync(jgs2_solver* s) {
  TRYS(check_usable(s));
  HIP_CHECK(hipStreamSynchronize(s->batch->stream));
  TRYS(check_bp_overflow(s->batch));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_read_contact_candidates(jgs2_solver* s, uint32_t* packed, uint64_t capacity, uint32_t* vt_count,
                                         uint64_t* total, float* reach) {
  TRYS(check_usable(s));
  if (!vt_count || !total || !reach || (capacity && !packed)) return fail(JGS2_ERROR_TYPE, "NULL argument.");
  jgs2_batch* b = s->batch;
  HIP_CHECK(hipStreamSynchronize(b->stream));
  TRYS(check_bp_overflow(b));
  uint32_t n = s->NC, vt = s->NVT;
  float r = INFINITY;
  if (s->bp) {
    uint4 ctl;
    float4 rc;
    HIP_CHECK(hipMemcpy(&ctl, b->dev.bp_ctl + s->slot, sizeof(uint4), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(&rc, b->dev.bp_reach + s->slot, sizeof(float4), hipMemcpyDeviceToHost));
    n = ctl.z;
    vt = ctl.y;
    r = rc.x;
  }
  *vt_count = vt;
  *total = n;
  *reach = r;
  if (packed && n) {
    std::vector<uint4> rec(n);
    HIP_CHECK(hipMemcpy(rec.data(), b->dev.cand + s->cbase(), n * sizeof(uint4), hipMemcpyDeviceToHost));
    const uint32_t VB = uint32_t(s->vbase());
    for (uint64_t c = 0; c < std::min<uint64_t>(n, capacity); ++c) {
      const uint32_t ids[4] = {rec[c].x, rec[c].y, rec[c].z, rec[c].w};
      for (int l = 0; l < 4; ++l) packed[4 * c + l] = s->vold[ids[l] - VB];
    }
  }
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_estimate_contact_capacity(const jgs2_scene_input* input, const jgs2_settings* overrides,
                                           uint64_t mask, uint64_t* capacity) {
  if (!input || !capacity) return fail(JGS2_ERROR_TYPE, "NULL argument.");
  jgs2_settings defaults, resolved;
  jgs2_default_settings(&defaults);
  TRYS(merge_settings(defaults, overrides, mask, resolved));
  jgs2_solver probe;
  TRYS(prepare_scene(*input, resolved, &probe));
  *capacity = probe.NC;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_get_contact_info(jgs2_solver* s, jgs2_contact_info* out) {
  TRYS(check_usable(s));
  if (!out) return fail(JGS2_ERROR_TYPE, "NULL info.");
  jgs2_batch* b = s->batch;
  HIP_CHECK(hipStreamSynchronize(b->stream));
  std::memset(out, 0, sizeof(*out));
  out->candidates = s->NC;
  out->vertex_triangle = s->NVT;
  out->capacity = s->NC;
  out->reach = INFINITY;
  if (s->bp) {
    uint4 ctl;
    float4 rc;
    HIP_CHECK(hipMemcpy(&ctl, b->dev.bp_ctl + s->slot, sizeof(uint4), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(&rc, b->dev.bp_reach + s->slot, sizeof(float4), hipMemcpyDeviceToHost));
    out->candidates = ctl.z;
    out->vertex_triangle = ctl.y;
    out->builds = uint64_t(rc.z);
    out->reach = rc.x;
    out->broad_phase = 1;
  }
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_rebuild_contact_candidates(jgs2_solver* s) {
  TRYS(check_usable(s));
  if (!s->bp) return fail(JGS2_ERROR_STATE, "The scene has no broad-phase contact.");
  jgs2_batch* b = s->batch;
  const uint4 ctl = make_uint4(kBpScene | kBpForce, 0, 0, 0);
  HIP_CHECK(hipStreamSynchronize(b->stream));
  HIP_CHECK(hipMemcpy(&b->dev.bp_ctl[s->slot].x, &ctl.x, sizeof(uint32_t), hipMemcpyHostToDevice));
  PassParams p = make_params(b, b->defaults, false);
  p.src = 0;
  p.tgt = 0;
  launch_broad_phase(b->dev, p, false, false, b->stream);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(b->stream));
  TRYS(check_bp_overflow(b));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_get_info(jgs2_solver* s, jgs2_solver_info* out) {
  TRYS(check_usable(s));
  if (!out) return fail(JGS2_ERROR_TYPE, "NULL info.");
  out->vertex_count = s->V;
  out->tet_count = s->T;
  out->cubature_k = s->K;
  out->body_count = s->bodies;
  out->schedule_color_count = s->color_count;
  out->globalization_enabled = s->globalization ? 1u : 0u;
  out->free_vertex_count = s->free_vertices;
  out->scene_scale = s->scene_scale;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_read_schedule_colors(jgs2_solver* s, uint32_t* colors, size_t count) {
  TRYS(check_usable(s));
  if (!colors || count != s->V) return fail(JGS2_ERROR_RANGE, "Schedule colors need exactly %u entries.", s->V);
  std::memcpy(colors, s->colors.data(), 4 * size_t(s->V));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_debug_local_systems(jgs2_solver* s, const jgs2_settings* ov, uint64_t mask, float* g, size_t gc,
                                     float* h, size_t hc) {
  TRYS(check_usable(s));
  if (!g || !h || gc != 3 * size_t(s->V) || hc != 9 * size_t(s->V))
    return fail(JGS2_ERROR_RANGE, "Local systems need 3V gradient and 9V Hessian floats.");
  jgs2_batch* b = s->batch;
  jgs2_settings r;
  TRYS(resolve_batch_step(b, ov, mask, "debugLocalSystems", r));
  DeviceState& d = b->dev;
  const size_t Vc = size_t(b->va.capacity) * kSceneBlock;
  if (!d.debug_g) {
    TRYS(dalloc(&d.debug_g, 3 * Vc));
    TRYS(dalloc(&d.debug_h, 9 * Vc));
  }
  PassParams p = make_params(b, r, false);
  p.src = 0;
  p.tgt = 0;
  hipStream_t q = b->stream;
  if (p.contact) {
    // Contact state of a frame start at the current pose: old = x, lagged
    // contact and active rows. The next frame's predict and lag pass rewrite
    // all of it, so stepping continues unaffected.
    HIP_CHECK(hipMemcpyAsync(d.old, d.pos[0], size_t(p.vertex_count) * sizeof(float4), hipMemcpyDeviceToDevice, q));
    launch_lag_contact(d, p, q);
    launch_build_active_rows(d, p, q);
  }
  if (b->corotated) launch_tet_polar(d, p, q);
  launch_vertex_polar(d, p, q);
  launch_debug_local_systems(d, p, q);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(q));
  const size_t Vp = size_t(s->vblk.n) * kSceneBlock;
  std::vector<float> gs(3 * Vp), hs(9 * Vp);
  HIP_CHECK(hipMemcpy(gs.data(), d.debug_g + 3 * s->vbase(), gs.size() * sizeof(float), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(hs.data(), d.debug_h + 9 * s->vbase(), hs.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (uint32_t ov = 0; ov < s->V; ++ov) {
    std::memcpy(g + 3 * size_t(ov), &gs[3 * size_t(s->vnew[ov])], 3 * sizeof(float));
    std::memcpy(h + 9 * size_t(ov), &hs[9 * size_t(s->vnew[ov])], 9 * sizeof(float));
  }
  b->graph_valid = false;
  clear_error();
  return JGS2_OK;
}

namespace {

// Doubles of one evaluation's device records: 36 per vertex, 4 per
// tetrahedron (+9 per tetrahedron for co-rotated frames), 2 per candidate
// slot, 32 of summary.
size_t diagnostics_doubles(const jgs2_solver* s) {
  const size_t V = s->V, T = s->T, C = size_t(s->cblk.n) * kSceneBlock;
  return 36 * V + 4 * T + (s->corotated ? 9 * T : 0) + 2 * C + 32;
}

// The evaluator's four passes over one scene at positions x against the
// predicted positions xp, with its records in `buf` (diagnostics_doubles),
// then the summary. Synchronizes the batch's stream.
jgs2_status evaluate_diagnostics(jgs2_solver* s, const jgs2_settings& r, const float4* x, const float4* xp, double* buf,
                                 jgs2_diagnostics* out, DiagParams* used) {
  jgs2_batch* b = s->batch;
  DeviceState& d = b->dev;
  hipStream_t q = b->stream;
  const size_t V = s->V, T = s->T, C = size_t(s->cblk.n) * kSceneBlock;
  DiagParams p{};
  p.x = x;
  p.xp = xp;
  p.vb = uint32_t(s->vbase());
  p.V = uint32_t(V);
  p.tb = uint32_t(s->tbase());
  p.T = uint32_t(T);
  p.cb = uint32_t(s->cbase());
  p.C = s->NC || s->bp ? uint32_t(C) : 0u;
  p.scene = s->slot;
  p.objective_flags = s->objective_flags;
  p.corotated = s->corotated ? 1u : 0u;
  p.ih2 = 1.0 / (r.timestep * r.timestep);
  p.floor_y = r.floor_height;
  p.floor_k = r.floor_stiffness;
  p.kappa = r.ipc_barrier_stiffness;
  p.dhat = r.ipc_activation_distance;
  p.polar_eps = r.rotation_epsilon;
  p.vrec = buf;
  p.trec = p.vrec + 36 * V;
  p.trot = p.trec + 4 * T;
  p.crec = p.trot + (s->corotated ? 9 * T : 0);
  p.summary = p.crec + 2 * C;
  launch_diagnostics(d, p, q);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(q));
  double m[32];
  HIP_CHECK(hipMemcpy(m, p.summary, sizeof m, hipMemcpyDeviceToHost));
  jgs2_diagnostics o{};
  o.energy = m[0];
  o.inertia = m[1];
  o.elasticity = m[2];
  o.external_force = m[3];
  o.quadratic_target = m[4];
  o.floor_contact = m[5];
  o.contact_barrier = m[6];
  o.gradient_norm_inertia = std::sqrt(std::max(m[7], 0.0));
  o.gradient_norm_material = std::sqrt(std::max(m[8], 0.0));
  o.gradient_norm_external_force = std::sqrt(std::max(m[9], 0.0));
  o.gradient_norm_target = std::sqrt(std::max(m[10], 0.0));
  o.gradient_norm_floor = std::sqrt(std::max(m[11], 0.0));
  o.gradient_norm = std::sqrt(std::max(m[12], 0.0));
  o.residual_denominator = o.gradient_norm_inertia + o.gradient_norm_material + o.gradient_norm_external_force +
                           o.gradient_norm_target + o.gradient_norm_floor;
  o.relative_residual = o.gradient_norm / std::max(1.0, o.residual_denominator);
  o.maximum_update = m[14] > 0 ? m[13] : 0.0;
  o.minimum_deformation_determinant = T ? m[15] : 0.0;
  o.minimum_floor_distance = m[16];
  o.minimum_contact_distance = m[20];
  o.active_floor_contacts = uint32_t(m[17]);
  o.active_vertices = uint32_t(m[18]);
  o.active_contact_pairs = uint32_t(m[21]);
  o.finite = m[19] == 0 && m[22] == 0 && std::isfinite(m[0]) ? 1u : 0u;
  o.maximum_update_valid = m[14] > 0 ? 1u : 0u;
  o.minimum_determinant_valid = T && std::isfinite(m[15]) ? 1u : 0u;
  o.floor_metrics_valid = r.floor_stiffness > 0 ? 1u : 0u;
  *out = o;
  if (used) *used = p;
  return JGS2_OK;
}

jgs2_status resolve_diagnostics_settings(jgs2_solver* s, const jgs2_settings* ov, uint64_t mask, const char* api,
                                         jgs2_settings& r) {
  TRYS(resolve_batch_step(s->batch, ov, mask, api, r));
  if (!(r.floor_stiffness >= 0)) return fail(JGS2_ERROR_RANGE, "floorStiffness must be nonnegative.");
  return JGS2_OK;
}

}  // namespace

jgs2_status jgs2_read_oracle_diagnostics(jgs2_solver* s, const jgs2_settings* ov, uint64_t mask, const float* positions,
                                         const float* predicted, jgs2_diagnostics* out, double* vertex_records,
                                         size_t vertex_record_count, double* tet_records, size_t tet_record_count) {
  TRYS(check_usable(s));
  if (!out) return fail(JGS2_ERROR_TYPE, "readOracleDiagnostics needs a summary destination.");
  if (vertex_records && vertex_record_count != 36 * size_t(s->V))
    return fail(JGS2_ERROR_RANGE, "Oracle diagnostics vertex records need 36V doubles.");
  if (tet_records && tet_record_count != 4 * size_t(s->T))
    return fail(JGS2_ERROR_RANGE, "Oracle diagnostics tetrahedron records need 4T doubles.");
  jgs2_batch* b = s->batch;
  jgs2_settings r;
  TRYS(resolve_diagnostics_settings(s, ov, mask, "readOracleDiagnostics", r));
  DeviceState& d = b->dev;
  hipStream_t q = b->stream;
  const size_t Vc = size_t(b->va.capacity) * kSceneBlock;
  const size_t V = s->V, T = s->T;
  const size_t need = diagnostics_doubles(s);
  if (!d.diag_x) {
    TRYS(dalloc(&d.diag_x, Vc));
    TRYS(dalloc(&d.diag_xp, Vc));
  }
  if (d.diag_doubles < need) {
    if (d.diag_buf) HIP_CHECK(hipFree(d.diag_buf));
    d.diag_buf = nullptr;
    d.diag_doubles = 0;
    TRYS(dalloc(&d.diag_buf, need));
    d.diag_doubles = need;
  }
  HIP_CHECK(hipStreamSynchronize(q));
  auto stage = [&](const float* src, float4* dst) -> jgs2_status {
    std::vector<float4> staged(size_t(s->vblk.n) * kSceneBlock, make_float4(0, 0, 0, 0));
    for (uint32_t ov = 0; ov < s->V; ++ov)
      staged[s->vnew[ov]] = make_float4(src[4 * ov], src[4 * ov + 1], src[4 * ov + 2], src[4 * ov + 3]);
    HIP_CHECK(hipMemcpy(dst + s->vbase(), staged.data(), staged.size() * sizeof(float4), hipMemcpyHostToDevice));
    return JGS2_OK;
  };
  if (positions) TRYS(stage(positions, d.diag_x));
  if (predicted) TRYS(stage(predicted, d.diag_xp));
  DiagParams p{};
  TRYS(evaluate_diagnostics(s, r, positions ? d.diag_x : d.pos[0], predicted ? d.diag_xp : d.predicted, d.diag_buf,
                            out, &p));
  if (vertex_records) {
    std::vector<double> all(36 * size_t(s->vblk.n) * kSceneBlock);
    HIP_CHECK(hipMemcpy(all.data(), p.vrec, 36 * V * sizeof(double), hipMemcpyDeviceToHost));
    for (uint32_t ovx = 0; ovx < s->V; ++ovx)
      std::memcpy(vertex_records + 36 * size_t(ovx), &all[36 * size_t(s->vnew[ovx])], 36 * sizeof(double));
  }
  if (tet_records) {
    std::vector<double> all(4 * T);
    HIP_CHECK(hipMemcpy(all.data(), p.trec, 4 * T * sizeof(double), hipMemcpyDeviceToHost));
    for (uint32_t ot = 0; ot < s->T; ++ot)
      std::memcpy(tet_records + 4 * size_t(ot), &all[4 * size_t(s->tnew[ot])], 4 * sizeof(double));
  }
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_read_oracle_diagnostics_noalloc(jgs2_solver* s, const jgs2_settings* ov, uint64_t mask,
                                                 jgs2_diagnostics* out) {
  TRYS(check_usable(s));
  if (!out) return fail(JGS2_ERROR_TYPE, "readOracleDiagnostics needs a summary destination.");
  jgs2_settings r;
  TRYS(resolve_diagnostics_settings(s, ov, mask, "readOracleDiagnostics", r));
  const size_t need = diagnostics_doubles(s);
  if (need > kDiagScratchDoubles)
    return fail(JGS2_ERROR_RANGE,
                "Allocation-free diagnostics hold at most %zu doubles of records (36 per vertex, 4 per tetrahedron, "
                "9 more per co-rotated tetrahedron, 2 per candidate slot, 32); this scene needs %zu.",
                kDiagScratchDoubles, need);
  double* buf = nullptr;
  hipError_t e = diagnostics_scratch(s->batch->device, &buf);
  if (e != hipSuccess)
    return fail(JGS2_ERROR_DEVICE, "Looking up the diagnostics buffer failed: %s", hipGetErrorString(e));
  TRYS(evaluate_diagnostics(s, r, s->batch->dev.pos[0], s->batch->dev.predicted, buf, out, nullptr));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_profile_begin(jgs2_batch* b) {
  TRYS(check_batch(b));
  if (b->events.empty()) {
    b->events.resize(8192);
    for (auto& e : b->events) HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventCreate(&b->frame_begin));
    HIP_CHECK(hipEventCreate(&b->frame_end));
  }
  HIP_CHECK(hipStreamSynchronize(b->stream));
  b->profiling = true;
  b->event_used = 0;
  HIP_CHECK(hipEventRecord(b->frame_begin, b->stream));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_profile_end(jgs2_batch* b, double* ms_per_launch, uint64_t* launches, double* total_ms) {
  TRYS(check_batch(b));
  if (!b->profiling) return fail(JGS2_ERROR_STATE, "Profiling was not started.");
  HIP_CHECK(hipEventRecord(b->frame_end, b->stream));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  double sum = 0;
  for (size_t i = 0; i + 1 < b->event_used; i += 2) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, b->events[i], b->events[i + 1]));
    sum += ms;
  }
  const uint64_t n = b->event_used / 2;
  float total = 0;
  HIP_CHECK(hipEventElapsedTime(&total, b->frame_begin, b->frame_end));
  if (ms_per_launch) *ms_per_launch = n ? sum / double(n) : 0.0;
  if (launches) *launches = n;
  if (total_ms) *total_ms = total;
  b->profiling = false;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_profile_begin(jgs2_solver* s) {
  TRYS(check_usable(s));
  return jgs2_batch_profile_begin(s->batch);
}

jgs2_status jgs2_profile_end(jgs2_solver* s, double* ms_per_launch, uint64_t* launches, double* total_ms) {
  TRYS(check_usable(s));
  return jgs2_batch_profile_end(s->batch, ms_per_launch, launches, total_ms);
}

// ---------------------------------------------------------------- batches
jgs2_status jgs2_batch_create(const jgs2_batch_config* config, const jgs2_settings* overrides, uint64_t mask,
                              jgs2_batch** out) {
  if (!config || !out) return fail(JGS2_ERROR_TYPE, "jgs2_batch_create requires a config and an output handle.");
  *out = nullptr;
  if (config->max_scenes < 1 || config->max_vertices < 1 || config->max_tets < 1)
    return fail(JGS2_ERROR_RANGE, "A batch needs room for at least one scene, vertex and tetrahedron.");
  if (config->cubature_k > 64) return fail(JGS2_ERROR_RANGE, "cubature_k must be at most 64; got %u.", config->cubature_k);
  jgs2_settings defaults;
  jgs2_default_settings(&defaults);
  jgs2_settings resolved;
  TRYS(merge_settings(defaults, overrides, mask, resolved));
  TRYS(batch_create(*config, resolved, out));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_destroy(jgs2_batch* b) {
  if (!b) return JGS2_OK;
  batch_destroy(b);
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_admit(jgs2_batch* b, const jgs2_scene_input* input, jgs2_solver** scene) {
  TRYS(check_batch(b));
  if (!input || !scene) return fail(JGS2_ERROR_TYPE, "jgs2_batch_admit requires an input and an output handle.");
  *scene = nullptr;
  jgs2_solver* s = new (std::nothrow) jgs2_solver();
  if (!s) return fail(JGS2_ERROR_MEMORY, "Out of host memory.");
  jgs2_status st = prepare_scene(*input, b->defaults, s);
  if (st != JGS2_OK) {
    delete s;
    return st;
  }
  return place_scene(b, s, *input, scene);
}

jgs2_status jgs2_batch_retire(jgs2_batch* b, jgs2_solver* scene) {
  TRYS(check_batch(b));
  TRYS(check_usable(scene));
  if (scene->batch != b || scene->owns_batch) return fail(JGS2_ERROR_STATE, "The scene is not resident in this batch.");
  TRYS(retire(b, scene));
  delete scene;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_step_frames(jgs2_batch* b, uint32_t frames, const jgs2_settings* ov, uint64_t mask) {
  TRYS(check_batch(b));
  TRYS(validate_frames(frames));
  jgs2_settings r;
  TRYS(resolve_batch_step(b, ov, mask, "stepFrames", r));
  return submit(b, r, normalize_odd(r.iterations), frames, true);
}

jgs2_status jgs2_batch_step_frames_exact_iterations(jgs2_batch* b, uint32_t frames, uint32_t iterations,
                                                    const jgs2_settings* ov, uint64_t mask) {
  TRYS(check_batch(b));
  TRYS(validate_frames(frames));
  if (iterations < 1) return fail(JGS2_ERROR_RANGE, "iterations must be a positive safe integer; got %u.", iterations);
  jgs2_settings r;
  TRYS(resolve_batch_step(b, ov, mask, "stepFramesExactIterations", r));
  return submit(b, r, iterations, frames, false);
}

jgs2_status jgs2_batch_sync(jgs2_batch* b) {
  TRYS(check_batch(b));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  TRYS(check_bp_overflow(b));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_get_info(jgs2_batch* b, jgs2_batch_info* out) {
  TRYS(check_batch(b));
  if (!out) return fail(JGS2_ERROR_TYPE, "NULL info.");
  std::memset(out, 0, sizeof(*out));
  out->scene_count = b->active_scenes();
  out->max_scenes = uint32_t(b->slots.size());
  out->solve_block = b->solve_block;
  out->solve_lanes = b->solve_serial_only ? 1u : b->solve_lanes;
  out->xcd_remap = b->xcd_remap;
  out->use_graphs = b->use_graphs ? 1u : 0u;
  uint64_t used = 0;
  for (auto* s : b->slots)
    if (s) used += s->vblk.n;
  out->vertex_capacity = b->va.capacity * kSceneBlock;
  out->vertex_blocks_used = used;
  out->vertex_extent = b->va.high() * kSceneBlock;
  out->tet_extent = b->ta.high() * kSceneBlock;
  out->arena_bytes = b->arena_bytes;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_arena_bytes(const jgs2_batch_config* c, uint64_t* bytes) {
  if (!c || !bytes) return fail(JGS2_ERROR_TYPE, "NULL argument.");
  // The allocation list of a batch of this configuration, counted only.
  jgs2_batch probe;
  probe.cfg = *c;
  probe.corotated = c->corotated != 0;
  TRYS(batch_allocate(&probe, true));
  *bytes = probe.arena_bytes;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_use_graphs(jgs2_batch* b, int enable) {
  TRYS(check_batch(b));
  b->use_graphs = enable != 0;
  b->graph_valid = false;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_graph_stats(jgs2_batch* b, uint64_t* instantiations, uint64_t* updates) {
  TRYS(check_batch(b));
  if (!instantiations || !updates) return fail(JGS2_ERROR_TYPE, "NULL argument.");
  *instantiations = b->graph_instantiations;
  *updates = b->graph_updates;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_set_launch(jgs2_batch* b, uint32_t solve_block, uint32_t xcd_remap, uint32_t solve_lanes) {
  TRYS(check_batch(b));
  if (solve_block != 64 && solve_block != 128 && solve_block != 256)
    return fail(JGS2_ERROR_RANGE, "solve_block must be 64, 128 or 256; got %u.", solve_block);
  if (solve_lanes != 1 && solve_lanes != 4 && solve_lanes != 8 && solve_lanes != 16)
    return fail(JGS2_ERROR_RANGE, "solve_lanes must be 1, 4, 8 or 16; got %u.", solve_lanes);
  b->solve_block = solve_block;
  b->solve_lanes = solve_lanes == 1 ? b->solve_lanes : solve_lanes;
  b->solve_serial_only = solve_lanes == 1;
  HIP_CHECK(hipStreamSynchronize(b->stream));
  for (auto* s : b->slots)
    if (s) {
      const uint32_t lanes = scene_lanes(b, s);
      HIP_CHECK(hipMemcpy(b->dev.scene_lanes + s->slot, &lanes, sizeof(uint32_t), hipMemcpyHostToDevice));
    }
  b->xcd_remap = xcd_remap ? 1u : 0u;
  b->graph_valid = false;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_batch_autotune(jgs2_batch* b, const jgs2_settings* ov, uint64_t mask, uint32_t repeats,
                                double* best_ms) {
  TRYS(check_batch(b));
  if (b->active_scenes() == 0) return fail(JGS2_ERROR_STATE, "The batch holds no scene.");
  if (repeats < 1 || repeats > 64) return fail(JGS2_ERROR_RANGE, "repeats must be in [1, 64]; got %u.", repeats);
  // The timed solve starts from the last frame's pose and line-search
  // control; before any frame it would time the early-exit path only.
  if (b->stepped_version != b->version)
    return fail(JGS2_ERROR_STATE, "Step at least one frame after the last admission or retirement before autotuning.");
  jgs2_settings r;
  TRYS(resolve_batch_step(b, ov, mask, "autotune", r));
  // Times the Jacobi local solve alone from the accepted pose into the
  // second position buffer, with the vertex frames the last frame left.
  // The buffers it writes (trial positions, vertex energies, local
  // diagnostics) are rewritten by the next frame before they are read, so
  // stepping is unaffected; only jgs2_read_local_diagnostics() shows the
  // timing passes until the next frame.
  // The six shapes are timed round robin, one launch each per round, so
  // every shape sees the same state and clock; each keeps its median.
  constexpr uint32_t kShapes = 6;
  const uint32_t blocks[3] = {64, 128, 256};
  std::vector<hipEvent_t> ev(2 * size_t(kShapes) * repeats, nullptr);
  hipError_t err = hipSuccess;
  for (auto& e : ev)
    if (err == hipSuccess) err = hipEventCreate(&e);
  const uint32_t keep_block = b->solve_block, keep_remap = b->xcd_remap;
  // Each timed launch follows a 512 MiB stream (past the 256 MB
  // last-level cache) when the arena's cubature records are that large.
  const uint64_t cub_records = uint64_t(b->va.capacity) * kSceneBlock * 10 * std::max<uint32_t>(b->cfg.cubature_k, 1);
  const uint32_t flush_records = uint32_t(std::min<uint64_t>(cub_records, uint64_t(32) << 20));
  for (uint32_t k = 0; k <= repeats && err == hipSuccess; ++k)  // round 0 warms every shape
    for (uint32_t c = 0; c < kShapes; ++c) {
      b->solve_block = blocks[c % 3];
      b->xcd_remap = c / 3;
      PassParams p = make_params(b, r, false);
      p.src = 0;
      p.tgt = 1;
      if (k == 0) {
        launch_solve_config(b->dev, p, b->stream);
        continue;
      }
      hipEvent_t* e = &ev[2 * (size_t(k - 1) * kShapes + c)];
      launch_cache_flush(b->dev, flush_records, b->stream);
      (void)hipEventRecord(e[0], b->stream);
      launch_solve_config(b->dev, p, b->stream);
      (void)hipEventRecord(e[1], b->stream);
    }
  if (err == hipSuccess) err = hipStreamSynchronize(b->stream);
  double best = INFINITY;
  uint32_t best_block = keep_block, best_remap = keep_remap;
  double med[kShapes];
  for (uint32_t c = 0; c < kShapes && err == hipSuccess; ++c) {
    std::vector<float> t(repeats);
    for (uint32_t k = 0; k < repeats && err == hipSuccess; ++k) {
      const hipEvent_t* e = &ev[2 * (size_t(k) * kShapes + c)];
      err = hipEventElapsedTime(&t[k], e[0], e[1]);
    }
    std::nth_element(t.begin(), t.begin() + repeats / 2, t.end());
    med[c] = t[repeats / 2];
    if (err == hipSuccess && med[c] < best) {
      best = med[c];
      best_block = blocks[c % 3];
      best_remap = c / 3;
    }
  }
  // Within 2 % of the fastest the XCD order wins: it keeps the vertices that
  // share elements on one L2 and moves about half the L2 misses at the same
  // time, which leaves the memory system to the other passes of a frame.
  if (err == hipSuccess && best_remap == 0u) {
    uint32_t pick = kShapes;
    for (uint32_t c = 3; c < kShapes; ++c)
      if (med[c] <= 1.02 * best && (pick == kShapes || med[c] < med[pick])) pick = c;
    if (pick < kShapes) {
      best = med[pick];
      best_block = blocks[pick % 3];
      best_remap = 1u;
    }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (err == hipSuccess) err = hipGetLastError();
  b->solve_block = err == hipSuccess ? best_block : keep_block;
  b->xcd_remap = err == hipSuccess ? best_remap : keep_remap;
  b->graph_valid = false;
  if (err != hipSuccess) return fail(JGS2_ERROR_DEVICE, "Autotuning failed: %s", hipGetErrorString(err));
  if (best_ms) *best_ms = best;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_solver_batch(jgs2_solver* s, jgs2_batch** out) {
  TRYS(check_usable(s));
  if (!out) return fail(JGS2_ERROR_TYPE, "NULL output.");
  *out = s->batch;
  clear_error();
  return JGS2_OK;
}

}  // extern "C"
