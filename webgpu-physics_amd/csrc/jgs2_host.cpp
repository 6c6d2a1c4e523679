// libjgs2hip host side: the device half of JGS2GpuSolver
// (src/simulation/gpu/jgs2-solver.ts) behind a C ABI. Validation, coloring,
// packing and frame orchestration run here; every pass of the frame is a
// HIP kernel in jgs2_kernels.hip enqueued on the solver's stream.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

float round_f32(double v) { return float(v); }

}  // namespace
}  // namespace jgs2

using namespace jgs2;

struct jgs2_solver {
  bool destroyed = false;
  DeviceState dev;
  hipStream_t stream = nullptr;
  jgs2_settings defaults{};
  double preprocessing_timestep = 0;
  float scene_scale = 0;
  bool globalization = true;
  bool corotated = false;
  uint32_t V = 0, T = 0, K = 0, bodies = 1, color_count = 1, free_vertices = 0;
  std::vector<uint32_t> colors;         // by caller vertex id
  std::vector<uint32_t> color_offsets;  // color_count + 1
  // Device order: vertices and tetrahedra are stored along a Morton curve of
  // their rest positions so that neighbors share cache lines and L2; the
  // boundary maps ids both ways.
  std::vector<uint32_t> vnew, vold, tnew, told;
  std::vector<float> objectives;        // 8 per vertex (host mirror)
  std::vector<uint8_t> pinned;
  uint32_t objective_flags = 0;
  uint32_t active_forces = 0, active_targets = 0;
  // Profiling of the local-solve kernel.
  bool profiling = false;
  std::vector<hipEvent_t> events;
  size_t event_used = 0;
  hipEvent_t frame_begin = nullptr, frame_end = nullptr;
  bool frame_timing = false;
};

namespace {

void free_device(DeviceState& d) {
  void* ptrs[] = {d.vtx_rest, d.vtx_info, d.tet_idx, d.tet_dm, d.tet_k, d.adjacency, d.cub, d.objectives,
                  d.color_order, d.colors, d.pos[0], d.pos[1], d.predicted, d.velocity, d.old, d.vrot, d.vrot_next, d.trot,
                  d.body_corr, d.final_update, d.local_diag, d.tet_glob, d.vertex_energy, d.conv, d.control,
                  d.history, d.partials, d.debug_g, d.debug_h, d.tri_ix, d.tri_inv, d.tri_area, d.hinge_ix,
                  d.hinge_rest, d.tri_rows, d.tri_elems, d.hinge_rows, d.hinge_elems, d.cand, d.lag_n, d.lag_w,
                  d.cscratch, d.src_c, d.src_w, d.tgt_c, d.tgt_w, d.contact_ctl, d.cand_rows, d.cand_ids,
                  d.active_count, d.active_ids};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  d = DeviceState{};
}

template <class T> jgs2_status dalloc(T** ptr, size_t count) {
  *ptr = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T));
  if (e != hipSuccess)
    return fail(JGS2_ERROR_MEMORY, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
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

// Vertex -> element CSR in device vertex order: the caller-order CSR
// (buildIncidentElementCsr, cloth-layout.ts:183-207) with its rows moved to
// their device vertex; each row stays ascending in element id.
jgs2_status incident_csr(const uint32_t* idx, uint32_t width, uint32_t count, const std::vector<uint32_t>& vold,
                         std::vector<uint32_t>& rows, std::vector<uint32_t>& elems) {
  const uint32_t V = uint32_t(vold.size());
  std::vector<uint32_t> crow(size_t(V) + 1), celem(size_t(count) * width);
  jgs2_status st = jgs2_build_element_incidence(V, idx, width, count, crow.data(), celem.data());
  if (st != JGS2_OK) return st;
  rows.assign(size_t(V) + 1, 0);
  elems.clear();
  elems.reserve(celem.size());
  for (uint32_t nv = 0; nv < V; ++nv) {
    const uint32_t ov = vold[nv];
    elems.insert(elems.end(), celem.begin() + crow[ov], celem.begin() + crow[ov + 1]);
    rows[nv + 1] = uint32_t(elems.size());
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

// ---------------------------------------------------------------- pass parameters
PassParams make_params(const jgs2_solver* s, const jgs2_settings& st, bool stop) {
  PassParams p{};
  p.vertex_count = s->V;
  p.tet_count = s->T;
  p.cubature_k = s->K;
  p.body_count = s->bodies;
  p.objective_flags = s->objective_flags;
  p.globalization = s->globalization ? 1u : 0u;
  p.stop_enabled = (s->globalization && stop) ? 1u : 0u;
  p.corotated = s->corotated ? 1u : 0u;
  p.tri_count = s->dev.NT;
  p.hinge_count = s->dev.NH;
  // contactEnabled (jgs2-solver.ts:2213-2214).
  p.contact = (s->dev.NC > 0 && st.ipc_barrier_stiffness > 0) ? 1u : 0u;
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
  p.scene_scale = s->scene_scale;
  return p;
}

void record_solve(jgs2_solver* s, const PassParams& p) {
  if (s->profiling && s->event_used + 2 <= s->events.size()) {
    (void)hipEventRecord(s->events[s->event_used], s->stream);
    launch_solve(s->dev, p, s->stream);
    (void)hipEventRecord(s->events[s->event_used + 1], s->stream);
    s->event_used += 2;
    return;
  }
  launch_solve(s->dev, p, s->stream);
}

// encodeFrame (jgs2-solver.ts:2207-2494), one HIP launch per pass.
void enqueue_frame(jgs2_solver* s, const jgs2_settings& st, uint32_t iterations, bool stop) {
  const DeviceState& d = s->dev;
  hipStream_t q = s->stream;
  PassParams base = make_params(s, st, stop);
  const bool gs = st.schedule == JGS2_SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL;
  base.src = 0;
  base.tgt = 0;
  const bool contact = base.contact != 0u;
  if (s->globalization) launch_clear_frame(d, base, q);
  launch_predict(d, base, q);
  if (contact) launch_lag_contact(d, base, q);
  if (s->globalization) launch_reduce_candidate(d, base, q);  // source feasibility preflight
  for (uint32_t it = 0; it < iterations; ++it) {
    PassParams p = base;
    p.src = (it % 2 == 0) ? 1u : 0u;
    p.tgt = (it % 2 == 0) ? 0u : 1u;
    if (contact) launch_build_active_rows(d, p, q);
    if (gs) {
      launch_copy_position(d, p, q);
      for (uint32_t c = 0; c < s->color_count; ++c) {
        PassParams pc = p;
        pc.src = p.tgt;
        pc.colored = 1;
        pc.color_begin = s->color_offsets[c];
        pc.color_end = s->color_offsets[c + 1];
        if (s->corotated) launch_tet_polar(d, pc, q);
        launch_vertex_polar(d, pc, q);
        record_solve(s, pc);
      }
      if (s->globalization && !contact) launch_vertex_energy(d, p, q);
    } else {
      // Sweeps after the first read the frames the previous convergence pass
      // computed at this sweep's source pose.
      const bool fused = s->globalization && it > 0;
      if (s->corotated) launch_tet_polar(d, p, q);
      if (!fused) launch_vertex_polar(d, p, q);
      PassParams ps = p;
      ps.fused_rot = fused ? 1u : 0u;
      record_solve(s, ps);
      if (s->globalization && (d.NT || d.NH) && !contact) {
        PassParams pe = p;
        pe.energy_after_sweep = 1;
        launch_vertex_energy(d, pe, q);
      }
    }
    if (contact) {  // safe step, validation, then the energy pair (jgs2-solver.ts:2347-2400)
      launch_contact_step(d, p, q);
      if (s->globalization) launch_vertex_energy(d, p, q);
    }
    if (s->globalization) {
      launch_reduce_candidate(d, p, q);
      launch_apply_candidate(d, p, q);
      if (contact) launch_promote_contact(d, p, q);
      PassParams pc = p;
      pc.compute_next_rot = (!gs && it + 1 < iterations) ? 1u : 0u;
      launch_convergence(d, pc, q);
    }
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

jgs2_status check_usable(jgs2_solver* s) {
  if (!s) return fail(JGS2_ERROR_TYPE, "NULL solver.");
  if (s->destroyed) return fail(JGS2_ERROR_STATE, "JGS2GpuSolver has been destroyed.");
  return JGS2_OK;
}

// resolveStepSettings (jgs2-solver.ts:1982-2009).
jgs2_status resolve_step(jgs2_solver* s, const jgs2_settings* ov, uint64_t mask, const char* api, jgs2_settings& out) {
  if (ov && (mask & JGS2_SET_TIMESTEP)) {
    const double pre = s->preprocessing_timestep, run = ov->timestep;
    if (!(std::isfinite(run) && run > 0 && float(pre) == float(run)))
      return fail(JGS2_ERROR_RANGE,
                  "JGS2 was precomputed for timestep %.17g, but %s() requested %.17g. Regenerate the precomputation "
                  "before changing timestep.",
                  pre, api, run);
  }
  if (s->dev.NC > 0 && ov && (mask & JGS2_SET_IPC_MINIMUM_DISTANCE) &&
      float(ov->ipc_minimum_distance) != float(s->defaults.ipc_minimum_distance))
    return fail(JGS2_ERROR_RANGE,
                "ipcMinimumDistance is fixed when an IPC solver is created; recreate the solver before changing it.");
  jgs2_status st = merge_settings(s->defaults, ov, mask, out);
  if (st != JGS2_OK) return st;
  // Objectives and mesh contact keep their solved centre-of-mass motion
  // (jgs2-solver.ts:1994-2008).
  const bool contact = s->dev.NC > 0 && out.ipc_barrier_stiffness > 0;
  if (s->globalization && (s->objective_flags != 0 || contact) && out.horizontal_body_correction)
    out.horizontal_body_correction = 0;
  return JGS2_OK;
}

uint32_t normalize_odd(int32_t iterations) {  // normalizeOddIterationCount (layout.ts:742-748)
  int32_t i = std::max(1, iterations);
  return uint32_t(i % 2 == 0 ? i + 1 : i);
}

jgs2_status submit(jgs2_solver* s, const jgs2_settings& st, uint32_t iterations, uint32_t frames, bool stop) {
  if (s->globalization) {
    if (iterations > kHistoryCapacity)
      return fail(JGS2_ERROR_RANGE, "Stable JGS2 records at most %u nonlinear iterations per frame; got %u.",
                  kHistoryCapacity, iterations);
    const uint32_t sub = st.schedule == JGS2_SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL ? s->color_count : 1u;
    if (uint64_t(iterations) * sub > uint64_t(JGS2_MAX_GLOBALIZED_ITERATIONS_PER_SUBMISSION / frames))
      return fail(JGS2_ERROR_RANGE,
                  "Stable JGS2 may encode at most %u total nonlinear iterations per submission; got %u frames * %u "
                  "iterations.",
                  JGS2_MAX_GLOBALIZED_ITERATIONS_PER_SUBMISSION, frames, iterations * sub);
  }
  for (uint32_t f = 0; f < frames; ++f) enqueue_frame(s, st, iterations, stop);
  HIP_CHECK(hipGetLastError());
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

jgs2_status upload_objectives(jgs2_solver* s, size_t first_vertex, size_t vertex_count) {
  // The host mirror is in caller vertex ids; device records are in device order.
  std::vector<float> staged(8 * vertex_count);
  if (vertex_count == s->V) {
    for (uint32_t nv = 0; nv < s->V; ++nv)
      std::memcpy(&staged[8 * size_t(nv)], &s->objectives[8 * size_t(s->vold[nv])], 32);
    HIP_CHECK(hipMemcpyAsync(s->dev.objectives, staged.data(), 32 * size_t(s->V), hipMemcpyHostToDevice, s->stream));
  } else {
    for (size_t i = 0; i < vertex_count; ++i) {
      const size_t ov = first_vertex + i;
      HIP_CHECK(hipMemcpyAsync(s->dev.objectives + 2 * size_t(s->vnew[ov]), &s->objectives[8 * ov], 32,
                               hipMemcpyHostToDevice, s->stream));
    }
  }
  HIP_CHECK(hipStreamSynchronize(s->stream));
  recount_objectives(s);
  return JGS2_OK;
}

jgs2_status objective_supported(jgs2_solver* s) {
  jgs2_status st = check_usable(s);
  if (st != JGS2_OK) return st;
  if (!s->globalization)
    return fail(JGS2_ERROR_STATE, "Vertex objectives require the stable Neo-Hookean production path.");
  return JGS2_OK;
}

// Copies a device array to the caller. Per-vertex arrays (`stride` floats
// per vertex) are returned in caller vertex ids.
jgs2_status read_buffer(jgs2_solver* s, const void* src, float* dst, size_t count, size_t expected, const char* what,
                        size_t stride = 0) {
  jgs2_status st = check_usable(s);
  if (st != JGS2_OK) return st;
  if (!dst) return fail(JGS2_ERROR_TYPE, "NULL destination for %s.", what);
  if (count != expected) return fail(JGS2_ERROR_RANGE, "%s needs exactly %zu floats; got %zu.", what, expected, count);
  HIP_CHECK(hipStreamSynchronize(s->stream));
  if (stride == 0) {
    HIP_CHECK(hipMemcpy(dst, src, expected * sizeof(float), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> staged(expected);
    HIP_CHECK(hipMemcpy(staged.data(), src, expected * sizeof(float), hipMemcpyDeviceToHost));
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
  const jgs2_scene_input& in = *input;
  jgs2_status st = validate_input(in);
  if (st != JGS2_OK) return st;
  jgs2_settings defaults;
  jgs2_default_settings(&defaults);
  jgs2_settings resolved;
  st = merge_settings(defaults, overrides, mask, resolved);
  if (st != JGS2_OK) return st;

  const uint32_t V = in.vertex_count, T = in.tet_count, K = in.cubature_k;
  const bool corotated = in.tet_meta[3] == JGS2_MATERIAL_COROTATED_LINEAR;
  const bool globalization = !corotated;
  const uint32_t NT = in.cloth ? in.cloth->triangle_count : 0u, NH = in.cloth ? in.cloth->hinge_count : 0u;
  if (in.cloth) {
    st = validate_cloth(*in.cloth, V);
    if (st != JGS2_OK) return st;
    if (NT > 0 && !globalization)  // jgs2-solver.ts:998-1001
      return fail(JGS2_ERROR_RANGE, "Triangle cloth requires the stable Neo-Hookean production path.");
  }
  const uint32_t NC = in.contact ? in.contact->vertex_triangle_count + in.contact->edge_edge_count : 0u;
  if (in.contact) {
    st = validate_contact(*in.contact, V);
    if (st != JGS2_OK) return st;
  }
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
  std::vector<uint32_t> colors;
  uint32_t color_count = 0;
  st = color_vertices(in, colors, color_count);
  if (st != JGS2_OK) return st;
  if (color_count > kMaxColor + 1)
    return fail(JGS2_ERROR_RANGE,
                "Colored Gauss-Seidel supports at most %u colors; the small-scene conflict graph needs %u.",
                kMaxColor + 1, color_count);
  std::vector<float> objectives(8 * size_t(V), 0.0f);
  for (uint32_t v = 0; v < V; ++v) {  // packJGS2VertexObjectives (layout.ts:680-716)
    float* o = &objectives[8 * size_t(v)];
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
  bool any_objective = false;
  for (uint32_t v = 0; v < V && !any_objective; ++v) {
    const float* o = &objectives[8 * size_t(v)];
    any_objective = o[0] != 0 || o[1] != 0 || o[2] != 0 || o[7] > 0;
  }
  if (!globalization && any_objective)
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
  int device_count = 0;
  if (hipGetDeviceCount(&device_count) != hipSuccess || device_count < 1)
    return fail(JGS2_ERROR_DEVICE, "No HIP device is available; libjgs2hip has no CPU fallback.");

  jgs2_solver* s = new (std::nothrow) jgs2_solver();
  if (!s) return fail(JGS2_ERROR_MEMORY, "Out of host memory.");
  s->defaults = resolved;
  s->preprocessing_timestep = resolved.timestep;
  s->globalization = globalization;
  s->corotated = corotated;
  s->V = V;
  s->T = T;
  s->K = K;
  uint32_t max_body = 0, free_count = 0;
  s->pinned.resize(V);
  for (uint32_t v = 0; v < V; ++v) {
    max_body = std::max(max_body, in.vertex_info[4 * size_t(v) + 3]);
    s->pinned[v] = in.vertex_info[4 * size_t(v) + 2] != 0;
    free_count += s->pinned[v] ? 0u : 1u;
  }
  s->bodies = max_body + 1;
  s->free_vertices = free_count;
  s->colors = colors;
  s->color_count = color_count;
  s->objectives = objectives;
  recount_objectives(s);
  // Initial dynamic scene scale (jgs2-solver.ts:732-766).
  {
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
  // Space-filling device order of vertices and tetrahedra.
  morton_orders(in, s->vnew, s->vold, s->tnew, s->told);
  // Color-sorted vertex order (vertex-coloring.ts:92-107), in device ids.
  s->color_offsets.assign(color_count + 1, 0);
  for (uint32_t v = 0; v < V; ++v) s->color_offsets[colors[v] + 1] += 1;
  for (uint32_t c = 0; c < color_count; ++c) s->color_offsets[c + 1] += s->color_offsets[c];
  std::vector<uint32_t> order(V);
  std::vector<uint32_t> dev_colors(V);
  {
    std::vector<uint32_t> cursor(s->color_offsets.begin(), s->color_offsets.end() - 1);
    for (uint32_t nv = 0; nv < V; ++nv) {
      const uint32_t c = colors[s->vold[nv]];
      order[cursor[c]++] = nv;
      dev_colors[nv] = c;
    }
  }

  auto cleanup = [&](jgs2_status code) {
    free_device(s->dev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return code;
  };
#define TRY(expr)                       \
  do {                                  \
    jgs2_status t_ = (expr);            \
    if (t_ != JGS2_OK) return cleanup(t_); \
  } while (0)
#define TRY_HIP(expr)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return cleanup(fail(JGS2_ERROR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_))); \
  } while (0)

  TRY_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  DeviceState& d = s->dev;
  d.V = V;
  d.T = T;
  d.K = K;
  d.bodies = s->bodies;
  // Static records in device order.
  const std::vector<uint32_t>& vnew = s->vnew;
  const std::vector<uint32_t>& vold = s->vold;
  const std::vector<uint32_t>& tnew = s->tnew;
  const std::vector<uint32_t>& told = s->told;
  std::vector<uint32_t> adjacency(in.adjacency_length);
  std::vector<float4> vtx_rest(V);
  std::vector<uint4> vtx_info(V);
  {
    uint32_t cursor = 0;
    for (uint32_t nv = 0; nv < V; ++nv) {
      const uint32_t ov = vold[nv];
      const float* r = &in.vertex_rest[4 * size_t(ov)];
      const uint32_t* i = &in.vertex_info[4 * size_t(ov)];
      vtx_rest[nv] = make_float4(r[0], r[1], r[2], r[3]);
      vtx_info[nv] = make_uint4(cursor, i[1], i[2], i[3]);
      // Row order stays ascending in caller tet ids (the reference's
      // summation order); entries are device tet ids.
      for (uint32_t k = 0; k < i[1]; ++k) adjacency[cursor + k] = tnew[in.adjacency[i[0] + k]];
      cursor += i[1];
    }
  }
  std::vector<uint4> tet_idx(T);
  std::vector<float4> tet_dm(3 * size_t(T));
  std::vector<float> tet_k;
  if (corotated) tet_k.resize(144 * size_t(T));
  for (uint32_t nt = 0; nt < T; ++nt) {
    const uint32_t ot = told[nt];
    const uint32_t* i = &in.tet_indices[4 * size_t(ot)];
    const float* m = &in.tet_inverse_dm[12 * size_t(ot)];
    const float* a = &in.tet_meta[4 * size_t(ot)];
    tet_idx[nt] = make_uint4(vnew[i[0]], vnew[i[1]], vnew[i[2]], vnew[i[3]]);
    tet_dm[3 * size_t(nt)] = make_float4(m[0], m[1], m[2], a[0]);
    tet_dm[3 * size_t(nt) + 1] = make_float4(m[4], m[5], m[6], a[1]);
    tet_dm[3 * size_t(nt) + 2] = make_float4(m[8], m[9], m[10], a[2]);
    if (corotated) std::memcpy(&tet_k[144 * size_t(nt)], &in.tet_rest_stiffness[144 * size_t(ot)], 144 * sizeof(float));
  }
  const size_t R = size_t(V) * K;
  std::vector<uint4> cub(10 * std::max<size_t>(R, 1));
  for (uint32_t nv = 0; nv < V; ++nv) {
    for (uint32_t k = 0; k < K; ++k) {
      const size_t ro = size_t(vold[nv]) * K + k, rn = size_t(nv) * K + k;
      uint32_t wbits;
      std::memcpy(&wbits, &in.cubature_weights[ro], 4);
      const uint32_t t = in.cubature_tet_ids[ro];
      cub[10 * rn] = make_uint4(t < T ? tnew[t] : 0xffffffffu, wbits, 0, 0);
      std::memcpy(&cub[10 * rn + 1], &in.cubature_basis[36 * ro], 36 * sizeof(float));
    }
  }
  std::vector<float4> dev_objectives(2 * size_t(V));
  for (uint32_t nv = 0; nv < V; ++nv) {
    const float* o = &s->objectives[8 * size_t(vold[nv])];
    dev_objectives[2 * size_t(nv)] = make_float4(o[0], o[1], o[2], o[3]);
    dev_objectives[2 * size_t(nv) + 1] = make_float4(o[4], o[5], o[6], o[7]);
  }
  std::vector<float4> positions(V), velocities(V, make_float4(0, 0, 0, 0)), rot_v(3 * size_t(V)),
      rot_t(3 * size_t(T));
  for (uint32_t nv = 0; nv < V; ++nv) {
    const uint32_t ov = vold[nv];
    const float* x = &in.positions[4 * size_t(ov)];
    positions[nv] = make_float4(x[0], x[1], x[2], 1.0f);
    if (in.velocities)
      velocities[nv] = make_float4(in.velocities[4 * size_t(ov)], in.velocities[4 * size_t(ov) + 1],
                                   in.velocities[4 * size_t(ov) + 2], 0.0f);
    rot_v[3 * size_t(nv)] = make_float4(1, 0, 0, 0);
    rot_v[3 * size_t(nv) + 1] = make_float4(0, 1, 0, 0);
    rot_v[3 * size_t(nv) + 2] = make_float4(0, 0, 1, 0);
  }
  for (uint32_t t = 0; t < T; ++t) {
    rot_t[3 * size_t(t)] = make_float4(1, 0, 0, 0);
    rot_t[3 * size_t(t) + 1] = make_float4(0, 1, 0, 0);
    rot_t[3 * size_t(t) + 2] = make_float4(0, 0, 1, 0);
  }
  // Cloth records in device vertex ids; elements keep caller order.
  std::vector<uint4> tri_ix(NT), hinge_ix(NH);
  std::vector<float4> tri_inv(NT), hinge_rest(NH);
  std::vector<float> tri_area(NT);
  std::vector<uint32_t> tri_rows, tri_elems, hinge_rows, hinge_elems;
  if (in.cloth) {
    const jgs2_cloth_input& c = *in.cloth;
    for (uint32_t t = 0; t < NT; ++t) {
      const uint32_t* ix = &c.triangle_indices[3 * size_t(t)];
      const uint32_t owner = std::min(ix[0], std::min(ix[1], ix[2]));  // S:1838-1840
      tri_ix[t] = make_uint4(vnew[ix[0]], vnew[ix[1]], vnew[ix[2]], vnew[owner]);
      const float* b = &c.triangle_inverse_rest_bases[4 * size_t(t)];
      tri_inv[t] = make_float4(b[0], b[1], b[2], b[3]);
      tri_area[t] = c.triangle_rest_areas[t];
    }
    for (uint32_t h = 0; h < NH; ++h) {
      const uint32_t* ix = &c.hinge_indices[4 * size_t(h)];
      const uint32_t owner = std::min(std::min(ix[0], ix[1]), std::min(ix[2], ix[3]));  // S:1842-1844
      hinge_ix[h] = make_uint4(vnew[ix[0]], vnew[ix[1]], vnew[ix[2]], vnew[ix[3]]);
      uint32_t ob = vnew[owner];
      float of;
      std::memcpy(&of, &ob, 4);
      hinge_rest[h] = make_float4(c.hinge_rest_angles[h], c.hinge_rest_edge_lengths[h], of, 0.0f);
    }
    if (jgs2_status e = incident_csr(c.triangle_indices, 3, NT, vold, tri_rows, tri_elems)) return cleanup(e);
    if (jgs2_status e = incident_csr(c.hinge_indices, 4, NH, vold, hinge_rows, hinge_elems)) return cleanup(e);
    // packJGS2GpuClothArena material words (cloth-layout.ts:335-343): f64
    // expressions rounded to f32.
    const double E = c.young_modulus, nu = c.poisson_ratio;
    d.cloth_lam = float(E * nu / (1 - nu * nu));
    d.cloth_mu = float(E / (2 * (1 + nu)));
    d.cloth_thick = float(c.thickness);
    d.cloth_bend = float(c.bending_stiffness);
    d.NT = NT;
    d.NH = NH;
  }
  // IPC candidate records in device vertex ids; the contact parameters are
  // fixed at create (jgs2-solver.ts:1096-1105).
  std::vector<uint4> cand(NC);
  std::vector<uint32_t> cand_rows, cand_ids;
  if (NC > 0) {
    for (uint32_t c = 0; c < NC; ++c) {
      const uint32_t* ix = &in.contact->packed_indices[4 * size_t(c)];
      cand[c] = make_uint4(vnew[ix[0]], vnew[ix[1]], vnew[ix[2]], vnew[ix[3]]);
    }
    if (jgs2_status e = incident_csr(in.contact->packed_indices, 4, NC, vold, cand_rows, cand_ids)) return cleanup(e);
    d.NC = NC;
    d.NVT = in.contact->vertex_triangle_count;
  }
  d.ipc_dmin = float(resolved.ipc_minimum_distance);
  d.ipc_mu = float(resolved.ipc_friction_coefficient);
  d.ipc_eps = float(resolved.ipc_friction_velocity_epsilon);
  d.ipc_safety = float(resolved.ipc_step_safety);
  TRY(dalloc(&d.vtx_rest, V));
  TRY(dalloc(&d.vtx_info, V));
  TRY(dalloc(&d.tet_idx, T));
  TRY(dalloc(&d.tet_dm, 3 * size_t(T)));
  if (corotated) TRY(dalloc(&d.tet_k, 144 * size_t(T)));
  TRY(dalloc(&d.adjacency, std::max<uint64_t>(in.adjacency_length, 1)));
  TRY(dalloc(&d.cub, cub.size()));
  TRY(dalloc(&d.objectives, 2 * size_t(V)));
  TRY(dalloc(&d.color_order, V));
  TRY(dalloc(&d.colors, V));
  TRY(dalloc(&d.pos[0], V));
  TRY(dalloc(&d.pos[1], V));
  TRY(dalloc(&d.predicted, V));
  TRY(dalloc(&d.velocity, V));
  TRY(dalloc(&d.old, V));
  TRY(dalloc(&d.vrot, 3 * size_t(V)));
  TRY(dalloc(&d.vrot_next, 3 * size_t(V)));
  TRY(dalloc(&d.trot, 3 * size_t(T)));
  TRY(dalloc(&d.body_corr, s->bodies));
  TRY(dalloc(&d.final_update, V));
  TRY(dalloc(&d.local_diag, 4 * size_t(V)));
  TRY(dalloc(&d.tet_glob, 2 * (size_t(T) + NT)));
  if (in.cloth) {
    TRY(dalloc(&d.tri_ix, NT));
    TRY(dalloc(&d.tri_inv, NT));
    TRY(dalloc(&d.tri_area, NT));
    TRY(dalloc(&d.hinge_ix, NH));
    TRY(dalloc(&d.hinge_rest, NH));
    TRY(dalloc(&d.tri_rows, size_t(V) + 1));
    TRY(dalloc(&d.tri_elems, tri_elems.size()));
    TRY(dalloc(&d.hinge_rows, size_t(V) + 1));
    TRY(dalloc(&d.hinge_elems, hinge_elems.size()));
  }
  TRY(dalloc(&d.vertex_energy, V));
  TRY(dalloc(&d.conv, 5 * size_t(V)));
  TRY(dalloc(&d.control, 4));
  TRY(dalloc(&d.history, size_t(kHistoryCapacity) * 8));
  if (NC > 0) {
    TRY(dalloc(&d.cand, NC));
    for (float4** a : {&d.lag_n, &d.lag_w, &d.cscratch, &d.src_c, &d.src_w, &d.tgt_c, &d.tgt_w}) TRY(dalloc(a, NC));
    TRY(dalloc(&d.contact_ctl, 1));
    TRY(dalloc(&d.cand_rows, size_t(V) + 1));
    TRY(dalloc(&d.cand_ids, 4 * size_t(NC)));
    TRY(dalloc(&d.active_count, V));
    TRY(dalloc(&d.active_ids, 4 * size_t(NC)));
  }
  d.partial_capacity = partial_blocks_needed(V, T + NT, NC);
  TRY(dalloc(reinterpret_cast<char**>(&d.partials), partial_bytes(d.partial_capacity)));
  hipStream_t q = s->stream;
  TRY_HIP(hipMemcpyAsync(d.vtx_rest, vtx_rest.data(), V * sizeof(float4), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.vtx_info, vtx_info.data(), V * sizeof(uint4), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.tet_idx, tet_idx.data(), T * sizeof(uint4), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.tet_dm, tet_dm.data(), tet_dm.size() * sizeof(float4), hipMemcpyHostToDevice, q));
  if (corotated)
    TRY_HIP(hipMemcpyAsync(d.tet_k, tet_k.data(), 144 * size_t(T) * sizeof(float), hipMemcpyHostToDevice, q));
  if (in.adjacency_length)
    TRY_HIP(hipMemcpyAsync(d.adjacency, adjacency.data(), in.adjacency_length * 4, hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.cub, cub.data(), cub.size() * sizeof(uint4), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.objectives, dev_objectives.data(), 32 * size_t(V), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.color_order, order.data(), 4 * size_t(V), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.colors, dev_colors.data(), 4 * size_t(V), hipMemcpyHostToDevice, q));
  if (in.cloth) {
    auto up = [&](void* dst, const void* src, size_t bytes) {
      return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, q) : hipSuccess;
    };
    TRY_HIP(up(d.tri_ix, tri_ix.data(), NT * sizeof(uint4)));
    TRY_HIP(up(d.tri_inv, tri_inv.data(), NT * sizeof(float4)));
    TRY_HIP(up(d.tri_area, tri_area.data(), NT * sizeof(float)));
    TRY_HIP(up(d.hinge_ix, hinge_ix.data(), NH * sizeof(uint4)));
    TRY_HIP(up(d.hinge_rest, hinge_rest.data(), NH * sizeof(float4)));
    TRY_HIP(up(d.tri_rows, tri_rows.data(), tri_rows.size() * 4));
    TRY_HIP(up(d.tri_elems, tri_elems.data(), tri_elems.size() * 4));
    TRY_HIP(up(d.hinge_rows, hinge_rows.data(), hinge_rows.size() * 4));
    TRY_HIP(up(d.hinge_elems, hinge_elems.data(), hinge_elems.size() * 4));
  }
  if (NC > 0) {  // packIpcContactBuffer initial records (ipc-contact-layout.ts:302-341)
    std::vector<float4> zero4(NC, make_float4(0, 0, 0, 0)), scratch(NC, make_float4(1, 0, 0, 0)),
        none(NC, make_float4(0, 0, 0, -1));
    const float4 ctl = make_float4(0, 0, 0, 1);
    std::vector<uint32_t> zero_counts(V, 0);
    TRY_HIP(hipMemcpyAsync(d.cand, cand.data(), NC * sizeof(uint4), hipMemcpyHostToDevice, q));
    for (float4* a : {d.lag_n, d.lag_w, d.src_w, d.tgt_w})
      TRY_HIP(hipMemcpyAsync(a, zero4.data(), NC * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.cscratch, scratch.data(), NC * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.src_c, none.data(), NC * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.tgt_c, none.data(), NC * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.contact_ctl, &ctl, sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.cand_rows, cand_rows.data(), cand_rows.size() * 4, hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.cand_ids, cand_ids.data(), cand_ids.size() * 4, hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.active_count, zero_counts.data(), 4 * size_t(V), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.active_ids, cand_ids.data(), cand_ids.size() * 4, hipMemcpyHostToDevice, q));
    TRY_HIP(hipStreamSynchronize(q));
  }
  for (float4* region : {d.pos[0], d.pos[1], d.predicted, d.old})
    TRY_HIP(hipMemcpyAsync(region, positions.data(), V * sizeof(float4), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.velocity, velocities.data(), V * sizeof(float4), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.vrot, rot_v.data(), rot_v.size() * sizeof(float4), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.vrot_next, rot_v.data(), rot_v.size() * sizeof(float4), hipMemcpyHostToDevice, q));
  TRY_HIP(hipMemcpyAsync(d.trot, rot_t.data(), rot_t.size() * sizeof(float4), hipMemcpyHostToDevice, q));
  {
    std::vector<float4> zeros(std::max<size_t>({size_t(5) * V, size_t(4) * V, 2 * (size_t(T) + NT), size_t(kHistoryCapacity) * 8,
                                                size_t(s->bodies)}),
                              make_float4(0, 0, 0, 0));
    TRY_HIP(hipMemcpyAsync(d.body_corr, zeros.data(), s->bodies * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.final_update, zeros.data(), V * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.local_diag, zeros.data(), 4 * size_t(V) * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.tet_glob, zeros.data(), 2 * (size_t(T) + NT) * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.vertex_energy, zeros.data(), V * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.conv, zeros.data(), 5 * size_t(V) * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.control, zeros.data(), 4 * sizeof(float4), hipMemcpyHostToDevice, q));
    TRY_HIP(hipMemcpyAsync(d.history, zeros.data(), size_t(kHistoryCapacity) * 8 * sizeof(float4),
                           hipMemcpyHostToDevice, q));
    TRY_HIP(hipStreamSynchronize(q));
  }
  // GPU preflight of the uploaded f32 source (validateGpuInitialStableSource,
  // jgs2-solver.ts:820-893).
  if (globalization) {
    PassParams p = make_params(s, resolved, false);
    p.src = 0;
    p.tgt = 0;
    launch_reduce_candidate(d, p, q);
    TRY_HIP(hipGetLastError());
    float control[4];
    TRY_HIP(hipMemcpyAsync(control, d.control, sizeof(control), hipMemcpyDeviceToHost, q));
    TRY_HIP(hipStreamSynchronize(q));
    const uint32_t required = 1u | 2u | 4u;
    const float bits = control[3];
    const bool ok = std::isfinite(control[0]) && std::isfinite(control[1]) && std::isfinite(control[2]) &&
                    control[0] > kDeterminantFloor && control[1] > kDeterminantFloor && control[2] > kDeterminantFloor &&
                    bits == std::floor(bits) && (uint32_t(bits) & required) == required;
    if (!ok)
      return cleanup(fail(JGS2_ERROR_INFEASIBLE,
                          "The uploaded stable Neo-Hookean source pose is not feasible under the production GPU f32 "
                          "determinant calculation; source J=%g, candidate J=%g, accepted J=%g.",
                          double(control[0]), double(control[1]), double(control[2])));
  }
#undef TRY
#undef TRY_HIP
  *out = s;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_destroy(jgs2_solver* s) {
  if (!s) return JGS2_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_device(s->dev);
  for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
  if (s->frame_begin) (void)hipEventDestroy(s->frame_begin);
  if (s->frame_end) (void)hipEventDestroy(s->frame_end);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  s->destroyed = true;
  delete s;
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_step(jgs2_solver* s, const jgs2_settings* ov, uint64_t mask) {
  jgs2_status st = check_usable(s);
  if (st != JGS2_OK) return st;
  jgs2_settings r;
  if ((st = resolve_step(s, ov, mask, "step", r)) != JGS2_OK) return st;
  return submit(s, r, normalize_odd(r.iterations), 1, true);
}

jgs2_status jgs2_step_frames(jgs2_solver* s, uint32_t frames, const jgs2_settings* ov, uint64_t mask) {
  jgs2_status st = check_usable(s);
  if (st != JGS2_OK) return st;
  if ((st = validate_frames(frames)) != JGS2_OK) return st;
  jgs2_settings r;
  if ((st = resolve_step(s, ov, mask, "stepFrames", r)) != JGS2_OK) return st;
  return submit(s, r, normalize_odd(r.iterations), frames, true);
}

jgs2_status jgs2_step_exact_iterations(jgs2_solver* s, uint32_t iterations, const jgs2_settings* ov, uint64_t mask) {
  jgs2_status st = check_usable(s);
  if (st != JGS2_OK) return st;
  if (iterations < 1) return fail(JGS2_ERROR_RANGE, "iterations must be a positive safe integer; got %u.", iterations);
  jgs2_settings r;
  if ((st = resolve_step(s, ov, mask, "stepExactIterations", r)) != JGS2_OK) return st;
  return submit(s, r, iterations, 1, false);
}

jgs2_status jgs2_step_frames_exact_iterations(jgs2_solver* s, uint32_t frames, uint32_t iterations,
                                              const jgs2_settings* ov, uint64_t mask) {
  jgs2_status st = check_usable(s);
  if (st != JGS2_OK) return st;
  if ((st = validate_frames(frames)) != JGS2_OK) return st;
  if (iterations < 1) return fail(JGS2_ERROR_RANGE, "iterations must be a positive safe integer; got %u.", iterations);
  jgs2_settings r;
  if ((st = resolve_step(s, ov, mask, "stepFramesExactIterations", r)) != JGS2_OK) return st;
  return submit(s, r, iterations, frames, false);
}

jgs2_status jgs2_set_external_force(jgs2_solver* s, uint32_t vertex, const float force[3]) {
  jgs2_status st = objective_supported(s);
  if (st != JGS2_OK) return st;
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
  jgs2_status st = objective_supported(s);
  if (st != JGS2_OK) return st;
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
  jgs2_status st = objective_supported(s);
  if (st != JGS2_OK) return st;
  std::vector<float> zeros(3 * size_t(s->V), 0.0f);
  return jgs2_replace_external_forces(s, zeros.data(), zeros.size());
}

jgs2_status jgs2_set_quadratic_target(jgs2_solver* s, uint32_t vertex, const float target[3], float stiffness) {
  jgs2_status st = objective_supported(s);
  if (st != JGS2_OK) return st;
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
  jgs2_status st = objective_supported(s);
  if (st != JGS2_OK) return st;
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
  jgs2_status st = objective_supported(s);
  if (st != JGS2_OK) return st;
  if (vertex >= s->V) return fail(JGS2_ERROR_RANGE, "Objective vertex %u is outside vertexCount %u.", vertex, s->V);
  s->objectives[8 * size_t(vertex) + 7] = 0;
  return upload_objectives(s, vertex, 1);
}

jgs2_status jgs2_release_all_quadratic_targets(jgs2_solver* s) {
  jgs2_status st = objective_supported(s);
  if (st != JGS2_OK) return st;
  for (uint32_t v = 0; v < s->V; ++v) s->objectives[8 * size_t(v) + 7] = 0;
  return upload_objectives(s, 0, s->V);
}

jgs2_status jgs2_read_positions(jgs2_solver* s, float* dst, size_t count) {
  if (jgs2_status st = check_usable(s)) return st;
  return read_buffer(s, s->dev.pos[0], dst, count, 4 * size_t(s->V), "readPositions", 4);
}
jgs2_status jgs2_read_velocities(jgs2_solver* s, float* dst, size_t count) {
  if (jgs2_status st = check_usable(s)) return st;
  return read_buffer(s, s->dev.velocity, dst, count, 4 * size_t(s->V), "readVelocities", 4);
}
jgs2_status jgs2_read_predicted_positions(jgs2_solver* s, float* dst, size_t count) {
  if (jgs2_status st = check_usable(s)) return st;
  return read_buffer(s, s->dev.predicted, dst, count, 4 * size_t(s->V), "readPredictedPositions", 4);
}
jgs2_status jgs2_read_final_iteration_updates(jgs2_solver* s, float* dst, size_t count) {
  if (jgs2_status st = check_usable(s)) return st;
  return read_buffer(s, s->dev.final_update, dst, count, 4 * size_t(s->V), "readFinalIterationUpdates", 4);
}
jgs2_status jgs2_read_vertex_rotations(jgs2_solver* s, float* dst, size_t count) {
  if (jgs2_status st = check_usable(s)) return st;
  return read_buffer(s, s->dev.vrot, dst, count, 12 * size_t(s->V), "readVertexRotations", 12);
}
jgs2_status jgs2_read_local_diagnostics(jgs2_solver* s, float* dst, size_t count) {
  if (jgs2_status st = check_usable(s)) return st;
  if (!s->globalization) return fail(JGS2_ERROR_STATE, "Globalization diagnostics require the stable material path.");
  return read_buffer(s, s->dev.local_diag, dst, count, 16 * size_t(s->V), "readLocalDiagnostics", 16);
}
jgs2_status jgs2_read_globalization_control(jgs2_solver* s, float* dst, size_t count) {
  if (jgs2_status st = check_usable(s)) return st;
  return read_buffer(s, s->dev.control, dst, count, 16, "readGlobalizationControl");
}
jgs2_status jgs2_read_globalization_history(jgs2_solver* s, float* dst, size_t count) {
  if (jgs2_status st = check_usable(s)) return st;
  return read_buffer(s, s->dev.history, dst, count, size_t(kHistoryCapacity) * 32, "readGlobalizationHistory");
}

jgs2_status jgs2_sync(jgs2_solver* s) {
  if (jgs2_status st = check_usable(s)) return st;
  HIP_CHECK(hipStreamSynchronize(s->stream));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_get_info(jgs2_solver* s, jgs2_solver_info* out) {
  if (jgs2_status st = check_usable(s)) return st;
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
  if (jgs2_status st = check_usable(s)) return st;
  if (!colors || count != s->V) return fail(JGS2_ERROR_RANGE, "Schedule colors need exactly %u entries.", s->V);
  std::memcpy(colors, s->colors.data(), 4 * size_t(s->V));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_debug_local_systems(jgs2_solver* s, const jgs2_settings* ov, uint64_t mask, float* g, size_t gc,
                                     float* h, size_t hc) {
  if (jgs2_status st = check_usable(s)) return st;
  if (!g || !h || gc != 3 * size_t(s->V) || hc != 9 * size_t(s->V))
    return fail(JGS2_ERROR_RANGE, "Local systems need 3V gradient and 9V Hessian floats.");
  jgs2_settings r;
  if (jgs2_status st = resolve_step(s, ov, mask, "debugLocalSystems", r)) return st;
  DeviceState& d = s->dev;
  if (!d.debug_g) {
    if (jgs2_status st = dalloc(&d.debug_g, 3 * size_t(s->V))) return st;
    if (jgs2_status st = dalloc(&d.debug_h, 9 * size_t(s->V))) return st;
  }
  PassParams p = make_params(s, r, false);
  p.src = 0;
  p.tgt = 0;
  if (p.contact) {
    // Contact state of a frame start at the current pose: old = x, lagged
    // contact and active rows. The next frame's predict and lag pass rewrite
    // all of it, so stepping continues unaffected.
    HIP_CHECK(hipMemcpyAsync(d.old, d.pos[0], s->V * sizeof(float4), hipMemcpyDeviceToDevice, s->stream));
    launch_lag_contact(d, p, s->stream);
    launch_build_active_rows(d, p, s->stream);
  }
  if (s->corotated) launch_tet_polar(d, p, s->stream);
  launch_vertex_polar(d, p, s->stream);
  launch_debug_local_systems(d, p, s->stream);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s->stream));
  std::vector<float> gs(gc), hs(hc);
  HIP_CHECK(hipMemcpy(gs.data(), d.debug_g, gc * sizeof(float), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(hs.data(), d.debug_h, hc * sizeof(float), hipMemcpyDeviceToHost));
  for (uint32_t ov = 0; ov < s->V; ++ov) {
    std::memcpy(g + 3 * size_t(ov), &gs[3 * size_t(s->vnew[ov])], 3 * sizeof(float));
    std::memcpy(h + 9 * size_t(ov), &hs[9 * size_t(s->vnew[ov])], 9 * sizeof(float));
  }
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_profile_begin(jgs2_solver* s) {
  if (jgs2_status st = check_usable(s)) return st;
  if (s->events.empty()) {
    s->events.resize(8192);
    for (auto& e : s->events) HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventCreate(&s->frame_begin));
    HIP_CHECK(hipEventCreate(&s->frame_end));
  }
  HIP_CHECK(hipStreamSynchronize(s->stream));
  s->profiling = true;
  s->event_used = 0;
  HIP_CHECK(hipEventRecord(s->frame_begin, s->stream));
  clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_profile_end(jgs2_solver* s, double* ms_per_launch, uint64_t* launches, double* total_ms) {
  if (jgs2_status st = check_usable(s)) return st;
  if (!s->profiling) return fail(JGS2_ERROR_STATE, "jgs2_profile_begin was not called.");
  HIP_CHECK(hipEventRecord(s->frame_end, s->stream));
  HIP_CHECK(hipStreamSynchronize(s->stream));
  double sum = 0;
  for (size_t i = 0; i + 1 < s->event_used; i += 2) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, s->events[i], s->events[i + 1]));
    sum += ms;
  }
  const uint64_t n = s->event_used / 2;
  float total = 0;
  HIP_CHECK(hipEventElapsedTime(&total, s->frame_begin, s->frame_end));
  if (ms_per_launch) *ms_per_launch = n ? sum / double(n) : 0.0;
  if (launches) *launches = n;
  if (total_ms) *total_ms = total;
  s->profiling = false;
  clear_error();
  return JGS2_OK;
}

}  // extern "C"
