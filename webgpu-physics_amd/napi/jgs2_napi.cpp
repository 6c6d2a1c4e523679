// N-API addon over libjgs2hip: the binding a Node host of the reference's
// `JGS2GpuSolver` (src/simulation/gpu/jgs2-solver.ts:905-3051) loads in place
// of its WebGPU body. Each function takes the solver handle first and maps
// one C entry point (include/jgs2hip.h); jgs2hip.js wraps them in the
// reference's class. Inputs use the reference's camelCase field names
// (JGS2GpuInput, layout.ts:41-90; JGS2StepSettings, jgs2-solver.ts:64-100);
// typed arrays are read in place (the library deep-copies at create).
//
// Errors: the reference throws RangeError / TypeError / Error
// (layout.ts:223-416, jgs2-solver.ts:429-529, :3046-3050); a nonzero
// jgs2_status is thrown as the matching class with the library's message.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "jgs2hip.h"

namespace {

struct Handle {
  jgs2_solver* solver = nullptr;
  bool destroyed = false;
};

// Sets a pending JS exception; callers return nullptr right after.
void throw_status(napi_env env, jgs2_status st) {
  char msg[2048];
  jgs2_last_error(msg, sizeof msg);
  if (st == JGS2_ERROR_RANGE || st == JGS2_ERROR_INFEASIBLE || st == JGS2_ERROR_MEMORY)
    napi_throw_range_error(env, nullptr, msg);
  else if (st == JGS2_ERROR_TYPE)
    napi_throw_type_error(env, nullptr, msg);
  else
    napi_throw_error(env, nullptr, msg);
}

bool pending(napi_env env) {
  bool p = false;
  napi_is_exception_pending(env, &p);
  return p;
}

bool type_error(napi_env env, const std::string& msg) {
  napi_throw_type_error(env, nullptr, msg.c_str());
  return false;
}

std::vector<napi_value> args(napi_env env, napi_callback_info info, size_t n) {
  std::vector<napi_value> v(n);
  size_t argc = n;
  napi_get_cb_info(env, info, &argc, v.data(), nullptr, nullptr);
  napi_value undef;
  napi_get_undefined(env, &undef);
  for (size_t i = argc; i < n; ++i) v[i] = undef;
  return v;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_undefined || t == napi_null;
}

bool get_property(napi_env env, napi_value obj, const char* name, napi_value* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return false;
  napi_get_named_property(env, obj, name, out);
  return !is_nullish(env, *out);
}

bool get_number(napi_env env, napi_value v, const char* what, double* out) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return type_error(env, std::string(what) + " must be a number.");
  napi_get_value_double(env, v, out);
  return true;
}

bool get_u32(napi_env env, napi_value v, const char* what, uint32_t* out) {
  double d;
  if (!get_number(env, v, what, &d)) return false;
  if (!(d >= 0 && d <= 4294967295.0) || d != double(uint64_t(d))) {
    napi_throw_range_error(env, nullptr, (std::string(what) + " must be a nonnegative integer.").c_str());
    return false;
  }
  *out = uint32_t(d);
  return true;
}

constexpr size_t kAnyLength = SIZE_MAX;

bool range_error(napi_env env, const std::string& msg) {
  napi_throw_range_error(env, nullptr, msg.c_str());
  return false;
}

// A typed array property of the expected element type and, unless
// kAnyLength, of exactly `expected` elements (the C ABI takes bare
// pointers, so every length is checked here); absent -> nullptr unless
// required.
template <class T>
bool typed(napi_env env, napi_value obj, const char* name, napi_typedarray_type want, bool required, const T** out,
           size_t* length = nullptr, size_t expected = kAnyLength) {
  *out = nullptr;
  if (length) *length = 0;
  napi_value v;
  if (!get_property(env, obj, name, &v)) {
    if (required) return type_error(env, std::string(name) + " is required.");
    return true;
  }
  bool is = false;
  napi_is_typedarray(env, v, &is);
  napi_typedarray_type t;
  size_t n = 0, offset = 0;
  void* data = nullptr;
  napi_value buf;
  if (is) napi_get_typedarray_info(env, v, &t, &n, &data, &buf, &offset);
  if (!is || t != want)
    return type_error(env, std::string(name) + (want == napi_float32_array ? " must be a Float32Array."
                                                                          : " must be a Uint32Array."));
  if (expected != kAnyLength && n != expected)  // layout.ts:130-138
    return range_error(env, std::string(name) + " must contain " + std::to_string(expected) + " values; got " +
                                std::to_string(n) + ".");
  *out = static_cast<const T*>(data);
  if (length) *length = n;
  return true;
}

// Partial<JGS2StepSettings> -> jgs2_settings + the mask of present keys.
bool settings_from(napi_env env, napi_value obj, jgs2_settings* s, uint64_t* mask) {
  jgs2_default_settings(s);
  *mask = 0;
  if (is_nullish(env, obj)) return true;
  napi_valuetype t;
  napi_typeof(env, obj, &t);
  if (t != napi_object) return type_error(env, "Step settings must be an object.");
  napi_value v;
  if (get_property(env, obj, "schedule", &v)) {
    char buf[64] = {0};
    size_t n = 0;
    napi_value text = v;
    napi_coerce_to_string(env, v, &text);  // validateStepSettings prints String(schedule)
    napi_get_value_string_utf8(env, text, buf, sizeof buf, &n);
    if (std::strcmp(buf, "jacobi") == 0) s->schedule = JGS2_SCHEDULE_JACOBI;
    else if (std::strcmp(buf, "graph-colored-gauss-seidel") == 0) s->schedule = JGS2_SCHEDULE_GRAPH_COLORED_GAUSS_SEIDEL;
    else {
      napi_throw_range_error(env, nullptr, (std::string("Unknown JGS2 schedule: ") + buf + ".").c_str());
      return false;
    }
    *mask |= JGS2_SET_SCHEDULE;
  }
  if (get_property(env, obj, "gravity", &v)) {
    bool arr = false;
    napi_is_array(env, v, &arr);
    uint32_t len = 0;
    if (arr) napi_get_array_length(env, v, &len);
    if (!arr || len != 3) return type_error(env, "gravity must be an array of three numbers.");
    for (uint32_t i = 0; i < 3; ++i) {
      napi_value e;
      napi_get_element(env, v, i, &e);
      if (!get_number(env, e, "gravity", &s->gravity[i])) return false;
    }
    *mask |= JGS2_SET_GRAVITY;
  }
  if (get_property(env, obj, "iterations", &v)) {
    double d;
    if (!get_number(env, v, "iterations", &d)) return false;
    s->iterations = int32_t(d);
    *mask |= JGS2_SET_ITERATIONS;
  }
  struct Num {
    const char* name;
    double* field;
    uint64_t bit;
  };
  const Num nums[] = {
      {"timestep", &s->timestep, JGS2_SET_TIMESTEP},
      {"floorHeight", &s->floor_height, JGS2_SET_FLOOR_HEIGHT},
      {"floorStiffness", &s->floor_stiffness, JGS2_SET_FLOOR_STIFFNESS},
      {"velocityDamping", &s->velocity_damping, JGS2_SET_VELOCITY_DAMPING},
      {"regularization", &s->regularization, JGS2_SET_REGULARIZATION},
      {"rotationEpsilon", &s->rotation_epsilon, JGS2_SET_ROTATION_EPSILON},
      {"maxStep", &s->max_step, JGS2_SET_MAX_STEP},
      {"residualTolerance", &s->residual_tolerance, JGS2_SET_RESIDUAL_TOLERANCE},
      {"normalizedUpdateTolerance", &s->normalized_update_tolerance, JGS2_SET_NORMALIZED_UPDATE_TOLERANCE},
      {"contactTangentialDamping", &s->contact_tangential_damping, JGS2_SET_CONTACT_TANGENTIAL_DAMPING},
      {"contactMargin", &s->contact_margin, JGS2_SET_CONTACT_MARGIN},
      {"ipcActivationDistance", &s->ipc_activation_distance, JGS2_SET_IPC_ACTIVATION_DISTANCE},
      {"ipcMinimumDistance", &s->ipc_minimum_distance, JGS2_SET_IPC_MINIMUM_DISTANCE},
      {"ipcBarrierStiffness", &s->ipc_barrier_stiffness, JGS2_SET_IPC_BARRIER_STIFFNESS},
      {"ipcFrictionCoefficient", &s->ipc_friction_coefficient, JGS2_SET_IPC_FRICTION_COEFFICIENT},
      {"ipcFrictionVelocityEpsilon", &s->ipc_friction_velocity_epsilon, JGS2_SET_IPC_FRICTION_VELOCITY_EPSILON},
      {"ipcStepSafety", &s->ipc_step_safety, JGS2_SET_IPC_STEP_SAFETY},
  };
  for (const Num& q : nums)
    if (get_property(env, obj, q.name, &v)) {
      if (!get_number(env, v, q.name, q.field)) return false;
      *mask |= q.bit;
    }
  struct Flag {
    const char* name;
    int32_t* field;
    uint64_t bit;
  };
  const Flag flags[] = {{"horizontalBodyCorrection", &s->horizontal_body_correction, JGS2_SET_HORIZONTAL_BODY_CORRECTION},
                        {"parityMode", &s->parity_mode, JGS2_SET_PARITY_MODE}};
  for (const Flag& q : flags)
    if (get_property(env, obj, q.name, &v)) {
      bool b = false;
      if (napi_get_value_bool(env, v, &b) != napi_ok) return type_error(env, std::string(q.name) + " must be a boolean.");
      *q.field = b ? 1 : 0;
      *mask |= q.bit;
    }
  return true;
}

Handle* unwrap(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "Expected a JGS2GpuSolver handle.");
    return nullptr;
  }
  Handle* h = static_cast<Handle*>(p);
  if (h->destroyed) {
    napi_throw_error(env, nullptr, "JGS2GpuSolver has been destroyed.");
    return nullptr;
  }
  return h;
}

napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

// create(input: JGS2GpuInput, settings?): External handle.
napi_value Create(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  napi_valuetype t;
  napi_typeof(env, a[0], &t);
  if (t != napi_object) {
    napi_throw_type_error(env, nullptr, "JGS2GpuInput must be an object.");
    return nullptr;
  }
  jgs2_scene_input in{};
  napi_value v, obj = a[0];
  if (!get_property(env, obj, "vertexCount", &v) || !get_u32(env, v, "vertexCount", &in.vertex_count)) {
    if (!pending(env)) napi_throw_type_error(env, nullptr, "vertexCount is required.");
    return nullptr;
  }
  if (!get_property(env, obj, "tetCount", &v) || !get_u32(env, v, "tetCount", &in.tet_count)) {
    if (!pending(env)) napi_throw_type_error(env, nullptr, "tetCount is required.");
    return nullptr;
  }
  if (!get_property(env, obj, "cubatureK", &v) || !get_u32(env, v, "cubatureK", &in.cubature_k)) {
    if (!pending(env)) napi_throw_type_error(env, nullptr, "cubatureK is required.");
    return nullptr;
  }
  size_t adjacency_length = 0;
  const size_t V = in.vertex_count, T = in.tet_count, VK = V * in.cubature_k;
  const bool ok =
      typed(env, obj, "positions", napi_float32_array, true, &in.positions, nullptr, 4 * V) &&
      typed(env, obj, "velocities", napi_float32_array, false, &in.velocities, nullptr, 4 * V) &&
      typed(env, obj, "vertexRest", napi_float32_array, true, &in.vertex_rest, nullptr, 4 * V) &&
      typed(env, obj, "vertexInfo", napi_uint32_array, true, &in.vertex_info, nullptr, 4 * V) &&
      typed(env, obj, "tetIndices", napi_uint32_array, true, &in.tet_indices, nullptr, 4 * T) &&
      typed(env, obj, "tetInverseDm", napi_float32_array, true, &in.tet_inverse_dm, nullptr, 12 * T) &&
      typed(env, obj, "tetMeta", napi_float32_array, true, &in.tet_meta, nullptr, 4 * T) &&
      typed(env, obj, "tetRestStiffness", napi_float32_array, false, &in.tet_rest_stiffness, nullptr, 144 * T) &&
      typed(env, obj, "adjacency", napi_uint32_array, true, &in.adjacency, &adjacency_length) &&
      typed(env, obj, "cubatureTetIds", napi_uint32_array, true, &in.cubature_tet_ids, nullptr, VK) &&
      typed(env, obj, "cubatureWeights", napi_float32_array, true, &in.cubature_weights, nullptr, VK) &&
      typed(env, obj, "cubatureBasis", napi_float32_array, true, &in.cubature_basis, nullptr, 36 * VK);
  if (!ok) return nullptr;
  in.adjacency_length = adjacency_length;
  napi_value o;
  if (get_property(env, obj, "objectives", &o)) {
    if (!typed(env, o, "externalForces", napi_float32_array, false, &in.external_forces, nullptr, 3 * V) ||
        !typed(env, o, "targetPositions", napi_float32_array, false, &in.target_positions, nullptr, 3 * V) ||
        !typed(env, o, "targetStiffnesses", napi_float32_array, false, &in.target_stiffnesses, nullptr, V))
      return nullptr;
  }
  jgs2_cloth_input cloth{};
  if (get_property(env, obj, "cloth", &o)) {
    size_t nt = 0, nh = 0;
    double num;
    if (!typed(env, o, "triangleIndices", napi_uint32_array, true, &cloth.triangle_indices, &nt) ||
        !typed(env, o, "hingeIndices", napi_uint32_array, true, &cloth.hinge_indices, &nh))
      return nullptr;
    if (nt % 3 || nh % 4) {
      range_error(env, "cloth.triangleIndices needs 3 and cloth.hingeIndices 4 values per element.");
      return nullptr;
    }
    if (!typed(env, o, "triangleInverseRestBases", napi_float32_array, true, &cloth.triangle_inverse_rest_bases,
               nullptr, 4 * (nt / 3)) ||
        !typed(env, o, "triangleRestAreas", napi_float32_array, true, &cloth.triangle_rest_areas, nullptr, nt / 3) ||
        !typed(env, o, "hingeRestAngles", napi_float32_array, true, &cloth.hinge_rest_angles, nullptr, nh / 4) ||
        !typed(env, o, "hingeRestEdgeLengths", napi_float32_array, true, &cloth.hinge_rest_edge_lengths, nullptr,
               nh / 4))
      return nullptr;
    cloth.triangle_count = uint32_t(nt / 3);
    cloth.hinge_count = uint32_t(nh / 4);
    struct F {
      const char* name;
      double* field;
    };
    const F fs[] = {{"youngModulus", &cloth.young_modulus},
                    {"poissonRatio", &cloth.poisson_ratio},
                    {"thickness", &cloth.thickness},
                    {"bendingStiffness", &cloth.bending_stiffness}};
    for (const F& f : fs) {
      if (!get_property(env, o, f.name, &v) || !get_number(env, v, f.name, &num)) {
        if (!pending(env)) napi_throw_type_error(env, nullptr, (std::string("cloth.") + f.name + " is required.").c_str());
        return nullptr;
      }
      *f.field = num;
    }
    in.cloth = &cloth;
  }
  jgs2_contact_input contact{};
  napi_value probe;
  if (get_property(env, obj, "contactCandidates", &o) && get_property(env, o, "surfaceTriangles", &probe)) {
    // Device broad phase over a surface (no counterpart in the reference):
    // {surfaceTriangles, surfaceEdges, candidateCapacity?, reachMargin?}.
    size_t nt = 0, ne = 0;
    double num = 0;
    if (!typed(env, o, "surfaceTriangles", napi_uint32_array, true, &contact.surface_triangles, &nt) ||
        !typed(env, o, "surfaceEdges", napi_uint32_array, false, &contact.surface_edges, &ne))
      return nullptr;
    if (nt % 3 || ne % 2) {
      range_error(env, "contactCandidates.surfaceTriangles needs 3 and surfaceEdges 2 values per element.");
      return nullptr;
    }
    contact.surface_triangle_count = uint32_t(nt / 3);
    contact.surface_edge_count = uint32_t(ne / 2);
    if (get_property(env, o, "candidateCapacity", &v)) {
      if (!get_number(env, v, "candidateCapacity", &num)) return nullptr;
      if (!(num >= 0 && num < 9007199254740992.0) || num != double(uint64_t(num))) {
        range_error(env, "candidateCapacity must be a nonnegative integer.");
        return nullptr;
      }
      contact.candidate_capacity = uint64_t(num);
    }
    if (get_property(env, o, "reachMargin", &v)) {
      if (!get_number(env, v, "reachMargin", &num)) return nullptr;
      contact.reach_margin = num;
    }
    in.contact = &contact;
  } else if (get_property(env, obj, "contactCandidates", &o)) {
    if (!get_property(env, o, "vertexTriangleCount", &v) ||
        !get_u32(env, v, "vertexTriangleCount", &contact.vertex_triangle_count) ||
        !get_property(env, o, "edgeEdgeCount", &v) || !get_u32(env, v, "edgeEdgeCount", &contact.edge_edge_count)) {
      if (!pending(env)) napi_throw_type_error(env, nullptr, "contactCandidates needs both counts.");
      return nullptr;
    }
    const size_t candidates = size_t(contact.vertex_triangle_count) + contact.edge_edge_count;
    if (!typed(env, o, "packedIndices", napi_uint32_array, true, &contact.packed_indices, nullptr, 4 * candidates))
      return nullptr;
    in.contact = &contact;
  }
  jgs2_settings s;
  uint64_t mask = 0;
  if (!settings_from(env, a[1], &s, &mask)) return nullptr;
  jgs2_solver* solver = nullptr;
  if (jgs2_status st = jgs2_create(&in, &s, mask, &solver)) {
    throw_status(env, st);
    return nullptr;
  }
  Handle* h = new Handle{solver, false};
  napi_value out;
  napi_create_external(
      env, h,
      [](napi_env, void* p, void*) {
        Handle* q = static_cast<Handle*>(p);
        if (!q->destroyed) jgs2_destroy(q->solver);
        delete q;
      },
      nullptr, &out);
  return out;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  void* p = nullptr;
  if (napi_get_value_external(env, a[0], &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "Expected a JGS2GpuSolver handle.");
    return nullptr;
  }
  Handle* h = static_cast<Handle*>(p);
  if (!h->destroyed) {  // destroy() twice is a no-op (jgs2-solver.ts:3029-3032)
    h->destroyed = true;
    if (jgs2_status st = jgs2_destroy(h->solver)) {
      throw_status(env, st);
      return nullptr;
    }
  }
  return undefined(env);
}

// step(handle, overrides?) / stepFrames(handle, n, overrides?) /
// stepExactIterations(handle, iterations, overrides?) /
// stepFramesExactIterations(handle, n, iterations, overrides?).
template <int Counts, bool Exact>
napi_value Step(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2 + Counts);
  Handle* h = unwrap(env, a[0]);
  if (!h) return nullptr;
  uint32_t c[2] = {1, 0};
  for (int i = 0; i < Counts; ++i)
    if (!get_u32(env, a[1 + i], i == 0 && !(Exact && Counts == 1) ? "frameCount" : "iterations", &c[i])) return nullptr;
  jgs2_settings s;
  uint64_t mask = 0;
  if (!settings_from(env, a[1 + Counts], &s, &mask)) return nullptr;
  jgs2_status st;
  if (!Exact && Counts == 0) st = jgs2_step(h->solver, &s, mask);
  else if (!Exact) st = jgs2_step_frames(h->solver, c[0], &s, mask);
  else if (Counts == 1) st = jgs2_step_exact_iterations(h->solver, c[0], &s, mask);
  else st = jgs2_step_frames_exact_iterations(h->solver, c[0], c[1], &s, mask);
  if (st) {
    throw_status(env, st);
    return nullptr;
  }
  return undefined(env);
}

bool vec3(napi_env env, napi_value v, const char* what, float out[3]) {
  bool arr = false, ta = false;
  napi_is_array(env, v, &arr);
  napi_is_typedarray(env, v, &ta);
  if (!arr && !ta) return type_error(env, std::string(what) + " must hold three numbers.");
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    napi_get_element(env, v, i, &e);
    double d;
    if (!get_number(env, e, what, &d)) return false;
    out[i] = float(d);
  }
  return true;
}

napi_value SetExternalForce(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 3);
  Handle* h = unwrap(env, a[0]);
  uint32_t vtx;
  float f[3];
  if (!h || !get_u32(env, a[1], "vertex", &vtx) || !vec3(env, a[2], "force", f)) return nullptr;
  if (jgs2_status st = jgs2_set_external_force(h->solver, vtx, f)) return throw_status(env, st), nullptr;
  return undefined(env);
}

napi_value ReplaceExternalForces(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  Handle* h = unwrap(env, a[0]);
  if (!h) return nullptr;
  napi_value holder;
  napi_create_object(env, &holder);
  napi_set_named_property(env, holder, "forces", a[1]);
  const float* f = nullptr;
  size_t n = 0;
  if (!typed(env, holder, "forces", napi_float32_array, true, &f, &n)) return nullptr;
  if (jgs2_status st = jgs2_replace_external_forces(h->solver, f, n)) return throw_status(env, st), nullptr;
  return undefined(env);
}

template <jgs2_status (*Fn)(jgs2_solver*)>
napi_value NoArg(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  Handle* h = unwrap(env, a[0]);
  if (!h) return nullptr;
  if (jgs2_status st = Fn(h->solver)) return throw_status(env, st), nullptr;
  return undefined(env);
}

napi_value SetQuadraticTarget(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 4);
  Handle* h = unwrap(env, a[0]);
  uint32_t vtx;
  float p[3];
  double k;
  if (!h || !get_u32(env, a[1], "vertex", &vtx) || !vec3(env, a[2], "target", p) ||
      !get_number(env, a[3], "stiffness", &k))
    return nullptr;
  if (jgs2_status st = jgs2_set_quadratic_target(h->solver, vtx, p, float(k))) return throw_status(env, st), nullptr;
  return undefined(env);
}

napi_value ReplaceQuadraticTargets(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 3);
  Handle* h = unwrap(env, a[0]);
  if (!h) return nullptr;
  napi_value holder;
  napi_create_object(env, &holder);
  napi_set_named_property(env, holder, "positions", a[1]);
  napi_set_named_property(env, holder, "stiffnesses", a[2]);
  const float *p = nullptr, *k = nullptr;
  size_t np = 0, nk = 0;
  if (!typed(env, holder, "positions", napi_float32_array, true, &p, &np) ||
      !typed(env, holder, "stiffnesses", napi_float32_array, true, &k, &nk))
    return nullptr;
  if (jgs2_status st = jgs2_replace_quadratic_targets(h->solver, p, np, k, nk)) return throw_status(env, st), nullptr;
  return undefined(env);
}

napi_value ReleaseQuadraticTarget(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  Handle* h = unwrap(env, a[0]);
  uint32_t vtx;
  if (!h || !get_u32(env, a[1], "vertex", &vtx)) return nullptr;
  if (jgs2_status st = jgs2_release_quadratic_target(h->solver, vtx)) return throw_status(env, st), nullptr;
  return undefined(env);
}

// Reads: fresh Float32Array copies (jgs2-solver.ts:2996-3027).
template <jgs2_status (*Fn)(jgs2_solver*, float*, size_t), size_t PerVertex, size_t Fixed>
napi_value Read(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  Handle* h = unwrap(env, a[0]);
  if (!h) return nullptr;
  jgs2_solver_info si;
  if (jgs2_status st = jgs2_get_info(h->solver, &si)) return throw_status(env, st), nullptr;
  const size_t n = PerVertex * size_t(si.vertex_count) + Fixed;
  void* data = nullptr;
  napi_value buf, arr;
  napi_create_arraybuffer(env, n * sizeof(float), &data, &buf);
  if (jgs2_status st = Fn(h->solver, static_cast<float*>(data), n)) return throw_status(env, st), nullptr;
  napi_create_typedarray(env, napi_float32_array, n, buf, 0, &arr);
  return arr;
}

napi_value Info(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  Handle* h = unwrap(env, a[0]);
  if (!h) return nullptr;
  jgs2_solver_info si;
  if (jgs2_status st = jgs2_get_info(h->solver, &si)) return throw_status(env, st), nullptr;
  napi_value o;
  napi_create_object(env, &o);
  auto put = [&](const char* name, double x) {
    napi_value v;
    napi_create_double(env, x, &v);
    napi_set_named_property(env, o, name, v);
  };
  put("vertexCount", si.vertex_count);
  put("tetCount", si.tet_count);
  put("cubatureK", si.cubature_k);
  put("bodyCount", si.body_count);
  put("scheduleColorCount", si.schedule_color_count);
  put("freeVertexCount", si.free_vertex_count);
  put("sceneScale", si.scene_scale);
  napi_value b;
  napi_get_boolean(env, si.globalization_enabled != 0, &b);
  napi_set_named_property(env, o, "globalizationEnabled", b);
  return o;
}

// readOracleDiagnostics(handle, overrides?, positions?, predicted?, records?):
// the FP64 summary under the C struct's names plus, with `records`, the
// per-vertex (36 doubles) and per-tetrahedron (4 doubles) records as
// Float64Arrays (include/jgs2hip.h states both layouts).
napi_value ReadOracleDiagnostics(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 5);
  Handle* h = unwrap(env, a[0]);
  if (!h) return nullptr;
  jgs2_settings s;
  uint64_t mask = 0;
  if (!settings_from(env, a[1], &s, &mask)) return nullptr;
  jgs2_solver_info si;
  if (jgs2_status st = jgs2_get_info(h->solver, &si)) return throw_status(env, st), nullptr;
  napi_value holder;
  napi_create_object(env, &holder);
  napi_set_named_property(env, holder, "positions", a[2]);
  napi_set_named_property(env, holder, "predicted", a[3]);
  const float *pos = nullptr, *pred = nullptr;
  size_t npos = 0, npred = 0;
  if (!typed(env, holder, "positions", napi_float32_array, false, &pos, &npos) ||
      !typed(env, holder, "predicted", napi_float32_array, false, &pred, &npred))
    return nullptr;
  if ((pos && npos != 4 * size_t(si.vertex_count)) || (pred && npred != 4 * size_t(si.vertex_count))) {
    napi_throw_range_error(env, nullptr, "Diagnostic poses need four floats per vertex.");
    return nullptr;
  }
  bool records = false;
  if (!is_nullish(env, a[4])) napi_get_value_bool(env, a[4], &records);
  const size_t nv = records ? 36 * size_t(si.vertex_count) : 0, nt = records ? 4 * size_t(si.tet_count) : 0;
  void *vdata = nullptr, *tdata = nullptr;
  napi_value vbuf = nullptr, tbuf = nullptr;
  if (records) {
    napi_create_arraybuffer(env, nv * sizeof(double), &vdata, &vbuf);
    napi_create_arraybuffer(env, (nt ? nt : 1) * sizeof(double), &tdata, &tbuf);
  }
  jgs2_diagnostics d{};
  if (jgs2_status st = jgs2_read_oracle_diagnostics(h->solver, &s, mask, pos, pred, &d, static_cast<double*>(vdata), nv,
                                                    nt ? static_cast<double*>(tdata) : nullptr, nt))
    return throw_status(env, st), nullptr;
  napi_value o;
  napi_create_object(env, &o);
  auto put = [&](const char* name, double x) {
    napi_value v;
    napi_create_double(env, x, &v);
    napi_set_named_property(env, o, name, v);
  };
  const std::pair<const char*, double> fields[] = {
      {"energy", d.energy},
      {"inertia", d.inertia},
      {"elasticity", d.elasticity},
      {"externalForce", d.external_force},
      {"quadraticTarget", d.quadratic_target},
      {"floorContact", d.floor_contact},
      {"contactBarrier", d.contact_barrier},
      {"gradientNormInertia", d.gradient_norm_inertia},
      {"gradientNormMaterial", d.gradient_norm_material},
      {"gradientNormExternalForce", d.gradient_norm_external_force},
      {"gradientNormTarget", d.gradient_norm_target},
      {"gradientNormFloor", d.gradient_norm_floor},
      {"gradientNorm", d.gradient_norm},
      {"residualDenominator", d.residual_denominator},
      {"relativeResidual", d.relative_residual},
      {"maximumUpdate", d.maximum_update},
      {"minimumDeformationDeterminant", d.minimum_deformation_determinant},
      {"minimumFloorDistance", d.minimum_floor_distance},
      {"minimumContactDistance", d.minimum_contact_distance},
      {"activeFloorContacts", double(d.active_floor_contacts)},
      {"activeVertices", double(d.active_vertices)},
      {"activeContactPairs", double(d.active_contact_pairs)},
      {"finite", double(d.finite)},
      {"maximumUpdateValid", double(d.maximum_update_valid)},
      {"minimumDeterminantValid", double(d.minimum_determinant_valid)},
      {"floorMetricsValid", double(d.floor_metrics_valid)},
  };
  for (const auto& f : fields) put(f.first, f.second);
  if (records) {
    napi_value va, ta;
    napi_create_typedarray(env, napi_float64_array, nv, vbuf, 0, &va);
    napi_create_typedarray(env, napi_float64_array, nt, tbuf, 0, &ta);
    napi_set_named_property(env, o, "vertexRecords", va);
    napi_set_named_property(env, o, "tetrahedronRecords", ta);
  }
  return o;
}

napi_value DeviceCount(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_int32(env, jgs2_device_count(), &v);
  return v;
}

// The C entry points this binding calls, for hosts that audit the binding.
const char* const kBound[] = {
    "jgs2_create", "jgs2_destroy", "jgs2_step", "jgs2_step_frames", "jgs2_step_exact_iterations",
    "jgs2_step_frames_exact_iterations", "jgs2_set_external_force", "jgs2_replace_external_forces",
    "jgs2_clear_external_forces", "jgs2_set_quadratic_target", "jgs2_replace_quadratic_targets",
    "jgs2_release_quadratic_target", "jgs2_release_all_quadratic_targets", "jgs2_read_positions",
    "jgs2_read_velocities", "jgs2_read_predicted_positions", "jgs2_read_final_iteration_updates",
    "jgs2_read_vertex_rotations", "jgs2_read_local_diagnostics", "jgs2_read_globalization_control",
    "jgs2_read_globalization_history", "jgs2_read_oracle_diagnostics", "jgs2_sync", "jgs2_get_info",
    "jgs2_device_count", "jgs2_last_error", "jgs2_default_settings"};

napi_value Bound(napi_env env, napi_callback_info) {
  napi_value arr;
  const size_t n = sizeof(kBound) / sizeof(kBound[0]);
  napi_create_array_with_length(env, n, &arr);
  for (size_t i = 0; i < n; ++i) {
    napi_value s;
    napi_create_string_utf8(env, kBound[i], NAPI_AUTO_LENGTH, &s);
    napi_set_element(env, arr, uint32_t(i), s);
  }
  return arr;
}

napi_value Init(napi_env env, napi_value exports) {
  const struct {
    const char* name;
    napi_callback fn;
  } fns[] = {
      {"create", Create},
      {"destroy", Destroy},
      {"step", Step<0, false>},
      {"stepFrames", Step<1, false>},
      {"stepExactIterations", Step<1, true>},
      {"stepFramesExactIterations", Step<2, true>},
      {"setExternalForce", SetExternalForce},
      {"replaceExternalForces", ReplaceExternalForces},
      {"clearExternalForces", NoArg<jgs2_clear_external_forces>},
      {"setQuadraticTarget", SetQuadraticTarget},
      {"replaceQuadraticTargets", ReplaceQuadraticTargets},
      {"releaseQuadraticTarget", ReleaseQuadraticTarget},
      {"releaseAllQuadraticTargets", NoArg<jgs2_release_all_quadratic_targets>},
      {"readPositions", Read<jgs2_read_positions, 4, 0>},
      {"readVelocities", Read<jgs2_read_velocities, 4, 0>},
      {"readPredictedPositions", Read<jgs2_read_predicted_positions, 4, 0>},
      {"readFinalIterationUpdates", Read<jgs2_read_final_iteration_updates, 4, 0>},
      {"readVertexRotations", Read<jgs2_read_vertex_rotations, 12, 0>},
      {"readLocalDiagnostics", Read<jgs2_read_local_diagnostics, 16, 0>},
      {"readGlobalizationControl", Read<jgs2_read_globalization_control, 0, 16>},
      {"readGlobalizationHistory", Read<jgs2_read_globalization_history, 0, 64 * 32>},
      {"readOracleDiagnostics", ReadOracleDiagnostics},
      {"awaitIdle", NoArg<jgs2_sync>},
      {"info", Info},
      {"deviceCount", DeviceCount},
      {"boundEntryPoints", Bound},
  };
  for (const auto& f : fns) {
    napi_value v;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v);
    napi_set_named_property(env, exports, f.name, v);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
