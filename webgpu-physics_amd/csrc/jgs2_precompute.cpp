// Cubature precompute: per free vertex, K co-rotated cubature records
// (sample tetrahedron, nonnegative weight, four rest-basis blocks) for the
// JGS2 local solve. It is the step the reference runs on the host once per
// scene, buildPrecomputedScene (src/simulation/cpu/precompute.ts:218-380),
// laid out so that it scales with the mesh:
//
//  * the rest Hessian M/h^2 + K (src/simulation/cpu/fem.ts:89-335) is held as
//    3x3 blocks in CSR over free vertices, never dense;
//  * each vertex's exact basis U = H^-1 S (S^T H^-1 S)^-1 (precompute.ts:
//    135-200) is solved either over the whole mesh (rings = 0, the
//    reference's definition: one factorization shared by every vertex) or over
//    the vertices within `rings` edges of the source, with the basis zero
//    outside (a local dense factorization per vertex, O(ring^3));
//  * candidate tetrahedra are those touching the basis support: every other
//    tetrahedron's training column is exactly zero, so restricting the
//    candidates changes nothing but the cost;
//  * the low-frequency modes of the pose corpus (cubature.ts:224-297,
//    nonlinear-cubature.ts:338-620) come from block inverse iteration whose
//    solves use the dense factor for small systems and block-Jacobi
//    preconditioned FP64 conjugate gradients otherwise, on the GPU when one is
//    present (jgs2_precompute_kernels.hip: every mode advances in one pass
//    over the resident Hessian);
//  * greedy nonnegative selection (cubature.ts:386-691) uses a memoized Gram
//    matrix and evaluates the NNLS masks shared by every candidate of a step
//    once, replaying them in mask order so every comparison is the
//    reference's;
//  * vertices train in parallel on a thread pool.
//
// Stable Neo-Hookean scenes train on the nonlinear pose corpus with co-rotated
// bases and f32 packing (nonlinear-cubature-training.ts:311-843);
// co-rotated linear scenes train on displacement modes (cubature.ts:299-369,
// 714-826).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>

#include "jgs2_internal.h"

namespace jgs2 {
namespace {

using Vec = std::vector<double>;
constexpr double kEps = 1e-14;
constexpr size_t kDenseModeLimit = 1536;  // free coordinates factored densely (serial O(n^3))

struct Error {
  jgs2_status status;
  std::string message;
};

[[noreturn]] void raise(jgs2_status status, const std::string& message) { throw Error{status, message}; }

// ------------------------------------------------------------ 3x3 helpers
double det3(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

void inv3(const double* m, double* o) {
  const double d = det3(m);
  double s = 0.0;
  for (int i = 0; i < 9; ++i) s = std::max(s, std::fabs(m[i]));
  if (s == 0.0 || std::fabs(d) <= kEps * s * s * s) raise(JGS2_ERROR_RANGE, "Cannot invert a singular 3 by 3 matrix.");
  const double r = 1.0 / d;
  const double t[9] = {(m[4] * m[8] - m[5] * m[7]) * r, (m[2] * m[7] - m[1] * m[8]) * r,
                       (m[1] * m[5] - m[2] * m[4]) * r, (m[5] * m[6] - m[3] * m[8]) * r,
                       (m[0] * m[8] - m[2] * m[6]) * r, (m[2] * m[3] - m[0] * m[5]) * r,
                       (m[3] * m[7] - m[4] * m[6]) * r, (m[1] * m[6] - m[0] * m[7]) * r,
                       (m[0] * m[4] - m[1] * m[3]) * r};
  std::memcpy(o, t, sizeof t);
}

void mul3(const double* a, const double* b, double* o) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) t[r * 3 + c] = a[r * 3] * b[c] + a[r * 3 + 1] * b[3 + c] + a[r * 3 + 2] * b[6 + c];
  std::memcpy(o, t, sizeof t);
}

// a * b^T
void mul3t(const double* a, const double* b, double* o) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) t[r * 3 + c] = a[r * 3] * b[c * 3] + a[r * 3 + 1] * b[c * 3 + 1] + a[r * 3 + 2] * b[c * 3 + 2];
  std::memcpy(o, t, sizeof t);
}

double dotn(const double* a, const double* b, size_t n) {
  double r = 0.0;
  for (size_t i = 0; i < n; ++i) r += a[i] * b[i];
  return r;
}

// Math.hypot as ECMAScript engines evaluate it: scale by the largest
// magnitude, Kahan-compensated sum of squares.
double hypot_js(const double* v, int n) {
  double mx = 0.0;
  for (int i = 0; i < n; ++i) mx = std::max(mx, std::fabs(v[i]));
  if (mx == 0.0) return 0.0;
  if (std::isinf(mx)) return INFINITY;
  double sum = 0.0, comp = 0.0;
  for (int i = 0; i < n; ++i) {
    const double q = std::fabs(v[i]) / mx;
    const double term = q * q - comp;
    const double next = sum + term;
    comp = (next - sum) - term;
    sum = next;
  }
  return std::sqrt(sum) * mx;
}

// Dense Cholesky-Banachiewicz with the reference's positivity floor
// (math.ts:56-99) and its triangular solves (math.ts:101-132).
bool cholesky(Vec& a, size_t n) {
  double largest = 0.0;
  for (size_t i = 0; i < n; ++i) largest = std::max(largest, std::fabs(a[i * n + i]));
  const double floor = std::max(largest, 1.0) * 1e-13;
  for (size_t r = 0; r < n; ++r) {
    double* lr = &a[r * n];
    for (size_t c = 0; c <= r; ++c) {
      const double* lc = &a[c * n];
      double v = lr[c];
      for (size_t k = 0; k < c; ++k) v -= lr[k] * lc[k];
      if (r == c) {
        if (!(v > floor) || !std::isfinite(v)) return false;
        lr[c] = std::sqrt(v);
      } else {
        lr[c] = v / lc[c];
      }
    }
    for (size_t c = r + 1; c < n; ++c) lr[c] = 0.0;
  }
  return true;
}

void cholesky_solve(const Vec& l, size_t n, double* x) {
  for (size_t r = 0; r < n; ++r) {
    double v = x[r];
    const double* lr = &l[r * n];
    for (size_t c = 0; c < r; ++c) v -= lr[c] * x[c];
    x[r] = v / lr[r];
  }
  for (size_t r = n; r-- > 0;) {
    double v = x[r];
    for (size_t c = r + 1; c < n; ++c) v -= l[c * n + r] * x[c];
    x[r] = v / l[r * n + r];
  }
}

// Gaussian elimination with partial pivoting; false below the reference's
// absolute pivot floor (math.ts:134-199). In place, n <= 12.
bool small_solve(double* a, double* b, int n) {
  for (int p = 0; p < n; ++p) {
    int best = p;
    double mag = std::fabs(a[p * n + p]);
    for (int r = p + 1; r < n; ++r)
      if (std::fabs(a[r * n + p]) > mag) {
        mag = std::fabs(a[r * n + p]);
        best = r;
      }
    if (mag <= kEps || !std::isfinite(mag)) return false;
    if (best != p) {
      for (int c = p; c < n; ++c) std::swap(a[p * n + c], a[best * n + c]);
      std::swap(b[p], b[best]);
    }
    const double d = a[p * n + p];
    for (int r = p + 1; r < n; ++r) {
      const double m = a[r * n + p] / d;
      a[r * n + p] = 0.0;
      for (int c = p + 1; c < n; ++c) a[r * n + c] -= m * a[p * n + c];
      b[r] -= m * b[p];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double v = b[r];
    for (int c = r + 1; c < n; ++c) v -= a[r * n + c] * b[c];
    b[r] = v / a[r * n + r];
  }
  return true;
}

// ------------------------------------------------------------ thread pool
// Runs body(i, worker) for i in [0, count) on `threads` workers (the caller
// is worker 0). The error of the lowest failing index is rethrown, so the
// reported failure does not depend on scheduling.
template <class F>
void parallel_for(size_t count, unsigned threads, F&& body) {
  if (count == 0) return;
  threads = std::max(1u, std::min<unsigned>(threads, unsigned(std::min<size_t>(count, 1u << 16))));
  std::atomic<size_t> next{0};
  std::mutex m;
  Error first{JGS2_OK, {}};
  size_t first_index = std::numeric_limits<size_t>::max();
  auto record = [&](size_t i, Error e) {
    std::lock_guard<std::mutex> g(m);
    if (i < first_index) {
      first_index = i;
      first = std::move(e);
    }
  };
  auto worker = [&](unsigned id) {
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= count) return;
      try {
        body(i, id);
      } catch (const Error& e) {
        record(i, e);
      } catch (const std::bad_alloc&) {
        record(i, Error{JGS2_ERROR_MEMORY, "Cubature precompute: host allocation failed."});
      } catch (const std::exception& e) {
        record(i, Error{JGS2_ERROR_RANGE, e.what()});
      }
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; ++t) pool.emplace_back(worker, t);
  worker(0);
  for (auto& t : pool) t.join();
  if (first.status != JGS2_OK) throw first;
}

// ------------------------------------------------------------ the scene
struct Precompute {
  uint32_t V = 0, T = 0, K = 0;
  const double* x = nullptr;
  const uint32_t* tets = nullptr;
  std::vector<uint8_t> fixed;
  std::vector<uint32_t> mat;
  std::vector<jgs2_precompute_material> materials;
  double h = 0.0;
  bool snh = false;
  unsigned threads = 1;
  uint32_t rings = 0;
  double cg_tolerance = 1e-12;
  double residual_bound = 0.01;
  bool exact_all_elements = false;
  uint32_t mode_solver = 0;

  Vec vol, dminv, shape;  // per tetrahedron: volume, Dm^-1 (9), shape gradients (12)
  Vec mass;
  std::vector<uint32_t> incident;
  std::vector<uint32_t> vt_start, vt;      // vertex -> incident tetrahedra, ascending
  std::vector<uint32_t> vv_start, vv;      // vertex -> neighbor vertices, ascending (self excluded)
  std::vector<int32_t> active_of;          // vertex -> free index, -1 when fixed
  std::vector<uint32_t> active;            // free index -> vertex
  // Rest Hessian over free vertices, 3x3 blocks in CSR (diagonal included).
  std::vector<uint32_t> hb_start, hb_col;  // columns are free indices
  Vec hb;                                  // 9 per block
  size_t n = 0;                            // 3 * free vertices
  Vec dense_l;                             // dense factor when n <= kDenseModeLimit
  bool dense = false;
  std::shared_ptr<GpuCg> gpu;              // device CG for the mode solves of large systems
  mutable double gpu_ms = 0.0;
  mutable uint64_t gpu_iterations = 0;
};

void validate_and_load(Precompute& p, const jgs2_precompute_input& in) {
  if (in.vertex_count == 0 || in.tetrahedron_count == 0)
    raise(JGS2_ERROR_RANGE, "A cubature precompute needs at least one vertex and one tetrahedron.");
  if (!in.positions || !in.tetrahedra || !in.fixed || !in.material_ids || !in.materials || in.material_count == 0)
    raise(JGS2_ERROR_TYPE, "Cubature precompute input arrays must not be NULL.");
  if (in.cubature_samples == 0 || (in.cubature_samples > 12 && !p.exact_all_elements))
    raise(JGS2_ERROR_RANGE, "Cubature samples must be between 1 and 12.");
  if (!(in.timestep > 0) || !std::isfinite(in.timestep)) raise(JGS2_ERROR_RANGE, "The timestep must be finite and positive.");
  p.V = in.vertex_count;
  p.T = in.tetrahedron_count;
  p.K = in.cubature_samples;
  p.x = in.positions;
  p.tets = in.tetrahedra;
  p.h = in.timestep;
  for (size_t i = 0; i < 3 * size_t(p.V); ++i)
    if (!std::isfinite(p.x[i])) raise(JGS2_ERROR_RANGE, "Rest positions must be finite.");
  for (size_t i = 0; i < 4 * size_t(p.T); ++i)
    if (p.tets[i] >= p.V) raise(JGS2_ERROR_RANGE, "Tetrahedron references a vertex out of range.");
  p.fixed.assign(in.fixed, in.fixed + p.V);
  p.mat.assign(in.material_ids, in.material_ids + p.T);
  p.materials.assign(in.materials, in.materials + in.material_count);
  for (const auto& m : p.materials) {
    if (m.model != JGS2_MATERIAL_COROTATED_LINEAR && m.model != JGS2_MATERIAL_STABLE_NEO_HOOKEAN)
      raise(JGS2_ERROR_RANGE, "Material has invalid model.");
    if (!(m.density > 0) || !std::isfinite(m.density)) raise(JGS2_ERROR_RANGE, "Material has invalid density.");
    if (!(m.young_modulus > 0) || !std::isfinite(m.young_modulus))
      raise(JGS2_ERROR_RANGE, "Material has invalid Young's modulus.");
    if (!(m.poisson_ratio > -1 && m.poisson_ratio < 0.5)) raise(JGS2_ERROR_RANGE, "Material has invalid Poisson ratio.");
  }
  bool any_snh = false, any_linear = false;
  for (uint32_t t = 0; t < p.T; ++t) {
    if (p.mat[t] >= p.materials.size()) raise(JGS2_ERROR_RANGE, "Tetrahedron references an unknown material.");
    (p.materials[p.mat[t]].model == JGS2_MATERIAL_STABLE_NEO_HOOKEAN ? any_snh : any_linear) = true;
  }
  if (any_snh && any_linear)
    raise(JGS2_ERROR_RANGE, "Mixed stable Neo-Hookean/co-rotated nonlinear Cubature preprocessing is not implemented.");
  p.snh = any_snh;
}

void tet_stiffness(const Precompute& p, uint32_t t, double* K) {  // fem.ts:89-165
  const double* g = &p.shape[12 * size_t(t)];
  double B[72] = {0};
  for (int v = 0; v < 4; ++v) {
    const double gx = g[v * 3], gy = g[v * 3 + 1], gz = g[v * 3 + 2];
    const int c = v * 3;
    B[c] = gx;
    B[12 + c + 1] = gy;
    B[24 + c + 2] = gz;
    B[36 + c] = gy;
    B[36 + c + 1] = gx;
    B[48 + c + 1] = gz;
    B[48 + c + 2] = gy;
    B[60 + c] = gz;
    B[60 + c + 2] = gx;
  }
  const auto& m = p.materials[p.mat[t]];
  const double lambda = (m.young_modulus * m.poisson_ratio) / ((1 + m.poisson_ratio) * (1 - 2 * m.poisson_ratio));
  const double mu = m.young_modulus / (2 * (1 + m.poisson_ratio));
  double D[36] = {0};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) D[r * 6 + c] = r == c ? lambda + 2 * mu : lambda;
  D[21] = D[28] = D[35] = mu;
  double DB[72];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 12; ++c) {
      double v = 0.0;
      for (int k = 0; k < 6; ++k) v += D[r * 6 + k] * B[k * 12 + c];
      DB[r * 12 + c] = v;
    }
  const double vol = p.vol[t];
  for (int r = 0; r < 12; ++r)
    for (int c = r; c < 12; ++c) {
      double v = 0.0;
      for (int k = 0; k < 6; ++k) v += B[k * 12 + r] * DB[k * 12 + c];
      v *= vol;
      K[r * 12 + c] = v;
      K[c * 12 + r] = v;
    }
}

void build_rest(Precompute& p) {
  p.vol.assign(p.T, 0.0);
  p.dminv.assign(9 * size_t(p.T), 0.0);
  p.shape.assign(12 * size_t(p.T), 0.0);
  parallel_for(p.T, p.threads, [&](size_t t, unsigned) {
    const uint32_t* q = p.tets + 4 * t;
    double m[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m[r * 3 + c] = p.x[3 * size_t(q[c + 1]) + r] - p.x[3 * size_t(q[0]) + r];
    const double vol = std::fabs(det3(m)) / 6.0;
    if (!(vol > 1e-14) || !std::isfinite(vol)) raise(JGS2_ERROR_RANGE, "Cannot assemble a degenerate tetrahedron.");
    double* inv = &p.dminv[9 * t];
    inv3(m, inv);
    double* g = &p.shape[12 * t];
    for (int c = 0; c < 3; ++c) {
      g[3 + c] = inv[c];
      g[6 + c] = inv[3 + c];
      g[9 + c] = inv[6 + c];
      g[c] = -g[3 + c] - g[6 + c] - g[9 + c];
    }
    p.vol[t] = vol;
  });
  p.mass.assign(p.V, 0.0);
  p.incident.assign(p.V, 0);
  for (uint32_t t = 0; t < p.T; ++t) {
    const double vm = (p.materials[p.mat[t]].density * p.vol[t]) / 4.0;
    for (int l = 0; l < 4; ++l) p.mass[p.tets[4 * size_t(t) + l]] += vm;
  }
  // vertex -> tetrahedra
  p.vt_start.assign(p.V + 1, 0);
  for (size_t i = 0; i < 4 * size_t(p.T); ++i) p.vt_start[p.tets[i] + 1] += 1;
  for (uint32_t v = 0; v < p.V; ++v) {
    p.incident[v] = p.vt_start[v + 1];
    p.vt_start[v + 1] += p.vt_start[v];
  }
  p.vt.assign(p.vt_start[p.V], 0);
  {  // a tetrahedron listing a vertex twice is degenerate and already rejected
    std::vector<uint32_t> fill(p.vt_start.begin(), p.vt_start.end() - 1);
    for (uint32_t t = 0; t < p.T; ++t)
      for (int l = 0; l < 4; ++l) p.vt[fill[p.tets[4 * size_t(t) + l]]++] = t;
  }
  // vertex -> neighbors
  p.vv_start.assign(p.V + 1, 0);
  std::vector<std::vector<uint32_t>> rows(p.V);
  parallel_for(p.V, p.threads, [&](size_t v, unsigned) {
    std::vector<uint32_t>& r = rows[v];
    for (uint32_t k = p.vt_start[v]; k < p.vt_start[v + 1]; ++k)
      for (int l = 0; l < 4; ++l) {
        const uint32_t u = p.tets[4 * size_t(p.vt[k]) + l];
        if (u != v) r.push_back(u);
      }
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  });
  for (uint32_t v = 0; v < p.V; ++v) p.vv_start[v + 1] = p.vv_start[v] + uint32_t(rows[v].size());
  p.vv.resize(p.vv_start[p.V]);
  for (uint32_t v = 0; v < p.V; ++v) std::copy(rows[v].begin(), rows[v].end(), p.vv.begin() + p.vv_start[v]);
  // free indices
  p.active_of.assign(p.V, -1);
  p.active.clear();
  for (uint32_t v = 0; v < p.V; ++v)
    if (!p.fixed[v]) {
      p.active_of[v] = int32_t(p.active.size());
      p.active.push_back(v);
    }
  if (p.active.empty()) raise(JGS2_ERROR_RANGE, "A simulation scene must have at least one movable vertex.");
  p.n = 3 * p.active.size();
  // block CSR of the rest Hessian: row a holds a's free neighbors and itself
  const size_t A = p.active.size();
  p.hb_start.assign(A + 1, 0);
  for (size_t a = 0; a < A; ++a) {
    const uint32_t v = p.active[a];
    uint32_t count = 1;
    for (uint32_t k = p.vv_start[v]; k < p.vv_start[v + 1]; ++k) count += p.fixed[p.vv[k]] ? 0 : 1;
    p.hb_start[a + 1] = p.hb_start[a] + count;
  }
  p.hb_col.assign(p.hb_start[A], 0);
  for (size_t a = 0; a < A; ++a) {
    const uint32_t v = p.active[a];
    uint32_t o = p.hb_start[a];
    bool self = false;
    for (uint32_t k = p.vv_start[v]; k <= p.vv_start[v + 1]; ++k) {
      const uint32_t u = k < p.vv_start[v + 1] ? p.vv[k] : std::numeric_limits<uint32_t>::max();
      if (!self && u > v) {
        p.hb_col[o++] = uint32_t(a);
        self = true;
      }
      if (k < p.vv_start[v + 1] && !p.fixed[u]) p.hb_col[o++] = uint32_t(p.active_of[u]);
    }
  }
  p.hb.assign(9 * size_t(p.hb_start[A]), 0.0);
  auto block = [&](uint32_t a, uint32_t b) -> double* {
    const uint32_t* lo = &p.hb_col[p.hb_start[a]];
    const uint32_t* hi = &p.hb_col[p.hb_start[a + 1]];
    return &p.hb[9 * size_t(std::lower_bound(lo, hi, b) - p.hb_col.data())];
  };
  // tetrahedra in index order, so every entry sums its contributions in the
  // reference's order (fem.ts:286-315)
  for (uint32_t t = 0; t < p.T; ++t) {
    double Kt[144];
    tet_stiffness(p, t, Kt);
    for (int lr = 0; lr < 4; ++lr) {
      const int32_t ar = p.active_of[p.tets[4 * size_t(t) + lr]];
      if (ar < 0) continue;
      for (int lc = 0; lc < 4; ++lc) {
        const int32_t ac = p.active_of[p.tets[4 * size_t(t) + lc]];
        if (ac < 0) continue;
        double* b = block(uint32_t(ar), uint32_t(ac));
        for (int rc = 0; rc < 3; ++rc)
          for (int cc = 0; cc < 3; ++cc) b[rc * 3 + cc] += Kt[(lr * 3 + rc) * 12 + lc * 3 + cc];
      }
    }
  }
  const double ih2 = 1.0 / (p.h * p.h);
  for (size_t a = 0; a < A; ++a) {
    const double inertia = p.mass[p.active[a]] * ih2;
    if (!(inertia > 0) || !std::isfinite(inertia)) raise(JGS2_ERROR_RANGE, "Movable vertex has invalid lumped mass.");
    double* b = block(uint32_t(a), uint32_t(a));
    for (int c = 0; c < 3; ++c) b[c * 4] += inertia;
  }
  p.dense = p.mode_solver == 1 || (p.mode_solver == 0 && p.n <= kDenseModeLimit);
  if (!p.dense && (p.mode_solver == 3 || (p.mode_solver == 0 && GpuCg::available()))) {
    Vec pre(9 * A);
    for (size_t a = 0; a < A; ++a) {
      const uint32_t* lo = &p.hb_col[p.hb_start[a]];
      const size_t k = size_t(std::lower_bound(lo, lo + (p.hb_start[a + 1] - p.hb_start[a]), uint32_t(a)) - p.hb_col.data());
      inv3(&p.hb[9 * k], &pre[9 * a]);
    }
    p.gpu = std::make_shared<GpuCg>();
    const jgs2_status st = p.gpu->upload(uint32_t(A), p.hb_start.data(), p.hb_col.data(), p.hb.data(), pre.data());
    if (st != JGS2_OK) raise(st, "Cubature precompute: the device mode solver could not be set up.");
  }
  if (p.dense) {
    p.dense_l.assign(p.n * p.n, 0.0);
    for (size_t a = 0; a < A; ++a)
      for (uint32_t k = p.hb_start[a]; k < p.hb_start[a + 1]; ++k) {
        const size_t b = p.hb_col[k];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) p.dense_l[(3 * a + r) * p.n + 3 * b + c] = p.hb[9 * size_t(k) + r * 3 + c];
      }
    if (!cholesky(p.dense_l, p.n)) raise(JGS2_ERROR_RANGE, "Rest Hessian is not positive definite.");
  }
}

// y = H x over free coordinates.
void spmv(const Precompute& p, const double* x, double* y) {
  parallel_for((p.active.size() + 255) / 256, p.threads, [&](size_t chunk, unsigned) {
    const size_t end = std::min(p.active.size(), (chunk + 1) * 256);
    for (size_t a = chunk * 256; a < end; ++a) {
      double s[3] = {0, 0, 0};
      for (uint32_t k = p.hb_start[a]; k < p.hb_start[a + 1]; ++k) {
        const double* b = &p.hb[9 * size_t(k)];
        const double* xv = x + 3 * size_t(p.hb_col[k]);
        for (int r = 0; r < 3; ++r) s[r] += b[r * 3] * xv[0] + b[r * 3 + 1] * xv[1] + b[r * 3 + 2] * xv[2];
      }
      std::memcpy(y + 3 * a, s, sizeof s);
    }
  });
}

// Solve H x = b in place: the dense factor, or block-Jacobi preconditioned
// conjugate gradients to a relative residual of cg_tolerance.
void global_solve(const Precompute& p, double* b) {
  if (p.dense) {
    cholesky_solve(p.dense_l, p.n, b);
    return;
  }
  const size_t n = p.n, A = p.active.size();
  Vec pre(9 * A);
  for (size_t a = 0; a < A; ++a) {
    const uint32_t* lo = &p.hb_col[p.hb_start[a]];
    const size_t k = size_t(std::lower_bound(lo, lo + (p.hb_start[a + 1] - p.hb_start[a]), uint32_t(a)) - p.hb_col.data());
    inv3(&p.hb[9 * k], &pre[9 * a]);
  }
  auto precond = [&](const double* r, double* z) {
    for (size_t a = 0; a < A; ++a) {
      const double* m = &pre[9 * a];
      const double* ra = r + 3 * a;
      for (int i = 0; i < 3; ++i) z[3 * a + i] = m[i * 3] * ra[0] + m[i * 3 + 1] * ra[1] + m[i * 3 + 2] * ra[2];
    }
  };
  Vec x(n, 0.0), r(b, b + n), z(n), d(n), q(n);
  const double bnorm = std::sqrt(dotn(b, b, n));
  if (bnorm == 0.0) return;
  precond(r.data(), z.data());
  d = z;
  double rz = dotn(r.data(), z.data(), n);
  for (size_t it = 0; it < 20 * n + 100; ++it) {
    spmv(p, d.data(), q.data());
    const double alpha = rz / dotn(d.data(), q.data(), n);
    for (size_t i = 0; i < n; ++i) {
      x[i] += alpha * d[i];
      r[i] -= alpha * q[i];
    }
    if (std::sqrt(dotn(r.data(), r.data(), n)) <= p.cg_tolerance * bnorm) break;
    precond(r.data(), z.data());
    const double rz2 = dotn(r.data(), z.data(), n);
    const double beta = rz2 / rz;
    rz = rz2;
    for (size_t i = 0; i < n; ++i) d[i] = z[i] + beta * d[i];
  }
  std::memcpy(b, x.data(), n * sizeof(double));
}

// ------------------------------------------------------------ modes
void sign_by_largest(Vec& v) {  // cubature.ts:55-67
  size_t k = 0;
  for (size_t i = 1; i < v.size(); ++i)
    if (std::fabs(v[i]) > std::fabs(v[k])) k = i;
  if (v[k] < 0)
    for (double& e : v) e = -e;
}

void orthonormalize(std::vector<Vec>& vs) {  // cubature.ts:69-128
  const size_t n = vs.empty() ? 0 : vs[0].size();
  for (size_t i = 0; i < vs.size(); ++i) {
    Vec& c = vs[i];
    for (int pass = 0; pass < 2; ++pass)
      for (size_t j = 0; j < i; ++j) {
        const double pr = dotn(c.data(), vs[j].data(), n);
        for (size_t r = 0; r < n; ++r) c[r] -= pr * vs[j][r];
      }
    double norm = std::sqrt(dotn(c.data(), c.data(), n));
    if (norm <= 1e-11) {
      bool found = false;
      for (size_t off = 0; off < n && !found; ++off) {
        Vec t(n, 0.0);
        t[(i + off) % n] = 1.0;
        for (size_t j = 0; j < i; ++j) {
          const double pr = dotn(t.data(), vs[j].data(), n);
          for (size_t r = 0; r < n; ++r) t[r] -= pr * vs[j][r];
        }
        const double tn = std::sqrt(dotn(t.data(), t.data(), n));
        if (tn > 1e-11) {
          for (double& e : t) e /= tn;
          c = t;
          found = true;
        }
      }
      if (!found) raise(JGS2_ERROR_RANGE, "Could not construct the requested low-frequency basis.");
      norm = 1.0;
    }
    for (double& e : c) e /= norm;
    sign_by_largest(c);
  }
}

// Symmetric eigendecomposition of a small matrix by Jacobi rotations about
// the largest off-diagonal entry (cubature.ts:130-222); eigenvalues
// ascending (ties by index), vectors in columns.
void small_eigen(Vec a, size_t d, Vec& vals, Vec& vecs) {
  Vec v(d * d, 0.0);
  for (size_t i = 0; i < d; ++i) v[i * d + i] = 1.0;
  for (size_t it = 0; it < d * d * 16; ++it) {
    size_t p = 0, q = 1;
    double big = 0.0;
    for (size_t r = 0; r < d; ++r)
      for (size_t c = r + 1; c < d; ++c)
        if (std::fabs(a[r * d + c]) > big) {
          big = std::fabs(a[r * d + c]);
          p = r;
          q = c;
        }
    if (big <= 1e-12 * std::max({std::fabs(a[p * d + p]), std::fabs(a[q * d + q]), 1.0})) break;
    const double app = a[p * d + p], aqq = a[q * d + q], apq = a[p * d + q];
    const double th = 0.5 * std::atan2(2 * apq, aqq - app);
    const double c = std::cos(th), s = std::sin(th);
    for (size_t i = 0; i < d; ++i) {
      if (i == p || i == q) continue;
      const double ip = a[i * d + p], iq = a[i * d + q];
      a[i * d + p] = a[p * d + i] = c * ip - s * iq;
      a[i * d + q] = a[q * d + i] = s * ip + c * iq;
    }
    a[p * d + p] = c * c * app - 2 * s * c * apq + s * s * aqq;
    a[q * d + q] = s * s * app + 2 * s * c * apq + c * c * aqq;
    a[p * d + q] = a[q * d + p] = 0.0;
    for (size_t r = 0; r < d; ++r) {
      const double rp = v[r * d + p], rq = v[r * d + q];
      v[r * d + p] = c * rp - s * rq;
      v[r * d + q] = s * rp + c * rq;
    }
  }
  std::vector<size_t> order(d);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t l, size_t r) { return a[l * d + l] < a[r * d + r]; });
  vals.resize(d);
  vecs.assign(d * d, 0.0);
  for (size_t c = 0; c < d; ++c) {
    vals[c] = a[order[c] * d + order[c]];
    for (size_t r = 0; r < d; ++r) vecs[r * d + c] = v[r * d + order[c]];
  }
}

// Lowest rest-Hessian modes (cubature.ts:224-297), full 3 V coordinates.
std::vector<Vec> low_modes(const Precompute& p, size_t requested, size_t iterations) {
  const size_t n = p.n;
  const size_t count = std::min(requested, n);
  std::vector<Vec> m(count, Vec(n));
  for (size_t k = 0; k < count; ++k)
    for (size_t r = 0; r < n; ++r)
      m[k][r] = std::sin(double(r + 1) * double(k + 1) * 0.754877666) +
                0.5 * std::cos(double(r + 1) * double(k + 3) * 0.569840291);
  orthonormalize(m);
  for (size_t it = 0; it < iterations; ++it) {
    if (p.dense) {
      parallel_for(count, p.threads, [&](size_t k, unsigned) { global_solve(p, m[k].data()); });
    } else if (p.gpu) {
      for (size_t first = 0; first < count; first += 32) {  // up to 32 right-hand sides per device solve
        const uint32_t cols = uint32_t(std::min<size_t>(32, count - first));
        Vec block(n * cols);
        for (size_t r = 0; r < n; ++r)
          for (uint32_t j = 0; j < cols; ++j) block[r * cols + j] = m[first + j][r];
        uint32_t iters = 0;
        double ms = 0.0;
        const jgs2_status st = p.gpu->solve(block.data(), cols, p.cg_tolerance, 20000, &iters, &ms);
        if (st != JGS2_OK) raise(st, "Cubature precompute: the device mode solve failed.");
        p.gpu_ms += ms;
        p.gpu_iterations += iters;
        for (size_t r = 0; r < n; ++r)
          for (uint32_t j = 0; j < cols; ++j) m[first + j][r] = block[r * cols + j];
      }
    } else {
      for (size_t k = 0; k < count; ++k) global_solve(p, m[k].data());
    }
    orthonormalize(m);
  }
  std::vector<Vec> hm(count, Vec(n));
  for (size_t k = 0; k < count; ++k) spmv(p, m[k].data(), hm[k].data());
  Vec proj(count * count);
  for (size_t r = 0; r < count; ++r)
    for (size_t c = 0; c < count; ++c) proj[r * count + c] = dotn(m[r].data(), hm[c].data(), n);
  Vec vals, vecs;
  small_eigen(proj, count, vals, vecs);
  std::vector<Vec> out;
  for (size_t k = 0; k < count; ++k) {
    Vec v(n, 0.0);
    for (size_t b = 0; b < count; ++b) {
      const double coef = vecs[b * count + k];
      for (size_t r = 0; r < n; ++r) v[r] += m[b][r] * coef;
    }
    sign_by_largest(v);
    Vec full(3 * size_t(p.V), 0.0);
    for (size_t a = 0; a < p.active.size(); ++a)
      for (int c = 0; c < 3; ++c) full[3 * size_t(p.active[a]) + c] = v[3 * a + c];
    out.push_back(std::move(full));
  }
  return out;
}

bool append_normalized(std::vector<Vec>& dst, Vec c) {  // nonlinear-cubature.ts:310-336
  for (const Vec& e : dst) {
    const double pr = dotn(c.data(), e.data(), c.size());
    for (size_t i = 0; i < c.size(); ++i) c[i] -= pr * e[i];
  }
  const double norm = std::sqrt(dotn(c.data(), c.data(), c.size()));
  if (!(norm > 1e-10) || !std::isfinite(norm)) return false;
  for (double& e : c) e /= norm;
  for (double e : c) {
    if (std::fabs(e) <= 1e-12) continue;
    if (e < 0)
      for (double& w : c) w = -w;
    break;
  }
  dst.push_back(std::move(c));
  return true;
}

// Translations and rotations of every connected component without a pinned
// vertex (nonlinear-cubature.ts:338-422), components ordered by their
// lowest vertex.
std::vector<Vec> rigid_modes(const Precompute& p) {
  std::vector<uint32_t> comp(p.V, std::numeric_limits<uint32_t>::max());
  std::vector<std::vector<uint32_t>> comps;
  for (uint32_t s = 0; s < p.V; ++s) {
    if (comp[s] != std::numeric_limits<uint32_t>::max()) continue;
    std::vector<uint32_t> members{s}, stack{s};
    comp[s] = uint32_t(comps.size());
    while (!stack.empty()) {
      const uint32_t v = stack.back();
      stack.pop_back();
      for (uint32_t k = p.vv_start[v]; k < p.vv_start[v + 1]; ++k)
        if (comp[p.vv[k]] == std::numeric_limits<uint32_t>::max()) {
          comp[p.vv[k]] = comp[s];
          members.push_back(p.vv[k]);
          stack.push_back(p.vv[k]);
        }
    }
    std::sort(members.begin(), members.end());
    comps.push_back(std::move(members));
  }
  std::vector<Vec> out;
  for (const auto& vs : comps) {
    bool pinned = false;
    for (uint32_t v : vs) pinned |= p.fixed[v] != 0;
    if (pinned) continue;
    double total = 0.0, ctr[3] = {0, 0, 0};
    for (uint32_t v : vs) {
      total += p.mass[v];
      for (int c = 0; c < 3; ++c) ctr[c] += p.mass[v] * p.x[3 * size_t(v) + c];
    }
    for (double& c : ctr) c /= total;
    std::vector<Vec> cand(6, Vec(3 * size_t(p.V), 0.0));
    for (uint32_t v : vs) {
      const double* q = p.x + 3 * size_t(v);
      const double x = q[0] - ctr[0], y = q[1] - ctr[1], z = q[2] - ctr[2];
      for (int c = 0; c < 3; ++c) cand[size_t(c)][3 * size_t(v) + size_t(c)] = 1.0;
      const double rot[3][3] = {{0, -z, y}, {z, 0, -x}, {-y, x, 0}};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) cand[3 + size_t(r)][3 * size_t(v) + size_t(c)] = rot[r][c];
    }
    for (auto& c : cand) append_normalized(out, c);
  }
  return out;
}

// ------------------------------------------------------------ poses
double min_det_at(const Precompute& p, const double* x) {
  double m = INFINITY;
  for (uint32_t t = 0; t < p.T; ++t) {
    const uint32_t* q = p.tets + 4 * size_t(t);
    double d[9], F[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) d[r * 3 + c] = x[3 * size_t(q[c + 1]) + r] - x[3 * size_t(q[0]) + r];
    mul3(d, &p.dminv[9 * size_t(t)], F);
    m = std::min(m, det3(F));
  }
  return m;
}

struct Pose {
  Vec x, xp;
  double min_j = 0.0;
};

Pose scaled_pose(const Precompute& p, const Vec& raw, double amplitude, double blend, double scale) {
  // max(max |grad u|_F over tetrahedra, max |u_v| / scene scale)
  // (nonlinear-cubature.ts:146-182, 218-256)
  double g = 0.0, u = 0.0;
  for (uint32_t t = 0; t < p.T; ++t) {
    const uint32_t* q = p.tets + 4 * size_t(t);
    double d[9], F[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) d[r * 3 + c] = raw[3 * size_t(q[c + 1]) + r] - raw[3 * size_t(q[0]) + r];
    mul3(d, &p.dminv[9 * size_t(t)], F);
    double f = 0.0;
    for (double e : F) f += e * e;
    g = std::max(g, std::sqrt(f));
  }
  for (uint32_t v = 0; v < p.V; ++v) u = std::max(u, hypot_js(&raw[3 * size_t(v)], 3));
  const double measure = std::max(g, u / scale);
  if (!(measure > 1e-14) || !std::isfinite(measure)) raise(JGS2_ERROR_RANGE, "Nonlinear Cubature mode has zero direction.");
  const double k = amplitude / measure;
  Pose out;
  out.x.assign(p.x, p.x + 3 * size_t(p.V));
  out.xp = out.x;
  for (size_t e = 0; e < out.x.size(); ++e) {
    const double d = raw[e] * k;
    out.x[e] += d;
    out.xp[e] += blend * d;
  }
  out.min_j = min_det_at(p, out.x.data());
  if (!(out.min_j >= 0.5)) raise(JGS2_ERROR_RANGE, "Nonlinear Cubature pose has minimum J below 0.5.");
  return out;
}

// The training poses of the nonlinear corpus (nonlinear-cubature.ts:462-565):
// the lowest 8 deformational modes, each at both signs, scaled to amplitude
// 0.12 of the direction measure, predicted positions at rest.
std::vector<Pose> training_poses(const Precompute& p) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < p.V; ++v)
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], p.x[3 * size_t(v) + c]);
      hi[c] = std::max(hi[c], p.x[3 * size_t(v) + c]);
    }
  const double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const double scale = hypot_js(ext, 3);
  if (!(scale > 0)) raise(JGS2_ERROR_RANGE, "Nonlinear Cubature scene scale must be finite and positive.");
  const size_t wanted = 8 + 4;
  const std::vector<Vec> rigid = rigid_modes(p);
  std::vector<Vec> raw = low_modes(p, wanted + rigid.size(), 24);
  std::vector<Vec> modes;
  if (rigid.empty()) {
    modes = std::move(raw);
  } else {
    for (Vec& c : raw) {
      for (const Vec& r : rigid) {
        const double pr = dotn(c.data(), r.data(), c.size());
        for (size_t i = 0; i < c.size(); ++i) c[i] -= pr * r[i];
      }
      append_normalized(modes, c);
      if (modes.size() >= wanted) break;
    }
    if (modes.empty()) raise(JGS2_ERROR_RANGE, "Nonlinear Cubature free components have no deformational modes.");
  }
  std::vector<Pose> out;
  for (size_t m = 0; m < std::min<size_t>(8, modes.size()); ++m)
    for (double sign : {1.0, -1.0}) {
      Vec dir(modes[m].size());
      for (size_t i = 0; i < dir.size(); ++i) dir[i] = sign * modes[m][i];
      out.push_back(scaled_pose(p, dir, 0.12, 0.0, scale));
    }
  return out;
}

// Polar frame of the rest-volume-weighted vertex deformation gradient
// (deformation-gradient.ts:22-172): up to 20 averaging steps R <- (R + R^-T)/2
// and a right-handed column orthonormalization (oracle.ts:102-217).
void unit_or(const double* v, const double* fb, double* o) {
  const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (!(n > 1e-14) || !std::isfinite(n)) {
    std::memcpy(o, fb, 3 * sizeof(double));
    return;
  }
  o[0] = v[0] / n;
  o[1] = v[1] / n;
  o[2] = v[2] / n;
}
void cross(const double* a, const double* b, double* o) {
  const double t[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  std::memcpy(o, t, sizeof t);
}
void perp(const double* axis, double* o) {
  const bool x = std::fabs(axis[0]) > 0.8;
  const double cand[3] = {x ? 0.0 : 1.0, x ? 1.0 : 0.0, 0.0};
  const double z[3] = {0, 0, 1};
  double c[3];
  cross(axis, cand, c);
  unit_or(c, z, o);
}

void polar_frame(const double* F, double* R) {
  double r[9];
  std::memcpy(r, F, sizeof r);
  for (int it = 0; it < 20; ++it) {
    const double d = det3(r);
    if (!(std::fabs(d) > 1e-14) || !std::isfinite(d)) break;
    double inv[9], nx[9];
    inv3(r, inv);
    double change = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        nx[i * 3 + j] = 0.5 * (r[i * 3 + j] + inv[j * 3 + i]);
        change = std::max(change, std::fabs(nx[i * 3 + j] - r[i * 3 + j]));
      }
    std::memcpy(r, nx, sizeof r);
    if (change <= 1e-14) break;
  }
  const double k0[3] = {r[0], r[3], r[6]}, k1[3] = {r[1], r[4], r[7]}, k2[3] = {r[2], r[5], r[8]};
  const double ex[3] = {1, 0, 0}, ez[3] = {0, 0, 1};
  double c0[3], c1[3], c2[3], t[3], pp[3];
  unit_or(k0, ex, c0);
  const double pr = c0[0] * k1[0] + c0[1] * k1[1] + c0[2] * k1[2];
  const double cand[3] = {k1[0] - pr * c0[0], k1[1] - pr * c0[1], k1[2] - pr * c0[2]};
  perp(c0, pp);
  unit_or(cand, pp, c1);
  cross(c0, c1, t);
  unit_or(t, ez, c2);
  if (c2[0] * k2[0] + c2[1] * k2[1] + c2[2] * k2[2] < 0)
    for (int i = 0; i < 3; ++i) {
      c1[i] = -c1[i];
      c2[i] = -c2[i];
    }
  cross(c2, c0, t);
  perp(c0, pp);
  unit_or(t, pp, c1);
  const double out[9] = {c0[0], c1[0], c2[0], c0[1], c1[1], c2[1], c0[2], c1[2], c2[2]};
  std::memcpy(R, out, sizeof out);
}

// Frames of every vertex at each pose.
std::vector<Vec> pose_frames(const Precompute& p, const std::vector<Pose>& poses) {
  // stencil: sorted source vertices and volume-weighted shape gradients
  std::vector<uint32_t> st_start(p.V + 1, 0), st_src;
  Vec st_coef;
  {
    std::vector<std::vector<std::pair<uint32_t, std::array<double, 3>>>> rows(p.V);
    parallel_for(p.V, p.threads, [&](size_t c, unsigned) {
      std::map<uint32_t, std::array<double, 3>> acc;
      double total = 0.0;
      for (uint32_t k = p.vt_start[c]; k < p.vt_start[c + 1]; ++k) {
        const uint32_t t = p.vt[k];
        const double vol = p.vol[t];
        total += vol;
        for (int l = 0; l < 4; ++l) {
          auto& e = acc.emplace(p.tets[4 * size_t(t) + l], std::array<double, 3>{0, 0, 0}).first->second;
          for (int i = 0; i < 3; ++i) e[size_t(i)] += vol * p.shape[12 * size_t(t) + size_t(l * 3 + i)];
        }
      }
      if (acc.empty()) raise(JGS2_ERROR_RANGE, "Vertex has no incident tetrahedron for its deformation frame.");
      for (auto& kv : acc) {
        for (double& e : kv.second) e /= total;
        rows[c].push_back(kv);
      }
    });
    for (uint32_t c = 0; c < p.V; ++c) {
      for (auto& kv : rows[c]) {
        st_src.push_back(kv.first);
        st_coef.insert(st_coef.end(), kv.second.begin(), kv.second.end());
      }
      st_start[c + 1] = uint32_t(st_src.size());
    }
  }
  std::vector<Vec> out(poses.size(), Vec(9 * size_t(p.V)));
  for (size_t q = 0; q < poses.size(); ++q) {
    const double* x = poses[q].x.data();
    parallel_for(p.V, p.threads, [&](size_t c, unsigned) {
      double F[9] = {0};
      for (uint32_t k = st_start[c]; k < st_start[c + 1]; ++k)
        for (int w = 0; w < 3; ++w) {
          const double pos = x[3 * size_t(st_src[k]) + size_t(w)];
          for (int m = 0; m < 3; ++m) F[w * 3 + m] += pos * st_coef[3 * size_t(k) + size_t(m)];
        }
      polar_frame(F, &out[q][9 * c]);
    });
  }
  return out;
}

// ------------------------------------------------------------ selection
// Training columns of one source vertex: `rows` values per candidate.
struct Columns {
  size_t rows = 0;
  std::vector<uint32_t> tet;   // candidate -> tetrahedron, ascending
  Vec col;                     // rows per candidate
  Vec target;
  Vec gram;                    // memoized dots; NaN until computed
  Vec tdot;
  const double* c(size_t i) const { return &col[i * rows]; }
  void prepare() {
    const size_t n = tet.size();
    gram.assign(n * n, std::numeric_limits<double>::quiet_NaN());
    tdot.resize(n);
    for (size_t i = 0; i < n; ++i) tdot[i] = dotn(c(i), target.data(), rows);
  }
  double g(size_t i, size_t j) {
    double& v = gram[i * tet.size() + j];
    if (std::isnan(v)) {
      v = dotn(c(i), c(j), rows);
      gram[j * tet.size() + i] = v;
    }
    return v;
  }
};

struct MaskFit {
  bool ok = false;
  double err = 0.0;
  double w[12] = {0};
};

// One NNLS mask (cubature.ts:386-433): unconstrained least squares over the
// mask's columns; rejected when singular or any weight is below -1e-10.
MaskFit fit_mask(Columns& C, const size_t* sel, int n, unsigned mask, double* scratch) {
  MaskFit f;
  int act[12], d = 0;
  for (int i = 0; i < n; ++i)
    if (mask & (1u << i)) act[d++] = i;
  double A[144], b[12];
  for (int r = 0; r < d; ++r) {
    b[r] = C.tdot[sel[act[r]]];
    for (int c = 0; c < d; ++c) A[r * d + c] = C.g(sel[act[r]], sel[act[c]]);
  }
  if (!small_solve(A, b, d)) return f;
  for (int i = 0; i < d; ++i)
    if (b[i] < -1e-10) return f;
  for (int i = 0; i < d; ++i) f.w[act[i]] = std::max(0.0, b[i]);
  std::memcpy(scratch, C.target.data(), C.rows * sizeof(double));
  for (int i = 0; i < n; ++i) {
    if (f.w[i] == 0.0) continue;
    const double* col = C.c(sel[i]);
    for (size_t r = 0; r < C.rows; ++r) scratch[r] -= col[r] * f.w[i];
  }
  f.err = dotn(scratch, scratch, C.rows);
  f.ok = true;
  return f;
}

// Exhaustive NNLS over masks 1..2^n-1 in order, keeping a strictly better
// fit (by 1e-14 relative); `shared` holds precomputed fits of the masks that
// do not contain position `fresh` (fresh < 0: none precomputed).
void nnls(Columns& C, const size_t* sel, int n, int fresh, const std::vector<MaskFit>* shared, double* w, double& err,
          double* scratch) {
  double best = dotn(C.target.data(), C.target.data(), C.rows);
  double bw[12] = {0};
  for (unsigned mask = 1; mask < (1u << n); ++mask) {
    const bool reuse = shared && fresh >= 0 && !(mask & (1u << fresh));
    const MaskFit f = reuse ? (*shared)[mask] : fit_mask(C, sel, n, mask, scratch);
    if (!f.ok) continue;
    if (f.err < best - 1e-14 * std::max(1.0, best)) {
      best = f.err;
      std::memcpy(bw, f.w, sizeof bw);
    }
  }
  std::memcpy(w, bw, sizeof(double) * size_t(n));
  err = best;
}

struct Chosen {
  std::vector<size_t> sel;  // candidate indices, slot order
  Vec w;
  double normalized = 0.0;
  double residual_sq = 0.0;
};

double residual_sq_of(Columns& C, const std::vector<size_t>& sel, const double* w, double* scratch) {
  std::memcpy(scratch, C.target.data(), C.rows * sizeof(double));
  for (size_t i = 0; i < sel.size(); ++i) {
    const double* col = C.c(sel[i]);
    for (size_t r = 0; r < C.rows; ++r) scratch[r] -= col[r] * w[i];
  }
  return dotn(scratch, scratch, C.rows);
}

// Greedy growth and one-for-one replacement passes (cubature.ts:508-691).
Chosen greedy_select(Columns& C, size_t max_columns, double tolerance, size_t passes) {
  Chosen out;
  Vec scratch(C.rows);
  std::vector<size_t> cand;
  for (size_t i = 0; i < C.tet.size(); ++i)
    if (dotn(C.c(i), C.c(i), C.rows) > 1e-18) cand.push_back(i);
  const size_t limit = std::min(max_columns, cand.size());
  const double tn = std::sqrt(dotn(C.target.data(), C.target.data(), C.rows));
  if (tn <= 1e-12) return out;
  if (cand.empty() || limit == 0) {
    out.normalized = 1.0;
    return out;
  }
  auto in = [&](size_t c) { return std::find(out.sel.begin(), out.sel.end(), c) != out.sel.end(); };
  double cur = tn * tn;
  {
    const Vec r0 = C.target;
    cur = dotn(r0.data(), r0.data(), C.rows);
  }
  std::vector<MaskFit> shared;
  for (size_t step = 0; step < limit; ++step) {
    const int n = int(out.sel.size()) + 1;
    std::vector<size_t> trial = out.sel;
    trial.push_back(0);
    shared.assign(size_t(1) << n, MaskFit{});
    for (unsigned mask = 1; mask < (1u << (n - 1)); ++mask) shared[mask] = fit_mask(C, trial.data(), n, mask, scratch.data());
    long best = -1;
    double be = cur, bw[12] = {0};
    for (size_t c : cand) {
      if (in(c)) continue;
      trial.back() = c;
      double w[12], e;
      nnls(C, trial.data(), n, n - 1, &shared, w, e, scratch.data());
      const double ex = residual_sq_of(C, trial, w, scratch.data());
      const double tol = 1e-14 * std::max(1.0, be);
      if (ex < be - tol || (std::fabs(ex - be) <= tol && best >= 0 && C.tet[c] < C.tet[size_t(best)])) {
        best = long(c);
        be = ex;
        std::memcpy(bw, w, sizeof bw);
      }
    }
    if (best < 0) break;
    out.sel.push_back(size_t(best));
    out.w.assign(bw, bw + n);
    cur = be;
    if (std::sqrt(cur) / tn < tolerance) break;
  }
  for (size_t pass = 0; pass < passes && !out.sel.empty(); ++pass) {
    const int n = int(out.sel.size());
    long slot = -1, best = -1;
    double be = cur, bw[12] = {0};
    for (int sl = 0; sl < n; ++sl) {
      std::vector<size_t> trial = out.sel;
      shared.assign(size_t(1) << n, MaskFit{});
      for (unsigned mask = 1; mask < (1u << n); ++mask)
        if (!(mask & (1u << sl))) shared[mask] = fit_mask(C, trial.data(), n, mask, scratch.data());
      for (size_t c : cand) {
        if (in(c)) continue;
        trial[size_t(sl)] = c;
        double w[12], e;
        nnls(C, trial.data(), n, sl, &shared, w, e, scratch.data());
        const double ex = residual_sq_of(C, trial, w, scratch.data());
        const double tol = 1e-14 * std::max(1.0, be);
        if (ex < be - tol || (std::fabs(ex - be) <= tol && best >= 0 && C.tet[c] < C.tet[size_t(best)])) {
          slot = sl;
          best = long(c);
          be = ex;
          std::memcpy(bw, w, sizeof bw);
        }
      }
    }
    if (slot < 0) break;
    out.sel[size_t(slot)] = size_t(best);
    out.w.assign(bw, bw + n);
    cur = be;
  }
  out.residual_sq = cur;
  out.normalized = std::sqrt(cur) / tn;
  return out;
}

void column_rank(const Columns& C, uint32_t& nonzero, uint32_t& rank) {  // training.ts:594-629
  double mx = 0.0;
  for (size_t i = 0; i < C.tet.size(); ++i) mx = std::max(mx, std::sqrt(dotn(C.c(i), C.c(i), C.rows)));
  nonzero = rank = 0;
  if (mx == 0.0) return;
  const double th = 1e-9 * mx;
  std::vector<Vec> basis;
  for (size_t i = 0; i < C.tet.size(); ++i) {
    nonzero += std::sqrt(dotn(C.c(i), C.c(i), C.rows)) > th ? 1 : 0;
    Vec v(C.c(i), C.c(i) + C.rows);
    for (int pass = 0; pass < 2; ++pass)
      for (const Vec& d : basis) {
        const double pr = dotn(v.data(), d.data(), C.rows);
        for (size_t r = 0; r < C.rows; ++r) v[r] -= pr * d[r];
      }
    const double nv = std::sqrt(dotn(v.data(), v.data(), C.rows));
    if (nv <= th) continue;
    for (double& e : v) e /= nv;
    basis.push_back(std::move(v));
    if (basis.size() == C.rows) {
      for (size_t j = i + 1; j < C.tet.size(); ++j) nonzero += std::sqrt(dotn(C.c(j), C.c(j), C.rows)) > th ? 1 : 0;
      break;
    }
  }
  rank = uint32_t(basis.size());
}

// ------------------------------------------------------------ per vertex
struct Support {
  std::vector<uint32_t> verts;       // free vertices of the basis support, ascending
  std::vector<int32_t> local;        // vertex -> position in verts (-1 outside); sized V, reset after use
  Vec U;                             // 9 per support vertex (row = coordinate)
};

// Exact basis of `src` over its support (rings = 0: every free vertex).
void solve_basis(const Precompute& p, uint32_t src, Support& s) {
  s.verts.clear();
  if (p.rings == 0) {
    s.verts = p.active;
  } else {
    s.verts.push_back(src);
    s.local[src] = 0;
    size_t frontier = 0;
    for (uint32_t ring = 0; ring < p.rings; ++ring) {
      const size_t end = s.verts.size();
      for (size_t i = frontier; i < end; ++i) {
        const uint32_t v = s.verts[i];
        for (uint32_t k = p.vv_start[v]; k < p.vv_start[v + 1]; ++k) {
          const uint32_t u = p.vv[k];
          if (p.fixed[u] || s.local[u] >= 0) continue;
          s.local[u] = 0;
          s.verts.push_back(u);
        }
      }
      frontier = end;
    }
    std::sort(s.verts.begin(), s.verts.end());
  }
  for (size_t i = 0; i < s.verts.size(); ++i) s.local[s.verts[i]] = int32_t(i);
  const size_t m = 3 * s.verts.size();
  const size_t base = 3 * size_t(s.local[src]);
  Vec X(3 * m, 0.0);  // three right-hand sides, column-major
  if (p.rings == 0) {
    for (int c = 0; c < 3; ++c) {
      double* col = &X[size_t(c) * m];
      col[base + size_t(c)] = 1.0;
      global_solve(p, col);
    }
  } else {
    Vec L(m * m, 0.0);
    for (size_t i = 0; i < s.verts.size(); ++i) {
      const size_t a = size_t(p.active_of[s.verts[i]]);
      for (uint32_t k = p.hb_start[a]; k < p.hb_start[a + 1]; ++k) {
        const int32_t j = s.local[p.active[p.hb_col[k]]];
        if (j < 0) continue;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) L[(3 * i + size_t(r)) * m + 3 * size_t(j) + size_t(c)] = p.hb[9 * size_t(k) + size_t(r * 3 + c)];
      }
    }
    if (!cholesky(L, m)) raise(JGS2_ERROR_RANGE, "Rest Hessian is not positive definite.");
    for (int c = 0; c < 3; ++c) {
      double* col = &X[size_t(c) * m];
      col[base + size_t(c)] = 1.0;
      cholesky_solve(L, m, col);
    }
  }
  double comp[9], schur[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) comp[r * 3 + c] = 0.5 * (X[size_t(c) * m + base + size_t(r)] + X[size_t(r) * m + base + size_t(c)]);
  inv3(comp, schur);
  s.U.assign(9 * s.verts.size(), 0.0);
  for (size_t i = 0; i < s.verts.size(); ++i)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double v = 0.0;
        for (int k = 0; k < 3; ++k) v += X[size_t(k) * m + 3 * i + size_t(r)] * schur[k * 3 + c];
        s.U[9 * i + size_t(r * 3 + c)] = v;
      }
}

std::vector<uint32_t> support_tets(const Precompute& p, const Support& s) {
  std::vector<uint32_t> t;
  for (uint32_t v : s.verts)
    for (uint32_t k = p.vt_start[v]; k < p.vt_start[v + 1]; ++k) t.push_back(p.vt[k]);
  std::sort(t.begin(), t.end());
  t.erase(std::unique(t.begin(), t.end()), t.end());
  return t;
}

double snh_gradient(const Precompute& p, uint32_t t, const double* x, double* g) {
  const auto& m = p.materials[p.mat[t]];
  if (!(m.poisson_ratio >= 0 && m.poisson_ratio < 0.5))
    raise(JGS2_ERROR_RANGE, "Stable Neo-Hookean material requires Poisson ratio in [0, 0.5).");
  const double lmu = m.young_modulus / (2 * (1 + m.poisson_ratio));
  const double llam = (m.young_modulus * m.poisson_ratio) / ((1 + m.poisson_ratio) * (1 - 2 * m.poisson_ratio));
  const double mu = (4.0 / 3.0) * lmu;
  const double lambda = llam + (5.0 / 6.0) * lmu;
  const double alpha = 1 + mu / lambda - mu / (4 * lambda);
  const uint32_t* q = p.tets + 4 * size_t(t);
  double d[9], F[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) d[r * 3 + c] = x[3 * size_t(q[c + 1]) + r] - x[3 * size_t(q[0]) + r];
  mul3(d, &p.dminv[9 * size_t(t)], F);
  double I1 = 0.0;
  for (double e : F) I1 += e * e;
  const double J = det3(F);
  const double cof[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                         F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                         F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
  const double a = mu * (1 - 1 / (I1 + 1)), b = lambda * (J - alpha);
  double P[9];
  for (int i = 0; i < 9; ++i) P[i] = a * F[i] + b * cof[i];
  const double* sg = &p.shape[12 * size_t(t)];
  for (int l = 0; l < 4; ++l)
    for (int r = 0; r < 3; ++r)
      g[l * 3 + r] = p.vol[t] * (P[r * 3] * sg[l * 3] + P[r * 3 + 1] * sg[l * 3 + 1] + P[r * 3 + 2] * sg[l * 3 + 2]);
  return J;
}

struct VertexResult {
  std::vector<uint32_t> tets;
  std::vector<float> w;
  std::vector<float> basis;  // 36 per record
  double residual = 0.0;
  uint32_t nonzero = 0, rank = 0, packed_rank = 0;
};

// Columns of the stable Neo-Hookean training system (nonlinear-cubature-
// training.ts:311-575) for the candidates; `packed` uses the f32-rounded
// basis. Returns the per-pose target norms (0 for trivial poses).
void snh_columns(const Precompute& p, const std::vector<Pose>& poses, const std::vector<Vec>& frames, uint32_t src,
                 const Support& s, const Vec& U, const Vec* norms, Columns& C, Vec& pose_norm, int& valid) {
  const size_t P = poses.size(), n = C.tet.size();
  C.rows = 3 * P;
  C.col.assign(n * C.rows, 0.0);
  C.target.assign(C.rows, 0.0);
  pose_norm.assign(P, 0.0);
  valid = 0;
  const double ih2 = 1.0 / (p.h * p.h);
  Vec cur(U.size());
  Vec rem(3 * n);
  for (size_t q = 0; q < P; ++q) {
    const Vec& R = frames[q];
    double rs_t[9];
    std::memcpy(rs_t, &R[9 * size_t(src)], sizeof rs_t);
    for (size_t i = 0; i < s.verts.size(); ++i) {  // B_u = R_u U_u R_src^T
      double left[9];
      mul3(&R[9 * size_t(s.verts[i])], &U[9 * i], left);
      mul3t(left, rs_t, &cur[9 * i]);
    }
    const double* x = poses[q].x.data();
    const double* xp = poses[q].xp.data();
    double pt[3] = {0, 0, 0}, scale_sq = 0.0;
    for (size_t k = 0; k < n; ++k) {
      const uint32_t t = C.tet[k];
      double g[12];
      snh_gradient(p, t, x, g);
      double r3[3] = {0, 0, 0};
      int slot = -1;
      for (int l = 0; l < 4; ++l) {
        const uint32_t v = p.tets[4 * size_t(t) + l];
        const double di = (p.mass[v] * ih2) / double(p.incident[v]);
        for (int c = 0; c < 3; ++c) g[l * 3 + c] += di * (x[3 * size_t(v) + c] - xp[3 * size_t(v) + c]);
        if (v == src) slot = l;
      }
      // B^T g summed over the element rows in order (training.ts:404-411)
      for (int r = 0; r < 3; ++r)
        for (int l = 0; l < 4; ++l) {
          const int32_t li = s.local[p.tets[4 * size_t(t) + l]];
          if (li < 0) continue;
          const double* B = &cur[9 * size_t(li)];
          for (int c = 0; c < 3; ++c) r3[r] += B[c * 3 + r] * g[l * 3 + c];
        }
      if (slot >= 0)
        for (int r = 0; r < 3; ++r) r3[r] -= g[slot * 3 + r];
      for (int r = 0; r < 3; ++r) {
        rem[3 * k + size_t(r)] = r3[r];
        pt[r] += r3[r];
      }
      scale_sq += r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2];
    }
    const double pn = std::sqrt(pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]);
    pose_norm[q] = pn;
    if (scale_sq == 0.0) continue;
    if (pn <= 1e-12 * std::sqrt(scale_sq))
      raise(JGS2_ERROR_INFEASIBLE, "Nonlinear Cubature training pose has a cancellation-dominated complementary target.");
    ++valid;
    const double nn = norms ? (*norms)[q] : pn;
    for (int c = 0; c < 3; ++c) {
      const size_t row = 3 * q + size_t(c);
      C.target[row] = pt[c] / nn;
      for (size_t k = 0; k < n; ++k) C.col[k * C.rows + row] = rem[3 * k + size_t(c)] / nn;
    }
  }
}

// Everything of one source vertex's stable Neo-Hookean training except the
// selection: basis, candidate columns and target, the f32-packed columns,
// and their ranks.
struct Prepared {
  Columns C;
  Vec packed;                   // packed columns, C.rows per candidate
  Vec norms;
  int valid = 0;
  double tn = 0.0;
  uint32_t nonzero = 0, rank = 0, packed_rank = 0;
  std::vector<uint32_t> verts;  // basis support, ascending
  Vec U;                        // 9 per support vertex
};

void prepare_vertex_snh(const Precompute& p, const std::vector<Pose>& poses, const std::vector<Vec>& frames,
                        uint32_t src, Support& s, Prepared& P) {
  solve_basis(p, src, s);
  Columns& C = P.C;
  C.tet = support_tets(p, s);
  if (p.T <= 12 && p.K < p.T) {  // the subset audit needs every column
    C.tet.resize(p.T);
    std::iota(C.tet.begin(), C.tet.end(), 0u);
  }
  snh_columns(p, poses, frames, src, s, s.U, nullptr, C, P.norms, P.valid);
  P.tn = std::sqrt(dotn(C.target.data(), C.target.data(), C.rows));
  if (P.tn <= 1e-12)
    raise(JGS2_ERROR_INFEASIBLE, "Nonlinear Cubature source vertex has no valid nontrivial training target.");
  column_rank(C, P.nonzero, P.rank);
  Vec Uf(s.U.size());
  for (size_t i = 0; i < Uf.size(); ++i) Uf[i] = double(float(s.U[i]));
  Columns Pk;
  Pk.tet = C.tet;
  Vec unused;
  int pvalid = 0;
  snh_columns(p, poses, frames, src, s, Uf, &P.norms, Pk, unused, pvalid);
  if (pvalid != P.valid) raise(JGS2_ERROR_INFEASIBLE, "Packing changed nonlinear Cubature pose validity.");
  uint32_t pnz = 0;
  column_rank(Pk, pnz, P.packed_rank);
  P.packed = std::move(Pk.col);
  P.verts = s.verts;
  P.U = s.U;
}

// Greedy selection and, for the reference's tiny fixtures, the subset audit
// (training.ts:761-808) on the host.
Chosen select_host(const Precompute& p, Prepared& P) {
  Columns& C = P.C;
  C.prepare();
  Chosen ch = greedy_select(C, p.K, 0.01, 12);
  if (p.T <= 12 && p.K < p.T) {  // over candidate (= tetrahedron) order
    Vec scratch(C.rows);
    std::vector<size_t> sub;
    std::function<void(size_t)> visit = [&](size_t next) {
      if (sub.size() == p.K) {
        double w[12], e;
        nnls(C, sub.data(), int(sub.size()), -1, nullptr, w, e, scratch.data());
        const double rn = std::sqrt(residual_sq_of(C, sub, w, scratch.data()));
        const double nr = P.tn > 1e-12 ? rn / P.tn : 0.0;
        if (nr < ch.normalized - 1e-14) {
          ch.normalized = nr;
          ch.sel = sub;
          ch.w.assign(w, w + sub.size());
        }
        return;
      }
      for (size_t t = next; t + (p.K - sub.size()) <= C.tet.size(); ++t) {
        sub.push_back(t);
        visit(t + 1);
        sub.pop_back();
      }
    };
    visit(0);
  }
  return ch;
}

// Positive weights, f32 packing, the packed residual against the f64 target
// and the reference's bound (training.ts:810-843, precompute.ts:313-323).
VertexResult finish_vertex_snh(const Precompute& p, const Prepared& P, const Chosen& ch, uint32_t src) {
  VertexResult out;
  out.nonzero = P.nonzero;
  out.rank = P.rank;
  out.packed_rank = P.packed_rank;
  const Columns& C = P.C;
  Vec res = C.target;
  for (size_t i = 0; i < ch.sel.size(); ++i) {
    if (!(ch.w[i] > 1e-10)) continue;
    const float wf = float(ch.w[i]);
    out.tets.push_back(C.tet[ch.sel[i]]);
    out.w.push_back(wf);
    const double* col = &P.packed[ch.sel[i] * C.rows];
    for (size_t r = 0; r < res.size(); ++r) res[r] += -double(wf) * col[r];
  }
  out.residual = P.tn > 1e-12 ? std::sqrt(dotn(res.data(), res.data(), res.size())) / P.tn : 0.0;
  if (out.residual > p.residual_bound + 1e-12) {
    char msg[240];
    std::snprintf(msg, sizeof msg,
                  "Stable Neo-Hookean Cubature source vertex %u has packed training residual %.17g; required at most %.17g.",
                  src, out.residual, p.residual_bound);
    raise(JGS2_ERROR_INFEASIBLE, msg);
  }
  for (uint32_t t : out.tets)
    for (int l = 0; l < 4; ++l) {
      const uint32_t v = p.tets[4 * size_t(t) + l];
      const auto it = std::lower_bound(P.verts.begin(), P.verts.end(), v);
      const bool in = it != P.verts.end() && *it == v;
      const size_t li = size_t(it - P.verts.begin());
      for (int i = 0; i < 9; ++i) out.basis.push_back(in ? float(P.U[9 * li + size_t(i)]) : 0.0f);
    }
  return out;
}

VertexResult train_vertex_snh(const Precompute& p, const std::vector<Pose>& poses, const std::vector<Vec>& frames,
                              uint32_t src, Support& s) {
  Prepared P;
  prepare_vertex_snh(p, poses, frames, src, s, P);
  const Chosen ch = select_host(p, P);
  return finish_vertex_snh(p, P, ch, src);
}

// Co-rotated linear training over displacement modes (cubature.ts:312-369, 714-826).
VertexResult train_vertex_linear(const Precompute& p, const std::vector<Vec>& modes, uint32_t src, Support& s) {
  solve_basis(p, src, s);
  Columns C;
  C.tet = support_tets(p, s);
  const size_t n = C.tet.size(), P = modes.size();
  C.rows = 3 * P;
  C.col.assign(n * C.rows, 0.0);
  C.target.assign(C.rows, 0.0);
  const double ih2 = 1.0 / (p.h * p.h);
  for (size_t q = 0; q < P; ++q) {
    const double* u = modes[q].data();
    double pt[3] = {0, 0, 0};
    for (size_t k = 0; k < n; ++k) {
      const uint32_t t = C.tet[k];
      double Kt[144], g[12];
      tet_stiffness(p, t, Kt);
      for (int r = 0; r < 12; ++r) {
        double v = 0.0;
        for (int c = 0; c < 12; ++c) {
          const uint32_t vx = p.tets[4 * size_t(t) + c / 3];
          v += Kt[r * 12 + c] * u[3 * size_t(vx) + size_t(c % 3)];
          if (r == c) v += (p.mass[vx] * ih2 * u[3 * size_t(vx) + size_t(c % 3)]) / double(p.incident[vx]);
        }
        g[r] = v;
      }
      double r3[3] = {0, 0, 0};
      int slot = -1;
      for (int l = 0; l < 4; ++l) {
        const uint32_t vx = p.tets[4 * size_t(t) + l];
        if (vx == src && slot < 0) slot = l;
        const int32_t li = s.local[vx];
        if (li < 0) continue;
        const double* B = &s.U[9 * size_t(li)];
        for (int rc = 0; rc < 3; ++rc)
          for (int c = 0; c < 3; ++c) r3[c] += B[rc * 3 + c] * g[l * 3 + rc];
      }
      if (slot >= 0)
        for (int c = 0; c < 3; ++c) r3[c] -= g[slot * 3 + c];
      for (int c = 0; c < 3; ++c) {
        C.col[k * C.rows + 3 * q + size_t(c)] = r3[c];
        pt[c] += r3[c];
      }
    }
    const double pn = std::sqrt(pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]);
    if (pn > 1e-12)
      for (int c = 0; c < 3; ++c) {
        const size_t row = 3 * q + size_t(c);
        C.target[row] = pt[c] / pn;
        for (size_t k = 0; k < n; ++k) C.col[k * C.rows + row] /= pn;
      }
  }
  VertexResult out;
  bool any = false;
  for (size_t k = 0; k < n && !any; ++k) any = dotn(C.c(k), C.c(k), C.rows) > 1e-18;
  if (std::sqrt(dotn(C.target.data(), C.target.data(), C.rows)) <= 1e-12 || !any) return out;
  C.prepare();
  const Chosen ch = greedy_select(C, p.K, 0.01, 4);
  column_rank(C, out.nonzero, out.rank);
  out.packed_rank = out.rank;
  for (size_t i = 0; i < ch.sel.size(); ++i)
    if (ch.w[i] > 1e-10) {
      const uint32_t t = C.tet[ch.sel[i]];
      out.tets.push_back(t);
      out.w.push_back(float(ch.w[i]));
      for (int l = 0; l < 4; ++l) {
        const int32_t li = s.local[p.tets[4 * size_t(t) + l]];
        for (int e = 0; e < 9; ++e) out.basis.push_back(li < 0 ? 0.0f : float(s.U[9 * size_t(li) + size_t(e)]));
      }
    }
  out.residual = ch.normalized;
  return out;
}

}  // namespace
}  // namespace jgs2

extern "C" jgs2_status jgs2_precompute_cubature(const jgs2_precompute_input* input,
                                                const jgs2_precompute_options* options,
                                                const jgs2_precompute_output* output) {
  using namespace jgs2;
  clear_error();
  if (!input || !output || !output->cubature_tet_ids || !output->cubature_weights || !output->cubature_basis)
    return fail(JGS2_ERROR_TYPE, "jgs2_precompute_cubature: input and output arrays must not be NULL.");
  try {
    Precompute p;
    jgs2_precompute_options opt{};
    if (options) opt = *options;
    if (opt.mode_solver > 3)
      raise(JGS2_ERROR_RANGE,
            "mode_solver must be 0 (automatic), 1 (dense), 2 (host conjugate gradients) or 3 (device conjugate gradients).");
    p.rings = opt.basis_rings;
    p.mode_solver = opt.mode_solver;
    p.cg_tolerance = opt.cg_tolerance > 0 ? opt.cg_tolerance : 1e-12;
    if (!(opt.residual_bound >= 0) || !std::isfinite(opt.residual_bound))
      raise(JGS2_ERROR_RANGE, "residual_bound must be finite and nonnegative.");
    p.residual_bound = opt.residual_bound > 0 ? opt.residual_bound : 0.01;
    p.threads = opt.threads ? opt.threads : std::max(1u, std::thread::hardware_concurrency());
    p.exact_all_elements = opt.exact_all_elements != 0;
    if (p.exact_all_elements && p.rings != 0)
      raise(JGS2_ERROR_RANGE, "Exact all-element cubature uses the whole-mesh basis (basis_rings 0).");
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto seconds = [](clock::time_point a, clock::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    validate_and_load(p, *input);
    build_rest(p);
    const auto t1 = clock::now();
    const size_t V = p.V, K = p.K;
    std::fill(output->cubature_tet_ids, output->cubature_tet_ids + V * K, 0xffffffffu);
    std::fill(output->cubature_weights, output->cubature_weights + V * K, 0.0f);
    std::fill(output->cubature_basis, output->cubature_basis + V * K * 36, 0.0f);
    if (output->vertex_residuals) std::fill(output->vertex_residuals, output->vertex_residuals + V, NAN);
    if (output->vertex_ranks) std::fill(output->vertex_ranks, output->vertex_ranks + 3 * V, 0u);
    if (p.exact_all_elements) {
      // buildPhase0ForceFreeScene (canonical-phase0.ts:144-205): every free
      // vertex's exact basis at the four vertices of every tetrahedron,
      // tetrahedra in order, weight 1, no training.
      if (K != p.T)
        raise(JGS2_ERROR_RANGE, "Exact all-element cubature packs every tetrahedron: cubature_samples must equal "
                                "tetrahedron_count.");
      std::vector<Support> supports(p.threads);
      parallel_for(p.active.size(), p.threads, [&](size_t a, unsigned worker) {
        const uint32_t v = p.active[a];
        Support& s = supports[worker];
        if (s.local.size() != p.V) s.local.assign(p.V, -1);
        solve_basis(p, v, s);
        for (uint32_t t = 0; t < p.T; ++t) {
          output->cubature_tet_ids[v * K + t] = t;
          output->cubature_weights[v * K + t] = 1.0f;
          float* b = &output->cubature_basis[(v * K + t) * 36];
          for (int l = 0; l < 4; ++l) {
            const int32_t li = s.local[p.tets[4 * size_t(t) + l]];
            for (int e = 0; e < 9; ++e) b[9 * l + e] = li < 0 ? 0.0f : float(s.U[9 * size_t(li) + size_t(e)]);
          }
        }
        if (output->vertex_residuals) output->vertex_residuals[v] = 0.0;
        for (uint32_t u : s.verts) s.local[u] = -1;
      });
      if (output->timings) {
        const double t[8] = {seconds(t0, t1), 0.0, seconds(t1, clock::now()), p.gpu_ms, double(p.gpu_iterations),
                             p.gpu ? 1.0 : 0.0, 0.0, 0.0};
        std::memcpy(output->timings, t, sizeof t);
      }
      return JGS2_OK;
    }
    std::vector<Pose> poses;
    std::vector<Vec> frames, modes;
    bool has_tet_vertex = false;
    for (uint32_t v : p.active) has_tet_vertex |= p.incident[v] > 0;
    if (p.snh && has_tet_vertex) {
      poses = training_poses(p);
      frames = pose_frames(p, poses);
    } else if (!p.snh) {
      modes = low_modes(p, 8, 24);
    }
    const auto t2 = clock::now();
    std::vector<Support> supports(p.threads);
    auto store = [&](uint32_t v, const VertexResult& r) {
      for (size_t k = 0; k < r.tets.size(); ++k) {
        output->cubature_tet_ids[v * K + k] = r.tets[k];
        output->cubature_weights[v * K + k] = r.w[k];
        std::memcpy(&output->cubature_basis[(v * K + k) * 36], &r.basis[36 * k], 36 * sizeof(float));
      }
      if (output->vertex_residuals) output->vertex_residuals[v] = r.residual;
      if (output->vertex_ranks) {
        output->vertex_ranks[3 * size_t(v)] = r.nonzero;
        output->vertex_ranks[3 * size_t(v) + 1] = r.rank;
        output->vertex_ranks[3 * size_t(v) + 2] = r.packed_rank;
      }
    };
    // Selection on the GPU: stable Neo-Hookean, past the reference's tiny
    // fixtures (no subset audit), at most 8 samples.
    const bool device_ok = p.snh && p.T > 12 && p.K <= 8;
    if (opt.training > 2) raise(JGS2_ERROR_RANGE, "training must be 0 (automatic), 1 (host) or 2 (GPU).");
    if (opt.training == 2 && !device_ok)
      raise(JGS2_ERROR_RANGE, "GPU training needs a stable Neo-Hookean scene of more than 12 tetrahedra and K <= 8.");
    if (opt.training == 2 && !GpuCg::available()) raise(JGS2_ERROR_DEVICE, "GPU training needs a HIP device.");
    const bool device = device_ok && (opt.training == 2 || (opt.training == 0 && p.rings > 0 &&
                                                             p.active.size() >= 256 && GpuCg::available()));
    double select_ms = 0.0;
    if (device) {
      std::vector<uint32_t> list;
      for (uint32_t v : p.active) {
        if (p.incident[v] == 0) {
          VertexResult r;
          store(v, r);
        } else {
          list.push_back(v);
        }
      }
      const size_t chunk = 16384;  // several waves per SIMD for the selection kernel
      for (size_t first = 0; first < list.size(); first += chunk) {
        const size_t n = std::min(chunk, list.size() - first);
        std::vector<Prepared> prep(n);
        parallel_for(n, p.threads, [&](size_t i, unsigned worker) {
          Support& s = supports[worker];
          if (s.local.size() != p.V) s.local.assign(p.V, -1);
          prepare_vertex_snh(p, poses, frames, list[first + i], s, prep[i]);
          for (uint32_t u : s.verts) s.local[u] = -1;
        });
        const uint32_t rows = uint32_t(prep[0].C.rows);
        std::vector<GpuSelectJob> jobs(n);
        size_t col_total = 0, scratch_total = 0;
        for (size_t i = 0; i < n; ++i) {
          const uint32_t cnt = uint32_t(prep[i].C.tet.size());
          jobs[i] = GpuSelectJob{col_total, uint64_t(i) * rows, scratch_total, cnt, 0u, prep[i].tn};
          col_total += size_t(cnt) * rows;
          scratch_total += size_t(cnt) * (1 + K);
        }
        Vec cols(col_total), tgts(n * rows);
        parallel_for(n, p.threads, [&](size_t i, unsigned) {
          std::memcpy(&cols[jobs[i].col_offset], prep[i].C.col.data(), prep[i].C.col.size() * sizeof(double));
          std::memcpy(&tgts[i * rows], prep[i].C.target.data(), rows * sizeof(double));
        });
        std::vector<int32_t> sel(n * K);
        Vec w(n * K), cur(n);
        std::vector<uint32_t> cnt(n);
        double ms = 0.0;
        const jgs2_status st = gpu_greedy_select(cols.data(), cols.size(), tgts.data(), tgts.size(), jobs.data(),
                                                 uint32_t(n), rows, uint32_t(K), 0.01, 12, sel.data(), w.data(),
                                                 cur.data(), cnt.data(), &ms);
        if (st != JGS2_OK) raise(st, "Cubature precompute: the device selection failed.");
        select_ms += ms;
        parallel_for(n, p.threads, [&](size_t i, unsigned) {
          Chosen ch;
          for (uint32_t k = 0; k < cnt[i]; ++k) {
            ch.sel.push_back(size_t(sel[i * K + k]));
            ch.w.push_back(w[i * K + k]);
          }
          ch.normalized = std::sqrt(cur[i]) / prep[i].tn;
          store(list[first + i], finish_vertex_snh(p, prep[i], ch, list[first + i]));
        });
      }
    } else {
      parallel_for(p.active.size(), p.threads, [&](size_t a, unsigned worker) {
        const uint32_t v = p.active[a];
        Support& s = supports[worker];
        if (s.local.size() != p.V) s.local.assign(p.V, -1);
        VertexResult r;
        if (p.snh && p.incident[v] == 0) {
          r.residual = 0.0;
        } else if (p.snh) {
          r = train_vertex_snh(p, poses, frames, v, s);
        } else {
          r = train_vertex_linear(p, modes, v, s);
        }
        for (uint32_t u : s.verts) s.local[u] = -1;
        store(v, r);
      });
    }
    if (output->timings) {
      const double t[8] = {seconds(t0, t1), seconds(t1, t2), seconds(t2, clock::now()), p.gpu_ms,
                           double(p.gpu_iterations), p.gpu ? 1.0 : 0.0, select_ms, device ? 1.0 : 0.0};
      std::memcpy(output->timings, t, sizeof t);
    }
    return JGS2_OK;
  } catch (const Error& e) {
    return fail(e.status, "%s", e.message.c_str());
  } catch (const std::bad_alloc&) {
    return fail(JGS2_ERROR_MEMORY, "jgs2_precompute_cubature: host allocation failed.");
  }
}
