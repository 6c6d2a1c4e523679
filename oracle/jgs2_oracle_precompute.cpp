// CPU restatement of the reference's cubature precompute (TEST
// INFRASTRUCTURE ONLY: loaded by tests/ through oracle/jgs2_oracle.py; the
// product never links it).
//
// It follows, in FP64 and in the reference's operation order, the dense
// algorithm of buildPrecomputedScene (src/simulation/cpu/precompute.ts:218-380):
//   * rest data, lumped masses and the rest Hessian M/h^2 + K with its dense
//     Cholesky factor (src/simulation/cpu/fem.ts:41-335, math.ts:56-132);
//   * the exact vertex basis U_v = H^-1 S (S^T H^-1 S)^-1
//     (precompute.ts:135-200);
//   * low-frequency modes by block inverse iteration and a Jacobi Ritz step
//     (src/simulation/cpu/cubature.ts:55-297);
//   * the nonlinear pose corpus (src/simulation/cpu/nonlinear-cubature.ts:
//     95-620) for stable Neo-Hookean scenes;
//   * per source vertex: co-rotated remainder-gradient columns over every
//     tetrahedron and training pose, greedy nonnegative column selection with
//     exhaustive small NNLS, replacement passes, the small-fixture subset
//     audit, f32 packing and the packed residual and column ranks
//     (src/simulation/cpu/nonlinear-cubature-training.ts:311-843,
//     cubature.ts:370-691);
//   * the co-rotated linear selection over displacement modes
//     (cubature.ts:299-369, 693-826).
// Vertex frames use the CPU polar iteration (src/simulation/cpu/oracle.ts:
// 102-217: up to 20 Newton steps, then right-handed column orthonormalization)
// over the rest-volume-weighted vertex stencil
// (src/simulation/cpu/deformation-gradient.ts:22-172).
//
// The only liberty taken: dot products between training columns and with the
// target are memoized. They are the same products summed in the same order,
// so every NNLS sees bit-identical normal equations.
//
// Pinning: tests/test_precompute_oracle.py checks the pose-corpus anchors
// and every packed selection (tetrahedra, f32 weights, ranks, residuals) of
// the reference's 12-tetrahedron beam fixture (manifests/phase1-cubature.v1.json),
// the exact basis against its defining identities (H U = S (S^T H^-1 S)^-1,
// identity block at the source), the modes against a direct eigensolver, and
// the NNLS against brute force on small random systems.
// Parity unpinned: the co-rotated linear selection (cubature.ts:714-826) has
// no golden vector; only the reference's bounds on it are tested.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using Vec = std::vector<double>;
constexpr double kMatrixEpsilon = 1e-14;  // math.ts:1

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::string g_error;

// ---------------------------------------------------------------- math.ts
double det3(const double* m) {  // math.ts:3-9
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

void inv3(const double* m, double* out) {  // math.ts:11-35
  const double d = det3(m);
  double scale = 0.0;
  for (int i = 0; i < 9; ++i) scale = std::max(scale, std::fabs(m[i]));
  if (scale == 0.0 || std::fabs(d) <= kMatrixEpsilon * scale * scale * scale)
    throw Failure("Cannot invert a singular 3 by 3 matrix.");
  const double r = 1.0 / d;
  out[0] = (m[4] * m[8] - m[5] * m[7]) * r;
  out[1] = (m[2] * m[7] - m[1] * m[8]) * r;
  out[2] = (m[1] * m[5] - m[2] * m[4]) * r;
  out[3] = (m[5] * m[6] - m[3] * m[8]) * r;
  out[4] = (m[0] * m[8] - m[2] * m[6]) * r;
  out[5] = (m[2] * m[3] - m[0] * m[5]) * r;
  out[6] = (m[3] * m[7] - m[4] * m[6]) * r;
  out[7] = (m[1] * m[6] - m[0] * m[7]) * r;
  out[8] = (m[0] * m[4] - m[1] * m[3]) * r;
}

void mul3(const double* a, const double* b, double* out) {  // math.ts:37-54
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int k = 0; k < 3; ++k) v += a[r * 3 + k] * b[k * 3 + c];
      t[r * 3 + c] = v;
    }
  std::memcpy(out, t, sizeof t);
}

void transpose3(const double* m, double* out) {
  const double t[9] = {m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]};
  std::memcpy(out, t, sizeof t);
}

double dot(const double* a, const double* b, size_t n) {  // math.ts:201-211
  double r = 0.0;
  for (size_t i = 0; i < n; ++i) r += a[i] * b[i];
  return r;
}
double sq(const Vec& v) { return dot(v.data(), v.data(), v.size()); }

Vec cholesky(const Vec& a, size_t n) {  // math.ts:56-99
  Vec lower(n * n, 0.0);
  double largest = 0.0;
  for (size_t i = 0; i < n; ++i) largest = std::max(largest, std::fabs(a[i * n + i]));
  const double tolerance = std::max(largest, 1.0) * 1e-13;
  for (size_t r = 0; r < n; ++r)
    for (size_t c = 0; c <= r; ++c) {
      double v = a[r * n + c];
      for (size_t k = 0; k < c; ++k) v -= lower[r * n + k] * lower[c * n + k];
      if (r == c) {
        if (!(v > tolerance) || !std::isfinite(v))
          throw Failure("Rest Hessian is not positive definite at diagonal " + std::to_string(r) + ".");
        lower[r * n + c] = std::sqrt(v);
      } else {
        lower[r * n + c] = v / lower[c * n + c];
      }
    }
  return lower;
}

Vec chol_solve(const Vec& lower, const Vec& rhs, size_t n) {  // math.ts:101-132
  Vec x(n);
  for (size_t r = 0; r < n; ++r) {
    double v = rhs[r];
    for (size_t c = 0; c < r; ++c) v -= lower[r * n + c] * x[c];
    x[r] = v / lower[r * n + r];
  }
  for (size_t rr = n; rr-- > 0;) {
    double v = x[rr];
    for (size_t c = rr + 1; c < n; ++c) v -= lower[c * n + rr] * x[c];
    x[rr] = v / lower[rr * n + rr];
  }
  return x;
}

bool dense_solve(Vec a, Vec b, size_t n, Vec& x) {  // math.ts:134-199
  for (size_t p = 0; p < n; ++p) {
    size_t best = p;
    double mag = std::fabs(a[p * n + p]);
    for (size_t r = p + 1; r < n; ++r) {
      const double m = std::fabs(a[r * n + p]);
      if (m > mag) { mag = m; best = r; }
    }
    if (mag <= kMatrixEpsilon || !std::isfinite(mag)) return false;
    if (best != p) {
      for (size_t c = p; c < n; ++c) std::swap(a[p * n + c], a[best * n + c]);
      std::swap(b[p], b[best]);
    }
    const double d = a[p * n + p];
    for (size_t r = p + 1; r < n; ++r) {
      const double m = a[r * n + p] / d;
      a[r * n + p] = 0.0;
      for (size_t c = p + 1; c < n; ++c) a[r * n + c] -= m * a[p * n + c];
      b[r] -= m * b[p];
    }
  }
  for (size_t rr = n; rr-- > 0;) {
    double v = b[rr];
    for (size_t c = rr + 1; c < n; ++c) v -= a[rr * n + c] * b[c];
    b[rr] = v / a[rr * n + rr];
  }
  x = b;
  return true;
}

// ---------------------------------------------------------------- scene
struct Material { double density, young, poisson; };

struct Scene {
  uint32_t V = 0, T = 0;
  const double* x = nullptr;         // 3 V rest positions
  const uint32_t* tets = nullptr;    // 4 T
  const uint32_t* fixed = nullptr;   // V
  const uint32_t* mat = nullptr;     // T
  std::vector<Material> materials;
  double h = 0.0;
  bool snh = false;
  // fem.ts:167-238
  Vec volume, inv_dm, stiffness, mass;
  std::vector<uint32_t> incident;
  // fem.ts:240-335
  std::vector<int> dof;
  std::vector<uint32_t> active;
  size_t n = 0;
  Vec hessian, lower;
};

void shape_gradients(const double* inv, double* g) {  // fem.ts:70-87
  for (int c = 0; c < 3; ++c) {
    g[3 + c] = inv[c];
    g[6 + c] = inv[3 + c];
    g[9 + c] = inv[6 + c];
    g[c] = -g[3 + c] - g[6 + c] - g[9 + c];
  }
}

void rest_data(Scene& s) {  // fem.ts:41-209
  s.volume.assign(s.T, 0.0);
  s.inv_dm.assign(9 * size_t(s.T), 0.0);
  s.stiffness.assign(144 * size_t(s.T), 0.0);
  for (uint32_t t = 0; t < s.T; ++t) {
    const uint32_t* q = s.tets + 4 * size_t(t);
    const double* a = s.x + 3 * size_t(q[0]);
    double m[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m[r * 3 + c] = s.x[3 * size_t(q[c + 1]) + r] - a[r];
    const double vol = std::fabs(det3(m)) / 6.0;
    if (!(vol > 1e-14) || !std::isfinite(vol)) throw Failure("Cannot assemble a degenerate tetrahedron.");
    double inv[9];
    inv3(m, inv);
    double g[12];
    shape_gradients(inv, g);
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
    const Material& mt = s.materials[s.mat[t]];
    const double lambda = (mt.young * mt.poisson) / ((1 + mt.poisson) * (1 - 2 * mt.poisson));
    const double mu = mt.young / (2 * (1 + mt.poisson));
    double D[36] = {0};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) D[r * 6 + c] = r == c ? lambda + 2 * mu : lambda;
    D[21] = mu;
    D[28] = mu;
    D[35] = mu;
    double DB[72];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 12; ++c) {
        double v = 0.0;
        for (int k = 0; k < 6; ++k) v += D[r * 6 + k] * B[k * 12 + c];
        DB[r * 12 + c] = v;
      }
    double* K = &s.stiffness[144 * size_t(t)];
    for (int r = 0; r < 12; ++r)
      for (int c = r; c < 12; ++c) {
        double v = 0.0;
        for (int k = 0; k < 6; ++k) v += B[k * 12 + r] * DB[k * 12 + c];
        v *= vol;
        K[r * 12 + c] = v;
        K[c * 12 + r] = v;
      }
    s.volume[t] = vol;
    std::memcpy(&s.inv_dm[9 * size_t(t)], inv, sizeof inv);
  }
  s.mass.assign(s.V, 0.0);
  s.incident.assign(s.V, 0);
  for (uint32_t t = 0; t < s.T; ++t) {
    const double vm = (s.materials[s.mat[t]].density * s.volume[t]) / 4.0;
    for (int l = 0; l < 4; ++l) s.mass[s.tets[4 * size_t(t) + l]] += vm;
  }
  for (size_t i = 0; i < 4 * size_t(s.T); ++i) s.incident[s.tets[i]] += 1;
}

void rest_system(Scene& s) {  // fem.ts:240-335
  s.dof.assign(s.V, -1);
  s.active.clear();
  for (uint32_t v = 0; v < s.V; ++v)
    if (s.fixed[v] == 0) {
      s.dof[v] = int(s.active.size() * 3);
      s.active.push_back(v);
    }
  if (s.active.empty()) throw Failure("A simulation scene must have at least one movable vertex.");
  const size_t n = s.active.size() * 3;
  s.n = n;
  s.hessian.assign(n * n, 0.0);
  for (uint32_t t = 0; t < s.T; ++t) {
    const double* K = &s.stiffness[144 * size_t(t)];
    for (int lr = 0; lr < 4; ++lr) {
      const int rb = s.dof[s.tets[4 * size_t(t) + lr]];
      if (rb < 0) continue;
      for (int lc = 0; lc < 4; ++lc) {
        const int cb = s.dof[s.tets[4 * size_t(t) + lc]];
        if (cb < 0) continue;
        for (int rc = 0; rc < 3; ++rc)
          for (int cc = 0; cc < 3; ++cc)
            s.hessian[size_t(rb + rc) * n + size_t(cb + cc)] += K[(lr * 3 + rc) * 12 + lc * 3 + cc];
      }
    }
  }
  const double ih2 = 1.0 / (s.h * s.h);
  for (uint32_t v : s.active) {
    const size_t b = size_t(s.dof[v]);
    const double inertia = s.mass[v] * ih2;
    if (!(inertia > 0) || !std::isfinite(inertia)) throw Failure("Movable vertex has invalid lumped mass.");
    for (int c = 0; c < 3; ++c) s.hessian[(b + c) * n + b + c] += inertia;
  }
  s.lower = cholesky(s.hessian, n);
}

// precompute.ts:135-200. Returns the full 9 V basis (zero blocks at fixed
// vertices); row-major 3x3 block per vertex, row = coordinate of that vertex.
Vec exact_basis(const Scene& s, uint32_t vertex) {
  const size_t n = s.n;
  const size_t base = size_t(s.dof[vertex]);
  Vec sel(n * 3);
  for (int c = 0; c < 3; ++c) {
    Vec rhs(n, 0.0);
    rhs[base + c] = 1.0;
    const Vec sol = chol_solve(s.lower, rhs, n);
    for (size_t r = 0; r < n; ++r) sel[r * 3 + c] = sol[r];
  }
  double comp[9], schur[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) comp[r * 3 + c] = 0.5 * (sel[(base + r) * 3 + c] + sel[(base + c) * 3 + r]);
  inv3(comp, schur);
  Vec basis(9 * size_t(s.V), 0.0);
  for (uint32_t v : s.active) {
    const size_t ab = size_t(s.dof[v]);
    for (int row = 0; row < 3; ++row)
      for (int c = 0; c < 3; ++c) {
        double val = 0.0;
        for (int k = 0; k < 3; ++k) val += sel[(ab + row) * 3 + k] * schur[k * 3 + c];
        basis[9 * size_t(v) + row * 3 + c] = val;
      }
  }
  return basis;
}

// ---------------------------------------------------------------- cubature.ts
void canonicalize_sign(Vec& v) {  // cubature.ts:55-67
  size_t largest = 0;
  for (size_t i = 1; i < v.size(); ++i)
    if (std::fabs(v[i]) > std::fabs(v[largest])) largest = i;
  if (v[largest] < 0)
    for (double& x : v) x *= -1;
}

std::vector<Vec> orthonormalize_block(const std::vector<Vec>& cands, size_t n) {  // cubature.ts:69-128
  std::vector<Vec> out;
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    Vec c = cands[ci];
    for (int pass = 0; pass < 2; ++pass)
      for (const Vec& b : out) {
        const double p = dot(c.data(), b.data(), n);
        for (size_t r = 0; r < n; ++r) c[r] -= p * b[r];
      }
    double norm = std::sqrt(sq(c));
    if (norm <= 1e-11) {
      bool found = false;
      for (size_t off = 0; off < n && !found; ++off) {
        Vec trial(n, 0.0);
        trial[(ci + off) % n] = 1.0;
        for (const Vec& b : out) {
          const double p = dot(trial.data(), b.data(), n);
          for (size_t r = 0; r < n; ++r) trial[r] -= p * b[r];
        }
        const double tn = std::sqrt(sq(trial));
        if (tn > 1e-11) {
          for (double& x : trial) x /= tn;
          c = trial;
          found = true;
        }
      }
      if (!found) throw Failure("Could not construct the requested low-frequency basis.");
      norm = 1.0;
    }
    for (double& x : c) x /= norm;
    canonicalize_sign(c);
    out.push_back(c);
  }
  return out;
}

void jacobi_eigen(const Vec& input, size_t d, Vec& values, Vec& vectors) {  // cubature.ts:130-222
  Vec a = input;
  Vec vec(d * d, 0.0);
  for (size_t i = 0; i < d; ++i) vec[i * d + i] = 1.0;
  for (size_t it = 0; it < d * d * 16; ++it) {
    size_t p = 0, q = 1;
    double largest = 0.0;
    for (size_t r = 0; r < d; ++r)
      for (size_t c = r + 1; c < d; ++c) {
        const double m = std::fabs(a[r * d + c]);
        if (m > largest) { largest = m; p = r; q = c; }
      }
    const double ds = std::max({std::fabs(a[p * d + p]), std::fabs(a[q * d + q]), 1.0});
    if (largest <= 1e-12 * ds) break;
    const double app = a[p * d + p], aqq = a[q * d + q], apq = a[p * d + q];
    const double angle = 0.5 * std::atan2(2 * apq, aqq - app);
    const double co = std::cos(angle), si = std::sin(angle);
    for (size_t i = 0; i < d; ++i) {
      if (i == p || i == q) continue;
      const double aip = a[i * d + p], aiq = a[i * d + q];
      const double rp = co * aip - si * aiq, rq = si * aip + co * aiq;
      a[i * d + p] = rp;
      a[p * d + i] = rp;
      a[i * d + q] = rq;
      a[q * d + i] = rq;
    }
    a[p * d + p] = co * co * app - 2 * si * co * apq + si * si * aqq;
    a[q * d + q] = si * si * app + 2 * si * co * apq + co * co * aqq;
    a[p * d + q] = 0.0;
    a[q * d + p] = 0.0;
    for (size_t r = 0; r < d; ++r) {
      const double vip = vec[r * d + p], viq = vec[r * d + q];
      vec[r * d + p] = co * vip - si * viq;
      vec[r * d + q] = si * vip + co * viq;
    }
  }
  std::vector<size_t> order(d);
  for (size_t i = 0; i < d; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t l, size_t r) { return a[l * d + l] < a[r * d + r]; });
  values.assign(d, 0.0);
  vectors.assign(d * d, 0.0);
  for (size_t c = 0; c < d; ++c) {
    values[c] = a[order[c] * d + order[c]];
    for (size_t r = 0; r < d; ++r) vectors[r * d + c] = vec[r * d + order[c]];
  }
}

// cubature.ts:224-297: full-coordinate (3 V) modes, zero at fixed vertices.
std::vector<Vec> low_frequency_modes(const Scene& s, size_t requested, size_t iterations) {
  const size_t n = s.n;
  const size_t count = std::min(requested, n);
  std::vector<Vec> modes(count, Vec(n));
  for (size_t m = 0; m < count; ++m)
    for (size_t r = 0; r < n; ++r)
      modes[m][r] = std::sin(double(r + 1) * double(m + 1) * 0.754877666) +
                    0.5 * std::cos(double(r + 1) * double(m + 3) * 0.569840291);
  modes = orthonormalize_block(modes, n);
  for (size_t it = 0; it < iterations; ++it) {
    for (Vec& m : modes) m = chol_solve(s.lower, m, n);
    modes = orthonormalize_block(modes, n);
  }
  Vec proj(count * count);
  std::vector<Vec> hm(count, Vec(n, 0.0));
  for (size_t m = 0; m < count; ++m)
    for (size_t r = 0; r < n; ++r)
      for (size_t c = 0; c < n; ++c) hm[m][r] += s.hessian[r * n + c] * modes[m][c];
  for (size_t r = 0; r < count; ++r)
    for (size_t c = 0; c < count; ++c) proj[r * count + c] = dot(modes[r].data(), hm[c].data(), n);
  Vec values, vectors;
  jacobi_eigen(proj, count, values, vectors);
  std::vector<Vec> out;
  for (size_t m = 0; m < count; ++m) {
    Vec v(n, 0.0);
    for (size_t b = 0; b < count; ++b) {
      const double coef = vectors[b * count + m];
      for (size_t r = 0; r < n; ++r) v[r] += modes[b][r] * coef;
    }
    canonicalize_sign(v);
    Vec full(3 * size_t(s.V), 0.0);
    for (uint32_t vx : s.active)
      for (int c = 0; c < 3; ++c) full[3 * size_t(vx) + c] = v[size_t(s.dof[vx]) + c];
    out.push_back(std::move(full));
  }
  return out;
}

// ---------------------------------------------------------------- SNH
struct SnhParameters { double lambda, mu, alpha; };

SnhParameters snh_parameters(const Material& m) {  // stable-neo-hookean.ts:68-110
  if (!(m.poisson >= 0 && m.poisson < 0.5))
    throw Failure("Stable Neo-Hookean material requires Poisson ratio in [0, 0.5).");
  const double lame_mu = m.young / (2 * (1 + m.poisson));
  const double lame_lambda = (m.young * m.poisson) / ((1 + m.poisson) * (1 - 2 * m.poisson));
  const double mu = (4.0 / 3.0) * lame_mu;
  const double lambda = lame_lambda + (5.0 / 6.0) * lame_mu;
  return {lambda, mu, 1 + mu / lambda - mu / (4 * lambda)};
}

void deformation_gradient(const Scene& s, uint32_t t, const double* x, double* F) {  // stable-neo-hookean.ts:283-313
  const uint32_t* q = s.tets + 4 * size_t(t);
  double d[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) d[r * 3 + c] = x[3 * size_t(q[c + 1]) + r] - x[3 * size_t(q[0]) + r];
  mul3(d, &s.inv_dm[9 * size_t(t)], F);
}

// stable-neo-hookean.ts:225-281 (first Piola only) and :384-475 (gradient).
double snh_gradient(const Scene& s, uint32_t t, const double* x, double* g) {
  const SnhParameters p = snh_parameters(s.materials[s.mat[t]]);
  double F[9];
  deformation_gradient(s, t, x, F);
  const double J = det3(F);
  double I1 = 0.0;
  for (int i = 0; i < 9; ++i) I1 += F[i] * F[i];
  const double cof[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                         F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                         F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
  const double shear = p.mu * (1 - 1 / (I1 + 1));
  const double volume = p.lambda * (J - p.alpha);
  double P[9];
  for (int i = 0; i < 9; ++i) P[i] = shear * F[i] + volume * cof[i];
  double gr[12];
  shape_gradients(&s.inv_dm[9 * size_t(t)], gr);
  for (int l = 0; l < 4; ++l)
    for (int r = 0; r < 3; ++r) {
      double v = 0.0;
      for (int c = 0; c < 3; ++c) v += P[r * 3 + c] * gr[l * 3 + c];
      g[l * 3 + r] = s.volume[t] * v;
    }
  return J;
}

double min_determinant(const Scene& s, const double* x) {
  double m = INFINITY;
  for (uint32_t t = 0; t < s.T; ++t) {
    double F[9];
    deformation_gradient(s, t, x, F);
    m = std::min(m, det3(F));
  }
  return m;
}

// ---------------------------------------------------------------- frames
struct Stencil { std::vector<uint32_t> start, src; Vec coef; };

Stencil vertex_stencil(const Scene& s) {  // deformation-gradient.ts:22-100
  std::vector<std::vector<uint32_t>> inc(s.V);
  for (uint32_t t = 0; t < s.T; ++t)
    for (int l = 0; l < 4; ++l) inc[s.tets[4 * size_t(t) + l]].push_back(t);
  Stencil st;
  st.start.assign(s.V + 1, 0);
  for (uint32_t c = 0; c < s.V; ++c) {
    if (inc[c].empty()) throw Failure("Vertex has no incident tetrahedron for its deformation frame.");
    double total = 0.0;
    std::map<uint32_t, std::array<double, 3>> co;
    for (uint32_t t : inc[c]) {
      const double vol = s.volume[t];
      total += vol;
      double g[12];
      shape_gradients(&s.inv_dm[9 * size_t(t)], g);
      for (int l = 0; l < 4; ++l) {
        auto& e = co.emplace(s.tets[4 * size_t(t) + l], std::array<double, 3>{0.0, 0.0, 0.0}).first->second;
        for (int k = 0; k < 3; ++k) e[k] += vol * g[l * 3 + k];
      }
    }
    for (const auto& kv : co) {
      st.src.push_back(kv.first);
      for (int k = 0; k < 3; ++k) st.coef.push_back(kv.second[k] / total);
    }
    st.start[c + 1] = uint32_t(st.src.size());
  }
  return st;
}

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
void cross3(const double* a, const double* b, double* o) {
  const double t[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  std::memcpy(o, t, sizeof t);
}
void normalized3(const double* v, const double* fallback, double* o) {  // oracle.ts:118-131
  const double n = norm3(v);
  if (!(n > 1e-14) || !std::isfinite(n)) {
    std::memcpy(o, fallback, 3 * sizeof(double));
    return;
  }
  o[0] = v[0] / n;
  o[1] = v[1] / n;
  o[2] = v[2] / n;
}
void perpendicular(const double* axis, double* o) {  // oracle.ts:133-139
  const double cand[3] = {std::fabs(axis[0]) > 0.8 ? 0.0 : 1.0, std::fabs(axis[0]) > 0.8 ? 1.0 : 0.0, 0.0};
  double c[3];
  cross3(axis, cand, c);
  const double fb[3] = {0, 0, 1};
  normalized3(c, fb, o);
}

void orthonormalize_rotation(const double* m, double* out) {  // oracle.ts:141-179
  const double i0[3] = {m[0], m[3], m[6]}, i1[3] = {m[1], m[4], m[7]}, i2[3] = {m[2], m[5], m[8]};
  const double e0[3] = {1, 0, 0}, e2[3] = {0, 0, 1};
  double c0[3], c1[3], c2[3], pp[3], cr[3];
  normalized3(i0, e0, c0);
  const double proj = c0[0] * i1[0] + c0[1] * i1[1] + c0[2] * i1[2];
  const double cand[3] = {i1[0] - proj * c0[0], i1[1] - proj * c0[1], i1[2] - proj * c0[2]};
  perpendicular(c0, pp);
  normalized3(cand, pp, c1);
  cross3(c0, c1, cr);
  normalized3(cr, e2, c2);
  if (c2[0] * i2[0] + c2[1] * i2[1] + c2[2] * i2[2] < 0) {
    for (int k = 0; k < 3; ++k) {
      c1[k] = -c1[k];
      c2[k] = -c2[k];
    }
  }
  cross3(c2, c0, cr);
  perpendicular(c0, pp);
  normalized3(cr, pp, c1);
  const double r[9] = {c0[0], c1[0], c2[0], c0[1], c1[1], c2[1], c0[2], c1[2], c2[2]};
  std::memcpy(out, r, sizeof r);
}

void polar_rotation(const double* m, double* out) {  // oracle.ts:182-217
  double r[9];
  std::memcpy(r, m, sizeof r);
  for (int it = 0; it < 20; ++it) {
    const double d = det3(r);
    if (!(std::fabs(d) > 1e-14) || !std::isfinite(d)) break;
    double inv[9], next[9];
    inv3(r, inv);
    double change = 0.0;
    for (int row = 0; row < 3; ++row)
      for (int c = 0; c < 3; ++c) {
        const int i = row * 3 + c;
        next[i] = 0.5 * (r[i] + inv[c * 3 + row]);
        change = std::max(change, std::fabs(next[i] - r[i]));
      }
    std::memcpy(r, next, sizeof r);
    if (change <= 1e-14) break;
  }
  orthonormalize_rotation(r, out);
}

Vec vertex_frames(const Stencil& st, uint32_t V, const double* x) {  // deformation-gradient.ts:103-172
  Vec out(9 * size_t(V));
  for (uint32_t c = 0; c < V; ++c) {
    double F[9] = {0};
    for (uint32_t rec = st.start[c]; rec < st.start[c + 1]; ++rec)
      for (int w = 0; w < 3; ++w) {
        const double pos = x[3 * size_t(st.src[rec]) + w];
        for (int m = 0; m < 3; ++m) F[w * 3 + m] += pos * st.coef[3 * size_t(rec) + m];
      }
    polar_rotation(F, &out[9 * size_t(c)]);
  }
  return out;
}

// ---------------------------------------------------------------- corpus
struct Pose { Vec x, xp; double min_j; };

// Math.hypot as ECMAScript engines evaluate it (V8: scale by the largest
// magnitude, Kahan-compensated sum of squares, sqrt times the scale).
double hypot3(double a, double b, double c) {
  const double v[3] = {std::fabs(a), std::fabs(b), std::fabs(c)};
  const double mx = std::max({v[0], v[1], v[2]});
  if (mx == 0.0) return 0.0;
  if (std::isinf(mx)) return INFINITY;
  double sum = 0.0, comp = 0.0;
  for (double x : v) {
    const double n = x / mx;
    const double summand = n * n - comp;
    const double pre = sum + summand;
    comp = (pre - sum) - summand;
    sum = pre;
  }
  return std::sqrt(sum) * mx;
}

double max_vertex_norm(const Vec& d) {
  double m = 0.0;
  for (size_t v = 0; v < d.size() / 3; ++v) m = std::max(m, hypot3(d[3 * v], d[3 * v + 1], d[3 * v + 2]));
  return m;
}

double scene_scale(const Scene& s) {  // nonlinear-cubature.ts:95-130
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < s.V; ++v)
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], s.x[3 * size_t(v) + c]);
      hi[c] = std::max(hi[c], s.x[3 * size_t(v) + c]);
    }
  return hypot3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
}

double direction_measure(const Scene& s, const Vec& d, double scale) {  // nonlinear-cubature.ts:146-182
  double mg = 0.0;
  for (uint32_t t = 0; t < s.T; ++t) {
    const uint32_t* q = s.tets + 4 * size_t(t);
    double sh[9], g[9];
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) sh[r * 3 + c] = d[3 * size_t(q[c + 1]) + r] - d[3 * size_t(q[0]) + r];
    mul3(sh, &s.inv_dm[9 * size_t(t)], g);
    double f = 0.0;
    for (double v : g) f += v * v;
    mg = std::max(mg, std::sqrt(f));
  }
  return std::max(mg, max_vertex_norm(d) / scale);
}

Pose scaled_pose(const Scene& s, const Vec& raw, double amplitude, double blend, double scale) {  // :218-256
  const double measure = direction_measure(s, raw, scale);
  if (!(measure > 1e-14) || !std::isfinite(measure)) throw Failure("Nonlinear Cubature pose has zero direction.");
  const double k = amplitude / measure;
  Pose p;
  p.x.assign(s.x, s.x + 3 * size_t(s.V));
  p.xp = p.x;
  for (size_t e = 0; e < p.x.size(); ++e) {
    const double disp = raw[e] * k;
    p.x[e] += disp;
    p.xp[e] += blend * disp;
  }
  p.min_j = min_determinant(s, p.x.data());  // :184-216
  if (!(p.min_j >= 0.5)) throw Failure("Nonlinear Cubature pose has minimum J below 0.5.");
  return p;
}

bool append_mode(std::vector<Vec>& dst, Vec c) {  // nonlinear-cubature.ts:310-336
  for (const Vec& e : dst) {
    const double p = dot(c.data(), e.data(), c.size());
    for (size_t i = 0; i < c.size(); ++i) c[i] -= p * e[i];
  }
  const double n = std::sqrt(sq(c));
  if (!(n > 1e-10) || !std::isfinite(n)) return false;
  for (double& v : c) v /= n;
  for (double v : c) {
    if (std::fabs(v) <= 1e-12) continue;
    if (v < 0)
      for (double& w : c) w *= -1;
    break;
  }
  dst.push_back(std::move(c));
  return true;
}

std::vector<Vec> rigid_modes(const Scene& s) {  // nonlinear-cubature.ts:338-422
  std::vector<uint32_t> parent(s.V);
  for (uint32_t v = 0; v < s.V; ++v) parent[v] = v;
  std::function<uint32_t(uint32_t)> find = [&](uint32_t v) {
    uint32_t r = v;
    while (parent[r] != r) r = parent[r];
    while (parent[v] != v) {
      const uint32_t nx = parent[v];
      parent[v] = r;
      v = nx;
    }
    return r;
  };
  for (uint32_t t = 0; t < s.T; ++t)
    for (int l = 1; l < 4; ++l) {
      const uint32_t a = find(s.tets[4 * size_t(t)]), b = find(s.tets[4 * size_t(t) + l]);
      if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
  std::map<uint32_t, std::vector<uint32_t>> comps;
  for (uint32_t v = 0; v < s.V; ++v) comps[find(v)].push_back(v);
  std::vector<std::vector<uint32_t>> ordered;
  for (auto& kv : comps) ordered.push_back(kv.second);
  std::sort(ordered.begin(), ordered.end(), [](const auto& l, const auto& r) { return l[0] < r[0]; });
  std::vector<Vec> out;
  for (const auto& vs : ordered) {
    bool pinned = false;
    for (uint32_t v : vs) pinned |= s.fixed[v] != 0;
    if (pinned) continue;
    double total = 0.0, ctr[3] = {0, 0, 0};
    for (uint32_t v : vs) {
      total += s.mass[v];
      for (int c = 0; c < 3; ++c) ctr[c] += s.mass[v] * s.x[3 * size_t(v) + c];
    }
    for (double& c : ctr) c /= total;
    std::vector<Vec> cand(6, Vec(3 * size_t(s.V), 0.0));
    for (uint32_t v : vs) {
      const double x = s.x[3 * size_t(v)] - ctr[0], y = s.x[3 * size_t(v) + 1] - ctr[1],
                   z = s.x[3 * size_t(v) + 2] - ctr[2];
      for (int c = 0; c < 3; ++c) cand[c][3 * size_t(v) + c] = 1.0;
      const double r3[3] = {0, -z, y}, r4[3] = {z, 0, -x}, r5[3] = {-y, x, 0};
      for (int c = 0; c < 3; ++c) {
        cand[3][3 * size_t(v) + c] = r3[c];
        cand[4][3 * size_t(v) + c] = r4[c];
        cand[5][3 * size_t(v) + c] = r5[c];
      }
    }
    for (auto& c : cand) append_mode(out, c);
  }
  return out;
}

const double kRotations[4][9] = {{1, 0, 0, 0, 0, -1, 0, 1, 0},
                                 {0, 0, 1, 0, 1, 0, -1, 0, 0},
                                 {0, -1, 0, 1, 0, 0, 0, 0, 1},
                                 {0, 0, 1, 1, 0, 0, 0, 1, 0}};  // nonlinear-cubature.ts:449-454

Vec rotate_positions(const Vec& x, const double* ctr, const double* R) {  // :284-308
  Vec out(x.size());
  for (size_t v = 0; v < x.size() / 3; ++v) {
    const double rel[3] = {x[3 * v] - ctr[0], x[3 * v + 1] - ctr[1], x[3 * v + 2] - ctr[2]};
    for (int r = 0; r < 3; ++r) out[3 * v + r] = ctr[r] + R[r * 3] * rel[0] + R[r * 3 + 1] * rel[1] + R[r * 3 + 2] * rel[2];
  }
  return out;
}

// nonlinear-cubature.ts:462-620 with the default options.
void pose_corpus(const Scene& s, std::vector<Pose>& training, std::vector<Pose>& validation) {
  const double scale = scene_scale(s);
  bool has_fixed = false;
  for (uint32_t v = 0; v < s.V; ++v) has_fixed |= s.fixed[v] != 0;
  const size_t requested = 8 + 4;
  const std::vector<Vec> rigid = rigid_modes(s);
  const std::vector<Vec> raw = low_frequency_modes(s, requested + rigid.size(), 24);
  std::vector<Vec> modes;
  if (rigid.empty()) {
    modes = raw;
  } else {  // :424-447
    for (const Vec& src : raw) {
      Vec c = src;
      for (const Vec& r : rigid) {
        const double p = dot(c.data(), r.data(), c.size());
        for (size_t i = 0; i < c.size(); ++i) c[i] -= p * r[i];
      }
      append_mode(modes, c);
      if (modes.size() >= requested) break;
    }
    if (modes.empty()) throw Failure("Nonlinear Cubature free components have no deformational modes.");
  }
  training.clear();
  validation.clear();
  for (size_t m = 0; m < std::min<size_t>(8, modes.size()); ++m)
    for (double sign : {1.0, -1.0}) {
      Vec dir(modes[m].size());
      for (size_t i = 0; i < dir.size(); ++i) dir[i] = sign * modes[m][i];
      training.push_back(scaled_pose(s, dir, 0.12, 0.0, scale));
    }
  for (size_t idx = 0; idx < 8; ++idx) {
    Vec comb(3 * size_t(s.V), 0.0);
    for (size_t m = 0; m < modes.size(); ++m) {
      const double coef = std::sin(double(idx + 1) * double(m + 1) * 0.754877666) +
                          0.5 * std::cos(double(idx + 1) * double(m + 3) * 0.569840291);
      for (size_t e = 0; e < comb.size(); ++e) comb[e] += coef * modes[m][e];
    }
    validation.push_back(scaled_pose(s, comb, idx < 4 ? 0.06 : 0.18, 0.25, scale));
  }
  if (!has_fixed) {
    double total = 0.0, ctr[3] = {0, 0, 0};
    for (uint32_t v = 0; v < s.V; ++v) {
      total += s.mass[v];
      for (int c = 0; c < 3; ++c) ctr[c] += s.mass[v] * s.x[3 * size_t(v) + c];
    }
    for (double& c : ctr) c /= total;
    for (int r = 0; r < 4; ++r) {
      Pose p;
      p.x = rotate_positions(validation[size_t(r)].x, ctr, kRotations[r]);
      p.xp = rotate_positions(validation[size_t(r)].xp, ctr, kRotations[r]);
      p.min_j = min_determinant(s, p.x.data());
      if (!(p.min_j >= 0.5)) throw Failure("Nonlinear Cubature pose has minimum J below 0.5.");
      validation.push_back(std::move(p));
    }
  }
}

// ---------------------------------------------------------------- selection
struct Columns {
  size_t rows = 0;
  std::vector<Vec> col;
  Vec target;
  // Memoized dot products (bit-identical to recomputing them).
  mutable std::vector<double> gram, tdot;
  mutable std::vector<uint8_t> have;
  void prepare() {
    const size_t c = col.size();
    gram.assign(c * c, 0.0);
    have.assign(c * c, 0);
    tdot.resize(c);
    for (size_t i = 0; i < c; ++i) tdot[i] = dot(col[i].data(), target.data(), rows);
  }
  double g(size_t i, size_t j) const {
    const size_t k = i * col.size() + j;
    if (!have[k]) {
      gram[k] = dot(col[i].data(), col[j].data(), rows);
      have[k] = 1;
    }
    return gram[k];
  }
};

Vec residual_of(const Columns& C, const std::vector<size_t>& sel, const Vec& w) {  // cubature.ts:370-383
  Vec r = C.target;
  for (size_t k = 0; k < sel.size(); ++k) {
    const Vec& c = C.col[sel[k]];
    for (size_t i = 0; i < r.size(); ++i) r[i] -= c[i] * w[k];
  }
  return r;
}

Vec small_nnls(const Columns& C, const std::vector<size_t>& sel) {  // cubature.ts:386-433
  const size_t n = sel.size();
  Vec best(n, 0.0);
  double best_err = sq(C.target);
  for (size_t mask = 1; mask < (size_t(1) << n); ++mask) {
    std::vector<size_t> act;
    for (size_t c = 0; c < n; ++c)
      if (mask & (size_t(1) << c)) act.push_back(c);
    const size_t d = act.size();
    Vec A(d * d), b(d), x;
    for (size_t r = 0; r < d; ++r) {
      b[r] = C.tdot[sel[act[r]]];
      for (size_t c = 0; c < d; ++c) A[r * d + c] = C.g(sel[act[r]], sel[act[c]]);
    }
    if (!dense_solve(A, b, d, x)) continue;
    bool negative = false;
    for (double v : x) negative |= v < -1e-10;
    if (negative) continue;
    Vec w(n, 0.0);
    for (size_t i = 0; i < d; ++i) w[act[i]] = std::max(0.0, x[i]);
    const double err = sq(residual_of(C, sel, w));
    if (err < best_err - 1e-14 * std::max(1.0, best_err)) {
      best_err = err;
      best = w;
    }
  }
  return best;
}

struct Selection { std::vector<size_t> sel; Vec w; double target_norm = 0, normalized = 0; };

// cubature.ts:508-691
Selection greedy(const Columns& C, size_t max_columns, double tolerance, size_t passes) {
  std::vector<size_t> cand;
  for (size_t i = 0; i < C.col.size(); ++i)
    if (sq(C.col[i]) > 1e-18) cand.push_back(i);
  const size_t limit = std::min(max_columns, cand.size());
  if (limit > 12) throw Failure("Exhaustive NNLS supports at most 12 selected columns.");
  Selection out;
  out.target_norm = std::sqrt(sq(C.target));
  if (out.target_norm <= 1e-12) return out;
  if (cand.empty() || limit == 0) {
    out.normalized = 1.0;
    return out;
  }
  auto contains = [](const std::vector<size_t>& v, size_t x) { return std::find(v.begin(), v.end(), x) != v.end(); };
  Vec residual = C.target;
  for (size_t step = 0; step < limit; ++step) {
    long best = -1;
    Vec bw, br;
    double be = sq(residual);
    for (size_t c : cand) {
      if (contains(out.sel, c)) continue;
      std::vector<size_t> trial = out.sel;
      trial.push_back(c);
      const Vec w = small_nnls(C, trial);
      const Vec r = residual_of(C, trial, w);
      const double e = sq(r);
      const double tol = 1e-14 * std::max(1.0, be);
      if (e < be - tol || (std::fabs(e - be) <= tol && best >= 0 && long(c) < best)) {
        best = long(c);
        bw = w;
        br = r;
        be = e;
      }
    }
    if (best < 0) break;
    out.sel.push_back(size_t(best));
    out.w = bw;
    residual = br;
    if (std::sqrt(sq(residual)) / out.target_norm < tolerance) break;
  }
  for (size_t pass = 0; pass < passes && !out.sel.empty(); ++pass) {
    long slot = -1, best = -1;
    Vec bw, br;
    double be = sq(residual);
    for (size_t sl = 0; sl < out.sel.size(); ++sl)
      for (size_t c : cand) {
        if (contains(out.sel, c)) continue;
        std::vector<size_t> trial = out.sel;
        trial[sl] = c;
        const Vec w = small_nnls(C, trial);
        const Vec r = residual_of(C, trial, w);
        const double e = sq(r);
        const double tol = 1e-14 * std::max(1.0, be);
        if (e < be - tol || (std::fabs(e - be) <= tol && best >= 0 && long(c) < best)) {
          slot = long(sl);
          best = long(c);
          bw = w;
          br = r;
          be = e;
        }
      }
    if (slot < 0) break;
    out.sel[size_t(slot)] = size_t(best);
    out.w = bw;
    residual = br;
  }
  out.normalized = std::sqrt(sq(residual)) / out.target_norm;
  return out;
}

void column_rank(const std::vector<Vec>& cols, double& nonzero, double& rank) {  // training.ts:594-629
  double mx = 0.0;
  for (const Vec& c : cols) mx = std::max(mx, std::sqrt(sq(c)));
  nonzero = 0;
  rank = 0;
  if (mx == 0.0) return;
  const double th = 1e-9 * mx;
  for (const Vec& c : cols) nonzero += std::sqrt(sq(c)) > th ? 1 : 0;
  std::vector<Vec> basis;
  for (const Vec& col : cols) {
    Vec c = col;
    for (int pass = 0; pass < 2; ++pass)
      for (const Vec& d : basis) {
        double p = 0.0;
        for (size_t r = 0; r < c.size(); ++r) p += c[r] * d[r];
        for (size_t r = 0; r < c.size(); ++r) c[r] -= p * d[r];
      }
    const double n = std::sqrt(sq(c));
    if (n <= th) continue;
    for (double& v : c) v /= n;
    basis.push_back(std::move(c));
  }
  rank = double(basis.size());
}

// nonlinear-cubature-training.ts:311-470 (remainder gradient only) for every
// tetrahedron at one pose, given the co-rotated current basis.
void remainder_gradients(const Scene& s, const Vec& x, const Vec& xp, const Vec& current, uint32_t source,
                         std::vector<std::array<double, 3>>& out) {
  const double ih2 = 1.0 / (s.h * s.h);
  out.assign(s.T, {0.0, 0.0, 0.0});
  for (uint32_t t = 0; t < s.T; ++t) {
    double g[12];
    snh_gradient(s, t, x.data(), g);
    int slot = -1;
    double blocks[36];
    for (int l = 0; l < 4; ++l) {
      const uint32_t v = s.tets[4 * size_t(t) + l];
      std::memcpy(blocks + 9 * l, &current[9 * size_t(v)], 9 * sizeof(double));
      const double di = (s.mass[v] * ih2) / double(s.incident[v]);
      for (int c = 0; c < 3; ++c) g[l * 3 + c] += di * (x[3 * size_t(v) + c] - xp[3 * size_t(v) + c]);
      if (v == source) slot = l;
    }
    double proj[3] = {0, 0, 0};
    for (int r = 0; r < 3; ++r)
      for (int e = 0; e < 12; ++e) proj[r] += blocks[e * 3 + r] * g[e];
    if (slot >= 0)
      for (int r = 0; r < 3; ++r) proj[r] += -1.0 * g[slot * 3 + r];
    out[t] = {proj[0], proj[1], proj[2]};
  }
}

struct Trained {
  std::vector<uint32_t> tets;
  Vec weights;          // packed (f32 values)
  Vec basis;            // f32 values, 36 per sample
  double stats[8] = {0};
};

// nonlinear-cubature-training.ts:489-575.
// The vertex frames of a pose do not depend on the source vertex; they are
// computed once per pose (identical values to recomputing them).
void stacked_system(const Scene& s, const std::vector<Pose>& poses, const Vec& basis, uint32_t source,
                    const std::vector<Vec>& pose_frames, const Vec* norms, Columns& C, Vec& pose_norms, int& valid,
                    int& trivial) {
  C.rows = poses.size() * 3;
  C.col.assign(s.T, Vec(C.rows, 0.0));
  C.target.assign(C.rows, 0.0);
  pose_norms.assign(poses.size(), 0.0);
  valid = trivial = 0;
  for (size_t p = 0; p < poses.size(); ++p) {
    const Vec& frames = pose_frames[p];
    double inv_src[9];
    transpose3(&frames[9 * size_t(source)], inv_src);
    Vec current(9 * size_t(s.V));
    for (uint32_t v = 0; v < s.V; ++v) {
      double left[9];
      mul3(&frames[9 * size_t(v)], &basis[9 * size_t(v)], left);
      mul3(left, inv_src, &current[9 * size_t(v)]);
    }
    std::vector<std::array<double, 3>> rem;
    remainder_gradients(s, poses[p].x, poses[p].xp, current, source, rem);
    double pt[3] = {0, 0, 0};
    for (const auto& r : rem)
      for (int c = 0; c < 3; ++c) pt[c] += 1.0 * r[c];
    double scale_sq = 0.0;
    for (const auto& r : rem) scale_sq += r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double cscale = std::sqrt(scale_sq);
    const double pn = std::sqrt(pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]);
    pose_norms[p] = pn;
    if (cscale == 0.0) {
      ++trivial;
      continue;
    }
    if (pn <= 1e-12 * cscale) throw Failure("Nonlinear Cubature training pose has a cancellation-dominated target.");
    ++valid;
    const double nn = norms ? (*norms)[p] : pn;
    if (!(nn > 0) || !std::isfinite(nn)) throw Failure("Nonlinear Cubature training pose has an invalid norm.");
    for (int c = 0; c < 3; ++c) {
      const size_t row = p * 3 + size_t(c);
      C.target[row] = pt[c] / nn;
      for (uint32_t t = 0; t < s.T; ++t) C.col[t][row] = rem[t][size_t(c)] / nn;
    }
  }
}

Vec round_f32(const Vec& v) {
  Vec o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = double(float(v[i]));
  return o;
}

// nonlinear-cubature-training.ts:731-843 for one source vertex.
Trained train_snh(const Scene& s, const std::vector<Pose>& poses, const std::vector<Vec>& frames, uint32_t source,
                  size_t K) {
  const Vec basis = exact_basis(s, source);
  Columns C;
  Vec norms;
  int valid = 0, trivial = 0;
  stacked_system(s, poses, basis, source, frames, nullptr, C, norms, valid, trivial);
  if (std::sqrt(sq(C.target)) <= 1e-12) throw Failure("Nonlinear Cubature source vertex has no valid training target.");
  C.prepare();
  Selection sel = greedy(C, K, 0.01, 12);
  std::vector<size_t> chosen = sel.sel;
  Vec weights = sel.w;
  double normalized = sel.normalized;
  double nonzero, rank;
  column_rank(C.col, nonzero, rank);
  if (s.T <= 12 && K < s.T) {  // subset audit, training.ts:775-808
    const double tn = std::sqrt(sq(C.target));
    std::vector<size_t> subset;
    std::function<void(size_t)> visit = [&](size_t next) {
      if (subset.size() == K) {
        const Vec w = small_nnls(C, subset);
        const double rn = std::sqrt(sq(residual_of(C, subset, w)));
        const double nr = tn > 1e-12 ? rn / tn : 0.0;
        if (nr < normalized - 1e-14) {
          normalized = nr;
          chosen = subset;
          weights = w;
        }
        return;
      }
      const size_t remaining = K - subset.size();
      for (size_t t = next; t + remaining <= s.T; ++t) {
        subset.push_back(t);
        visit(t + 1);
        subset.pop_back();
      }
    };
    visit(0);
  }
  Trained out;
  Vec kept;
  for (size_t i = 0; i < chosen.size(); ++i)
    if (weights[i] > 1e-10) {
      out.tets.push_back(uint32_t(chosen[i]));
      kept.push_back(weights[i]);
    }
  const Vec packed_basis = round_f32(basis);
  out.weights = round_f32(kept);
  Columns P;
  Vec unused;
  int pvalid = 0, ptrivial = 0;
  stacked_system(s, poses, packed_basis, source, frames, &norms, P, unused, pvalid, ptrivial);
  if (pvalid != valid || ptrivial != trivial) throw Failure("Packing changed nonlinear Cubature pose validity.");
  double pnonzero, prank;
  column_rank(P.col, pnonzero, prank);
  Vec residual = C.target;  // training.ts:578-592 with the f64 target
  for (size_t k = 0; k < out.tets.size(); ++k) {
    const Vec& c = P.col[out.tets[k]];
    for (size_t i = 0; i < residual.size(); ++i) residual[i] += -out.weights[k] * c[i];
  }
  const double tn = std::sqrt(sq(C.target));
  const double packed_residual = tn > 1e-12 ? std::sqrt(sq(residual)) / tn : 0.0;
  for (uint32_t t : out.tets)
    for (int l = 0; l < 4; ++l) {
      const uint32_t v = s.tets[4 * size_t(t) + l];
      for (int i = 0; i < 9; ++i) out.basis.push_back(packed_basis[9 * size_t(v) + i]);
    }
  const double st8[8] = {packed_residual, normalized, double(valid), double(trivial), nonzero, rank, pnonzero, prank};
  std::memcpy(out.stats, st8, sizeof st8);
  return out;
}

// cubature.ts:312-369 and :714-826 for one source vertex.
Trained train_linear(const Scene& s, const std::vector<Vec>& poses, uint32_t source, size_t K) {
  const Vec basis = exact_basis(s, source);
  const double ih2 = 1.0 / (s.h * s.h);
  Columns C;
  C.rows = poses.size() * 3;
  C.col.assign(s.T, Vec(C.rows, 0.0));
  C.target.assign(C.rows, 0.0);
  for (size_t p = 0; p < poses.size(); ++p) {
    const Vec& pose = poses[p];
    double pt[3] = {0, 0, 0};
    for (uint32_t t = 0; t < s.T; ++t) {
      const double* Kt = &s.stiffness[144 * size_t(t)];
      double eg[12];
      for (int r = 0; r < 12; ++r) {
        double v = 0.0;
        for (int c = 0; c < 12; ++c) {
          const uint32_t vx = s.tets[4 * size_t(t) + c / 3];
          const int co = c % 3;
          v += Kt[r * 12 + c] * pose[3 * size_t(vx) + co];
          if (r == c) v += (s.mass[vx] * ih2 * pose[3 * size_t(vx) + co]) / double(s.incident[vx]);
        }
        eg[r] = v;
      }
      double red[3] = {0, 0, 0};
      int slot = -1;
      for (int l = 0; l < 4; ++l) {
        const uint32_t vx = s.tets[4 * size_t(t) + l];
        if (vx == source && slot < 0) slot = l;
        for (int rc = 0; rc < 3; ++rc) {
          const double gv = eg[l * 3 + rc];
          for (int c = 0; c < 3; ++c) red[c] += basis[9 * size_t(vx) + rc * 3 + c] * gv;
        }
      }
      if (slot >= 0)
        for (int c = 0; c < 3; ++c) red[c] -= eg[slot * 3 + c];
      for (int c = 0; c < 3; ++c) {
        C.col[t][p * 3 + size_t(c)] = red[c];
        pt[c] += red[c];
      }
    }
    const double pn = std::sqrt(pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]);
    if (pn > 1e-12)
      for (int c = 0; c < 3; ++c) {
        const size_t row = p * 3 + size_t(c);
        C.target[row] = pt[c] / pn;
        for (uint32_t t = 0; t < s.T; ++t) C.col[t][row] /= pn;
      }
  }
  Trained out;
  bool any = false;
  for (const Vec& c : C.col) any |= sq(c) > 1e-18;
  if (std::sqrt(sq(C.target)) <= 1e-12 || !any) return out;
  C.prepare();
  const Selection sel = greedy(C, K, 0.01, 4);
  Vec kept;
  for (size_t i = 0; i < sel.sel.size(); ++i)
    if (sel.w[i] > 1e-10) {
      out.tets.push_back(uint32_t(sel.sel[i]));
      kept.push_back(sel.w[i]);
    }
  out.weights = round_f32(kept);
  for (uint32_t t : out.tets)
    for (int l = 0; l < 4; ++l) {
      const uint32_t v = s.tets[4 * size_t(t) + l];
      for (int i = 0; i < 9; ++i) out.basis.push_back(double(float(basis[9 * size_t(v) + i])));
    }
  out.stats[0] = sel.normalized;
  out.stats[1] = sel.normalized;
  return out;
}

Scene make_scene(uint32_t V, const double* x, uint32_t T, const uint32_t* tets, const uint32_t* fixed,
                 const uint32_t* mat, uint32_t M, const double* materials, const uint32_t* models, double h) {
  Scene s;
  s.V = V;
  s.T = T;
  s.x = x;
  s.tets = tets;
  s.fixed = fixed;
  s.mat = mat;
  s.h = h;
  for (uint32_t m = 0; m < M; ++m) s.materials.push_back({materials[3 * m], materials[3 * m + 1], materials[3 * m + 2]});
  bool any_snh = false, any_linear = false;
  for (uint32_t t = 0; t < T; ++t) {
    if (mat[t] >= M) throw Failure("Tetrahedron references an unknown material.");
    (models[mat[t]] ? any_snh : any_linear) = true;
  }
  if (any_snh && any_linear)
    throw Failure("Mixed stable Neo-Hookean/co-rotated nonlinear Cubature preprocessing is not implemented.");
  s.snh = any_snh;
  rest_data(s);
  rest_system(s);
  return s;
}

}  // namespace

extern "C" {

const char* jgs2o_precompute_error() { return g_error.c_str(); }

// Pose corpus of a stable Neo-Hookean scene: training then validation poses,
// each as positions (3 V), predicted positions (3 V) and minimum det F.
// Returns the number of training poses (validation count in *validation),
// or -1 on error.
int jgs2o_cubature_corpus(uint32_t V, const double* x, uint32_t T, const uint32_t* tets, const uint32_t* fixed,
                          const uint32_t* mat, uint32_t M, const double* materials, const uint32_t* models,
                          double h, uint32_t capacity, double* positions, double* predicted, double* min_j,
                          uint32_t* validation) {
  try {
    const Scene s = make_scene(V, x, T, tets, fixed, mat, M, materials, models, h);
    std::vector<Pose> tr, va;
    pose_corpus(s, tr, va);
    std::vector<const Pose*> all;
    for (const Pose& p : tr) all.push_back(&p);
    for (const Pose& p : va) all.push_back(&p);
    if (all.size() > capacity) throw Failure("corpus capacity too small");
    for (size_t i = 0; i < all.size(); ++i) {
      std::memcpy(positions + i * 3 * size_t(V), all[i]->x.data(), 3 * size_t(V) * sizeof(double));
      std::memcpy(predicted + i * 3 * size_t(V), all[i]->xp.data(), 3 * size_t(V) * sizeof(double));
      min_j[i] = all[i]->min_j;
    }
    *validation = uint32_t(va.size());
    return int(tr.size());
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// Low-frequency rest modes (full 3 V coordinates) as the reference builds
// them for a given count and iteration number. Returns the count or -1.
int jgs2o_low_frequency_modes(uint32_t V, const double* x, uint32_t T, const uint32_t* tets, const uint32_t* fixed,
                              const uint32_t* mat, uint32_t M, const double* materials, const uint32_t* models,
                              double h, uint32_t count, uint32_t iterations, double* out, double* hessian) {
  try {
    const Scene s = make_scene(V, x, T, tets, fixed, mat, M, materials, models, h);
    const std::vector<Vec> modes = low_frequency_modes(s, count, iterations);
    for (size_t m = 0; m < modes.size(); ++m) std::memcpy(out + m * 3 * size_t(V), modes[m].data(), 3 * size_t(V) * 8);
    if (hessian) std::memcpy(hessian, s.hessian.data(), s.hessian.size() * 8);
    return int(modes.size());
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// Exact basis (9 V, row-major blocks) of one free vertex.
int jgs2o_exact_basis(uint32_t V, const double* x, uint32_t T, const uint32_t* tets, const uint32_t* fixed,
                      const uint32_t* mat, uint32_t M, const double* materials, const uint32_t* models, double h,
                      uint32_t vertex, double* out) {
  try {
    const Scene s = make_scene(V, x, T, tets, fixed, mat, M, materials, models, h);
    if (vertex >= V || s.dof[vertex] < 0) throw Failure("vertex is fixed or out of range");
    const Vec b = exact_basis(s, vertex);
    std::memcpy(out, b.data(), b.size() * 8);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// Greedy nonnegative column selection (cubature.ts:508-691) of `count`
// columns of `rows` values. Writes the selected indices and weights in slot
// order and {target norm, residual norm, normalized residual}; returns the
// number selected or -1.
int jgs2o_select_columns(uint32_t count, uint32_t rows, const double* columns, const double* target,
                         uint32_t max_columns, double tolerance, uint32_t passes, uint32_t* selected,
                         double* weights, double* norms) {
  try {
    Columns C;
    C.rows = rows;
    for (uint32_t c = 0; c < count; ++c) C.col.emplace_back(columns + size_t(c) * rows, columns + size_t(c + 1) * rows);
    C.target.assign(target, target + rows);
    C.prepare();
    const Selection sel = greedy(C, max_columns, tolerance, passes);
    const Vec r = residual_of(C, sel.sel, sel.w);
    for (size_t k = 0; k < sel.sel.size(); ++k) {
      selected[k] = uint32_t(sel.sel[k]);
      weights[k] = sel.w[k];
    }
    norms[0] = sel.target_norm;
    norms[1] = std::sqrt(sq(r));
    norms[2] = sel.normalized;
    return int(sel.sel.size());
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// Exact nonnegative fit of every given column (cubature.ts:475-506):
// weights, then {residual norm, normalized residual}.
int jgs2o_fit_columns(uint32_t count, uint32_t rows, const double* columns, const double* target, double* weights,
                      double* norms) {
  if (count > 12) {
    g_error = "Exact NNLS supports at most 12 columns.";
    return -1;
  }
  Columns C;
  C.rows = rows;
  for (uint32_t c = 0; c < count; ++c) C.col.emplace_back(columns + size_t(c) * rows, columns + size_t(c + 1) * rows);
  C.target.assign(target, target + rows);
  C.prepare();
  std::vector<size_t> all(count);
  for (uint32_t c = 0; c < count; ++c) all[c] = c;
  const Vec w = small_nnls(C, all);
  const double rn = std::sqrt(sq(residual_of(C, all, w)));
  const double tn = std::sqrt(sq(C.target));
  for (uint32_t c = 0; c < count; ++c) weights[c] = w[c];
  norms[0] = rn;
  norms[1] = tn > 1e-12 ? rn / tn : 0.0;
  return 0;
}

// Packed cubature records of every free vertex, K per vertex (the
// reference requires packed residuals <= 0.01; `bound` sets that limit):
// tetrahedron
// ids (0xffffffff where unused), f32 weights, 36 f32 basis values per record,
// and 8 statistics per vertex (packed residual, f64 residual, valid and
// trivial training poses, nonzero candidates, column rank, packed nonzero
// candidates, packed column rank; NaN for fixed vertices).
int jgs2o_precompute_cubature(uint32_t V, const double* x, uint32_t T, const uint32_t* tets, const uint32_t* fixed,
                              const uint32_t* mat, uint32_t M, const double* materials, const uint32_t* models,
                              double h, uint32_t K, double bound, uint32_t* out_tets, float* out_weights,
                              float* out_basis, double* out_stats) {
  try {
    const Scene s = make_scene(V, x, T, tets, fixed, mat, M, materials, models, h);
    std::fill(out_tets, out_tets + size_t(V) * K, 0xffffffffu);
    std::fill(out_weights, out_weights + size_t(V) * K, 0.0f);
    std::fill(out_basis, out_basis + size_t(V) * K * 36, 0.0f);
    std::fill(out_stats, out_stats + size_t(V) * 8, NAN);
    std::vector<Pose> training, validation;
    std::vector<Vec> linear;
    std::vector<Vec> frames;
    bool has_active_tet_vertex = false;
    for (uint32_t v : s.active) has_active_tet_vertex |= s.incident[v] > 0;
    if (s.snh && has_active_tet_vertex) {
      pose_corpus(s, training, validation);
      const Stencil st = vertex_stencil(s);
      for (const Pose& p : training) frames.push_back(vertex_frames(st, s.V, p.x.data()));
    } else if (!s.snh) {
      linear = low_frequency_modes(s, 8, 24);  // precompute.ts:20-21, 256-263
    }
    for (uint32_t v : s.active) {
      Trained tr;
      if (s.snh && s.incident[v] == 0) {
        for (double& q : tr.stats) q = 0.0;
      } else if (s.snh) {
        tr = train_snh(s, training, frames, v, K);
        if (tr.stats[0] > bound + 1e-12)
          throw Failure("Stable Neo-Hookean Cubature source vertex " + std::to_string(v) +
                        " has packed training residual above the bound.");
      } else {
        tr = train_linear(s, linear, v, K);
      }
      for (size_t k = 0; k < tr.tets.size(); ++k) {
        out_tets[size_t(v) * K + k] = tr.tets[k];
        out_weights[size_t(v) * K + k] = float(tr.weights[k]);
        for (int i = 0; i < 36; ++i) out_basis[(size_t(v) * K + k) * 36 + size_t(i)] = float(tr.basis[k * 36 + size_t(i)]);
      }
      std::memcpy(out_stats + size_t(v) * 8, tr.stats, sizeof tr.stats);
    }
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

}  // extern "C"
