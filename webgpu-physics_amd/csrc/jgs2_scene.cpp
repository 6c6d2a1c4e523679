// Host scene setup for libjgs2hip: cuboid meshes, rest data, lumped masses,
// material parameters, canonical CSR adjacency, a cubature record generator
// and the deterministic greedy coloring. Pure host code.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "jgs2_scene_internal.h"

namespace {

// Seeded uniform generator in [0, 1): the reference's corpus PRNG
// (src/scenes/phase1.ts:106-115).
struct CorpusRandom {
  uint32_t state;
  explicit CorpusRandom(uint32_t seed) : state(seed) {}
  double next() {
    state += 0x6d2b79f5u;
    uint32_t v = state;
    v = (v ^ (v >> 15)) * (v | 1u);
    v ^= v + (v ^ (v >> 7)) * (v | 61u);
    return double(v ^ (v >> 14)) / 4294967296.0;
  }
};

double determinant3(const double* m) {  // src/simulation/cpu/math.ts:3-9
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

bool invert3(const double* m, double* out) {  // src/simulation/cpu/math.ts:11-35
  double d = determinant3(m);
  double scale = 0.0;
  for (int i = 0; i < 9; ++i) scale = std::max(scale, std::fabs(m[i]));
  if (scale == 0.0 || std::fabs(d) <= 1e-14 * scale * scale * scale) return false;
  double r = 1.0 / d;
  out[0] = (m[4] * m[8] - m[5] * m[7]) * r;
  out[1] = (m[2] * m[7] - m[1] * m[8]) * r;
  out[2] = (m[1] * m[5] - m[2] * m[4]) * r;
  out[3] = (m[5] * m[6] - m[3] * m[8]) * r;
  out[4] = (m[0] * m[8] - m[2] * m[6]) * r;
  out[5] = (m[2] * m[3] - m[0] * m[5]) * r;
  out[6] = (m[3] * m[7] - m[4] * m[6]) * r;
  out[7] = (m[1] * m[6] - m[0] * m[7]) * r;
  out[8] = (m[0] * m[4] - m[1] * m[3]) * r;
  return true;
}

// Row-major edge matrix [b-a, c-a, d-a] as columns (src/simulation/cpu/fem.ts:41-60).
void rest_matrix(const std::vector<double>& x, const uint32_t* t, double* m) {
  const double* a = &x[3 * size_t(t[0])];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[3 * r + c] = x[3 * size_t(t[c + 1]) + r] - a[r];
}

// 12x12 small-strain stiffness B^T C B V (src/simulation/cpu/fem.ts:89-165).
void linear_stiffness(const double* inv, double volume, double young, double poisson, float* out) {
  double grad[12];
  for (int c = 0; c < 3; ++c) {
    grad[3 + c] = inv[c];
    grad[6 + c] = inv[3 + c];
    grad[9 + c] = inv[6 + c];
    grad[c] = -grad[3 + c] - grad[6 + c] - grad[9 + c];
  }
  double B[6][12] = {};
  for (int v = 0; v < 4; ++v) {
    double gx = grad[3 * v], gy = grad[3 * v + 1], gz = grad[3 * v + 2];
    int col = 3 * v;
    B[0][col] = gx;
    B[1][col + 1] = gy;
    B[2][col + 2] = gz;
    B[3][col] = gy;
    B[3][col + 1] = gx;
    B[4][col + 1] = gz;
    B[4][col + 2] = gy;
    B[5][col] = gz;
    B[5][col + 2] = gx;
  }
  double lam = young * poisson / ((1 + poisson) * (1 - 2 * poisson));
  double mu = young / (2 * (1 + poisson));
  double C[6][6] = {};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[r][c] = r == c ? lam + 2 * mu : lam;
  C[3][3] = C[4][4] = C[5][5] = mu;
  double CB[6][12];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 12; ++c) {
      double s = 0;
      for (int i = 0; i < 6; ++i) s += C[r][i] * B[i][c];
      CB[r][c] = s;
    }
  for (int r = 0; r < 12; ++r)
    for (int c = r; c < 12; ++c) {
      double s = 0;
      for (int i = 0; i < 6; ++i) s += B[i][r] * CB[i][c];
      s *= volume;
      out[r * 12 + c] = float(s);
      out[c * 12 + r] = float(s);
    }
}

}  // namespace

namespace jgs2 {

void cuboid_grid(const uint32_t cells[3], const double origin[3], const double size[3], uint32_t material,
                 uint32_t body, bool fixed, Mesh& m) {
  const uint64_t nx = cells[0], ny = cells[1], nz = cells[2];
  const uint64_t sy = nx + 1, sz = (nx + 1) * (ny + 1);
  const size_t V = size_t((nx + 1) * (ny + 1) * (nz + 1)), T = size_t(nx * ny * nz * 6);
  auto vid = [&](uint64_t x, uint64_t y, uint64_t z) { return uint32_t(x + y * sy + z * sz); };
  m.x.assign(3 * V, 0.0);
  m.fixed.assign(V, fixed ? 1 : 0);
  m.body.assign(V, body);
  m.tets.resize(4 * T);
  m.material.assign(T, material);
  for (uint64_t z = 0; z <= nz; ++z)  // mesh.ts:131-146
    for (uint64_t y = 0; y <= ny; ++y)
      for (uint64_t xx = 0; xx <= nx; ++xx) {
        const uint32_t v = vid(xx, y, z);
        const uint64_t g[3] = {xx, y, z};
        for (int a = 0; a < 3; ++a) m.x[3 * size_t(v) + a] = origin[a] + (size[a] * double(g[a])) / double(cells[a]);
      }
  size_t o = 0;
  for (uint64_t z = 0; z < nz; ++z)  // mesh.ts:152-191
    for (uint64_t y = 0; y < ny; ++y)
      for (uint64_t xx = 0; xx < nx; ++xx) {
        const uint32_t v000 = vid(xx, y, z), v100 = vid(xx + 1, y, z), v010 = vid(xx, y + 1, z),
                       v110 = vid(xx + 1, y + 1, z), v001 = vid(xx, y, z + 1), v101 = vid(xx + 1, y, z + 1),
                       v011 = vid(xx, y + 1, z + 1), v111 = vid(xx + 1, y + 1, z + 1);
        const uint32_t cell[24] = {v000, v100, v110, v111, v000, v110, v010, v111, v000, v010, v011, v111,
                                   v000, v011, v001, v111, v000, v001, v101, v111, v000, v101, v100, v111};
        std::memcpy(&m.tets[o], cell, sizeof(cell));
        o += 24;
      }
}

jgs2_status orient(Mesh& m) {
  for (uint32_t t = 0; t < m.tet_count(); ++t) {
    uint32_t* ix = &m.tets[4 * size_t(t)];
    double r[9];
    rest_matrix(m.x, ix, r);
    const double det = determinant3(r);
    if (std::fabs(det) <= 1e-14) return fail(JGS2_ERROR_RANGE, "Tetrahedron %u is degenerate.", t);
    if (det < 0) std::swap(ix[1], ix[2]);
  }
  return JGS2_OK;
}

void append(Mesh& into, const Mesh& m) {
  const uint32_t off = into.vertex_count();
  into.x.insert(into.x.end(), m.x.begin(), m.x.end());
  into.fixed.insert(into.fixed.end(), m.fixed.begin(), m.fixed.end());
  into.body.insert(into.body.end(), m.body.begin(), m.body.end());
  into.material.insert(into.material.end(), m.material.begin(), m.material.end());
  for (uint32_t i : m.tets) into.tets.push_back(i + off);
}

jgs2_status pack_scene(const Mesh& m, const std::vector<Material>& materials, uint32_t cubature_k,
                       bool rest_stiffness, double cubature_h, jgs2_scene* s) {
  const uint32_t V = m.vertex_count(), T = m.tet_count();
  s->vertex_count = V;
  s->tet_count = T;
  s->cubature_k = cubature_k;
  s->tet_indices = m.tets;
  const std::vector<double>& x = m.x;
  // Rest data (fem.ts:167-209), lumped masses (fem.ts:211-238) and the
  // packed tetrahedron records (scenes/index.ts:1019-1050).
  s->tet_inverse_dm.assign(12 * size_t(T), 0.0f);
  s->tet_meta.resize(4 * size_t(T));
  if (rest_stiffness) s->tet_rest_stiffness.resize(144 * size_t(T));
  std::vector<double> mass(V, 0.0);
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t* ix = &s->tet_indices[4 * size_t(t)];
    double r[9], inv[9];
    rest_matrix(x, ix, r);
    const double volume = std::fabs(determinant3(r)) / 6.0;
    if (!(volume > 1e-14) || !invert3(r, inv)) return fail(JGS2_ERROR_RANGE, "Cannot assemble a degenerate tetrahedron.");
    for (int c = 0; c < 3; ++c)
      for (int q = 0; q < 3; ++q) s->tet_inverse_dm[12 * size_t(t) + 4 * c + q] = float(inv[3 * q + c]);
    if (m.material[t] >= materials.size()) return fail(JGS2_ERROR_RANGE, "Tetrahedron %u has no material.", t);
    const Material& mat = materials[m.material[t]];
    const double E = mat.young, nu = mat.poisson;
    double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    double mu = E / (2 * (1 + nu));
    if (mat.model == JGS2_MATERIAL_STABLE_NEO_HOOKEAN) {
      // computeStableNeoHookeanParameters (stable-neo-hookean.ts:68-109).
      const double adj_mu = (4.0 / 3.0) * mu;
      const double adj_lam = lam + (5.0 / 6.0) * mu;
      lam = adj_lam;
      mu = adj_mu;
    }
    s->tet_meta[4 * size_t(t)] = float(volume);
    s->tet_meta[4 * size_t(t) + 1] = float(lam);
    s->tet_meta[4 * size_t(t) + 2] = float(mu);
    s->tet_meta[4 * size_t(t) + 3] = float(mat.model);
    if (rest_stiffness) linear_stiffness(inv, volume, E, nu, &s->tet_rest_stiffness[144 * size_t(t)]);
    const double vm = mat.density * volume / 4.0;
    for (int l = 0; l < 4; ++l) mass[ix[l]] += vm;
  }
  // Canonical CSR and vertex records (scenes/index.ts:975-1017).
  std::vector<uint32_t> starts(V), counts(V);
  s->adjacency.resize(4 * size_t(T));
  jgs2_build_adjacency(V, T, s->tet_indices.data(), starts.data(), counts.data(), s->adjacency.data());
  s->positions.resize(4 * size_t(V));
  s->vertex_rest.resize(4 * size_t(V));
  s->vertex_info.resize(4 * size_t(V));
  for (uint32_t v = 0; v < V; ++v) {
    for (int a = 0; a < 3; ++a) {
      s->positions[4 * size_t(v) + a] = float(x[3 * size_t(v) + a]);
      s->vertex_rest[4 * size_t(v) + a] = float(x[3 * size_t(v) + a]);
    }
    s->positions[4 * size_t(v) + 3] = 1.0f;
    s->vertex_rest[4 * size_t(v) + 3] = float(mass[v]);
    s->vertex_info[4 * size_t(v)] = starts[v];
    s->vertex_info[4 * size_t(v) + 1] = counts[v];
    s->vertex_info[4 * size_t(v) + 2] = m.fixed[v];
    s->vertex_info[4 * size_t(v) + 3] = m.body[v];
  }
  // Cubature records. The reference trains them offline per scene with a
  // dense exact basis and a nonnegative least-squares selection
  // (src/simulation/cpu/precompute.ts:141-380); the step itself only
  // consumes fixed-width records (layout.ts:24-27). This generator writes
  // records of the same form and role, deterministically and in O(V):
  //   * up to K-2 samples on incident tetrahedra (weight 1), spread evenly
  //     over the vertex's CSR row;
  //   * the remaining samples on second-ring tetrahedra that are not
  //     incident (weight 1/4), which the step treats without the exact-block
  //     subtraction;
  //   * each sample's four Ubar blocks are phi(r) (I + a d e^T) with
  //     r = |x_j - x_i| / h, phi = exp(-r^2), d the unit offset, e a fixed
  //     axis and a = 0.1, so Ubar is I at the source vertex, decays with
  //     distance, is zero on pinned vertices, and is not symmetric.
  const uint32_t K = cubature_k;
  const size_t R = size_t(V) * K;
  s->cubature_tet_ids.assign(R, 0xffffffffu);
  s->cubature_weights.assign(R, 0.0f);
  s->cubature_basis.assign(R * 36, 0.0f);
  if (K > 0) {
    const double hh = cubature_h;
    const double e[3] = {1.0 / std::sqrt(14.0), 2.0 / std::sqrt(14.0), 3.0 / std::sqrt(14.0)};
    for (uint32_t v = 0; v < V; ++v) {
      if (m.fixed[v] || counts[v] == 0) continue;
      const uint32_t* row = &s->adjacency[starts[v]];
      const uint32_t n = counts[v];
      const uint32_t incident_slots = K >= 2 ? std::min<uint32_t>(K - 2, n) : std::min<uint32_t>(K, n);
      uint32_t written = 0;
      auto write_record = [&](uint32_t tet, float weight) {
        size_t rec = size_t(v) * K + written;
        s->cubature_tet_ids[rec] = tet;
        s->cubature_weights[rec] = weight;
        const uint32_t* ix = &s->tet_indices[4 * size_t(tet)];
        for (int l = 0; l < 4; ++l) {
          float* b = &s->cubature_basis[rec * 36 + 9 * size_t(l)];
          uint32_t j = ix[l];
          if (m.fixed[j]) continue;
          double off[3], r2 = 0;
          for (int a = 0; a < 3; ++a) {
            off[a] = x[3 * size_t(j) + a] - x[3 * size_t(v) + a];
            r2 += off[a] * off[a];
          }
          double r = std::sqrt(r2);
          double phi = std::exp(-r2 / (hh * hh));
          double dir[3] = {0, 0, 0};
          if (r > 0)
            for (int a = 0; a < 3; ++a) dir[a] = off[a] / r;
          for (int rr = 0; rr < 3; ++rr)
            for (int cc = 0; cc < 3; ++cc) b[3 * rr + cc] = float(phi * ((rr == cc ? 1.0 : 0.0) + 0.1 * dir[rr] * e[cc]));
        }
        ++written;
      };
      for (uint32_t k = 0; k < incident_slots; ++k) write_record(row[(size_t(k) * n) / incident_slots], 1.0f);
      // Second ring: tetrahedra incident to the vertices of the first
      // incident tetrahedra, not incident to v, in CSR order.
      for (uint32_t a = 0; a < n && written < K; ++a) {
        const uint32_t* ix = &s->tet_indices[4 * size_t(row[a])];
        for (int l = 0; l < 4 && written < K; ++l) {
          uint32_t u = ix[l];
          if (u == v) continue;
          const uint32_t* urow = &s->adjacency[starts[u]];
          for (uint32_t b = 0; b < counts[u] && written < K; ++b) {
            uint32_t cand = urow[(b * 7u + v) % counts[u]];
            const uint32_t* cx = &s->tet_indices[4 * size_t(cand)];
            bool incident = cx[0] == v || cx[1] == v || cx[2] == v || cx[3] == v;
            bool seen = false;
            for (uint32_t q = 0; q < written; ++q) seen |= s->cubature_tet_ids[size_t(v) * K + q] == cand;
            if (!incident && !seen) {
              write_record(cand, 0.25f);
              break;
            }
          }
        }
      }
    }
  }
  return JGS2_OK;
}

}  // namespace jgs2

extern "C" {

jgs2_status jgs2_build_adjacency(uint32_t vertex_count, uint32_t tet_count, const uint32_t* tet_indices,
                                 uint32_t* starts, uint32_t* counts, uint32_t* tets) {
  if (!tet_indices || !starts || !counts || !tets) return jgs2::fail(JGS2_ERROR_TYPE, "NULL adjacency argument.");
  // buildAdjacency (src/scenes/index.ts:811-837): counts, exclusive prefix,
  // then tetrahedra in ascending order into each vertex row.
  std::fill(counts, counts + vertex_count, 0u);
  for (size_t i = 0; i < size_t(tet_count) * 4; ++i) {
    if (tet_indices[i] >= vertex_count)
      return jgs2::fail(JGS2_ERROR_RANGE, "tetIndices[%zu] references vertex %u, but vertexCount is %u.", i,
                        tet_indices[i], vertex_count);
    counts[tet_indices[i]] += 1;
  }
  uint32_t total = 0;
  for (uint32_t v = 0; v < vertex_count; ++v) {
    starts[v] = total;
    total += counts[v];
  }
  std::vector<uint32_t> cursor(starts, starts + vertex_count);
  for (uint32_t t = 0; t < tet_count; ++t)
    for (int l = 0; l < 4; ++l) tets[cursor[tet_indices[4 * size_t(t) + l]]++] = t;
  jgs2::clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_build_element_incidence(uint32_t vertex_count, const uint32_t* element_indices, uint32_t width,
                                         uint32_t element_count, uint32_t* rows, uint32_t* elements) {
  if (!rows || (element_count && (!element_indices || !elements)))
    return jgs2::fail(JGS2_ERROR_TYPE, "NULL incidence argument.");
  if (width != 3 && width != 4) return jgs2::fail(JGS2_ERROR_RANGE, "Element width must be 3 or 4; got %u.", width);
  const size_t n = size_t(element_count) * width;
  std::fill(rows, rows + size_t(vertex_count) + 1, 0u);
  for (size_t i = 0; i < n; ++i) {
    if (element_indices[i] >= vertex_count)
      return jgs2::fail(JGS2_ERROR_RANGE, "Element index %u is outside vertexCount %u.", element_indices[i],
                        vertex_count);
    rows[element_indices[i] + 1] += 1;
  }
  for (uint32_t v = 0; v < vertex_count; ++v) rows[v + 1] += rows[v];
  std::vector<uint32_t> cursor(rows, rows + vertex_count);
  for (uint32_t e = 0; e < element_count; ++e)
    for (uint32_t l = 0; l < width; ++l) elements[cursor[element_indices[size_t(width) * e + l]]++] = e;
  jgs2::clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_greedy_vertex_coloring(uint32_t vertex_count, const uint32_t* group_offsets, uint32_t group_count,
                                        const uint32_t* group_members, uint32_t* colors, uint32_t* color_count) {
  if (!group_offsets || !colors || !color_count || (group_count && !group_members))
    return jgs2::fail(JGS2_ERROR_TYPE, "NULL coloring argument.");
  // Vertex -> group incidence, so each vertex visits its conflict partners
  // without materializing the pair graph (src/simulation/cpu/vertex-coloring.ts:20-110
  // builds explicit neighbor sets; the greedy result depends only on the
  // graph, which is identical).
  std::vector<uint32_t> incident_count(vertex_count + 1, 0);
  for (uint32_t g = 0; g < group_count; ++g) {
    for (uint32_t i = group_offsets[g]; i < group_offsets[g + 1]; ++i) {
      uint32_t v = group_members[i];
      if (v >= vertex_count)
        return jgs2::fail(JGS2_ERROR_RANGE, "conflictGroups[%u][%u] must be a valid vertex id", g,
                          i - group_offsets[g]);
      incident_count[v + 1] += 1;
    }
  }
  for (uint32_t v = 0; v < vertex_count; ++v) incident_count[v + 1] += incident_count[v];
  std::vector<uint32_t> incident(incident_count[vertex_count]);
  std::vector<uint32_t> cursor(incident_count.begin(), incident_count.end() - 1);
  for (uint32_t g = 0; g < group_count; ++g)
    for (uint32_t i = group_offsets[g]; i < group_offsets[g + 1]; ++i) incident[cursor[group_members[i]]++] = g;
  const uint32_t none = 0xffffffffu;
  std::fill(colors, colors + vertex_count, none);
  std::vector<uint32_t> stamp_of(size_t(vertex_count) + 1, 0);
  uint32_t stamp = 0, count = 0;
  for (uint32_t v = 0; v < vertex_count; ++v) {
    ++stamp;
    for (uint32_t k = incident_count[v]; k < incident_count[v + 1]; ++k) {
      uint32_t g = incident[k];
      for (uint32_t i = group_offsets[g]; i < group_offsets[g + 1]; ++i) {
        uint32_t u = group_members[i];
        if (u != v && colors[u] != none) stamp_of[colors[u]] = stamp;
      }
    }
    uint32_t c = 0;
    while (c < count && stamp_of[c] == stamp) ++c;
    colors[v] = c;
    count = std::max(count, c + 1);
  }
  *color_count = count;
  jgs2::clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_scene_build_cuboid(const jgs2_cuboid_desc* d, jgs2_scene** out) {
  if (!d || !out) return jgs2::fail(JGS2_ERROR_TYPE, "NULL cuboid description.");
  *out = nullptr;
  for (int a = 0; a < 3; ++a) {
    if (d->cells[a] < 1) return jgs2::fail(JGS2_ERROR_RANGE, "cells[%d] must be a positive integer.", a);
    if (!(d->size[a] > 0) || !std::isfinite(d->size[a]))
      return jgs2::fail(JGS2_ERROR_RANGE, "Cuboid dimensions must be finite and positive.");
  }
  if (!(d->density > 0) || !(d->young_modulus[0] > 0) || !(d->young_modulus[1] > 0 || d->layer_cells == 0))
    return jgs2::fail(JGS2_ERROR_RANGE, "Material density and Young's modulus must be positive.");
  if (!(d->poisson_ratio >= 0 && d->poisson_ratio < 0.5))
    return jgs2::fail(JGS2_ERROR_RANGE, "Stable Neo-Hookean material requires Poisson ratio in [0, 0.5).");
  if (d->material_model != JGS2_MATERIAL_COROTATED_LINEAR && d->material_model != JGS2_MATERIAL_STABLE_NEO_HOOKEAN)
    return jgs2::fail(JGS2_ERROR_RANGE, "Unknown material model %d.", d->material_model);
  if (!(d->jitter_fraction >= 0 && d->jitter_fraction < 0.25))
    return jgs2::fail(JGS2_ERROR_RANGE, "jitter_fraction must lie in [0, 0.25).");
  const uint64_t nx = d->cells[0], ny = d->cells[1], nz = d->cells[2];
  const uint64_t V64 = (nx + 1) * (ny + 1) * (nz + 1);
  const uint64_t T64 = nx * ny * nz * 6;
  if (V64 >= 0xffffffffull || T64 * 4 >= 0xffffffffull)
    return jgs2::fail(JGS2_ERROR_RANGE, "Cuboid is too large for 32-bit indices.");
  jgs2::Mesh m;
  jgs2::cuboid_grid(d->cells, d->origin, d->size, 0, 0, false, m);
  const uint32_t V = m.vertex_count(), T = m.tet_count();
  // Pinned face and seeded vertex jitter (before orientation).
  const uint64_t sy = nx + 1, sz = (nx + 1) * (ny + 1);
  double h[3] = {d->size[0] / double(nx), d->size[1] / double(ny), d->size[2] / double(nz)};
  for (uint32_t v = 0; v < V; ++v) {
    const uint64_t g[3] = {v % sy, (v % sz) / sy, v / sz};
    if (d->pin_axis < 3 && g[d->pin_axis] == 0) m.fixed[v] = 1;
  }
  if (d->jitter_fraction > 0) {
    CorpusRandom rng(d->seed);
    for (uint32_t v = 0; v < V; ++v)
      for (int a = 0; a < 3; ++a) m.x[3 * size_t(v) + a] += (2.0 * rng.next() - 1.0) * d->jitter_fraction * h[a];
  }
  // Alternating material layers by cell index along `layer_axis`.
  if (d->layer_cells) {
    for (uint32_t t = 0; t < T; ++t) {
      const uint64_t cell = t / 6;
      const uint64_t g[3] = {cell % nx, (cell / nx) % ny, cell / (nx * ny)};
      m.material[t] = uint32_t((g[d->layer_axis % 3] / d->layer_cells) & 1u);
    }
  }
  jgs2_status st = jgs2::orient(m);
  if (st != JGS2_OK) return st;
  std::vector<jgs2::Material> mats = {
      {d->material_model, d->density, d->young_modulus[0], d->poisson_ratio},
      {d->material_model, d->density, d->layer_cells ? d->young_modulus[1] : d->young_modulus[0], d->poisson_ratio}};
  jgs2_scene* s = new (std::nothrow) jgs2_scene();
  if (!s) return jgs2::fail(JGS2_ERROR_MEMORY, "Out of host memory.");
  st = jgs2::pack_scene(m, mats, d->cubature_k, d->pack_rest_stiffness != 0, std::cbrt(h[0] * h[1] * h[2]), s);
  if (st != JGS2_OK) {
    delete s;
    return st;
  }
  *out = s;
  jgs2::clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_scene_build_mesh(const jgs2_mesh_desc* d, jgs2_scene** out) {
  if (!d || !out || !d->positions || !d->tetrahedra || !d->materials)
    return jgs2::fail(JGS2_ERROR_TYPE, "NULL mesh description or array.");
  *out = nullptr;
  const uint32_t V = d->vertex_count, T = d->tet_count;
  if (V < 4 || T < 1) return jgs2::fail(JGS2_ERROR_RANGE, "A mesh needs at least one tetrahedron.");
  if (d->material_count < 1) return jgs2::fail(JGS2_ERROR_RANGE, "A mesh needs at least one material.");
  jgs2::Mesh m;
  m.x.assign(d->positions, d->positions + 3 * size_t(V));
  for (double c : m.x)
    if (!std::isfinite(c)) return jgs2::fail(JGS2_ERROR_RANGE, "Mesh positions must be finite.");
  m.tets.assign(d->tetrahedra, d->tetrahedra + 4 * size_t(T));
  for (uint32_t t = 0; t < T; ++t)
    for (int a = 0; a < 4; ++a) {
      const uint32_t i = m.tets[4 * size_t(t) + a];
      if (i >= V) return jgs2::fail(JGS2_ERROR_RANGE, "Tetrahedron %u references vertex %u of %u.", t, i, V);
      for (int b = 0; b < a; ++b)
        if (m.tets[4 * size_t(t) + b] == i)
          return jgs2::fail(JGS2_ERROR_RANGE, "Tetrahedron %u repeats vertex %u.", t, i);
    }
  m.material.assign(T, 0u);
  if (d->material_ids) m.material.assign(d->material_ids, d->material_ids + T);
  m.fixed.assign(V, 0u);
  if (d->fixed)
    for (uint32_t v = 0; v < V; ++v) m.fixed[v] = d->fixed[v] ? 1u : 0u;
  m.body.assign(V, 0u);
  if (d->body_ids) m.body.assign(d->body_ids, d->body_ids + V);
  std::vector<jgs2::Material> mats;
  for (uint32_t i = 0; i < d->material_count; ++i) {
    const jgs2_precompute_material& q = d->materials[i];
    if (!(q.density > 0) || !(q.young_modulus > 0))
      return jgs2::fail(JGS2_ERROR_RANGE, "Material density and Young's modulus must be positive.");
    if (!(q.poisson_ratio >= 0 && q.poisson_ratio < 0.5))
      return jgs2::fail(JGS2_ERROR_RANGE, "Stable Neo-Hookean material requires Poisson ratio in [0, 0.5).");
    if (q.model != JGS2_MATERIAL_COROTATED_LINEAR && q.model != JGS2_MATERIAL_STABLE_NEO_HOOKEAN)
      return jgs2::fail(JGS2_ERROR_RANGE, "Unknown material model %u.", q.model);
    mats.push_back({int32_t(q.model), q.density, q.young_modulus, q.poisson_ratio});
  }
  double h = d->cubature_length;
  if (!(h > 0)) {
    double volume = 0;
    for (uint32_t t = 0; t < T; ++t) {
      double r[9];
      rest_matrix(m.x, &m.tets[4 * size_t(t)], r);
      volume += std::fabs(determinant3(r)) / 6.0;
    }
    h = std::cbrt(6.0 * volume / double(T));
  }
  jgs2_scene* s = new (std::nothrow) jgs2_scene();
  if (!s) return jgs2::fail(JGS2_ERROR_MEMORY, "Out of host memory.");
  jgs2_status st = jgs2::pack_scene(m, mats, d->cubature_k, d->pack_rest_stiffness != 0, h, s);
  if (st != JGS2_OK) {
    delete s;
    return st;
  }
  *out = s;
  jgs2::clear_error();
  return JGS2_OK;
}

}  // extern "C"

namespace {

// precomputeStVKTriangleRest (src/simulation/cpu/cloth.ts:115-140).
bool triangle_rest(const double* x0, const double* x1, const double* x2, double inv[4], double* area) {
  double e01[3], e02[3];
  for (int a = 0; a < 3; ++a) {
    e01[a] = x1[a] - x0[a];
    e02[a] = x2[a] - x0[a];
  }
  double c[3] = {e01[1] * e02[2] - e01[2] * e02[1], e01[2] * e02[0] - e01[0] * e02[2], e01[0] * e02[1] - e01[1] * e02[0]};
  double twice = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  double len = std::sqrt(e01[0] * e01[0] + e01[1] * e01[1] + e01[2] * e01[2]);
  if (!(twice > 1e-12) || !(len > 0)) return false;
  double t[3], n[3];
  for (int a = 0; a < 3; ++a) {
    t[a] = e01[a] / len;
    n[a] = c[a] / twice;
  }
  double b[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
  double mx = e02[0] * t[0] + e02[1] * t[1] + e02[2] * t[2];
  double my = e02[0] * b[0] + e02[1] * b[1] + e02[2] * b[2];
  inv[0] = 1.0 / len;
  inv[1] = -mx / (len * my);
  inv[2] = 0.0;
  inv[3] = 1.0 / my;
  *area = 0.5 * twice;
  return true;
}

// Signed dihedral angle atan2(t.(n0 x n1), n0.n1) of the oriented triangles
// (0,1,2) and (1,0,3) (src/simulation/cpu/cloth.ts:258-340), with edge length.
bool hinge_rest(const double* x0, const double* x1, const double* x2, const double* x3, double* angle, double* edge_len) {
  double e[3], to2[3], to3[3];
  for (int a = 0; a < 3; ++a) {
    e[a] = x1[a] - x0[a];
    to2[a] = x2[a] - x0[a];
    to3[a] = x3[a] - x1[a];
  }
  auto cross = [](const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
  };
  auto norm = [](const double* u) { return std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]); };
  double rev[3] = {-e[0], -e[1], -e[2]}, n0[3], n1[3], nc[3];
  cross(e, to2, n0);
  cross(rev, to3, n1);
  double l = norm(e), l0 = norm(n0), l1 = norm(n1);
  if (!(l > 0) || !(l0 > 0) || !(l1 > 0)) return false;
  for (int a = 0; a < 3; ++a) {
    n0[a] /= l0;
    n1[a] /= l1;
  }
  cross(n0, n1, nc);
  double sine = (e[0] * nc[0] + e[1] * nc[1] + e[2] * nc[2]) / l;
  double cosine = n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2];
  *angle = std::atan2(sine, cosine);
  *edge_len = l;
  return true;
}

}  // namespace

extern "C" jgs2_status jgs2_scene_build_cloth_drape(const jgs2_cloth_desc* d, jgs2_scene** out) {
  if (!d || !out) return jgs2::fail(JGS2_ERROR_TYPE, "NULL cloth description.");
  *out = nullptr;
  if (d->cells < 1 || d->cells > 4096) return jgs2::fail(JGS2_ERROR_RANGE, "cloth cells must lie in [1, 4096].");
  if (!(d->size[0] > 0) || !(d->size[1] > 0) || !(d->density > 0) || !(d->young_modulus > 0) || !(d->thickness > 0))
    return jgs2::fail(JGS2_ERROR_RANGE, "Cloth density, youngModulus and thickness must be finite and positive.");
  if (!(d->poisson_ratio >= 0 && d->poisson_ratio < 0.5))
    return jgs2::fail(JGS2_ERROR_RANGE, "Cloth poissonRatio must be in [0, 0.5).");
  if (!(d->bending_stiffness >= 0)) return jgs2::fail(JGS2_ERROR_RANGE, "Cloth bendingStiffness must be finite and nonnegative.");
  if (d->fold_column > d->cells || d->fold_column == 1)
    return jgs2::fail(JGS2_ERROR_RANGE, "Cloth fold_column must be 0 or in [2, cells].");
  // The collider: a fixed one-cell stable Neo-Hookean block
  // (CLOTH_COLLIDER_MATERIAL, src/scenes/index.ts:227-234).
  jgs2_cuboid_desc c{};
  c.cells[0] = c.cells[1] = c.cells[2] = 1;
  for (int a = 0; a < 3; ++a) {
    c.origin[a] = d->collider_origin[a];
    c.size[a] = d->collider_size[a];
  }
  c.pin_axis = 3;
  c.material_model = JGS2_MATERIAL_STABLE_NEO_HOOKEAN;
  c.density = 1000;
  c.poisson_ratio = 0.34;
  c.young_modulus[0] = c.young_modulus[1] = 110000;
  c.cubature_k = d->cubature_k;
  jgs2_scene* s = nullptr;
  jgs2_status st = jgs2_scene_build_cuboid(&c, &s);
  if (st != JGS2_OK) return st;
  const uint32_t n = d->cells, VC = s->vertex_count, side = n + 1;
  const uint32_t V = VC + side * side;
  s->vertex_count = V;
  // The collider is fixed: every block vertex pinned, no cubature records.
  for (uint32_t v = 0; v < VC; ++v) s->vertex_info[4 * size_t(v) + 2] = 1;
  std::fill(s->cubature_tet_ids.begin(), s->cubature_tet_ids.end(), 0xffffffffu);
  std::fill(s->cubature_weights.begin(), s->cubature_weights.end(), 0.0f);
  std::fill(s->cubature_basis.begin(), s->cubature_basis.end(), 0.0f);
  const size_t K = s->cubature_k;
  s->cubature_tet_ids.resize(size_t(V) * K, 0xffffffffu);
  s->cubature_weights.resize(size_t(V) * K, 0.0f);
  s->cubature_basis.resize(size_t(V) * K * 36, 0.0f);
  std::vector<double> x(3 * size_t(side) * side);
  for (uint32_t z = 0; z <= n; ++z)
    for (uint32_t xx = 0; xx <= n; ++xx) {
      const size_t i = size_t(xx) + size_t(z) * side;
      const bool folded = d->fold_column != 0 && xx >= d->fold_column;
      const double column = folded ? 2.0 * double(d->fold_column - 1) - double(xx) : double(xx);
      x[3 * i] = d->origin[0] + d->size[0] * column / double(n);
      x[3 * i + 1] = d->origin[1] + (folded ? d->fold_lift : 0.0);
      x[3 * i + 2] = d->origin[2] + d->size[1] * double(z) / double(n);
    }
  // Sheet vertices: pinned at (0, 0) and (n/2, 0) of the first row, body 1.
  const uint32_t adjacency_end = uint32_t(s->adjacency.size());
  for (uint32_t i = 0; i < side * side; ++i) {
    const uint32_t xx = i % side, z = i / side;
    const bool pinned = z == 0 && (xx == 0 || xx == n / 2);
    for (int a = 0; a < 3; ++a) {
      s->positions.push_back(float(x[3 * size_t(i) + a]));
      s->vertex_rest.push_back(float(x[3 * size_t(i) + a]));
    }
    s->positions.push_back(1.0f);
    s->vertex_rest.push_back(0.0f);  // mass added from triangles below
    s->vertex_info.push_back(adjacency_end);
    s->vertex_info.push_back(0);
    s->vertex_info.push_back(pinned ? 1u : 0u);
    s->vertex_info.push_back(1);
  }
  // Triangles and cloth masses (precompute.ts:46-93).
  std::vector<double> mass(side * side, 0.0);
  for (uint32_t z = 0; z < n; ++z)
    for (uint32_t xx = 0; xx < n; ++xx) {
      const uint32_t i00 = xx + z * side, i10 = xx + 1 + z * side, i01 = xx + (z + 1) * side, i11 = xx + 1 + (z + 1) * side;
      const uint32_t tris[6] = {i00, i01, i10, i10, i01, i11};
      for (int k = 0; k < 2; ++k) {
        const uint32_t a = tris[3 * k], b = tris[3 * k + 1], cc = tris[3 * k + 2];
        double inv[4], area;
        if (!triangle_rest(&x[3 * size_t(a)], &x[3 * size_t(b)], &x[3 * size_t(cc)], inv, &area)) {
          delete s;
          return jgs2::fail(JGS2_ERROR_RANGE, "Cloth triangle must be nondegenerate.");
        }
        s->cloth_triangles.push_back(VC + a);
        s->cloth_triangles.push_back(VC + b);
        s->cloth_triangles.push_back(VC + cc);
        for (int q = 0; q < 4; ++q) s->cloth_inverse.push_back(float(inv[q]));
        s->cloth_areas.push_back(float(area));
        const double vm = d->density * d->thickness * area / 3.0;
        mass[a] += vm;
        mass[b] += vm;
        mass[cc] += vm;
      }
    }
  for (uint32_t i = 0; i < side * side; ++i) s->vertex_rest[4 * size_t(VC + i) + 3] = float(mass[i]);
  // Hinges from oriented half-edges in triangle order (scenes/index.ts:894-946).
  struct Half {
    uint32_t first, second, opposite;
  };
  std::vector<std::pair<uint64_t, Half>> open;
  std::vector<uint64_t> closed;
  const size_t NT = s->cloth_areas.size();
  std::unordered_map<uint64_t, Half> open_map;
  std::unordered_set<uint64_t> closed_set;
  for (size_t t = 0; t < NT; ++t) {
    const uint32_t* ix = &s->cloth_triangles[3 * t];
    const uint32_t hs[3][3] = {{ix[0], ix[1], ix[2]}, {ix[1], ix[2], ix[0]}, {ix[2], ix[0], ix[1]}};
    for (auto& h : hs) {
      const uint64_t key = (uint64_t(std::min(h[0], h[1])) << 32) | std::max(h[0], h[1]);
      if (closed_set.count(key)) {
        delete s;
        return jgs2::fail(JGS2_ERROR_RANGE, "Cloth edge is non-manifold.");
      }
      auto it = open_map.find(key);
      if (it == open_map.end()) {
        open_map[key] = Half{h[0], h[1], h[2]};
        continue;
      }
      const Half prev = it->second;
      if (prev.first != h[1] || prev.second != h[0]) {
        delete s;
        return jgs2::fail(JGS2_ERROR_RANGE, "Cloth triangles sharing an edge must have consistent orientation.");
      }
      s->cloth_hinges.push_back(prev.first);
      s->cloth_hinges.push_back(prev.second);
      s->cloth_hinges.push_back(prev.opposite);
      s->cloth_hinges.push_back(h[2]);
      open_map.erase(it);
      closed_set.insert(key);
    }
  }
  const size_t NH = s->cloth_hinges.size() / 4;
  for (size_t h = 0; h < NH; ++h) {
    const uint32_t* ix = &s->cloth_hinges[4 * h];
    double angle = 0, len = 0;
    if (!hinge_rest(&x[3 * size_t(ix[0] - VC)], &x[3 * size_t(ix[1] - VC)], &x[3 * size_t(ix[2] - VC)],
                    &x[3 * size_t(ix[3] - VC)], &angle, &len)) {
      delete s;
      return jgs2::fail(JGS2_ERROR_RANGE, "Cloth hinge %zu is degenerate.", h);
    }
    s->hinge_angles.push_back(float(angle));
    s->hinge_lengths.push_back(float(len));
  }
  // Contact surface (surfaceWithCloth, src/simulation/cpu/precompute.ts:
  // 98-133): the collider's boundary faces and edges, then every cloth
  // triangle and its edges not listed yet.
  {
    jgs2::Mesh collider;
    collider.x.resize(3 * size_t(VC));
    for (uint32_t v = 0; v < VC; ++v)
      for (int a = 0; a < 3; ++a) collider.x[3 * size_t(v) + a] = s->vertex_rest[4 * size_t(v) + a];
    collider.tets = s->tet_indices;
    collider.material.assign(s->tet_count, 0u);
    collider.fixed.assign(VC, 1u);
    collider.body.assign(VC, 0u);
    st = jgs2::boundary_surface(collider, s->surface_triangles, s->surface_edges);
    if (st != JGS2_OK) {
      delete s;
      return st;
    }
    std::unordered_set<uint64_t> listed;
    for (size_t e = 0; e + 1 < s->surface_edges.size(); e += 2)
      listed.insert((uint64_t(s->surface_edges[e]) << 32) | s->surface_edges[e + 1]);
    for (size_t t = 0; t < NT; ++t) {
      const uint32_t* ix = &s->cloth_triangles[3 * t];
      s->surface_triangles.insert(s->surface_triangles.end(), ix, ix + 3);
      for (int e = 0; e < 3; ++e) {
        const uint32_t a = std::min(ix[e], ix[(e + 1) % 3]), b = std::max(ix[e], ix[(e + 1) % 3]);
        if (listed.insert((uint64_t(a) << 32) | b).second) {
          s->surface_edges.push_back(a);
          s->surface_edges.push_back(b);
        }
      }
    }
  }
  s->has_cloth = true;
  jgs2_cloth_input& ci = s->cloth;
  ci.triangle_count = uint32_t(NT);
  ci.triangle_indices = s->cloth_triangles.data();
  ci.triangle_inverse_rest_bases = s->cloth_inverse.data();
  ci.triangle_rest_areas = s->cloth_areas.data();
  ci.hinge_count = uint32_t(NH);
  ci.hinge_indices = s->cloth_hinges.data();
  ci.hinge_rest_angles = s->hinge_angles.data();
  ci.hinge_rest_edge_lengths = s->hinge_lengths.data();
  ci.young_modulus = d->young_modulus;
  ci.poisson_ratio = d->poisson_ratio;
  ci.thickness = d->thickness;
  ci.bending_stiffness = d->bending_stiffness;
  *out = s;
  jgs2::clear_error();
  return JGS2_OK;
}

extern "C" {

jgs2_status jgs2_scene_get_input(const jgs2_scene* s, jgs2_scene_input* out) {
  if (!s || !out) return jgs2::fail(JGS2_ERROR_TYPE, "NULL scene.");
  std::memset(out, 0, sizeof(*out));
  out->vertex_count = s->vertex_count;
  out->tet_count = s->tet_count;
  out->cubature_k = s->cubature_k;
  out->positions = s->positions.data();
  out->vertex_rest = s->vertex_rest.data();
  out->vertex_info = s->vertex_info.data();
  out->tet_indices = s->tet_indices.data();
  out->tet_inverse_dm = s->tet_inverse_dm.data();
  out->tet_meta = s->tet_meta.data();
  out->tet_rest_stiffness = s->tet_rest_stiffness.empty() ? nullptr : s->tet_rest_stiffness.data();
  out->adjacency = s->adjacency.data();
  out->adjacency_length = s->adjacency.size();
  out->cubature_tet_ids = s->cubature_tet_ids.data();
  out->cubature_weights = s->cubature_weights.data();
  out->cubature_basis = s->cubature_basis.data();
  out->external_forces = s->forces.empty() ? nullptr : s->forces.data();
  out->target_positions = s->targets.empty() ? nullptr : s->targets.data();
  out->target_stiffnesses = s->stiffnesses.empty() ? nullptr : s->stiffnesses.data();
  out->cloth = s->has_cloth ? &s->cloth : nullptr;
  out->contact = s->has_contact ? &s->contact : nullptr;
  jgs2::clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_scene_set_objectives(jgs2_scene* s, const float* forces, const float* targets,
                                      const float* stiffnesses) {
  if (!s) return jgs2::fail(JGS2_ERROR_TYPE, "NULL scene.");
  if ((targets == nullptr) != (stiffnesses == nullptr))
    return jgs2::fail(JGS2_ERROR_RANGE,
                      "objectives.targetPositions and objectives.targetStiffnesses must be provided together.");
  const size_t V = s->vertex_count;
  if (forces) s->forces.assign(forces, forces + 3 * V); else s->forces.clear();
  if (targets) {
    s->targets.assign(targets, targets + 3 * V);
    s->stiffnesses.assign(stiffnesses, stiffnesses + V);
  } else {
    s->targets.clear();
    s->stiffnesses.clear();
  }
  jgs2::clear_error();
  return JGS2_OK;
}

void jgs2_scene_destroy(jgs2_scene* s) { delete s; }

}  // extern "C"
