// Host contact setup for libjgs2hip: boundary surfaces, the static IPC
// candidate superset and the reference's IPC trough scene. Pure host code.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <thread>
#include <vector>

#include "jgs2_scene_internal.h"

namespace jgs2 {

namespace {

struct FaceRecord {
  uint32_t count;
  std::array<uint32_t, 3> face;
};

// outwardFace (mesh.ts:338-364): the face winding whose normal points away
// from the opposite vertex.
std::array<uint32_t, 3> outward_face(const std::vector<double>& x, std::array<uint32_t, 3> f, uint32_t opposite) {
  const double* a = &x[3 * size_t(f[0])];
  const double* b = &x[3 * size_t(f[1])];
  const double* c = &x[3 * size_t(f[2])];
  const double* o = &x[3 * size_t(opposite)];
  const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  const double toward = n[0] * (o[0] - a[0]) + n[1] * (o[1] - a[1]) + n[2] * (o[2] - a[2]);
  return toward > 0 ? std::array<uint32_t, 3>{f[0], f[2], f[1]} : f;
}

}  // namespace

jgs2_status boundary_surface(const Mesh& m, std::vector<uint32_t>& triangles, std::vector<uint32_t>& edges) {
  std::map<std::array<uint32_t, 3>, FaceRecord> faces;
  for (uint32_t t = 0; t < m.tet_count(); ++t) {
    const uint32_t* ix = &m.tets[4 * size_t(t)];
    const uint32_t a = ix[0], b = ix[1], c = ix[2], d = ix[3];
    const uint32_t local[4][4] = {{b, c, d, a}, {a, d, c, b}, {a, b, d, c}, {a, c, b, d}};
    for (const auto& f : local) {
      std::array<uint32_t, 3> key{f[0], f[1], f[2]};
      std::sort(key.begin(), key.end());
      auto it = faces.find(key);
      if (it != faces.end()) {
        if (++it->second.count > 2)
          return fail(JGS2_ERROR_RANGE, "Non-manifold face %u,%u,%u belongs to over two tets.", key[0], key[1], key[2]);
      } else {
        faces.emplace(key, FaceRecord{1u, outward_face(m.x, {f[0], f[1], f[2]}, f[3])});
      }
    }
  }
  // Boundary faces in ascending key order, then their unique sorted edges.
  triangles.clear();
  std::vector<std::array<uint32_t, 2>> pairs;
  for (const auto& kv : faces) {
    if (kv.second.count != 1) continue;
    const auto& f = kv.second.face;
    triangles.insert(triangles.end(), f.begin(), f.end());
    for (int e = 0; e < 3; ++e) {
      const uint32_t u = f[e], v = f[(e + 1) % 3];
      pairs.push_back({std::min(u, v), std::max(u, v)});
    }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  edges.clear();
  for (const auto& p : pairs) edges.insert(edges.end(), p.begin(), p.end());
  return JGS2_OK;
}

namespace {

// minimumStaticIpcContactDistance (ipc-contact.ts:349-391) in f64 over the
// caller's packed positions.
double clamp01(double v) {
  if (std::isnan(v) || v <= 0) return 0;
  return v >= 1 ? 1 : v;
}
double dist3(const double* a, const double* b) {
  const double d = std::hypot(a[0] - b[0], a[1] - b[1], a[2] - b[2]);
  return std::isfinite(d) ? d : 1.7976931348623157e308;
}
void lerp3(const double* s, const double* e, double t, double* out) {
  for (int a = 0; a < 3; ++a) out[a] = s[a] + t * (e[a] - s[a]);
  if (!(std::isfinite(out[0]) && std::isfinite(out[1]) && std::isfinite(out[2])))
    std::memcpy(out, t <= 0.5 ? s : e, 3 * sizeof(double));
}
double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double segment_distance(const double* q, const double* s, const double* e) {  // ipc-contact.ts:130-150
  const double edge[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
  const double els = dot3(edge, edge);
  double t = 0;
  if (std::isfinite(els) && els > 1e-20) {
    const double qs[3] = {q[0] - s[0], q[1] - s[1], q[2] - s[2]};
    t = clamp01(dot3(qs, edge) / els);
  }
  double p[3];
  lerp3(s, e, t, p);
  return dist3(q, p);
}
double vertex_triangle_distance(const double* p, const double* t0, const double* t1, const double* t2) {  // :194-275
  const double e01[3] = {t1[0] - t0[0], t1[1] - t0[1], t1[2] - t0[2]};
  const double e02[3] = {t2[0] - t0[0], t2[1] - t0[1], t2[2] - t0[2]};
  const double e12[3] = {t2[0] - t1[0], t2[1] - t1[1], t2[2] - t1[2]};
  const double s01 = dot3(e01, e01), s02 = dot3(e02, e02), s12 = dot3(e12, e12);
  const double largest = std::max(s01, std::max(s02, s12));
  const double an[3] = {e01[1] * e02[2] - e01[2] * e02[1], e01[2] * e02[0] - e01[0] * e02[2],
                        e01[0] * e02[1] - e01[1] * e02[0]};
  const double as = dot3(an, an);
  const double thr = std::max(1e-20 * 1e-20, 1e-6 * largest * largest);
  if (std::isfinite(as) && as > thr) {
    const double d01 = dot3(e01, e02);
    const double po[3] = {p[0] - t0[0], p[1] - t0[1], p[2] - t0[2]};
    const double d20 = dot3(po, e01), d21 = dot3(po, e02);
    const double den = s01 * s02 - d01 * d01;
    const double b1 = (s02 * d20 - d01 * d21) / den;
    const double b2 = (s01 * d21 - d01 * d20) / den;
    const double b0 = 1 - b1 - b2;
    if (std::isfinite(b0) && std::isfinite(b1) && std::isfinite(b2) && b0 >= 0 && b1 >= 0 && b2 >= 0) {
      const double c[3] = {b0 * t0[0] + b1 * t1[0] + b2 * t2[0], b0 * t0[1] + b1 * t1[1] + b2 * t2[1],
                           b0 * t0[2] + b1 * t1[2] + b2 * t2[2]};
      if (std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2])) return dist3(p, c);
    }
  }
  double best = segment_distance(p, t0, t1);
  best = std::min(best, segment_distance(p, t1, t2));
  best = std::min(best, segment_distance(p, t2, t0));
  return best;
}
double edge_edge_distance(const double* a0, const double* a1, const double* b0, const double* b1) {  // :277-347
  const double da[3] = {a1[0] - a0[0], a1[1] - a0[1], a1[2] - a0[2]};
  const double db[3] = {b1[0] - b0[0], b1[1] - b0[1], b1[2] - b0[2]};
  const double off[3] = {a0[0] - b0[0], a0[1] - b0[1], a0[2] - b0[2]};
  const double la = dot3(da, da), lb = dot3(db, db);
  const double thr = std::max(1e-20, 1e-6 * std::max(la, lb));
  const bool adeg = !(la > thr), bdeg = !(lb > thr);
  double pa = 0, pb = 0;
  if (adeg && !bdeg) {
    pb = clamp01(dot3(db, off) / lb);
  } else if (!adeg) {
    const double proj_a = dot3(da, off);
    if (bdeg) {
      pa = clamp01(-proj_a / la);
    } else {
      const double coupling = dot3(da, db), proj_b = dot3(db, off);
      const double den = la * lb - coupling * coupling;
      if (den > 1e-6 * la * lb) pa = clamp01((coupling * proj_b - proj_a * lb) / den);
      const double pbu = (coupling * pa + proj_b) / lb;
      if (pbu < 0) {
        pb = 0;
        pa = clamp01(-proj_a / la);
      } else if (pbu > 1) {
        pb = 1;
        pa = clamp01((coupling - proj_a) / la);
      } else {
        pb = clamp01(pbu);
      }
    }
  }
  double pa3[3], pb3[3];
  lerp3(a0, a1, pa, pa3);
  lerp3(b0, b1, pb, pb3);
  return dist3(pa3, pb3);
}

}  // namespace

double min_static_contact_distance(const float* positions4, const jgs2_contact_input& c) {
  double minimum = 1.7976931348623157e308;
  const uint32_t n = c.vertex_triangle_count + c.edge_edge_count;
  for (uint32_t k = 0; k < n; ++k) {
    double p[4][3];
    for (int l = 0; l < 4; ++l)
      for (int a = 0; a < 3; ++a) p[l][a] = positions4[4 * size_t(c.packed_indices[4 * size_t(k) + l]) + a];
    const double d = k < c.vertex_triangle_count ? vertex_triangle_distance(p[0], p[1], p[2], p[3])
                                                 : edge_edge_distance(p[0], p[1], p[2], p[3]);
    minimum = std::min(minimum, d);
  }
  return minimum;
}

jgs2_status canonical_surface(uint32_t V, const uint32_t* triangles, uint32_t triangle_count, const uint32_t* edges,
                              uint32_t edge_count, Surface& out) {
  if ((triangle_count && !triangles) || (edge_count && !edges))
    return fail(JGS2_ERROR_TYPE, "NULL contact-surface array.");
  out = Surface{};
  for (uint32_t t = 0; t < triangle_count; ++t) {  // canonicalTriangles (ipc-contact.ts:500-527)
    std::array<uint32_t, 3> f{triangles[3 * size_t(t)], triangles[3 * size_t(t) + 1], triangles[3 * size_t(t) + 2]};
    std::sort(f.begin(), f.end());
    for (uint32_t v : f)
      if (v >= V) return fail(JGS2_ERROR_RANGE, "Surface vertex index %u is outside [0, %u).", v, V);
    if (f[0] == f[1] || f[1] == f[2])
      return fail(JGS2_ERROR_RANGE, "Surface triangles must have three distinct vertices.");
    out.tris.push_back(f);
  }
  std::sort(out.tris.begin(), out.tris.end());
  out.tris.erase(std::unique(out.tris.begin(), out.tris.end()), out.tris.end());
  for (uint32_t e = 0; e < edge_count; ++e) {  // canonicalEdges (:546-569)
    const uint32_t a = edges[2 * size_t(e)], b = edges[2 * size_t(e) + 1];
    if (a >= V || b >= V) return fail(JGS2_ERROR_RANGE, "Surface vertex index %u is outside [0, %u).", std::max(a, b), V);
    if (a == b) return fail(JGS2_ERROR_RANGE, "Surface edges must have two distinct vertices.");
    out.edges.push_back({std::min(a, b), std::max(a, b)});
  }
  for (const auto& f : out.tris) {
    out.edges.push_back({f[0], f[1]});
    out.edges.push_back({f[1], f[2]});
    out.edges.push_back({f[0], f[2]});
  }
  std::sort(out.edges.begin(), out.edges.end());
  out.edges.erase(std::unique(out.edges.begin(), out.edges.end()), out.edges.end());
  out.ring_start.assign(size_t(V) + 1, 0);  // buildOneRing (:571-584)
  for (const auto& e : out.edges) {
    out.ring_start[e[0] + 1] += 1;
    out.ring_start[e[1] + 1] += 1;
  }
  for (uint32_t v = 0; v < V; ++v) out.ring_start[v + 1] += out.ring_start[v];
  out.ring.assign(out.ring_start[V], 0);
  std::vector<uint32_t> cursor(out.ring_start.begin(), out.ring_start.end() - 1);
  for (const auto& e : out.edges) {
    out.ring[cursor[e[0]]++] = e[1];
    out.ring[cursor[e[1]]++] = e[0];
  }
  for (uint32_t v = 0; v < V; ++v) std::sort(out.ring.begin() + out.ring_start[v], out.ring.begin() + out.ring_start[v + 1]);
  return JGS2_OK;
}

namespace {

bool in_closed_ring(const Surface& s, uint32_t a, uint32_t b) {  // isInClosedOneRing (:586-592)
  return a == b || std::binary_search(s.ring.begin() + s.ring_start[a], s.ring.begin() + s.ring_start[a + 1], b);
}

struct Box {
  float lo[3], hi[3];
  bool wild;
};

float box_lower_bound(const Box& a, const Box& b) {  // ipcAabbDistanceLowerBound (S:525-550)
  float g[3];
  for (int k = 0; k < 3; ++k) g[k] = std::fmax(std::fmax(b.lo[k] - a.hi[k], a.lo[k] - b.hi[k]), 0.0f);
  const float xx = g[0] * g[0], yy = g[1] * g[1], zz = g[2] * g[2];
  return std::sqrt((xx + yy) + zz);
}

Box box_of(const float* positions4, std::initializer_list<uint32_t> ids) {
  Box b{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}, false};
  for (uint32_t v : ids)
    for (int k = 0; k < 3; ++k) {
      const float c = positions4[4 * size_t(v) + k];
      if (!std::isfinite(c)) b.wild = true;
      b.lo[k] = std::fmin(b.lo[k], c);
      b.hi[k] = std::fmax(b.hi[k], c);
    }
  return b;
}

// A uniform grid over the finite points' bounds: every item is listed in
// each cell its box, grown by `grow`, overlaps.
struct Grid {
  double origin[3] = {0, 0, 0}, cell = 1;
  int64_t dims[3] = {1, 1, 1};
  std::vector<uint64_t> keys;  // (cell << 32) | item, sorted
  void cells_of(const Box& b, double grow, int64_t lo[3], int64_t hi[3]) const {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::clamp<int64_t>(int64_t(std::floor((double(b.lo[k]) - grow - origin[k]) / cell)), 0, dims[k] - 1);
      hi[k] = std::clamp<int64_t>(int64_t(std::floor((double(b.hi[k]) + grow - origin[k]) / cell)), 0, dims[k] - 1);
    }
  }
  uint64_t index(int64_t x, int64_t y, int64_t z) const { return uint64_t((z * dims[1] + y) * dims[0] + x); }
  void insert(const std::vector<Box>& boxes, double grow) {
    keys.clear();
    for (size_t i = 0; i < boxes.size(); ++i) {
      if (boxes[i].wild) continue;
      int64_t lo[3], hi[3];
      cells_of(boxes[i], grow, lo, hi);
      for (int64_t z = lo[2]; z <= hi[2]; ++z)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
          for (int64_t x = lo[0]; x <= hi[0]; ++x) keys.push_back(index(x, y, z) << 32 | uint64_t(i));
    }
    std::sort(keys.begin(), keys.end());
  }
  template <class F> void visit(uint64_t c, F&& f) const {
    auto it = std::lower_bound(keys.begin(), keys.end(), c << 32);
    for (; it != keys.end() && (*it >> 32) == c; ++it) f(uint32_t(*it & 0xffffffffu));
  }
};

}  // namespace

bool superset_nonempty(const Surface& s, uint32_t V) {
  for (const auto& f : s.tris) {  // a vertex outside a triangle's closed rings
    uint64_t near = 3 + (s.ring_start[f[0] + 1] - s.ring_start[f[0]]) + (s.ring_start[f[1] + 1] - s.ring_start[f[1]]) +
                    (s.ring_start[f[2] + 1] - s.ring_start[f[2]]);
    if (near < V) return true;
    for (uint32_t v = 0; v < V; ++v)
      if (!in_closed_ring(s, v, f[0]) && !in_closed_ring(s, v, f[1]) && !in_closed_ring(s, v, f[2])) return true;
  }
  for (size_t i = 0; i < s.edges.size(); ++i)
    for (size_t j = i + 1; j < s.edges.size(); ++j) {
      const auto& a = s.edges[i];
      const auto& b = s.edges[j];
      if (!in_closed_ring(s, a[0], b[0]) && !in_closed_ring(s, a[0], b[1]) && !in_closed_ring(s, a[1], b[0]) &&
          !in_closed_ring(s, a[1], b[1]))
        return true;
    }
  return false;
}

SurfaceScan scan_surface_pairs(const Surface& s, uint32_t V, const float* positions4, float reach,
                               float distance_reach) {
  std::vector<Box> pts(V), tris(s.tris.size()), edges(s.edges.size());
  for (uint32_t v = 0; v < V; ++v) pts[v] = box_of(positions4, {v});
  for (size_t t = 0; t < s.tris.size(); ++t) tris[t] = box_of(positions4, {s.tris[t][0], s.tris[t][1], s.tris[t][2]});
  for (size_t e = 0; e < s.edges.size(); ++e) edges[e] = box_of(positions4, {s.edges[e][0], s.edges[e][1]});
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const Box& b : pts)
    if (!b.wild)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], double(b.lo[k]));
        hi[k] = std::max(hi[k], double(b.hi[k]));
      }
  const double grow = std::isfinite(reach) ? double(reach) * (1 + 1e-6) + 1e-12 : INFINITY;
  const bool everything = !std::isfinite(grow) || !(lo[0] <= hi[0]);
  Grid tri_grid, edge_grid;
  if (!everything) {
    const double extent = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-9});
    const double cell = std::max(grow, extent / 128.0);
    for (Grid* g : {&tri_grid, &edge_grid}) {
      g->cell = cell;
      for (int k = 0; k < 3; ++k) {
        g->origin[k] = lo[k];
        g->dims[k] = std::max<int64_t>(1, int64_t(std::ceil((hi[k] - lo[k]) / cell)) + 1);
      }
    }
    tri_grid.insert(tris, grow);
    edge_grid.insert(edges, grow);
  }
  // Primitives with a non-finite point are tested against everything.
  std::vector<uint32_t> wild_tris, wild_edges;
  for (uint32_t t = 0; t < tris.size(); ++t)
    if (tris[t].wild) wild_tris.push_back(t);
  for (uint32_t e = 0; e < edges.size(); ++e)
    if (edges[e].wild) wild_edges.push_back(e);
  // A pair counts when its AABB lower bound is within `reach` (or a point
  // is non-finite) and no vertex of one lies in the other's closed one-ring;
  // pairs within `distance_reach` also take the exact distance. The bound
  // is tested first: it is the cheaper test and rejects most grid pairs.
  auto consider = [&](SurfaceScan& out, bool vt, const Box& a, const Box& b, std::initializer_list<uint32_t> ids) {
    const bool wild = a.wild || b.wild;
    const float lb = box_lower_bound(a, b);
    if (!(wild || !(lb > reach))) return;
    const uint32_t* q = ids.begin();
    const bool excluded = vt ? (in_closed_ring(s, q[0], q[1]) || in_closed_ring(s, q[0], q[2]) ||
                                in_closed_ring(s, q[0], q[3]))
                             : (in_closed_ring(s, q[0], q[2]) || in_closed_ring(s, q[0], q[3]) ||
                                in_closed_ring(s, q[1], q[2]) || in_closed_ring(s, q[1], q[3]));
    if (excluded) return;
    (vt ? out.vertex_triangle : out.edge_edge) += 1;
    if (wild || lb > distance_reach) return;
    double p[4][3];
    int n = 0;
    for (uint32_t v : ids) {
      for (int k = 0; k < 3; ++k) p[n][k] = positions4[4 * size_t(v) + k];
      ++n;
    }
    const double d = vt ? vertex_triangle_distance(p[0], p[1], p[2], p[3]) : edge_edge_distance(p[0], p[1], p[2], p[3]);
    out.min_distance = std::min(out.min_distance, d);
  };
  // VT: each point against the triangles listed in its cell.
  auto scan_points = [&](uint32_t begin, uint32_t end, SurfaceScan& out) {
    for (uint32_t v = begin; v < end; ++v) {
      auto test = [&](uint32_t t) {
        const auto& f = s.tris[t];
        consider(out, true, pts[v], tris[t], {v, f[0], f[1], f[2]});
      };
      if (everything || pts[v].wild) {
        for (uint32_t t = 0; t < s.tris.size(); ++t) test(t);
        continue;
      }
      int64_t c0[3], c1[3];
      tri_grid.cells_of(pts[v], 0.0, c0, c1);
      tri_grid.visit(tri_grid.index(c0[0], c0[1], c0[2]), test);
      for (uint32_t t : wild_tris) test(t);
    }
  };
  // EE: each edge against the later edges listed in the cells it overlaps.
  auto scan_edges = [&](uint32_t begin, uint32_t end, SurfaceScan& out) {
    std::vector<uint32_t> stamp(s.edges.size(), 0xffffffffu);
    for (uint32_t i = begin; i < end; ++i) {
      const auto& a = s.edges[i];
      auto test = [&](uint32_t j) {
        if (j <= i || stamp[j] == i) return;
        stamp[j] = i;
        const auto& b = s.edges[j];
        consider(out, false, edges[i], edges[j], {a[0], a[1], b[0], b[1]});
      };
      if (everything || edges[i].wild) {
        for (uint32_t j = i + 1; j < s.edges.size(); ++j) test(j);
        continue;
      }
      int64_t c0[3], c1[3];
      edge_grid.cells_of(edges[i], 0.0, c0, c1);
      for (int64_t z = c0[2]; z <= c1[2]; ++z)
        for (int64_t y = c0[1]; y <= c1[1]; ++y)
          for (int64_t x = c0[0]; x <= c1[0]; ++x) edge_grid.visit(edge_grid.index(x, y, z), test);
      for (uint32_t j : wild_edges) test(j);
    }
  };
  // Chunks of points and edges over the host threads; the counts are sums
  // and the distance a minimum, so the result does not depend on the split.
  const size_t work = size_t(V) + s.edges.size();
  const unsigned threads =
      work < 65536 ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint32_t chunks = 8 * threads;
  std::vector<SurfaceScan> parts(2 * size_t(chunks));
  auto run_chunk = [&](uint32_t c) {
    const uint32_t k = c % chunks;
    if (c < chunks)
      scan_points(uint32_t(uint64_t(V) * k / chunks), uint32_t(uint64_t(V) * (k + 1) / chunks), parts[c]);
    else
      scan_edges(uint32_t(s.edges.size() * k / chunks), uint32_t(s.edges.size() * (k + 1) / chunks), parts[c]);
  };
  if (threads == 1) {
    for (uint32_t c = 0; c < 2 * chunks; ++c) run_chunk(c);
  } else {
    std::vector<std::thread> pool;
    std::atomic<uint32_t> next{0};
    for (unsigned t = 0; t < threads; ++t)
      pool.emplace_back([&] {
        for (uint32_t c = next++; c < 2 * chunks; c = next++) run_chunk(c);
      });
    for (auto& th : pool) th.join();
  }
  SurfaceScan out;
  for (const SurfaceScan& p : parts) {
    out.vertex_triangle += p.vertex_triangle;
    out.edge_edge += p.edge_edge;
    out.min_distance = std::min(out.min_distance, p.min_distance);
  }
  return out;
}

namespace {

// rotateEulerXyz (mesh.ts:210-235): x, then y, then z.
void rotate_euler_xyz(double* p, const double r[3]) {
  const double cx = std::cos(r[0]), sx = std::sin(r[0]), cy = std::cos(r[1]), sy = std::sin(r[1]);
  const double cz = std::cos(r[2]), sz = std::sin(r[2]);
  const double ax[3] = {p[0], cx * p[1] - sx * p[2], sx * p[1] + cx * p[2]};
  const double ay[3] = {cy * ax[0] + sy * ax[2], ax[1], -sy * ax[0] + cy * ax[2]};
  p[0] = cz * ay[0] - sz * ay[1];
  p[1] = sz * ay[0] + cz * ay[1];
  p[2] = ay[2];
}

}  // namespace

}  // namespace jgs2

extern "C" {

jgs2_status jgs2_build_static_contact_candidates(uint32_t vertex_count, const uint32_t* triangles,
                                                 uint32_t triangle_count, const uint32_t* edges, uint32_t edge_count,
                                                 uint32_t* packed, uint64_t capacity, uint32_t* vt_count,
                                                 uint32_t* ee_count) {
  if (!vt_count || !ee_count || (triangle_count && !triangles) || (edge_count && !edges))
    return jgs2::fail(JGS2_ERROR_TYPE, "NULL contact-candidate argument.");
  const uint32_t V = vertex_count;
  // canonicalTriangles / canonicalEdges (ipc-contact.ts:500-569).
  std::vector<std::array<uint32_t, 3>> tris;
  for (uint32_t t = 0; t < triangle_count; ++t) {
    std::array<uint32_t, 3> f{triangles[3 * size_t(t)], triangles[3 * size_t(t) + 1], triangles[3 * size_t(t) + 2]};
    std::sort(f.begin(), f.end());
    for (uint32_t v : f)
      if (v >= V) return jgs2::fail(JGS2_ERROR_RANGE, "Surface vertex index %u is outside [0, %u).", v, V);
    if (f[0] == f[1] || f[1] == f[2])
      return jgs2::fail(JGS2_ERROR_RANGE, "Surface triangles must have three distinct vertices.");
    tris.push_back(f);
  }
  std::sort(tris.begin(), tris.end());
  tris.erase(std::unique(tris.begin(), tris.end()), tris.end());
  std::vector<std::array<uint32_t, 2>> es;
  auto add_edge = [&](uint32_t a, uint32_t b) -> jgs2_status {
    if (a >= V || b >= V) return jgs2::fail(JGS2_ERROR_RANGE, "Surface vertex index %u is outside [0, %u).", std::max(a, b), V);
    if (a == b) return jgs2::fail(JGS2_ERROR_RANGE, "Surface edges must have two distinct vertices.");
    es.push_back({std::min(a, b), std::max(a, b)});
    return JGS2_OK;
  };
  for (uint32_t e = 0; e < edge_count; ++e)
    if (jgs2_status st = add_edge(edges[2 * size_t(e)], edges[2 * size_t(e) + 1])) return st;
  for (const auto& f : tris) {
    add_edge(f[0], f[1]);
    add_edge(f[1], f[2]);
    add_edge(f[2], f[0]);
  }
  std::sort(es.begin(), es.end());
  es.erase(std::unique(es.begin(), es.end()), es.end());
  // One-ring adjacency (buildOneRing, :571-584) as sorted rows.
  std::vector<std::vector<uint32_t>> ring(V);
  for (const auto& e : es) {
    ring[e[0]].push_back(e[1]);
    ring[e[1]].push_back(e[0]);
  }
  for (auto& r : ring) std::sort(r.begin(), r.end());
  auto closed = [&](uint32_t a, uint32_t b) {  // isInClosedOneRing (:586-592)
    return a == b || std::binary_search(ring[a].begin(), ring[a].end(), b);
  };
  // buildStaticIpcContactCandidates (:600-665): VT point-major over sorted
  // triangles, then EE pairs (i < j) over sorted edges.
  uint64_t written = 0;
  auto emit = [&](uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    if (packed && written < capacity) {
      packed[4 * written] = a;
      packed[4 * written + 1] = b;
      packed[4 * written + 2] = c;
      packed[4 * written + 3] = d;
    }
    ++written;
  };
  for (uint32_t v = 0; v < V; ++v)
    for (const auto& f : tris) {
      if (closed(v, f[0]) || closed(v, f[1]) || closed(v, f[2])) continue;
      emit(v, f[0], f[1], f[2]);
    }
  const uint64_t vt = written;
  for (size_t i = 0; i < es.size(); ++i)
    for (size_t j = i + 1; j < es.size(); ++j) {
      const auto& a = es[i];
      const auto& b = es[j];
      if (closed(a[0], b[0]) || closed(a[0], b[1]) || closed(a[1], b[0]) || closed(a[1], b[1])) continue;
      emit(a[0], a[1], b[0], b[1]);
    }
  if (written > 0xffffffffull / 4)
    return jgs2::fail(JGS2_ERROR_RANGE, "The static IPC superset has %llu candidates, beyond 32-bit indexing.",
                      (unsigned long long)written);
  *vt_count = uint32_t(vt);
  *ee_count = uint32_t(written - vt);
  if (packed && written > capacity)
    return jgs2::fail(JGS2_ERROR_RANGE, "Candidate capacity %llu is below the %llu candidates.",
                      (unsigned long long)capacity, (unsigned long long)written);
  jgs2::clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_scene_build_trough(const jgs2_trough_desc* d, jgs2_scene** out) {
  if (!d || !out) return jgs2::fail(JGS2_ERROR_TYPE, "NULL trough description.");
  *out = nullptr;
  if (d->box_count < 1 || d->box_count > 20) return jgs2::fail(JGS2_ERROR_RANGE, "box_count must lie in [1, 20].");
  if (d->box_cells < 1 || d->box_cells > 32) return jgs2::fail(JGS2_ERROR_RANGE, "box_cells must lie in [1, 32].");
  // troughMesh (src/scenes/index.ts:551-572): a pinned 2x1x1 cuboid warped
  // into a V whose two top faces share the apex column.
  jgs2::Mesh scene;
  {
    const uint32_t cells[3] = {2, 1, 1};
    const double origin[3] = {-2.2, 0, -1.5}, size[3] = {4.4, 1, 3};
    jgs2::cuboid_grid(cells, origin, size, 0, 0, true, scene);
    if (jgs2_status st = jgs2::orient(scene)) return st;
    for (uint32_t v = 0; v < scene.vertex_count(); ++v) {
      double* p = &scene.x[3 * size_t(v)];
      const double layer = p[1];
      p[1] = 0.05 + 0.65 * std::fabs(p[0]) - 0.18 * (1 - layer);
    }
  }
  // troughBox (:574-596): boxes of 0.46 m in layers of four slots, each with
  // a small fixed rotation; a nonzero seed replaces the placements by
  // seeded offsets within each slot and a seeded yaw.
  uint32_t state = d->seed;
  auto next = [&]() {  // the corpus PRNG (src/scenes/phase1.ts:106-115)
    state += 0x6d2b79f5u;
    uint32_t v = state;
    v = (v ^ (v >> 15)) * (v | 1u);
    v ^= v + (v ^ (v >> 7)) * (v | 61u);
    return double(v ^ (v >> 14)) / 4294967296.0;
  };
  for (uint32_t i = 0; i < d->box_count; ++i) {
    jgs2::Mesh box;
    const uint32_t cells[3] = {d->box_cells, d->box_cells, d->box_cells};
    const double origin[3] = {-0.23, -0.23, -0.23}, size[3] = {0.46, 0.46, 0.46};
    jgs2::cuboid_grid(cells, origin, size, 1 + i % 3, i + 1, false, box);
    if (jgs2_status st = jgs2::orient(box)) return st;
    const uint32_t slot = i % 4, layer = i / 4;
    double t[3] = {slot % 2 == 0 ? -0.37 : 0.37, 2.15 + 0.72 * layer, slot < 2 ? -0.38 : 0.38};
    double r[3] = {(double(i % 3) - 1) * 0.04, (double(i % 5) - 2) * 0.035, (i % 2 == 0 ? -1.0 : 1.0) * 0.055};
    if (d->seed != 0) {
      t[0] += (2 * next() - 1) * 0.06;
      t[2] += (2 * next() - 1) * 0.06;
      r[1] = (2 * next() - 1) * 3.141592653589793;
    }
    for (uint32_t v = 0; v < box.vertex_count(); ++v) {
      double* p = &box.x[3 * size_t(v)];
      jgs2::rotate_euler_xyz(p, r);
      for (int a = 0; a < 3; ++a) p[a] += t[a];
    }
    if (jgs2_status st = jgs2::orient(box)) return st;  // transformTetrahedralMesh (mesh.ts:237-282)
    jgs2::append(scene, box);
  }
  // TROUGH_MATERIAL and TROUGH_BOX_MATERIALS (:191-225).
  std::vector<jgs2::Material> mats = {{JGS2_MATERIAL_STABLE_NEO_HOOKEAN, 1000, 120000, 0.34},
                                      {JGS2_MATERIAL_STABLE_NEO_HOOKEAN, 850, 70000, 0.35},
                                      {JGS2_MATERIAL_STABLE_NEO_HOOKEAN, 850, 70000, 0.35},
                                      {JGS2_MATERIAL_STABLE_NEO_HOOKEAN, 850, 70000, 0.35}};
  jgs2_scene* s = new (std::nothrow) jgs2_scene();
  if (!s) return jgs2::fail(JGS2_ERROR_MEMORY, "Out of host memory.");
  const double h = 0.46 / double(d->box_cells);
  jgs2_status st = jgs2::pack_scene(scene, mats, d->cubature_k, false, h, s);
  if (st == JGS2_OK) st = jgs2::boundary_surface(scene, s->surface_triangles, s->surface_edges);
  if (st == JGS2_OK && d->build_candidates) {
    uint32_t vt = 0, ee = 0;
    const uint32_t nt = uint32_t(s->surface_triangles.size() / 3), ne = uint32_t(s->surface_edges.size() / 2);
    st = jgs2_build_static_contact_candidates(s->vertex_count, s->surface_triangles.data(), nt,
                                              s->surface_edges.data(), ne, nullptr, 0, &vt, &ee);
    if (st == JGS2_OK) {
      s->contact_indices.resize(4 * (size_t(vt) + ee));
      st = jgs2_build_static_contact_candidates(s->vertex_count, s->surface_triangles.data(), nt,
                                                s->surface_edges.data(), ne, s->contact_indices.data(),
                                                uint64_t(vt) + ee, &vt, &ee);
      s->has_contact = true;
      s->contact.vertex_triangle_count = vt;
      s->contact.edge_edge_count = ee;
      s->contact.packed_indices = s->contact_indices.data();
    }
  }
  if (st != JGS2_OK) {
    delete s;
    return st;
  }
  *out = s;
  jgs2::clear_error();
  return JGS2_OK;
}

jgs2_status jgs2_scene_get_surface(const jgs2_scene* s, const uint32_t** triangles, uint32_t* triangle_count,
                                   const uint32_t** edges, uint32_t* edge_count) {
  if (!s || !triangles || !triangle_count || !edges || !edge_count) return jgs2::fail(JGS2_ERROR_TYPE, "NULL surface argument.");
  *triangles = s->surface_triangles.data();
  *triangle_count = uint32_t(s->surface_triangles.size() / 3);
  *edges = s->surface_edges.data();
  *edge_count = uint32_t(s->surface_edges.size() / 2);
  jgs2::clear_error();
  return JGS2_OK;
}

}  // extern "C"
