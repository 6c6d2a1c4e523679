// Internal helpers shared by the host translation units of libjgs2hip.
#pragma once

#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "jgs2hip.h"

namespace jgs2 {

// FP64 conjugate gradients on the current HIP device for the cubature
// precompute's global rest-Hessian solves (jgs2_precompute_kernels.hip): the
// block-CSR Hessian over free vertices is uploaded once, then solve() runs
// block-Jacobi preconditioned CG on up to 32 right-hand sides together
// (n x m row-major, overwritten by the solutions).
class GpuCg {
 public:
  GpuCg();
  ~GpuCg();
  GpuCg(const GpuCg&) = delete;
  GpuCg& operator=(const GpuCg&) = delete;
  static bool available();
  jgs2_status upload(uint32_t rows, const uint32_t* start, const uint32_t* col, const double* blocks,
                     const double* block_jacobi);
  jgs2_status solve(double* b, uint32_t m, double tol, uint32_t max_iterations, uint32_t* iterations,
                    double* milliseconds);

 private:
  struct Impl;
  Impl* impl_;
};

// Greedy nonnegative column selection of the cubature training on the GPU
// (jgs2_precompute_kernels.hip), one job per source vertex: `count` columns
// of `rows` doubles at columns[col_offset], the target at
// targets[tgt_offset], its norm, and (1 + K) * count doubles of device
// scratch at scratch_offset. Outputs per job: K candidate indices (-1
// unused), K weights, the final squared residual and the number selected.
struct GpuSelectJob {
  uint64_t col_offset;
  uint64_t tgt_offset;
  uint64_t scratch_offset;
  uint32_t count;
  uint32_t pad;
  double target_norm;
};
jgs2_status gpu_greedy_select(const double* columns, size_t column_doubles, const double* targets,
                              size_t target_doubles, const GpuSelectJob* jobs, uint32_t job_count, uint32_t rows,
                              uint32_t K, double tol, uint32_t passes, int32_t* out_sel, double* out_w,
                              double* out_cur, uint32_t* out_n, double* milliseconds);

// Records the calling thread's last error message and returns `status`.
jgs2_status fail(jgs2_status status, const char* format, ...);
void clear_error();

// minimumStaticIpcContactDistance (src/simulation/cpu/ipc-contact.ts:349-391)
// in f64 over positions with 4 floats per vertex.
double min_static_contact_distance(const float* positions4, const jgs2_contact_input& contact);

// Canonical contact surface of a broad-phase scene: canonicalTriangles,
// canonicalEdges and buildOneRing (src/simulation/cpu/ipc-contact.ts:
// 500-584). Triangles and edges ascending; ring rows ascending (CSR).
struct Surface {
  std::vector<std::array<uint32_t, 3>> tris;
  std::vector<std::array<uint32_t, 2>> edges;
  std::vector<uint32_t> ring_start, ring;
};
jgs2_status canonical_surface(uint32_t vertex_count, const uint32_t* triangles, uint32_t triangle_count,
                              const uint32_t* edges, uint32_t edge_count, Surface& out);
// Whether the static superset of the surface holds any pair.
bool superset_nonempty(const Surface& s, uint32_t vertex_count);
// Host scan of the superset's pairs whose f32 AABB distance lower bound at
// `positions4` is within `reach` (a point that is not finite counts against
// everything): the pair count, and the minimum f64 distance over pairs whose
// bound is within `distance_reach` (+inf when there are none).
struct SurfaceScan {
  uint64_t vertex_triangle = 0, edge_edge = 0;
  double min_distance = INFINITY;
};
SurfaceScan scan_surface_pairs(const Surface& s, uint32_t vertex_count, const float* positions4, float reach,
                               float distance_reach);

}  // namespace jgs2
