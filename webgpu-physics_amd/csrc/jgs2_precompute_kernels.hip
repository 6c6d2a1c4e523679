// FP64 conjugate gradients on the GPU for the cubature precompute's global
// solves: the block inverse iteration that yields the low-frequency rest
// modes (src/simulation/cpu/cubature.ts:224-297) solves H Y = X for every
// mode at each of its iterations. The rest Hessian H = M/h^2 + K stays
// resident in HBM as 3x3 blocks in CSR over free vertices; all m right-hand
// sides advance together so each iteration reads H once.
//
// Layout: vectors are n x m row-major (coordinate-major, the m right-hand
// sides of one coordinate adjacent), so a wave's lanes read consecutive
// doubles. Reductions are two-stage and fixed-order (per-block partials, then
// one pass over the blocks), so results do not depend on scheduling. Each
// right-hand side runs its own CG recurrence; once converged its step sizes
// are zero and it no longer changes.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "jgs2_internal.h"

namespace jgs2 {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxRhs = 32;
constexpr int kDotBlocks = 1024;

// y = H x for m right-hand sides; one thread per (free vertex, rhs).
__global__ void __launch_bounds__(kThreads)
k_bsr_spmm(const uint32_t* __restrict__ start, const uint32_t* __restrict__ col, const double* __restrict__ blocks,
           const double* __restrict__ x, double* __restrict__ y, uint32_t rows, uint32_t m) {
  const uint32_t per = kThreads / m;
  const uint32_t j = threadIdx.x % m, local = threadIdx.x / m;
  const uint32_t a = blockIdx.x * per + local;
  if (local >= per || a >= rows) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t k = start[a]; k < start[a + 1]; ++k) {
    const double* b = blocks + 9 * size_t(k);
    const double* xv = x + size_t(3) * col[k] * m + j;
    const double x0 = xv[0], x1 = xv[m], x2 = xv[2 * m];
    s0 += b[0] * x0 + b[1] * x1 + b[2] * x2;
    s1 += b[3] * x0 + b[4] * x1 + b[5] * x2;
    s2 += b[6] * x0 + b[7] * x1 + b[8] * x2;
  }
  double* yv = y + size_t(3) * a * m + j;
  yv[0] = s0;
  yv[m] = s1;
  yv[2 * m] = s2;
}

// Partial sums of <u_j, v_j> (and <w_j, w_j> when w is given) over a
// grid-strided set of coordinates, in a fixed order per thread and a fixed
// LDS tree per block.
__global__ void __launch_bounds__(kThreads)
k_dot_partial(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w, size_t n,
              uint32_t m, double* __restrict__ partial) {
  __shared__ double su[kThreads], sw[kThreads];
  const uint32_t per = kThreads / m;
  const uint32_t j = threadIdx.x % m, local = threadIdx.x / m;
  double acc = 0.0, accw = 0.0;
  if (local < per)
    for (size_t i = size_t(blockIdx.x) * per + local; i < n; i += size_t(gridDim.x) * per) {
      const size_t e = i * m + j;
      acc += u[e] * v[e];
      if (w) accw += w[e] * w[e];
    }
  su[threadIdx.x] = acc;
  sw[threadIdx.x] = accw;
  __syncthreads();
  if (threadIdx.x < m) {
    double s = 0.0, t = 0.0;
    for (uint32_t r = 0; r < per; ++r) {
      s += su[r * m + threadIdx.x];
      t += sw[r * m + threadIdx.x];
    }
    partial[(size_t(blockIdx.x) * m + threadIdx.x) * 2] = s;
    partial[(size_t(blockIdx.x) * m + threadIdx.x) * 2 + 1] = t;
  }
}

// The scalar kernels below run as one block of kLanes x kMaxRhs threads:
// thread (lane, j) sums partials lane, lane + kLanes, ... of rhs j, then a
// fixed LDS tree over the lanes finishes the sum, so the order is fixed.
constexpr int kLanes = 32;

__device__ void sum_partials(const double* __restrict__ partial, uint32_t blocks, uint32_t m, double* s0, double* s1,
                             double& a, double& b) {
  const uint32_t j = threadIdx.x % kMaxRhs, lane = threadIdx.x / kMaxRhs;
  double x = 0.0, y = 0.0;
  if (j < m)
    for (uint32_t blk = lane; blk < blocks; blk += kLanes) {
      x += partial[(size_t(blk) * m + j) * 2];
      y += partial[(size_t(blk) * m + j) * 2 + 1];
    }
  s0[threadIdx.x] = x;
  s1[threadIdx.x] = y;
  __syncthreads();
  for (uint32_t stride = kLanes / 2; stride > 0; stride /= 2) {
    if (lane < stride) {
      s0[threadIdx.x] += s0[threadIdx.x + stride * kMaxRhs];
      s1[threadIdx.x] += s1[threadIdx.x + stride * kMaxRhs];
    }
    __syncthreads();
  }
  a = s0[j];
  b = s1[j];
}

// out[2 j], out[2 j + 1]: the two sums of rhs j.
__global__ void __launch_bounds__(kLanes* kMaxRhs)
k_dot_final(const double* __restrict__ partial, uint32_t blocks, uint32_t m, double* __restrict__ out) {
  __shared__ double s0[kLanes * kMaxRhs], s1[kLanes * kMaxRhs];
  double a, b;
  sum_partials(partial, blocks, m, s0, s1, a, b);
  const uint32_t j = threadIdx.x;
  if (j >= m) return;
  out[2 * j] = a;
  out[2 * j + 1] = b;
}

// Per-rhs CG scalars, state[j] = {rz, alpha, beta, done}.
// alpha = rz / <d, q>; zero once converged or when <d, q> is not positive.
__global__ void __launch_bounds__(kLanes* kMaxRhs)
k_cg_alpha(const double* __restrict__ partial, uint32_t blocks, double* __restrict__ state, uint32_t m) {
  __shared__ double s0[kLanes * kMaxRhs], s1[kLanes * kMaxRhs];
  double dq, unused;
  sum_partials(partial, blocks, m, s0, s1, dq, unused);
  const uint32_t j = threadIdx.x;
  if (j >= m) return;
  double* s = state + 4 * j;
  s[1] = (s[3] != 0.0 || !(dq > 0.0)) ? 0.0 : s[0] / dq;
}

// beta = <r, z>_new / <r, z>; converged (done, alpha = beta = 0) once
// <r, r> <= threshold[j] = (tol |b_j|)^2.
__global__ void __launch_bounds__(kLanes* kMaxRhs)
k_cg_beta(const double* __restrict__ partial, uint32_t blocks, double* __restrict__ state,
          const double* __restrict__ threshold, uint32_t m) {
  __shared__ double s0[kLanes * kMaxRhs], s1[kLanes * kMaxRhs];
  double rz_new, rr;
  sum_partials(partial, blocks, m, s0, s1, rz_new, rr);
  const uint32_t j = threadIdx.x;
  if (j >= m) return;
  double* s = state + 4 * j;
  if (s[3] != 0.0) {
    s[2] = 0.0;
    return;
  }
  if (rr <= threshold[j] || !(s[0] > 0.0)) {
    s[3] = 1.0;
    s[2] = 0.0;
    s[1] = 0.0;
    return;
  }
  s[2] = rz_new / s[0];
  s[0] = rz_new;
}

// x += alpha d, r -= alpha q, z = P r (block-Jacobi) for one free vertex and rhs.
__global__ void __launch_bounds__(kThreads)
k_cg_update(double* __restrict__ x, double* __restrict__ r, double* __restrict__ z, const double* __restrict__ d,
            const double* __restrict__ q, const double* __restrict__ pre, const double* __restrict__ state,
            uint32_t rows, uint32_t m) {
  const uint32_t per = kThreads / m;
  const uint32_t j = threadIdx.x % m, local = threadIdx.x / m;
  const uint32_t a = blockIdx.x * per + local;
  if (local >= per || a >= rows) return;
  const double alpha = state[4 * j + 1];
  const size_t e = size_t(3) * a * m + j;
  double rv[3];
  for (int c = 0; c < 3; ++c) {
    const size_t i = e + size_t(c) * m;
    x[i] += alpha * d[i];
    rv[c] = r[i] - alpha * q[i];
    r[i] = rv[c];
  }
  const double* p = pre + 9 * size_t(a);
  for (int c = 0; c < 3; ++c) z[e + size_t(c) * m] = p[c * 3] * rv[0] + p[c * 3 + 1] * rv[1] + p[c * 3 + 2] * rv[2];
}

// z = P r (the first preconditioned residual) for one free vertex and rhs.
__global__ void __launch_bounds__(kThreads)
k_cg_precondition(const double* __restrict__ r, double* __restrict__ z, const double* __restrict__ pre, uint32_t rows,
                  uint32_t m) {
  const uint32_t per = kThreads / m;
  const uint32_t j = threadIdx.x % m, local = threadIdx.x / m;
  const uint32_t a = blockIdx.x * per + local;
  if (local >= per || a >= rows) return;
  const size_t e = size_t(3) * a * m + j;
  const double r0 = r[e], r1 = r[e + m], r2 = r[e + 2 * m];
  const double* p = pre + 9 * size_t(a);
  for (int c = 0; c < 3; ++c) z[e + size_t(c) * m] = p[c * 3] * r0 + p[c * 3 + 1] * r1 + p[c * 3 + 2] * r2;
}

// d = z + beta d (d = z when beta is 0).
__global__ void __launch_bounds__(kThreads)
k_cg_direction(double* __restrict__ d, const double* __restrict__ z, const double* __restrict__ state, size_t n,
               uint32_t m) {
  const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n * m) return;
  const double beta = state[4 * (i % m) + 2];
  d[i] = beta == 0.0 ? z[i] : z[i] + beta * d[i];  // the first direction never reads d
}

bool hip_ok(hipError_t e) { return e == hipSuccess; }

}  // namespace

struct GpuCg::Impl {
  uint32_t rows = 0;
  size_t n = 0;
  uint32_t* start = nullptr;
  uint32_t* col = nullptr;
  double* blocks = nullptr;
  double* pre = nullptr;
  hipStream_t stream = nullptr;
  bool ready = false;
};

GpuCg::GpuCg() : impl_(new Impl) {}

GpuCg::~GpuCg() {
  if (impl_->stream) (void)hipStreamSynchronize(impl_->stream);
  (void)hipFree(impl_->start);
  (void)hipFree(impl_->col);
  (void)hipFree(impl_->blocks);
  (void)hipFree(impl_->pre);
  if (impl_->stream) (void)hipStreamDestroy(impl_->stream);
  delete impl_;
}

bool GpuCg::available() {
  int count = 0;
  return hipGetDeviceCount(&count) == hipSuccess && count > 0;
}

jgs2_status GpuCg::upload(uint32_t rows, const uint32_t* start, const uint32_t* col, const double* blocks,
                          const double* block_jacobi) {
  Impl& g = *impl_;
  g.rows = rows;
  g.n = size_t(3) * rows;
  const size_t nnzb = start[rows];
  if (!hip_ok(hipStreamCreate(&g.stream))) return fail(JGS2_ERROR_DEVICE, "hipStreamCreate failed.");
  if (!hip_ok(hipMalloc(&g.start, (rows + 1) * sizeof(uint32_t))) || !hip_ok(hipMalloc(&g.col, nnzb * sizeof(uint32_t))) ||
      !hip_ok(hipMalloc(&g.blocks, 9 * nnzb * sizeof(double))) || !hip_ok(hipMalloc(&g.pre, 9 * size_t(rows) * sizeof(double))))
    return fail(JGS2_ERROR_MEMORY, "Cubature precompute: device allocation for the rest Hessian failed.");
  if (!hip_ok(hipMemcpy(g.start, start, (rows + 1) * sizeof(uint32_t), hipMemcpyHostToDevice)) ||
      !hip_ok(hipMemcpy(g.col, col, nnzb * sizeof(uint32_t), hipMemcpyHostToDevice)) ||
      !hip_ok(hipMemcpy(g.blocks, blocks, 9 * nnzb * sizeof(double), hipMemcpyHostToDevice)) ||
      !hip_ok(hipMemcpy(g.pre, block_jacobi, 9 * size_t(rows) * sizeof(double), hipMemcpyHostToDevice)))
    return fail(JGS2_ERROR_DEVICE, "Cubature precompute: upload of the rest Hessian failed.");
  g.ready = true;
  return JGS2_OK;
}

// Solves H Y = B for m <= 32 right-hand sides given n x m row-major; the
// solutions overwrite `b`. Each rhs stops at |r| <= tol |b| (or after
// max_iterations); `iterations` reports the iterations run.
jgs2_status GpuCg::solve(double* b, uint32_t m, double tol, uint32_t max_iterations, uint32_t* iterations,
                         double* milliseconds) {
  Impl& g = *impl_;
  if (!g.ready) return fail(JGS2_ERROR_STATE, "Cubature precompute: the rest Hessian is not on the device.");
  if (m == 0 || m > kMaxRhs) return fail(JGS2_ERROR_RANGE, "Cubature precompute: 1 to 32 right-hand sides per solve.");
  const size_t nm = g.n * m;
  double *x = nullptr, *r = nullptr, *z = nullptr, *d = nullptr, *q = nullptr, *partial = nullptr, *dots = nullptr,
         *state = nullptr, *threshold = nullptr;
  auto release = [&]() {
    for (double* p : {x, r, z, d, q, partial, dots, state, threshold}) (void)hipFree(p);
  };
  if (!hip_ok(hipMalloc(&x, nm * 8)) || !hip_ok(hipMalloc(&r, nm * 8)) || !hip_ok(hipMalloc(&z, nm * 8)) ||
      !hip_ok(hipMalloc(&d, nm * 8)) || !hip_ok(hipMalloc(&q, nm * 8)) ||
      !hip_ok(hipMalloc(&partial, size_t(kDotBlocks) * m * 2 * 8)) || !hip_ok(hipMalloc(&dots, 2 * m * 8)) ||
      !hip_ok(hipMalloc(&state, 4 * m * 8)) || !hip_ok(hipMalloc(&threshold, m * 8))) {
    release();
    return fail(JGS2_ERROR_MEMORY, "Cubature precompute: device allocation for the CG vectors failed.");
  }
  // |b_j|^2 on the host (the right-hand sides are there already)
  std::vector<double> bn(m, 0.0), th(m), st(4 * m, 0.0), dv(2 * m, 0.0);
  for (size_t i = 0; i < g.n; ++i)
    for (uint32_t j = 0; j < m; ++j) bn[j] += b[i * m + j] * b[i * m + j];
  for (uint32_t j = 0; j < m; ++j) {
    th[j] = tol * tol * bn[j];
    st[4 * j + 3] = bn[j] == 0.0 ? 1.0 : 0.0;  // a zero rhs is solved by x = 0
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const uint32_t per = kThreads / m;
  const uint32_t row_blocks = (g.rows + per - 1) / per;
  const uint32_t dot_blocks = uint32_t(std::min<size_t>(kDotBlocks, (g.n + per - 1) / per));
  const uint32_t flat_blocks = uint32_t((nm + kThreads - 1) / kThreads);
  // state = {rz, alpha, beta, done}; beta = 0 makes the first direction d = z
  bool ok = hip_ok(hipMemcpyAsync(r, b, nm * 8, hipMemcpyHostToDevice, g.stream)) &&
            hip_ok(hipMemcpyAsync(threshold, th.data(), m * 8, hipMemcpyHostToDevice, g.stream)) &&
            hip_ok(hipMemcpyAsync(state, st.data(), 4 * m * 8, hipMemcpyHostToDevice, g.stream)) &&
            hip_ok(hipMemsetAsync(x, 0, nm * 8, g.stream));
  (void)hipEventRecord(e0, g.stream);
  k_cg_precondition<<<row_blocks, kThreads, 0, g.stream>>>(r, z, g.pre, g.rows, m);
  k_cg_direction<<<flat_blocks, kThreads, 0, g.stream>>>(d, z, state, g.n, m);
  k_dot_partial<<<dot_blocks, kThreads, 0, g.stream>>>(r, z, nullptr, g.n, m, partial);
  k_dot_final<<<1, kLanes * kMaxRhs, 0, g.stream>>>(partial, dot_blocks, m, dots);
  ok = ok && hip_ok(hipMemcpyAsync(dv.data(), dots, 2 * m * 8, hipMemcpyDeviceToHost, g.stream)) &&
       hip_ok(hipStreamSynchronize(g.stream));
  for (uint32_t j = 0; j < m; ++j) st[4 * j] = dv[2 * j];
  ok = ok && hip_ok(hipMemcpyAsync(state, st.data(), 4 * m * 8, hipMemcpyHostToDevice, g.stream));
  // One CG iteration is seven launches; kChunk iterations run as one hipGraph
  // (no memory-set nodes), after which the host reads the convergence flags.
  // Iterations past a rhs's convergence leave it unchanged (alpha = 0).
  constexpr uint32_t kChunk = 16;
  auto iteration = [&]() {
    k_bsr_spmm<<<row_blocks, kThreads, 0, g.stream>>>(g.start, g.col, g.blocks, d, q, g.rows, m);
    k_dot_partial<<<dot_blocks, kThreads, 0, g.stream>>>(d, q, nullptr, g.n, m, partial);
    k_cg_alpha<<<1, kLanes * kMaxRhs, 0, g.stream>>>(partial, dot_blocks, state, m);
    k_cg_update<<<row_blocks, kThreads, 0, g.stream>>>(x, r, z, d, q, g.pre, state, g.rows, m);
    k_dot_partial<<<dot_blocks, kThreads, 0, g.stream>>>(r, z, r, g.n, m, partial);
    k_cg_beta<<<1, kLanes * kMaxRhs, 0, g.stream>>>(partial, dot_blocks, state, threshold, m);
    k_cg_direction<<<flat_blocks, kThreads, 0, g.stream>>>(d, z, state, g.n, m);
  };
  // JGS2_CG_GRAPHS=0 launches every iteration eagerly (same kernels, same
  // order, same results), for tools that trace kernels outside graphs only.
  const char* env = std::getenv("JGS2_CG_GRAPHS");
  const bool graphs = !(env && env[0] == '0');
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (ok && graphs && hip_ok(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal))) {
    for (uint32_t k = 0; k < kChunk; ++k) iteration();
    if (!hip_ok(hipStreamEndCapture(g.stream, &graph)) || !hip_ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)))
      exec = nullptr;
  }
  (void)hipGetLastError();
  uint32_t it = 0;
  std::vector<double> host_state(4 * m);
  while (ok && it < max_iterations) {
    if (exec && it + kChunk <= max_iterations) {
      ok = hip_ok(hipGraphLaunch(exec, g.stream));
      it += kChunk;
    } else {
      const uint32_t steps = std::min(kChunk, max_iterations - it);
      for (uint32_t k = 0; k < steps; ++k) iteration();
      it += steps;
    }
    ok = ok && hip_ok(hipMemcpyAsync(host_state.data(), state, 4 * m * 8, hipMemcpyDeviceToHost, g.stream)) &&
         hip_ok(hipStreamSynchronize(g.stream));
    bool all = true;
    for (uint32_t j = 0; j < m; ++j) all = all && host_state[4 * j + 3] != 0.0;
    if (all) break;
  }
  // An error can leave the loop with a replay still queued: drain the
  // stream before the exec goes away.
  (void)hipStreamSynchronize(g.stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventRecord(e1, g.stream);
  ok = ok && hip_ok(hipMemcpyAsync(b, x, nm * 8, hipMemcpyDeviceToHost, g.stream)) &&
       hip_ok(hipStreamSynchronize(g.stream)) && hip_ok(hipGetLastError());
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  release();
  if (iterations) *iterations = it;
  if (milliseconds) *milliseconds = ms;
  if (!ok) return fail(JGS2_ERROR_DEVICE, "Cubature precompute: conjugate gradients on the device failed.");
  return JGS2_OK;
}


// ------------------------------------------------------------ greedy NNLS selection
// The cubature training's greedy nonnegative column selection
// (src/simulation/cpu/cubature.ts:386-691) for many source vertices at once:
// one wave per vertex, lanes over candidate columns. Every comparison is the
// reference's: the NNLS masks run in order per candidate, and one lane scans
// the candidates in order with the 1e-14 relative tolerance and the
// lowest-index tie rule. Floating-point contraction is off and every sum runs
// in the host's order, so the selections are the host's bit for bit.
namespace {

template <int kSelK>
struct MaskResult {
  double err;
  double w[kSelK];
  int ok;
};

__device__ double sel_dot(const double* a, const double* b, uint32_t rows) {
#pragma clang fp contract(off)
  double r = 0.0;
  for (uint32_t i = 0; i < rows; ++i) r += a[i] * b[i];
  return r;
}

// Gaussian elimination with partial pivoting (math.ts:134-199), n <= kSelK.
__device__ __forceinline__ bool sel_solve(double* a, double* b, int n) {
#pragma clang fp contract(off)
  for (int p = 0; p < n; ++p) {
    int best = p;
    double mag = fabs(a[p * n + p]);
    for (int r = p + 1; r < n; ++r)
      if (fabs(a[r * n + p]) > mag) {
        mag = fabs(a[r * n + p]);
        best = r;
      }
    if (mag <= 1e-14 || !isfinite(mag)) return false;
    if (best != p) {
      for (int c = p; c < n; ++c) {
        const double t = a[p * n + c];
        a[p * n + c] = a[best * n + c];
        a[best * n + c] = t;
      }
      const double t = b[p];
      b[p] = b[best];
      b[best] = t;
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

// One mask of a trial set of n columns `col[i]` with Gram entries g(i, j)
// and target dots td[i]: least squares over the mask, rejected when singular
// or a weight is below -1e-10; the error is |t - sum w_i col_i|^2 formed
// row by row in column order.
template <int kSelK, class G>
__device__ MaskResult<kSelK> sel_fit(const double* const* col, const double* td, G g, int n, unsigned mask,
                                     const double* t, uint32_t rows) {
#pragma clang fp contract(off)
  MaskResult<kSelK> f;
  f.ok = 0;
  f.err = 0.0;
  for (int i = 0; i < kSelK; ++i) f.w[i] = 0.0;
  int act[kSelK], d = 0;
  for (int i = 0; i < n; ++i)
    if (mask & (1u << i)) act[d++] = i;
  double A[kSelK * kSelK], b[kSelK];
  for (int r = 0; r < d; ++r) {
    b[r] = td[act[r]];
    for (int c = 0; c < d; ++c) A[r * d + c] = g(act[r], act[c]);
  }
  if (!sel_solve(A, b, d)) return f;
  for (int i = 0; i < d; ++i)
    if (b[i] < -1e-10) return f;
  for (int i = 0; i < d; ++i) f.w[act[i]] = fmax(0.0, b[i]);
  double err = 0.0;
  for (uint32_t r = 0; r < rows; ++r) {
    double v = t[r];
    for (int i = 0; i < n; ++i)
      if (f.w[i] != 0.0) v -= col[i][r] * f.w[i];
    err += v * v;
  }
  f.err = err;
  f.ok = 1;
  return f;
}

__device__ double sel_residual_sq(const double* const* col, const double* w, int n, const double* t, uint32_t rows) {
#pragma clang fp contract(off)
  double err = 0.0;
  for (uint32_t r = 0; r < rows; ++r) {
    double v = t[r];
    for (int i = 0; i < n; ++i) v -= col[i][r] * w[i];
    err += v * v;
  }
  return err;
}

template <int kSelK>
struct SelState {
  int sel[kSelK];
  double w[kSelK];
  int n;
  double cur;
  int done;
  // scan results
  int best, best_slot;
  double be, bw[kSelK];
};

template <int kSelK>
__global__ void __launch_bounds__(64)
k_greedy_select(const double* __restrict__ cols, const double* __restrict__ tgts, const GpuSelectJob* __restrict__ jobs,
                uint32_t rows, uint32_t K, double tol, uint32_t passes, double* __restrict__ scratch,
                int32_t* __restrict__ out_sel, double* __restrict__ out_w, double* __restrict__ out_cur,
                uint32_t* __restrict__ out_n) {
#pragma clang fp contract(off)
  __shared__ MaskResult<kSelK> tab[1 << kSelK];
  __shared__ double gsel[kSelK * kSelK], tdsel[kSelK];
  __shared__ SelState<kSelK> st;
  const GpuSelectJob job = jobs[blockIdx.x];
  const uint32_t lane = threadIdx.x;
  const double* C = cols + job.col_offset;
  const double* t = tgts + job.tgt_offset;
  const uint32_t count = job.count;
  double* esc = scratch + job.scratch_offset;  // per candidate: 1 + K doubles
  // usable candidates: squared norm above 1e-18
  uint32_t usable = 0;
  for (uint32_t c = lane; c < count; c += 64) {
    const bool u = sel_dot(C + size_t(c) * rows, C + size_t(c) * rows, rows) > 1e-18;
    esc[size_t(c) * (1 + K)] = u ? 0.0 : -1.0;  // -1 marks an unusable column
    usable += u ? 1u : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) usable += __shfl_down(usable, off, 64);
  usable = __shfl(usable, 0, 64);
  const double tt = sel_dot(t, t, rows);
  if (lane == 0) {
    st.n = 0;
    st.cur = tt;
    st.done = 0;
  }
  __syncthreads();
  const uint32_t limit = usable < K ? usable : K;
  const double tn = job.target_norm;
  auto selected = [&](uint32_t c) {
    for (int i = 0; i < st.n; ++i)
      if (st.sel[i] == int(c)) return true;
    return false;
  };
  // Gram of the selected columns and their target dots
  auto refresh_gram = [&]() {
    for (int e = int(lane); e < st.n * st.n; e += 64) {
      const int i = e / st.n, j = e % st.n;
      gsel[i * kSelK + j] = sel_dot(C + size_t(st.sel[i]) * rows, C + size_t(st.sel[j]) * rows, rows);
    }
    for (int i = int(lane); i < st.n; i += 64) tdsel[i] = sel_dot(C + size_t(st.sel[i]) * rows, t, rows);
    __syncthreads();
  };
  if (tn > 1e-12 && limit > 0) {
    for (uint32_t step = 0; step < limit; ++step) {
      refresh_gram();
      const int n = st.n + 1;
      // masks without the new (last) position, shared by every candidate
      const double* cs[kSelK];
      for (int i = 0; i < st.n; ++i) cs[i] = C + size_t(st.sel[i]) * rows;
      for (unsigned mask = lane + 1; mask < (1u << (n - 1)); mask += 64)
        tab[mask] = sel_fit<kSelK>(cs, tdsel, [&](int i, int j) { return gsel[i * kSelK + j]; }, n, mask, t, rows);
      __syncthreads();
      for (uint32_t c = lane; c < count; c += 64) {
        double* ec = esc + size_t(c) * (1 + K);
        if (ec[0] < 0.0 || selected(c)) continue;
        const double* colc = C + size_t(c) * rows;
        cs[n - 1] = colc;
        double gc[kSelK];
        for (int i = 0; i < n - 1; ++i) gc[i] = sel_dot(colc, cs[i], rows);
        gc[n - 1] = sel_dot(colc, colc, rows);
        double td[kSelK];
        for (int i = 0; i < n - 1; ++i) td[i] = tdsel[i];
        td[n - 1] = sel_dot(colc, t, rows);
        auto g = [&](int i, int j) {
          if (i == n - 1) return gc[j];
          if (j == n - 1) return gc[i];
          return gsel[i * kSelK + j];
        };
        double best = tt, bw[kSelK];
        for (int i = 0; i < kSelK; ++i) bw[i] = 0.0;
        for (unsigned mask = 1; mask < (1u << n); ++mask) {
          const MaskResult<kSelK> f =
              (mask & (1u << (n - 1))) ? sel_fit<kSelK>(cs, td, g, n, mask, t, rows) : tab[mask];
          if (!f.ok) continue;
          if (f.err < best - 1e-14 * fmax(1.0, best)) {
            best = f.err;
            for (int i = 0; i < kSelK; ++i) bw[i] = f.w[i];
          }
        }
        ec[0] = sel_residual_sq(cs, bw, n, t, rows);
        for (int i = 0; i < n; ++i) ec[1 + i] = bw[i];
      }
      __syncthreads();
      if (lane == 0) {
        int best = -1;
        double be = st.cur;
        for (uint32_t c = 0; c < count; ++c) {
          const double* ec = esc + size_t(c) * (1 + K);
          if (ec[0] < 0.0 || selected(c)) continue;
          const double ex = ec[0], tl = 1e-14 * fmax(1.0, be);
          if (ex < be - tl || (fabs(ex - be) <= tl && best >= 0 && int(c) < best)) {
            best = int(c);
            be = ex;
          }
        }
        if (best < 0) {
          st.done = 1;
        } else {
          const double* ec = esc + size_t(best) * (1 + K);
          st.sel[st.n] = best;
          for (int i = 0; i < n; ++i) st.w[i] = ec[1 + i];
          st.n = n;
          st.cur = be;
          if (sqrt(be) / tn < tol) st.done = 1;
        }
      }
      __syncthreads();
      if (st.done) break;
      // the marks of usable columns were overwritten by errors: restore them
      for (uint32_t c = lane; c < count; c += 64) {
        double* ec = esc + size_t(c) * (1 + K);
        if (ec[0] >= 0.0) ec[0] = 0.0;
      }
      __syncthreads();
    }
    // one-for-one replacement passes
    for (uint32_t pass = 0; pass < passes && st.n > 0; ++pass) {
      for (uint32_t c = lane; c < count; c += 64) {
        double* ec = esc + size_t(c) * (1 + K);
        if (ec[0] >= 0.0) ec[0] = 0.0;
      }
      refresh_gram();
      const int n = st.n;
      if (lane == 0) {
        st.best = -1;
        st.best_slot = -1;
        st.be = st.cur;
      }
      __syncthreads();
      for (int sl = 0; sl < n; ++sl) {
        const double* cs[kSelK];
        for (int i = 0; i < n; ++i) cs[i] = C + size_t(st.sel[i]) * rows;
        for (unsigned mask = lane + 1; mask < (1u << n); mask += 64)
          if (!(mask & (1u << sl)))
            tab[mask] = sel_fit<kSelK>(cs, tdsel, [&](int i, int j) { return gsel[i * kSelK + j]; }, n, mask, t, rows);
        __syncthreads();
        for (uint32_t c = lane; c < count; c += 64) {
          double* ec = esc + size_t(c) * (1 + K);
          if (ec[0] < 0.0 || selected(c)) continue;
          const double* colc = C + size_t(c) * rows;
          const double* ct[kSelK];
          for (int i = 0; i < n; ++i) ct[i] = i == sl ? colc : cs[i];
          double gc[kSelK];
          for (int i = 0; i < n; ++i) gc[i] = i == sl ? sel_dot(colc, colc, rows) : sel_dot(colc, cs[i], rows);
          double td[kSelK];
          for (int i = 0; i < n; ++i) td[i] = i == sl ? sel_dot(colc, t, rows) : tdsel[i];
          auto g = [&](int i, int j) {
            if (i == sl) return gc[j];
            if (j == sl) return gc[i];
            return gsel[i * kSelK + j];
          };
          double best = tt, bw[kSelK];
          for (int i = 0; i < kSelK; ++i) bw[i] = 0.0;
          for (unsigned mask = 1; mask < (1u << n); ++mask) {
            const MaskResult<kSelK> f = (mask & (1u << sl)) ? sel_fit<kSelK>(ct, td, g, n, mask, t, rows) : tab[mask];
            if (!f.ok) continue;
            if (f.err < best - 1e-14 * fmax(1.0, best)) {
              best = f.err;
              for (int i = 0; i < kSelK; ++i) bw[i] = f.w[i];
            }
          }
          ec[0] = sel_residual_sq(ct, bw, n, t, rows);
          for (int i = 0; i < n; ++i) ec[1 + i] = bw[i];
        }
        __syncthreads();
        if (lane == 0) {
          for (uint32_t c = 0; c < count; ++c) {
            const double* ec = esc + size_t(c) * (1 + K);
            if (ec[0] < 0.0 || selected(c)) continue;
            const double ex = ec[0], tl = 1e-14 * fmax(1.0, st.be);
            if (ex < st.be - tl || (fabs(ex - st.be) <= tl && st.best >= 0 && int(c) < st.best)) {
              st.best = int(c);
              st.best_slot = sl;
              st.be = ex;
              for (int i = 0; i < n; ++i) st.bw[i] = ec[1 + i];
            }
          }
        }
        __syncthreads();
      }
      if (st.best_slot < 0) break;
      __syncthreads();
      if (lane == 0) {
        st.sel[st.best_slot] = st.best;
        for (int i = 0; i < n; ++i) st.w[i] = st.bw[i];
        st.cur = st.be;
      }
      __syncthreads();
    }
  }
  if (lane == 0) {
    out_n[blockIdx.x] = uint32_t(st.n);
    out_cur[blockIdx.x] = st.cur;
    for (uint32_t i = 0; i < K; ++i) {
      out_sel[size_t(blockIdx.x) * K + i] = int(i) < st.n ? st.sel[i] : -1;
      out_w[size_t(blockIdx.x) * K + i] = int(i) < st.n ? st.w[i] : 0.0;
    }
  }
}

}  // namespace

jgs2_status gpu_greedy_select(const double* columns, size_t column_doubles, const double* targets,
                              size_t target_doubles, const GpuSelectJob* jobs, uint32_t job_count, uint32_t rows,
                              uint32_t K, double tol, uint32_t passes, int32_t* out_sel, double* out_w,
                              double* out_cur, uint32_t* out_n, double* milliseconds) {
  if (K == 0 || K > 8u) return fail(JGS2_ERROR_RANGE, "Device selection supports 1 to 8 samples.");
  if (job_count == 0) return JGS2_OK;
  size_t scratch_doubles = 0;
  for (uint32_t j = 0; j < job_count; ++j)
    scratch_doubles = std::max(scratch_doubles, size_t(jobs[j].scratch_offset) + size_t(jobs[j].count) * (1 + K));
  double *dc = nullptr, *dt = nullptr, *ds = nullptr, *dw = nullptr, *dcur = nullptr;
  int32_t* dsel = nullptr;
  uint32_t* dn = nullptr;
  GpuSelectJob* dj = nullptr;
  hipStream_t q = nullptr;
  auto release = [&]() {
    (void)hipFree(dc);
    (void)hipFree(dt);
    (void)hipFree(ds);
    (void)hipFree(dw);
    (void)hipFree(dcur);
    (void)hipFree(dsel);
    (void)hipFree(dn);
    (void)hipFree(dj);
    if (q) (void)hipStreamDestroy(q);
  };
  if (!hip_ok(hipStreamCreate(&q)) || !hip_ok(hipMalloc(&dc, column_doubles * 8)) ||
      !hip_ok(hipMalloc(&dt, target_doubles * 8)) || !hip_ok(hipMalloc(&ds, scratch_doubles * 8)) ||
      !hip_ok(hipMalloc(&dw, size_t(job_count) * K * 8)) || !hip_ok(hipMalloc(&dcur, size_t(job_count) * 8)) ||
      !hip_ok(hipMalloc(&dsel, size_t(job_count) * K * 4)) || !hip_ok(hipMalloc(&dn, size_t(job_count) * 4)) ||
      !hip_ok(hipMalloc(&dj, size_t(job_count) * sizeof(GpuSelectJob)))) {
    release();
    return fail(JGS2_ERROR_MEMORY, "Cubature precompute: device allocation for the selection failed.");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  bool ok = hip_ok(hipMemcpyAsync(dc, columns, column_doubles * 8, hipMemcpyHostToDevice, q)) &&
            hip_ok(hipMemcpyAsync(dt, targets, target_doubles * 8, hipMemcpyHostToDevice, q)) &&
            hip_ok(hipMemcpyAsync(dj, jobs, size_t(job_count) * sizeof(GpuSelectJob), hipMemcpyHostToDevice, q));
  (void)hipEventRecord(e0, q);
  if (ok && K <= 6)  // smaller private arrays and mask table for the usual K
    hipLaunchKernelGGL(k_greedy_select<6>, dim3(job_count), dim3(64), 0, q, dc, dt, dj, rows, K, tol, passes, ds, dsel,
                       dw, dcur, dn);
  else if (ok)
    hipLaunchKernelGGL(k_greedy_select<8>, dim3(job_count), dim3(64), 0, q, dc, dt, dj, rows, K, tol, passes, ds, dsel,
                       dw, dcur, dn);
  (void)hipEventRecord(e1, q);
  ok = ok && hip_ok(hipGetLastError()) &&
       hip_ok(hipMemcpyAsync(out_sel, dsel, size_t(job_count) * K * 4, hipMemcpyDeviceToHost, q)) &&
       hip_ok(hipMemcpyAsync(out_w, dw, size_t(job_count) * K * 8, hipMemcpyDeviceToHost, q)) &&
       hip_ok(hipMemcpyAsync(out_cur, dcur, size_t(job_count) * 8, hipMemcpyDeviceToHost, q)) &&
       hip_ok(hipMemcpyAsync(out_n, dn, size_t(job_count) * 4, hipMemcpyDeviceToHost, q)) &&
       hip_ok(hipStreamSynchronize(q));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  release();
  if (milliseconds) *milliseconds = ms;
  if (!ok) return fail(JGS2_ERROR_DEVICE, "Cubature precompute: the device selection failed.");
  return JGS2_OK;
}

}  // namespace jgs2
