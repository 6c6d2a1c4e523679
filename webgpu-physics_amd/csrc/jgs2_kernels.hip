// HIP kernels of the JGS2 step for gfx950 (MI355X).
//
// One kernel per pass of the reference's frame (src/simulation/gpu/
// jgs2-solver.ts:2207-2494). Each kernel cites the WGSL entry point it
// implements in src/simulation/gpu/jgs2-shader.ts ("S:line"). The arithmetic
// is FP32 like the reference. Where a closed form is exactly equal in real
// arithmetic to the reference's literal expression and cheaper, the kernel
// uses the closed form (stated at each site); results then agree with the
// reference to f32 rounding, not bit for bit.
//
// Launch shape: 256 threads (four 64-wide wavefronts) per block; vertex and
// element kernels map blocks to contiguous index ranges per XCD so that the
// neighbors a vertex gathers share one XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "jgs2_device.h"

namespace jgs2 {
namespace {

constexpr uint32_t kBlock = 256;
constexpr float kF32Max = 3.0e38f;
constexpr float kUnitRoundoff = 0.000000059604644775390625f;
constexpr float kPositionResolutionMultiplier = 8.0f;
constexpr float kRelativeEigenvalueFloor = 0.0000152587890625f;
constexpr float kMaxNormalizedShift = 0.001f;
constexpr float kArmijoC1 = 0.0001f;
constexpr uint32_t kMaxBacktracks = 12;
constexpr uint32_t kTrialCount = 13;
constexpr float kDeterminantFloor = 0.0001f;
constexpr float kTwoPiOverThree = 2.0943951023931953f;
constexpr uint32_t kEmptyTet = 0xffffffffu;

enum : uint32_t {
  kAccepted = 0, kPinned = 1, kZeroGradient = 2, kShiftLimit = 3, kNonDescent = 4,
  kLineSearchFailed = 5, kNonfinite = 6, kCholeskyFailed = 7,
};
enum : uint32_t {
  kSourceGeometryValid = 1u, kCandidateGeometryValid = 2u, kAcceptedMinimumValid = 4u,
  kLocalNumericsValid = 8u, kValidityMask = 15u, kStopLatch = 16u, kControlMask = 31u,
  kSourceEnergyValid = 16u, kCandidateEnergyValid = 32u, kReductionMask = 1u | 2u | 8u | 16u | 32u,
};
enum : uint32_t { kForce = 1u, kTarget = 2u };

// ------------------------------------------------------------------ small math
// Vector sums of products are fused multiply-adds written out by the types
// themselves: `s * a` is kept unevaluated (sf3) until it meets an addition,
// which then issues one fmaf per component. The rounding of every
// expression is therefore fixed by the source, identical in every kernel
// variant (the device code is compiled with -ffp-contract=on, which fuses
// only within one scalar expression), and a 3x3 product costs 27 fused
// operations instead of 27 multiplies and 18 adds.
struct f3 {
  float x, y, z;
};
struct sf3 {  // s * a, not yet rounded
  float s;
  f3 a;
  __device__ __forceinline__ operator f3() const { return f3{s * a.x, s * a.y, s * a.z}; }
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ f3 fma3(float s, f3 a, f3 b) {  // s * a + b, one rounding per component
  return f3{fmaf(s, a.x, b.x), fmaf(s, a.y, b.y), fmaf(s, a.z, b.z)};
}
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ sf3 operator*(float s, f3 a) { return sf3{s, a}; }
__device__ __forceinline__ sf3 operator*(f3 a, float s) { return sf3{s, a}; }
__device__ __forceinline__ sf3 operator-(sf3 m) { return sf3{-m.s, m.a}; }
__device__ __forceinline__ f3 operator+(f3 b, sf3 m) { return fma3(m.s, m.a, b); }
__device__ __forceinline__ f3 operator+(sf3 m, f3 b) { return fma3(m.s, m.a, b); }
__device__ __forceinline__ f3 operator+(sf3 m, sf3 n) { return fma3(n.s, n.a, f3(m)); }
__device__ __forceinline__ f3 operator-(f3 b, sf3 m) { return fma3(-m.s, m.a, b); }
__device__ __forceinline__ f3 operator-(sf3 m, f3 b) { return fma3(m.s, m.a, -b); }
__device__ __forceinline__ f3 operator-(sf3 m, sf3 n) { return fma3(-n.s, n.a, f3(m)); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Column-major 3x3 like WGSL mat3x3f.
struct m3 {
  f3 c0, c1, c2;
};
__device__ __forceinline__ m3 mzero() { return {mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)}; }
__device__ __forceinline__ m3 mid() { return {mk(1, 0, 0), mk(0, 1, 0), mk(0, 0, 1)}; }
__device__ __forceinline__ f3 mul(const m3& m, f3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }
__device__ __forceinline__ m3 mul(const m3& a, const m3& b) { return {mul(a, b.c0), mul(a, b.c1), mul(a, b.c2)}; }
__device__ __forceinline__ m3 tr(const m3& m) {
  return {mk(m.c0.x, m.c1.x, m.c2.x), mk(m.c0.y, m.c1.y, m.c2.y), mk(m.c0.z, m.c1.z, m.c2.z)};
}
__device__ __forceinline__ f3 tmul(const m3& m, f3 v) {  // transpose(m) * v
  return mk(dot(m.c0, v), dot(m.c1, v), dot(m.c2, v));
}
__device__ __forceinline__ m3 madd(const m3& a, const m3& b) { return {a.c0 + b.c0, a.c1 + b.c1, a.c2 + b.c2}; }
__device__ __forceinline__ m3 msub(const m3& a, const m3& b) { return {a.c0 - b.c0, a.c1 - b.c1, a.c2 - b.c2}; }
__device__ __forceinline__ m3 mscale(float s, const m3& a) { return {s * a.c0, s * a.c1, s * a.c2}; }
__device__ __forceinline__ m3 mfma(float s, const m3& a, const m3& b) {  // s * a + b
  return {fma3(s, a.c0, b.c0), fma3(s, a.c1, b.c1), fma3(s, a.c2, b.c2)};
}
__device__ __forceinline__ float frob(const m3& a, const m3& b) {
  return fmaf(a.c2.z, b.c2.z, fmaf(a.c2.y, b.c2.y, fmaf(a.c2.x, b.c2.x, fmaf(a.c1.z, b.c1.z,
         fmaf(a.c1.y, b.c1.y, fmaf(a.c1.x, b.c1.x, fmaf(a.c0.z, b.c0.z, fmaf(a.c0.y, b.c0.y, a.c0.x * b.c0.x))))))));
}
__device__ __forceinline__ float mdet(const m3& f) { return dot(f.c0, cross(f.c1, f.c2)); }
__device__ __forceinline__ m3 cofactor(const m3& f) { return {cross(f.c1, f.c2), cross(f.c2, f.c0), cross(f.c0, f.c1)}; }

__device__ __forceinline__ bool finite_s(float v) { return v == v && fabsf(v) <= kF32Max; }
__device__ __forceinline__ bool finite3(f3 v) { return finite_s(v.x) && finite_s(v.y) && finite_s(v.z); }
__device__ __forceinline__ bool finite4(float4 v) {
  return finite_s(v.x) && finite_s(v.y) && finite_s(v.z) && finite_s(v.w);
}
__device__ __forceinline__ bool finite_m(const m3& m) { return finite3(m.c0) && finite3(m.c1) && finite3(m.c2); }
__device__ __forceinline__ bool finite_nonneg_int(float v, float maximum) {  // S:919-922
  return finite_s(v) && v >= 0.0f && v <= maximum && v == floorf(v);
}
__device__ __forceinline__ uint32_t control_bits(float packed) {  // S:888-898
  if (!finite_s(packed) || packed < 0.0f || packed > float(kControlMask) || packed != floorf(packed)) return 0u;
  return uint32_t(packed);
}
__device__ __forceinline__ float scaled_length3(f3 v) {  // jgs2-globalization-wgsl.ts:287-294
  float mc = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
  if (!(mc > 0.0f)) return 0.0f;
  f3 n = v / mc;
  return mc * sqrtf(dot(n, n));
}

// XCD-aware block mapping: consecutive logical blocks land on one XCD so
// that vertices sharing tetrahedra share an L2. Bijective for any grid.
__device__ __forceinline__ uint32_t logical_block(uint32_t remap) {
  const uint32_t grid = gridDim.x, bid = blockIdx.x;
  if (!remap) return bid;
  const uint32_t q = grid / 8u, r = grid % 8u;
  const uint32_t xcd = bid % 8u, j = bid / 8u;
  const uint32_t base = xcd < r ? xcd * (q + 1u) : r * (q + 1u) + (xcd - r) * q;
  return base + j;
}

// Vertex frames the local solve reads: those of this iteration's polar
// pass, or the ones the previous iteration's convergence pass computed at
// the same positions.
__device__ __forceinline__ const float4* rot_source(const DeviceState& s, const PassParams& p) {
  return p.fused_rot ? s.vrot_next : s.vrot;
}

__device__ __forceinline__ bool stop_latched(const DeviceState& s, const PassParams& p, uint32_t sc) {  // S:909-917
  if (!p.stop_enabled) return false;
  return (control_bits(s.control[4 * size_t(sc)].w) & kStopLatch) != 0u;
}
// Whether a lane sweep with kind filter `f` solves scene `sc` (0: every
// scene; 1: scenes without cloth or contact; 2: the others with
// tetrahedral terms at an unpinned vertex; 3: those of them that hold
// cubature records; 4: the others without; 5: every other, colored sweeps
// only).
// Bit k of a filter's mask selects the scenes of kind k (a set of the four
// kKind bits): without cloth or contact 0x1111, with cubature 0xF0F0, with
// tetrahedral terms 0xFF00.
// A Jacobi sweep over a kind_blocks list visits only that kind's blocks.
__device__ __forceinline__ bool kind_selected(const DeviceState& s, const PassParams& p, uint32_t sc) {
  const uint32_t f = p.kind_filter;
  if (f == 0u || !p.colored) return true;
  static_assert(kKindCloth == 1u && kKindContact == 2u && kKindCubature == 4u && kKindTets == 8u, "kind bits");
  const uint32_t m = f == 1u ? 0x1111u : f == 2u ? 0xEE00u : f == 3u ? 0xE0E0u : f == 4u ? 0x00EEu : 0xEEEEu;
  return ((m >> (s.scene_kind[sc] & 15u)) & 1u) != 0u;
}
// Thread index of a Jacobi lane sweep over vertex blocks: with a kind
// filter the launch covers only the blocks of that kind_blocks list
// (`per` launch blocks of `threads` threads per 256-vertex block).
__device__ __forceinline__ uint32_t lane_sweep_gid(const DeviceState& s, const PassParams& p, uint32_t lb,
                                                   uint32_t threads, uint32_t lanes) {
  if (p.kind_filter == 0u || p.colored) return lb * threads + threadIdx.x;
  const uint32_t per = kSceneBlock * lanes / threads;
  const uint32_t f = p.kind_filter;
  const uint32_t first = (f > 1u ? p.kind_blocks1 : 0u) + (f > 2u ? p.kind_blocks2 : 0u) + (f > 3u ? p.kind_blocks3 : 0u);
  return s.kind_blocks[first + lb / per] * (kSceneBlock * lanes) + (lb % per) * threads + threadIdx.x;
}
// Scene of a vertex (kNoScene in free arena blocks).
__device__ __forceinline__ uint32_t vscene(const DeviceState& s, uint32_t v) { return s.vblock_scene[v / kSceneBlock]; }

// ------------------------------------------------------------------ element access
struct Tet {
  uint4 ix;
  m3 inv;           // Dm^-1 columns
  float vol, lam, mu;
};
__device__ __forceinline__ Tet load_tet(const DeviceState& s, uint32_t t) {
  // One 64-byte record: a single cache-line gather per tetrahedron.
  const uint4* r = s.tet_rec + 4 * size_t(t);
  Tet e;
  e.ix = r[0];
  const uint4 a = r[1], b = r[2], c = r[3];
  e.inv = {mk(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)),
           mk(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z)),
           mk(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z))};
  e.vol = __uint_as_float(a.w);
  e.lam = __uint_as_float(b.w);
  e.mu = __uint_as_float(c.w);
  return e;
}
// Shape gradients: rows of Dm^-1 and their negative sum (S:1080-1088).
__device__ __forceinline__ f3 shape_gradient(const Tet& e, uint32_t slot) {
  f3 g1 = mk(e.inv.c0.x, e.inv.c1.x, e.inv.c2.x);
  f3 g2 = mk(e.inv.c0.y, e.inv.c1.y, e.inv.c2.y);
  f3 g3 = mk(e.inv.c0.z, e.inv.c1.z, e.inv.c2.z);
  if (slot == 1) return g1;
  if (slot == 2) return g2;
  if (slot == 3) return g3;
  return -g1 - g2 - g3;
}
__device__ __forceinline__ uint32_t local_slot(uint4 ix, uint32_t v) {  // S:1045-1051
  if (ix.x == v) return 0;
  if (ix.y == v) return 1;
  if (ix.z == v) return 2;
  if (ix.w == v) return 3;
  return 4;
}
__device__ __forceinline__ f3 load_pos(const float4* pos, uint32_t v) { return xyz(pos[v]); }
// F = Ds Dm^-1 (S:1062-1074).
__device__ __forceinline__ m3 deformation(const Tet& e, const float4* pos) {
  f3 p0 = load_pos(pos, e.ix.x);
  f3 e1 = load_pos(pos, e.ix.y) - p0, e2 = load_pos(pos, e.ix.z) - p0, e3 = load_pos(pos, e.ix.w) - p0;
  m3 ds = {e1, e2, e3};
  return mul(ds, e.inv);
}
// F with the pose of vertex v taken from a register instead of memory
// (the same value: an incident tetrahedron's gather skips one load).
__device__ __forceinline__ f3 pos_or(const float4* pos, uint32_t i, uint32_t v, f3 xv) {
  return i == v ? xv : load_pos(pos, i);
}
__device__ __forceinline__ m3 deformation_at(const Tet& e, const float4* pos, uint32_t v, f3 xv) {
  f3 p0 = pos_or(pos, e.ix.x, v, xv);
  f3 e1 = pos_or(pos, e.ix.y, v, xv) - p0, e2 = pos_or(pos, e.ix.z, v, xv) - p0, e3 = pos_or(pos, e.ix.w, v, xv) - p0;
  m3 ds = {e1, e2, e3};
  return mul(ds, e.inv);
}
__device__ __forceinline__ m3 load_rot(const float4* r, uint32_t i) {
  return {xyz(r[3 * size_t(i)]), xyz(r[3 * size_t(i) + 1]), xyz(r[3 * size_t(i) + 2])};
}
__device__ __forceinline__ void store_rot(float4* r, uint32_t i, const m3& m) {
  r[3 * size_t(i)] = make_float4(m.c0.x, m.c0.y, m.c0.z, 0.0f);
  r[3 * size_t(i) + 1] = make_float4(m.c1.x, m.c1.y, m.c1.z, 0.0f);
  r[3 * size_t(i) + 2] = make_float4(m.c2.x, m.c2.y, m.c2.z, 0.0f);
}
// The per-vertex cubature record (frame, pose, inertia residual and
// weight) in one 64-byte line; see DeviceState::vcub.
struct VCub {
  m3 R;
  f3 x, diff;  // diff = x - x_pred
  float w;     // m / (h^2 n); 0 for a pinned vertex
  bool pinned;
};
// Layout: {x, w}, {R.c0, diff.x}, {R.c1, diff.y}, {R.c2, diff.z}, so a
// sample tetrahedron's pose is one aligned load per vertex.
__device__ __forceinline__ VCub load_vcub(const float4* vc, uint32_t v) {
  const float4* r = vc + 4 * size_t(v);
  const float4 a = r[0], b = r[1], c = r[2], d = r[3];
  VCub o;
  o.R = {xyz(b), xyz(c), xyz(d)};
  o.x = xyz(a);
  o.diff = mk(b.w, c.w, d.w);
  o.pinned = a.w < 0.0f;
  o.w = o.pinned ? 0.0f : a.w;
  return o;
}
__device__ __forceinline__ f3 vcub_pose(const float4* vc, uint32_t v) { return xyz(vc[4 * size_t(v)]); }
// Inertia weight of a vertex's distributed share (S:2401-2408).
__device__ __forceinline__ float share_weight(const DeviceState& s, const PassParams& p, uint32_t v, uint4 inf) {
  const float n = float(inf.y > 1u ? inf.y : 1u);
  return s.vtx_rest[v].w * p.inv_dt2 / n;
}
__device__ __forceinline__ void store_vcub(const DeviceState& s, const PassParams& p, float4* vc, uint32_t v, uint4 inf,
                                           const m3& R, f3 x) {
  const f3 diff = x - xyz(s.predicted[v]);
  const float w = inf.z != 0u ? -1.0f : share_weight(s, p, v, inf);
  float4* r = vc + 4 * size_t(v);
  r[0] = make_float4(x.x, x.y, x.z, w);
  r[1] = make_float4(R.c0.x, R.c0.y, R.c0.z, diff.x);
  r[2] = make_float4(R.c1.x, R.c1.y, R.c1.z, diff.y);
  r[3] = make_float4(R.c2.x, R.c2.y, R.c2.z, diff.z);
}
// The record paired with this sweep's frames (rot_source).
__device__ __forceinline__ const float4* vcub_source(const DeviceState& s, const PassParams& p) {
  return p.fused_rot ? s.vcub[1] : s.vcub[0];
}

// ------------------------------------------------------------------ stable Neo-Hookean
// Invariants shared by density, stress and tangent (stable-neo-hookean-wgsl.ts:44-94).
struct SnhPoint {
  m3 F, cof;
  float J, den, shear, volscale;
};
__device__ __forceinline__ SnhPoint snh_point(const m3& F, float lam, float mu) {
  SnhPoint s;
  s.F = F;
  s.cof = cofactor(F);
  float i1 = frob(F, F);
  s.den = i1 + 1.0f;
  s.J = dot(F.c0, s.cof.c0);
  s.shear = mu * (1.0f - 1.0f / s.den);
  s.volscale = lam * (s.J - 1.0f) - 0.75f * mu;
  return s;
}
__device__ __forceinline__ float snh_energy(const m3& F, float lam, float mu) {  // SN:44-55
  float i1 = frob(F, F);
  float jm1 = mdet(F) - 1.0f;
  return 0.5f * mu * (i1 - 3.0f) - 0.75f * mu * jm1 + 0.5f * lam * jm1 * jm1 - 0.5f * mu * logf((i1 + 1.0f) * 0.25f);
}
// P * g = shear F g + volscale cof g (SN:57-69 applied to a vector).
__device__ __forceinline__ f3 piola_times(const SnhPoint& s, f3 g) {
  return s.shear * mul(s.F, g) + s.volscale * mul(s.cof, g);
}
// ------------------------------------------------------------------ globalization (jgs2-globalization-wgsl.ts)
struct Sym3 {
  float a00, a11, a22, a01, a02, a12;
};
__device__ __forceinline__ float frobenius_over_sqrt_three(const Sym3& m) {  // GL:136-162
  float me = fmaxf(fmaxf(fmaxf(fabsf(m.a00), fabsf(m.a11)), fabsf(m.a22)),
                   fmaxf(fmaxf(fabsf(m.a01), fabsf(m.a02)), fabsf(m.a12)));
  if (!(me > 0.0f)) return 0.0f;
  float a00 = m.a00 / me, a11 = m.a11 / me, a22 = m.a22 / me, a01 = m.a01 / me, a02 = m.a02 / me, a12 = m.a12 / me;
  float sq = a00 * a00 + a11 * a11 + a22 * a22 + 2.0f * (a01 * a01 + a02 * a02 + a12 * a12);
  return me * sqrtf(sq / 3.0f);
}
__device__ __forceinline__ float min_eigenvalue(const Sym3& m) {  // GL:165-208
  float off = m.a01 * m.a01 + m.a02 * m.a02 + m.a12 * m.a12;
  if (off == 0.0f) return fminf(m.a00, fminf(m.a11, m.a22));
  float mean = (m.a00 + m.a11 + m.a22) / 3.0f;
  float d00 = m.a00 - mean, d11 = m.a11 - mean, d22 = m.a22 - mean;
  float centered = d00 * d00 + d11 * d11 + d22 * d22 + 2.0f * off;
  float spectral = sqrtf(centered / 6.0f);
  if (!(spectral > 0.0f)) return fminf(m.a00, fminf(m.a11, m.a22));
  m3 nc = {mk(d00, m.a01, m.a02) / spectral, mk(m.a01, d11, m.a12) / spectral, mk(m.a02, m.a12, d22) / spectral};
  float hd = fminf(fmaxf(0.5f * mdet(nc), -1.0f), 1.0f);
  float angle = acosf(hd) / 3.0f;
  float largest = mean + 2.0f * spectral * cosf(angle);
  float smallest = mean + 2.0f * spectral * cosf(angle + kTwoPiOverThree);
  float middle = 3.0f * mean - largest - smallest;
  return fminf(largest, fminf(middle, smallest));
}
struct SolveOut {
  f3 dir;
  float gdotp, scale, min_eig, shift, residual;
  uint32_t status;
};
// jgs2_globalization_solve (GL:328-492); `H` is the accumulated local Hessian.
__device__ SolveOut globalization_solve(const m3& H, f3 g, float inertia_scale) {
  SolveOut o{mk(0, 0, 0), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, kNonfinite};
  // classify_shift (GL:223-285)
  if (!finite_m(H) || !finite_s(inertia_scale) || !(inertia_scale > 0.0f)) return o;
  Sym3 s{H.c0.x, H.c1.y, H.c2.z, 0.5f * (H.c0.y + H.c1.x), 0.5f * (H.c0.z + H.c2.x), 0.5f * (H.c1.z + H.c2.y)};
  float lad = fmaxf(fabsf(s.a00), fmaxf(fabsf(s.a11), fabsf(s.a22)));
  float scale = fmaxf(inertia_scale, fmaxf(lad, frobenius_over_sqrt_three(s)));
  if (!finite_s(scale) || !(scale > 0.0f)) return o;
  Sym3 n{s.a00 / scale, s.a11 / scale, s.a22 / scale, s.a01 / scale, s.a02 / scale, s.a12 / scale};
  float me = min_eigenvalue(n);
  if (!finite_s(me)) return o;
  float shift = fmaxf(0.0f, kRelativeEigenvalueFloor - me);
  o.scale = scale;
  o.min_eig = me;
  o.shift = shift;
  if (shift > kMaxNormalizedShift) {
    o.status = kShiftLimit;
    return o;
  }
  if (!finite3(g)) {
    o.status = kNonfinite;
    return o;
  }
  float gs = fmaxf(fabsf(g.x), fmaxf(fabsf(g.y), fabsf(g.z)));
  if (!(gs > 0.0f)) {
    o.status = kZeroGradient;
    return o;
  }
  float a00 = n.a00 + shift, a11 = n.a11 + shift, a22 = n.a22 + shift;
  f3 rhs = (-g) / scale;
  if (!finite3(rhs)) {
    o.status = kNonfinite;
    return o;
  }
  o.status = kCholeskyFailed;
  if (!finite_s(a00) || !(a00 > 0.0f)) return o;
  float l00 = sqrtf(a00);
  float l10 = n.a01 / l00, l20 = n.a02 / l00;
  float p1 = a11 - l10 * l10;
  if (!finite_s(p1) || !(p1 > 0.0f)) return o;
  float l11 = sqrtf(p1);
  float l21 = (n.a12 - l20 * l10) / l11;
  float p2 = a22 - l20 * l20 - l21 * l21;
  if (!finite_s(p2) || !(p2 > 0.0f)) return o;
  float l22 = sqrtf(p2);
  float y0 = rhs.x / l00;
  float y1 = (rhs.y - l10 * y0) / l11;
  float y2 = (rhs.z - l20 * y0 - l21 * y1) / l22;
  float x2 = y2 / l22;
  float x1 = (y1 - l21 * x2) / l11;
  float x0 = (y0 - l10 * x1 - l20 * x2) / l00;
  f3 d = mk(x0, x1, x2);
  if (!finite3(d)) return o;
  float gd = dot(g, d);
  if (!finite_s(gd)) {
    o.status = kNonDescent;
    return o;
  }
  if (!(gd < 0.0f)) {
    o.gdotp = gd;
    o.status = kNonDescent;
    return o;
  }
  f3 r = mk(a00 * d.x + n.a01 * d.y + n.a02 * d.z, n.a01 * d.x + a11 * d.y + n.a12 * d.z,
            n.a02 * d.x + n.a12 * d.y + a22 * d.z) - rhs;
  float res = scaled_length3(r) / fmaxf(1.0f, scaled_length3(rhs));
  o.gdotp = gd;
  if (!finite_s(res)) {
    o.status = kCholeskyFailed;
    return o;
  }
  o.dir = d;
  o.residual = res;
  o.status = kAccepted;
  return o;
}
__device__ __forceinline__ float log_one_plus(float v) {  // GL:496-506
  if (fabsf(v) <= 0.001f) return v * (1.0f + v * (-0.5f + v * (0.3333333333333333f + v * (-0.25f + 0.2f * v))));
  return logf(1.0f + v);
}
// Cancellation-safe SNH density difference for F1 = F0 + D (GL:510-559).
__device__ __forceinline__ float snh_delta_general(const m3& i, const m3& dd, float lam, float mu) {
  m3 f = madd(i, dd);
  float ii = frob(i, i);
  float id = frob(dd, madd(f, i));
  float jd = dot(dd.c0, cross(i.c1, i.c2)) + dot(i.c0, cross(dd.c1, i.c2)) + dot(i.c0, cross(i.c1, dd.c2)) +
             dot(dd.c0, cross(dd.c1, i.c2)) + dot(dd.c0, cross(i.c1, dd.c2)) + dot(i.c0, cross(dd.c1, dd.c2)) +
             dot(dd.c0, cross(dd.c1, dd.c2));
  float ij = mdet(i), fj = mdet(f);
  float lr = log_one_plus(id / (ii + 1.0f));
  return 0.5f * mu * id - 0.75f * mu * jd + 0.5f * lam * jd * (fj + ij - 2.0f) - 0.5f * mu * lr;
}

// ------------------------------------------------------------------ per-vertex helpers
__device__ __forceinline__ float dist_inertia_weight(const DeviceState& s, const PassParams& p, uint32_t v) {  // S:2401-2408
  uint4 inf = s.vtx_info[v];
  if (inf.z != 0u) return 0.0f;
  uint32_t n = inf.y > 1u ? inf.y : 1u;
  return s.vtx_rest[v].w * p.inv_dt2 / float(n);
}
__device__ __forceinline__ f3 objective_gradient(const DeviceState& s, const PassParams& p, uint32_t v, f3 x) {  // S:1926-1938
  f3 r = mk(0, 0, 0);
  if (p.objective_flags & kForce) r = r - xyz(s.objectives[2 * size_t(v)]);
  if (p.objective_flags & kTarget) {
    float4 t = s.objectives[2 * size_t(v) + 1];
    if (t.w > 0.0f) r = r + t.w * (x - xyz(t));
  }
  return r;
}
__device__ __forceinline__ float target_stiffness(const DeviceState& s, const PassParams& p, uint32_t v) {
  return (p.objective_flags & kTarget) ? s.objectives[2 * size_t(v) + 1].w : 0.0f;
}
__device__ __forceinline__ float quadratic_delta(f3 base, f3 target, f3 d, float w) {  // S:1891-1901
  f3 r = base - target;
  return w * (dot(r, d) + 0.5f * dot(d, d));
}
__device__ __forceinline__ float objective_delta(const DeviceState& s, const PassParams& p, uint32_t v, f3 base, f3 d) {  // S:1903-1924
  float r = 0.0f;
  if (p.objective_flags & kForce) r -= dot(xyz(s.objectives[2 * size_t(v)]), d);
  if (p.objective_flags & kTarget) {
    float4 t = s.objectives[2 * size_t(v) + 1];
    if (t.w > 0.0f) r += quadratic_delta(base, xyz(t), d, t.w);
  }
  return r;
}
__device__ __forceinline__ float floor_delta(const PassParams& p, float by, float dy) {  // S:1944-1956
  if (!(p.floor_k > 0.0f)) return 0.0f;
  float ip = fminf(0.0f, by - p.floor_y);
  float fp = fminf(0.0f, by + dy - p.floor_y);
  return 0.5f * p.floor_k * (fp - ip) * (fp + ip);
}
// ------------------------------------------------------------------ cloth (S:1225-1845)
// StVK membrane triangles and quadratic dihedral hinges on the shared vertex
// state. `pos` is read with vertex `ov` replaced by `op`, the restricted
// trial's moved vertex (S:1520-1531); kNoVertex reads `pos` unchanged.
constexpr uint32_t kNoVertex = 0xffffffffu;
__device__ __forceinline__ f3 cloth_pos(const float4* pos, uint32_t v, uint32_t ov, f3 op) {
  return v == ov ? op : load_pos(pos, v);
}
__device__ __forceinline__ float len3(f3 v) { return sqrtf(dot(v, v)); }
__device__ __forceinline__ void tri_shape_gradient(float4 inv, uint32_t l, float& x, float& y) {  // S:1533-1544
  if (l == 1) {
    x = inv.x;
    y = inv.y;
  } else if (l == 2) {
    x = inv.z;
    y = inv.w;
  } else {
    x = -inv.x - inv.z;
    y = -inv.y - inv.w;
  }
}
__device__ __forceinline__ uint32_t tri_slot(uint4 ix, uint32_t v) {  // S:1505-1510
  return ix.x == v ? 0u : (ix.y == v ? 1u : (ix.z == v ? 2u : 3u));
}
__device__ __forceinline__ uint32_t hinge_slot(uint4 ix, uint32_t v) {  // S:1512-1518
  return ix.x == v ? 0u : (ix.y == v ? 1u : (ix.z == v ? 2u : (ix.w == v ? 3u : 4u)));
}
struct Membrane {
  float energy;
  f3 f0, f1, stress;  // deformation columns and the Green strain's StVK stress (xx, xy, yy)
};
__device__ Membrane membrane_at(const DeviceState& s, uint32_t tri, const float4* pos, uint32_t ov, f3 op) {  // S:1546-1604
  const uint4 ix = s.tri_ix[tri];
  const float4 inv = s.tri_inv[tri];
  const f3 x0 = cloth_pos(pos, ix.x, ov, op);
  const f3 e01 = cloth_pos(pos, ix.y, ov, op) - x0;
  const f3 e02 = cloth_pos(pos, ix.z, ov, op) - x0;
  Membrane m;
  m.f0 = e01 * inv.x + e02 * inv.z;
  m.f1 = e01 * inv.y + e02 * inv.w;
  const float e00 = 0.5f * (dot(m.f0, m.f0) - 1.0f);
  const float e01s = 0.5f * dot(m.f0, m.f1);
  const float e11 = 0.5f * (dot(m.f1, m.f1) - 1.0f);
  const float trace = e00 + e11;
  const float mu = s.cloth_mu, lam = s.cloth_lam;
  m.stress = mk(2.0f * mu * e00 + lam * trace, 2.0f * mu * e01s, 2.0f * mu * e11 + lam * trace);
  const float density = mu * (e00 * e00 + 2.0f * e01s * e01s + e11 * e11) + 0.5f * lam * trace * trace;
  m.energy = (s.tri_area[tri] * s.cloth_thick) * density;
  return m;
}
__device__ __forceinline__ f3 membrane_gradient(const DeviceState& s, uint32_t tri, const Membrane& m, uint32_t slot) {
  float sx, sy;
  tri_shape_gradient(s.tri_inv[tri], slot, sx, sy);
  const f3 p0 = m.f0 * m.stress.x + m.f1 * m.stress.y;
  const f3 p1 = m.f0 * m.stress.y + m.f1 * m.stress.z;
  return (s.tri_area[tri] * s.cloth_thick) * (p0 * sx + p1 * sy);
}
// Diagonal 3x3 block of the membrane Hessian for one corner, symmetrized
// (S:1608-1658).
__device__ m3 membrane_local_hessian(const DeviceState& s, uint32_t tri, const Membrane& ev, uint32_t slot) {
  float sx, sy;
  tri_shape_gradient(s.tri_inv[tri], slot, sx, sy);
  const float aw = s.tri_area[tri] * s.cloth_thick;
  const float mu = s.cloth_mu, lam = s.cloth_lam;
  const float f0c[3] = {ev.f0.x, ev.f0.y, ev.f0.z}, f1c[3] = {ev.f1.x, ev.f1.y, ev.f1.z};
  f3 col[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const f3 axis = mk(c == 0 ? 1.0f : 0.0f, c == 1 ? 1.0f : 0.0f, c == 2 ? 1.0f : 0.0f);
    const f3 df0 = axis * sx, df1 = axis * sy;
    const float de00 = sx * f0c[c];
    const float de01 = 0.5f * (sx * f1c[c] + sy * f0c[c]);
    const float de11 = sy * f1c[c];
    const float dtr = de00 + de11;
    const f3 ds = mk(2.0f * mu * de00 + lam * dtr, 2.0f * mu * de01, 2.0f * mu * de11 + lam * dtr);
    const f3 dp0 = df0 * ev.stress.x + df1 * ev.stress.y + ev.f0 * ds.x + ev.f1 * ds.y;
    const f3 dp1 = df0 * ev.stress.y + df1 * ev.stress.z + ev.f0 * ds.y + ev.f1 * ds.z;
    col[c] = aw * (dp0 * sx + dp1 * sy);
  }
  const m3 r = {col[0], col[1], col[2]};
  return mscale(0.5f, madd(r, tr(r)));
}
__device__ __forceinline__ f3 tri_raw_normal(const DeviceState& s, uint32_t tri, const float4* pos, uint32_t ov, f3 op) {  // S:1660-1686
  const uint4 ix = s.tri_ix[tri];
  const f3 x0 = cloth_pos(pos, ix.x, ov, op);
  return cross(cloth_pos(pos, ix.y, ov, op) - x0, cloth_pos(pos, ix.z, ov, op) - x0);
}
// Signed dihedral angle of triangles (0,1,2) and (1,0,3) and the gradient
// of one corner `slot` (4 = none). Validity covers every corner's gradient,
// like the reference, which forms all twelve entries (S:1700-1810).
struct Dihedral {
  float angle;
  bool valid;
  f3 grad;
};
__device__ Dihedral dihedral_at(const DeviceState& s, uint32_t h, const float4* pos, uint32_t ov, f3 op, uint32_t slot,
                                float4* store = nullptr) {
  const uint4 ix = s.hinge_ix[h];
  const f3 x0 = cloth_pos(pos, ix.x, ov, op), x1 = cloth_pos(pos, ix.y, ov, op);
  const f3 x2 = cloth_pos(pos, ix.z, ov, op), x3 = cloth_pos(pos, ix.w, ov, op);
  Dihedral out{0.0f, false, mk(0, 0, 0)};
  const f3 edge = x1 - x0;
  const float el = len3(edge);
  const f3 to2 = x2 - x0, rev = -edge, to3 = x3 - x1;
  const f3 rn0 = cross(edge, to2), rn1 = cross(rev, to3);
  const float nl0 = len3(rn0), nl1 = len3(rn1);
  const float es0 = fmaxf(el, fmaxf(len3(to2), len3(to2 - edge)));
  const float es1 = fmaxf(el, fmaxf(len3(to3), len3(to3 - rev)));
  const bool ok = finite_s(el) && el > 0.0f && finite_s(nl0) && finite_s(nl1) && nl0 > 1.0e-12f * es0 * es0 &&
                  nl1 > 1.0e-12f * es1 * es1;
  if (!ok) return out;
  const f3 tangent = edge / el, n0 = rn0 / nl0, n1 = rn1 / nl1;
  const f3 nc = cross(n0, n1);
  const float sine = dot(tangent, nc), cosine = dot(n0, n1);
  const float den = sine * sine + cosine * cosine;
  if (!finite_s(den) || !(den > 0.0f)) return out;
  bool fin = true;
  f3 keep = mk(0, 0, 0);
  // The twelve entries only decide validity when no corner gradient is
  // wanted, and then they are finite whenever el >= 1e-20: with unit
  // directions |dt| <= 2.01 / el, |dn0r| <= |to2| + |edge| <= 2 es0 and
  // nl0 >= max(1e-12 es0^2, smallest f32) give |dn0| <= 1.1e29 (likewise
  // dn1, es0 and es1 >= el), and den = sine^2 + cosine^2 of unit normals
  // is 1 to rounding, so every entry is below 1e30. Shorter edges form
  // them all, as the reference does.
  uint32_t corners = (slot < 4u || store != nullptr || !(el >= 1.0e-20f)) ? 4u : 0u;
  if (corners && el >= 1.0e-20f) {
    // The corner gradients in closed form, equal in exact arithmetic to the
    // twelve entries below (pinned against the FP64 oracle): with N0 = rn0,
    // N1 = rn1 and a_k = N_k / |N_k|^2, grad_2 = -|e| a_0, grad_3 = -|e| a_1,
    // grad_0 = ((x1 - x2).e a_0 + (x1 - x3).e a_1) / |e|,
    // grad_1 = ((x2 - x0).e a_0 + (x3 - x0).e a_1) / |e|. When all are finite
    // so are the entries (el >= 1e-20, above), and validity is unchanged;
    // otherwise the entries are formed as the reference does.
    const f3 a0 = rn0 / dot(rn0, rn0), a1 = rn1 / dot(rn1, rn1);
    const f3 g2 = (-el) * a0, g3 = (-el) * a1;
    const f3 g0 = (dot(x1 - x2, edge) / el) * a0 + (dot(x1 - x3, edge) / el) * a1;
    const f3 g1 = (dot(x2 - x0, edge) / el) * a0 + (dot(x3 - x0, edge) / el) * a1;
    if (finite3(g0) && finite3(g1) && finite3(g2) && finite3(g3)) {
      corners = 0u;
      keep = slot == 0u ? g0 : slot == 1u ? g1 : slot == 2u ? g2 : slot == 3u ? g3 : mk(0, 0, 0);
      if (store) {
        store[0] = make_float4(g0.x, g0.y, g0.z, 0.0f);
        store[1] = make_float4(g1.x, g1.y, g1.z, 0.0f);
        store[2] = make_float4(g2.x, g2.y, g2.z, 0.0f);
        store[3] = make_float4(g3.x, g3.y, g3.z, 0.0f);
      }
    }
  }
#pragma unroll 1
  for (uint32_t dv = 0; dv < corners; ++dv) {
    float gc[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const f3 u = mk(ax == 0 ? 1.0f : 0.0f, ax == 1 ? 1.0f : 0.0f, ax == 2 ? 1.0f : 0.0f);
      const f3 z = mk(0, 0, 0);
      const f3 d0 = dv == 0 ? u : z, d1 = dv == 1 ? u : z, d2 = dv == 2 ? u : z, d3 = dv == 3 ? u : z;
      const f3 de = d1 - d0;
      const f3 dt = (de - tangent * dot(tangent, de)) / el;
      const f3 dto2 = d2 - d0, dto3 = d3 - d1;
      const f3 dn0r = cross(de, to2) + cross(edge, dto2);
      const f3 dn1r = cross(-de, to3) + cross(rev, dto3);
      const f3 dn0 = (dn0r - n0 * dot(n0, dn0r)) / nl0;
      const f3 dn1 = (dn1r - n1 * dot(n1, dn1r)) / nl1;
      const float dcos = dot(dn0, n1) + dot(n0, dn1);
      const float dsin = dot(dt, nc) + dot(tangent, cross(dn0, n1) + cross(n0, dn1));
      gc[ax] = (cosine * dsin - sine * dcos) / den;
    }
    const f3 g = mk(gc[0], gc[1], gc[2]);
    fin = fin && finite3(g);
    if (dv == slot) keep = g;
    if (store) store[dv] = make_float4(g.x, g.y, g.z, 0.0f);
  }
  const float angle = atan2f(sine, cosine);
  if (!finite_s(angle) || !fin) return out;
  out.angle = angle;
  out.valid = true;
  out.grad = keep;
  return out;
}
__device__ __forceinline__ float wrapped_angle_difference(float a, float r) {  // S:1812-1815
  const float d = a - r;
  return atan2f(sinf(d), cosf(d));
}
__device__ __forceinline__ float bending_energy_of(const DeviceState& s, uint32_t h, const Dihedral& ev) {  // S:1817-1836
  if (!ev.valid) return kF32Max;
  const float4 r = s.hinge_rest[h];
  const float ad = wrapped_angle_difference(ev.angle, r.x);
  const float k = s.cloth_bend * r.y;
  return 0.5f * k * ad * ad;
}
__device__ __forceinline__ float bending_energy_at(const DeviceState& s, uint32_t h, const float4* pos, uint32_t ov, f3 op) {
  return bending_energy_of(s, h, dihedral_at(s, h, pos, ov, op, 4u));
}
// Element records at a sweep's source pose, computed once per element by
// k_cloth_base (p.cloth_base set) instead of once per incident vertex; the
// same functions produce them, so they equal the inline evaluations.
//   tri_base[3t]  {f0, energy}, {f1, source energy}, {stress, stretch test bits}
//   hinge_base[4h] {grad0, angle}, {grad1, valid}, {grad2, bending energy}, {grad3, source energy}
// The source energies are the energies the records held before a pass
// moved them to the target pose (k_cloth_base mode 1), for the energy pass.
// kRecords: the caller launches only with p.cloth_base set, so the inline
// evaluation is not compiled into it (and does not raise its register count).
template <bool kRecords = false>
__device__ __forceinline__ Membrane membrane_base(const DeviceState& s, const PassParams& p, uint32_t tri,
                                                  const float4* pos) {
  if (!kRecords && !p.cloth_base) return membrane_at(s, tri, pos, kNoVertex, mk(0, 0, 0));
  const float4* r = s.tri_base + 3 * size_t(tri);
  const float4 a = r[0], b = r[1], c = r[2];
  Membrane m;
  m.energy = a.w;
  m.f0 = xyz(a);
  m.f1 = xyz(b);
  m.stress = xyz(c);
  return m;
}
template <bool kRecords = false>
__device__ __forceinline__ Dihedral dihedral_base(const DeviceState& s, const PassParams& p, uint32_t h,
                                                  const float4* pos, uint32_t slot) {
  if (!kRecords && !p.cloth_base) return dihedral_at(s, h, pos, kNoVertex, mk(0, 0, 0), slot);
  const float4* r = s.hinge_base + 4 * size_t(h);
  Dihedral d;
  d.angle = r[0].w;
  d.valid = r[1].w != 0.0f;
  d.grad = d.valid && slot < 4u ? xyz(r[slot]) : mk(0, 0, 0);
  return d;
}
__device__ __forceinline__ float bending_base(const DeviceState& s, const PassParams& p, uint32_t h, const float4* pos) {
  if (!p.cloth_base) return bending_energy_at(s, h, pos, kNoVertex, mk(0, 0, 0));
  return s.hinge_base[4 * size_t(h) + 2].w;
}
// Membrane and bending contributions to one vertex's force (S:4028-4085).
__device__ __forceinline__ f3 cloth_gradient(const DeviceState& s, const PassParams& p, const float4* pos, uint32_t v) {
  f3 g = mk(0, 0, 0);
  for (uint32_t k = s.tri_row[v].x, ke = k + s.tri_row[v].y; k < ke; ++k) {
    const uint32_t tri = s.tri_elems[k];
    const uint32_t slot = tri_slot(s.tri_ix[tri], v);
    if (slot < 3) g = g + membrane_gradient(s, tri, membrane_base(s, p, tri, pos), slot);
  }
  for (uint32_t k = s.hinge_row[v].x, ke = k + s.hinge_row[v].y; k < ke; ++k) {
    const uint32_t h = s.hinge_elems[k];
    const uint32_t slot = hinge_slot(s.hinge_ix[h], v);
    if (slot >= 4) continue;
    const Dihedral d = dihedral_base(s, p, h, pos, slot);
    if (!d.valid) {
      g = g + mk(kF32Max, kF32Max, kF32Max);
    } else {
      const float4 r = s.hinge_rest[h];
      const float ad = wrapped_angle_difference(d.angle, r.x);
      g = g + (s.cloth_bend * r.y * ad) * d.grad;
    }
  }
  return g;
}

// ------------------------------------------------------------------ IPC contact (IC, S:296-874)
// Contact weights satisfy sum_i w_i x_i = point_a - point_b over a record's
// four vertices; VT records are (point, t0, t1, t2), EE (a0, a1, b0, b1)
// (ipc-contact-wgsl.ts:1-17). Bulk arithmetic is FP32 like the reference;
// the barrier's log term and the safe-step fraction are formed in FP64.
constexpr float kIpcLengthSquaredFloor = 1.0e-20f;
constexpr float kIpcParallelRelativeEpsilon = 1.0e-6f;
struct Contact {
  float distance;
  f3 normal;
  bool valid;
  float4 weights;
};
__device__ __forceinline__ f3 vmin3(f3 a, f3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ __forceinline__ f3 vmax3(f3 a, f3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
__device__ __forceinline__ float clamp01f(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
__device__ __forceinline__ f3 ipc_safe_unit(f3 v, f3 fb) {  // IC:56-73
  const float ls = dot(v, v);
  if (finite_s(ls) && ls > kIpcLengthSquaredFloor) return v * (1.0f / sqrtf(ls));
  const float fls = dot(fb, fb);
  if (finite_s(fls) && fls > kIpcLengthSquaredFloor) return fb * (1.0f / sqrtf(fls));
  return mk(1, 0, 0);
}
__device__ __forceinline__ f3 ipc_orthogonal_unit(f3 d) {  // IC:75-86
  const f3 a = mk(fabsf(d.x), fabsf(d.y), fabsf(d.z));
  f3 axis = mk(1, 0, 0);
  if (a.y <= a.x && a.y <= a.z) axis = mk(0, 1, 0);
  else if (a.z <= a.x && a.z <= a.y) axis = mk(0, 0, 1);
  return ipc_safe_unit(cross(d, axis), mk(1, 0, 0));
}
__device__ __forceinline__ Contact ipc_invalid() { return Contact{0.0f, mk(1, 0, 0), false, make_float4(0, 0, 0, 0)}; }
__device__ __forceinline__ Contact ipc_make_contact(f3 a, f3 b, float4 w, f3 fb) {  // IC:100-125
  if (!finite3(a) || !finite3(b) || !finite4(w)) return ipc_invalid();
  const f3 gap = a - b;
  const float sq = dot(gap, gap);
  if (!finite_s(sq) || sq < 0.0f) return ipc_invalid();
  return Contact{sqrtf(sq), ipc_safe_unit(gap, fb), true, w};
}
__device__ __forceinline__ void ipc_closest_on_segment(f3 q, f3 s0, f3 e, f3& pt, float& t, float& sq) {  // IC:127-141
  const f3 edge = e - s0;
  const float els = dot(edge, edge);
  t = 0.0f;
  if (finite_s(els) && els > kIpcLengthSquaredFloor) t = clamp01f(dot(q - s0, edge) / els);
  pt = s0 + t * edge;
  const f3 d = q - pt;
  sq = dot(d, d);
}
__device__ Contact ipc_point_triangle(f3 p, f3 t0, f3 t1, f3 t2) {  // IC:146-224
  const f3 e01 = t1 - t0, e02 = t2 - t0, e12 = t2 - t1;
  const float s01 = dot(e01, e01), s02 = dot(e02, e02), s12 = dot(e12, e12);
  const float largest = fmaxf(s01, fmaxf(s02, s12));
  const f3 an = cross(e01, e02);
  const float as = dot(an, an);
  const float threshold = fmaxf(kIpcLengthSquaredFloor * kIpcLengthSquaredFloor,
                                kIpcParallelRelativeEpsilon * largest * largest);
  f3 longest = e01;
  if (s02 > s01 && s02 >= s12) longest = e02;
  else if (s12 > s01 && s12 > s02) longest = e12;
  const f3 fb = ipc_safe_unit(an, ipc_orthogonal_unit(longest));
  if (finite_s(as) && as > threshold) {
    const float d01 = dot(e01, e02);
    const f3 po = p - t0;
    const float d20 = dot(po, e01), d21 = dot(po, e02);
    const float den = s01 * s02 - d01 * d01;
    const float b1 = (s02 * d20 - d01 * d21) / den;
    const float b2 = (s01 * d21 - d01 * d20) / den;
    const float b0 = 1.0f - b1 - b2;
    if (b0 >= 0.0f && b1 >= 0.0f && b2 >= 0.0f)
      return ipc_make_contact(p, b0 * t0 + b1 * t1 + b2 * t2, make_float4(1.0f, -b0, -b1, -b2), fb);
  }
  f3 p01, p12, p20;
  float u01, u12, u20, q01, q12, q20;
  ipc_closest_on_segment(p, t0, t1, p01, u01, q01);
  ipc_closest_on_segment(p, t1, t2, p12, u12, q12);
  ipc_closest_on_segment(p, t2, t0, p20, u20, q20);
  f3 best = p01;
  float bq = q01;
  f3 bary = mk(1.0f - u01, u01, 0.0f);
  if (q12 < bq) {
    best = p12;
    bq = q12;
    bary = mk(0.0f, 1.0f - u12, u12);
  }
  if (q20 < bq) {
    best = p20;
    bary = mk(u20, 0.0f, 1.0f - u20);
  }
  return ipc_make_contact(p, best, make_float4(1.0f, -bary.x, -bary.y, -bary.z), fb);
}
__device__ Contact ipc_edge_edge(f3 a0, f3 a1, f3 b0, f3 b1) {  // IC:228-308
  const f3 da = a1 - a0, db = b1 - b0, off = a0 - b0;
  const float la = dot(da, da), lb = dot(db, db);
  const float dthr = fmaxf(kIpcLengthSquaredFloor, kIpcParallelRelativeEpsilon * fmaxf(la, lb));
  const bool adeg = !(la > dthr), bdeg = !(lb > dthr);
  float pa = 0.0f, pb = 0.0f;
  if (adeg && !bdeg) {
    pb = clamp01f(dot(db, off) / lb);
  } else if (!adeg) {
    const float proj_a = dot(da, off);
    if (bdeg) {
      pa = clamp01f(-proj_a / la);
    } else {
      const float coupling = dot(da, db), proj_b = dot(db, off);
      const float den = la * lb - coupling * coupling;
      if (den > kIpcParallelRelativeEpsilon * la * lb) pa = clamp01f((coupling * proj_b - proj_a * lb) / den);
      const float pbu = (coupling * pa + proj_b) / lb;
      if (pbu < 0.0f) {
        pb = 0.0f;
        pa = clamp01f(-proj_a / la);
      } else if (pbu > 1.0f) {
        pb = 1.0f;
        pa = clamp01f((coupling - proj_a) / la);
      } else {
        pb = pbu;
      }
    }
  }
  const f3 longest = lb > la ? db : da;
  const f3 fb = ipc_safe_unit(cross(da, db), ipc_orthogonal_unit(longest));
  return ipc_make_contact(a0 + pa * da, b0 + pb * db, make_float4(1.0f - pa, pa, -(1.0f - pb), -pb), fb);
}
__device__ __forceinline__ bool ipc_barrier_active(float d, float dhat) {  // IC:310-314
  return finite_s(d) && finite_s(dhat) && d > 0.0f && d < dhat;
}
__device__ __forceinline__ float ipc_log_ratio(float d, float dhat) { return float(log(double(d) / double(dhat))); }
__device__ __forceinline__ float ipc_barrier(float d, float dhat) {  // IC:316-325
  if (!ipc_barrier_active(d, dhat)) return 0.0f;
  const float o = d - dhat;
  return -(o * o) * ipc_log_ratio(d, dhat);
}
__device__ __forceinline__ float ipc_barrier_d1(float d, float dhat) {  // IC:327-337
  if (!ipc_barrier_active(d, dhat)) return 0.0f;
  const float o = d - dhat;
  return -2.0f * o * ipc_log_ratio(d, dhat) - o * o / d;
}
__device__ __forceinline__ float ipc_barrier_d2(float d, float dhat) {  // IC:339-351
  if (!ipc_barrier_active(d, dhat)) return 0.0f;
  const float o = d - dhat;
  return -2.0f * ipc_log_ratio(d, dhat) - 4.0f * o / d + o * o / (d * d);
}
__device__ __forceinline__ float ipc_safe_step_cap(float ds, float dmin, float lip, float safety) {  // IC:398-422
  if (!finite_s(ds) || !finite_s(dmin) || !finite_s(lip) || !finite_s(safety) || dmin < 0.0f || ds <= dmin ||
      lip < 0.0f)
    return 0.0f;
  if (lip == 0.0f) return 1.0f;
  const double f = fmin(fmax(double(safety), 0.0), 1.0) * (double(ds) - double(dmin)) / double(lip);
  return float(fmin(fmax(f, 0.0), 1.0));
}
__device__ __forceinline__ float ipc_friction_f1(float slip, float eps) {  // IC:439-446
  const float s = fmaxf(slip, 0.0f);
  if (!(eps > 0.0f) || s >= eps) return s > 0.0f ? 1.0f : 0.0f;
  const float n = s / eps;
  return n * (2.0f - n);
}
__device__ __forceinline__ float ipc_friction_dissipation(float slip, float eps, float mu, float force) {  // IC:448-472
  if (!finite_s(slip) || !finite_s(eps) || !finite_s(mu) || !finite_s(force)) return 0.0f;
  const float s = fmaxf(slip, 0.0f);
  const float scale = fmaxf(mu, 0.0f) * fmaxf(force, 0.0f);
  if (!(eps > 0.0f)) return scale * s;
  if (s < eps) {
    const float n = s / eps;
    return scale * eps * n * n * (1.0f - n / 3.0f);
  }
  return scale * (s - eps / 3.0f);
}
__device__ __forceinline__ float ipc_rounding_reach(const PassParams& p, float scene_scale) {  // S:509-513
  return fmaxf(fmaxf(scene_scale, p.ipc_dhat), 1.0e-12f) * kUnitRoundoff * kPositionResolutionMultiplier;
}
__device__ __forceinline__ float ipc_aabb_lower_bound(bool vt, f3 p0, f3 p1, f3 p2, f3 p3) {  // S:525-550
  f3 gap;
  if (vt) {
    const f3 mn = vmin3(p1, vmin3(p2, p3)), mx = vmax3(p1, vmax3(p2, p3));
    gap = vmax3(vmax3(mn - p0, p0 - mx), mk(0, 0, 0));
  } else {
    const f3 mn0 = vmin3(p0, p1), mx0 = vmax3(p0, p1), mn1 = vmin3(p2, p3), mx1 = vmax3(p2, p3);
    gap = vmax3(vmax3(mn1 - mx0, mn0 - mx1), mk(0, 0, 0));
  }
  return sqrtf(dot(gap, gap));
}
// ipcContactFromPositions / ipcContactAtThreshold / ipcContactAtOverride
// (S:564-650): the AABB gap stands in for distances at or beyond `threshold`.
__device__ Contact ipc_contact_at(const DeviceState& s, uint32_t c, const float4* pos, float threshold,
                                  uint32_t mv = kNoVertex, f3 mp = f3{0, 0, 0}) {
  const uint4 ix = s.cand[c];
  const bool vt = (s.cand_meta[c] & 1u) == 0u;
  const f3 p0 = cloth_pos(pos, ix.x, mv, mp), p1 = cloth_pos(pos, ix.y, mv, mp);
  const f3 p2 = cloth_pos(pos, ix.z, mv, mp), p3 = cloth_pos(pos, ix.w, mv, mp);
  const float lb = ipc_aabb_lower_bound(vt, p0, p1, p2, p3);
  if (finite_s(lb) && lb >= threshold) return Contact{lb, mk(1, 0, 0), true, make_float4(0, 0, 0, 0)};
  return vt ? ipc_point_triangle(p0, p1, p2, p3) : ipc_edge_edge(p0, p1, p2, p3);
}
__device__ __forceinline__ float ipc_weight_for_vertex(uint4 ix, float4 w, uint32_t v) {  // S:652-663
  float r = 0.0f;
  if (ix.x == v) r += w.x;
  if (ix.y == v) r += w.y;
  if (ix.z == v) r += w.z;
  if (ix.w == v) r += w.w;
  return r;
}
__device__ __forceinline__ f3 ipc_weighted_displacement(const DeviceState& s, uint4 ix, float4 w, const float4* pos,
                                                        uint32_t mv, f3 mp) {  // S:665-705
  f3 r = mk(0, 0, 0);
  r = r + w.x * (cloth_pos(pos, ix.x, mv, mp) - load_pos(s.old, ix.x));
  r = r + w.y * (cloth_pos(pos, ix.y, mv, mp) - load_pos(s.old, ix.y));
  r = r + w.z * (cloth_pos(pos, ix.z, mv, mp) - load_pos(s.old, ix.z));
  r = r + w.w * (cloth_pos(pos, ix.w, mv, mp) - load_pos(s.old, ix.w));
  return r;
}
__device__ __forceinline__ void ipc_store_cached(const DeviceState& s, uint32_t c, bool target, const Contact& k) {  // S:386-419
  const bool valid = k.valid && finite_s(k.distance) && k.distance >= 0.0f && finite3(k.normal) && finite4(k.weights);
  (target ? s.tgt_c : s.src_c)[c] =
      valid ? make_float4(k.normal.x, k.normal.y, k.normal.z, k.distance) : make_float4(0, 0, 0, -1);
  (target ? s.tgt_w : s.src_w)[c] = valid ? k.weights : make_float4(0, 0, 0, 0);
}
__device__ __forceinline__ Contact ipc_cached(const DeviceState& s, uint32_t c, bool target) {  // S:421-448
  const float4 nd = (target ? s.tgt_c : s.src_c)[c];
  const float4 w = (target ? s.tgt_w : s.src_w)[c];
  const bool valid = nd.w >= 0.0f && finite4(nd) && finite4(w);
  return Contact{valid ? nd.w : 0.0f, valid ? xyz(nd) : mk(1, 0, 0), valid, valid ? w : make_float4(0, 0, 0, 0)};
}
// Eight bytes per candidate that decide whether a row needs its full
// records: the cached source distance (-1 when ipc_cached would report the
// contact invalid) and the lagged normal force. Rewritten wherever the
// source cache or the lagged force changes.
__device__ __forceinline__ void ipc_store_summary(const DeviceState& s, uint32_t c) {
  const float4 nd = s.src_c[c], w = s.src_w[c];
  const bool valid = nd.w >= 0.0f && finite4(nd) && finite4(w);
  s.src_d[c] = make_float2(valid ? nd.w : -1.0f, s.lag_n[c].w);
}
__device__ __forceinline__ bool ipc_friction_active(const DeviceState& s, uint32_t c) {
  return s.ipc_mu > 0.0f && s.lag_n[c].w > 0.0f;
}
// A candidate's energy at contact `k` with the lagged record `lag` and the
// friction displacement `rel` (used only with lagged friction).
__device__ __forceinline__ float ipc_energy_with_rel(const DeviceState& s, const PassParams& p, const Contact& k,
                                                     float4 lag, f3 rel) {  // S:707-740
  if (!k.valid || !(k.distance > s.ipc_dmin)) return kF32Max;
  float e = p.ipc_kappa * ipc_barrier(k.distance, p.ipc_dhat);
  if (s.ipc_mu > 0.0f && lag.w > 0.0f) {
    const f3 n = ipc_safe_unit(xyz(lag), k.normal);
    const f3 tang = rel - n * dot(n, rel);
    const float slip_velocity = sqrtf(dot(tang, tang)) * p.inv_dt;
    e += p.dt * ipc_friction_dissipation(slip_velocity, s.ipc_eps, s.ipc_mu, lag.w);
  }
  return e;
}
__device__ float ipc_energy_from_contact(const DeviceState& s, const PassParams& p, uint32_t c, uint4 ix,
                                         const Contact& k, const float4* pos, uint32_t mv, f3 mp) {  // S:707-740
  if (!k.valid || !(k.distance > s.ipc_dmin)) return kF32Max;
  const float4 lag = s.lag_n[c];
  const bool friction = s.ipc_mu > 0.0f && lag.w > 0.0f;
  return ipc_energy_with_rel(s, p, k, lag, friction ? ipc_weighted_displacement(s, ix, s.lag_w[c], pos, mv, mp) : mk(0, 0, 0));
}
// The friction displacement from a candidate's sweep record (fric): the
// same terms, in the same order, as ipc_weighted_displacement at the
// sweep's source pose with vertex mv's term (x_mv' - old_mv) given.
__device__ __forceinline__ f3 ipc_record_displacement(const float4* fr, uint4 ix, float4 w, uint32_t mv, f3 mv_term) {
  f3 r = mk(0, 0, 0);
  r = r + w.x * (ix.x == mv ? mv_term : xyz(fr[0]));
  r = r + w.y * (ix.y == mv ? mv_term : xyz(fr[1]));
  r = r + w.z * (ix.z == mv ? mv_term : xyz(fr[2]));
  r = r + w.w * (ix.w == mv ? mv_term : xyz(fr[3]));
  return r;
}
// The record slot (0-3) of the vertex that owns a candidate's energy
// (S:742-748: the first unpinned vertex, else the first). Pins are fixed at
// admission, so the slot is computed when a record is written and kept in
// bits 1-2 of cand_meta, beside the kind bit (bit 0: edge-edge).
__device__ __forceinline__ uint32_t ipc_owner_slot(const DeviceState& s, uint4 ix) {
  if (s.vtx_info[ix.x].z == 0u) return 0u;
  if (s.vtx_info[ix.y].z == 0u) return 1u;
  if (s.vtx_info[ix.z].z == 0u) return 2u;
  if (s.vtx_info[ix.w].z == 0u) return 3u;
  return 0u;
}
// cand_meta of a record: bit 0 edge-edge, bits 1-2 the energy owner's
// slot, bit 3 set when one of its vertices is pinned (the active-row pass
// checks pinned motion only for those).
constexpr uint32_t kMetaPinned = 8u;
__device__ __forceinline__ uint32_t ipc_meta(const DeviceState& s, uint4 ix, bool ee) {
  const bool pinned = s.vtx_info[ix.x].z != 0u || s.vtx_info[ix.y].z != 0u || s.vtx_info[ix.z].z != 0u ||
                      s.vtx_info[ix.w].z != 0u;
  return (ee ? 1u : 0u) | (ipc_owner_slot(s, ix) << 1) | (pinned ? kMetaPinned : 0u);
}
__device__ __forceinline__ uint32_t ipc_energy_owner(uint4 ix, uint32_t meta) {
  const uint32_t slot = (meta >> 1) & 3u;
  return slot == 0u ? ix.x : slot == 1u ? ix.y : slot == 2u ? ix.z : ix.w;
}
// Barrier and lagged-friction terms of one vertex (ipcLocalContribution,
// S:786-874) over its active candidate row.
// One active candidate's barrier and friction terms for vertex v.
// `records`: pos is a Jacobi sweep's source pose and the friction records
// (fric) hold it.
__device__ void ipc_row_contribution(const DeviceState& s, const PassParams& p, uint32_t v, const float4* pos,
                                     uint32_t c, f3& g, m3& H, bool records = false) {
  {
    if (!p.colored) {  // the same tests on the eight-byte summary first
      const float2 q = s.src_d[c];
      if (q.x < 0.0f || (!ipc_barrier_active(q.x, p.ipc_dhat) && !(s.ipc_mu > 0.0f && q.y > 0.0f))) return;
    }
    const Contact k = p.colored ? ipc_contact_at(s, c, pos, p.ipc_dhat) : ipc_cached(s, c, false);
    if (!k.valid) return;
    const bool ba = ipc_barrier_active(k.distance, p.ipc_dhat);
    const float4 lag = s.lag_n[c];
    const bool fa = s.ipc_mu > 0.0f && lag.w > 0.0f;
    if (!ba && !fa) return;
    const uint4 ix = s.cand[c];
    const float w = ipc_weight_for_vertex(ix, k.weights, v);
    if (w != 0.0f && ba) {
      const float d1 = ipc_barrier_d1(k.distance, p.ipc_dhat);
      const float d2 = ipc_barrier_d2(k.distance, p.ipc_dhat);
      // ipcNormalGradientScalar / ipcPsdNormalHessianScalar (IC:354-385).
      const float gs = (finite_s(w) && finite_s(p.ipc_kappa) && finite_s(d1) && p.ipc_kappa >= 0.0f) ? w * p.ipc_kappa * d1 : 0.0f;
      g = g + gs * k.normal;
      float hs = 0.0f;
      if (finite_s(w) && finite_s(p.ipc_kappa) && finite_s(d2) && p.ipc_kappa > 0.0f && w != 0.0f && d2 > 0.0f) {
        hs = p.ipc_kappa * w * w * d2;
        if (!finite_s(hs)) hs = kF32Max;
      }
      const f3 u = ipc_safe_unit(k.normal, mk(1, 0, 0));
      const float a = fmaxf(hs, 0.0f);
      H.c0 = H.c0 + (a * u.x) * u;
      H.c1 = H.c1 + (a * u.y) * u;
      H.c2 = H.c2 + (a * u.z) * u;
    }
    const float lw = ipc_weight_for_vertex(ix, s.lag_w[c], v);
    if (fa && lw != 0.0f) {
      const f3 n = ipc_safe_unit(xyz(lag), k.normal);
      const f3 rel = records ? ipc_record_displacement(s.fric + 4 * size_t(c), ix, s.lag_w[c], kNoVertex, mk(0, 0, 0))
                             : ipc_weighted_displacement(s, ix, s.lag_w[c], pos, kNoVertex, mk(0, 0, 0));
      const f3 tang = rel - n * dot(n, rel);
      const float slip = sqrtf(dot(tang, tang));
      const float fs = s.ipc_mu * lag.w;
      f3 dir = mk(0, 0, 0);
      if (slip > 1.0e-12f) dir = tang / slip;
      const float f1 = ipc_friction_f1(slip * p.inv_dt, s.ipc_eps);
      g = g + (lw * fs * f1) * dir;
      float tc = 2.0f * fs / fmaxf(s.ipc_eps * p.dt, 1.0e-12f);
      if (slip > 1.0e-12f) tc = fs * f1 / slip;
      const float cf = lw * lw * fmaxf(tc, 0.0f);
      // (I - n n^T) scaled; n is the unit lagged normal (IC:387-395).
      const f3 un = ipc_safe_unit(n, mk(1, 0, 0));
      H.c0 = H.c0 + cf * (mk(1, 0, 0) - un.x * un);
      H.c1 = H.c1 + cf * (mk(0, 1, 0) - un.y * un);
      H.c2 = H.c2 + cf * (mk(0, 0, 1) - un.z * un);
    }
  }
}
__device__ void ipc_local_contribution(const DeviceState& s, const PassParams& p, uint32_t v, const float4* pos,
                                       f3& g, m3& H) {
  const uint32_t begin = s.cand_row[v].x, end = begin + s.active_count[v];
#pragma unroll 1
  for (uint32_t i = begin; i < end; ++i) ipc_row_contribution(s, p, v, pos, s.active_ids[i], g, H);
}
// The signed relative area stretch test of an assembled energy
// (S:2585-2600): the triangle's rest area is positive and its stretch at
// `pos` finite and above the floor.
__device__ __forceinline__ bool tri_stretch_ok(const DeviceState& s, uint32_t tri, const float4* pos) {
  const float tra = 2.0f * s.tri_area[tri];
  const float stretch = len3(tri_raw_normal(s, tri, pos, kNoVertex, mk(0, 0, 0))) / tra;
  return (tra > 0.0f) && finite_s(stretch) && (stretch > kDeterminantFloor);
}
template <bool kCloth, bool kIpc>
__device__ __forceinline__ float vertex_energy_at(const DeviceState& s, const PassParams& p, const float4* pos, uint32_t v,
                                                  bool is_target) {  // S:2557-2645
  bool pin = s.vtx_info[v].z != 0u;
  f3 x = load_pos(pos, v);
  float e = 0.0f;
  if (!pin) {
    float in = s.vtx_rest[v].w * p.inv_dt2;
    f3 r = x - xyz(s.predicted[v]);
    e += 0.5f * in * dot(r, r);
  }
  if (p.objective_flags != 0u && !pin) {
    if (p.objective_flags & kForce) e -= dot(xyz(s.objectives[2 * size_t(v)]), x);
    if (p.objective_flags & kTarget) {
      float4 t = s.objectives[2 * size_t(v) + 1];
      if (t.w > 0.0f) {
        f3 tr = x - xyz(t);
        e += 0.5f * t.w * dot(tr, tr);
      }
    }
  }
  if (!pin && p.floor_k > 0.0f) {
    float pen = fminf(0.0f, x.y - p.floor_y);
    e += 0.5f * p.floor_k * pen * pen;
  }
  if constexpr (kCloth) {
    // Cloth elements are summed once, by their owner vertex (S:2585-2643).
    for (uint32_t k = s.tri_row[v].x, ke = k + s.tri_row[v].y; k < ke; ++k) {
      const uint32_t tri = s.tri_elems[k];
      if (s.tri_ix[tri].w != v) continue;
      if (p.target_records) {
        // After k_cloth_base mode 1 the records hold the target's energy and
        // stretch test, the source's in spare words.
        const float* r = reinterpret_cast<const float*>(s.tri_base + 3 * size_t(tri));
        if (!(__float_as_uint(r[11]) & (is_target ? 1u : 2u))) return kF32Max;
        e += r[is_target ? 3 : 7];
        continue;
      }
      if (!tri_stretch_ok(s, tri, pos)) return kF32Max;
      // At the source pose the element records hold these energies.
      e += is_target ? membrane_at(s, tri, pos, kNoVertex, mk(0, 0, 0)).energy : membrane_base(s, p, tri, pos).energy;
    }
    for (uint32_t k = s.hinge_row[v].x, ke = k + s.hinge_row[v].y; k < ke; ++k) {
      const uint32_t h = s.hinge_elems[k];
      if (__float_as_uint(s.hinge_rest[h].z) != v) continue;
      if (p.target_records) e += s.hinge_base[4 * size_t(h) + (is_target ? 2 : 3)].w;
      else e += is_target ? bending_energy_at(s, h, pos, kNoVertex, mk(0, 0, 0)) : bending_base(s, p, h, pos);
    }
  }
  if constexpr (kIpc) {  // S:2643 with ipcCachedContactAt (S:450-460); summed by k_ipc_energy_lanes
    const float2 owned = s.ipc_energy[v];
    e += is_target ? owned.y : owned.x;
  }
  return e;
}
template <bool kCloth, bool kIpc>
__device__ __forceinline__ void store_vertex_energy(const DeviceState& s, const PassParams& p, uint32_t v) {  // S:2647-2658
  float se = vertex_energy_at<kCloth, kIpc>(s, p, s.pos[p.src], v, false);
  float ce = vertex_energy_at<kCloth, kIpc>(s, p, s.pos[p.tgt], v, true);
  bool sv = finite_s(se), cv = finite_s(ce);
  s.vertex_energy[v] = make_float4(sv ? se : 0.0f, cv ? ce : 0.0f, sv ? 1.0f : 0.0f, cv ? 1.0f : 0.0f);
}

// ------------------------------------------------------------------ co-rotated (legacy) element terms
__device__ __forceinline__ m3 stiffness_block(const DeviceState& s, uint32_t t, uint32_t rv, uint32_t cv) {  // S:1022-1043
  const float* k = s.tet_k + size_t(t) * 144;
  uint32_t r0 = rv * 3, c0 = cv * 3;
  return {mk(k[r0 * 12 + c0], k[(r0 + 1) * 12 + c0], k[(r0 + 2) * 12 + c0]),
          mk(k[r0 * 12 + c0 + 1], k[(r0 + 1) * 12 + c0 + 1], k[(r0 + 2) * 12 + c0 + 1]),
          mk(k[r0 * 12 + c0 + 2], k[(r0 + 1) * 12 + c0 + 2], k[(r0 + 2) * 12 + c0 + 2])};
}
// tetGradient, co-rotated branch (S:1171-1197).
__device__ void corot_gradients(const DeviceState& s, const PassParams& p, uint32_t t, uint4 ix, f3 out[4]) {
  const float4* pos = s.pos[p.src];
  uint32_t id[4] = {ix.x, ix.y, ix.z, ix.w};
  f3 cur[4], rst[4];
  for (int l = 0; l < 4; ++l) {
    cur[l] = load_pos(pos, id[l]);
    rst[l] = xyz(s.vtx_rest[id[l]]);
  }
  f3 cc = 0.25f * (((cur[0] + cur[1]) + cur[2]) + cur[3]);
  f3 rc = 0.25f * (((rst[0] + rst[1]) + rst[2]) + rst[3]);
  m3 R = load_rot(s.trot, t);
  f3 disp[4];
  for (int l = 0; l < 4; ++l) disp[l] = tmul(R, cur[l] - cc) - (rst[l] - rc);
  for (int r = 0; r < 4; ++r) {
    f3 lg = mk(0, 0, 0);
    for (int c = 0; c < 4; ++c) lg = lg + mul(stiffness_block(s, t, r, c), disp[c]);
    out[r] = mul(R, lg);
  }
}

// ------------------------------------------------------------------ the local system
struct Local {
  f3 g;
  m3 H;
};
__device__ __forceinline__ uint32_t tet_vertex(uint4 ix, int l) {
  return l == 0 ? ix.x : (l == 1 ? ix.y : (l == 2 ? ix.z : ix.w));
}
// Row-major Ubar block l of a cubature record (S:1204-1212), as columns.
__device__ __forceinline__ m3 ubar_block(const float* ub, int l) {
  const float* b = ub + 9 * l;
  return {mk(b[0], b[3], b[6]), mk(b[1], b[4], b[7]), mk(b[2], b[5], b[8])};
}
// Ubar_l q for a row-major block.
__device__ __forceinline__ f3 ubar_times(const float* ub, int l, f3 q) {
  const float* b = ub + 9 * l;
  return mk(b[0] * q.x + b[1] * q.y + b[2] * q.z, b[3] * q.x + b[4] * q.y + b[5] * q.z,
            b[6] * q.x + b[7] * q.y + b[8] * q.z);
}
// Incident SNH block in closed form: with a = F g, b = cof(F) g,
//   vol DP(F)[e_c g^T] g = vol (shear |g|^2 e_c + 2 mu/den^2 a_c a + lambda b_c b),
// because det(F + x g^T) is affine in x, so the directional-cofactor term of
// the tangent (stable-neo-hookean-wgsl.ts:72-94) vanishes on this direction.
__device__ __forceinline__ void add_incident_block(const SnhPoint& sp, const Tet& e, f3 gr, float sign, f3& g, m3& H) {
  const f3 a = mul(sp.F, gr), b = mul(sp.cof, gr);
  g = g + (sign * e.vol) * (sp.shear * a + sp.volscale * b);
  const float cs = sign * e.vol * sp.shear * dot(gr, gr);
  const float ca = sign * e.vol * 2.0f * e.mu / (sp.den * sp.den);
  const float cb = sign * e.vol * e.lam;
  H.c0 = H.c0 + mk(cs, 0, 0) + (ca * a.x) * a + (cb * b.x) * b;
  H.c1 = H.c1 + mk(0, cs, 0) + (ca * a.y) * a + (cb * b.y) * b;
  H.c2 = H.c2 + mk(0, 0, cs) + (ca * a.z) * a + (cb * b.z) * b;
}
// <A, d cof(F)[B]>: the second derivative of det F in directions A, B.
__device__ __forceinline__ float det_second(const m3& F, const m3& A, const m3& B) {
  return dot(A.c0, cross(B.c1, F.c2) + cross(F.c1, B.c2)) + dot(A.c1, cross(B.c2, F.c0) + cross(F.c2, B.c0)) +
         dot(A.c2, cross(B.c0, F.c1) + cross(F.c0, B.c1));
}

// One cubature record's contribution to (g, H) for stable Neo-Hookean
// (S:3198-3263 with cubatureHessian S:2339-2366 and the distributed inertia
// and objective terms S:2401-2477). The reduced Hessian entry (r, c) is
// vol <D_r, DP(F)[D_c]> with D_c = sum_l (B_l e_c) grad_l^T, B_l = R_{v_l}
// Ubar_l R_s^T; blocks are streamed so only the three D_c stay live.
__device__ __forceinline__ void cubature_record_snh(const DeviceState& s, const PassParams& p, uint32_t v, f3 x,
                                                    uint32_t t, const float* ub, const m3& Rs_t, float w, Local& L) {
  const Tet e = load_tet(s, t);
  const float4* vc = vcub_source(s, p);
  // The sample tetrahedron's pose from its four vertex records (no local
  // arrays: a runtime-indexed array would live in scratch memory).
  const f3 x0 = vcub_pose(vc, e.ix.x), x1 = vcub_pose(vc, e.ix.y), x2 = vcub_pose(vc, e.ix.z), x3 = vcub_pose(vc, e.ix.w);
  const m3 ds = {x1 - x0, x2 - x0, x3 - x0};
  const SnhPoint sp = snh_point(mul(ds, e.inv), e.lam, e.mu);
  f3 pg = mk(0, 0, 0);
  float i00 = 0, i11 = 0, i22 = 0, i01 = 0, i02 = 0, i12 = 0;
  m3 D0 = mzero(), D1 = mzero(), D2 = mzero();
  int ss = 4;
#pragma unroll 1
  for (int l = 0; l < 4; ++l) {
    const uint32_t vl = tet_vertex(e.ix, l);
    if (vl == v) ss = l;
    const VCub ql = load_vcub(vc, vl);
    const m3 B = mul(mul(ql.R, ubar_block(ub, l)), Rs_t);
    const f3 gr = shape_gradient(e, l);
    const float wl = ql.w;
    f3 sg = e.vol * piola_times(sp, gr) + wl * ql.diff;
    float wt = wl;
    if (p.objective_flags != 0u && !ql.pinned) {
      const uint4 il = s.vtx_info[vl];
      const float n = float(il.y > 1u ? il.y : 1u);
      sg = sg + objective_gradient(s, p, vl, ql.x) / n;
      if (p.objective_flags & kTarget) wt += s.objectives[2 * size_t(vl) + 1].w / n;
    }
    pg = pg + tmul(B, sg);
    D0.c0 = D0.c0 + gr.x * B.c0;
    D0.c1 = D0.c1 + gr.y * B.c0;
    D0.c2 = D0.c2 + gr.z * B.c0;
    D1.c0 = D1.c0 + gr.x * B.c1;
    D1.c1 = D1.c1 + gr.y * B.c1;
    D1.c2 = D1.c2 + gr.z * B.c1;
    D2.c0 = D2.c0 + gr.x * B.c2;
    D2.c1 = D2.c1 + gr.y * B.c2;
    D2.c2 = D2.c2 + gr.z * B.c2;
    i00 += wt * dot(B.c0, B.c0);
    i11 += wt * dot(B.c1, B.c1);
    i22 += wt * dot(B.c2, B.c2);
    i01 += wt * dot(B.c0, B.c1);
    i02 += wt * dot(B.c0, B.c2);
    i12 += wt * dot(B.c1, B.c2);
  }
  const float ca = 2.0f * e.mu / (sp.den * sp.den);
  const float f0 = frob(sp.F, D0), f1 = frob(sp.F, D1), f2 = frob(sp.F, D2);
  const float c0 = frob(sp.cof, D0), c1 = frob(sp.cof, D1), c2 = frob(sp.cof, D2);
  auto entry = [&](const m3& A, const m3& B, float fa, float fb, float ga, float gb) {
    return e.vol * (sp.shear * frob(A, B) + ca * fa * fb + e.lam * ga * gb + sp.volscale * det_second(sp.F, A, B));
  };
  const float h00 = entry(D0, D0, f0, f0, c0, c0) + i00;
  const float h11 = entry(D1, D1, f1, f1, c1, c1) + i11;
  const float h22 = entry(D2, D2, f2, f2, c2, c2) + i22;
  const float h01 = entry(D0, D1, f0, f1, c0, c1) + i01;
  const float h02 = entry(D0, D2, f0, f2, c0, c2) + i02;
  const float h12 = entry(D1, D2, f1, f2, c1, c2) + i12;
  m3 ph = {mk(h00, h01, h02), mk(h01, h11, h12), mk(h02, h12, h22)};
  if (ss < 4) {
    // Remove the exact source block and the source's distributed inertia and
    // objective shares (S:3237-3259).
    const uint4 inf = s.vtx_info[v];
    const float n = float(inf.y > 1u ? inf.y : 1u);
    const float w_src = s.vtx_rest[v].w * p.inv_dt2 / n;
    add_incident_block(sp, e, shape_gradient(e, ss), -1.0f, pg, ph);
    pg = pg - w_src * (x - xyz(s.predicted[v]));
    float diag = w_src;
    if (p.objective_flags != 0u) {
      pg = pg - objective_gradient(s, p, v, x) / n;
      if (p.objective_flags & kTarget) diag += target_stiffness(s, p, v) / n;
    }
    ph.c0.x -= diag;
    ph.c1.y -= diag;
    ph.c2.z -= diag;
  }
  L.g = L.g + w * pg;
  L.H = mfma(w, ph, L.H);
}

// Legacy co-rotated record (S:2368-2398 with S:3226-3259).
__device__ void cubature_record_corot(const DeviceState& s, const PassParams& p, uint32_t v, f3 x, uint32_t t,
                                      const float* ub, const m3& Rs_t, float w, Local& L) {
  const float4* pos = s.pos[p.src];
  const Tet e = load_tet(s, t);
  m3 B[4];
  float wl[4];
  for (int l = 0; l < 4; ++l) {
    const uint32_t vl = tet_vertex(e.ix, l);
    B[l] = mul(mul(load_rot(rot_source(s, p), vl), ubar_block(ub, l)), Rs_t);
    wl[l] = dist_inertia_weight(s, p, vl);
  }
  f3 tg[4];
  corot_gradients(s, p, t, e.ix, tg);
  f3 pg = mk(0, 0, 0);
  for (int l = 0; l < 4; ++l) {
    const uint32_t vl = tet_vertex(e.ix, l);
    pg = pg + tmul(B[l], tg[l] + wl[l] * (load_pos(pos, vl) - xyz(s.predicted[vl])));
  }
  const m3 R = load_rot(s.trot, t);
  f3 cols[3];
  for (int c = 0; c < 3; ++c) {
    f3 ld[4], wp[4];
    for (int l = 0; l < 4; ++l) ld[l] = tmul(R, c == 0 ? B[l].c0 : (c == 1 ? B[l].c1 : B[l].c2));
    for (int rv = 0; rv < 4; ++rv) {
      f3 lp = mk(0, 0, 0);
      for (int cv = 0; cv < 4; ++cv) lp = lp + mul(stiffness_block(s, t, rv, cv), ld[cv]);
      wp[rv] = mul(R, lp);
    }
    cols[c].x = dot(B[0].c0, wp[0]) + dot(B[1].c0, wp[1]) + dot(B[2].c0, wp[2]) + dot(B[3].c0, wp[3]);
    cols[c].y = dot(B[0].c1, wp[0]) + dot(B[1].c1, wp[1]) + dot(B[2].c1, wp[2]) + dot(B[3].c1, wp[3]);
    cols[c].z = dot(B[0].c2, wp[0]) + dot(B[1].c2, wp[1]) + dot(B[2].c2, wp[2]) + dot(B[3].c2, wp[3]);
  }
  m3 ph = {cols[0], cols[1], cols[2]};
  for (int l = 0; l < 4; ++l) {
    const m3& Bl = B[l];
    ph.c0 = ph.c0 + wl[l] * tmul(Bl, Bl.c0);
    ph.c1 = ph.c1 + wl[l] * tmul(Bl, Bl.c1);
    ph.c2 = ph.c2 + wl[l] * tmul(Bl, Bl.c2);
  }
  const uint32_t ss = local_slot(e.ix, v);
  if (ss < 4) {
    const float w_src = dist_inertia_weight(s, p, v);
    pg = pg - (tg[ss] + w_src * (x - xyz(s.predicted[v])));
    ph = msub(ph, madd(mul(mul(R, stiffness_block(s, t, ss, ss)), tr(R)), mscale(w_src, mid())));
  }
  L.g = L.g + w * pg;
  L.H = mfma(w, ph, L.H);
}

// Gather of jgs2Solve (S:3074-3263): inertia, objectives, the exact
// incident elements, the floor penalty and the cubature correction.
template <bool kCorot, bool kCloth, bool kIpc>
__device__ Local gather_local(const DeviceState& s, const PassParams& p, uint32_t v) {
  const float4* pos = s.pos[p.src];
  const f3 x = load_pos(pos, v);
  const uint4 inf = s.vtx_info[v];
  const float inertia = s.vtx_rest[v].w * p.inv_dt2;
  Local L;
  L.g = inertia * (x - xyz(s.predicted[v]));
  L.H = mscale(inertia + target_stiffness(s, p, v), mid());
  if (p.objective_flags != 0u) L.g = L.g + objective_gradient(s, p, v, x);
#pragma unroll 1
  for (uint32_t k = 0; k < inf.y; ++k) {  // S:3097-3111
    const uint32_t t = s.adjacency[inf.x + k];
    const Tet e = load_tet(s, t);
    const uint32_t slot = local_slot(e.ix, v);
    if (slot >= 4) continue;
    if constexpr (!kCorot) {
      add_incident_block(snh_point(deformation(e, pos), e.lam, e.mu), e, shape_gradient(e, slot), 1.0f, L.g, L.H);
    } else {
      f3 tg[4];
      corot_gradients(s, p, t, e.ix, tg);
      const m3 R = load_rot(s.trot, t);
      L.g = L.g + tg[slot];
      L.H = madd(L.H, mul(mul(R, stiffness_block(s, t, slot, slot)), tr(R)));
    }
  }
  if constexpr (kCloth) {
#pragma unroll 1
    for (uint32_t k = s.tri_row[v].x, ke = k + s.tri_row[v].y; k < ke; ++k) {  // S:3117-3144
      const uint32_t tri = s.tri_elems[k];
      const uint32_t slot = tri_slot(s.tri_ix[tri], v);
      if (slot >= 3) continue;
      const Membrane m = membrane_base(s, p, tri, pos);
      L.g = L.g + membrane_gradient(s, tri, m, slot);
      L.H = madd(L.H, membrane_local_hessian(s, tri, m, slot));
    }
#pragma unroll 1
    for (uint32_t k = s.hinge_row[v].x, ke = k + s.hinge_row[v].y; k < ke; ++k) {  // S:3145-3184
      const uint32_t h = s.hinge_elems[k];
      const uint32_t slot = hinge_slot(s.hinge_ix[h], v);
      if (slot >= 4) continue;
      const Dihedral d = dihedral_base(s, p, h, pos, slot);
      if (!d.valid) {
        L.g = L.g + mk(kF32Max, kF32Max, kF32Max);
      } else {
        const float4 r = s.hinge_rest[h];
        const float ad = wrapped_angle_difference(d.angle, r.x);
        const float k2 = s.cloth_bend * r.y;
        L.g = L.g + (k2 * ad) * d.grad;
        L.H = m3{fma3(k2 * d.grad.x, d.grad, L.H.c0), fma3(k2 * d.grad.y, d.grad, L.H.c1), fma3(k2 * d.grad.z, d.grad, L.H.c2)};
      }
    }
  }
  if (p.floor_k > 0.0f && x.y < p.floor_y) {  // S:3188-3191
    L.g.y += p.floor_k * (x.y - p.floor_y);
    L.H.c1.y += p.floor_k;
  }
  if constexpr (kIpc) {  // S:3193-3195
    f3 ig = mk(0, 0, 0);
    m3 ih = mzero();
    ipc_local_contribution(s, p, v, pos, ig, ih);
    L.g = L.g + ig;
    L.H = madd(L.H, ih);
  }
  if (p.cubature_k == 0) return L;
  const m3 Rs_t = tr(load_vcub(vcub_source(s, p), v).R);
#pragma unroll 1
  for (uint32_t k = 0; k < p.cubature_k; ++k) {  // S:3198-3263
    const size_t rec = size_t(v) * p.cubature_k + k;
    const uint4 head = s.cub[10 * rec];
    const float w = __uint_as_float(head.y);
    if (head.x == kEmptyTet || head.x >= p.tet_count || !(w > 0.0f)) continue;
    const float* ub = reinterpret_cast<const float*>(s.cub + 10 * rec + 1);
    if constexpr (!kCorot)
      cubature_record_snh(s, p, v, x, head.x, ub, Rs_t, w, L);
    else
      cubature_record_corot(s, p, v, x, head.x, ub, Rs_t, w, L);
  }
  return L;
}

// Restricted trial model (S:1980-2337) for displacement d of vertex v:
// minimum determinant over changed elements and the restricted energy delta.
struct Trial {
  float min_det;
  bool geometry_valid;
  float energy;
};
template <bool kCloth, bool kIpc>
__device__ Trial restricted_trial(const DeviceState& s, const PassParams& p, uint32_t v, f3 x, f3 d, float initial_min) {
  const float4* pos = s.pos[p.src];
  const uint4 inf = s.vtx_info[v];
  Trial tr{initial_min, true, 0.0f};
  const f3 pred = xyz(s.predicted[v]);
  float energy = quadratic_delta(x, pred, d, s.vtx_rest[v].w * p.inv_dt2) + floor_delta(p, x.y, d.y);
  if (p.objective_flags != 0u) energy += objective_delta(s, p, v, x, d);
  if constexpr (kIpc) {
    // Contact feasibility of the moved vertex (S:1997-2037) and the contact
    // energy delta (S:2140-2193) in one pass over its active row.
    const f3 tp = x + d;
    const float dl = sqrtf(dot(d, d));
    const uint32_t begin = s.cand_row[v].x, end = begin + s.active_count[v];
#pragma unroll 1
    for (uint32_t i = begin; i < end; ++i) {
      const uint32_t c = s.active_ids[i];
      if (!p.colored) {  // the same tests on the eight-byte summary first
        const float2 q = s.src_d[c];
        if (q.x >= 0.0f && q.x - dl > s.ipc_dmin && !(q.x - dl < p.ipc_dhat) && !(s.ipc_mu > 0.0f && q.y > 0.0f))
          continue;
      }
      const Contact sc = p.colored ? ipc_contact_at(s, c, pos, p.ipc_dhat) : ipc_cached(s, c, false);
      if (!sc.valid) {
        tr.geometry_valid = false;
        return tr;
      }
      const bool near = !(sc.distance - dl > s.ipc_dmin);
      const bool may_change = sc.distance - dl < p.ipc_dhat;
      const bool fa = ipc_friction_active(s, c);
      if (!near && !may_change && !fa) continue;
      Contact tc = Contact{p.ipc_dhat, mk(1, 0, 0), true, make_float4(0, 0, 0, 0)};
      if (near || may_change) tc = ipc_contact_at(s, c, pos, p.ipc_dhat, v, tp);
      if (near) {
        if (!tc.valid) {
          tr.geometry_valid = false;
          return tr;
        }
        if (!(tc.distance > s.ipc_dmin)) {
          tr.min_det = 0.0f;
          return tr;
        }
      }
      if (!may_change && !fa) continue;
      const uint4 ix = s.cand[c];
      energy += ipc_energy_from_contact(s, p, c, ix, tc, pos, v, tp) -
                ipc_energy_from_contact(s, p, c, ix, sc, pos, kNoVertex, mk(0, 0, 0));
    }
  }
  // Incident tetrahedra: F' = F + d g^T. In closed form det(F') = J + d.b
  // and <D, F' + F> = 2 d.a + |d|^2 |g|^2 with a = F g, b = cof g; the pair
  // and triple terms of the determinant delta (GL:510-524) vanish for rank one.
  for (uint32_t k = 0; k < inf.y; ++k) {
    const uint32_t t = s.adjacency[inf.x + k];
    const Tet e = load_tet(s, t);
    const f3 gr = shape_gradient(e, local_slot(e.ix, v));
    const m3 F = deformation(e, pos);
    const f3 c12 = cross(F.c1, F.c2), c20 = cross(F.c2, F.c0), c01 = cross(F.c0, F.c1);
    const float J = dot(F.c0, c12);
    const f3 b = gr.x * c12 + gr.y * c20 + gr.z * c01;
    const float jd = dot(d, b);
    const float det_trial = J + jd;
    if (!finite_s(det_trial)) {
      tr.geometry_valid = false;
      return tr;
    }
    tr.min_det = fminf(tr.min_det, det_trial);
    const float id = 2.0f * dot(d, mul(F, gr)) + dot(d, d) * dot(gr, gr);
    const float lr = log_one_plus(id / (frob(F, F) + 1.0f));
    energy += e.vol * (0.5f * e.mu * id - 0.75f * e.mu * jd + 0.5f * e.lam * jd * (2.0f * J + jd - 2.0f) - 0.5f * e.mu * lr);
  }
  if constexpr (kCloth) {
    const f3 tp = x + d;
#pragma unroll 1
    for (uint32_t k = s.tri_row[v].x, ke = k + s.tri_row[v].y; k < ke; ++k) {
      const uint32_t tri = s.tri_elems[k];
      if (tri_slot(s.tri_ix[tri], v) >= 3) continue;
      // Signed relative area stretch (S:2062-2116).
      const f3 sn = tri_raw_normal(s, tri, pos, kNoVertex, mk(0, 0, 0));
      const f3 tn = tri_raw_normal(s, tri, pos, v, tp);
      const float tra = 2.0f * s.tri_area[tri];
      const float ss = len3(sn) / tra;
      const float tm = len3(tn) / tra;
      const float ts = dot(sn, tn) > 0.0f ? tm : -tm;
      if (!(tra > 0.0f) || !finite_s(ss) || !finite_s(ts)) {
        tr.geometry_valid = false;
        return tr;
      }
      tr.min_det = fminf(tr.min_det, fminf(ss, ts));
      // Membrane energy delta (S:2203-2236).
      energy += membrane_at(s, tri, pos, v, tp).energy - membrane_base(s, p, tri, pos).energy;
    }
#pragma unroll 1
    for (uint32_t k = s.hinge_row[v].x, ke = k + s.hinge_row[v].y; k < ke; ++k) {  // S:2237-2271
      const uint32_t h = s.hinge_elems[k];
      if (hinge_slot(s.hinge_ix[h], v) >= 4) continue;
      energy += bending_energy_at(s, h, pos, v, tp) - bending_base(s, p, h, pos);
    }
  }
  if (p.cubature_k > 0) {
    const float4* vc = vcub_source(s, p);
    const f3 ds = tmul(load_vcub(vc, v).R, d);  // R_s^T d
    const float n_src = float(inf.y > 1u ? inf.y : 1u);
    const float w_src = s.vtx_rest[v].w * p.inv_dt2 / n_src;
    for (uint32_t k = 0; k < p.cubature_k; ++k) {
      const size_t rec = size_t(v) * p.cubature_k + k;
      const uint4 head = s.cub[10 * rec];
      const float w = __uint_as_float(head.y);
      if (head.x == kEmptyTet || head.x >= p.tet_count || !(w > 0.0f)) continue;
      const Tet e = load_tet(s, head.x);
      const float* ub = reinterpret_cast<const float*>(s.cub + 10 * rec + 1);
      const f3 x0 = vcub_pose(vc, e.ix.x), x1 = vcub_pose(vc, e.ix.y), x2 = vcub_pose(vc, e.ix.z),
               x3 = vcub_pose(vc, e.ix.w);
      const m3 F = mul(m3{x1 - x0, x2 - x0, x3 - x0}, e.inv);
      m3 Dd = mzero();
      float compl_e = 0.0f;
      int ss = 4;
#pragma unroll 1
      for (int l = 0; l < 4; ++l) {
        const uint32_t vl = tet_vertex(e.ix, l);
        if (vl == v) ss = l;
        const VCub ql = load_vcub(vc, vl);
        const f3 bd = mul(ql.R, ubar_times(ub, l, ds));  // B_l d
        const f3 gl = shape_gradient(e, l);
        Dd.c0 = Dd.c0 + gl.x * bd;
        Dd.c1 = Dd.c1 + gl.y * bd;
        Dd.c2 = Dd.c2 + gl.z * bd;
        if (!ql.pinned) {
          // quadratic_delta(x_l, x_pred_l, bd, w_l) from the stored x - x_pred.
          compl_e += ql.w * (dot(ql.diff, bd) + 0.5f * dot(bd, bd));
          if (p.objective_flags != 0u) {
            const uint4 il = s.vtx_info[vl];
            const float n = float(il.y > 1u ? il.y : 1u);
            compl_e += objective_delta(s, p, vl, ql.x, bd) / n;
          }
        }
      }
      const float det_trial = mdet(madd(F, Dd));
      if (!finite_s(det_trial)) {
        tr.geometry_valid = false;
        return tr;
      }
      tr.min_det = fminf(tr.min_det, det_trial);
      compl_e += snh_delta_general(F, Dd, e.lam, e.mu) * e.vol;
      if (ss < 4) {
        const f3 gs = shape_gradient(e, ss);
        const f3 c12 = cross(F.c1, F.c2), c20 = cross(F.c2, F.c0), c01 = cross(F.c0, F.c1);
        const float J = dot(F.c0, c12);
        const f3 b = gs.x * c12 + gs.y * c20 + gs.z * c01;
        const float jd = dot(d, b);
        const float id = 2.0f * dot(d, mul(F, gs)) + dot(d, d) * dot(gs, gs);
        const float lr = log_one_plus(id / (frob(F, F) + 1.0f));
        compl_e -= e.vol * (0.5f * e.mu * id - 0.75f * e.mu * jd + 0.5f * e.lam * jd * (2.0f * J + jd - 2.0f) - 0.5f * e.mu * lr);
        compl_e -= quadratic_delta(x, pred, d, w_src);
        if (p.objective_flags != 0u) compl_e -= objective_delta(s, p, v, x, d) / n_src;
      }
      energy += w * compl_e;
    }
  }
  tr.energy = energy;
  return tr;
}

__device__ __forceinline__ void store_local_diag(const DeviceState& s, uint32_t v, f3 dir, float alpha, float min_eig,
                                                 float scale, float dshift, float nshift, float de, float bound, float gd,
                                                 float res, float minj, uint32_t bt, uint32_t ev, uint32_t st) {  // S:2518-2555
  float4* b = s.local_diag + 4 * size_t(v);
  b[0] = make_float4(dir.x, dir.y, dir.z, alpha);
  b[1] = make_float4(min_eig, scale, dshift, nshift);
  b[2] = make_float4(de, bound, gd, res);
  b[3] = make_float4(minj, float(bt), float(ev), float(st));
}

// ------------------------------------------------------------------ kernels
__global__ void __launch_bounds__(kBlock) k_clear_frame(DeviceState s, PassParams p) {
  // Control records are cleared by predict's first invocation (S:2687-2691).
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < 4 * p.scene_count) s.control[i] = make_float4(0, 0, 0, 0);
}

__global__ void __launch_bounds__(kBlock) k_predict(DeviceState s, PassParams p) {  // S:2661-2692
  const uint32_t v = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (v >= p.vertex_count || vscene(s, v) == kNoScene) return;
  const f3 o = load_pos(s.pos[0], v);
  s.old[v] = make_float4(o.x, o.y, o.z, 1.0f);
  if (s.bp_nv && (s.bp_ctl[vscene(s, v)].x & kBpScene)) s.bp_cpose[v] = make_float4(o.x, o.y, o.z, 1.0f);
  const bool pinned = s.vtx_info[v].z != 0u;
  f3 pr = xyz(s.vtx_rest[v]);
  if (!pinned) {
    const f3 dv = p.damping * xyz(s.velocity[v]);
    pr = o + p.dt * dv + (p.dt * p.dt) * mk(p.gx, p.gy, p.gz);
  } else {
    s.velocity[v] = make_float4(0, 0, 0, 0);
  }
  s.predicted[v] = make_float4(pr.x, pr.y, pr.z, 1.0f);
  const f3 src = p.globalization ? o : pr;
  s.pos[1][v] = make_float4(src.x, src.y, src.z, 1.0f);
  s.final_update[v] = make_float4(0, 0, 0, 0);
}

__device__ m3 polar_rotation(m3 r, float eps);

__global__ void __launch_bounds__(kBlock) k_tet_polar(DeviceState s, PassParams p) {  // S:2932-2957
  const uint32_t t = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (t >= p.tet_count) return;
  const uint32_t sc = s.tblock_scene[t / kSceneBlock];
  if (sc == kNoScene || t >= s.scene_end[sc].y || stop_latched(s, p, sc)) return;
  if (p.colored && p.color_index >= s.scene_cnt[sc].z) return;  // the scene's sweep has ended
  const Tet e = load_tet(s, t);
  store_rot(s.trot, t, polar_rotation(deformation(e, s.pos[p.src]), p.polar_eps));
}

__device__ __forceinline__ f3 safe_normalize(f3 v, f3 fb, float eps) {  // S:924-930
  float sq = dot(v, v);
  if (sq > eps * eps && sq < 1.0e30f) return v * (1.0f / sqrtf(sq));
  return fb;
}
__device__ __forceinline__ f3 perpendicular_to(f3 a, float eps) {  // S:932-938
  f3 c = fabsf(a.x) > 0.8f ? mk(0, 1, 0) : mk(1, 0, 0);
  return safe_normalize(cross(a, c), mk(0, 0, 1), eps);
}
__device__ m3 polar_rotation(m3 r, float eps) {  // S:940-978
  for (int it = 0; it < 7; ++it) {
    float d = mdet(r);
    float mag = fabsf(d);
    if (!(mag > eps) || mag > 1.0e20f) break;
    float rc = 1.0f / d;
    m3 it_ = mscale(rc, cofactor(r));
    r = mscale(0.5f, madd(r, it_));
  }
  f3 c0 = safe_normalize(r.c0, mk(1, 0, 0), eps);
  f3 c1c = r.c1 - c0 * dot(c0, r.c1);
  f3 c1 = safe_normalize(c1c, perpendicular_to(c0, eps), eps);
  f3 c2 = safe_normalize(cross(c0, c1), mk(0, 0, 1), eps);
  if (dot(c2, r.c2) < 0.0f) {
    c1 = -c1;
    c2 = -c2;
  }
  c1 = safe_normalize(cross(c2, c0), perpendicular_to(c0, eps), eps);
  return {c0, c1, c2};
}

__global__ void __launch_bounds__(kBlock) k_vertex_polar(DeviceState s, PassParams p) {  // S:2960-2989
  const uint32_t v = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (v >= p.vertex_count) return;
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || stop_latched(s, p, sc)) return;
  // A scene with fewer colors than the batch has finished its sweep: its
  // frames stay those of its own last color pass.
  if (p.colored && p.color_index >= s.scene_cnt[sc].z) return;
  const uint4 inf = s.vtx_info[v];
  const f3 x = load_pos(s.pos[p.src], v);
  if (inf.y == 0u) {
    store_rot(s.vrot, v, mid());
    store_vcub(s, p, s.vcub[0], v, inf, mid(), x);
    return;
  }
  m3 avg = mzero();
  float tv = 0.0f;
  for (uint32_t k = 0; k < inf.y; ++k) {
    const uint32_t t = s.adjacency[inf.x + k];
    const Tet e = load_tet(s, t);
    avg = mfma(e.vol, deformation(e, s.pos[p.src]), avg);
    tv += e.vol;
  }
  avg = tv > 0.0f ? mscale(1.0f / tv, avg) : mid();
  const m3 R = polar_rotation(avg, p.polar_eps);
  store_rot(s.vrot, v, R);
  store_vcub(s, p, s.vcub[0], v, inf, R, x);
}

// ------------------------------------------------------------------ cooperative local solve
// The local solve of one vertex split over LPV lanes of a wavefront. A
// vertex's items -- active contact rows, incident tetrahedra, cloth
// triangles, hinges and cubature records, in that canonical order -- are
// dealt round robin (item j to lane j mod LPV); each lane sums its items'
// gradient and Hessian terms, the lanes combine them with xor butterflies
// (commutative pairwise adds, so every lane ends with the same bits), and
// every lane then runs the same 3x3 Newton step and line-search decisions.
// The line-search trial splits the same way; the earliest failing item in
// canonical order decides an infeasible trial exactly as the sequential
// early returns do (S:1980-2337).
template <int LPV> __device__ __forceinline__ float lanes_sum(float x) {
#pragma unroll
  for (int m = LPV / 2; m >= 1; m >>= 1) x += __shfl_xor(x, m, LPV);
  return x;
}
template <int LPV> __device__ __forceinline__ float lanes_min(float x) {
#pragma unroll
  for (int m = LPV / 2; m >= 1; m >>= 1) x = fminf(x, __shfl_xor(x, m, LPV));
  return x;
}
template <int LPV> __device__ __forceinline__ uint32_t lanes_min_u(uint32_t x) {
#pragma unroll
  for (int m = LPV / 2; m >= 1; m >>= 1) {
    const uint32_t o = __shfl_xor(x, m, LPV);
    x = o < x ? o : x;
  }
  return x;
}
template <int LPV> __device__ __forceinline__ f3 lanes_sum3(f3 a) {
  return mk(lanes_sum<LPV>(a.x), lanes_sum<LPV>(a.y), lanes_sum<LPV>(a.z));
}
// First item of a type (prefix = items before it in canonical order) that
// lane `sub` handles.
template <int LPV> __device__ __forceinline__ uint32_t lane_first(uint32_t sub, uint32_t prefix) {
  return (sub + LPV - prefix % LPV) % LPV;
}
// Active contact rows above which a vertex's gather terms come from
// k_ipc_local_lanes (16 lanes) instead of the solve's own lanes.
constexpr uint32_t kIpcLocalRows = 128;

template <bool kCloth, bool kIpc, int LPV, bool kTets = true>
__device__ Local gather_local_lanes(const DeviceState& s, const PassParams& p, uint32_t v, uint32_t sub) {
  const float4* pos = s.pos[p.src];
  const f3 x = load_pos(pos, v);
  const uint4 inf = s.vtx_info[v];
  Local L;
  L.g = mk(0, 0, 0);
  L.H = mzero();
  if (sub == 0) {  // the vertex's own terms (S:3083-3096, S:3188-3191)
    const float inertia = s.vtx_rest[v].w * p.inv_dt2;
    L.g = inertia * (x - xyz(s.predicted[v]));
    L.H = mscale(inertia + target_stiffness(s, p, v), mid());
    if (p.objective_flags != 0u) L.g = L.g + objective_gradient(s, p, v, x);
    if (p.floor_k > 0.0f && x.y < p.floor_y) {
      L.g.y += p.floor_k * (x.y - p.floor_y);
      L.H.c1.y += p.floor_k;
    }
  }
  uint32_t prefix = 0;
  if constexpr (kIpc) {  // S:3193-3195
    const uint32_t begin = s.cand_row[v].x, n = s.active_count[v];
    if (p.ipc_local && n > kIpcLocalRows) {  // summed over the rows by k_ipc_local_lanes at this source pose
      if (sub == 0) {
        const float4 a = s.ipc_local[3 * size_t(v)], b = s.ipc_local[3 * size_t(v) + 1],
                     c = s.ipc_local[3 * size_t(v) + 2];
        L.g = L.g + mk(a.x, a.y, a.z);
        L.H.c0 = L.H.c0 + mk(a.w, b.x, b.y);
        L.H.c1 = L.H.c1 + mk(b.z, b.w, c.x);
        L.H.c2 = L.H.c2 + mk(c.y, c.z, c.w);
      }
    } else {
      const bool records = p.fric_records && !p.colored;
#pragma unroll 1
      for (uint32_t k = lane_first<LPV>(sub, prefix); k < n; k += LPV)
        ipc_row_contribution(s, p, v, pos, s.active_ids[begin + k], L.g, L.H, records);
    }
    prefix += n;
  }
  if constexpr (kTets) {
#pragma unroll 1
    for (uint32_t k = lane_first<LPV>(sub, prefix); k < inf.y; k += LPV) {  // S:3097-3111
      const Tet e = load_tet(s, s.adjacency[inf.x + k]);
      const uint32_t slot = local_slot(e.ix, v);
      if (slot >= 4) continue;
      add_incident_block(snh_point(deformation_at(e, pos, v, x), e.lam, e.mu), e, shape_gradient(e, slot), 1.0f, L.g, L.H);
    }
  }
  prefix += inf.y;
  if constexpr (kCloth) {
    const uint2 tr = s.tri_row[v], hr = s.hinge_row[v];
#pragma unroll 1
    for (uint32_t k = lane_first<LPV>(sub, prefix); k < tr.y; k += LPV) {  // S:3117-3144
      const uint32_t tri = s.tri_elems[tr.x + k];
      const uint32_t slot = tri_slot(s.tri_ix[tri], v);
      if (slot >= 3) continue;
      const Membrane m = membrane_base(s, p, tri, pos);
      L.g = L.g + membrane_gradient(s, tri, m, slot);
      L.H = madd(L.H, membrane_local_hessian(s, tri, m, slot));
    }
    prefix += tr.y;
#pragma unroll 1
    for (uint32_t k = lane_first<LPV>(sub, prefix); k < hr.y; k += LPV) {  // S:3145-3184
      const uint32_t h = s.hinge_elems[hr.x + k];
      const uint32_t slot = hinge_slot(s.hinge_ix[h], v);
      if (slot >= 4) continue;
      const Dihedral d = dihedral_base(s, p, h, pos, slot);
      if (!d.valid) {
        L.g = L.g + mk(kF32Max, kF32Max, kF32Max);
      } else {
        const float4 r = s.hinge_rest[h];
        const float ad = wrapped_angle_difference(d.angle, r.x);
        const float k2 = s.cloth_bend * r.y;
        L.g = L.g + (k2 * ad) * d.grad;
        L.H = m3{fma3(k2 * d.grad.x, d.grad, L.H.c0), fma3(k2 * d.grad.y, d.grad, L.H.c1), fma3(k2 * d.grad.z, d.grad, L.H.c2)};
      }
    }
    prefix += hr.y;
  }
  // The cubature records (S:3198-3263). The tetrahedral variant deals them
  // over its lanes behind the incident tetrahedra; the variants that also
  // carry cloth or contact terms read them from k_cub_local_lanes (run by
  // launch_solve before the sweep at this source pose), which keeps their
  // registers within three waves per SIMD without spilling.
  if constexpr (!kTets) {
  } else if constexpr (kCloth || kIpc) {
    if (p.cub_local && sub == 0 && (s.scene_kind[vscene(s, v)] & kKindCubature)) {
      const float4 a = s.cub_local[3 * size_t(v)], b = s.cub_local[3 * size_t(v) + 1],
                   c = s.cub_local[3 * size_t(v) + 2];
      if (c.w > 0.0f) {  // the vertex has a record in use
        L.g = L.g + mk(a.x, a.y, a.z);
        L.H.c0 = L.H.c0 + mk(a.w, b.x, b.y);
        L.H.c1 = L.H.c1 + mk(b.x, b.z, b.w);
        L.H.c2 = L.H.c2 + mk(b.y, b.w, c.x);
      }
    }
  } else {
    const uint32_t k0 = lane_first<LPV>(sub, prefix);
    if (k0 < p.cubature_k) {
      const m3 Rs_t = tr(load_vcub(vcub_source(s, p), v).R);
#pragma unroll 1
      for (uint32_t k = k0; k < p.cubature_k; k += LPV) {
        const size_t rec = size_t(v) * p.cubature_k + k;
        const uint4 head = s.cub[10 * rec];
        const float w = __uint_as_float(head.y);
        if (head.x == kEmptyTet || head.x >= p.tet_count || !(w > 0.0f)) continue;
        cubature_record_snh(s, p, v, x, head.x, reinterpret_cast<const float*>(s.cub + 10 * rec + 1), Rs_t, w, L);
      }
    }
  }
  L.g = lanes_sum3<LPV>(L.g);
  L.H = {lanes_sum3<LPV>(L.H.c0), lanes_sum3<LPV>(L.H.c1), lanes_sum3<LPV>(L.H.c2)};
  return L;
}

// The cubature correction's gather terms (S:3198-3263) of the vertices the
// lane sweep solves next, at its source pose: one record per lane, eight
// lanes per vertex, the lane sums combined by an xor butterfly (a fixed
// order). Inside the sweep a vertex's K records were dealt over its four
// lanes behind the incident tetrahedra, so two lanes evaluated two records
// each in series; here every record of a vertex with K <= 8 is evaluated at
// once, and the sweep reads one record per vertex: {g, H00}, {H01, H02,
// H11, H12}, {H22, 0, 0, records in use} (H is symmetric). Vertices the
// sweep does not solve (pinned, latched, an infeasible source) are skipped.
__global__ void __launch_bounds__(kBlock) k_cub_local_lanes(DeviceState s, PassParams p) {
  constexpr int kLanes = 8;
  const uint32_t gid = lane_sweep_gid(s, p, logical_block(p.xcd_remap), kBlock, kLanes);
  const uint32_t lane = gid / kLanes, sub = gid % kLanes;
  uint32_t v;
  if (p.colored) {
    if (p.color_begin + lane >= p.color_end) return;  // uniform per 8-lane group
    v = s.color_order[p.color_begin + lane];
  } else {
    if (lane >= p.vertex_count) return;
    v = lane;
  }
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || s.scene_lanes[sc] != p.solve_lanes || !kind_selected(s, p, sc) ||
      stop_latched(s, p, sc) || s.vtx_info[v].z != 0u)
    return;
  {  // the sweep's feasibility gate (k_solve_lanes)
    const float4 pre = s.control[4 * size_t(sc)];
    const uint32_t pv = control_bits(pre.w) & kValidityMask;
    const bool amv = (pv & kAcceptedMinimumValid) != 0u && finite_s(pre.z);
    if (!(amv && pre.z > kDeterminantFloor)) return;
  }
  const f3 x = load_pos(s.pos[p.src], v);
  Local L;
  L.g = mk(0, 0, 0);
  L.H = mzero();
  float used = 0.0f;
  if (sub < p.cubature_k) {
    const m3 Rs_t = tr(load_vcub(vcub_source(s, p), v).R);
#pragma unroll 1
    for (uint32_t k = sub; k < p.cubature_k; k += kLanes) {
      const size_t rec = size_t(v) * p.cubature_k + k;
      const uint4 head = s.cub[10 * rec];
      const float w = __uint_as_float(head.y);
      if (head.x == kEmptyTet || head.x >= p.tet_count || !(w > 0.0f)) continue;
      cubature_record_snh(s, p, v, x, head.x, reinterpret_cast<const float*>(s.cub + 10 * rec + 1), Rs_t, w, L);
      used += 1.0f;
    }
  }
  L.g = lanes_sum3<kLanes>(L.g);
  L.H = {lanes_sum3<kLanes>(L.H.c0), lanes_sum3<kLanes>(L.H.c1), lanes_sum3<kLanes>(L.H.c2)};
  used = lanes_sum<kLanes>(used);
  if (sub == 0) {
    s.cub_local[3 * size_t(v)] = make_float4(L.g.x, L.g.y, L.g.z, L.H.c0.x);
    s.cub_local[3 * size_t(v) + 1] = make_float4(L.H.c1.x, L.H.c2.x, L.H.c1.y, L.H.c2.y);
    s.cub_local[3 * size_t(v) + 2] = make_float4(L.H.c2.z, 0.0f, 0.0f, used);
  }
}

// One lane's share of the restricted trial. `fail` is the canonical index of
// the lane's first failing item times two, plus one for invalid geometry
// (zero for an infeasible contact distance); ~0 when none failed.
struct TrialLanes {
  float energy, min_det;
  uint32_t fail;
};
template <bool kCloth, bool kIpc, int LPV, bool kTets = true>
__device__ TrialLanes trial_lanes_part(const DeviceState& s, const PassParams& p, uint32_t v, uint32_t sub, f3 x, f3 d,
                                       float initial_min) {
  const float4* pos = s.pos[p.src];
  const uint4 inf = s.vtx_info[v];
  TrialLanes tr{0.0f, initial_min, 0xffffffffu};
  const f3 pred = xyz(s.predicted[v]);
  if (sub == 0) {
    tr.energy = quadratic_delta(x, pred, d, s.vtx_rest[v].w * p.inv_dt2) + floor_delta(p, x.y, d.y);
    if (p.objective_flags != 0u) tr.energy += objective_delta(s, p, v, x, d);
  }
  uint32_t prefix = 0;
  if constexpr (kIpc) {  // S:1997-2037, S:2140-2193
    const f3 tp = x + d;
    const float dl = sqrtf(dot(d, d));
    const uint32_t begin = s.cand_row[v].x, n = s.active_count[v];
    // Jacobi sweeps read a row's source distance from its summary, and a
    // friction row's source energy and displacement terms from its sweep
    // record; the moved vertex's own term is formed here.
    const bool records = p.fric_records && !p.colored;
    const f3 own = records ? tp - load_pos(s.old, v) : mk(0, 0, 0);
#pragma unroll 1
    for (uint32_t k = lane_first<LPV>(sub, prefix); k < n; k += LPV) {
      const uint32_t c = s.active_ids[begin + k];
      float2 q = make_float2(0.0f, 0.0f);
      if (!p.colored) {  // the same tests on the eight-byte summary first
        q = s.src_d[c];
        if (q.x >= 0.0f && q.x - dl > s.ipc_dmin && !(q.x - dl < p.ipc_dhat) && !(s.ipc_mu > 0.0f && q.y > 0.0f))
          continue;
      }
      if (records) {
        if (!(q.x >= 0.0f)) {  // the cached source contact is invalid
          tr.fail = 2u * (prefix + k) + 1u;
          return tr;
        }
        const bool near = !(q.x - dl > s.ipc_dmin);
        const bool may_change = q.x - dl < p.ipc_dhat;
        const bool fa = s.ipc_mu > 0.0f && q.y > 0.0f;
        if (!near && !may_change && !fa) continue;
        Contact tc = Contact{p.ipc_dhat, mk(1, 0, 0), true, make_float4(0, 0, 0, 0)};
        if (near || may_change) tc = ipc_contact_at(s, c, pos, p.ipc_dhat, v, tp);
        if (near) {
          if (!tc.valid) {
            tr.fail = 2u * (prefix + k) + 1u;
            return tr;
          }
          if (!(tc.distance > s.ipc_dmin)) {
            tr.fail = 2u * (prefix + k);
            return tr;
          }
        }
        if (!may_change && !fa) continue;
        if (fa) {
          const float4* fr = s.fric + 4 * size_t(c);
          const f3 rel = ipc_record_displacement(fr, s.cand[c], s.lag_w[c], v, own);
          tr.energy += ipc_energy_with_rel(s, p, tc, s.lag_n[c], rel) - fr[0].w;
        } else {
          // The source contact's energy is its barrier term alone.
          const Contact sc = Contact{q.x, mk(1, 0, 0), true, make_float4(0, 0, 0, 0)};
          tr.energy += ipc_energy_with_rel(s, p, tc, make_float4(0, 0, 0, 0), mk(0, 0, 0)) -
                       ipc_energy_with_rel(s, p, sc, make_float4(0, 0, 0, 0), mk(0, 0, 0));
        }
        continue;
      }
      const Contact sc = p.colored ? ipc_contact_at(s, c, pos, p.ipc_dhat) : ipc_cached(s, c, false);
      if (!sc.valid) {
        tr.fail = 2u * (prefix + k) + 1u;
        return tr;
      }
      const bool near = !(sc.distance - dl > s.ipc_dmin);
      const bool may_change = sc.distance - dl < p.ipc_dhat;
      const bool fa = ipc_friction_active(s, c);
      if (!near && !may_change && !fa) continue;
      Contact tc = Contact{p.ipc_dhat, mk(1, 0, 0), true, make_float4(0, 0, 0, 0)};
      if (near || may_change) tc = ipc_contact_at(s, c, pos, p.ipc_dhat, v, tp);
      if (near) {
        if (!tc.valid) {
          tr.fail = 2u * (prefix + k) + 1u;
          return tr;
        }
        if (!(tc.distance > s.ipc_dmin)) {
          tr.fail = 2u * (prefix + k);
          return tr;
        }
      }
      if (!may_change && !fa) continue;
      const uint4 ix = s.cand[c];
      tr.energy += ipc_energy_from_contact(s, p, c, ix, tc, pos, v, tp) -
                   ipc_energy_from_contact(s, p, c, ix, sc, pos, kNoVertex, mk(0, 0, 0));
    }
    prefix += n;
  }
  if constexpr (kTets) {
#pragma unroll 1
    for (uint32_t k = lane_first<LPV>(sub, prefix); k < inf.y; k += LPV) {  // S:2038-2061, S:2120-2139
      const Tet e = load_tet(s, s.adjacency[inf.x + k]);
      const f3 gr = shape_gradient(e, local_slot(e.ix, v));
      const m3 F = deformation_at(e, pos, v, x);
      const f3 c12 = cross(F.c1, F.c2), c20 = cross(F.c2, F.c0), c01 = cross(F.c0, F.c1);
      const float J = dot(F.c0, c12);
      const f3 b = gr.x * c12 + gr.y * c20 + gr.z * c01;
      const float jd = dot(d, b);
      const float det_trial = J + jd;
      if (!finite_s(det_trial)) {
        tr.fail = 2u * (prefix + k) + 1u;
        return tr;
      }
      tr.min_det = fminf(tr.min_det, det_trial);
      const float id = 2.0f * dot(d, mul(F, gr)) + dot(d, d) * dot(gr, gr);
      const float lr = log_one_plus(id / (frob(F, F) + 1.0f));
      tr.energy += e.vol * (0.5f * e.mu * id - 0.75f * e.mu * jd + 0.5f * e.lam * jd * (2.0f * J + jd - 2.0f) - 0.5f * e.mu * lr);
    }
  }
  prefix += inf.y;
  if constexpr (kCloth) {
    const f3 tp = x + d;
    const uint2 trr = s.tri_row[v], hr = s.hinge_row[v];
#pragma unroll 1
    for (uint32_t k = lane_first<LPV>(sub, prefix); k < trr.y; k += LPV) {
      const uint32_t tri = s.tri_elems[trr.x + k];
      if (tri_slot(s.tri_ix[tri], v) >= 3) continue;
      // Signed relative area stretch (S:2062-2116).
      const f3 sn = tri_raw_normal(s, tri, pos, kNoVertex, mk(0, 0, 0));
      const f3 tn = tri_raw_normal(s, tri, pos, v, tp);
      const float tra = 2.0f * s.tri_area[tri];
      const float ss = len3(sn) / tra;
      const float tm = len3(tn) / tra;
      const float ts = dot(sn, tn) > 0.0f ? tm : -tm;
      if (!(tra > 0.0f) || !finite_s(ss) || !finite_s(ts)) {
        tr.fail = 2u * (prefix + k) + 1u;
        return tr;
      }
      tr.min_det = fminf(tr.min_det, fminf(ss, ts));
      // Membrane energy delta (S:2203-2236).
      tr.energy += membrane_at(s, tri, pos, v, tp).energy - membrane_base(s, p, tri, pos).energy;
    }
    prefix += trr.y;
#pragma unroll 1
    for (uint32_t k = lane_first<LPV>(sub, prefix); k < hr.y; k += LPV) {  // S:2237-2271
      const uint32_t h = s.hinge_elems[hr.x + k];
      if (hinge_slot(s.hinge_ix[h], v) >= 4) continue;
      tr.energy += bending_energy_at(s, h, pos, v, tp) - bending_base(s, p, h, pos);
    }
    prefix += hr.y;
  }
  if constexpr (kTets) {
    const uint32_t k0 = lane_first<LPV>(sub, prefix);
    if (k0 < p.cubature_k) {  // S:2272-2337
      const float4* vc = vcub_source(s, p);
      const f3 ds = tmul(load_vcub(vc, v).R, d);  // R_s^T d
      const float n_src = float(inf.y > 1u ? inf.y : 1u);
      const float w_src = s.vtx_rest[v].w * p.inv_dt2 / n_src;
#pragma unroll 1
      for (uint32_t k = k0; k < p.cubature_k; k += LPV) {
        const size_t rec = size_t(v) * p.cubature_k + k;
        const uint4 head = s.cub[10 * rec];
        const float w = __uint_as_float(head.y);
        if (head.x == kEmptyTet || head.x >= p.tet_count || !(w > 0.0f)) continue;
        const Tet e = load_tet(s, head.x);
        const float* ub = reinterpret_cast<const float*>(s.cub + 10 * rec + 1);
        const f3 x0 = vcub_pose(vc, e.ix.x), x1 = vcub_pose(vc, e.ix.y), x2 = vcub_pose(vc, e.ix.z),
                 x3 = vcub_pose(vc, e.ix.w);
        const m3 F = mul(m3{x1 - x0, x2 - x0, x3 - x0}, e.inv);
        m3 Dd = mzero();
        float compl_e = 0.0f;
        int ss = 4;
#pragma unroll 1
        for (int l = 0; l < 4; ++l) {
          const uint32_t vl = tet_vertex(e.ix, l);
          if (vl == v) ss = l;
          const VCub ql = load_vcub(vc, vl);
          const f3 bd = mul(ql.R, ubar_times(ub, l, ds));  // B_l d
          const f3 gl = shape_gradient(e, l);
          Dd.c0 = Dd.c0 + gl.x * bd;
          Dd.c1 = Dd.c1 + gl.y * bd;
          Dd.c2 = Dd.c2 + gl.z * bd;
          if (!ql.pinned) {
            compl_e += ql.w * (dot(ql.diff, bd) + 0.5f * dot(bd, bd));
            if (p.objective_flags != 0u) {
              const uint4 il = s.vtx_info[vl];
              const float n = float(il.y > 1u ? il.y : 1u);
              compl_e += objective_delta(s, p, vl, ql.x, bd) / n;
            }
          }
        }
        const float det_trial = mdet(madd(F, Dd));
        if (!finite_s(det_trial)) {
          tr.fail = 2u * (prefix + k) + 1u;
          return tr;
        }
        tr.min_det = fminf(tr.min_det, det_trial);
        compl_e += snh_delta_general(F, Dd, e.lam, e.mu) * e.vol;
        if (ss < 4) {
          const f3 gs = shape_gradient(e, ss);
          const f3 c12 = cross(F.c1, F.c2), c20 = cross(F.c2, F.c0), c01 = cross(F.c0, F.c1);
          const float J = dot(F.c0, c12);
          const f3 b = gs.x * c12 + gs.y * c20 + gs.z * c01;
          const float jd = dot(d, b);
          const float id = 2.0f * dot(d, mul(F, gs)) + dot(d, d) * dot(gs, gs);
          const float lr = log_one_plus(id / (frob(F, F) + 1.0f));
          compl_e -= e.vol * (0.5f * e.mu * id - 0.75f * e.mu * jd + 0.5f * e.lam * jd * (2.0f * J + jd - 2.0f) - 0.5f * e.mu * lr);
          compl_e -= quadratic_delta(x, pred, d, w_src);
          if (p.objective_flags != 0u) compl_e -= objective_delta(s, p, v, x, d) / n_src;
        }
        tr.energy += w * compl_e;
      }
    }
  }
  return tr;
}
template <bool kCloth, bool kIpc, int LPV, bool kTets = true>
__device__ Trial trial_lanes(const DeviceState& s, const PassParams& p, uint32_t v, uint32_t sub, f3 x, f3 d,
                             float initial_min) {
  const TrialLanes part = trial_lanes_part<kCloth, kIpc, LPV, kTets>(s, p, v, sub, x, d, initial_min);
  const uint32_t fail = lanes_min_u<LPV>(part.fail);
  const float min_det = lanes_min<LPV>(part.min_det);
  const float energy = lanes_sum<LPV>(part.energy);
  if (fail != 0xffffffffu) {
    if (fail & 1u) return Trial{0.0f, false, 0.0f};
    return Trial{0.0f, true, 0.0f};  // an infeasible contact distance
  }
  return Trial{min_det, true, energy};
}

// vertexPolarRotation with LPV lanes per vertex splitting the incident
// tetrahedra of the volume-weighted average.
template <int LPV>
__global__ void __launch_bounds__(kBlock) k_vertex_polar_lanes(DeviceState s, PassParams p) {
  const uint32_t gid = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  const uint32_t v = gid / LPV, sub = gid % LPV;
  if (v >= p.vertex_count) return;
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || stop_latched(s, p, sc)) return;
  if (p.colored && p.color_index >= s.scene_cnt[sc].z) return;
  const uint4 inf = s.vtx_info[v];
  const f3 x = load_pos(s.pos[p.src], v);
  m3 avg = mzero();
  float tv = 0.0f;
#pragma unroll 1
  for (uint32_t k = sub; k < inf.y; k += LPV) {
    const Tet e = load_tet(s, s.adjacency[inf.x + k]);
    avg = mfma(e.vol, deformation_at(e, s.pos[p.src], v, x), avg);
    tv += e.vol;
  }
  avg = {lanes_sum3<LPV>(avg.c0), lanes_sum3<LPV>(avg.c1), lanes_sum3<LPV>(avg.c2)};
  tv = lanes_sum<LPV>(tv);
  if (sub != 0) return;
  m3 R = mid();
  if (inf.y != 0u) R = polar_rotation(tv > 0.0f ? mscale(1.0f / tv, avg) : mid(), p.polar_eps);
  store_rot(s.vrot, v, R);
  store_vcub(s, p, s.vcub[0], v, inf, R, x);
}

// jgs2Solve (S:2992-3487) with LPV lanes per vertex (Jacobi: every vertex;
// graph-colored Gauss-Seidel: the vertices of the active color, in place).
// Three waves per SIMD: the register budget drops from 177 to 168 VGPRs at the
// price of a few spilled registers (32 B/lane of scratch), which measures about
// 15% faster than two waves on both the mixed batch and the layered box; four
// waves spill over 200 B/lane. kTets false: the launch covers only scenes
// whose unpinned vertices have no incident tetrahedra and no cubature
// records, so those terms (zero there) are not compiled; the cloth and
// contact form then fits 168 VGPRs without spilling.
template <bool kCloth, bool kIpc, int LPV, bool kTets = true>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_solve_lanes(DeviceState s, PassParams p) {
  constexpr bool kEnergyInSolve = !kCloth && !kIpc;
  // In a mixed batch the vertex-energy pass after the sweep covers every
  // scene, so the tetrahedral variant leaves the energy to it.
  const bool energy_here = !p.colored && p.kind_filter == 0u;
  const uint32_t gid = lane_sweep_gid(s, p, logical_block(p.xcd_remap), blockDim.x, LPV);
  const uint32_t lane = gid / LPV, sub = gid % LPV;
  uint32_t v;
  if (p.colored) {
    if (p.color_begin + lane >= p.color_end) return;
    v = s.color_order[p.color_begin + lane];
  } else {
    if (lane >= p.vertex_count) return;
    v = lane;
  }
  // Every branch below is uniform over a vertex's LPV lanes.
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || s.scene_lanes[sc] != uint32_t(LPV) || !kind_selected(s, p, sc)) return;
  float4* tgt = s.pos[p.tgt];
  if (stop_latched(s, p, sc)) {
    if (sub == 0) tgt[v] = s.pos[p.src][v];
    return;
  }
  if (p.fused_rot && sub < 3) {  // this sweep's vertexPolarRotation result (S:2960-2989)
    s.vrot[3 * size_t(v) + sub] = s.vrot_next[3 * size_t(v) + sub];
  }
  if (s.vtx_info[v].z != 0u) {  // S:2997-3003
    if (sub == 0) {
      const float4 r = s.vtx_rest[v];
      tgt[v] = make_float4(r.x, r.y, r.z, 1.0f);
      s.final_update[v] = make_float4(0, 0, 0, 1);
      store_local_diag(s, v, mk(0, 0, 0), 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, kPinned);
      if (kEnergyInSolve && energy_here) store_vertex_energy<false, false>(s, p, v);
    }
    return;
  }
  const f3 x = load_pos(s.pos[p.src], v);
  float initial_min = kF32Max;
  {
    const float4 pre = s.control[4 * size_t(sc)];
    const uint32_t pv = control_bits(pre.w) & kValidityMask;
    const bool amv = (pv & kAcceptedMinimumValid) != 0u && finite_s(pre.z);
    if (amv) initial_min = pre.z;
    if (!(amv && pre.z > kDeterminantFloor)) {
      if (sub == 0) {
        tgt[v] = s.pos[p.src][v];
        s.final_update[v] = make_float4(0, 0, 0, 1);
        store_local_diag(s, v, mk(0, 0, 0), 0, 0, 0, 0, 0, 0, 0, 0, 0, amv ? pre.z : 0.0f, 0, 0,
                         amv ? kLineSearchFailed : kNonfinite);
        if (kEnergyInSolve && energy_here) store_vertex_energy<false, false>(s, p, v);
      }
      return;
    }
  }
  const Local L = gather_local_lanes<kCloth, kIpc, LPV, kTets>(s, p, v, sub);
  const float inertia = s.vtx_rest[v].w * p.inv_dt2;
  // Globalized Newton direction and line search (S:3265-3471).
  const SolveOut dr = globalization_solve(L.H, L.g, inertia);
  uint32_t status = kNonfinite;
  const f3 full = x + dr.dir;
  const f3 fd = full - x;
  const float xs = fmaxf(fabsf(x.x), fmaxf(fabsf(x.y), fabsf(x.z)));
  const float resolution = fmaxf(s.scene_scale[sc], fmaxf(xs, 1.0e-12f)) * kUnitRoundoff * kPositionResolutionMultiplier;
  const bool below = dr.status == kAccepted && finite3(full) && finite3(fd) && scaled_length3(fd) <= resolution;
  if (dr.status == kAccepted) status = below ? kZeroGradient : kLineSearchFailed;
  else if (dr.status == kZeroGradient) status = kZeroGradient;
  else if (dr.status == kShiftLimit) status = kShiftLimit;
  else if (dr.status == kCholeskyFailed) status = kCholeskyFailed;
  else if (dr.status == kNonDescent) status = kNonDescent;
  float alpha = 0.0f, accepted_de = 0.0f, bound = 0.0f, min_trial = 0.0f;
  uint32_t backtracks = 0, evaluations = 0;
  f3 accepted_d = mk(0, 0, 0), accepted_x = x;
  if (dr.status == kAccepted && !below) {
    float ta = 1.0f;
    const float dl = scaled_length3(dr.dir);
    if (p.max_step > 0.0f && dl > p.max_step) ta = p.max_step / dl;
    for (uint32_t b = 0; b <= kMaxBacktracks; ++b) {
      const f3 tp = x + ta * dr.dir;
      const f3 td = tp - x;
      if (!finite3(tp) || !finite3(td)) {
        status = kNonfinite;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      const float tl = scaled_length3(td);
      if (p.max_step > 0.0f && tl > p.max_step) {
        ta *= 0.5f;
        continue;
      }
      if (tl <= resolution) {
        status = kZeroGradient;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      const float tgd = dot(L.g, td);
      if (!finite_s(tgd)) {
        status = kNonfinite;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      if (!(tgd < 0.0f)) {
        status = kZeroGradient;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      const Trial trial = trial_lanes<kCloth, kIpc, LPV, kTets>(s, p, v, sub, x, td, initial_min);
      if (!trial.geometry_valid) {
        status = kNonfinite;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      min_trial = trial.min_det;
      if (trial.min_det > kDeterminantFloor) {
        evaluations += 1;
        if (!finite_s(trial.energy)) {
          status = kNonfinite;
          backtracks = b;
          break;
        }
        const float tb = kArmijoC1 * tgd;
        if (trial.energy <= tb) {
          alpha = ta;
          accepted_de = trial.energy;
          bound = tb;
          backtracks = b;
          status = kAccepted;
          accepted_d = td;
          accepted_x = tp;
          break;
        }
      }
      ta *= 0.5f;
    }
    if (status == kLineSearchFailed) backtracks = kMaxBacktracks;
  }
  if (sub == 0) {
    tgt[v] = make_float4(accepted_x.x, accepted_x.y, accepted_x.z, 1.0f);
    s.final_update[v] = make_float4(scaled_length3(accepted_d), 0, 0, 1);
    store_local_diag(s, v, dr.dir, alpha, dr.min_eig * dr.scale, dr.scale, dr.shift * dr.scale, dr.shift, accepted_de,
                     bound, dr.gdotp, dr.residual, min_trial, backtracks, evaluations, status);
    if (kEnergyInSolve && energy_here) store_vertex_energy<false, false>(s, p, v);
  }
}

// Cloth element records (membrane state; dihedral angle, validity, four
// corner gradients and bending energy) at positions `which`, once per
// element, for the gathers and trials of the vertices that share it.
// Modes: 0 the records at `which`; 1 the same, each element's previous
// energy first kept in its spare word (the records held the sweep's source
// pose; latched scenes included, whose target equals their source); 2 the
// elements of scenes whose candidate k_apply_candidate rejected (their
// target was restored to the source pose the records held before mode 1).
__device__ __forceinline__ void cloth_tri_record(const DeviceState& s, const PassParams& p, const float4* pos,
                                                 uint32_t t, uint32_t mode) {
  const uint32_t sc = s.rblock_scene[t / kSceneBlock];
  if (sc == kNoScene || t >= s.scene_end[sc].z || (mode != 1u && stop_latched(s, p, sc))) return;
  if (mode == 2u && s.control[4 * size_t(sc) + 1].x == 1.0f) return;
  float4* r = s.tri_base + 3 * size_t(t);
  // Word 7: the energy at the pose the records held; word 11: bit 0 the
  // stretch test here, bit 1 the one at that pose.
  const float se = mode == 1u ? r[0].w : 0.0f;
  const uint32_t sf = mode == 1u ? (__float_as_uint(r[2].w) & 1u) << 1 : 0u;
  const Membrane m = membrane_at(s, t, pos, kNoVertex, mk(0, 0, 0));
  const uint32_t fl = sf | (tri_stretch_ok(s, t, pos) ? 1u : 0u);
  r[0] = make_float4(m.f0.x, m.f0.y, m.f0.z, m.energy);
  r[1] = make_float4(m.f1.x, m.f1.y, m.f1.z, se);
  r[2] = make_float4(m.stress.x, m.stress.y, m.stress.z, __uint_as_float(fl));
}
__device__ __forceinline__ void cloth_hinge_record(const DeviceState& s, const PassParams& p, const float4* pos,
                                                   uint32_t h, uint32_t mode) {
  const uint4 ix = s.hinge_ix[h];
  if (ix.x >= p.vertex_count) return;  // an entry no resident scene owns
  const uint32_t sc = vscene(s, ix.x);
  if (sc == kNoScene || (mode != 1u && stop_latched(s, p, sc))) return;
  if (mode == 2u && s.control[4 * size_t(sc) + 1].x == 1.0f) return;
  float4* r = s.hinge_base + 4 * size_t(h);
  const float se = mode == 1u ? r[2].w : 0.0f;
  const Dihedral d = dihedral_at(s, h, pos, kNoVertex, mk(0, 0, 0), 4u, r);
  r[0].w = d.angle;
  r[1].w = d.valid ? 1.0f : 0.0f;
  r[2].w = bending_energy_of(s, h, d);
  r[3].w = se;
}
__global__ void __launch_bounds__(kBlock) k_cloth_base(DeviceState s, PassParams p, uint32_t which, uint32_t tri_blocks,
                                                       uint32_t mode) {
  const float4* pos = s.pos[which];
  if (blockIdx.x < tri_blocks) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < p.tri_count) cloth_tri_record(s, p, pos, t, mode);
    return;
  }
  const uint32_t h = (blockIdx.x - tri_blocks) * kBlock + threadIdx.x;
  if (h < p.hinge_count) cloth_hinge_record(s, p, pos, h, mode);
}
// Mode 2 over a small grid: a block first scans the scenes' accept flags
// and leaves when every scene accepted its candidate (the common case), so
// a sweep without rejections pays one launch of kClothRejectBlocks blocks.
constexpr uint32_t kClothRejectBlocks = 1024;
__global__ void __launch_bounds__(kBlock) k_cloth_base_rejected(DeviceState s, PassParams p, uint32_t which) {
  __shared__ uint32_t any;
  if (threadIdx.x == 0) any = 0u;
  __syncthreads();
  for (uint32_t sc = threadIdx.x; sc < p.scene_count; sc += kBlock)  // resident scenes have vertices
    if (s.scene_cnt[sc].x != 0u && s.control[4 * size_t(sc) + 1].x != 1.0f && !stop_latched(s, p, sc)) any = 1u;
  __syncthreads();
  if (!any) return;  // uniform per block
  const float4* pos = s.pos[which];
  for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < p.tri_count; t += gridDim.x * kBlock)
    cloth_tri_record(s, p, pos, t, 2u);
  for (uint32_t h = blockIdx.x * kBlock + threadIdx.x; h < p.hinge_count; h += gridDim.x * kBlock)
    cloth_hinge_record(s, p, pos, h, 2u);
}

// jgs2Solve (S:2992-3487), one thread per vertex (Jacobi) or per vertex of
// the active color (graph-colored Gauss-Seidel, in place).
// With cloth the assembled vertex energy reads neighbor positions, so it is
// evaluated by a separate pass after the sweep (k_vertex_energy).
template <bool kCorot, bool kCloth, bool kIpc>
__global__ void __launch_bounds__(kBlock) k_solve(DeviceState s, PassParams p) {
  constexpr bool kEnergyInSolve = !kCloth && !kIpc;
  const uint32_t lane = logical_block(p.xcd_remap) * blockDim.x + threadIdx.x;
  uint32_t v;
  if (p.colored) {
    if (p.color_begin + lane >= p.color_end) return;
    v = s.color_order[p.color_begin + lane];
  } else {
    if (lane >= p.vertex_count) return;
    v = lane;
  }
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || (!kCorot && s.scene_lanes[sc] != 1u)) return;
  float4* tgt = s.pos[p.tgt];
  if (stop_latched(s, p, sc)) {
    tgt[v] = s.pos[p.src][v];
    return;
  }
  if (p.fused_rot) {  // this sweep's vertexPolarRotation result (S:2960-2989)
    const float4* n = s.vrot_next + 3 * size_t(v);
    float4* o = s.vrot + 3 * size_t(v);
    o[0] = n[0];
    o[1] = n[1];
    o[2] = n[2];
  }
  const uint4 inf = s.vtx_info[v];
  if (inf.z != 0u) {
    const float4 r = s.vtx_rest[v];
    tgt[v] = make_float4(r.x, r.y, r.z, 1.0f);
    s.final_update[v] = make_float4(0, 0, 0, 1);
    if (p.globalization) {
      store_local_diag(s, v, mk(0, 0, 0), 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, kPinned);
      if (kEnergyInSolve && !p.colored) store_vertex_energy<false, false>(s, p, v);
    }
    return;
  }
  const f3 x = load_pos(s.pos[p.src], v);
  float initial_min = kF32Max;
  if (p.globalization) {
    const float4 pre = s.control[4 * size_t(sc)];
    const uint32_t pv = control_bits(pre.w) & kValidityMask;
    const bool amv = (pv & kAcceptedMinimumValid) != 0u && finite_s(pre.z);
    if (amv) initial_min = pre.z;
    if (!(amv && pre.z > kDeterminantFloor)) {
      tgt[v] = s.pos[p.src][v];
      s.final_update[v] = make_float4(0, 0, 0, 1);
      store_local_diag(s, v, mk(0, 0, 0), 0, 0, 0, 0, 0, 0, 0, 0, 0, amv ? pre.z : 0.0f, 0, 0,
                       amv ? kLineSearchFailed : kNonfinite);
      if (kEnergyInSolve && !p.colored) store_vertex_energy<false, false>(s, p, v);
      return;
    }
  }
  const Local L = gather_local<kCorot, kCloth, kIpc>(s, p, v);
  const float inertia = s.vtx_rest[v].w * p.inv_dt2;
  if constexpr (kCorot) {  // legacy regularized solve (S:2479-2512, S:3475-3486)
    const m3& m = L.H;
    const f3 rhs = -L.g;
    float dsc = fmaxf(1.0f, fmaxf(fabsf(m.c0.x), fmaxf(fabsf(m.c1.y), fabsf(m.c2.z))));
    float shift = fmaxf(p.regularization * dsc, 1.0e-9f);
    float a00 = m.c0.x + shift, a11 = m.c1.y + shift, a22 = m.c2.z + shift;
    float a01 = 0.5f * (m.c0.y + m.c1.x), a02 = 0.5f * (m.c0.z + m.c2.x), a12 = 0.5f * (m.c1.z + m.c2.y);
    float l00 = sqrtf(fmaxf(a00, shift));
    float l10 = a01 / l00, l20 = a02 / l00;
    float l11 = sqrtf(fmaxf(a11 - l10 * l10, shift));
    float l21 = (a12 - l20 * l10) / l11;
    float l22 = sqrtf(fmaxf(a22 - l20 * l20 - l21 * l21, shift));
    float y0 = rhs.x / l00, y1 = (rhs.y - l10 * y0) / l11, y2 = (rhs.z - l20 * y0 - l21 * y1) / l22;
    float x2 = y2 / l22, x1 = (y1 - l21 * x2) / l11, x0 = (y0 - l10 * x1 - l20 * x2) / l00;
    f3 delta = mk(x0, x1, x2);
    float sq = dot(delta, delta);
    if (!(sq >= 0.0f && sq < 1.0e30f)) delta = mk(0, 0, 0);
    float dl2 = dot(delta, delta);
    if (p.max_step > 0.0f && dl2 > p.max_step * p.max_step) delta = delta * (p.max_step * (1.0f / sqrtf(dl2)));
    const f3 nx = x + delta;
    tgt[v] = make_float4(nx.x, nx.y, nx.z, 1.0f);
    s.final_update[v] = make_float4(sqrtf(dot(delta, delta)), 0, 0, 1);
    return;
  }
  // Globalized Newton direction and line search (S:3265-3471).
  const SolveOut dr = globalization_solve(L.H, L.g, inertia);
  uint32_t status = kNonfinite;
  const f3 full = x + dr.dir;
  const f3 fd = full - x;
  const float xs = fmaxf(fabsf(x.x), fmaxf(fabsf(x.y), fabsf(x.z)));
  const float resolution = fmaxf(s.scene_scale[sc], fmaxf(xs, 1.0e-12f)) * kUnitRoundoff * kPositionResolutionMultiplier;
  const bool below = dr.status == kAccepted && finite3(full) && finite3(fd) && scaled_length3(fd) <= resolution;
  if (dr.status == kAccepted) status = below ? kZeroGradient : kLineSearchFailed;
  else if (dr.status == kZeroGradient) status = kZeroGradient;
  else if (dr.status == kShiftLimit) status = kShiftLimit;
  else if (dr.status == kCholeskyFailed) status = kCholeskyFailed;
  else if (dr.status == kNonDescent) status = kNonDescent;
  float alpha = 0.0f, accepted_de = 0.0f, bound = 0.0f, min_trial = 0.0f;
  uint32_t backtracks = 0, evaluations = 0;
  f3 accepted_d = mk(0, 0, 0), accepted_x = x;
  if (dr.status == kAccepted && !below) {
    float ta = 1.0f;
    const float dl = scaled_length3(dr.dir);
    if (p.max_step > 0.0f && dl > p.max_step) ta = p.max_step / dl;
    for (uint32_t b = 0; b <= kMaxBacktracks; ++b) {
      const f3 tp = x + ta * dr.dir;
      const f3 td = tp - x;
      if (!finite3(tp) || !finite3(td)) {
        status = kNonfinite;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      const float tl = scaled_length3(td);
      if (p.max_step > 0.0f && tl > p.max_step) {
        ta *= 0.5f;
        continue;
      }
      if (tl <= resolution) {
        status = kZeroGradient;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      const float tgd = dot(L.g, td);
      if (!finite_s(tgd)) {
        status = kNonfinite;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      if (!(tgd < 0.0f)) {
        status = kZeroGradient;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      const Trial trial = restricted_trial<kCloth, kIpc>(s, p, v, x, td, initial_min);
      if (!trial.geometry_valid) {
        status = kNonfinite;
        min_trial = 0.0f;
        backtracks = b;
        break;
      }
      min_trial = trial.min_det;
      if (trial.min_det > kDeterminantFloor) {
        evaluations += 1;
        if (!finite_s(trial.energy)) {
          status = kNonfinite;
          backtracks = b;
          break;
        }
        const float tb = kArmijoC1 * tgd;
        if (trial.energy <= tb) {
          alpha = ta;
          accepted_de = trial.energy;
          bound = tb;
          backtracks = b;
          status = kAccepted;
          accepted_d = td;
          accepted_x = tp;
          break;
        }
      }
      ta *= 0.5f;
    }
    if (status == kLineSearchFailed) backtracks = kMaxBacktracks;
  }
  tgt[v] = make_float4(accepted_x.x, accepted_x.y, accepted_x.z, 1.0f);
  s.final_update[v] = make_float4(scaled_length3(accepted_d), 0, 0, 1);
  store_local_diag(s, v, dr.dir, alpha, dr.min_eig * dr.scale, dr.scale, dr.shift * dr.scale, dr.shift, accepted_de,
                   bound, dr.gdotp, dr.residual, min_trial, backtracks, evaluations, status);
  if (kEnergyInSolve && !p.colored) store_vertex_energy<false, false>(s, p, v);
}

// assembledVertexEnergy (S:3490-3496). After a Jacobi sweep with cloth it
// stands in for the in-solve evaluation and, like the solve, leaves the
// records alone once the stop latch is set (`energy_after_sweep`).
template <bool kCloth, bool kIpc>
__global__ void __launch_bounds__(kBlock) k_vertex_energy(DeviceState s, PassParams p) {
  const uint32_t v = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (v >= p.vertex_count) return;
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || (p.energy_after_sweep && stop_latched(s, p, sc))) return;
  store_vertex_energy<kCloth, kIpc>(s, p, v);
}
// ipcOwnedEnergy (S:750-779) of every vertex at the source and the target
// pose, for the vertex-energy pass. Sixteen lanes test a vertex's active
// rows round robin; the owned terms are then added in row order (ballot),
// so each sum is the sequential one bit for bit; any invalid row or
// non-finite term gives the f32 sentinel as the sequential loop's early
// return does.
// The owned contact energy of every active candidate at the source and the
// target pose, once per candidate (not once per each of its four rows): a
// record {energy at the source, energy at the target, flags, owner vertex}.
// Flags: bit w (w = 0 source, 1 target) the cached distance is invalid (the
// row makes its vertices' energies invalid), bit 2 + w the row is active
// there (barrier within reach or lagged friction), the energy then being
// the owner's term.
__global__ void __launch_bounds__(kBlock) k_ipc_energy_candidates(DeviceState s, PassParams p) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t sc = s.cblock_scene[blockIdx.x];
  if (sc == kNoScene || c >= s.scene_end[sc].w) return;
  if ((p.energy_after_sweep && stop_latched(s, p, sc)) || s.cand_keep[c] == 0u) return;
  const float2 q = s.src_d[c];
  const uint4 ix = s.cand[c];
  const uint32_t owner = ipc_energy_owner(ix, s.cand_meta[c]);
  const bool use_target = p.tgt != p.src && s.contact_ctl[sc].w != 0.0f;
  float e[2] = {0.0f, 0.0f};
  uint32_t flags = 0;
#pragma unroll
  for (int which = 0; which < 2; ++which) {
    const bool target = which == 1 && use_target;
    const float dq = target ? s.tgt_d[c].x : q.x;
    if (dq < 0.0f) {
      flags |= 1u << which;
    } else if (ipc_barrier_active(dq, p.ipc_dhat) || (s.ipc_mu > 0.0f && q.y > 0.0f)) {
      const Contact k = ipc_cached(s, c, target);
      e[which] = ipc_energy_from_contact(s, p, c, ix, k, which ? s.pos[p.tgt] : s.pos[p.src], kNoVertex, mk(0, 0, 0));
      flags |= 4u << which;
    }
  }
  s.ipc_erec[c] = make_float4(e[0], e[1], __uint_as_float(flags), __uint_as_float(owner));
}
// Per vertex, 16 lanes: the owned energies of its active rows summed in row
// order (S:2643, ipcCachedContactAt S:450-460), from the candidate records.
__global__ void __launch_bounds__(kBlock) k_ipc_energy_lanes(DeviceState s, PassParams p) {
  constexpr uint32_t kLanes = 16;
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t v = gid / kLanes, sub = gid % kLanes;
  if (v >= p.vertex_count) return;  // uniform per 16-lane group
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || (p.energy_after_sweep && stop_latched(s, p, sc))) return;
  const uint32_t group = (threadIdx.x % 64u) & ~(kLanes - 1u);
  const uint32_t first = s.cand_row[v].x, last = first + s.active_count[v];
  float r[2] = {0.0f, 0.0f};
  bool bad[2] = {false, false};
  for (uint32_t base = first; base < last; base += kLanes) {  // same trip count on the group's lanes
    const uint32_t i = base + sub;
    float4 rec = make_float4(0, 0, 0, 0);
    if (i < last) rec = s.ipc_erec[s.active_ids[i]];
    const uint32_t flags = __float_as_uint(rec.z);
    const bool mine = i < last && __float_as_uint(rec.w) == v;
#pragma unroll
    for (int which = 0; which < 2; ++which) {
      bad[which] = bad[which] || (flags & (1u << which)) != 0u;
      const bool owned = mine && (flags & (4u << which)) != 0u;
      const float e = which ? rec.y : rec.x;
      bad[which] = bad[which] || (owned && !finite_s(e));
      uint32_t mask = uint32_t(__ballot(owned) >> group) & 0xffffu;
      while (mask) {
        const uint32_t j = uint32_t(__builtin_ctz(mask));
        mask &= mask - 1u;
        r[which] += __shfl(e, int(group + j));
      }
    }
  }
  float out[2];
#pragma unroll
  for (int which = 0; which < 2; ++which) {
    const bool b = (uint32_t(__ballot(bad[which]) >> group) & 0xffffu) != 0u;
    out[which] = b ? kF32Max : r[which];
  }
  if (sub == 0) s.ipc_energy[v] = make_float2(out[0], out[1]);
}

// ------------------------------------------------------------------ candidate reduction
// Block partial of reduceCandidate (S:3612-3955). The reference reduces
// with one 128-lane workgroup; here every block reduces a contiguous range
// and a final single-block pass combines the partials in block order. Min,
// max and integer counts are order independent; the energy sums use a fixed
// order, so results are deterministic run to run.
struct CandPartial {
  float minj0, minj1, e0, e1, max_shift;
  uint32_t failures, max_bt, skipped, evals, bits;
};
__device__ __forceinline__ void cand_identity(CandPartial& r) {
  r = CandPartial{kF32Max, kF32Max, 0.0f, 0.0f, 0.0f, 0u, 0u, 0u, 0u, kReductionMask};
}
__device__ __forceinline__ void cand_add_energy(CandPartial& r, int which, float value, bool valid) {
  const uint32_t bit = which == 0 ? kSourceEnergyValid : kCandidateEnergyValid;
  float& acc = which == 0 ? r.e0 : r.e1;
  if (valid) {
    float sum = acc + value;
    if (finite_s(sum)) acc = sum;
    else {
      acc = 0.0f;
      r.bits &= ~bit;
    }
  } else {
    r.bits &= ~bit;
  }
}
__device__ __forceinline__ void cand_combine(CandPartial& l, const CandPartial& r) {
  l.minj0 = fminf(l.minj0, r.minj0);
  l.minj1 = fminf(l.minj1, r.minj1);
  float s0 = l.e0 + r.e0, s1 = l.e1 + r.e1;
  if (finite_s(s0)) l.e0 = s0;
  else {
    l.e0 = 0.0f;
    l.bits &= ~kSourceEnergyValid;
  }
  if (finite_s(s1)) l.e1 = s1;
  else {
    l.e1 = 0.0f;
    l.bits &= ~kCandidateEnergyValid;
  }
  l.max_shift = fmaxf(l.max_shift, r.max_shift);
  l.failures += r.failures;
  l.max_bt = l.max_bt > r.max_bt ? l.max_bt : r.max_bt;
  l.skipped += r.skipped;
  l.evals += r.evals;
  l.bits &= r.bits;
}
template <class P, class Combine>
__device__ __forceinline__ void block_reduce(P& mine, P* shared, Combine combine) {
  shared[threadIdx.x] = mine;
  __syncthreads();
  for (uint32_t stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      P a = shared[threadIdx.x];
      combine(a, shared[threadIdx.x + stride]);
      shared[threadIdx.x] = a;
    }
    __syncthreads();
  }
  mine = shared[0];
}

__device__ __forceinline__ bool local_status_failure(uint32_t st) {  // S:3603-3609
  return st == kShiftLimit || st == kNonDescent || st == kLineSearchFailed || st == kNonfinite || st == kCholeskyFailed;
}

// candidateTetrahedron (S:3499-3601) fused with the tetrahedral half of the
// candidate reduction; blocks past the tetrahedra reduce the vertex records.
// Blocks [0, tb) cover tetrahedra, [tb, tb + rb) cloth triangles and the
// rest vertices; each block lies inside one scene's range.
// `win` = {first tetrahedron block, count, first triangle block, count} and
// `v0` the first vertex block of the blocks this launch covers (the whole
// arena, or one scene's ranges for its admission preflight); partials land
// at their arena-wide positions either way.
__global__ void __launch_bounds__(kBlock) k_candidate_partials(DeviceState s, PassParams p, uint32_t tb, uint32_t rb,
                                                               uint4 win, uint32_t v0) {
  __shared__ CandPartial sh[kBlock];
  CandPartial r;
  cand_identity(r);
  uint32_t b = blockIdx.x;
  if (b < win.y) b = win.x + b;
  else if (b < win.y + win.w) b = tb + win.z + (b - win.y);
  else b = tb + rb + v0 + (b - win.y - win.w);
  uint32_t sc, base, kind;
  if (b < tb) {
    sc = s.tblock_scene[b];
    base = b * kBlock;
    kind = 0;
  } else if (b < tb + rb) {
    sc = s.rblock_scene[b - tb];
    base = (b - tb) * kBlock;
    kind = 1;
  } else {
    sc = s.vblock_scene[b - tb - rb];
    base = (b - tb - rb) * kBlock;
    kind = 2;
  }
  if (sc == kNoScene) return;  // uniform per block
  const bool latched = stop_latched(s, p, sc);
  if (kind == 0) {
    const uint32_t t = base + threadIdx.x;
    if (t < s.scene_end[sc].y) {
      float4 geo, en;
      if (!latched) {
        const Tet e = load_tet(s, t);
        const m3 sf = deformation(e, s.pos[p.src]);
        const m3 cf = deformation(e, s.pos[p.tgt]);
        const float sj = mdet(sf), cj = mdet(cf);
        const bool sgv = finite_s(sj), cgv = finite_s(cj);
        float se = 0.0f, ce = 0.0f;
        bool sev = false, cev = false;
        if (sgv && sj > kDeterminantFloor) {
          se = e.vol * snh_energy(sf, e.lam, e.mu);
          sev = finite_s(se);
        }
        if (cgv && cj > kDeterminantFloor) {
          ce = e.vol * snh_energy(cf, e.lam, e.mu);
          cev = finite_s(ce);
        }
        geo = make_float4(sgv ? sj : 0.0f, cgv ? cj : 0.0f, sgv ? 1.0f : 0.0f, cgv ? 1.0f : 0.0f);
        en = make_float4(sev ? se : 0.0f, cev ? ce : 0.0f, sev ? 1.0f : 0.0f, cev ? 1.0f : 0.0f);
        s.tet_glob[2 * size_t(t)] = geo;
        s.tet_glob[2 * size_t(t) + 1] = en;
      } else {
        geo = s.tet_glob[2 * size_t(t)];
        en = s.tet_glob[2 * size_t(t) + 1];
      }
      const bool sgv = geo.z == 1.0f && finite_s(geo.x);
      const bool cgv = geo.w == 1.0f && finite_s(geo.y);
      if (sgv) r.minj0 = fminf(r.minj0, geo.x);
      else r.bits &= ~kSourceGeometryValid;
      if (cgv) r.minj1 = fminf(r.minj1, geo.y);
      else r.bits &= ~kCandidateGeometryValid;
      cand_add_energy(r, 0, en.x, en.z == 1.0f && finite_s(en.x));
      cand_add_energy(r, 1, en.y, en.w == 1.0f && finite_s(en.y));
    }
  } else if (kind == 1) {
    const uint32_t tri = base + threadIdx.x;
    if (tri < s.scene_end[sc].z) {
      float4 geo;
      if (!latched) {
        // Cloth triangles: signed relative area stretch (S:3508-3549).
        const f3 sn = tri_raw_normal(s, tri, s.pos[p.src], kNoVertex, mk(0, 0, 0));
        const f3 cn = tri_raw_normal(s, tri, s.pos[p.tgt], kNoVertex, mk(0, 0, 0));
        const float tra = 2.0f * s.tri_area[tri];
        const float ss = len3(sn) / tra;
        const float cm = len3(cn) / tra;
        const float cs = dot(sn, cn) > 0.0f ? cm : -cm;
        const bool sgv = tra > 0.0f && finite3(sn) && finite_s(ss);
        const bool cgv = tra > 0.0f && finite3(cn) && finite_s(cs);
        geo = make_float4(sgv ? ss : 0.0f, cgv ? cs : 0.0f, sgv ? 1.0f : 0.0f, cgv ? 1.0f : 0.0f);
        s.tri_glob[2 * size_t(tri)] = geo;
        s.tri_glob[2 * size_t(tri) + 1] = make_float4(0.0f, 0.0f, 1.0f, 1.0f);
      } else {
        geo = s.tri_glob[2 * size_t(tri)];
      }
      const bool sgv = geo.z == 1.0f && finite_s(geo.x);
      const bool cgv = geo.w == 1.0f && finite_s(geo.y);
      if (sgv) r.minj0 = fminf(r.minj0, geo.x);
      else r.bits &= ~kSourceGeometryValid;
      if (cgv) r.minj1 = fminf(r.minj1, geo.y);
      else r.bits &= ~kCandidateGeometryValid;
      cand_add_energy(r, 0, 0.0f, true);
      cand_add_energy(r, 1, 0.0f, true);
    }
  } else {
    const uint32_t v = base + threadIdx.x;
    if (v < p.vertex_count) {
      const float4* l = s.local_diag + 4 * size_t(v);
      const float4 l0 = l[0], l1 = l[1], l2 = l[2], l3 = l[3];
      bool lnv = finite4(l0) && finite4(l1) && finite4(l2) && finite4(l3);
      const bool stv = finite_nonneg_int(l3.w, float(kCholeskyFailed));
      const uint32_t st = stv ? uint32_t(l3.w) : kNonfinite;
      lnv = lnv && stv && st != kNonfinite;
      if (local_status_failure(st)) r.failures += 1;
      if (finite_s(l1.w) && l1.w >= 0.0f) r.max_shift = fmaxf(r.max_shift, l1.w);
      else lnv = false;
      const bool btv = finite_nonneg_int(l3.y, float(kMaxBacktracks));
      const bool evv = finite_nonneg_int(l3.z, float(kTrialCount));
      const uint32_t bt = btv ? uint32_t(l3.y) : 0u;
      const uint32_t ev = evv ? uint32_t(l3.z) : 0u;
      if (!btv || !evv) lnv = false;
      r.max_bt = bt;
      r.evals = ev;
      uint32_t trials = 0;
      if (st == kAccepted) trials = bt + 1;
      else if (st == kLineSearchFailed && bt == kMaxBacktracks) trials = kTrialCount;
      else if (st == kZeroGradient && bt > 0) trials = bt;
      else if (st == kNonfinite) trials = bt + (l3.x > kDeterminantFloor ? 1u : 0u);
      r.skipped = trials - (trials < ev ? trials : ev);
      if (!lnv) r.bits &= ~kLocalNumericsValid;
      const float4 ve = s.vertex_energy[v];
      cand_add_energy(r, 0, ve.x, ve.z == 1.0f && finite_s(ve.x));
      cand_add_energy(r, 1, ve.y, ve.w == 1.0f && finite_s(ve.y));
    }
  }
  block_reduce(r, sh, cand_combine);
  if (threadIdx.x == 0) reinterpret_cast<CandPartial*>(s.partials)[b] = r;
}

// One block per scene: its tetrahedron, triangle and vertex partials in
// that order, then the control records (S:3612-3955 final lane).
__global__ void __launch_bounds__(kBlock) k_candidate_final(DeviceState s, PassParams p, uint32_t tb, uint32_t rb,
                                                             uint32_t first_scene) {
  __shared__ CandPartial sh[kBlock];
  const uint32_t sc = first_scene + blockIdx.x;
  if (s.scene_cnt[sc].z == 0u) return;
  const uint4 b0 = s.scene_blk0[sc], b1 = s.scene_blk1[sc];
  const uint32_t nt = b0.w - b0.z, nr = b1.y - b1.x, nv = b0.y - b0.x;
  CandPartial r;
  cand_identity(r);
  const CandPartial* parts = reinterpret_cast<const CandPartial*>(s.partials);
  for (uint32_t i = threadIdx.x; i < nt + nr + nv; i += kBlock) {
    const uint32_t at = i < nt ? b0.z + i : (i < nt + nr ? tb + b1.x + (i - nt) : tb + rb + b0.x + (i - nt - nr));
    cand_combine(r, parts[at]);
  }
  block_reduce(r, sh, cand_combine);
  if (threadIdx.x != 0 || stop_latched(s, p, sc)) return;
  float4* control = s.control + 4 * size_t(sc);
  uint32_t bits = r.bits;
  if (s.scene_cnt[sc].y == 0u) bits &= ~(kSourceGeometryValid | kCandidateGeometryValid);
  const bool sgv = (bits & kSourceGeometryValid) != 0u, cgv = (bits & kCandidateGeometryValid) != 0u;
  const bool sev = (bits & kSourceEnergyValid) != 0u, cev = (bits & kCandidateEnergyValid) != 0u;
  const bool accepted = sgv && r.minj0 > kDeterminantFloor && cgv && r.minj1 > kDeterminantFloor;
  const float am = accepted ? r.minj1 : r.minj0;
  const bool amv = accepted ? cgv : sgv;
  const float ae = accepted ? r.e1 : r.e0;
  const bool aev = accepted ? cev : sev;
  if (amv) bits |= kAcceptedMinimumValid;
  uint32_t hist = 0;
  const float packed = control[1].w;
  if (finite_nonneg_int(packed, float(kHistoryCapacity))) hist = uint32_t(packed);
  else bits &= ~kLocalNumericsValid;
  const uint32_t cb = bits & kValidityMask;
  control[0] = make_float4(sgv ? r.minj0 : 0.0f, cgv ? r.minj1 : 0.0f, amv ? am : 0.0f, float(cb));
  control[1] = make_float4(accepted ? 1.0f : 0.0f, accepted ? 0.0f : 1.0f, float(r.failures), float(hist));
  control[2] = make_float4(sev ? r.e0 : 0.0f, cev ? r.e1 : 0.0f, aev ? ae : 0.0f, (sev && cev) ? 1.0f : 0.0f);
  control[3] = make_float4(r.max_shift, float(r.max_bt), float(r.skipped), float(r.evals));
}

__global__ void __launch_bounds__(kBlock) k_apply_candidate(DeviceState s, PassParams p) {  // S:3958-3983
  const uint32_t v = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (v >= p.vertex_count) return;
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene) return;
  const bool accepted = s.control[4 * size_t(sc) + 1].x == 1.0f;
  if (p.contact && s.bp_nv && (s.bp_ctl[sc].x & kBpScene)) {  // the pose k_promote_contact caches
    const float alpha = s.contact_ctl[sc].w;
    if (accepted && finite_s(alpha) && alpha > 0.0f) s.bp_cpose[v] = s.bp_tval[v];
  }
  if (stop_latched(s, p, sc)) return;
  if (!accepted) {
    s.pos[p.tgt][v] = s.pos[p.src][v];
    s.final_update[v] = make_float4(0, 0, 0, 1);
    return;
  }
  const f3 u = load_pos(s.pos[p.tgt], v) - load_pos(s.pos[p.src], v);
  s.final_update[v] = make_float4(scaled_length3(u), 0, 0, 1);
}

// ------------------------------------------------------------------ convergence
struct ConvPartial {
  float ch[4];
  float fin[2];
  float max_update;
  uint32_t valid;
};
__device__ __forceinline__ void conv_identity(ConvPartial& r) {
  r = ConvPartial{{0, 0, 0, 0}, {0, 0}, 0.0f, 1u};
}
__device__ __forceinline__ void conv_max(ConvPartial& l, const ConvPartial& r) {
  for (int k = 0; k < 4; ++k) l.ch[k] = fmaxf(l.ch[k], r.ch[k]);
  l.fin[0] = fmaxf(l.fin[0], r.fin[0]);
  l.fin[1] = fmaxf(l.fin[1], r.fin[1]);
  l.max_update = fmaxf(l.max_update, r.max_update);
  l.valid &= r.valid;
}
__device__ __forceinline__ void conv_sum(ConvPartial& l, const ConvPartial& r) {
  for (int k = 0; k < 4; ++k) {
    float s = l.ch[k] + r.ch[k];
    if (finite_s(s)) l.ch[k] = s;
    else {
      l.ch[k] = 0.0f;
      l.valid = 0u;
    }
  }
  for (int k = 0; k < 2; ++k) {
    float s = l.fin[k] + r.fin[k];
    if (finite_s(s)) l.fin[k] = s;
    else {
      l.fin[k] = 0.0f;
      l.valid = 0u;
    }
  }
  l.valid &= r.valid;
}
__device__ __forceinline__ float max_abs(f3 v) { return fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))); }

// convergenceGradient (S:3986-4097) fused with scan one of reduceConvergence
// (S:4119-4189): component gradients at the candidate and their scales.
// With `compute_next_rot` the same gather also evaluates the next sweep's
// vertexPolarRotation (S:2960-2989): its source pose is this candidate.
template <bool kCloth, bool kIpc>
__global__ void __launch_bounds__(kBlock) k_convergence_gradient(DeviceState s, PassParams p) {
  __shared__ ConvPartial sh[kBlock];
  ConvPartial r;
  conv_identity(r);
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t sc = s.vblock_scene[blockIdx.x];
  if (sc == kNoScene) return;  // uniform per block
  const bool latched = stop_latched(s, p, sc);
  if (v < p.vertex_count) {
    float4* c = s.conv + 5 * size_t(v);
    f3 comp5[5];
    if (!latched) {
      const uint4 inf = s.vtx_info[v];
      const bool pinned = inf.z != 0u;
      const float4* pos = s.pos[p.tgt];
      const f3 x = load_pos(pos, v);
      f3 mat = mk(0, 0, 0);
      m3 avg = mzero();
      float tv = 0.0f;
      if (!pinned || p.compute_next_rot) {
        for (uint32_t k = 0; k < inf.y; ++k) {
          const uint32_t t = s.adjacency[inf.x + k];
          const Tet e = load_tet(s, t);
          const m3 F = deformation(e, pos);
          if (p.compute_next_rot) {
            avg = mfma(e.vol, F, avg);
            tv += e.vol;
          }
          const uint32_t slot = local_slot(e.ix, v);
          if (pinned || slot >= 4) continue;
          const SnhPoint sp = snh_point(F, e.lam, e.mu);
          mat = mat + e.vol * piola_times(sp, shape_gradient(e, slot));
        }
      }
      if constexpr (kCloth) {
        if (!pinned) mat = mat + cloth_gradient(s, p, pos, v);
      }
      if (p.compute_next_rot) {
        m3 R = mid();
        if (inf.y != 0u) R = polar_rotation(tv > 0.0f ? mscale(1.0f / tv, avg) : mid(), p.polar_eps);
        store_rot(s.vrot_next, v, R);
        store_vcub(s, p, s.vcub[1], v, inf, R, x);
      }
      if (pinned) {
        for (int k = 0; k < 5; ++k) comp5[k] = mk(0, 0, 0);
      } else {
        comp5[0] = (s.vtx_rest[v].w * p.inv_dt2) * (x - xyz(s.predicted[v]));
        comp5[1] = mat;
        comp5[2] = mk(0, 0, 0);
        comp5[3] = mk(0, 0, 0);
        if (p.objective_flags != 0u) {
          if (p.objective_flags & kForce) comp5[2] = -xyz(s.objectives[2 * size_t(v)]);
          if (p.objective_flags & kTarget) {
            const float4 t = s.objectives[2 * size_t(v) + 1];
            if (t.w > 0.0f) comp5[3] = t.w * (x - xyz(t));
          }
        }
        comp5[4] = mk(0, 0, 0);
        if (p.floor_k > 0.0f && x.y < p.floor_y) comp5[4].y = p.floor_k * (x.y - p.floor_y);
        if constexpr (kIpc) {  // S:4091
          f3 ig = mk(0, 0, 0);
          m3 ih = mzero();
          ipc_local_contribution(s, p, v, pos, ig, ih);
          comp5[4] = comp5[4] + ig;
        }
      }
      for (int k = 0; k < 5; ++k) c[k] = make_float4(comp5[k].x, comp5[k].y, comp5[k].z, 0.0f);
    } else {
      for (int k = 0; k < 5; ++k) comp5[k] = xyz(c[k]);
    }
    f3 total = mk(0, 0, 0);
    bool tv_ok = true;
    for (int k = 0; k < 5; ++k) {
      const bool ok = finite3(comp5[k]);
      const f3 val = ok ? comp5[k] : mk(0, 0, 0);
      const float vs = max_abs(val);
      if (k < 4) r.ch[k] = fmaxf(r.ch[k], vs);
      else r.fin[0] = fmaxf(r.fin[0], vs);
      total = total + val;
      tv_ok = tv_ok && ok;
    }
    tv_ok = tv_ok && finite3(total);
    if (tv_ok) r.fin[1] = fmaxf(r.fin[1], max_abs(total));
    else r.valid = 0u;
    const float up = s.final_update[v].x;
    if (finite_s(up) && up >= 0.0f) r.max_update = fmaxf(r.max_update, up);
    else r.valid = 0u;
  }
  block_reduce(r, sh, conv_max);
  if (threadIdx.x == 0) reinterpret_cast<ConvPartial*>(s.partials)[blockIdx.x] = r;
}

// The contact component of convergenceGradient (S:4091) at the candidate
// pose, 16 lanes per vertex: a vertex's active rows are strided over its 16
// lanes and the lane sums combined by an xor butterfly (a fixed order). A
// contact-heavy vertex has hundreds of rows; split over the convergence
// gather's 4 or 8 lanes they made that pass latency-bound. Pinned
// vertices and latched scenes write nothing (the gather reads no contact
// term for them).
__global__ void __launch_bounds__(kBlock) k_ipc_gradient_lanes(DeviceState s, PassParams p) {
  constexpr int kLanes = 16;
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t v = gid / kLanes, sub = gid % kLanes;
  if (v >= p.vertex_count) return;  // uniform per 16-lane group
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || stop_latched(s, p, sc) || s.vtx_info[v].z != 0u) return;
  const float4* pos = s.pos[p.tgt];
  f3 g = mk(0, 0, 0);
  m3 h = mzero();  // the rows' Hessian terms are not needed here
  const uint32_t begin = s.cand_row[v].x, n = s.active_count[v];
#pragma unroll 1
  for (uint32_t k = sub; k < n; k += kLanes) ipc_row_contribution(s, p, v, pos, s.active_ids[begin + k], g, h);
  g = lanes_sum3<kLanes>(g);
  if (sub == 0) s.ipc_grad[v] = make_float4(g.x, g.y, g.z, 0.0f);
}

// The contact rows' gradient and Hessian terms of the local solve's gather
// (S:3193-3195, ipcLocalContribution S:786-874) at a Jacobi sweep's source
// pose, for vertices with more than kIpcLocalRows active rows, summed by 16
// lanes per vertex like k_ipc_gradient_lanes; the lane sweep of the solve
// then reads one record per such vertex: {g, H c0.x}, {H c0.yz, H c1.xy},
// {H c1.z, H c2} (columns of H). Vertices with fewer rows keep them in the
// solve's own lanes, where they overlap the element terms.
__global__ void __launch_bounds__(kBlock) k_ipc_local_lanes(DeviceState s, PassParams p) {
  constexpr int kLanes = 16;
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t v = gid / kLanes, sub = gid % kLanes;
  if (v >= p.vertex_count) return;  // uniform per 16-lane group
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene || stop_latched(s, p, sc) || s.vtx_info[v].z != 0u) return;
  const uint32_t begin = s.cand_row[v].x, n = s.active_count[v];
  if (n <= kIpcLocalRows) return;  // uniform per 16-lane group
  const float4* pos = s.pos[p.src];
  f3 g = mk(0, 0, 0);
  m3 h = mzero();
  const bool records = p.fric_records != 0u;
#pragma unroll 1
  for (uint32_t k = sub; k < n; k += kLanes) ipc_row_contribution(s, p, v, pos, s.active_ids[begin + k], g, h, records);
  g = lanes_sum3<kLanes>(g);
  h = {lanes_sum3<kLanes>(h.c0), lanes_sum3<kLanes>(h.c1), lanes_sum3<kLanes>(h.c2)};
  if (sub == 0) {
    s.ipc_local[3 * size_t(v)] = make_float4(g.x, g.y, g.z, h.c0.x);
    s.ipc_local[3 * size_t(v) + 1] = make_float4(h.c0.y, h.c0.z, h.c1.x, h.c1.y);
    s.ipc_local[3 * size_t(v) + 2] = make_float4(h.c1.z, h.c2.x, h.c2.y, h.c2.z);
  }
}

// convergenceGradient with LPV lanes per vertex. A block still covers one
// 256-vertex arena block (its partial record), in LPV rounds of 256 / LPV
// vertices; the lanes of a vertex split its incident tetrahedra and combine
// the material gradient and the rotation average with xor butterflies.
// kRecords: cloth element records at the candidate are present (Jacobi
// sweeps), so the cloth terms are read from them only.
template <bool kCloth, bool kIpc, int LPV, bool kRecords = false>
__global__ void __launch_bounds__(kBlock) k_convergence_gradient_lanes(DeviceState s, PassParams p) {
  __shared__ ConvPartial sh[kBlock];
  ConvPartial r;
  conv_identity(r);
  const uint32_t sc = s.vblock_scene[blockIdx.x];
  if (sc == kNoScene) return;  // uniform per block
  const bool latched = stop_latched(s, p, sc);
  const uint32_t sub = threadIdx.x % LPV;
#pragma unroll 1
  for (uint32_t round = 0; round < LPV; ++round) {
    const uint32_t v = blockIdx.x * kBlock + round * (kBlock / LPV) + threadIdx.x / LPV;
    if (v >= p.vertex_count) continue;  // uniform per vertex
    float4* c = s.conv + 5 * size_t(v);
    f3 comp5[5];
    if (!latched) {
      const uint4 inf = s.vtx_info[v];
      const bool pinned = inf.z != 0u;
      const float4* pos = s.pos[p.tgt];
      const f3 x = load_pos(pos, v);
      f3 mat = mk(0, 0, 0);
      m3 avg = mzero();
      float tv = 0.0f;
      if (!pinned || p.compute_next_rot) {
#pragma unroll 1
        for (uint32_t k = sub; k < inf.y; k += LPV) {
          const Tet e = load_tet(s, s.adjacency[inf.x + k]);
          const m3 F = deformation_at(e, pos, v, x);
          if (p.compute_next_rot) {
            avg = mfma(e.vol, F, avg);
            tv += e.vol;
          }
          const uint32_t slot = local_slot(e.ix, v);
          if (pinned || slot >= 4) continue;
          const SnhPoint sp = snh_point(F, e.lam, e.mu);
          mat = mat + e.vol * piola_times(sp, shape_gradient(e, slot));
        }
        mat = lanes_sum3<LPV>(mat);
        if (p.compute_next_rot) {
          avg = {lanes_sum3<LPV>(avg.c0), lanes_sum3<LPV>(avg.c1), lanes_sum3<LPV>(avg.c2)};
          tv = lanes_sum<LPV>(tv);
        }
      }
      if constexpr (kCloth) {  // membrane and bending terms, split over the lanes too
        f3 cg = mk(0, 0, 0);
        if (!pinned) {
          const uint2 trr = s.tri_row[v], hr = s.hinge_row[v];
#pragma unroll 1
          for (uint32_t k = lane_first<LPV>(sub, 0); k < trr.y; k += LPV) {
            const uint32_t tri = s.tri_elems[trr.x + k];
            const uint32_t slot = tri_slot(s.tri_ix[tri], v);
            if (slot < 3) cg = cg + membrane_gradient(s, tri, membrane_base<kRecords>(s, p, tri, pos), slot);
          }
#pragma unroll 1
          for (uint32_t k = lane_first<LPV>(sub, trr.y); k < hr.y; k += LPV) {
            const uint32_t h = s.hinge_elems[hr.x + k];
            const uint32_t slot = hinge_slot(s.hinge_ix[h], v);
            if (slot >= 4) continue;
            const Dihedral d = dihedral_base<kRecords>(s, p, h, pos, slot);
            if (!d.valid) {
              cg = cg + mk(kF32Max, kF32Max, kF32Max);
            } else {
              const float4 r = s.hinge_rest[h];
              cg = cg + (s.cloth_bend * r.y * wrapped_angle_difference(d.angle, r.x)) * d.grad;
            }
          }
          cg = lanes_sum3<LPV>(cg);
          mat = mat + cg;
        }
      }
      f3 ig = mk(0, 0, 0);
      if constexpr (kIpc) {  // S:4091, summed over the active rows by k_ipc_gradient_lanes
        if (!pinned && sub == 0) ig = xyz(s.ipc_grad[v]);
      }
      if (sub == 0) {
        if (p.compute_next_rot) {
          m3 R = mid();
          if (inf.y != 0u) R = polar_rotation(tv > 0.0f ? mscale(1.0f / tv, avg) : mid(), p.polar_eps);
          store_rot(s.vrot_next, v, R);
          store_vcub(s, p, s.vcub[1], v, inf, R, x);
        }
        if (pinned) {
          for (int k = 0; k < 5; ++k) comp5[k] = mk(0, 0, 0);
        } else {
          comp5[0] = (s.vtx_rest[v].w * p.inv_dt2) * (x - xyz(s.predicted[v]));
          comp5[1] = mat;
          comp5[2] = mk(0, 0, 0);
          comp5[3] = mk(0, 0, 0);
          if (p.objective_flags != 0u) {
            if (p.objective_flags & kForce) comp5[2] = -xyz(s.objectives[2 * size_t(v)]);
            if (p.objective_flags & kTarget) {
              const float4 t = s.objectives[2 * size_t(v) + 1];
              if (t.w > 0.0f) comp5[3] = t.w * (x - xyz(t));
            }
          }
          comp5[4] = mk(0, 0, 0);
          if (p.floor_k > 0.0f && x.y < p.floor_y) comp5[4].y = p.floor_k * (x.y - p.floor_y);
          if constexpr (kIpc) comp5[4] = comp5[4] + ig;
        }
        for (int k = 0; k < 5; ++k) c[k] = make_float4(comp5[k].x, comp5[k].y, comp5[k].z, 0.0f);
      }
    } else if (sub == 0) {
      for (int k = 0; k < 5; ++k) comp5[k] = xyz(c[k]);
    }
    if (sub != 0) continue;
    f3 total = mk(0, 0, 0);
    bool tv_ok = true;
    for (int k = 0; k < 5; ++k) {
      const bool ok = finite3(comp5[k]);
      const f3 val = ok ? comp5[k] : mk(0, 0, 0);
      const float vs = max_abs(val);
      if (k < 4) r.ch[k] = fmaxf(r.ch[k], vs);
      else r.fin[0] = fmaxf(r.fin[0], vs);
      total = total + val;
      tv_ok = tv_ok && ok;
    }
    tv_ok = tv_ok && finite3(total);
    if (tv_ok) r.fin[1] = fmaxf(r.fin[1], max_abs(total));
    else r.valid = 0u;
    const float up = s.final_update[v].x;
    if (finite_s(up) && up >= 0.0f) r.max_update = fmaxf(r.max_update, up);
    else r.valid = 0u;
  }
  block_reduce(r, sh, conv_max);
  if (threadIdx.x == 0) reinterpret_cast<ConvPartial*>(s.partials)[blockIdx.x] = r;
}

// Per-scene scales: combine the scene's scan-one partials (one block per
// scene) into its scale record.
__global__ void __launch_bounds__(kBlock) k_convergence_scales(DeviceState s) {
  __shared__ ConvPartial sh[kBlock];
  const uint32_t sc = blockIdx.x;
  if (s.scene_cnt[sc].z == 0u) return;
  const uint4 b0 = s.scene_blk0[sc];
  ConvPartial r;
  conv_identity(r);
  const ConvPartial* parts = reinterpret_cast<const ConvPartial*>(s.partials);
  for (uint32_t i = b0.x + threadIdx.x; i < b0.y; i += kBlock) conv_max(r, parts[i]);
  block_reduce(r, sh, conv_max);
  if (threadIdx.x == 0) reinterpret_cast<ConvPartial*>(s.scene_conv)[sc] = r;
}

// Scan two (S:4211-4283): normalized squared sums with the global scales.
__global__ void __launch_bounds__(kBlock) k_convergence_sums(DeviceState s, PassParams p, uint32_t blocks) {
  __shared__ ConvPartial sh[kBlock];
  ConvPartial* parts = reinterpret_cast<ConvPartial*>(s.partials);
  const uint32_t scene = s.vblock_scene[blockIdx.x];
  if (scene == kNoScene) return;  // uniform per block
  const ConvPartial sc = reinterpret_cast<const ConvPartial*>(s.scene_conv)[scene];
  ConvPartial r;
  conv_identity(r);
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v < p.vertex_count) {
    const float4* c = s.conv + 5 * size_t(v);
    f3 total = mk(0, 0, 0);
    bool tv = true;
    for (int k = 0; k < 5; ++k) {
      const f3 raw = xyz(c[k]);
      const bool ok = finite3(raw);
      const f3 val = ok ? raw : mk(0, 0, 0);
      const float scale = k < 4 ? sc.ch[k] : sc.fin[0];
      float sq = 0.0f;
      if (ok && scale > 0.0f) {
        const f3 n = val / scale;
        sq = dot(n, n);
      }
      const bool sqv = finite_s(sq) && sq >= 0.0f;
      float& slot = k < 4 ? r.ch[k] : r.fin[0];
      const float sum = slot + sq;
      if (sqv && finite_s(sum)) slot = sum;
      else {
        slot = 0.0f;
        r.valid = 0u;
      }
      total = total + val;
      tv = tv && ok;
    }
    tv = tv && finite3(total);
    float tsq = 0.0f;
    if (tv && sc.fin[1] > 0.0f) {
      const f3 n = total / sc.fin[1];
      tsq = dot(n, n);
    }
    const float tsum = r.fin[1] + tsq;
    if (tv && finite_s(tsq) && tsq >= 0.0f && finite_s(tsum)) r.fin[1] = tsum;
    else {
      r.fin[1] = 0.0f;
      r.valid = 0u;
    }
  }
  block_reduce(r, sh, conv_sum);
  if (threadIdx.x == 0) parts[blocks + blockIdx.x] = r;
}

// Final pass of reduceConvergence (S:4320-4485): norms, history, stop latch.
// One block per scene: its scan-two partials, norms, history and latch.
__global__ void __launch_bounds__(kBlock) k_convergence_final(DeviceState s, PassParams p, uint32_t blocks) {
  __shared__ ConvPartial sh[kBlock];
  const uint32_t scene = blockIdx.x;
  if (s.scene_cnt[scene].z == 0u) return;
  const uint4 b0 = s.scene_blk0[scene];
  ConvPartial* parts = reinterpret_cast<ConvPartial*>(s.partials);
  const ConvPartial sc = reinterpret_cast<const ConvPartial*>(s.scene_conv)[scene];
  ConvPartial r;
  conv_identity(r);
  for (uint32_t i = b0.x + threadIdx.x; i < b0.y; i += kBlock) conv_sum(r, parts[blocks + i]);
  block_reduce(r, sh, conv_sum);
  if (threadIdx.x != 0 || stop_latched(s, p, scene)) return;
  float4* control = s.control + 4 * size_t(scene);
  const float scene_scale = s.scene_scale[scene];
  const float scene_vertices = float(s.scene_cnt[scene].x);
  float norms[5];
  for (int k = 0; k < 4; ++k) norms[k] = sc.ch[k] * sqrtf(r.ch[k]);
  norms[4] = sc.fin[0] * sqrtf(r.fin[0]);
  const float gnorm = sc.fin[1] * sqrtf(r.fin[1]);
  const bool raw_valid = sc.valid == 1u;
  bool cnv = raw_valid && r.valid == 1u;
  for (int k = 0; k < 5; ++k) cnv = cnv && finite_s(norms[k]) && norms[k] >= 0.0f;
  const float den = fmaxf(1.0f, norms[0] + norms[1] + norms[2] + norms[3] + norms[4]);
  const float rel = gnorm / den;
  const bool rnv = cnv && finite_s(gnorm) && gnorm >= 0.0f && finite_s(den) && den >= 1.0f && finite_s(rel) && rel >= 0.0f;
  const bool ssv = finite_s(scene_scale) && scene_scale >= 0.0f;
  const float nmu = sc.max_update / fmaxf(ssv ? scene_scale : 1.0e-12f, 1.0e-12f);
  const bool unv = raw_valid && ssv && finite_s(sc.max_update) && sc.max_update >= 0.0f && finite_s(nmu) && nmu >= 0.0f;
  const float4 geo = control[0];
  const uint32_t cb = control_bits(geo.w) & kValidityMask;
  const bool sgv = (cb & kSourceGeometryValid) && finite_s(geo.x);
  const bool cgv = (cb & kCandidateGeometryValid) && finite_s(geo.y);
  const bool amv = (cb & kAcceptedMinimumValid) && finite_s(geo.z);
  const bool lnv = (cb & kLocalNumericsValid) != 0u;
  const float4 acc = control[1];
  const bool aav = finite_nonneg_int(acc.x, 1.0f), arv = finite_nonneg_int(acc.y, 1.0f);
  const bool lfv = finite_nonneg_int(acc.z, scene_vertices);
  const bool hiv = finite_nonneg_int(acc.w, float(kHistoryCapacity));
  const bool aacc = aav && acc.x == 1.0f, arev = arv && acc.y == 1.0f;
  const uint32_t lf = lfv ? uint32_t(acc.z) : 0u, hi = hiv ? uint32_t(acc.w) : 0u;
  const float4 cm = control[3];
  const float trials_max = scene_vertices * float(kTrialCount);
  const bool cmv = finite_s(cm.x) && cm.x >= 0.0f && finite_nonneg_int(cm.y, float(kMaxBacktracks)) &&
                   finite_nonneg_int(cm.z, trials_max) && finite_nonneg_int(cm.w, trials_max);
  const bool fin = rnv && unv && sgv && cgv && amv && lnv && aav && arv && lfv && hiv && cmv;
  const bool rs = fin && rel <= p.residual_tol;
  const bool us = fin && nmu <= p.update_tol;
  const bool converged = rs && us && aacc && !arev && lf == 0u;
  if (hiv && hi < kHistoryCapacity) {
    float4* h = s.history + (size_t(scene) * kHistoryCapacity + hi) * 8;
    h[0] = control[0];
    h[1] = make_float4(acc.x, acc.y, acc.z, arev ? 1.0f : 0.0f);
    h[2] = control[2];
    h[3] = make_float4(cnv ? norms[0] : 0.0f, cnv ? norms[1] : 0.0f, cnv ? norms[2] : 0.0f, cnv ? norms[3] : 0.0f);
    h[4] = make_float4(cnv ? norms[4] : 0.0f, rnv ? gnorm : 0.0f, rnv ? den : 0.0f, rnv ? rel : 0.0f);
    h[5] = make_float4(unv ? sc.max_update : 0.0f, unv ? nmu : 0.0f, rs ? 1.0f : 0.0f, us ? 1.0f : 0.0f);
    h[6] = make_float4(converged ? 1.0f : 0.0f, fin ? 1.0f : 0.0f, float(hi), cmv ? cm.x : 0.0f);
    h[7] = cmv ? make_float4(cm.y, cm.z, cm.w, 1.0f) : make_float4(0, 0, 0, 1);
  }
  control[1].w = float(hi + 1u);
  if (converged && p.stop_enabled) control[0].w = float(cb | kStopLatch);
}

__global__ void __launch_bounds__(kBlock) k_copy_position(DeviceState s, PassParams p) {  // S:4489-4496
  const uint32_t v = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (v >= p.vertex_count || vscene(s, v) == kNoScene) return;
  const float4 x = s.pos[p.src][v];
  s.pos[p.tgt][v] = make_float4(x.x, x.y, x.z, 1.0f);
}

// bodyHorizontalCorrection (S:4499-4532): one block per body reduces its
// vertices in a fixed order.
// Objectives and mesh contact keep their solved centre-of-mass motion
// (jgs2-solver.ts:1994-2008): the correction is off for such a scene.
// body_rng = {scene vertex begin, end, objective flags, scene}.
__device__ __forceinline__ bool body_correction_off(const DeviceState& s, const PassParams& p, uint4 range) {
  if (!p.globalization) return false;
  if (range.z != 0u) return true;
  const uint4 b1 = s.scene_blk1[range.w];
  return p.contact != 0u && b1.z != b1.w;
}
__global__ void __launch_bounds__(kBlock) k_body_correction(DeviceState s, PassParams p) {
  __shared__ float sm[5][kBlock];
  __shared__ uint32_t sp[kBlock];
  const uint32_t body = blockIdx.x;
  const uint4 range = s.body_rng[body];  // the owning scene's vertex range
  if (range.w == kNoScene || body_correction_off(s, p, range)) return;  // uniform per block
  float tm = 0, sx = 0, sz = 0, px = 0, pz = 0;
  uint32_t pinned = 0;
  for (uint32_t v = range.x + threadIdx.x; v < range.y; v += kBlock) {
    const uint4 inf = s.vtx_info[v];
    if (inf.w != body) continue;
    pinned |= inf.z != 0u ? 1u : 0u;
    const float m = fmaxf(s.vtx_rest[v].w, 0.0f);
    const float4 x = s.pos[0][v], q = s.predicted[v];
    tm += m;
    sx += m * x.x;
    sz += m * x.z;
    px += m * q.x;
    pz += m * q.z;
  }
  sm[0][threadIdx.x] = tm;
  sm[1][threadIdx.x] = sx;
  sm[2][threadIdx.x] = sz;
  sm[3][threadIdx.x] = px;
  sm[4][threadIdx.x] = pz;
  sp[threadIdx.x] = pinned;
  __syncthreads();
  for (uint32_t stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      for (int k = 0; k < 5; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + stride];
      sp[threadIdx.x] |= sp[threadIdx.x + stride];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float cx = 0.0f, cz = 0.0f;
    if (sp[0] == 0u && sm[0][0] > 0.0f) {
      cx = (sm[3][0] - sm[1][0]) / sm[0][0];
      cz = (sm[4][0] - sm[2][0]) / sm[0][0];
    }
    s.body_corr[body] = make_float4(cx, 0.0f, cz, 0.0f);
  }
}
__global__ void __launch_bounds__(kBlock) k_apply_body_correction(DeviceState s, PassParams p) {  // S:4535-4551
  const uint32_t v = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (v >= p.vertex_count || vscene(s, v) == kNoScene) return;
  const uint4 inf = s.vtx_info[v];
  if (inf.z != 0u || inf.w >= p.body_count) return;
  if (body_correction_off(s, p, s.body_rng[inf.w])) return;
  const float4 c = s.body_corr[inf.w];
  float4 x = s.pos[0][v];
  x.x += c.x;
  x.z += c.z;
  s.pos[0][v] = x;
}

__global__ void __launch_bounds__(kBlock) k_finalize(DeviceState s, PassParams p) {  // S:4554-4589
  const uint32_t v = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (v >= p.vertex_count || vscene(s, v) == kNoScene) return;
  if (s.vtx_info[v].z != 0u) {
    if (!p.globalization) {
      const float4 r = s.vtx_rest[v];
      s.pos[0][v] = make_float4(r.x, r.y, r.z, 1.0f);
    }
    s.velocity[v] = make_float4(0, 0, 0, 0);
    return;
  }
  const f3 x = load_pos(s.pos[0], v);
  const f3 o = load_pos(s.old, v);
  f3 vel = (x - o) * p.inv_dt;
  if (p.floor_k > 0.0f && p.tangential_damping > 0.0f && x.y <= p.floor_y + p.contact_margin) {
    const float ts = expf(-p.tangential_damping * p.dt);
    vel.x *= ts;
    vel.z *= ts;
  }
  s.velocity[v] = make_float4(vel.x, vel.y, vel.z, 0.0f);
}

template <bool kCorot, bool kCloth, bool kIpc>
__global__ void __launch_bounds__(kBlock) k_debug_local_systems(DeviceState s, PassParams p) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= p.vertex_count || vscene(s, v) == kNoScene) return;
  float* g = s.debug_g + 3 * size_t(v);
  float* h = s.debug_h + 9 * size_t(v);
  if (s.vtx_info[v].z != 0u) {
    for (int k = 0; k < 3; ++k) g[k] = 0.0f;
    for (int k = 0; k < 9; ++k) h[k] = 0.0f;
    return;
  }
  const Local L = gather_local<kCorot, kCloth, kIpc>(s, p, v);
  g[0] = L.g.x;
  g[1] = L.g.y;
  g[2] = L.g.z;
  const f3 cols[3] = {L.H.c0, L.H.c1, L.H.c2};
  for (int c = 0; c < 3; ++c) {
    h[0 * 3 + c] = cols[c].x;
    h[1 * 3 + c] = cols[c].y;
    h[2 * 3 + c] = cols[c].z;
  }
}


// ------------------------------------------------------------------ IPC passes (S:2695-2929)
__device__ __forceinline__ float ipc_broad_threshold(const PassParams& p, float scene_scale) {  // S:515-517
  return p.ipc_dhat + ipc_rounding_reach(p, scene_scale);
}
// lagContact (S:2713-2744): source contact cache of the frame-start pose
// (posA) and the lagged normal force and weights for friction.
__device__ __forceinline__ uint32_t cscene(const DeviceState& s, uint32_t c) { return s.cblock_scene[c / kSceneBlock]; }
__global__ void __launch_bounds__(kBlock) k_lag_contact(DeviceState s, PassParams p) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= s.NC) return;
  const uint32_t sc = cscene(s, c);
  if (sc == kNoScene) return;
  if (c == s.scene_blk1[sc].z * kSceneBlock) s.contact_ctl[sc].w = 1.0f;
  if (c >= s.scene_end[sc].w) return;
  const Contact k = ipc_contact_at(s, c, s.pos[0], ipc_broad_threshold(p, s.scene_scale[sc]));
  ipc_store_cached(s, c, false, k);
  float4 nf = make_float4(0, 0, 0, 0), w = make_float4(0, 0, 0, 0);
  if (k.valid && k.distance > s.ipc_dmin) {
    const float force = fmaxf(0.0f, -p.ipc_kappa * ipc_barrier_d1(k.distance, p.ipc_dhat));
    if (finite_s(force)) {
      nf = make_float4(k.normal.x, k.normal.y, k.normal.z, force);
      w = k.weights;
    }
  }
  s.lag_n[c] = nf;
  s.lag_w[c] = w;
  ipc_store_summary(s, c);
}
// buildActiveContactRows (S:2747-2785): keep candidates the bounded step can
// bring into the barrier, with friction, or touched by unbounded pin motion.
// The predicate is evaluated once per candidate (not once per each of its
// four vertices), then each vertex's row is compacted by 16 lanes with a
// wave ballot; the active ids keep the row's ascending order.
__global__ void __launch_bounds__(kBlock) k_candidate_keep(DeviceState s, PassParams p) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= s.NC) return;
  const uint32_t sc = cscene(s, c);
  if (sc == kNoScene || c >= s.scene_end[sc].w) return;
  const bool bounded = p.max_step > 0.0f;
  const float reach = p.ipc_dhat + 2.0f * (fmaxf(p.max_step, 0.0f) + ipc_rounding_reach(p, s.scene_scale[sc]));
  const float bound = fmaxf(p.max_step, 0.0f);
  const float2 q = s.src_d[c];  // the cached source contact's validity and distance, and the lagged force
  bool include = q.x < 0.0f || !bounded || q.x <= reach || (s.ipc_mu > 0.0f && q.y > 0.0f);
  if (!include && (s.cand_meta[c] & kMetaPinned)) {  // ipcCandidatePinnedMotionExceedsBound (S:484-507)
    const float4* src = s.pos[p.src];
    const uint4 ix = s.cand[c];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const uint32_t vl = tet_vertex(ix, l);
      if (include || s.vtx_info[vl].z == 0u) continue;
      const f3 d = xyz(s.vtx_rest[vl]) - load_pos(src, vl);
      const float dl = sqrtf(dot(d, d));
      include = !finite3(d) || !finite_s(dl) || dl > bound;
    }
  }
  s.cand_keep[c] = include ? uint8_t(1) : uint8_t(0);
  if (p.fric_records && s.ipc_mu > 0.0f && q.y > 0.0f) {  // a friction row: its record for this sweep
    const float4* src = s.pos[p.src];
    const uint4 ix = s.cand[c];
    float4* fr = s.fric + 4 * size_t(c);
    const f3 d0 = load_pos(src, ix.x) - load_pos(s.old, ix.x), d1 = load_pos(src, ix.y) - load_pos(s.old, ix.y);
    const f3 d2 = load_pos(src, ix.z) - load_pos(s.old, ix.z), d3 = load_pos(src, ix.w) - load_pos(s.old, ix.w);
    const float e = ipc_energy_from_contact(s, p, c, ix, ipc_cached(s, c, false), src, kNoVertex, mk(0, 0, 0));
    fr[0] = make_float4(d0.x, d0.y, d0.z, e);
    fr[1] = make_float4(d1.x, d1.y, d1.z, 0.0f);
    fr[2] = make_float4(d2.x, d2.y, d2.z, 0.0f);
    fr[3] = make_float4(d3.x, d3.y, d3.z, 0.0f);
  }
}
__global__ void __launch_bounds__(kBlock) k_compact_active_rows(DeviceState s, PassParams p) {
  constexpr uint32_t kLanes = 16;
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t v = gid / kLanes, sub = gid % kLanes;
  if (v >= p.vertex_count) return;  // uniform per 16-lane group
  if (vscene(s, v) == kNoScene) return;
  const uint2 row = s.cand_row[v];
  const uint32_t first = row.x, last = row.x + row.y;
  const uint32_t group = (threadIdx.x % 64u) & ~(kLanes - 1u);
  uint32_t n = 0;
  for (uint32_t base = first; base < last; base += kLanes) {  // same trip count on the group's lanes
    const uint32_t i = base + sub;
    uint32_t c = 0;
    bool keep = false;
    if (i < last) {
      c = s.cand_ids[i];
      keep = s.cand_keep[c] != 0u;
    }
    const uint64_t wave = __ballot(keep);
    const uint32_t mine = uint32_t(wave >> group) & 0xffffu;
    if (keep) s.active_ids[first + n + __popc(mine & ((1u << sub) - 1u))] = c;
    n += __popc(mine);
  }
  if (sub == 0) s.active_count[v] = n;
}
// reduceContactStep (S:2878-2918), block half, fused into the two passes
// that produce the per-candidate records {alpha, source distance, trial
// distance, valid}: the minimum alpha and an all-valid flag of the block's
// candidates (min and a 0/1 product are order independent, so the partials
// are exact); one block per scene combines them (k_reduce_contact_final).
__device__ __forceinline__ void contact_min(float4& a, const float4& b) {
  a = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), fminf(a.w, b.w));
}
__device__ __forceinline__ void contact_block_partial(const DeviceState& s, bool live, float4 q) {
  __shared__ float4 sh[kBlock];
  float4 r = make_float4(1.0f, kF32Max, kF32Max, 1.0f);
  if (live) {
    const bool valid = q.w == 1.0f && finite_s(q.x) && finite_s(q.y) && finite_s(q.z);
    r = make_float4(fminf(1.0f, valid ? q.x : 0.0f), valid ? q.y : 0.0f, valid ? q.z : 0.0f, valid ? 1.0f : 0.0f);
  }
  block_reduce(r, sh, contact_min);
  if (threadIdx.x == 0) reinterpret_cast<float4*>(s.partials)[blockIdx.x] = r;
}
// The largest step length |x_T - x_S| (as the safe-step pass forms it) of
// each scene with candidates, in contact_ctl[scene].x: per-block maxima
// over the scene's vertex blocks, then one block per scene (a non-finite
// step gives +inf). A candidate's Lipschitz bound l_a + max(l_b..) (VT) or
// max + max (EE) is at most twice it, rounding included (f32 addition is
// monotone and 2M is exact), so a candidate whose cached source distance
// exceeds the minimum distance by twice the bound takes the full step and
// its record is known without loading its eight positions.
__global__ void __launch_bounds__(kBlock) k_step_bound_partials(DeviceState s, PassParams p) {
  __shared__ float sh[kBlock];
  const uint32_t sc = s.vblock_scene[blockIdx.x];
  if (sc == kNoScene) return;                 // uniform per block
  const uint4 b1 = s.scene_blk1[sc];
  if (b1.z == b1.w) return;                   // no candidates: uniform per block
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  float l = 0.0f;
  if (v < p.vertex_count) {
    const f3 d = load_pos(s.pos[p.tgt], v) - load_pos(s.pos[p.src], v);
    l = sqrtf(dot(d, d));
    if (!(l <= kF32Max)) l = INFINITY;       // NaN or inf: no candidate skips
  }
  block_reduce(l, sh, [](float& a, const float& b) { a = fmaxf(a, b); });
  if (threadIdx.x == 0) reinterpret_cast<float4*>(s.partials)[blockIdx.x] = make_float4(l, 0, 0, 0);
}
__global__ void __launch_bounds__(kBlock) k_step_bound_final(DeviceState s) {
  __shared__ float sh[kBlock];
  const uint32_t sc = blockIdx.x;
  const uint4 b0 = s.scene_blk0[sc], b1 = s.scene_blk1[sc];
  if (s.scene_cnt[sc].z == 0u || b1.z == b1.w) return;
  float m = 0.0f;
  const float4* parts = reinterpret_cast<const float4*>(s.partials);
  for (uint32_t i = b0.x + threadIdx.x; i < b0.y; i += kBlock) m = fmaxf(m, parts[i].x);
  block_reduce(m, sh, [](float& a, const float& b) { a = fmaxf(a, b); });
  if (threadIdx.x == 0) s.contact_ctl[sc].x = m;
}
// candidateContactStep (S:2788-2852): per-candidate conservative step cap.
__device__ __forceinline__ float4 candidate_contact_step(const DeviceState& s, const PassParams& p, uint32_t c,
                                                         uint32_t sc) {
  if (stop_latched(s, p, sc)) return make_float4(1, 0, 0, 1);
  {  // far from contact: the full step, the record it would compute
    const float2 q = s.src_d[c];
    const float m = s.contact_ctl[sc].x;
    if (q.x >= 0.0f && q.x - (m + m) > s.ipc_dmin) return make_float4(1.0f, q.x, q.x, 1.0f);
  }
  const uint4 ix = s.cand[c];
  const float4* src = s.pos[p.src];
  const float4* tgt = s.pos[p.tgt];
  const f3 d0 = load_pos(tgt, ix.x) - load_pos(src, ix.x), d1 = load_pos(tgt, ix.y) - load_pos(src, ix.y);
  const f3 d2 = load_pos(tgt, ix.z) - load_pos(src, ix.z), d3 = load_pos(tgt, ix.w) - load_pos(src, ix.w);
  const float l0 = sqrtf(dot(d0, d0)), l1 = sqrtf(dot(d1, d1)), l2 = sqrtf(dot(d2, d2)), l3 = sqrtf(dot(d3, d3));
  const float lip = (s.cand_meta[c] & 1u) == 0u ? l0 + fmaxf(l1, fmaxf(l2, l3)) : fmaxf(l0, l1) + fmaxf(l2, l3);
  Contact src_k = ipc_cached(s, c, false);
  bool full = src_k.valid && finite_s(lip) && src_k.distance - lip > s.ipc_dmin;
  if (!full) {
    src_k = ipc_contact_at(s, c, src, kF32Max);
    full = src_k.valid && finite_s(lip) && src_k.distance - lip > s.ipc_dmin;
  }
  const bool trial_valid = finite3(d0) && finite3(d1) && finite3(d2) && finite3(d3);
  const bool valid = src_k.valid && trial_valid && src_k.distance > s.ipc_dmin && finite_s(lip);
  float alpha = 0.0f;
  if (valid) alpha = full ? 1.0f : ipc_safe_step_cap(src_k.distance, s.ipc_dmin, lip, s.ipc_safety);
  return make_float4(alpha, src_k.valid ? src_k.distance : 0.0f, trial_valid ? src_k.distance : 0.0f,
                     valid ? 1.0f : 0.0f);
}
// Blocks of no scene write no partial (the final pass reads only a scene's
// own blocks); a block's scene is uniform, so they leave before any barrier.
__global__ void __launch_bounds__(kBlock) k_candidate_contact_step(DeviceState s, PassParams p) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t sc = s.cblock_scene[blockIdx.x];
  if (sc == kNoScene) return;  // uniform per block
  const bool live = c < s.scene_end[sc].w;
  contact_block_partial(s, live, live ? candidate_contact_step(s, p, c, sc) : make_float4(0, 0, 0, 0));
}
// validateContactStep (S:2855-2875): exact check of the scaled target pose.
__device__ __forceinline__ float4 validate_contact_step(const DeviceState& s, const PassParams& p, uint32_t c,
                                                        uint32_t sc) {
  const Contact k = ipc_contact_at(s, c, s.pos[p.tgt], ipc_broad_threshold(p, s.scene_scale[sc]));
  ipc_store_cached(s, c, true, k);
  // The target cache's distance as ipc_cached reads it (-1 when invalid):
  // ipc_store_cached keeps the contact exactly when this holds.
  const bool cached = k.valid && finite_s(k.distance) && k.distance >= 0.0f && finite3(k.normal) && finite4(k.weights);
  s.tgt_d[c] = make_float2(cached ? k.distance : -1.0f, s.src_d[c].y);
  const bool valid = k.valid && finite_s(k.distance) && k.distance > s.ipc_dmin;
  const float d = valid ? k.distance : 0.0f;
  return make_float4(1.0f, d, d, valid ? 1.0f : 0.0f);
}
__global__ void __launch_bounds__(kBlock) k_validate_contact_step(DeviceState s, PassParams p) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t sc = s.cblock_scene[blockIdx.x];
  if (sc == kNoScene) return;  // uniform per block
  const bool live = c < s.scene_end[sc].w;
  contact_block_partial(s, live, live ? validate_contact_step(s, p, c, sc) : make_float4(0, 0, 0, 0));
}
// One block per scene: the scene's safe alpha (zero if any record failed).
__global__ void __launch_bounds__(kBlock) k_reduce_contact_final(DeviceState s) {
  __shared__ float4 sh[kBlock];
  const uint32_t sc = blockIdx.x;
  const uint4 b1 = s.scene_blk1[sc];
  if (s.scene_cnt[sc].z == 0u || b1.z == b1.w) return;
  float4 r = make_float4(1.0f, kF32Max, kF32Max, 1.0f);
  const float4* parts = reinterpret_cast<const float4*>(s.partials);
  for (uint32_t i = b1.z + threadIdx.x; i < b1.w; i += kBlock) contact_min(r, parts[i]);
  block_reduce(r, sh, contact_min);
  if (threadIdx.x == 0) s.contact_ctl[sc].w = r.w == 1.0f ? fminf(fmaxf(r.x, 0.0f), 1.0f) : 0.0f;
}
// applyContactStep (S:2921-2929).
// The first application's result is the pose the validation caches
// (kept for broad-phase scenes, whose cache of a newly listed pair is
// evaluated there).
__global__ void __launch_bounds__(kBlock) k_apply_contact_step(DeviceState s, PassParams p, uint32_t first) {
  const uint32_t v = logical_block(p.xcd_remap) * kBlock + threadIdx.x;
  if (v >= p.vertex_count) return;
  const uint32_t sc = vscene(s, v);
  if (sc == kNoScene) return;
  const bool keep = first && s.bp_nv && (s.bp_ctl[sc].x & kBpScene);
  if (stop_latched(s, p, sc)) {
    if (keep) s.bp_tval[v] = s.pos[p.tgt][v];
    return;
  }
  const uint4 b1 = s.scene_blk1[sc];
  if (b1.z == b1.w) return;  // a scene without candidates takes no contact step
  const float a = fminf(fmaxf(s.contact_ctl[sc].w, 0.0f), 1.0f);
  const f3 src = load_pos(s.pos[p.src], v), c = load_pos(s.pos[p.tgt], v);
  const f3 x = src + a * (c - src);
  s.pos[p.tgt][v] = make_float4(x.x, x.y, x.z, 1.0f);
  if (keep) s.bp_tval[v] = make_float4(x.x, x.y, x.z, 1.0f);
}
// promoteAcceptedContactCache (S:2695-2710).
// Launched after the host swapped the source and target cache arrays and
// summaries (the validated target cache becomes the source without a copy;
// the validation wrote each target summary with the lagged force, which is
// unchanged within a frame). A scene that does not promote copies its
// previous source cache back from what are now the target arrays (nothing
// reads the target cache before the next validation rewrites it).
// A bounded grid: each workgroup first scans the scenes and leaves when
// every scene with candidates promoted (the common case), else strides over
// the candidate blocks of the scenes that did not.
constexpr uint32_t kPromoteGrid = 1024;
__device__ __forceinline__ bool contact_promotes(const DeviceState& s, uint32_t sc) {
  const float alpha = s.contact_ctl[sc].w;
  return s.control[4 * size_t(sc) + 1].x == 1.0f && finite_s(alpha) && alpha > 0.0f;
}
__global__ void __launch_bounds__(kBlock) k_promote_contact(DeviceState s, PassParams p) {
  __shared__ uint32_t any;
  if (threadIdx.x == 0) any = 0u;
  __syncthreads();
  for (uint32_t sc = threadIdx.x; sc < p.scene_count; sc += kBlock) {
    const uint4 b1 = s.scene_blk1[sc];
    if (b1.z != b1.w && s.scene_cnt[sc].x != 0u && !contact_promotes(s, sc)) any = 1u;
  }
  __syncthreads();
  if (!any) return;  // uniform per block
  for (uint32_t blk = blockIdx.x; blk < s.NC / kBlock; blk += gridDim.x) {
    const uint32_t sc = s.cblock_scene[blk];
    if (sc == kNoScene || contact_promotes(s, sc)) continue;  // uniform per candidate block
    const uint32_t c = blk * kBlock + threadIdx.x;
    if (c >= s.scene_end[sc].w) continue;
    s.src_c[c] = s.tgt_c[c];
    s.src_w[c] = s.tgt_w[c];
    s.src_d[c] = s.tgt_d[c];
  }
}
// Copies every live candidate's source cache into the target array, which
// the host then names the source again (a frame with an odd number of
// promotions ends with the arrays in their original roles).
__global__ void __launch_bounds__(kBlock) k_contact_cache_home(DeviceState s) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t sc = s.cblock_scene[blockIdx.x];
  if (sc == kNoScene || c >= s.scene_end[sc].w) return;
  s.tgt_c[c] = s.src_c[c];
  s.tgt_w[c] = s.src_w[c];
  s.tgt_d[c] = s.src_d[c];
}

// ------------------------------------------------------------------ broad-phase contact
// A broad-phase scene's candidates are the reference's static superset
// (buildStaticIpcContactCandidates, src/simulation/cpu/ipc-contact.ts:
// 600-665) restricted to pairs whose AABB distance lower bound (S:525-550) at
// the build pose is within the build's reach, in the superset's canonical
// order (VT point-major over sorted triangles, then EE pairs i < j over
// sorted edges), plus every pair with a non-finite point or a pin that
// snaps by more than half the step bound. A pair outside the list stays
// beyond the activation reach (S:519-523) until the next check point, where
// the list is rebuilt if motion since the build could have brought one
// within it; such a pair is never active, never limits the safe step and
// has no lagged friction, so stepping the list gives the superset's results
// bit for bit.

// ipcAabbDistanceLowerBound with one rounding per operation (no
// contraction), the sequence the oracle evaluates.
__device__ __forceinline__ float bp_lower_bound(f3 lo, f3 hi, f3 lo2, f3 hi2) {
  const float gx = fmaxf(fmaxf(lo2.x - hi.x, lo.x - hi2.x), 0.0f);
  const float gy = fmaxf(fmaxf(lo2.y - hi.y, lo.y - hi2.y), 0.0f);
  const float gz = fmaxf(fmaxf(lo2.z - hi.z, lo.z - hi2.z), 0.0f);
  return __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(gx, gx), __fmul_rn(gy, gy)), __fmul_rn(gz, gz)));
}
__device__ __forceinline__ bool bp_building(const DeviceState& s, uint32_t sc) {
  return sc != kNoScene && (s.bp_ctl[sc].x & (kBpScene | kBpNeed)) == (kBpScene | kBpNeed);
}
// Work lists of a build. bp_list has four segments (vertex, surface
// triangle, surface edge and candidate blocks of every broad-phase scene);
// after the rebuild decision, bp_work holds, per segment, the positions in
// that segment of the blocks whose scene rebuilds now (any order), and
// bp_work_n their counts. Every build pass loops over its segment's list
// with a bounded grid, so a check point at which no scene rebuilds costs one
// counter read per workgroup instead of a launch over every listed block.
enum : uint32_t { kBpSegV = 0, kBpSegS = 1, kBpSegE = 2, kBpSegC = 3 };
__device__ __forceinline__ uint32_t bp_seg_base(const DeviceState& s, uint32_t seg) {
  return seg == kBpSegV ? 0u : seg == kBpSegS ? s.bp_nv : seg == kBpSegE ? s.bp_nv + s.bp_ns : s.bp_nv + s.bp_ns + s.bp_ne;
}
// Calls f(item) for the work of segment `seg`, `per` items per listed
// block; item = (position in the segment) * per + sub-index, the block
// index the pass would have had in a launch over the whole segment. All
// threads of a workgroup visit the same items; LDS is free again between
// items.
template <class F>
__device__ __forceinline__ void bp_for_work(const DeviceState& s, uint32_t seg, uint32_t per, F&& f) {
  const uint32_t n = s.bp_work_n[seg] * per;
  const uint32_t* list = s.bp_work + bp_seg_base(s, seg);
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    f(list[w / per] * per + w % per);
    __syncthreads();
  }
}
// A point the reach argument does not cover: non-finite, or a pin whose
// snap to rest (S:2997-3003) may exceed half the step bound.
__device__ __forceinline__ bool bp_wild(const DeviceState& s, const PassParams& p, uint32_t v, f3 x) {
  if (!finite3(x)) return true;
  if (s.vtx_info[v].z == 0u) return false;
  const f3 d = xyz(s.vtx_rest[v]) - x;
  return !(sqrtf(dot(d, d)) <= 0.5f * fmaxf(p.max_step, 0.0f));
}
__device__ __forceinline__ uint32_t bp_bits(float a) { return finite_s(a) ? __float_as_uint(fabsf(a)) : 0x7f800000u; }
__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) {
#pragma unroll
  for (uint32_t m = 32; m >= 1; m >>= 1) v = max(v, uint32_t(__shfl_xor(int(v), int(m))));
  return v;
}

// Order-preserving integer key of a float (for atomic min/max) and back.
__device__ __forceinline__ uint32_t f2key(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
// Atomic max (or min) of `value` into counters[key], lanes of the wavefront
// with equal keys combined into one atomic.
template <bool kMax>
__device__ __forceinline__ void wave_keyed_extreme(uint32_t* counters, uint32_t key, uint32_t value, bool active) {
  const uint32_t lane = threadIdx.x % 64u;
  uint64_t pending = __ballot(active);
  while (pending) {
    const uint32_t leader = uint32_t(__builtin_ctzll(pending));
    const uint32_t lk = uint32_t(__shfl(int(key), int(leader)));
    const bool mine = active && key == lk && ((pending >> lane) & 1ull);
    uint32_t v = mine ? value : (kMax ? 0u : 0xffffffffu);
#pragma unroll
    for (uint32_t m = 32; m >= 1; m >>= 1) {
      const uint32_t o = uint32_t(__shfl_xor(int(v), int(m)));
      v = kMax ? max(v, o) : min(v, o);
    }
    if (lane == leader) {
      if (kMax) atomicMax(counters + lk, v);
      else atomicMin(counters + lk, v);
    }
    pending &= ~__ballot(mine);
  }
}
// Per-axis motion of every broad-phase scene since its last build and since
// the frame started, and each body's deviation from its reference vertex's
// motion since the build (maxima as float bits, order independent).
__global__ void __launch_bounds__(kBlock) k_bp_motion(DeviceState s, PassParams p) {
  const uint32_t vb = s.bp_list[blockIdx.x];
  const uint32_t sc = s.vblock_scene[vb];
  if (sc == kNoScene) return;  // uniform per block
  const uint32_t v = vb * kSceneBlock + threadIdx.x;
  const bool real = v - s.scene_blk0[sc].x * kSceneBlock < s.scene_cnt[sc].x;
  const bool bodies = s.bp_bodies[sc].z != 0u;
  // m[0..2]: largest x - x_build per axis (order keys), m[3..5]: largest
  // |x - x_frame| (float bits), m[6..8]: smallest x - x_build, kept as the
  // complement of its key so that every entry reduces by a maximum.
  uint32_t m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dev[3] = {0, 0, 0}, body = 0;
  if (real) {
    const f3 x = load_pos(s.pos[p.src], v), b = xyz(s.bp_x[v]), o = load_pos(s.old, v);
    const f3 d = x - b;
    const bool fin = finite3(d);  // a non-finite motion spans every axis
    m[0] = f2key(fin ? d.x : INFINITY), m[1] = f2key(fin ? d.y : INFINITY), m[2] = f2key(fin ? d.z : INFINITY);
    m[6] = ~f2key(fin ? d.x : -INFINITY), m[7] = ~f2key(fin ? d.y : -INFINITY), m[8] = ~f2key(fin ? d.z : -INFINITY);
    m[3] = bp_bits(x.x - o.x), m[4] = bp_bits(x.y - o.y), m[5] = bp_bits(x.z - o.z);
    if (bodies) {
      body = s.vtx_info[v].w;
      const uint32_t r = s.bp_bref[body];
      const f3 d = load_pos(s.pos[p.src], r) - xyz(s.bp_x[r]);
      const f3 e = (x - b) - d;
      dev[0] = bp_bits(e.x), dev[1] = bp_bits(e.y), dev[2] = bp_bits(e.z);
    }
  }
  // Wavefront maxima, then the block's in LDS: one lane per block updates
  // the scene's accumulators (every block of a large sheet hits the same
  // nine words).
  __shared__ uint32_t wm[kBlock / 64][9];
#pragma unroll
  for (int k = 0; k < 9; ++k) m[k] = wave_max_u(m[k]);
  if (threadIdx.x % 64u == 0u)
    for (int k = 0; k < 9; ++k) wm[threadIdx.x / 64u][k] = m[k];
  __syncthreads();
  if (threadIdx.x == 0u) {
    for (uint32_t w = 1; w < kBlock / 64; ++w)
      for (int k = 0; k < 9; ++k) m[k] = max(m[k], wm[w][k]);
    uint32_t* acc = reinterpret_cast<uint32_t*>(s.bp_motion + 3 * size_t(sc));
    for (int k = 0; k < 3; ++k) atomicMax(acc + k, m[k]);
    for (int k = 0; k < 3; ++k) atomicMax(acc + 4 + k, m[3 + k]);
    for (int k = 0; k < 3; ++k) atomicMin(acc + 8 + k, ~m[6 + k]);
  }
  if (bodies) {  // uniform per block
    uint32_t* bd = reinterpret_cast<uint32_t*>(s.bp_bdev);
#pragma unroll
    for (int k = 0; k < 3; ++k) wave_keyed_extreme<true>(bd + k, 4u * body, dev[k], real);
  }
}
// Per scene with the body rule: is every unlisted pair of every two bodies
// (or of one body) still beyond the activation reach?
__device__ bool bp_bodies_keep(const DeviceState& s, const PassParams& p, uint32_t sc, float reach, float act) {
  const uint4 bd = s.bp_bodies[sc];
  const float4* pos = s.pos[p.src];
  const float slack = 1.0e-4f * act;
  for (uint32_t a = bd.x; a < bd.x + bd.y; ++a) {
    const uint32_t ra = s.bp_bref[a];
    if (ra == kNoVertex) continue;
    const f3 da = load_pos(pos, ra) - xyz(s.bp_x[ra]);
    const uint4 ea4 = s.bp_bdev[a];
    const f3 ea = mk(__uint_as_float(ea4.x), __uint_as_float(ea4.y), __uint_as_float(ea4.z));
    const uint4 alo = s.bp_bbox[2 * size_t(a)], ahi = s.bp_bbox[2 * size_t(a) + 1];
    for (uint32_t b = a; b < bd.x + bd.y; ++b) {
      const uint32_t rb = s.bp_bref[b];
      if (rb == kNoVertex) continue;
      f3 bound;
      float base = reach;
      if (b == a) {
        bound = 2.0f * ea;
      } else {
        const f3 db = load_pos(pos, rb) - xyz(s.bp_x[rb]);
        const uint4 eb4 = s.bp_bdev[b];
        const f3 eb = mk(__uint_as_float(eb4.x), __uint_as_float(eb4.y), __uint_as_float(eb4.z));
        const f3 dd = da - db;
        bound = ea + eb + mk(fabsf(dd.x), fabsf(dd.y), fabsf(dd.z));
        const uint4 blo = s.bp_bbox[2 * size_t(b)], bhi = s.bp_bbox[2 * size_t(b) + 1];
        const float gap = bp_lower_bound(mk(key2f(alo.x), key2f(alo.y), key2f(alo.z)),
                                         mk(key2f(ahi.x), key2f(ahi.y), key2f(ahi.z)),
                                         mk(key2f(blo.x), key2f(blo.y), key2f(blo.z)),
                                         mk(key2f(bhi.x), key2f(bhi.y), key2f(bhi.z)));
        if (finite_s(gap)) base = fmaxf(base, gap);
      }
      const float grow = sqrtf(dot(bound, bound));
      if (!(base - grow > act + slack)) return false;
    }
  }
  return true;
}
// Per scene: rebuild now? AABB lower bounds are invariant under a common
// translation, so with every vertex's motion since the build inside
// [lo_a, hi_a] per axis a, the gap of two boxes along a shrinks by at most
// hi_a - lo_a and a lower bound by at most |hi - lo|: the list stays complete
// while |hi - lo| < reach - activation reach (at most 2|m| with m the
// per-axis largest |motion|; a sheet that falls or swings about its pins
// spans far less than twice its largest motion). A rebuild's reach keeps the
// margin past the activation reach and covers every pair that was inside
// the barrier at the frame start (the pairs with lagged friction).
template <bool kCheck>
__global__ void __launch_bounds__(kBlock) k_bp_decide(DeviceState s, PassParams p, uint32_t carry) {
  // Work-list counts (k_bp_worklist) and the long-row count (k_bp_row_sort).
  if (blockIdx.x == 0 && threadIdx.x < 5u) s.bp_work_n[threadIdx.x] = 0u;
  const uint32_t sc = blockIdx.x * kBlock + threadIdx.x;
  if (sc >= p.scene_count) return;
  uint4 ctl = s.bp_ctl[sc];
  if (!(ctl.x & kBpScene) || s.scene_cnt[sc].z == 0u) return;
  uint4* mot = s.bp_motion + 3 * size_t(sc);
  const uint4 mb = mot[0], mf = mot[1], ml = mot[2];
  mot[0] = make_uint4(0, 0, 0, 0);
  mot[1] = make_uint4(0, 0, 0, 0);
  mot[2] = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
  const bool bounded = p.max_step > 0.0f;
  const float act = bounded ? p.ipc_dhat + 2.0f * (p.max_step + ipc_rounding_reach(p, s.scene_scale[sc])) : INFINITY;
  float4 r = s.bp_reach[sc];
  bool need = (ctl.x & kBpForce) != 0u;
  if (kCheck && !need && r.x != INFINITY) {
    const float bx = key2f(mb.x) - key2f(ml.x), by = key2f(mb.y) - key2f(ml.y), bz = key2f(mb.z) - key2f(ml.z);
    const float grow = sqrtf(bx * bx + by * by + bz * bz);
    need = !(grow < r.x - act - 1.0e-4f * act);
    if (need && s.bp_bodies[sc].z != 0u) need = !bp_bodies_keep(s, p, sc, r.x, act);
  }
  if (s.bp_bodies[sc].z != 0u) {  // deviations restart; a rebuild also restarts the body boxes
    const uint4 bd = s.bp_bodies[sc];
    for (uint32_t b = bd.x; b < bd.x + bd.y; ++b) {
      s.bp_bdev[b] = make_uint4(0, 0, 0, 0);
      if (need) {
        s.bp_bbox[2 * size_t(b)] = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0);
        s.bp_bbox[2 * size_t(b) + 1] = make_uint4(0, 0, 0, 0);
      }
    }
  }
  if (need) {
    const float fx = __uint_as_float(mf.x), fy = __uint_as_float(mf.y), fz = __uint_as_float(mf.z);
    const float frame = 2.0f * sqrtf(fx * fx + fy * fy + fz * fz);
    float reach = bounded ? act + r.y : INFINITY;
    reach = fmaxf(reach, (p.ipc_dhat + frame) * 1.0001f);
    if (!finite_s(frame)) reach = INFINITY;
    r.x = reach;
    r.z += 1.0f;  // builds
    s.bp_reach[sc] = r;
    uint4 sz = s.bp_size[sc];
    sz.w = carry ? ctl.z : 0u;  // the list a carrying rebuild looks lagged data up in
    s.bp_size[sc] = sz;
    ctl.x = (ctl.x | kBpNeed) & ~kBpForce;
  } else {
    ctl.x &= ~kBpNeed;
  }
  s.bp_ctl[sc] = ctl;
}
// The work lists of the build the decision above asked for.
__global__ void __launch_bounds__(kBlock) k_bp_worklist(DeviceState s) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t total = s.bp_nv + s.bp_ns + s.bp_ne + s.bp_nc;
  uint32_t seg = kBpSegC;
  bool take = false;
  if (i < total) {
    seg = i < s.bp_nv ? kBpSegV : i < s.bp_nv + s.bp_ns ? kBpSegS : i < s.bp_nv + s.bp_ns + s.bp_ne ? kBpSegE : kBpSegC;
    const uint32_t* table = seg == kBpSegV ? s.vblock_scene
                            : seg == kBpSegS ? s.sblock_scene
                            : seg == kBpSegE ? s.eblock_scene
                                             : s.cblock_scene;
    take = bp_building(s, table[s.bp_list[i]]);
  }
  const uint32_t lane = threadIdx.x % 64u;
  for (uint32_t g = 0; g < 4u; ++g) {
    const bool mine = take && seg == g;
    const uint64_t m = __ballot(mine);
    if (!m) continue;
    const uint32_t leader = uint32_t(__builtin_ctzll(m));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(s.bp_work_n + g, uint32_t(__popcll(m)));
    base = uint32_t(__shfl(int(base), int(leader)));
    const uint32_t first = bp_seg_base(s, g);
    if (mine) s.bp_work[first + base + uint32_t(__popcll(m & ((1ull << lane) - 1ull)))] = i - first;
  }
}
// Triangle and edge AABBs at the source pose, and the AABB of each run of
// 64 consecutive primitives (w = 1 when a point is wild).
__device__ __forceinline__ void bp_chunk_box(f3 lo, f3 hi, float wild, float4* chunk) {
#pragma unroll
  for (uint32_t m = 32; m >= 1; m >>= 1) {
    lo.x = fminf(lo.x, __shfl_xor(lo.x, int(m)));
    lo.y = fminf(lo.y, __shfl_xor(lo.y, int(m)));
    lo.z = fminf(lo.z, __shfl_xor(lo.z, int(m)));
    hi.x = fmaxf(hi.x, __shfl_xor(hi.x, int(m)));
    hi.y = fmaxf(hi.y, __shfl_xor(hi.y, int(m)));
    hi.z = fmaxf(hi.z, __shfl_xor(hi.z, int(m)));
    wild = fmaxf(wild, __shfl_xor(wild, int(m)));
  }
  if (threadIdx.x % 64u == 0u) {
    chunk[0] = make_float4(lo.x, lo.y, lo.z, wild);
    chunk[1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
  }
}
__device__ void bp_bounds_item(const DeviceState& s, const PassParams& p, uint32_t item) {
  const bool tris = item < s.bp_ns;
  const uint32_t blk = s.bp_list[s.bp_nv + item];
  const uint32_t sc = tris ? s.sblock_scene[blk] : s.eblock_scene[blk];
  if (!bp_building(s, sc)) return;
  const uint32_t i = blk * kSceneBlock + threadIdx.x;
  const uint4 b2 = s.scene_blk2[sc];
  const uint32_t end = (tris ? b2.x : b2.z) * kSceneBlock + (tris ? s.bp_size[sc].x : s.bp_size[sc].y);
  const float4* pos = s.pos[p.src];
  f3 lo = mk(INFINITY, INFINITY, INFINITY), hi = mk(-INFINITY, -INFINITY, -INFINITY);
  float wild = 0.0f;
  if (i < end) {
    if (tris) {
      const uint4 ix = s.bp_tri[i];
      const f3 a = load_pos(pos, ix.x), b = load_pos(pos, ix.y), c = load_pos(pos, ix.z);
      lo = vmin3(a, vmin3(b, c));
      hi = vmax3(a, vmax3(b, c));
      wild = (bp_wild(s, p, ix.x, a) || bp_wild(s, p, ix.y, b) || bp_wild(s, p, ix.z, c)) ? 1.0f : 0.0f;
      s.bp_tbox[2 * size_t(i)] = make_float4(lo.x, lo.y, lo.z, wild);
      s.bp_tbox[2 * size_t(i) + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
    } else {
      const uint2 ix = s.bp_edge[i];
      const f3 a = load_pos(pos, ix.x), b = load_pos(pos, ix.y);
      lo = vmin3(a, b);
      hi = vmax3(a, b);
      wild = (bp_wild(s, p, ix.x, a) || bp_wild(s, p, ix.y, b)) ? 1.0f : 0.0f;
      s.bp_ebox[2 * size_t(i)] = make_float4(lo.x, lo.y, lo.z, wild);
      s.bp_ebox[2 * size_t(i) + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
    }
  }
  bp_chunk_box(lo, hi, wild, (tris ? s.bp_tchunk : s.bp_echunk) + 2 * size_t(i / 64u));
}
__global__ void __launch_bounds__(kBlock) k_bp_bounds(DeviceState s, PassParams p) {
  bp_for_work(s, kBpSegS, 1u, [&](uint32_t item) { bp_bounds_item(s, p, item); });
  bp_for_work(s, kBpSegE, 1u, [&](uint32_t item) { bp_bounds_item(s, p, s.bp_ns + item); });
}
// Superchunks: the AABB of 64 consecutive chunks (4096 primitives, 16
// blocks) of a scene, stored at the index of its first chunk. The block of a
// listed segment that starts a superchunk (scene-local block index a
// multiple of 16) reduces it with its first wavefront.
__device__ void bp_super_item(const DeviceState& s, uint32_t item) {
  const bool tris = item < s.bp_ns;
  const uint32_t blk = s.bp_list[s.bp_nv + item];
  const uint32_t sc = tris ? s.sblock_scene[blk] : s.eblock_scene[blk];
  if (!bp_building(s, sc)) return;
  const uint4 b2 = s.scene_blk2[sc];
  const uint32_t first = tris ? b2.x : b2.z;
  const uint32_t local = blk - first;
  if (local % 16u != 0u || threadIdx.x >= 64u) return;
  const uint32_t n = tris ? s.bp_size[sc].x : s.bp_size[sc].y, nchunk = (n + 63u) / 64u;
  const uint32_t c = 4u * local + threadIdx.x;  // scene-local chunk of this lane
  const size_t gc = size_t(blk) * 4u;          // global index of the superchunk's first chunk
  const float4* chunk = tris ? s.bp_tchunk : s.bp_echunk;
  f3 lo = mk(INFINITY, INFINITY, INFINITY), hi = mk(-INFINITY, -INFINITY, -INFINITY);
  float wild = 0.0f;
  if (c < nchunk) {
    const float4 a = chunk[2 * (gc + threadIdx.x)], b = chunk[2 * (gc + threadIdx.x) + 1];
    lo = xyz(a);
    hi = xyz(b);
    wild = a.w;
  }
  bp_chunk_box(lo, hi, wild, (tris ? s.bp_tsuper : s.bp_esuper) + 2 * gc);
}
__global__ void __launch_bounds__(kBlock) k_bp_super(DeviceState s) {
  bp_for_work(s, kBpSegS, 1u, [&](uint32_t item) { bp_super_item(s, item); });
  bp_for_work(s, kBpSegE, 1u, [&](uint32_t item) { bp_super_item(s, s.bp_ns + item); });
}
// One wavefront per query: a vertex (VT, canonical vertex order) or an edge
// (EE). Lanes test 64 superchunk boxes at a time, then the 64 chunk boxes
// of each surviving superchunk; the surviving chunks are
// visited in order and each lane tests one primitive, so matches come out
// in canonical order and a ballot gives each its slot. The count pass
// stores the number of matches of each query, the write pass the records.
template <bool kWrite>
__device__ void bp_vt_item(const DeviceState& s, const PassParams& p, uint32_t item) {
  const uint32_t vb = s.bp_list[item / 64u];
  const uint32_t sc = s.vblock_scene[vb];
  if (!bp_building(s, sc)) return;
  const uint32_t lane = threadIdx.x % 64u;
  const uint32_t qa = vb * kSceneBlock + (item % 64u) * 4u + threadIdx.x / 64u;
  const uint32_t q = qa - s.scene_blk0[sc].x * kSceneBlock;
  if (q >= s.scene_cnt[sc].x) {  // padding: no matches (uniform per wavefront)
    if (!kWrite && lane == 0u) s.bp_vcnt[qa] = 0u;
    return;
  }
  if (kWrite) {  // a query with no match writes nothing: its offset equals the next one's
    const uint32_t next = q + 1u < s.scene_cnt[sc].x ? s.bp_vcnt[qa + 1u]
                          : s.bp_size[sc].y ? s.bp_ecnt[s.scene_blk2[sc].z * kSceneBlock] : s.bp_ctl[sc].w;
    if (next == s.bp_vcnt[qa]) return;  // uniform per wavefront
  }
  const uint32_t dv = s.bp_qv[qa];
  const f3 x = load_pos(s.pos[p.src], dv);
  const bool wild = bp_wild(s, p, dv, x);
  const uint2 ring = s.ring_row[dv];
  const uint32_t my_ring = lane < ring.y ? s.ring_ids[ring.x + lane] : kNoVertex;
  const uint4 sz = s.bp_size[sc];
  const uint32_t t0 = s.scene_blk2[sc].x * kSceneBlock, nt = sz.x, nchunk = (nt + 63u) / 64u;
  const float reach = s.bp_reach[sc].x;
  const uint32_t cb = s.scene_blk1[sc].z * kSceneBlock, at = kWrite ? s.bp_vcnt[qa] : 0u;
  const uint32_t nsuper = (nchunk + 63u) / 64u;
  uint32_t n = 0;
  for (uint32_t s0 = 0; s0 < nsuper; s0 += 64u) {
    // Superchunk boxes contain their chunk boxes, so a chunk that passes
    // lies in a superchunk that passes; chunks are still visited in order.
    bool spass = false;
    if (s0 + lane < nsuper) {
      const size_t si = t0 / 64u + size_t(s0 + lane) * 64u;
      const float4 lo = s.bp_tsuper[2 * si], hi = s.bp_tsuper[2 * si + 1];
      spass = wild || lo.w != 0.0f || !(bp_lower_bound(x, x, xyz(lo), xyz(hi)) > reach);
    }
    uint64_t supers = __ballot(spass);
    while (supers) {
      const uint32_t c0 = (s0 + uint32_t(__builtin_ctzll(supers))) * 64u;
      supers &= supers - 1ull;
      bool pass = false;
      if (c0 + lane < nchunk) {
        const size_t ci = t0 / 64u + c0 + lane;
        const float4 lo = s.bp_tchunk[2 * ci], hi = s.bp_tchunk[2 * ci + 1];
        pass = wild || lo.w != 0.0f || !(bp_lower_bound(x, x, xyz(lo), xyz(hi)) > reach);
      }
      uint64_t chunks = __ballot(pass);
      while (chunks) {
        const uint32_t k = uint32_t(__builtin_ctzll(chunks));
        chunks &= chunks - 1ull;
        const uint32_t tl = (c0 + k) * 64u + lane;
        bool in = false;
        uint4 ix = make_uint4(kNoVertex, kNoVertex, kNoVertex, 0);
        if (tl < nt) {
          const size_t t = t0 + tl;
          const float4 lo = s.bp_tbox[2 * t], hi = s.bp_tbox[2 * t + 1];
          in = wild || lo.w != 0.0f || !(bp_lower_bound(x, x, xyz(lo), xyz(hi)) > reach);
          if (in) ix = s.bp_tri[t];
        }
        if (__ballot(in) == 0ull) continue;  // no triangle of the chunk within reach (uniform per wavefront)
        // isInClosedOneRing (ipc-contact.ts:586-592) against the query vertex.
        bool ringed = ix.x == dv || ix.y == dv || ix.z == dv;
        for (uint32_t j = 0; j < ring.y; ++j) {
          const uint32_t r = uint32_t(__shfl(int(my_ring), int(j)));
          ringed = ringed || r == ix.x || r == ix.y || r == ix.z;
        }
        in = in && !ringed;
        const uint64_t hits = __ballot(in);
        if (kWrite && in) {
          const uint32_t slot = at + n + uint32_t(__popcll(hits & ((1ull << lane) - 1ull)));
          if (slot < sz.z) {
            const uint4 rec = make_uint4(dv, ix.x, ix.y, ix.z);
            s.cand[cb + slot] = rec;
            s.cand_meta[cb + slot] = ipc_meta(s, rec, false);
            s.cand_key[cb + slot] = make_uint2(q, tl);
          }
        }
        n += uint32_t(__popcll(hits));
      }
    }
  }
  if (!kWrite && lane == 0u) s.bp_vcnt[qa] = n;
}
template <bool kWrite>
__global__ void __launch_bounds__(kBlock) k_bp_vt(DeviceState s, PassParams p) {
  bp_for_work(s, kBpSegV, 64u, [&](uint32_t item) { bp_vt_item<kWrite>(s, p, item); });
}
template <bool kWrite>
__device__ void bp_ee_item(const DeviceState& s, const PassParams& p, uint32_t item) {
  const uint32_t eb = s.bp_list[s.bp_nv + s.bp_ns + item / 64u];
  const uint32_t sc = s.eblock_scene[eb];
  if (!bp_building(s, sc)) return;
  const uint32_t lane = threadIdx.x % 64u;
  const uint32_t ea = eb * kSceneBlock + (item % 64u) * 4u + threadIdx.x / 64u;
  const uint32_t e0 = s.scene_blk2[sc].z * kSceneBlock;
  const uint32_t i = ea - e0;
  const uint4 sz = s.bp_size[sc];
  if (i >= sz.y) {  // padding: no matches (uniform per wavefront)
    if (!kWrite && lane == 0u) s.bp_ecnt[ea] = 0u;
    return;
  }
  if (kWrite) {  // a query with no match writes nothing: its offset equals the next one's
    const uint32_t next = i + 1u < sz.y ? s.bp_ecnt[ea + 1u] : s.bp_ctl[sc].w;
    if (next == s.bp_ecnt[ea]) return;  // uniform per wavefront
  }
  const uint2 e = s.bp_edge[ea];
  const float4 elo = s.bp_ebox[2 * size_t(ea)], ehi = s.bp_ebox[2 * size_t(ea) + 1];
  const bool wild = elo.w != 0.0f;
  const uint2 r0 = s.ring_row[e.x], r1 = s.ring_row[e.y];
  const uint32_t my0 = lane < r0.y ? s.ring_ids[r0.x + lane] : kNoVertex;
  const uint32_t my1 = lane < r1.y ? s.ring_ids[r1.x + lane] : kNoVertex;
  const uint32_t ne = sz.y, nchunk = (ne + 63u) / 64u;
  const float reach = s.bp_reach[sc].x;
  const uint32_t cb = s.scene_blk1[sc].z * kSceneBlock, at = kWrite ? s.bp_ecnt[ea] : 0u;
  const uint32_t cfirst = (i + 1u) / 64u, nsuper = (nchunk + 63u) / 64u;  // later edges only
  uint32_t n = 0;
  for (uint32_t s0 = cfirst / 64u; s0 < nsuper; s0 += 64u) {
    bool spass = false;
    if (s0 + lane < nsuper) {
      const size_t si = e0 / 64u + size_t(s0 + lane) * 64u;
      const float4 lo = s.bp_esuper[2 * si], hi = s.bp_esuper[2 * si + 1];
      spass = wild || lo.w != 0.0f || !(bp_lower_bound(xyz(elo), xyz(ehi), xyz(lo), xyz(hi)) > reach);
    }
    uint64_t supers = __ballot(spass);
    while (supers) {
      const uint32_t c0 = (s0 + uint32_t(__builtin_ctzll(supers))) * 64u;
      supers &= supers - 1ull;
      bool pass = false;
      if (c0 + lane < nchunk && c0 + lane >= cfirst) {
        const size_t ci = e0 / 64u + c0 + lane;
        const float4 lo = s.bp_echunk[2 * ci], hi = s.bp_echunk[2 * ci + 1];
        pass = wild || lo.w != 0.0f || !(bp_lower_bound(xyz(elo), xyz(ehi), xyz(lo), xyz(hi)) > reach);
      }
      uint64_t chunks = __ballot(pass);
      while (chunks) {
        const uint32_t k = uint32_t(__builtin_ctzll(chunks));
        chunks &= chunks - 1ull;
        const uint32_t jl = (c0 + k) * 64u + lane;
        bool in = false;
        uint2 o = make_uint2(kNoVertex, kNoVertex);
        if (jl > i && jl < ne) {
          const size_t j = e0 + jl;
          const float4 lo = s.bp_ebox[2 * j], hi = s.bp_ebox[2 * j + 1];
          in = wild || lo.w != 0.0f || !(bp_lower_bound(xyz(elo), xyz(ehi), xyz(lo), xyz(hi)) > reach);
          if (in) o = s.bp_edge[j];
        }
        if (__ballot(in) == 0ull) continue;  // no edge of the chunk within reach (uniform per wavefront)
        // Topologically adjacent edges (ipc-contact.ts:647-651): an endpoint of
        // one in the closed one-ring of an endpoint of the other.
        bool adjacent = o.x == e.x || o.x == e.y || o.y == e.x || o.y == e.y;
        for (uint32_t j = 0; j < r0.y; ++j) {
          const uint32_t r = uint32_t(__shfl(int(my0), int(j)));
          adjacent = adjacent || r == o.x || r == o.y;
        }
        for (uint32_t j = 0; j < r1.y; ++j) {
          const uint32_t r = uint32_t(__shfl(int(my1), int(j)));
          adjacent = adjacent || r == o.x || r == o.y;
        }
        in = in && !adjacent;
        const uint64_t hits = __ballot(in);
        if (kWrite && in) {
          const uint32_t slot = at + n + uint32_t(__popcll(hits & ((1ull << lane) - 1ull)));
          if (slot < sz.z) {
            const uint4 rec = make_uint4(e.x, e.y, o.x, o.y);
            s.cand[cb + slot] = rec;
            s.cand_meta[cb + slot] = ipc_meta(s, rec, true);
            s.cand_key[cb + slot] = make_uint2(i | 0x80000000u, jl);
          }
        }
        n += uint32_t(__popcll(hits));
      }
    }
  }
  if (!kWrite && lane == 0u) s.bp_ecnt[ea] = n;
}
template <bool kWrite>
__global__ void __launch_bounds__(kBlock) k_bp_ee(DeviceState s, PassParams p) {
  bp_for_work(s, kBpSegE, 64u, [&](uint32_t item) { bp_ee_item<kWrite>(s, p, item); });
}
// Scans: block sums of a count array over the listed blocks, one serial pass
// per scene over its blocks in order, then exclusive offsets in place.
enum : uint32_t { kBpScanVt = 0, kBpScanEe = 1, kBpScanRows = 2 };
__device__ __forceinline__ uint32_t block_sum_u(uint32_t v, uint32_t* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) sh[threadIdx.x] += sh[threadIdx.x + stride];
    __syncthreads();
  }
  const uint32_t r = sh[0];
  __syncthreads();
  return r;
}
__device__ __forceinline__ uint32_t block_exclusive_u(uint32_t v, uint32_t* sh) {  // Hillis-Steele
  sh[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t d = 1; d < kBlock; d <<= 1) {
    const uint32_t add = threadIdx.x >= d ? sh[threadIdx.x - d] : 0u;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  return sh[threadIdx.x] - v;
}
struct BpScan {
  uint32_t* data;
  uint32_t* blk;
  const uint32_t* scene_of;
};
__device__ __forceinline__ BpScan bp_scan_target(const DeviceState& s, uint32_t which) {
  if (which == kBpScanVt) return BpScan{s.bp_vcnt, s.bp_blk_v, s.vblock_scene};
  if (which == kBpScanEe) return BpScan{s.bp_ecnt, s.bp_blk_e, s.eblock_scene};
  return BpScan{s.bp_rowcnt, s.bp_blk_r, s.vblock_scene};
}
__device__ __forceinline__ uint32_t bp_list_block(const DeviceState& s, uint32_t which, uint32_t i) {
  return s.bp_list[which == kBpScanEe ? s.bp_nv + s.bp_ns + i : i];
}
__global__ void __launch_bounds__(kBlock) k_bp_block_sums(DeviceState s, uint32_t which) {
  __shared__ uint32_t sh[kBlock];
  const BpScan t = bp_scan_target(s, which);
  bp_for_work(s, which == kBpScanEe ? kBpSegE : kBpSegV, 1u, [&](uint32_t item) {
    const uint32_t blk = bp_list_block(s, which, item);
    if (!bp_building(s, t.scene_of[blk])) return;  // uniform per block
    const uint32_t v = t.data[blk * kSceneBlock + threadIdx.x];
    const uint32_t sum = block_sum_u(v, sh);
    if (threadIdx.x == 0) t.blk[blk] = sum;
  });
}
// One thread per scene. VT offsets run over the vertex blocks, EE offsets
// continue after the VT total; the row scan runs over the vertex blocks.
__global__ void __launch_bounds__(kBlock) k_bp_scene_scan(DeviceState s, PassParams p, uint32_t rows) {
  const uint32_t sc = blockIdx.x * kBlock + threadIdx.x;
  if (sc >= p.scene_count || !bp_building(s, sc) || s.scene_cnt[sc].z == 0u) return;
  const uint4 b0 = s.scene_blk0[sc], b2 = s.scene_blk2[sc];
  uint32_t run = 0;
  if (rows) {
    for (uint32_t b = b0.x; b < b0.y; ++b) {
      const uint32_t c = s.bp_blk_r[b];
      s.bp_blk_r[b] = run;
      run += c;
    }
    return;
  }
  for (uint32_t b = b0.x; b < b0.y; ++b) {
    const uint32_t c = s.bp_blk_v[b];
    s.bp_blk_v[b] = run;
    run += c;
  }
  const uint32_t vt = run;
  for (uint32_t b = b2.z; b < b2.w; ++b) {
    const uint32_t c = s.bp_blk_e[b];
    s.bp_blk_e[b] = run;
    run += c;
  }
  uint4 ctl = s.bp_ctl[sc];
  const uint32_t cap = s.bp_size[sc].z;
  ctl.y = min(vt, cap);
  ctl.z = min(run, cap);
  ctl.w = run;
  if (run > cap) ctl.x |= kBpOverflow;
  s.bp_ctl[sc] = ctl;
  uint4 end = s.scene_end[sc];
  end.w = s.scene_blk1[sc].z * kSceneBlock + ctl.z;
  s.scene_end[sc] = end;
}
__global__ void __launch_bounds__(kBlock) k_bp_scan_apply(DeviceState s, uint32_t which) {
  __shared__ uint32_t sh[kBlock];
  const BpScan t = bp_scan_target(s, which);
  bp_for_work(s, which == kBpScanEe ? kBpSegE : kBpSegV, 1u, [&](uint32_t item) {
    const uint32_t blk = bp_list_block(s, which, item);
    if (!bp_building(s, t.scene_of[blk])) return;  // uniform per block
    const size_t i = size_t(blk) * kSceneBlock + threadIdx.x;
    const uint32_t v = t.data[i];
    t.data[i] = t.blk[blk] + block_exclusive_u(v, sh);
  });
}
// The list before a rebuild, for lagged-friction look-ups.
__global__ void __launch_bounds__(kBlock) k_bp_save_prev(DeviceState s) {
  bp_for_work(s, kBpSegC, 1u, [&](uint32_t item) {
    const uint32_t blk = s.bp_list[s.bp_nv + s.bp_ns + s.bp_ne + item];
    const uint32_t sc = s.cblock_scene[blk];
    if (!bp_building(s, sc)) return;
    const uint32_t c = blk * kSceneBlock + threadIdx.x;
    if (c - s.scene_blk1[sc].z * kSceneBlock >= s.bp_size[sc].w) return;
    s.prev_key[c] = s.cand_key[c];
    s.prev_lag_n[c] = s.lag_n[c];
    s.prev_lag_w[c] = s.lag_w[c];
    s.prev_src_c[c] = s.src_c[c];
    s.prev_src_w[c] = s.src_w[c];
  });
}
// Vertex -> candidate rows: counts, offsets (scan), fill, sort.
__global__ void __launch_bounds__(kBlock) k_bp_row_zero(DeviceState s) {
  bp_for_work(s, kBpSegV, 1u, [&](uint32_t item) {
    const uint32_t vb = s.bp_list[item];
    if (!bp_building(s, s.vblock_scene[vb])) return;
    const uint32_t v = vb * kSceneBlock + threadIdx.x;
    s.bp_rowcnt[v] = 0u;
    s.bp_rowfill[v] = 0u;
  });
}
// Adds one to counters[key] for every active lane, lanes of the wavefront
// with equal keys combined into one atomic; returns each lane's value
// before its own increment (equal keys ranked in lane order).
__device__ __forceinline__ uint32_t wave_counted_add(uint32_t* counters, uint32_t key, bool active) {
  const uint32_t lane = threadIdx.x % 64u;
  uint64_t pending = __ballot(active);
  uint32_t rank = 0;
  while (pending) {
    const uint32_t leader = uint32_t(__builtin_ctzll(pending));
    const uint32_t lk = uint32_t(__shfl(int(key), int(leader)));
    const uint64_t same = __ballot(active && key == lk) & pending;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counters + lk, uint32_t(__popcll(same)));
    base = uint32_t(__shfl(int(base), int(leader)));
    if ((same >> lane) & 1ull) rank = base + uint32_t(__popcll(same & ((1ull << lane) - 1ull)));
    pending &= ~same;
  }
  return rank;
}
// Row counts (kFill false) or row entries (kFill true) of every listed
// candidate. Consecutive candidates share their first vertex (the VT point,
// the first EE edge), so those increments are combined per wavefront.
template <bool kFill>
__device__ void bp_rows_item(const DeviceState& s, const PassParams& p, uint32_t carry, uint32_t item) {
  const uint32_t blk = s.bp_list[s.bp_nv + s.bp_ns + s.bp_ne + item];
  const uint32_t sc = s.cblock_scene[blk];
  if (!bp_building(s, sc)) return;  // uniform per block
  const uint32_t c = blk * kSceneBlock + threadIdx.x;
  const bool live = c < s.scene_end[sc].w;
  const uint4 ix = live ? s.cand[c] : make_uint4(0, 0, 0, 0);
  const bool ee = live && (s.cand_meta[c] & 1u) == 1u;
  uint32_t* counter = kFill ? s.bp_rowfill : s.bp_rowcnt;
  const uint32_t r0 = wave_counted_add(counter, ix.x, live);
  const uint32_t r1e = wave_counted_add(counter, ix.y, ee);
  if (!live) return;
  const uint32_t r1 = ee ? r1e : atomicAdd(counter + ix.y, 1u);
  const uint32_t r2 = atomicAdd(counter + ix.z, 1u), r3 = atomicAdd(counter + ix.w, 1u);
  if (!kFill) return;
  s.cand_ids[s.cand_row[ix.x].x + r0] = c;
  s.cand_ids[s.cand_row[ix.y].x + r1] = c;
  s.cand_ids[s.cand_row[ix.z].x + r2] = c;
  s.cand_ids[s.cand_row[ix.w].x + r3] = c;
  if (!carry) return;
  // A rebuild inside a frame. A pair already listed keeps its lagged
  // friction data and source contact cache. A new pair had no lagged force
  // at the frame start (every such pair stays listed all frame), and its
  // cache is what the superset would hold: the contact at the pose of the
  // frame start's lag pass or of the last promoted validation
  // (S:2713-2744, S:2695-2710). The previous list is searched within the
  // query's range of it.
  const uint2 key = s.cand_key[c];
  const uint32_t cb = s.scene_blk1[sc].z * kSceneBlock, old_total = s.bp_size[sc].w;
  const uint32_t e0 = s.scene_blk2[sc].z * kSceneBlock, ne = s.bp_size[sc].y;
  uint32_t lo, hi;
  if (!(key.x & 0x80000000u)) {
    const uint32_t qa = s.scene_blk0[sc].x * kSceneBlock + key.x;
    lo = s.bp_vprev[qa];
    hi = key.x + 1u < s.scene_cnt[sc].x ? s.bp_vprev[qa + 1u] : (ne ? s.bp_eprev[e0] : old_total);
  } else {
    const uint32_t i = key.x & 0x7fffffffu;
    lo = s.bp_eprev[e0 + i];
    hi = i + 1u < ne ? s.bp_eprev[e0 + i + 1u] : old_total;
  }
  hi = min(hi, old_total);
  lo = min(lo, hi);
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2u;
    if (s.prev_key[cb + mid].y < key.y) lo = mid + 1u;
    else hi = mid;
  }
  const uint32_t at = cb + lo;
  if (lo < old_total && s.prev_key[at].x == key.x && s.prev_key[at].y == key.y) {
    s.lag_n[c] = s.prev_lag_n[at];
    s.lag_w[c] = s.prev_lag_w[at];
    s.src_c[c] = s.prev_src_c[at];
    s.src_w[c] = s.prev_src_w[at];
  } else {
    s.lag_n[c] = make_float4(0, 0, 0, 0);
    s.lag_w[c] = make_float4(0, 0, 0, 0);
    ipc_store_cached(s, c, false, ipc_contact_at(s, c, s.bp_cpose, ipc_broad_threshold(p, s.scene_scale[sc])));
  }
  ipc_store_summary(s, c);
}
template <bool kFill>
__global__ void __launch_bounds__(kBlock) k_bp_rows(DeviceState s, PassParams p, uint32_t carry) {
  bp_for_work(s, kBpSegC, 1u, [&](uint32_t item) { bp_rows_item<kFill>(s, p, carry, item); });
}
__global__ void __launch_bounds__(kBlock) k_bp_row_offsets(DeviceState s) {
  bp_for_work(s, kBpSegV, 1u, [&](uint32_t item) {
    const uint32_t vb = s.bp_list[item];
    const uint32_t sc = s.vblock_scene[vb];
    if (!bp_building(s, sc)) return;
    const uint32_t v = vb * kSceneBlock + threadIdx.x;
    // bp_rowcnt holds exclusive offsets after the scan; the count is the
    // difference to the next vertex (or to four times the list length).
    const uint32_t end = 4u * s.bp_ctl[sc].z;
    const uint32_t last = s.scene_blk0[sc].y * kSceneBlock;
    const uint32_t at = s.bp_rowcnt[v];
    const uint32_t next = v + 1u < last ? s.bp_rowcnt[v + 1u] : end;
    s.cand_row[v] = make_uint2(s.bp_rows_at[sc] + at, next - at);
  });
}
constexpr uint32_t kBpMaxRow = 8192;
constexpr uint64_t kBpWorkGrid = 16384;  // workgroups of a work-list pass
constexpr uint64_t kBpLongGrid = 2048;   // workgroups sorting rows longer than a wavefront's
constexpr uint32_t kBpWaveRow = 256;      // rows a wavefront sorts in registers
// Vertex rows sorted ascending in candidate id (the order the local solve
// sums contact rows in). A row of at most 256 entries is sorted by one
// wavefront in registers: entry i = 64 r + lane sits in register r of a
// lane, and a bitonic network exchanges partners i ^ j across lanes
// (__shfl_xor) for j < 64 and between a lane's registers for j >= 64. Longer
// rows go to a list that k_bp_row_sort_long sorts one block each in LDS.
__device__ __forceinline__ void bp_cas(uint32_t& a, uint32_t& b, bool ascending) {
  if ((a > b) == ascending) {
    const uint32_t t = a;
    a = b;
    b = t;
  }
}
__device__ void bp_row_sort_wave(const DeviceState& s, uint32_t item) {
  const uint32_t vb = s.bp_list[item / 64u];
  const uint32_t sc = s.vblock_scene[vb];
  if (!bp_building(s, sc)) return;
  const uint32_t lane = threadIdx.x % 64u;
  const uint32_t v = vb * kSceneBlock + (item % 64u) * 4u + threadIdx.x / 64u;
  const uint2 row = s.cand_row[v];
  const uint32_t n = row.y;
  if (n < 2u) return;  // uniform per wavefront
  if (n > kBpWaveRow) {
    if (lane == 0u) s.bp_long[atomicAdd(s.bp_work_n + 4, 1u)] = v;
    return;
  }
  uint32_t m = 64;
  while (m < n) m <<= 1;
  uint32_t key[4];
#pragma unroll
  for (uint32_t r = 0; r < 4u; ++r) {
    const uint32_t i = 64u * r + lane;
    key[r] = i < n ? s.cand_ids[row.x + i] : 0xffffffffu;
  }
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j == 128u) {
        bp_cas(key[0], key[2], ((0u + lane) & k) == 0u);
        bp_cas(key[1], key[3], ((64u + lane) & k) == 0u);
      } else if (j == 64u) {
        bp_cas(key[0], key[1], ((0u + lane) & k) == 0u);
        bp_cas(key[2], key[3], ((128u + lane) & k) == 0u);
      } else {
        const bool lower = (lane & j) == 0u;
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) {
          const uint32_t o = uint32_t(__shfl_xor(int(key[r]), int(j)));
          const bool ascending = ((64u * r + lane) & k) == 0u;
          key[r] = lower == ascending ? min(key[r], o) : max(key[r], o);
        }
      }
    }
  }
#pragma unroll
  for (uint32_t r = 0; r < 4u; ++r) {
    const uint32_t i = 64u * r + lane;
    if (i < n) s.cand_ids[row.x + i] = key[r];
  }
}
__global__ void __launch_bounds__(kBlock) k_bp_row_sort(DeviceState s) {
  bp_for_work(s, kBpSegV, 64u, [&](uint32_t item) { bp_row_sort_wave(s, item); });
}
// One block per long row: bitonic sort in LDS.
__device__ void bp_row_sort_item(const DeviceState& s, uint32_t* keys, uint32_t v) {
  const uint32_t sc = s.vblock_scene[v / kSceneBlock];
  const uint2 row = s.cand_row[v];
  const uint32_t n = row.y;
  if (n < 2u) return;
  if (n > kBpMaxRow) {
    if (threadIdx.x == 0) atomicOr(&s.bp_ctl[sc].x, kBpOverflow);
    return;
  }
  uint32_t m = 2;
  while (m < n) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += kBlock) keys[i] = i < n ? s.cand_ids[row.x + i] : 0xffffffffu;
  __syncthreads();
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < m; i += kBlock) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const uint32_t a = keys[i], b = keys[l];
          if ((a > b) == ((i & k) == 0u)) {
            keys[i] = b;
            keys[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < n; i += kBlock) s.cand_ids[row.x + i] = keys[i];
}
__global__ void __launch_bounds__(kBlock) k_bp_row_sort_long(DeviceState s) {
  __shared__ uint32_t keys[kBpMaxRow];
  const uint32_t n = s.bp_work_n[4];
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    bp_row_sort_item(s, keys, s.bp_long[w]);
    __syncthreads();
  }
}
// The build pose of every rebuilt scene, and its query offsets (the ranges
// a later rebuild inside a frame searches this list in).
__device__ void bp_commit_item(const DeviceState& s, const PassParams& p, uint32_t item) {
  const bool vertices = item < s.bp_nv;
  const uint32_t blk = s.bp_list[vertices ? item : s.bp_ns + item];
  const uint32_t i = blk * kSceneBlock + threadIdx.x;
  if (vertices) {
    const uint32_t sc = s.vblock_scene[blk];
    if (!bp_building(s, sc)) return;  // uniform per block
    const float4 x = s.pos[p.src][i];
    s.bp_x[i] = x;
    s.bp_vprev[i] = s.bp_vcnt[i];
    if (s.bp_bodies[sc].z != 0u) {  // body AABBs at the build pose
      const bool real = i - s.scene_blk0[sc].x * kSceneBlock < s.scene_cnt[sc].x;
      const uint32_t body = real ? s.vtx_info[i].w : 0u;
      uint32_t* box = reinterpret_cast<uint32_t*>(s.bp_bbox);
      wave_keyed_extreme<false>(box, 8u * body + 0u, f2key(x.x), real);
      wave_keyed_extreme<false>(box, 8u * body + 1u, f2key(x.y), real);
      wave_keyed_extreme<false>(box, 8u * body + 2u, f2key(x.z), real);
      wave_keyed_extreme<true>(box, 8u * body + 4u, f2key(x.x), real);
      wave_keyed_extreme<true>(box, 8u * body + 5u, f2key(x.y), real);
      wave_keyed_extreme<true>(box, 8u * body + 6u, f2key(x.z), real);
    }
  } else {
    if (!bp_building(s, s.eblock_scene[blk])) return;
    s.bp_eprev[i] = s.bp_ecnt[i];
  }
}
__global__ void __launch_bounds__(kBlock) k_bp_commit(DeviceState s, PassParams p) {
  bp_for_work(s, kBpSegV, 1u, [&](uint32_t item) { bp_commit_item(s, p, item); });
  bp_for_work(s, kBpSegE, 1u, [&](uint32_t item) { bp_commit_item(s, p, s.bp_nv + item); });
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }
inline bool has_cloth(const PassParams& p) { return p.tri_count != 0u || p.hinge_count != 0u; }

}  // namespace

uint32_t partial_blocks_needed(uint32_t vertices, uint32_t tets, uint32_t triangles, uint32_t candidates) {
  // Candidate partials use blocks(T) + blocks(triangles) + blocks(V);
  // convergence 2 blocks(V); the contact step blocks(candidates) and its
  // step bound blocks(V).
  const uint32_t a = blocks_for(tets) + blocks_for(triangles) + blocks_for(vertices);
  const uint32_t b = 2 * blocks_for(vertices);
  const uint32_t c = blocks_for(candidates);
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}
size_t partial_bytes(uint32_t blocks) {
  size_t a = sizeof(CandPartial), b = sizeof(ConvPartial);
  return size_t(blocks) * (a > b ? a : b);
}
size_t scene_conv_bytes(uint32_t scenes) { return size_t(scenes) * sizeof(ConvPartial); }

void launch_clear_frame(const DeviceState& s, const PassParams& p, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_frame, dim3(blocks_for(4 * p.scene_count)), dim3(kBlock), 0, st, s, p);
}
void launch_predict(const DeviceState& s, const PassParams& p, hipStream_t st) {
  hipLaunchKernelGGL(k_predict, dim3(blocks_for(p.vertex_count)), dim3(kBlock), 0, st, s, p);
}
void launch_tet_polar(const DeviceState& s, const PassParams& p, hipStream_t st) {
  hipLaunchKernelGGL(k_tet_polar, dim3(blocks_for(p.tet_count)), dim3(kBlock), 0, st, s, p);
}
void launch_vertex_polar(const DeviceState& s, const PassParams& p, hipStream_t st) {
  if (!p.corotated && p.solve_lanes >= 4u) {
    hipLaunchKernelGGL(k_vertex_polar_lanes<4>, dim3(blocks_for(4u * p.vertex_count)), dim3(kBlock), 0, st, s, p);
    return;
  }
  hipLaunchKernelGGL(k_vertex_polar, dim3(blocks_for(p.vertex_count)), dim3(kBlock), 0, st, s, p);
}
// Kernel variant for the step's features: co-rotated (no cloth or contact
// terms), or stable Neo-Hookean with or without the cloth and IPC terms.
#define JGS2_FEATURE_DISPATCH(KERNEL, GRID, ...)                                                      \
  do {                                                                                               \
    const bool cl = has_cloth(p), ipc = p.contact != 0u;                                            \
    if (p.corotated) hipLaunchKernelGGL((KERNEL<true, false, false>), GRID, dim3(kBlock), 0, st, __VA_ARGS__); \
    else if (cl && ipc) hipLaunchKernelGGL((KERNEL<false, true, true>), GRID, dim3(kBlock), 0, st, __VA_ARGS__); \
    else if (cl) hipLaunchKernelGGL((KERNEL<false, true, false>), GRID, dim3(kBlock), 0, st, __VA_ARGS__); \
    else if (ipc) hipLaunchKernelGGL((KERNEL<false, false, true>), GRID, dim3(kBlock), 0, st, __VA_ARGS__); \
    else hipLaunchKernelGGL((KERNEL<false, false, false>), GRID, dim3(kBlock), 0, st, __VA_ARGS__); \
  } while (0)

#define JGS2_FEATURE_DISPATCH_BLOCK(KERNEL, GRID, BLOCK, ...)                                        \
  do {                                                                                               \
    const bool cl = has_cloth(p), ipc = p.contact != 0u;                                            \
    if (p.corotated) hipLaunchKernelGGL((KERNEL<true, false, false>), GRID, BLOCK, 0, st, __VA_ARGS__); \
    else if (cl && ipc) hipLaunchKernelGGL((KERNEL<false, true, true>), GRID, BLOCK, 0, st, __VA_ARGS__); \
    else if (cl) hipLaunchKernelGGL((KERNEL<false, true, false>), GRID, BLOCK, 0, st, __VA_ARGS__); \
    else if (ipc) hipLaunchKernelGGL((KERNEL<false, false, true>), GRID, BLOCK, 0, st, __VA_ARGS__); \
    else hipLaunchKernelGGL((KERNEL<false, false, false>), GRID, BLOCK, 0, st, __VA_ARGS__); \
  } while (0)

#define JGS2_LANES_DISPATCH_P(LPV, GRID, BLOCK, P)                                                   \
  do {                                                                                               \
    const bool cl = has_cloth(P), ipc = (P).contact != 0u;                                          \
    if (cl && ipc) hipLaunchKernelGGL((k_solve_lanes<true, true, LPV>), GRID, BLOCK, 0, st, s, P);  \
    else if (cl) hipLaunchKernelGGL((k_solve_lanes<true, false, LPV>), GRID, BLOCK, 0, st, s, P);   \
    else if (ipc) hipLaunchKernelGGL((k_solve_lanes<false, true, LPV>), GRID, BLOCK, 0, st, s, P);  \
    else hipLaunchKernelGGL((k_solve_lanes<false, false, LPV>), GRID, BLOCK, 0, st, s, P);          \
  } while (0)
// The same for scenes without tetrahedral terms at unpinned vertices, four
// lanes per vertex.
#define JGS2_SHEET_DISPATCH_P(GRID, BLOCK, P)                                                           \
  do {                                                                                                  \
    const bool cl = has_cloth(P), ipc = (P).contact != 0u;                                             \
    if (cl && ipc) hipLaunchKernelGGL((k_solve_lanes<true, true, 4, false>), GRID, BLOCK, 0, st, s, P); \
    else if (cl) hipLaunchKernelGGL((k_solve_lanes<true, false, 4, false>), GRID, BLOCK, 0, st, s, P);  \
    else hipLaunchKernelGGL((k_solve_lanes<false, true, 4, false>), GRID, BLOCK, 0, st, s, P);          \
  } while (0)

// The local solve runs with the autotuned block size (64, 128 or 256) and,
// on the stable Neo-Hookean path, the autotuned lanes per vertex (1, 4, 8
// or 16; one lane runs the sequential kernel).
// Each scene is solved either with lanes (its scene_lanes entry equals
// p.solve_lanes) or one thread per vertex; a launch per class present skips
// the other class's scenes.
void launch_solve(const DeviceState& s, const PassParams& p, hipStream_t st) {
  const uint32_t n = p.colored ? p.color_end - p.color_begin : p.vertex_count;
  if (n == 0) return;
  const uint32_t block = (p.solve_block == 64u || p.solve_block == 128u) ? p.solve_block : kBlock;
  if (!p.corotated && (p.solve_classes & 1u) && p.solve_lanes > 1u) {
    const uint32_t lanes = p.solve_lanes;
    const dim3 grid((n * lanes + block - 1) / block);
    // The contact rows' gather terms first (Jacobi sweeps with contact): the
    // local solve is this pass and the lane sweep together.
    if (p.ipc_local && !p.colored)
      hipLaunchKernelGGL(k_ipc_local_lanes, dim3(blocks_for(16u * p.vertex_count)), dim3(kBlock), 0, st, s, p);
    PassParams q = p;
    dim3 qgrid = grid;
    // Launch blocks of a Jacobi sweep over one kind_blocks list.
    auto list_grid = [&](uint32_t listed, uint32_t threads, uint32_t lpv) {
      return p.colored ? dim3(uint32_t((uint64_t(n) * lpv + threads - 1) / threads))
                       : dim3(listed * (kSceneBlock * lpv / threads));
    };
    if (has_cloth(p) || p.contact != 0u) {
      // A batch with cloth or contact: scenes without either are swept by
      // the tetrahedral variant first, the others by the full variant, whose
      // cubature terms come from k_cub_local_lanes (eight lanes per vertex).
      // Jacobi sweeps launch over per-kind block lists and give the scenes
      // whose unpinned vertices have no tetrahedral or cubature terms (a
      // sheet over a pinned collider) a form of the full variant compiled
      // without those terms.
      const bool mixed = (p.solve_classes & 4u) && (p.solve_classes & 8u);
      if (mixed) {
        PassParams a = p;
        a.kind_filter = 1u;
        const dim3 ag = list_grid(p.kind_blocks1, block, lanes);
        if (ag.x > 0) {
          if (lanes == 4u) hipLaunchKernelGGL((k_solve_lanes<false, false, 4>), ag, dim3(block), 0, st, s, a);
          else if (lanes == 8u) hipLaunchKernelGGL((k_solve_lanes<false, false, 8>), ag, dim3(block), 0, st, s, a);
          else hipLaunchKernelGGL((k_solve_lanes<false, false, 16>), ag, dim3(block), 0, st, s, a);
        }
      }
      if (!p.colored) {
        PassParams w = p;
        w.kind_filter = 4u;
        const dim3 wg = list_grid(p.kind_blocks4, block, lanes);
        if (wg.x > 0) {
          if (lanes == 4u) JGS2_SHEET_DISPATCH_P(wg, dim3(block), w);
          else if (lanes == 8u) JGS2_LANES_DISPATCH_P(8, wg, dim3(block), w);
          else JGS2_LANES_DISPATCH_P(16, wg, dim3(block), w);
        }
        q.kind_filter = 2u;
        qgrid = list_grid(p.kind_blocks2, block, lanes);
      } else if (mixed) {
        q.kind_filter = 5u;
      }
      if (p.cubature_k > 0u && s.cub_local) {
        PassParams c = q;
        c.kind_filter = 3u;
        const dim3 cg = list_grid(p.kind_blocks3, kBlock, 8u);
        q.cub_local = 1u;
        if (cg.x > 0) hipLaunchKernelGGL(k_cub_local_lanes, cg, dim3(kBlock), 0, st, s, c);
      }
    }
    if (qgrid.x > 0) {
      if (lanes == 4u) JGS2_LANES_DISPATCH_P(4, qgrid, dim3(block), q);
      else if (lanes == 8u) JGS2_LANES_DISPATCH_P(8, qgrid, dim3(block), q);
      else JGS2_LANES_DISPATCH_P(16, qgrid, dim3(block), q);
    }
  }
  if (p.corotated || (p.solve_classes & 2u))
    JGS2_FEATURE_DISPATCH_BLOCK(k_solve, dim3((n + block - 1) / block), dim3(block), s, p);
}
void launch_solve_config(const DeviceState& s, const PassParams& p, hipStream_t st) { launch_solve(s, p, st); }
// Streams `n` records of the cubature array through the caches (the timed
// local solve of the autotuner then starts as cold as in a frame, where the
// other passes have replaced the cache contents). The sum lands in the
// scratch partials, which every reduction rewrites before reading.
__global__ void __launch_bounds__(kBlock) k_cache_flush(DeviceState s, uint32_t n) {
  __shared__ uint32_t sh[kBlock];
  uint32_t acc = 0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) acc ^= s.cub[i].y;
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t r = 0;
    for (uint32_t k = 0; k < kBlock; ++k) r ^= sh[k];
    reinterpret_cast<uint32_t*>(s.partials)[blockIdx.x] = r;
  }
}
void launch_cache_flush(const DeviceState& s, uint32_t records, hipStream_t st) {
  if (records == 0) return;
  const uint32_t blocks = std::min<uint32_t>(2048u, s.partial_capacity);
  hipLaunchKernelGGL(k_cache_flush, dim3(blocks), dim3(kBlock), 0, st, s, records);
}
void launch_vertex_energy(const DeviceState& s, const PassParams& p, hipStream_t st) {
  const dim3 grid(blocks_for(p.vertex_count));
  const bool cl = has_cloth(p), ipc = p.contact != 0u;
  if (ipc) {
    hipLaunchKernelGGL(k_ipc_energy_candidates, dim3(blocks_for(s.NC)), dim3(kBlock), 0, st, s, p);
    hipLaunchKernelGGL(k_ipc_energy_lanes, dim3(blocks_for(16u * p.vertex_count)), dim3(kBlock), 0, st, s, p);
  }
  if (cl && ipc) hipLaunchKernelGGL((k_vertex_energy<true, true>), grid, dim3(kBlock), 0, st, s, p);
  else if (cl) hipLaunchKernelGGL((k_vertex_energy<true, false>), grid, dim3(kBlock), 0, st, s, p);
  else if (ipc) hipLaunchKernelGGL((k_vertex_energy<false, true>), grid, dim3(kBlock), 0, st, s, p);
  else hipLaunchKernelGGL((k_vertex_energy<false, false>), grid, dim3(kBlock), 0, st, s, p);
}
void launch_reduce_candidate(const DeviceState& s, const PassParams& p, hipStream_t st) {
  const uint32_t tb = blocks_for(p.tet_count), rb = blocks_for(p.tri_count), vb = blocks_for(p.vertex_count);
  hipLaunchKernelGGL(k_candidate_partials, dim3(tb + rb + vb), dim3(kBlock), 0, st, s, p, tb, rb,
                     make_uint4(0, tb, 0, rb), 0u);
  hipLaunchKernelGGL(k_candidate_final, dim3(p.scene_count), dim3(kBlock), 0, st, s, p, tb, rb, 0u);
}
void launch_reduce_candidate_scene(const DeviceState& s, const PassParams& p, uint32_t scene, uint4 blk0, uint4 blk1,
                                   hipStream_t st) {
  const uint32_t tb = blocks_for(p.tet_count), rb = blocks_for(p.tri_count);
  const uint32_t nt = blk0.w - blk0.z, nr = blk1.y - blk1.x, nv = blk0.y - blk0.x;
  hipLaunchKernelGGL(k_candidate_partials, dim3(nt + nr + nv), dim3(kBlock), 0, st, s, p, tb, rb,
                     make_uint4(blk0.z, nt, blk1.x, nr), blk0.x);
  hipLaunchKernelGGL(k_candidate_final, dim3(1), dim3(kBlock), 0, st, s, p, tb, rb, scene);
}
void launch_apply_candidate(const DeviceState& s, const PassParams& p, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_candidate, dim3(blocks_for(p.vertex_count)), dim3(kBlock), 0, st, s, p);
}
void launch_convergence(const DeviceState& s, const PassParams& p, hipStream_t st) {
  const uint32_t vb = blocks_for(p.vertex_count);
  const bool cl = has_cloth(p), ipc = p.contact != 0u;
  if (p.solve_lanes == 4u || p.solve_lanes == 8u || p.solve_lanes == 16u) {
    if (ipc) hipLaunchKernelGGL(k_ipc_gradient_lanes, dim3(blocks_for(16u * p.vertex_count)), dim3(kBlock), 0, st, s, p);
    // The convergence gather splits a vertex over at most 8 lanes.
    const uint32_t lanes = p.solve_lanes >= 8u ? 8u : 4u;
    const bool rec = p.cloth_base != 0u;
    if (lanes == 8u) {
      if (cl && ipc && rec) hipLaunchKernelGGL((k_convergence_gradient_lanes<true, true, 8, true>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else if (cl && rec) hipLaunchKernelGGL((k_convergence_gradient_lanes<true, false, 8, true>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else if (cl && ipc) hipLaunchKernelGGL((k_convergence_gradient_lanes<true, true, 8>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else if (cl) hipLaunchKernelGGL((k_convergence_gradient_lanes<true, false, 8>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else if (ipc) hipLaunchKernelGGL((k_convergence_gradient_lanes<false, true, 8>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else hipLaunchKernelGGL((k_convergence_gradient_lanes<false, false, 8>), dim3(vb), dim3(kBlock), 0, st, s, p);
    } else {
      if (cl && ipc && rec) hipLaunchKernelGGL((k_convergence_gradient_lanes<true, true, 4, true>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else if (cl && rec) hipLaunchKernelGGL((k_convergence_gradient_lanes<true, false, 4, true>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else if (cl && ipc) hipLaunchKernelGGL((k_convergence_gradient_lanes<true, true, 4>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else if (cl) hipLaunchKernelGGL((k_convergence_gradient_lanes<true, false, 4>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else if (ipc) hipLaunchKernelGGL((k_convergence_gradient_lanes<false, true, 4>), dim3(vb), dim3(kBlock), 0, st, s, p);
      else hipLaunchKernelGGL((k_convergence_gradient_lanes<false, false, 4>), dim3(vb), dim3(kBlock), 0, st, s, p);
    }
  } else if (cl && ipc) hipLaunchKernelGGL((k_convergence_gradient<true, true>), dim3(vb), dim3(kBlock), 0, st, s, p);
  else if (cl) hipLaunchKernelGGL((k_convergence_gradient<true, false>), dim3(vb), dim3(kBlock), 0, st, s, p);
  else if (ipc) hipLaunchKernelGGL((k_convergence_gradient<false, true>), dim3(vb), dim3(kBlock), 0, st, s, p);
  else hipLaunchKernelGGL((k_convergence_gradient<false, false>), dim3(vb), dim3(kBlock), 0, st, s, p);
  hipLaunchKernelGGL(k_convergence_scales, dim3(p.scene_count), dim3(kBlock), 0, st, s);
  hipLaunchKernelGGL(k_convergence_sums, dim3(vb), dim3(kBlock), 0, st, s, p, vb);
  hipLaunchKernelGGL(k_convergence_final, dim3(p.scene_count), dim3(kBlock), 0, st, s, p, vb);
}
void launch_cloth_base(const DeviceState& s, const PassParams& p, uint32_t which, hipStream_t st, uint32_t mode) {
  const uint32_t tb = blocks_for(p.tri_count), hb = blocks_for(p.hinge_count);
  if (tb + hb == 0) return;
  if (mode == 2u) {
    hipLaunchKernelGGL(k_cloth_base_rejected, dim3(std::min(tb + hb, kClothRejectBlocks)), dim3(kBlock), 0, st, s, p,
                       which);
    return;
  }
  hipLaunchKernelGGL(k_cloth_base, dim3(tb + hb), dim3(kBlock), 0, st, s, p, which, tb, mode);
}
void launch_copy_position(const DeviceState& s, const PassParams& p, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_position, dim3(blocks_for(p.vertex_count)), dim3(kBlock), 0, st, s, p);
}
void launch_body_correction(const DeviceState& s, const PassParams& p, hipStream_t st) {
  if (p.body_count == 0) return;
  hipLaunchKernelGGL(k_body_correction, dim3(p.body_count), dim3(kBlock), 0, st, s, p);
  hipLaunchKernelGGL(k_apply_body_correction, dim3(blocks_for(p.vertex_count)), dim3(kBlock), 0, st, s, p);
}
void launch_finalize(const DeviceState& s, const PassParams& p, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(blocks_for(p.vertex_count)), dim3(kBlock), 0, st, s, p);
}
void launch_debug_local_systems(const DeviceState& s, const PassParams& p, hipStream_t st) {
  JGS2_FEATURE_DISPATCH(k_debug_local_systems, dim3(blocks_for(p.vertex_count)), s, p);
}
void launch_lag_contact(const DeviceState& s, const PassParams& p, hipStream_t st) {
  if (s.NC == 0) return;
  hipLaunchKernelGGL(k_lag_contact, dim3(blocks_for(s.NC)), dim3(kBlock), 0, st, s, p);
}
void launch_build_active_rows(const DeviceState& s, const PassParams& p, hipStream_t st) {
  if (s.NC == 0) return;
  hipLaunchKernelGGL(k_candidate_keep, dim3(blocks_for(s.NC)), dim3(kBlock), 0, st, s, p);
  hipLaunchKernelGGL(k_compact_active_rows, dim3(blocks_for(16u * p.vertex_count)), dim3(kBlock), 0, st, s, p);
}
// The six contact step dispatches of one iteration (jgs2-solver.ts:2347-2388):
// candidate caps, reduce, apply, exact validation, reduce, apply.
void launch_contact_step(const DeviceState& s, const PassParams& p, hipStream_t st) {
  const uint32_t cb = blocks_for(s.NC), vb = blocks_for(p.vertex_count);
  if (s.NC == 0) return;
  hipLaunchKernelGGL(k_step_bound_partials, dim3(vb), dim3(kBlock), 0, st, s, p);
  hipLaunchKernelGGL(k_step_bound_final, dim3(p.scene_count), dim3(kBlock), 0, st, s);
  hipLaunchKernelGGL(k_candidate_contact_step, dim3(cb), dim3(kBlock), 0, st, s, p);
  hipLaunchKernelGGL(k_reduce_contact_final, dim3(p.scene_count), dim3(kBlock), 0, st, s);
  hipLaunchKernelGGL(k_apply_contact_step, dim3(vb), dim3(kBlock), 0, st, s, p, 1u);
  hipLaunchKernelGGL(k_validate_contact_step, dim3(cb), dim3(kBlock), 0, st, s, p);
  hipLaunchKernelGGL(k_reduce_contact_final, dim3(p.scene_count), dim3(kBlock), 0, st, s);
  hipLaunchKernelGGL(k_apply_contact_step, dim3(vb), dim3(kBlock), 0, st, s, p, 0u);
}
void launch_promote_contact(const DeviceState& s, const PassParams& p, hipStream_t st) {
  if (s.NC == 0) return;
  hipLaunchKernelGGL(k_promote_contact, dim3(std::min(blocks_for(s.NC), kPromoteGrid)), dim3(kBlock), 0, st, s, p);
}
void launch_contact_cache_home(const DeviceState& s, hipStream_t st) {
  if (s.NC == 0) return;
  hipLaunchKernelGGL(k_contact_cache_home, dim3(blocks_for(s.NC)), dim3(kBlock), 0, st, s);
}
// Broad-phase check point: decide per scene, then (for the scenes that
// rebuild) AABBs, counts, offsets, records and vertex rows. Every kernel
// runs over the arena blocks of broad-phase scenes and skips scenes that do
// not rebuild, so a frame's launches are fixed (graph capture).
void launch_broad_phase(const DeviceState& s, const PassParams& p, bool check, bool carry, hipStream_t st) {
  if (s.bp_nv == 0) return;
  const dim3 B(kBlock);
  const dim3 scenes(blocks_for(p.scene_count));
  const uint32_t cr = carry ? 1u : 0u;
  if (check) {
    hipLaunchKernelGGL(k_bp_motion, dim3(s.bp_nv), B, 0, st, s, p);
    hipLaunchKernelGGL(k_bp_decide<true>, scenes, B, 0, st, s, p, cr);
  } else {
    hipLaunchKernelGGL(k_bp_decide<false>, scenes, B, 0, st, s, p, cr);
  }
  hipLaunchKernelGGL(k_bp_worklist, dim3(blocks_for(s.bp_nv + s.bp_ns + s.bp_ne + s.bp_nc)), B, 0, st, s);
  // The build passes loop over the work lists; a grid of at most
  // kBpWorkGrid workgroups keeps a check point without rebuilds cheap and
  // still fills the chip when every scene rebuilds.
  auto grid = [](uint64_t items) { return dim3(uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(items, kBpWorkGrid)))); };
  const uint64_t nv = s.bp_nv, ns = s.bp_ns, ne = s.bp_ne, nc = s.bp_nc;
  if (ns + ne) hipLaunchKernelGGL(k_bp_bounds, grid(ns + ne), B, 0, st, s, p);
  if (ns + ne) hipLaunchKernelGGL(k_bp_super, grid(ns + ne), B, 0, st, s);
  hipLaunchKernelGGL(k_bp_vt<false>, grid(64 * nv), B, 0, st, s, p);
  if (ne) hipLaunchKernelGGL(k_bp_ee<false>, grid(64 * ne), B, 0, st, s, p);
  hipLaunchKernelGGL(k_bp_block_sums, grid(nv), B, 0, st, s, uint32_t(kBpScanVt));
  if (ne) hipLaunchKernelGGL(k_bp_block_sums, grid(ne), B, 0, st, s, uint32_t(kBpScanEe));
  hipLaunchKernelGGL(k_bp_scene_scan, scenes, B, 0, st, s, p, 0u);
  hipLaunchKernelGGL(k_bp_scan_apply, grid(nv), B, 0, st, s, uint32_t(kBpScanVt));
  if (ne) hipLaunchKernelGGL(k_bp_scan_apply, grid(ne), B, 0, st, s, uint32_t(kBpScanEe));
  if (carry && nc) hipLaunchKernelGGL(k_bp_save_prev, grid(nc), B, 0, st, s);
  hipLaunchKernelGGL(k_bp_vt<true>, grid(64 * nv), B, 0, st, s, p);
  if (ne) hipLaunchKernelGGL(k_bp_ee<true>, grid(64 * ne), B, 0, st, s, p);
  hipLaunchKernelGGL(k_bp_row_zero, grid(nv), B, 0, st, s);
  if (nc) hipLaunchKernelGGL(k_bp_rows<false>, grid(nc), B, 0, st, s, p, 0u);
  hipLaunchKernelGGL(k_bp_block_sums, grid(nv), B, 0, st, s, uint32_t(kBpScanRows));
  hipLaunchKernelGGL(k_bp_scene_scan, scenes, B, 0, st, s, p, 1u);
  hipLaunchKernelGGL(k_bp_scan_apply, grid(nv), B, 0, st, s, uint32_t(kBpScanRows));
  hipLaunchKernelGGL(k_bp_row_offsets, grid(nv), B, 0, st, s);
  if (nc) hipLaunchKernelGGL(k_bp_rows<true>, grid(nc), B, 0, st, s, p, cr);
  hipLaunchKernelGGL(k_bp_row_sort, grid(64 * nv), B, 0, st, s);
  hipLaunchKernelGGL(k_bp_row_sort_long, grid(kBpLongGrid), B, 0, st, s);
  hipLaunchKernelGGL(k_bp_commit, grid(nv + ne), B, 0, st, s, p);
}


// ------------------------------------------------------------------ exact diagnostics (FP64)
// The reference's opt-in oracle diagnostics (gpu/jgs2-diagnostics.ts:
// 712-1110), evaluated for one scene in double precision from the f32
// state: per tetrahedron the elastic energy and det F; per vertex the
// implicit-Euler gradient by component, the diagonal Hessian block and the
// inertia, external-force, target and floor energies; the scene summary
// (energy components, component gradient norms, force-balance residual,
// largest final update, minimum det F, floor metrics). Contact scenes add
// the barrier energy sum_c kappa b(d_c) over the scene's candidates (the
// reference's evaluator has no contact term); cloth energies are not part
// of it (nor of the reference's).
namespace diag {

struct d3 { double x, y, z; };
struct dm3 { d3 c0, c1, c2; };  // columns
__device__ __forceinline__ d3 dk(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ d3 dload(const float4* p, uint32_t v) { const float4 q = p[v]; return {q.x, q.y, q.z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ d3 mv(const dm3& m, d3 v) { return v.x * m.c0 + v.y * m.c1 + v.z * m.c2; }
__device__ __forceinline__ d3 mtv(const dm3& m, d3 v) { return {ddot(m.c0, v), ddot(m.c1, v), ddot(m.c2, v)}; }
__device__ __forceinline__ dm3 mm(const dm3& a, const dm3& b) { return {mv(a, b.c0), mv(a, b.c1), mv(a, b.c2)}; }
__device__ __forceinline__ double ddet(const dm3& m) { return ddot(m.c0, dcross(m.c1, m.c2)); }
__device__ __forceinline__ double frob(const dm3& a, const dm3& b) { return ddot(a.c0, b.c0) + ddot(a.c1, b.c1) + ddot(a.c2, b.c2); }
// d det / dF, columns f1 x f2, f2 x f0, f0 x f1
__device__ __forceinline__ dm3 dcof(const dm3& F) { return {dcross(F.c1, F.c2), dcross(F.c2, F.c0), dcross(F.c0, F.c1)}; }
// its directional derivative along dF
__device__ __forceinline__ dm3 dcof_dir(const dm3& F, const dm3& H) {
  return {dcross(H.c1, F.c2) + dcross(F.c1, H.c2), dcross(H.c2, F.c0) + dcross(F.c2, H.c0),
          dcross(H.c0, F.c1) + dcross(F.c0, H.c1)};
}
__device__ __forceinline__ dm3 dscale(double s, const dm3& m) { return {s * m.c0, s * m.c1, s * m.c2}; }
__device__ __forceinline__ dm3 dadd(const dm3& a, const dm3& b) { return {a.c0 + b.c0, a.c1 + b.c1, a.c2 + b.c2}; }
__device__ __forceinline__ bool dfinite(double v) { return isfinite(v); }

struct DTet {
  uint4 ix;
  dm3 inv;
  double vol, lam, mu;
};
__device__ __forceinline__ DTet dtet(const DeviceState& s, uint32_t t) {
  const Tet e = load_tet(s, t);
  return {e.ix, {{e.inv.c0.x, e.inv.c0.y, e.inv.c0.z}, {e.inv.c1.x, e.inv.c1.y, e.inv.c1.z}, {e.inv.c2.x, e.inv.c2.y, e.inv.c2.z}},
          e.vol, e.lam, e.mu};
}
__device__ __forceinline__ d3 dshape(const DTet& e, uint32_t slot) {  // rows of Dm^-1, and minus their sum
  const d3 g1 = dk(e.inv.c0.x, e.inv.c1.x, e.inv.c2.x), g2 = dk(e.inv.c0.y, e.inv.c1.y, e.inv.c2.y),
           g3 = dk(e.inv.c0.z, e.inv.c1.z, e.inv.c2.z);
  if (slot == 1) return g1;
  if (slot == 2) return g2;
  if (slot == 3) return g3;
  return dk(0, 0, 0) - g1 - g2 - g3;
}
__device__ __forceinline__ dm3 ddeform(const DTet& e, const float4* x) {
  const d3 p0 = dload(x, e.ix.x);
  const dm3 ds = {dload(x, e.ix.y) - p0, dload(x, e.ix.z) - p0, dload(x, e.ix.w) - p0};
  return mm(ds, e.inv);
}
// Stable Neo-Hookean energy density and stress with the record's adjusted
// lambda, mu (stable-neo-hookean-wgsl.ts:15-94).
__device__ __forceinline__ double snh_psi(const dm3& F, double lam, double mu) {
  const double I1 = frob(F, F), J = ddet(F);
  return 0.5 * mu * (I1 - 3.0) - 0.75 * mu * (J - 1.0) + 0.5 * lam * (J - 1.0) * (J - 1.0) -
         0.5 * mu * log((I1 + 1.0) / 4.0);
}
__device__ __forceinline__ dm3 snh_P(const dm3& F, double lam, double mu) {
  const double I1 = frob(F, F), J = ddet(F);
  return dadd(dscale(mu * (1.0 - 1.0 / (I1 + 1.0)), F), dscale(lam * (J - 1.0) - 0.75 * mu, dcof(F)));
}
__device__ __forceinline__ dm3 snh_dP(const dm3& F, const dm3& H, double lam, double mu) {
  const double I1 = frob(F, F), J = ddet(F), q = I1 + 1.0;
  const dm3 C = dcof(F);
  dm3 r = dscale(mu * (1.0 - 1.0 / q), H);
  r = dadd(r, dscale(2.0 * mu * frob(F, H) / (q * q), F));
  r = dadd(r, dscale(lam * frob(C, H), C));
  return dadd(r, dscale(lam * (J - 1.0) - 0.75 * mu, dcof_dir(F, H)));
}
__device__ __forceinline__ d3 dunit_or(d3 v, d3 fb, double eps) {
  const double sq = ddot(v, v);
  if (sq > eps * eps && sq < 1.0e60) return (1.0 / sqrt(sq)) * v;
  return fb;
}
__device__ __forceinline__ d3 dperp(d3 a, double eps) {
  const d3 c = fabs(a.x) > 0.8 ? dk(0, 1, 0) : dk(1, 0, 0);
  return dunit_or(dcross(a, c), dk(0, 0, 1), eps);
}
// polarRotation (S:940-978) in double: seven averaging steps with the
// inverse transpose, then the right-handed column orthonormalization.
__device__ dm3 dpolar(dm3 r, double eps) {
  for (int it = 0; it < 7; ++it) {
    const double d = ddet(r);
    if (!(fabs(d) > eps) || fabs(d) > 1.0e40) break;
    r = dscale(0.5, dadd(r, dscale(1.0 / d, dcof(r))));
  }
  const d3 c0 = dunit_or(r.c0, dk(1, 0, 0), eps);
  const d3 c1c = r.c1 - ddot(c0, r.c1) * c0;
  d3 c1 = dunit_or(c1c, dperp(c0, eps), eps);
  d3 c2 = dunit_or(dcross(c0, c1), dk(0, 0, 1), eps);
  if (ddot(c2, r.c2) < 0.0) {
    c1 = -1.0 * c1;
    c2 = -1.0 * c2;
  }
  c1 = dunit_or(dcross(c2, c0), dperp(c0, eps), eps);
  return {c0, c1, c2};
}
__device__ __forceinline__ dm3 dstiff(const DeviceState& s, uint32_t t, uint32_t rv, uint32_t cv) {
  const float* k = s.tet_k + size_t(t) * 144;
  const uint32_t r0 = rv * 3, c0 = cv * 3;
  return {dk(k[r0 * 12 + c0], k[(r0 + 1) * 12 + c0], k[(r0 + 2) * 12 + c0]),
          dk(k[r0 * 12 + c0 + 1], k[(r0 + 1) * 12 + c0 + 1], k[(r0 + 2) * 12 + c0 + 1]),
          dk(k[r0 * 12 + c0 + 2], k[(r0 + 1) * 12 + c0 + 2], k[(r0 + 2) * 12 + c0 + 2])};
}
// Co-rotated displacements u_l = R^T (x_l - x_mean) - (X_l - X_mean).
__device__ void corot_disp(const DeviceState& s, const DTet& e, const float4* x, const dm3& R, d3 u[4]) {
  const uint32_t id[4] = {e.ix.x, e.ix.y, e.ix.z, e.ix.w};
  d3 cur[4], rst[4], cc = dk(0, 0, 0), rc = dk(0, 0, 0);
  for (int l = 0; l < 4; ++l) {
    cur[l] = dload(x, id[l]);
    rst[l] = dload(s.vtx_rest, id[l]);
    cc = cc + cur[l];
    rc = rc + rst[l];
  }
  cc = 0.25 * cc;
  rc = 0.25 * rc;
  for (int l = 0; l < 4; ++l) u[l] = mtv(R, cur[l] - cc) - (rst[l] - rc);
}

constexpr int kVertexRecord = 36;  // doubles per vertex record
constexpr int kTetRecord = 4;      // energy, det F, finite, 0

}  // namespace diag

// One thread per tetrahedron of the scene: energy, det F and (co-rotated)
// the frame the vertex pass rotates stiffness blocks with.
__global__ void __launch_bounds__(kBlock) k_diag_tets(DeviceState s, DiagParams q) {
  using namespace diag;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= q.T) return;
  const uint32_t t = q.tb + i;
  const DTet e = dtet(s, t);
  const dm3 F = ddeform(e, q.x);
  const double J = ddet(F);
  double energy;
  bool ok = dfinite(J);
  if (q.corotated) {
    const dm3 R = dpolar(F, q.polar_eps);
    double* rr = q.trot + 9 * size_t(i);
    rr[0] = R.c0.x; rr[1] = R.c0.y; rr[2] = R.c0.z; rr[3] = R.c1.x; rr[4] = R.c1.y; rr[5] = R.c1.z;
    rr[6] = R.c2.x; rr[7] = R.c2.y; rr[8] = R.c2.z;
    d3 u[4];
    corot_disp(s, e, q.x, R, u);
    energy = 0.0;
    for (uint32_t r = 0; r < 4; ++r) {
      d3 g = dk(0, 0, 0);
      for (uint32_t c = 0; c < 4; ++c) g = g + mv(dstiff(s, t, r, c), u[c]);
      energy += 0.5 * ddot(u[r], g);
    }
  } else {
    energy = e.vol * snh_psi(F, e.lam, e.mu);
  }
  ok = ok && dfinite(energy);
  double* rec = q.trec + kTetRecord * size_t(i);
  rec[0] = ok ? energy : 0.0;
  rec[1] = dfinite(J) ? J : 0.0;
  rec[2] = ok ? 1.0 : 0.0;
  rec[3] = 0.0;
}

// One thread per vertex of the scene (evaluateDiagnosticVertices, :791-950).
__global__ void __launch_bounds__(kBlock) k_diag_vertices(DeviceState s, DiagParams q) {
  using namespace diag;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= q.V) return;
  const uint32_t v = q.vb + i;
  const uint4 inf = s.vtx_info[v];
  const d3 x = dload(q.x, v), xp = dload(q.xp, v);
  const bool active = inf.z == 0u;
  const float4 fu = s.final_update[v];
  d3 g = dk(0, 0, 0), gi = dk(0, 0, 0), ge = dk(0, 0, 0), gf = dk(0, 0, 0), gt = dk(0, 0, 0), gfl = dk(0, 0, 0);
  dm3 H = {dk(0, 0, 0), dk(0, 0, 0), dk(0, 0, 0)};
  double Ei = 0.0, Ef = 0.0, Et = 0.0, Efl = 0.0, w = 0.0, upd = 0.0, upd_ok = 0.0, floor_on = 0.0;
  const double sfd = x.y - q.floor_y;
  bool ok = dfinite(x.x) && dfinite(x.y) && dfinite(x.z) && dfinite(xp.x) && dfinite(xp.y) && dfinite(xp.z);
  if (active) {
    w = double(s.vtx_rest[v].w) * q.ih2;
    const d3 dx = x - xp;
    gi = w * dx;
    Ei = 0.5 * w * ddot(dx, dx);
    H = {dk(w, 0, 0), dk(0, w, 0), dk(0, 0, w)};
    upd = fu.x;
    upd_ok = (fu.w >= 0.5f && fu.x >= 0.0f) ? 1.0 : 0.0;
    if (q.objective_flags & kForce) {
      const d3 f = dload(s.objectives, 2 * v);
      gf = dk(0, 0, 0) - f;
      Ef = -ddot(f, x);
    }
    if (q.objective_flags & kTarget) {
      const float4 tr = s.objectives[2 * size_t(v) + 1];
      if (tr.w > 0.0f) {
        const d3 r = x - dk(tr.x, tr.y, tr.z);
        gt = double(tr.w) * r;
        Et = 0.5 * double(tr.w) * ddot(r, r);
        H.c0.x += tr.w;
        H.c1.y += tr.w;
        H.c2.z += tr.w;
      }
    }
    for (uint32_t k = 0; k < inf.y; ++k) {
      const uint32_t t = s.adjacency[inf.x + k];
      const DTet e = dtet(s, t);
      const uint32_t slot = local_slot(e.ix, v);
      if (slot >= 4u) {
        ok = false;
        continue;
      }
      const dm3 F = ddeform(e, q.x);
      if (q.corotated) {
        const double* rr = q.trot + 9 * size_t(t - q.tb);
        const dm3 R = {dk(rr[0], rr[1], rr[2]), dk(rr[3], rr[4], rr[5]), dk(rr[6], rr[7], rr[8])};
        d3 u[4];
        corot_disp(s, e, q.x, R, u);
        d3 lg = dk(0, 0, 0);
        for (uint32_t c = 0; c < 4; ++c) lg = lg + mv(dstiff(s, t, slot, c), u[c]);
        ge = ge + mv(R, lg);
        const dm3 Kss = dstiff(s, t, slot, slot);
        const dm3 RK = mm(R, Kss);
        // R K_ss R^T
        const dm3 Rt = {dk(R.c0.x, R.c1.x, R.c2.x), dk(R.c0.y, R.c1.y, R.c2.y), dk(R.c0.z, R.c1.z, R.c2.z)};
        H = dadd(H, mm(RK, Rt));
      } else {
        const d3 gs = dshape(e, slot);
        ge = ge + e.vol * mv(snh_P(F, e.lam, e.mu), gs);
        // column c of the diagonal block: vol dP[e_c gs^T] gs
        for (int c = 0; c < 3; ++c) {
          dm3 dF = {dk(0, 0, 0), dk(0, 0, 0), dk(0, 0, 0)};
          // row c of dF is gs: dF.cj[c] = gs_j
          (c == 0 ? dF.c0.x : c == 1 ? dF.c0.y : dF.c0.z) = gs.x;
          (c == 0 ? dF.c1.x : c == 1 ? dF.c1.y : dF.c1.z) = gs.y;
          (c == 0 ? dF.c2.x : c == 1 ? dF.c2.y : dF.c2.z) = gs.z;
          const d3 col = e.vol * mv(snh_dP(F, dF, e.lam, e.mu), gs);
          if (c == 0) H.c0 = H.c0 + col;
          else if (c == 1) H.c1 = H.c1 + col;
          else H.c2 = H.c2 + col;
        }
      }
    }
    if (q.floor_k > 0.0 && sfd < 0.0) {
      floor_on = 1.0;
      Efl = 0.5 * q.floor_k * sfd * sfd;
      gfl = dk(0, q.floor_k * sfd, 0);
      H.c1.y += q.floor_k;
    }
    g = gi + ge + gf + gt + gfl;
  }
  double* r = q.vrec + kVertexRecord * size_t(i);
  const double gg = ddot(g, g);
  ok = ok && dfinite(gg) && dfinite(Ei) && dfinite(Ef) && dfinite(Et) && dfinite(Efl) && dfinite(H.c0.x + H.c1.y + H.c2.z);
  r[0] = g.x; r[1] = g.y; r[2] = g.z; r[3] = active ? 1.0 : 0.0;
  // row-major Hessian block
  r[4] = H.c0.x; r[5] = H.c1.x; r[6] = H.c2.x;
  r[7] = H.c0.y; r[8] = H.c1.y; r[9] = H.c2.y;
  r[10] = H.c0.z; r[11] = H.c1.z; r[12] = H.c2.z;
  r[13] = Ei; r[14] = Efl; r[15] = upd; r[16] = gg; r[17] = ok ? 1.0 : 0.0; r[18] = w; r[19] = upd_ok;
  r[20] = ddot(gi, gi); r[21] = ddot(ge, ge); r[22] = ddot(gfl, gfl); r[23] = sfd; r[24] = floor_on;
  r[25] = gf.x; r[26] = gf.y; r[27] = gf.z; r[28] = Ef;
  r[29] = gt.x; r[30] = gt.y; r[31] = gt.z; r[32] = Et;
  r[33] = ddot(gf, gf); r[34] = ddot(gt, gt); r[35] = 0.0;
}

// One thread per candidate slot of the scene: distance (f32 geometry, as the
// step evaluates it) and barrier energy kappa b(d) in double.
__global__ void __launch_bounds__(kBlock) k_diag_contact(DeviceState s, DiagParams q) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= q.C) return;
  const uint32_t c = q.cb + i;
  double* r = q.crec + 2 * size_t(i);
  r[0] = 0.0;
  r[1] = -1.0;  // not a live candidate
  if (c >= s.scene_end[q.scene].w) return;
  const Contact k = ipc_contact_at(s, c, q.x, q.dhat);
  if (!k.valid || !finite_s(k.distance)) {
    r[1] = -2.0;
    return;
  }
  const double d = k.distance, dh = q.dhat;
  r[1] = d;
  if (d < dh && d > 0.0) {
    const double o = d - dh;
    r[0] = q.kappa * (-(o * o) * log(d / dh));
  }
}

// Scene summary, one block: fixed-order strided sums and an LDS tree.
// summary[]: 0 energy, 1 inertia, 2 elasticity, 3 external force, 4 target,
// 5 floor, 6 contact barrier, 7-11 squared gradient norms (inertia,
// material, external force, target, floor), 12 |g|^2, 13 max update,
// 14 update valid count, 15 min det F, 16 min signed floor distance,
// 17 active floor contacts, 18 active vertices, 19 nonfinite records,
// 20 min contact distance, 21 active contact pairs, 22 invalid contacts.
__global__ void __launch_bounds__(kBlock) k_diag_reduce(DiagParams q) {
  using namespace diag;
  __shared__ double red[kBlock];
  double acc[23];
  for (int k = 0; k < 23; ++k) acc[k] = 0.0;
  acc[13] = -INFINITY;
  acc[15] = INFINITY;
  acc[16] = INFINITY;
  acc[20] = INFINITY;
  for (uint32_t i = threadIdx.x; i < q.V; i += kBlock) {
    const double* r = q.vrec + kVertexRecord * size_t(i);
    acc[19] += r[17] != 0.0 ? 0.0 : 1.0;
    if (r[3] == 0.0) continue;
    acc[1] += r[13];
    acc[3] += r[28];
    acc[4] += r[32];
    acc[5] += r[14];
    acc[7] += r[20];
    acc[8] += r[21];
    acc[9] += r[33];
    acc[10] += r[34];
    acc[11] += r[22];
    acc[12] += r[16];
    if (r[19] != 0.0) {
      acc[13] = fmax(acc[13], r[15]);
      acc[14] += 1.0;
    }
    acc[16] = fmin(acc[16], r[23]);
    acc[17] += r[24];
    acc[18] += 1.0;
  }
  for (uint32_t i = threadIdx.x; i < q.T; i += kBlock) {
    const double* r = q.trec + kTetRecord * size_t(i);
    acc[2] += r[0];
    acc[15] = fmin(acc[15], r[1]);
    acc[19] += r[2] != 0.0 ? 0.0 : 1.0;
  }
  for (uint32_t i = threadIdx.x; i < q.C; i += kBlock) {
    const double* r = q.crec + 2 * size_t(i);
    if (r[1] == -1.0) continue;
    if (r[1] == -2.0) {
      acc[22] += 1.0;
      continue;
    }
    acc[6] += r[0];
    acc[20] = fmin(acc[20], r[1]);
    acc[21] += r[0] > 0.0 ? 1.0 : 0.0;
  }
  for (int k = 1; k < 23; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (uint32_t stride = kBlock / 2; stride > 0; stride /= 2) {
      if (threadIdx.x < stride) {
        const double a = red[threadIdx.x], b = red[threadIdx.x + stride];
        red[threadIdx.x] = k == 13 ? fmax(a, b) : (k == 15 || k == 16 || k == 20) ? fmin(a, b) : a + b;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) q.summary[k] = red[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* m = q.summary;
    m[0] = m[1] + m[2] + m[3] + m[4] + m[5] + m[6];
  }
}

void launch_diagnostics(const DeviceState& s, const DiagParams& q, hipStream_t st) {
  if (q.T) hipLaunchKernelGGL(k_diag_tets, dim3((q.T + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, q);
  hipLaunchKernelGGL(k_diag_vertices, dim3((q.V + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, q);
  if (q.C) hipLaunchKernelGGL(k_diag_contact, dim3((q.C + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, q);
  hipLaunchKernelGGL(k_diag_reduce, dim3(1), dim3(kBlock), 0, st, q);
}

}  // namespace jgs2

// ------------------------------------------------------------------ fixed diagnostics buffer
// Working buffer of the allocation-free diagnostics entry point: the per-vertex,
// per-tetrahedron and per-candidate records and the summary of one evaluation
// (every word is written before it is read, so nothing needs clearing). Each
// device that loads the library holds its own copy until the process ends.
static __device__ __align__(256) double scratch[jgs2::kDiagScratchDoubles];

namespace jgs2 {
namespace {
constexpr int kScratchDevices = 64;
std::mutex scratch_mutex;
double* scratch_address[kScratchDevices] = {};
}  // namespace

// The array's name is only a host-side handle: kernels get the device
// address, looked up once per device with that device current.
hipError_t diagnostics_scratch(int device, double** ptr) {
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[device]) {
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e == hipSuccess && current != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (current >= 0 && current != device) (void)hipSetDevice(current);
    if (e != hipSuccess) return e;
    scratch_address[device] = static_cast<double*>(p);
  }
  *ptr = scratch_address[device];
  return hipSuccess;
}

}  // namespace jgs2
