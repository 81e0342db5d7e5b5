// Shared device helpers for the LEX dynamics kernels (gfx950, wave64).
//
// Layout (the reference's padded C-order, z contiguous; ghost widths 3, 3, 1):
//   scalar  (nx+6, ny+6, nz+2)   cell (i,j,k) at I=i+3, J=j+3, K=k+1
//   u       (nx+7, ny+6, nz+2)   face I lies between scalar cells I-1 and I
//   v       (nx+6, ny+7, nz+2)   face J lies between scalar cells J-1 and J
//   w       (nx+6, ny+6, nz+3)   level K lies between scalar levels K-1 and K
// Base-state profiles are horizontally uniform; kernels take them as z-columns of
// length nz+2 (setup_lex.py:76-137 builds them that way).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lex {

// Physical constants: namelist_n_constants.py:64-75 (rounded to FP32 the way JAX
// rounds weak Python scalars).
constexpr double kRd = 287.04;
constexpr double kRv = 461.5;
constexpr double kCp = 1004.64;
constexpr double kCv = kCp - kRd;
constexpr double kReps = kRv / kRd;
constexpr double kRepsm1 = kRv / kRd - 1.0;
constexpr double kG = 9.80616;
constexpr double kP00 = 100000.0;

struct Dims {
  int nx, ny, nz;
  float dx, dy, dz;
  float idx, idy, idz;  // reciprocals
  // strides (in elements) for each staggering; z stride is 1
  __host__ __device__ long sxc() const { return (long)(ny + 6) * (nz + 2); }
  __host__ __device__ long syc() const { return nz + 2; }
  __host__ __device__ long sxu() const { return (long)(ny + 6) * (nz + 2); }
  __host__ __device__ long sxv() const { return (long)(ny + 7) * (nz + 2); }
  __host__ __device__ long sxw() const { return (long)(ny + 6) * (nz + 3); }
  __host__ __device__ long syw() const { return nz + 3; }
  __host__ __device__ long nc() const { return (long)(nx + 6) * (ny + 6) * (nz + 2); }
  __host__ __device__ long nu() const { return (long)(nx + 7) * (ny + 6) * (nz + 2); }
  __host__ __device__ long nv() const { return (long)(nx + 6) * (ny + 7) * (nz + 2); }
  __host__ __device__ long nw() const { return (long)(nx + 6) * (ny + 6) * (nz + 3); }
};

// Element index of padded (I, J, K) in each staggering.  The arithmetic is 32-bit: every
// grid accepted at the C ABI has fewer than 2^30 elements in each padded field (see
// kMaxOffset32), so the products and sums fit an int, and an index that a caller takes
// into the halo and offsets back stays exact when negative (sign extension).
__device__ __forceinline__ long ic(const Dims& d, int I, int J, int K) {
  return (long)(I * ((d.ny + 6) * (d.nz + 2)) + J * (d.nz + 2) + K);
}
__device__ __forceinline__ long iu(const Dims& d, int I, int J, int K) {
  return (long)(I * ((d.ny + 6) * (d.nz + 2)) + J * (d.nz + 2) + K);
}
__device__ __forceinline__ long iv(const Dims& d, int I, int J, int K) {
  return (long)(I * ((d.ny + 7) * (d.nz + 2)) + J * (d.nz + 2) + K);
}
__device__ __forceinline__ long iw(const Dims& d, int I, int J, int K) {
  return (long)(I * ((d.ny + 6) * (d.nz + 3)) + J * (d.nz + 3) + K);
}

// Access at a 32-bit element offset.  The address is the (wave-uniform) base pointer
// plus a zero-extended 32-bit byte offset, which the compiler emits as
// `global_load/store ... v_off, s[base]` with no per-lane 64-bit address arithmetic.
// Callers keep offsets below 2^30 elements (every field is checked at the C ABI).
__device__ __forceinline__ float ld32(const float* __restrict__ base, uint32_t e) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (size_t)(e * 4u));
}
__device__ __forceinline__ void st32(float* __restrict__ base, uint32_t e, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (size_t)(e * 4u)) = v;
}
constexpr long kMaxOffset32 = 1L << 30;   // elements addressable by ld32 / st32
// An opaque copy of a loop-invariant 32-bit lane offset, taken inside the loop: without
// it the compiler hoists the zero-extended byte offset out of the loop as a 64-bit
// register pair and then adds it to each uniform base pointer per access (one 64-bit
// vector add per load); with it every access is `global_load/store v_off, s[base]`.
// The empty asm statement emits no instruction.
__device__ __forceinline__ uint32_t lane_off(uint32_t x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Fast reciprocal (v_rcp_f32, 1 ulp).  Used for the WENO weight ratios and the
// normalisation; results stay within the FP32-vs-FP64 tolerance of the reference.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// WENO5 face value from the upwind-ordered stencil (a, b, c | d, e), c upwind.
// Jiang & Shu (1996) candidates/indicators, Borges et al. (2008) weights with the
// squared ratio, as advection.py:311-326.  The reference caps the ratio at 1e30
// (fmin) before squaring; in FP32 that cap cannot change a result: any ratio above
// ~1.8e19 squares to +inf with or without it, and a ratio below 1e30 passes unchanged
// (tau and beta + eps are finite and beta + eps > 0, so the ratio is never NaN for
// finite data).  The FP32 kernels therefore omit it.
//
// Evaluated as c + sum_k a_k (q_k - c) / sum_k a_k, the same function as
// sum_k a_k q_k / sum_k a_k, in the first differences E = c - b, D = d - c and the
// second differences t1, t3, t5 centred at b, c, d, which every indicator and
// candidate shares:
//   q1 - c = t1/3 + E/2,  q2 - c = (D - (b - d))/6,  q3 - c = D/2 - t5/6,
//   a - 4b + 3c = t1 + 2E,  3c - 4d + e = t5 - 2D,
// with the indicators scaled by 4 (B_k = 4 beta_k, eps' = 4 eps, so tau/(beta_k + eps)
// is unchanged) and the linear weights folded into the candidates.  47 operations
// instead of 56, and the correction to c is formed without the cancellation of
// sum_k a_k q_k for a field far from zero (theta ~ 300 K).
__device__ __forceinline__ float weno5(float a, float b, float c, float d, float e, float weps) {
  const float k13 = (float)(13.0 / 3.0);   // 4 x 13/12
  const float E = c - b, D = d - c;
  const float t1 = a - 2.0f * b + c;
  const float t3 = b - 2.0f * c + d;
  const float t5 = c - 2.0f * d + e;
  const float t2 = t1 + 2.0f * E;          // a - 4b + 3c
  const float t4 = b - d;
  const float t6 = t5 - 2.0f * D;          // 3c - 4d + e
  const float B1 = (k13 * t1) * t1 + t2 * t2;
  const float B2 = (k13 * t3) * t3 + t4 * t4;
  const float B3 = (k13 * t5) * t5 + t6 * t6;
  const float eps4 = 4.0f * weps;
  const float tau = fabsf(B1 - B3);
  const float r1 = tau * frcp(B1 + eps4);
  const float r2 = tau * frcp(B2 + eps4);
  const float r3 = tau * frcp(B3 + eps4);
  const float e1 = 1.0f + r1 * r1, e2 = 1.0f + r2 * r2, e3 = 1.0f + r3 * r3;
  // d_k (q_k - c) with d = (0.1, 0.6, 0.3)
  const float g1 = (float)(0.1 / 3.0) * t1 + 0.05f * E;
  const float g2 = 0.1f * (D - t4);
  const float g3 = 0.15f * D - 0.05f * t5;
  const float num = e1 * g1 + e2 * g2 + e3 * g3;
  const float den = 0.1f * e1 + 0.6f * e2 + 0.3f * e3;
  return c + num * frcp(den);
}

// WENO3 face value from (a, b | c), b upwind (advection.py:329-340); the 1e30 cap is
// omitted for the reason given at weno5.  Weights (1/3, 2/3) (1 + (tau/(beta_k+eps))^2)
// on the candidates -a/2 + 3b/2 and (b + c)/2, evaluated as b + (correction) in the
// differences P = b - a, Q = c - b like weno5.
__device__ __forceinline__ float weno3(float a, float b, float c, float weps) {
  const float P = b - a, Q = c - b;
  const float b1 = P * P, b2 = Q * Q;
  const float tau = fabsf(b1 - b2);
  const float r1 = tau * frcp(b1 + weps);
  const float r2 = tau * frcp(b2 + weps);
  const float e1 = 1.0f + r1 * r1, e2 = 1.0f + r2 * r2;
  // (1/3) e1 (q1 - b) + (2/3) e2 (q2 - b) over (1/3) e1 + (2/3) e2, times 3 above and below
  return b + (e1 * (0.5f * P) + e2 * Q) * frcp(e1 + 2.0f * e2);
}


// Horizontal reconstruction at the flux point between q[m-1] and q[m] from the six
// values q[m-3..m+2] (advection.py:425-470): upwind side chosen by the mass-flux sign.
__device__ __forceinline__ float hrecon(float qm3, float qm2, float qm1, float q0, float qp1, float qp2,
                                        bool pos, float weps) {
  float a = pos ? qm3 : qp2;
  float b = pos ? qm2 : qp1;
  float c = pos ? qm1 : q0;
  float d = pos ? q0 : qm1;
  float e = pos ? qp1 : qm2;
  return weno5(a, b, c, d, e, weps);
}

// Vertical reconstruction at face f of a padded column p[0..L-1] (face f lies between
// p[f] and p[f+1]).  The table of advection.py:330-409 (scalars) and its u/v/w copies:
//   upward:   f=1 centred, f=2 WENO3, f=3..L-4 WENO5, f=L-3 WENO3
//   downward: f=1 WENO3, f=2..L-5 WENO5, f=L-4 WENO3, f=L-3 centred
// Faces 0 and L-2 carry zero (boundary-condition slots).  `p` has unit stride.
//
// Branch-free in the upwind side: the lanes of a wave run along z, and the sign of w
// changes from level to level, so an if/else on the sign would make every wave execute
// both WENO5 branches.  Instead the upwind-ordered stencil is selected from the six
// values p[f-2..f+3] (index clamped so the loads stay in the column) and one WENO5 is
// evaluated; the table's near-boundary rows (WENO3, centred, zero) are patched in a
// branch that only waves holding a face with f <= 2 or f >= L-4 enter.  Every face gets
// exactly the operands of the table above, so the result is the table's.  Needs L >= 8.
__device__ __forceinline__ float vrecon(const float* __restrict__ p, int L, int f, bool pos, float weps) {
  const int fc = min(max(f, 2), L - 4);
  const float q0 = p[fc - 2], q1 = p[fc - 1], q2 = p[fc], q3 = p[fc + 1], q4 = p[fc + 2], q5 = p[fc + 3];
  float r = weno5(pos ? q0 : q5, pos ? q1 : q4, pos ? q2 : q3, pos ? q3 : q2, pos ? q4 : q1, weps);
  if (f <= 2 || f >= L - 4) {
    // 1 <= f <= L-3 here unless the face is a zero slot: p[f-1..f+2] are in the column
    const int fe = min(max(f, 1), L - 3);
    const float a = p[fe - 1], b = p[fe], c = p[fe + 1], e = p[fe + 2];
    // upward rows f = 2 and L-3 take WENO3 of (p[f-1], p[f], p[f+1]); downward rows
    // f = 1 and L-4 take WENO3 of (p[f+2], p[f+1], p[f]); the centred rows (upward f = 1,
    // downward f = L-3) take 0.5 (p[f] + p[f+1])
    const float r3 = weno3(pos ? a : e, pos ? b : c, pos ? c : b, weps);
    const bool w3 = pos ? (f == 2 || f == L - 3) : (f == 1 || f == L - 4);
    const bool cen = pos ? (f == 1) : (f == L - 3);
    r = w3 ? r3 : (cen ? 0.5f * (b + c) : r);
    if (f <= 0 || f >= L - 2) r = 0.0f;
  }
  return r;
}

// z level of lane tz in a block of nzl = 64*ceil(nz/64) lanes along z.  When a column
// spans two or more waves, wave 0 holds the 32 lowest and the 32 highest levels (the
// rows of the vertical face table that need WENO3 / centred values live there) and the
// later waves the levels in between, so only wave 0 enters vrecon's near-boundary
// branch.  Lanes of later waves with no level left get a level >= nz (idle).
__device__ __forceinline__ int lane_level(int tz, int nzl, int nz) {
  if (nzl <= 64 || tz < 32) return tz;
  if (tz < 64) return nz - 64 + tz;
  const int k = tz - 32;
  return k < nz - 32 ? k : nz + (tz - 64);
}

// Coriolis force of a stage (pressure_gradient_coriolis.py:80-113): mode 0 off, 1
// calculate_coriolis_force (f times the four-point mean), 2 ..._with_lspgrad (a
// large-scale pressure gradient in balance with the geostrophic wind (ug, vg) is
// subtracted: f (mean - geostrophic component)).  `s4` is the four-point sum in the
// reference's order; the arithmetic follows the reference's expressions.
struct CorOpt {
  int mode;
  float f, ug, vg;
  __device__ __forceinline__ float of(float s4, float geo) const {
    return mode == 2 ? f * (0.25f * s4 - geo) : (f * 0.25f) * s4;
  }
};

// select(mass >= 0, mass*up, mass*dn) with only the upwind side evaluated.
__device__ __forceinline__ bool upwind_pos(float mass) { return mass >= 0.0f; }

// ----------------------------------------------------------------------------------
// Stores that also fill the periodic / copy ghosts (padding3*_array,
// one_step_integration.py:950-985).  Each ghost has exactly one interior source.
// ----------------------------------------------------------------------------------
__device__ __forceinline__ void ghost_list_x(int i, int n, int* out, int& cnt) {
  // scalar-type axis: interior i in [0,n); padded P = i+3; images at i+3+n (i<3) and i+3-n (i>=n-3)
  cnt = 0;
  out[cnt++] = i + 3;
  if (i < 3) out[cnt++] = i + 3 + n;
  if (i >= n - 3) out[cnt++] = i + 3 - n;
}

__device__ __forceinline__ void ghost_list_face(int f, int n, int* out, int& cnt) {
  // face-type axis (u along x, v along y): face f in [0,n]; padded f+3; ghosts skip
  // the duplicate face: padded 0..2 <- faces n-3..n-1, padded n+4..n+6 <- faces 1..3
  cnt = 0;
  out[cnt++] = f + 3;
  if (f >= n - 3 && f <= n - 1) out[cnt++] = f + 3 - n;
  if (f >= 1 && f <= 3) out[cnt++] = f + 3 + n;
}

// scalar field: interior (i,j,k)
__device__ __forceinline__ void store_c(const Dims& d, float* __restrict__ a, int i, int j, int k, float val) {
  int xs[3], ys[3], zs[2];
  int nxs, nys, nzs = 1;
  ghost_list_x(i, d.nx, xs, nxs);
  ghost_list_x(j, d.ny, ys, nys);
  zs[0] = k + 1;
  if (k == 0) zs[nzs++] = 0;
  if (k == d.nz - 1) zs[nzs++] = d.nz + 1;
  for (int a1 = 0; a1 < nxs; ++a1)
    for (int b1 = 0; b1 < nys; ++b1)
      for (int c1 = 0; c1 < nzs; ++c1) a[ic(d, xs[a1], ys[b1], zs[c1])] = val;
}

// u field: face f in [0,nx], (j,k) interior
__device__ __forceinline__ void store_u(const Dims& d, float* __restrict__ a, int f, int j, int k, float val) {
  int xs[3], ys[3], zs[2];
  int nxs, nys, nzs = 1;
  ghost_list_face(f, d.nx, xs, nxs);
  ghost_list_x(j, d.ny, ys, nys);
  zs[0] = k + 1;
  if (k == 0) zs[nzs++] = 0;
  if (k == d.nz - 1) zs[nzs++] = d.nz + 1;
  for (int a1 = 0; a1 < nxs; ++a1)
    for (int b1 = 0; b1 < nys; ++b1)
      for (int c1 = 0; c1 < nzs; ++c1) a[iu(d, xs[a1], ys[b1], zs[c1])] = val;
}

// v field: face f in [0,ny], (i,k) interior
__device__ __forceinline__ void store_v(const Dims& d, float* __restrict__ a, int i, int f, int k, float val) {
  int xs[3], ys[3], zs[2];
  int nxs, nys, nzs = 1;
  ghost_list_x(i, d.nx, xs, nxs);
  ghost_list_face(f, d.ny, ys, nys);
  zs[0] = k + 1;
  if (k == 0) zs[nzs++] = 0;
  if (k == d.nz - 1) zs[nzs++] = d.nz + 1;
  for (int a1 = 0; a1 < nxs; ++a1)
    for (int b1 = 0; b1 < nys; ++b1)
      for (int c1 = 0; c1 < nzs; ++c1) a[iv(d, xs[a1], ys[b1], zs[c1])] = val;
}

// w field: interior level l in [0,nz]; level l at K=l+1; z ghosts copy the nearest
// level (padding3_array on an (nx,ny,nz+1) array).  set_w_bc zeroing is the caller's.
__device__ __forceinline__ void store_w(const Dims& d, float* __restrict__ a, int i, int j, int l, float val) {
  int xs[3], ys[3], zs[2];
  int nxs, nys, nzs = 1;
  ghost_list_x(i, d.nx, xs, nxs);
  ghost_list_x(j, d.ny, ys, nys);
  zs[0] = l + 1;
  if (l == 0) zs[nzs++] = 0;
  if (l == d.nz) zs[nzs++] = d.nz + 2;
  for (int a1 = 0; a1 < nxs; ++a1)
    for (int b1 = 0; b1 < nys; ++b1)
      for (int c1 = 0; c1 < nzs; ++c1) a[iw(d, xs[a1], ys[b1], zs[c1])] = val;
}

// Interior index of a cell one period away at most (periodic halo cells and planes).
__device__ __forceinline__ int wrap_cell(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// a[idx] when ok, else 0, without a branch: the load always issues (at index 0 when
// !ok), so it is batched with the thread's other loads instead of waiting behind a branch.
__device__ __forceinline__ float ld_or0(const float* __restrict__ a, long idx, bool ok) {
  const float v = a[ok ? idx : 0];
  return ok ? v : 0.f;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2: hardware
// block h runs on XCD h % 8.  xcd_tile renumbers blocks so that each XCD receives one
// contiguous range of the logical grid (x fastest, then y, then z), keeping the halo
// lines shared by neighbouring tiles in one L2.  Bijective: the first 8*(n/8) blocks are
// permuted, the remainder keep their numbers.
struct TileId {
  int x, y, z;
};
__device__ __forceinline__ TileId xcd_tile() {
  const unsigned gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
  const unsigned h = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned per = n >> 3, full = per << 3;
  const unsigned lid = h < full ? (h & 7u) * per + (h >> 3) : h;
  const unsigned gxy = gx * gy;
  TileId t;
  t.z = (int)(lid / gxy);
  const unsigned r = lid - (unsigned)t.z * gxy;
  t.y = (int)(r / gx);
  t.x = (int)(r - (unsigned)t.y * gx);
  return t;
}

// Fast paths: away from the periodic x/y images a cell has its own position plus, at the
// first and last level, one z ghost.  The x/y test is uniform across a wave whose lanes
// run along z, so only boundary columns take the general ghost lists above.
__device__ __forceinline__ bool near_xy(const Dims& d, int i, int j) {
  return i < 3 || i >= d.nx - 3 || j < 3 || j >= d.ny - 3;
}
__device__ __forceinline__ void store_c_fast(const Dims& d, float* __restrict__ a, int i, int j, int k, float val) {
  if (near_xy(d, i, j)) {
    store_c(d, a, i, j, k, val);
    return;
  }
  const long b = ic(d, i + 3, j + 3, 0);
  a[b + k + 1] = val;
  if (k == 0) a[b] = val;
  if (k == d.nz - 1) a[b + d.nz + 1] = val;
}
// store_c_fast with the cell's 32-bit element offset oc = ic(d, i+3, j+3, k+1) known
__device__ __forceinline__ void store_c_fast32(const Dims& d, float* __restrict__ a, int i, int j, int k, uint32_t oc,
                                               float val) {
  if (near_xy(d, i, j)) {
    store_c(d, a, i, j, k, val);
    return;
  }
  st32(a, oc, val);
  if (k == 0) st32(a, oc - 1, val);
  if (k == d.nz - 1) st32(a, oc + 1, val);
}
__device__ __forceinline__ void store_u_fast(const Dims& d, float* __restrict__ a, int f, int j, int k, float val) {
  if (f < 4 || f >= d.nx - 3 || j < 3 || j >= d.ny - 3) {
    store_u(d, a, f, j, k, val);
    return;
  }
  const long b = iu(d, f + 3, j + 3, 0);
  a[b + k + 1] = val;
  if (k == 0) a[b] = val;
  if (k == d.nz - 1) a[b + d.nz + 1] = val;
}
__device__ __forceinline__ void store_v_fast(const Dims& d, float* __restrict__ a, int i, int f, int k, float val) {
  if (i < 3 || i >= d.nx - 3 || f < 4 || f >= d.ny - 3) {
    store_v(d, a, i, f, k, val);
    return;
  }
  const long b = iv(d, i + 3, f + 3, 0);
  a[b + k + 1] = val;
  if (k == 0) a[b] = val;
  if (k == d.nz - 1) a[b + d.nz + 1] = val;
}
__device__ __forceinline__ void store_w_fast(const Dims& d, float* __restrict__ a, int i, int j, int l, float val) {
  if (near_xy(d, i, j)) {
    store_w(d, a, i, j, l, val);
    return;
  }
  const long b = iw(d, i + 3, j + 3, 0);
  a[b + l + 1] = val;
  if (l == 0) a[b] = val;
  if (l == d.nz) a[b + d.nz + 2] = val;
}

}  // namespace lex
