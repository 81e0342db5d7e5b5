// Hand-written adjoint (vector-Jacobian products) of the compressible WS-RK3 step.
//
// Every forward kernel of the per-operator path (lex_ops.hip, lex_smag.hip) has a reverse
// kernel here, written without atomics so that every result is bit-reproducible: a
// thread either owns the adjoint it updates (gather form: it adds, in a fixed order, the
// terms of every forward point that read its input) or hands the contributions other
// threads share to slot planes that a gather kernel adds in a fixed order.  All reverse
// kernels ACCUMULATE into their adjoint outputs.  Ghost images written by the forward
// stores (padding3*_array) are reversed by k_fold: every ghost's adjoint is added to its
// interior source and the ghost cleared.
//
// Differentiation conventions are JAX's on the reference code: lax.select passes the
// gradient to the selected branch only; d|x|/dx = sign(x) (0 at 0); fmin(1e30, r) passes
// it to r only when r < 1e30; where(s2n2 > 0, s2n2, 0) blocks it where false.
#pragma once
#include "lex_common.hpp"

namespace lex {

// Slot planes: the deterministic form of a scatter.  A reverse kernel whose threads share
// targets writes each shared contribution into the plane of its slot (slot = which of
// the thread's targets it is), at the thread's own index in the launch box, and writes
// every slot of every thread in the box (zero where it has nothing), so no plane holds
// data of an earlier launch.  A gather kernel then adds, for every target point, the
// slots that reach it in one fixed order: no atomics, bit-reproducible results.
struct SlotPlanes {
  float* p;          // plane s starts at p + s * plane; null: this kernel contributes nothing
  long plane;        // elements per plane
  int gx, gy, gz;    // launch box: a in [0, gx), b in [0, gy), k in [0, gz); k fastest
  int zw;
  // zw = 1: the column stencil's contributions (slots 12..17 of the scalar, u and v reverses)
  // stay in the wave: the column is one 64-lane wave (gz <= 64, blocks of 64 lanes), whose
  // lanes add the slots that reach their own level with shuffles in slot order and update
  // the adjoint directly (one writer per level); the gathers skip those slots
  __device__ __forceinline__ long at(int a, int b, int k) const { return ((long)a * gy + b) * gz + k; }
  __device__ __forceinline__ void put(int s, long i, float v) const { p[s * plane + i] = v; }
  // slot s of box point (a, b, k), zero outside the box
  __device__ __forceinline__ float get(int s, int a, int b, int k) const {
    const bool ok = (unsigned)a < (unsigned)gx && (unsigned)b < (unsigned)gy && (unsigned)k < (unsigned)gz;
    return (ok && p) ? p[s * plane + at(a, b, k)] : 0.f;
  }
};

// sum over t of slot t (o[t]) of the lane k + 3 - t, for the lanes (levels) of [0, nz):
// the column-stencil contributions that reach this lane's level, in slot order
__device__ __forceinline__ float zw_sum(const float* o, int k, int nz) {
  float zs = 0.f;
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const int src = k + 3 - t;
    const float v = __shfl(o[t], src & 63, 64);
    if (src >= 0 && src < nz) zs += v;
  }
  return zs;
}

inline SlotPlanes slot_planes(float* p, int gx, int gy, int gz) {
  SlotPlanes sp;
  sp.p = p;
  sp.gx = gx; sp.gy = gy; sp.gz = gz;
  sp.plane = (((long)gx * gy * gz) + 63) / 64 * 64;
  sp.zw = 0;
  return sp;
}

// launch boxes and slot counts of the advection reverses
constexpr int kSlotsScalar = 18, kSlotsMom = 24;
inline SlotPlanes with_zw(SlotPlanes sp) {
  sp.zw = sp.gz <= 64 ? 1 : 0;
  return sp;
}
inline SlotPlanes slots_scalar(const Dims& d, float* p) { return with_zw(slot_planes(p, d.nx + 1, d.ny + 1, d.nz)); }
inline SlotPlanes slots_u(const Dims& d, float* p) { return with_zw(slot_planes(p, d.nx + 2, d.ny + 1, d.nz)); }
inline SlotPlanes slots_v(const Dims& d, float* p) { return with_zw(slot_planes(p, d.nx + 1, d.ny + 2, d.nz)); }
inline SlotPlanes slots_w(const Dims& d, float* p) { return slot_planes(p, d.nx + 1, d.ny + 1, d.nz + 1); }
// bytes of the scalar reverse's planes, and of the three momentum reverses' planes together
inline size_t slot_bytes_scalar(const Dims& d) { return (size_t)kSlotsScalar * slots_scalar(d, nullptr).plane * 4; }
inline size_t slot_bytes_mom(const Dims& d) {
  return (size_t)kSlotsMom * (slots_u(d, nullptr).plane + slots_v(d, nullptr).plane + slots_w(d, nullptr).plane) * 4;
}

// ---------------------------------------------------------------- WENO gradients
// R = weno5(a..e) and g[0..4] = dR/d(a, b, c, d, e).  With r_k = tau/(beta_k + eps),
// w_k = d_k (1 + r_k^2), R = sum w_k q_k / W:
//   dR/dr_k = (q_k - R) 2 d_k r_k / W,  dr_k/dtau = 1/(beta_k+eps),  dr_k/dbeta_k = -r_k/(beta_k+eps)
// Evaluated on differences of neighbouring values (exact in FP32 for close values), so
// fields with a large offset (theta ~ 300 K) keep their gradient accuracy.  Returns
// R - ref, formed as (c - ref) + (R - c) so a nearby ref leaves no cancellation error.
__device__ float weno5_grad(float a, float b, float c, float d, float e, float weps, float* g, float ref = 0.f) {
  const float c1312 = (float)(13.0 / 12.0);
  const float dab = a - b, dbc = b - c, dcd = c - d, dde = d - e;
  float t1 = dab - dbc, t2 = dab - 3.0f * dbc;
  float t3 = dbc - dcd, t4 = dbc + dcd;
  float t5 = dcd - dde, t6 = 3.0f * dcd - dde;
  float b1 = c1312 * t1 * t1 + 0.25f * t2 * t2;
  float b2 = c1312 * t3 * t3 + 0.25f * t4 * t4;
  float b3 = c1312 * t5 * t5 + 0.25f * t6 * t6;
  float df = b1 - b3, tau = fabsf(df);
  float sg = (df > 0.f) ? 1.f : ((df < 0.f) ? -1.f : 0.f);
  float e1 = b1 + weps, e2 = b2 + weps, e3 = b3 + weps;
  const float ie1 = frcp(e1), ie2 = frcp(e2), ie3 = frcp(e3);
  float x1 = tau * ie1, x2 = tau * ie2, x3 = tau * ie3;
  bool c1 = !(x1 < 1.0e30f), c2 = !(x2 < 1.0e30f), c3 = !(x3 < 1.0e30f);
  float r1 = fminf(1.0e30f, x1), r2 = fminf(1.0e30f, x2), r3 = fminf(1.0e30f, x3);
  float w1 = 0.1f * (1.0f + r1 * r1), w2 = 0.6f * (1.0f + r2 * r2), w3 = 0.3f * (1.0f + r3 * r3);
  const float s6 = (float)(1.0 / 6.0);
  // candidates relative to the upwind value c: h_k = q_k - c
  const float da = a - c, db = b - c, dd = d - c, de = e - c;
  float h1 = (float)(2.0 / 6.0) * da + (float)(-7.0 / 6.0) * db;
  float h2 = -s6 * db + (float)(2.0 / 6.0) * dd;
  float h3 = (float)(5.0 / 6.0) * dd - s6 * de;
  float W = w1 + w2 + w3;
  const float iW = frcp(W);
  float Rd = (w1 * h1 + w2 * h2 + w3 * h3) * iW;
  float R = (c - ref) + Rd;
  float C1 = c1 ? 0.f : (h1 - Rd) * 0.2f * r1 * iW * ie1;
  float C2 = c2 ? 0.f : (h2 - Rd) * 1.2f * r2 * iW * ie2;
  float C3 = c3 ? 0.f : (h3 - Rd) * 0.6f * r3 * iW * ie3;
  float Cs = C1 + C2 + C3;
  float E1 = sg * Cs - C1 * r1, E2 = -C2 * r2, E3 = -sg * Cs - C3 * r3;  // dR/dbeta_k
  float A1 = 2.f * c1312 * t1, B1 = 0.5f * t2, A2 = 2.f * c1312 * t3, B2 = 0.5f * t4;
  float A3 = 2.f * c1312 * t5, B3 = 0.5f * t6;
  float W1 = w1 * iW, W2 = w2 * iW, W3 = w3 * iW;
  g[0] = E1 * (A1 + B1) + W1 * (float)(2.0 / 6.0);
  g[1] = E1 * (-2.f * A1 - 4.f * B1) + E2 * (A2 + B2) + W1 * (float)(-7.0 / 6.0) - W2 * s6;
  g[2] = E1 * (A1 + 3.f * B1) + E2 * (-2.f * A2) + E3 * (A3 + 3.f * B3) + W1 * (float)(11.0 / 6.0) +
         W2 * (float)(5.0 / 6.0) + W3 * (float)(2.0 / 6.0);
  g[3] = E2 * (A2 - B2) + E3 * (-2.f * A3 - 4.f * B3) + W2 * (float)(2.0 / 6.0) + W3 * (float)(5.0 / 6.0);
  g[4] = E3 * (A3 + B3) - W3 * s6;
  return R;
}

// R = weno3(a, b, c) - ref and g[0..2]
__device__ float weno3_grad(float a, float b, float c, float weps, float* g, float ref = 0.f) {
  const float dab = a - b, dbc = b - c;
  float b1 = dab * dab, b2 = dbc * dbc;
  float df = b1 - b2, tau = fabsf(df);
  float sg = (df > 0.f) ? 1.f : ((df < 0.f) ? -1.f : 0.f);
  float e1 = b1 + weps, e2 = b2 + weps;
  const float ie1 = frcp(e1), ie2 = frcp(e2);
  float x1 = tau * ie1, x2 = tau * ie2;
  bool c1 = !(x1 < 1.0e30f), c2 = !(x2 < 1.0e30f);
  float r1 = fminf(1.0e30f, x1), r2 = fminf(1.0e30f, x2);
  float w1 = (float)(1.0 / 3.0) * (1.f + r1 * r1), w2 = (float)(2.0 / 3.0) * (1.f + r2 * r2);
  // candidates relative to the upwind value b: h1 = q1 - b, h2 = q2 - b
  float h1 = -0.5f * dab, h2 = -0.5f * dbc;
  float W = w1 + w2;
  const float iW = frcp(W);
  float Rd = (w1 * h1 + w2 * h2) * iW;
  float R = (b - ref) + Rd;
  float C1 = c1 ? 0.f : (h1 - Rd) * (float)(2.0 / 3.0) * r1 * iW * ie1;
  float C2 = c2 ? 0.f : (h2 - Rd) * (float)(4.0 / 3.0) * r2 * iW * ie2;
  float E1 = sg * (C1 + C2) - C1 * r1, E2 = -sg * (C1 + C2) - C2 * r2;
  float W1 = w1 * iW, W2 = w2 * iW;
  g[0] = E1 * 2.f * dab - 0.5f * W1;
  g[1] = -E1 * 2.f * dab + E2 * 2.f * dbc + 1.5f * W1 + 0.5f * W2;
  g[2] = -E2 * 2.f * dbc + 0.5f * W2;
  return R;
}

// hrecon (lex_common.hpp) minus ref, with its gradient w.r.t. q[m-3+t], t = 0..5.
// Branch-free in the upwind side (lanes run along z, where the mass-flux sign varies, so a
// branch would run both WENO5 gradients): the operands are selected, one weno5_grad is
// evaluated, and its derivatives are mapped back by the same selection.
__device__ float hrecon_grad(const float* q6, bool pos, float weps, float* g6, float ref = 0.f) {
  float g[5];
  const float R = weno5_grad(pos ? q6[0] : q6[5], pos ? q6[1] : q6[4], pos ? q6[2] : q6[3], pos ? q6[3] : q6[2],
                             pos ? q6[4] : q6[1], weps, g, ref);
  g6[0] = pos ? g[0] : 0.f;
  g6[1] = pos ? g[1] : g[4];
  g6[2] = pos ? g[2] : g[3];
  g6[3] = pos ? g[3] : g[2];
  g6[4] = pos ? g[4] : g[1];
  g6[5] = pos ? 0.f : g[0];
  return R;
}

// vrecon (lex_common.hpp) minus ref, with its gradient w.r.t. p[f-2+t], t = 0..5, from
// the six values q[t] = p[f-2+t] (indexes clamped to the column: the rows that reach
// past an end never use those entries).  Branch-free in the upwind side like
// hrecon_grad; the table's near-boundary rows are patched in a branch that only waves
// holding a face with f <= 2 or f >= L-4 enter.
__device__ float vrecon_grad6(const float* q, int L, int f, bool pos, float weps, float* g6, float ref = 0.f) {
  float R = hrecon_grad(q, pos, weps, g6, ref);
  if (f <= 2 || f >= L - 4) {
    // upward rows f = 2, L-3: weno3(p[f-1], p[f], p[f+1]) = (q1, q2, q3); downward rows
    // f = 1, L-4: weno3(p[f+2], p[f+1], p[f]) = (q4, q3, q2); centred rows (upward f = 1,
    // downward f = L-3): 0.5 (p[f] + p[f+1])
    float h[3];
    const float R3 = weno3_grad(pos ? q[1] : q[4], pos ? q[2] : q[3], pos ? q[3] : q[2], weps, h, ref);
    const bool w3 = pos ? (f == 2 || f == L - 3) : (f == 1 || f == L - 4);
    const bool cen = pos ? (f == 1) : (f == L - 3);
    const bool zero = f <= 0 || f >= L - 2;
    if (w3) {
      R = R3;
      g6[0] = 0.f;
      g6[1] = pos ? h[0] : 0.f;
      g6[2] = pos ? h[1] : h[2];
      g6[3] = pos ? h[2] : h[1];
      g6[4] = pos ? 0.f : h[0];
      g6[5] = 0.f;
    } else if (cen) {
      R = 0.5f * ((q[2] - ref) + (q[3] - ref));
      g6[0] = g6[1] = g6[4] = g6[5] = 0.f;
      g6[2] = g6[3] = 0.5f;
    }
    if (zero) {
      R = -ref;
      for (int t = 0; t < 6; ++t) g6[t] = 0.f;
    }
  }
  return R;
}

// vrecon_grad6 on a unit-stride column p[0..L-1]
__device__ float vrecon_grad(const float* __restrict__ p, int L, int f, bool pos, float weps, float* g6,
                            float ref = 0.f) {
  float q[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) q[t] = p[min(max(f - 2 + t, 0), L - 1)];
  return vrecon_grad6(q, L, f, pos, weps, g6, ref);
}

__device__ __forceinline__ float r8w_col(const Dims& d, const float* rc, int Kw) {
  return (Kw <= 0 || Kw >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kw - 1] + rc[Kw]);
}

// flow divergence div(rho0 u) at padded cell (I,J,K) (advection.py:200-226)
__device__ __forceinline__ float fd_at(const Dims& d, const float* __restrict__ rc, const float* __restrict__ u,
                                       const float* __restrict__ v, const float* __restrict__ w, int I, int J, int K) {
  float r = rc[K];
  return (r * u[iu(d, I + 1, J, K)] - r * u[iu(d, I, J, K)]) * d.idx +
         (r * v[iv(d, I, J + 1, K)] - r * v[iv(d, I, J, K)]) * d.idy +
         (r8w_col(d, rc, K + 1) * w[iw(d, I, J, K + 1)] - r8w_col(d, rc, K) * w[iw(d, I, J, K)]) * d.idz;
}

// adjoint of fd_at with weight g, as the six face-velocity weights (upper and lower face
// of each direction)
struct FdGrad {
  float u_hi, u_lo, v_hi, v_lo, w_hi, w_lo;
};
__device__ __forceinline__ FdGrad fd_grad(const Dims& d, const float* __restrict__ rc, int K, float g) {
  const float r = rc[K];
  FdGrad f;
  f.u_hi = g * r * d.idx;
  f.u_lo = -g * r * d.idx;
  f.v_hi = g * r * d.idy;
  f.v_lo = -g * r * d.idy;
  f.w_hi = g * r8w_col(d, rc, K + 1) * d.idz;
  f.w_lo = -g * r8w_col(d, rc, K) * d.idz;
  return f;
}

// ------------------------------------------------------------- A2 scalar advection
// T = [-div F + s div(rho0 u)] / rho0 at interior cells (advection.py:9-34);
// lam = dL/dT on the interior (nx, ny, nz); surface flux prescribed.
//
// Face-centric reverse.  Thread (i, j, k), i in [0, nx], j in [0, ny], k in [0, nz), owns
// the x face at padded I = i+3 (j < ny), the y face at J = j+3 (i < nx), the w level
// K = k+1 and, for k = nz-1, the lid level (i < nx, j < ny), and cell (I, J, K) when it
// is interior.  A face flux's adjoint is the difference of its two cells' lam/rho0 (zero
// outside the interior), so every WENO gradient is evaluated once per face, and the
// div(rho0 u) terms of both cells are collected at the face: each face-velocity adjoint
// (lu, lv, lw) has exactly one writer and is updated without atomics.
// lam/rho0 at interior cell (i, j, k) when ok, else 0 (branch-free load)
__device__ __forceinline__ float mu_cell(const Dims& d, const float* __restrict__ lam, const float* __restrict__ rc,
                                         int i, int j, int k, bool ok) {
  return ld_or0(lam, ((long)i * d.ny + j) * d.nz + k, ok) / rc[ok ? k + 1 : 1];
}

__global__ void k_adv_scalar_vjp(Dims d, const float* __restrict__ rc, const float* __restrict__ s,
                                 const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                 const float* __restrict__ lam, float weps, SlotPlanes sp, float* __restrict__ ls,
                                 float* __restrict__ lu, float* __restrict__ lv, float* __restrict__ lw) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (j > d.ny) return;  // whole wave
  const bool live = k < d.nz;
  if (!live && !sp.zw) return;
  if (!live) k = d.nz - 1;  // a lane past the column only takes part in the shuffles
  const int I = i + 3, J = j + 3, K = k + 1;
  const bool cell = i < d.nx && j < d.ny;
  const bool fx = j < d.ny, fy = i < d.nx, fz = cell && k > 0;
  const long sx = d.sxc(), sy = d.syc();
  const long c = ic(d, I, J, K);
  // Every operand is loaded up front (the x face of row ny and the y face of column nx
  // read ghost lines, which exist), the three reconstructions are evaluated for every
  // thread, and only the adjoint updates depend on which faces the thread owns.
  const float mc = mu_cell(d, lam, rc, i, j, k, cell);
  const float mlx = mu_cell(d, lam, rc, i - 1, j, k, i > 0 && fx);
  const float mly = mu_cell(d, lam, rc, i, j - 1, k, j > 0 && fy);
  const float mb = mu_cell(d, lam, rc, i, j, k - 1, fz);  // cell below
  float qx[6], qy[6], qz[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    qx[t] = s[c + (t - 3) * sx];
    qy[t] = s[c + (t - 3) * sy];
    qz[t] = s[c - K + min(max(k - 2 + t, 0), d.nz + 1)];
  }
  const float sc = qx[3];
  const long au = iu(d, I, J, K), av = iv(d, I, J, K), wb = iw(d, I, J, 0);
  const float ua = u[au], va = v[av], wa = w[wb + K];
  const float rcm = rc[K - 1], r = rc[K], rcp = rc[K + 1];
  const float fd = fd_at(d, rc, u, v, w, I, J, K);
  const float r8 = 0.5f * (rcm + r);
  float gx6[6], gy6[6], gz6[6];
  const float massx = r * ua, massy = r * va, massz = r8 * wa;
  const float Rcx = hrecon_grad(qx, upwind_pos(massx), weps, gx6, sc);
  const float Rcy = hrecon_grad(qy, upwind_pos(massy), weps, gy6, sc);
  const float Rcz = vrecon_grad6(qz, d.nz + 2, k, upwind_pos(massz), weps, gz6, sc);
  // slots: 0..5 the x stencil (cells I-3..I+2 of row J), 6..11 the y stencil, 12..17 the
  // column levels k-2..k+3; the cell's own div(rho0 u) term goes with slot 3 (cell I)
  float ox[6], oy[6], oz[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) ox[t] = oy[t] = oz[t] = 0.f;
  if (fx) {  // x face I: lower face of cell I, upper face of cell I-1
    const float gF = (mc - mlx) * d.idx;
    // d(flux)/du collects mc (R - s_I) - ml (R - s_{I-1}) = gF (R - s_I) - ml (s_I - s_{I-1}) / dx:
    // every factor is a small difference, not a difference of theta-sized products
    const float g = -mlx * d.idx * (sc - qx[2]) + gF * Rcx;
    const float cf = gF * massx;
#pragma unroll
    for (int t = 0; t < 6; ++t) ox[t] = cf * gx6[t];
    if (live && g != 0.f) lu[au] += r * g;
  }
  if (fy) {  // y face J
    const float gF = (mc - mly) * d.idy;
    const float g = -mly * d.idy * (sc - qy[2]) + gF * Rcy;
    const float cf = gF * massy;
#pragma unroll
    for (int t = 0; t < 6; ++t) oy[t] = cf * gy6[t];
    if (live && g != 0.f) lv[av] += r * g;
  }
  if (cell) {
    float g;
    if (k > 0) {  // flux face between levels k and k+1 (the surface face k = 0 is prescribed)
      const float gF = (mc - mb) * d.idz;
      g = -mb * d.idz * (sc - qz[2]) + gF * Rcz;
      const float cf = gF * massz;
#pragma unroll
      for (int t = 0; t < 6; ++t) oz[t] = cf * gz6[t];
    } else {
      g = -mc * sc * d.idz;
    }
    if (live && g != 0.f) lw[wb + K] += r8 * g;
    if (live && k == d.nz - 1) {  // lid level: div(rho0 u) of the top cell only
      const float gt = mc * sc * (0.5f * (r + rcp)) * d.idz;
      if (gt != 0.f) lw[wb + K + 1] += gt;
    }
    ox[3] += mc * fd;
  }
  if (sp.zw) {
    const float zs = zw_sum(oz, k, d.nz);   // every lane of the wave takes part
    if (live && cell) ls[c] += zs;
  }
  if (!live) return;
  const long p = sp.at(i, j, k);
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    sp.put(t, p, ox[t]);
    sp.put(6 + t, p, oy[t]);
    if (!sp.zw) sp.put(12 + t, p, oz[t]);
  }
}

// ls += the slots of k_adv_scalar_vjp that reach padded cell (I, J, K), in slot order
__global__ void k_gather_scalar(Dims d, SlotPlanes sp, float* __restrict__ ls) {
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int I = blockIdx.z;
  if (K >= d.nz + 2 || J >= d.ny + 6) return;
  float a = 0.f;
#pragma unroll
  for (int t = 0; t < 6; ++t) a += sp.get(t, I - t, J - 3, K - 1);
#pragma unroll
  for (int t = 0; t < 6; ++t) a += sp.get(6 + t, I - 3, J - t, K - 1);
  if (!sp.zw) {
#pragma unroll
    for (int t = 0; t < 6; ++t) a += sp.get(12 + t, I - 3, J - 3, K + 2 - t);
  }
  ls[ic(d, I, J, K)] += a;
}

// ------------------------------------------------------------ A3 momentum advection
// Flux-point-centric reverses: every flux point is shared by the two momentum points it
// separates, so its adjoint is the difference of their lam/rho0 and its WENO gradient is
// evaluated once.  The div(rho0 u) product term is reversed per scalar cell, collecting
// both momentum points that average it.
// lam/rho0 at a momentum point when ok, else 0 (branch-free loads)
__device__ __forceinline__ float mu_uface(const Dims& d, const float* __restrict__ lam, const float* __restrict__ rc,
                                          int f, int j, int k, bool ok) {
  return ld_or0(lam, ((long)f * d.ny + j) * d.nz + k, ok) / rc[ok ? k + 1 : 1];
}
__device__ __forceinline__ float mu_vface(const Dims& d, const float* __restrict__ lam, const float* __restrict__ rc,
                                          int i, int f, int k, bool ok) {
  return ld_or0(lam, ((long)i * (d.ny + 1) + f) * d.nz + k, ok) / rc[ok ? k + 1 : 1];
}
__device__ __forceinline__ float mu_wlev(const Dims& d, const float* __restrict__ lam, const float* __restrict__ rc,
                                         int i, int j, int l, bool ok) {
  const int lc = ok ? l : 1;
  return ld_or0(lam, ((long)i * d.ny + j) * (d.nz + 1) + l, ok) / (0.5f * (rc[lc] + rc[lc + 1]));
}

// reverse of advection_u (advection.py:37-74), lam: (nx+1, ny, nz).  Thread (a, b, k),
// a in [0, nx+1], b in [0, ny]: x flux at scalar cell a+3 (b < ny), y flux at corner
// (a+3, b+3) (a <= nx), z flux at the w level below u point (a+3, b+3, k+1) and that
// point's own term (a <= nx, b < ny), div(rho0 u) of cell (a+2, b+3, k+1) (b < ny).
__global__ void k_adv_u_vjp(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                            const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ lam,
                            SlotPlanes sp, float* __restrict__ lu) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y * blockDim.y + threadIdx.y;
  int a = blockIdx.z;
  if (b > d.ny) return;  // whole wave
  const bool live = k < d.nz;
  if (!live && !sp.zw) return;
  if (!live) k = d.nz - 1;  // a lane past the column only takes part in the shuffles
  const float weps = 1.0e-16f;
  const int J = b + 3, K = k + 1;
  const float r = rc[K];
  const bool upt = a <= d.nx && b < d.ny;
  const float mc = mu_uface(d, lam, rc, a, b, k, upt);
  const float mlx = mu_uface(d, lam, rc, a - 1, b, k, a > 0 && b < d.ny);
  const float mly = mu_uface(d, lam, rc, a, b - 1, k, b > 0 && a <= d.nx);
  const float mlz = mu_uface(d, lam, rc, a, b, k - 1, upt && k > 0);
  const long sx = d.sxu(), sy = d.syc();
  // slots (padded targets): 0..5 lu (a+t, J, K), 6..11 lu (a+3, b+t, K), 12..17 lu
  // (a+3, J, k-2+t); 18 lv (a+2, J+1, K), 19 lv (a+2, J, K), 20 lv (a+3, J, K);
  // 21 lw (a+2, J, K+1), 22 lw (a+2, J, K), 23 lw (a+3, J, K)
  float o[kSlotsMom];
#pragma unroll
  for (int q = 0; q < kSlotsMom; ++q) o[q] = 0.f;
  float q6[6], g6[6];
  if (b < d.ny) {
    // x flux at scalar cell a+3: lower flux of u face a, upper flux of u face a-1
    const float ml = mlx;
    const float gF = (mc - ml) * d.idx;
    const long base = iu(d, a, J, K);
    if (gF != 0.f) {
#pragma unroll
      for (int t = 0; t < 6; ++t) q6[t] = u[base + t * sx];
      const float mass = 0.5f * (q6[2] + q6[3]) * r;
      const float R = hrecon_grad(q6, upwind_pos(mass), weps, g6);
      const float cf = gF * mass;
#pragma unroll
      for (int t = 0; t < 6; ++t) o[t] = cf * g6[t];
      const float gm = gF * R * 0.5f * r;
      o[2] += gm;
      o[3] += gm;
    }
    // div(rho0 u) of cell a+2, averaged by u faces a-1 (its lower) and a (its upper)
    const float gl = ml * u[base + 2 * sx], gr = mc * u[base + 3 * sx];
    const FdGrad f = fd_grad(d, rc, K, 0.5f * (gl + gr));
    o[3] += f.u_hi;
    o[2] += f.u_lo;
    o[18] = f.v_hi;
    o[19] = f.v_lo;
    o[21] = f.w_hi;
    o[22] = f.w_lo;
  }
  if (a <= d.nx) {
    const int I = a + 3;
    {  // y flux at corner (I, J): lower flux of u (a, b), upper flux of u (a, b-1)
      const float gF = (mc - mly) * d.idy;
      if (gF != 0.f) {
        const long va = iv(d, I - 1, J, K), vb = iv(d, I, J, K);
        const float mass = 0.5f * (r * v[va] + r * v[vb]);
        const long base = iu(d, I, b, K);
#pragma unroll
        for (int t = 0; t < 6; ++t) q6[t] = u[base + t * sy];
        const float R = hrecon_grad(q6, upwind_pos(mass), weps, g6);
        const float cf = gF * mass;
#pragma unroll
        for (int t = 0; t < 6; ++t) o[6 + t] = cf * g6[t];
        const float gm = gF * R * 0.5f * r;
        o[19] += gm;
        o[20] = gm;
      }
    }
    if (upt) {
      const float* col = u + iu(d, I, J, 0);
      if (k > 0) {  // z flux at w level K between u levels K-1 and K
        const float gF = (mc - mlz) * d.idz;
        if (gF != 0.f) {
          const float rf = 0.5f * (rc[k] + rc[k + 1]);
          const long wa = iw(d, I - 1, J, K), wb = iw(d, I, J, K);
          const float mass = 0.5f * (rf * w[wa] + rf * w[wb]);
          const float R = vrecon_grad(col, d.nz + 2, k, upwind_pos(mass), weps, g6);
          const float cf = gF * mass;
#pragma unroll
          for (int t = 0; t < 6; ++t) o[12 + t] = cf * g6[t];
          const float gm = gF * R * 0.5f * rf;
          o[22] += gm;
          o[23] = gm;
        }
      }
      o[15] += mc * 0.5f * (fd_at(d, rc, u, v, w, I - 1, J, K) + fd_at(d, rc, u, v, w, I, J, K));
    }
  }
  if (sp.zw) {
    const float zs = zw_sum(o + 12, k, d.nz);   // every lane of the wave takes part
    if (live && upt) lu[iu(d, a + 3, J, K)] += zs;
  }
  if (!live) return;
  const long p = sp.at(a, b, k);
#pragma unroll
  for (int q = 0; q < kSlotsMom; ++q)
    if (q < 12 || q >= 18 || !sp.zw) sp.put(q, p, o[q]);
}

// reverse of advection_v (advection.py:77-114), lam: (nx, ny+1, nz).  Thread (a, b, k),
// a in [0, nx], b in [0, ny+1]: x flux at corner (a+3, b+3) (b <= ny), y flux at scalar
// cell b+3 (a < nx), z flux and own term of v point (a+3, b+3, k+1) (a < nx, b <= ny),
// div(rho0 u) of cell (a+3, b+2, k+1) (a < nx).
__global__ void k_adv_v_vjp(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                            const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ lam,
                            SlotPlanes sp, float* __restrict__ lv) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y * blockDim.y + threadIdx.y;
  int a = blockIdx.z;
  if (b > d.ny + 1) return;  // whole wave
  const bool live = k < d.nz;
  if (!live && !sp.zw) return;
  if (!live) k = d.nz - 1;  // a lane past the column only takes part in the shuffles
  const float weps = 1.0e-16f;
  const int I = a + 3, K = k + 1;
  const float r = rc[K];
  const bool vpt = a < d.nx && b <= d.ny;
  const float mc = mu_vface(d, lam, rc, a, b, k, vpt);
  const float mlx = mu_vface(d, lam, rc, a - 1, b, k, a > 0 && b <= d.ny);
  const float mly = mu_vface(d, lam, rc, a, b - 1, k, b > 0 && a < d.nx);
  const float mlz = mu_vface(d, lam, rc, a, b, k - 1, vpt && k > 0);
  const long sx = d.sxv(), sy = d.syc();
  // slots (padded targets): 0..5 lv (a+t, b+3, K), 6..11 lv (I, b+t, K), 12..17 lv
  // (I, b+3, k-2+t); 18 lu (I+1, b+2, K), 19 lu (I, b+2, K), 20 lu (I, b+3, K);
  // 21 lw (I, b+2, K+1), 22 lw (I, b+2, K), 23 lw (I, b+3, K)
  float o[kSlotsMom];
#pragma unroll
  for (int q = 0; q < kSlotsMom; ++q) o[q] = 0.f;
  float q6[6], g6[6];
  if (b <= d.ny) {  // x flux at corner (I, J): lower flux of v (a, b), upper flux of v (a-1, b)
    const int J = b + 3;
    const float gF = (mc - mlx) * d.idx;
    if (gF != 0.f) {
      const long ua = iu(d, I, J - 1, K), ub = iu(d, I, J, K);
      const float mass = 0.5f * (r * u[ua] + r * u[ub]);
      const long base = iv(d, a, J, K);
#pragma unroll
      for (int t = 0; t < 6; ++t) q6[t] = v[base + t * sx];
      const float R = hrecon_grad(q6, upwind_pos(mass), weps, g6);
      const float cf = gF * mass;
#pragma unroll
      for (int t = 0; t < 6; ++t) o[t] = cf * g6[t];
      const float gm = gF * R * 0.5f * r;
      o[19] = gm;
      o[20] = gm;
    }
  }
  if (a < d.nx) {
    {
      // y flux at scalar cell b+3: lower flux of v face b, upper flux of v face b-1
      const float ml = mly;
      const float gF = (mc - ml) * d.idy;
      const long base = iv(d, I, b, K);
      if (gF != 0.f) {
#pragma unroll
        for (int t = 0; t < 6; ++t) q6[t] = v[base + t * sy];
        const float mass = 0.5f * (q6[2] + q6[3]) * r;
        const float R = hrecon_grad(q6, upwind_pos(mass), weps, g6);
        const float cf = gF * mass;
#pragma unroll
        for (int t = 0; t < 6; ++t) o[6 + t] = cf * g6[t];
        const float gm = gF * R * 0.5f * r;
        o[8] += gm;
        o[9] += gm;
      }
      // div(rho0 u) of cell (I, b+2), averaged by v faces b-1 and b
      const float gl = ml * v[base + 2 * sy], gr = mc * v[base + 3 * sy];
      const FdGrad f = fd_grad(d, rc, K, 0.5f * (gl + gr));
      o[18] = f.u_hi;
      o[19] += f.u_lo;
      o[9] += f.v_hi;
      o[8] += f.v_lo;
      o[21] = f.w_hi;
      o[22] = f.w_lo;
    }
    if (vpt) {
      const int J = b + 3;
      const float* col = v + iv(d, I, J, 0);
      if (k > 0) {
        const float gF = (mc - mlz) * d.idz;
        if (gF != 0.f) {
          const float rf = 0.5f * (rc[k] + rc[k + 1]);
          const long wa = iw(d, I, J - 1, K), wb = iw(d, I, J, K);
          const float mass = 0.5f * (rf * w[wa] + rf * w[wb]);
          const float R = vrecon_grad(col, d.nz + 2, k, upwind_pos(mass), weps, g6);
          const float cf = gF * mass;
#pragma unroll
          for (int t = 0; t < 6; ++t) o[12 + t] = cf * g6[t];
          const float gm = gF * R * 0.5f * rf;
          o[22] += gm;
          o[23] = gm;
        }
      }
      o[15] += mc * 0.5f * (fd_at(d, rc, u, v, w, I, J - 1, K) + fd_at(d, rc, u, v, w, I, J, K));
    }
  }
  if (sp.zw) {
    const float zs = zw_sum(o + 12, k, d.nz);   // every lane of the wave takes part
    if (live && vpt) lv[iv(d, I, b + 3, K)] += zs;
  }
  if (!live) return;
  const long p = sp.at(a, b, k);
#pragma unroll
  for (int q = 0; q < kSlotsMom; ++q)
    if (q < 12 || q >= 18 || !sp.zw) sp.put(q, p, o[q]);
}

// reverse of advection_w (advection.py:117-149), lam: (nx, ny, nz+1); levels 0 and nz
// carry no advection.  Thread (a, b, t), a in [0, nx], b in [0, ny], t in [0, nz]: x flux
// at (a+3, b+3) on w level t+1 (b < ny), y flux there (a < nx), z flux at scalar level t,
// own term of w level t and div(rho0 u) of cell level t+1 (a < nx, b < ny).
__global__ void k_adv_w_vjp(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                            const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ lam,
                            SlotPlanes sp) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y * blockDim.y + threadIdx.y;
  int a = blockIdx.z;
  if (t > d.nz || b > d.ny) return;
  const float weps = 1.0e-16f;
  const int I = a + 3, J = b + 3, K = t + 1;
  const bool lev = t >= 1 && t <= d.nz - 1;  // advected w level
  const bool col_in = a < d.nx && b < d.ny;
  const float mc = mu_wlev(d, lam, rc, a, b, t, lev && col_in);
  const float mlx = mu_wlev(d, lam, rc, a - 1, b, t, lev && a > 0 && b < d.ny);
  const float mly = mu_wlev(d, lam, rc, a, b - 1, t, lev && b > 0 && a < d.nx);
  // w levels t-1 and t+1 of this column (advected for 1..nz-1)
  const float mlo = mu_wlev(d, lam, rc, a, b, t - 1, col_in && t >= 2 && t - 1 <= d.nz - 1);
  const float mup = mu_wlev(d, lam, rc, a, b, t + 1, col_in && t + 1 <= d.nz - 1);
  const long sx = d.sxw(), sy = d.syw();
  // slots (padded targets): 0..5 lw (a+q, J, K), 6..11 lw (I, b+q, K), 12..17 lw
  // (I, J, t-2+q); 18 lu (I, J, K-1), 19 lu (I, J, K), 20 lu (I+1, J, K);
  // 21 lv (I, J, K-1), 22 lv (I, J, K), 23 lv (I, J+1, K)
  float o[kSlotsMom];
#pragma unroll
  for (int q = 0; q < kSlotsMom; ++q) o[q] = 0.f;
  float q6[6], g6[6];
  if (lev && b < d.ny) {  // x flux: lower flux of w column a, upper flux of column a-1
    const float gF = (mc - mlx) * d.idx;
    if (gF != 0.f) {
      const long u0 = iu(d, I, J, K - 1), u1 = iu(d, I, J, K);
      const float mass = 0.5f * (rc[K - 1] * u[u0] + rc[K] * u[u1]);
      const long base = iw(d, a, J, K);
#pragma unroll
      for (int q = 0; q < 6; ++q) q6[q] = w[base + q * sx];
      const float R = hrecon_grad(q6, upwind_pos(mass), weps, g6);
      const float cf = gF * mass;
#pragma unroll
      for (int q = 0; q < 6; ++q) o[q] = cf * g6[q];
      o[18] = gF * R * 0.5f * rc[K - 1];
      o[19] = gF * R * 0.5f * rc[K];
    }
  }
  if (lev && a < d.nx) {  // y flux: lower flux of w row b, upper flux of row b-1
    const float gF = (mc - mly) * d.idy;
    if (gF != 0.f) {
      const long v0 = iv(d, I, J, K - 1), v1 = iv(d, I, J, K);
      const float mass = 0.5f * (rc[K - 1] * v[v0] + rc[K] * v[v1]);
      const long base = iw(d, I, b, K);
#pragma unroll
      for (int q = 0; q < 6; ++q) q6[q] = w[base + q * sy];
      const float R = hrecon_grad(q6, upwind_pos(mass), weps, g6);
      const float cf = gF * mass;
#pragma unroll
      for (int q = 0; q < 6; ++q) o[6 + q] = cf * g6[q];
      o[21] = gF * R * 0.5f * rc[K - 1];
      o[22] = gF * R * 0.5f * rc[K];
    }
  }
  if (col_in) {
    const float* col = w + iw(d, I, J, 0);
    if (t >= 1) {  // z flux at scalar level t: lower flux of w level t, upper flux of level t-1
      const float gF = (mc - mlo) * d.idz;
      if (gF != 0.f) {
        const int cz = t;
        const float mass = 0.5f * (col[cz] + col[cz + 1]) * rc[cz];
        const float R = vrecon_grad(col, d.nz + 3, cz, upwind_pos(mass), weps, g6);
        const float cf = gF * mass;
#pragma unroll
        for (int q = 0; q < 6; ++q) o[12 + q] = cf * g6[q];
        const float gm = gF * R * 0.5f * rc[cz];
        o[14] += gm;
        o[15] += gm;
      }
    }
    if (lev) o[15] += mc * 0.5f * (fd_at(d, rc, u, v, w, I, J, K - 1) + fd_at(d, rc, u, v, w, I, J, K));
    if (t < d.nz) {
      // div(rho0 u) of cell level K = t+1, averaged by w levels t (its lower) and t+1 (its upper)
      const float gsum = mc * col[K] + mup * col[K + 1];
      const FdGrad f = fd_grad(d, rc, K, 0.5f * gsum);
      o[20] = f.u_hi;
      o[19] += f.u_lo;
      o[23] = f.v_hi;
      o[22] += f.v_lo;
      o[16] += f.w_hi;
      o[15] += f.w_lo;
    }
  }
  const long p = sp.at(a, b, t);
#pragma unroll
  for (int q = 0; q < kSlotsMom; ++q) sp.put(q, p, o[q]);
}

// The momentum reverses' slots that reach each padded point, added in slot order
// (U, V, W: the slot planes of k_adv_u_vjp, k_adv_v_vjp, k_adv_w_vjp; a null plane set
// contributes nothing).
__global__ void k_gather_lu(Dims d, SlotPlanes U, SlotPlanes V, SlotPlanes W, float* __restrict__ lu) {
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int I = blockIdx.z;
  if (K >= d.nz + 2 || J >= d.ny + 6) return;
  float a = 0.f;
#pragma unroll
  for (int t = 0; t < 6; ++t) a += U.get(t, I - t, J - 3, K - 1);
#pragma unroll
  for (int t = 0; t < 6; ++t) a += U.get(6 + t, I - 3, J - t, K - 1);
  if (!U.zw) {
#pragma unroll
    for (int t = 0; t < 6; ++t) a += U.get(12 + t, I - 3, J - 3, K + 2 - t);
  }
  a += V.get(18, I - 4, J - 2, K - 1);
  a += V.get(19, I - 3, J - 2, K - 1);
  a += V.get(20, I - 3, J - 3, K - 1);
  a += W.get(18, I - 3, J - 3, K);
  a += W.get(19, I - 3, J - 3, K - 1);
  a += W.get(20, I - 4, J - 3, K - 1);
  lu[iu(d, I, J, K)] += a;
}

__global__ void k_gather_lv(Dims d, SlotPlanes U, SlotPlanes V, SlotPlanes W, float* __restrict__ lv) {
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int I = blockIdx.z;
  if (K >= d.nz + 2 || J >= d.ny + 7) return;
  float a = 0.f;
#pragma unroll
  for (int t = 0; t < 6; ++t) a += V.get(t, I - t, J - 3, K - 1);
#pragma unroll
  for (int t = 0; t < 6; ++t) a += V.get(6 + t, I - 3, J - t, K - 1);
  if (!V.zw) {
#pragma unroll
    for (int t = 0; t < 6; ++t) a += V.get(12 + t, I - 3, J - 3, K + 2 - t);
  }
  a += U.get(18, I - 2, J - 4, K - 1);
  a += U.get(19, I - 2, J - 3, K - 1);
  a += U.get(20, I - 3, J - 3, K - 1);
  a += W.get(21, I - 3, J - 3, K);
  a += W.get(22, I - 3, J - 3, K - 1);
  a += W.get(23, I - 3, J - 4, K - 1);
  lv[iv(d, I, J, K)] += a;
}

__global__ void k_gather_lw(Dims d, SlotPlanes U, SlotPlanes V, SlotPlanes W, float* __restrict__ lw) {
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int I = blockIdx.z;
  if (K >= d.nz + 3 || J >= d.ny + 6) return;
  float a = 0.f;
#pragma unroll
  for (int q = 0; q < 6; ++q) a += W.get(q, I - q, J - 3, K - 1);
#pragma unroll
  for (int q = 0; q < 6; ++q) a += W.get(6 + q, I - 3, J - q, K - 1);
#pragma unroll
  for (int q = 0; q < 6; ++q) a += W.get(12 + q, I - 3, J - 3, K + 2 - q);
  a += U.get(21, I - 2, J - 3, K - 2);
  a += U.get(22, I - 2, J - 3, K - 1);
  a += U.get(23, I - 3, J - 3, K - 1);
  a += V.get(21, I - 3, J - 2, K - 2);
  a += V.get(22, I - 3, J - 2, K - 1);
  a += V.get(23, I - 3, J - 3, K - 1);
  lw[iw(d, I, J, K)] += a;
}

// ------------------------------------------------------------------ B1 buoyancy
__device__ __forceinline__ void theta_rho_grad(float th, float qv, float& dth, float& dqv) {
  float inv = 1.0f / (1.0f + qv);
  dth = (1.0f + (float)kReps * qv) * inv;
  dqv = th * ((float)kReps - 1.0f) * inv * inv;
}

// theta_rho over the padded array: lth += ltr * d/dth, lqv += ltr * d/dqv (elementwise)
__global__ void k_theta_rho_vjp(long n, const float* __restrict__ th, const float* __restrict__ qv,
                                const float* __restrict__ ltr, float* __restrict__ lth, float* __restrict__ lqv) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float g = ltr[t];
  if (g == 0.f) return;
  float a, b;
  theta_rho_grad(th[t], qv[t], a, b);
  lth[t] += g * a;
  lqv[t] += g * b;
}

// b8w (interpolate_scalar2w + extrapolate_bottom_top of calculate_buoyancy,
// pressure_equations.py:123-172), lam on (nx, ny, nz+1) w levels -> theta, qv adjoints.
// Gather form, one thread per cell: w level l in 1..nz-1 averages cells l-1 and l, level 0
// extrapolates from cells 0, 1, 2 and level nz from nz-1, nz-2, nz-3 (1.875, -1.25, 0.375);
// each cell sums the weights of the levels that read it in one fixed order.
__global__ void k_b8w_vjp(Dims d, const float* __restrict__ th0c, const float* __restrict__ qv0c,
                          const float* __restrict__ th, const float* __restrict__ qv, const float* __restrict__ lam,
                          float* __restrict__ lth, float* __restrict__ lqv) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const int nz = d.nz;
  const float* lc = lam + ((long)i * d.ny + j) * (nz + 1);
  float W = 0.f;
  if (k <= 2) W += lc[0] * (k == 0 ? 1.875f : (k == 1 ? -1.25f : 0.375f));
  if (k >= 1) W += lc[k] * 0.5f;           // level k: upper cell of its average
  if (k <= nz - 2) W += lc[k + 1] * 0.5f;  // level k+1: lower cell of its average
  if (k >= nz - 3) W += lc[nz] * (k == nz - 1 ? 1.875f : (k == nz - 2 ? -1.25f : 0.375f));
  if (W == 0.f) return;
  const int K = k + 1;
  const long c = ic(d, i + 3, j + 3, K);
  const float tr0 = th0c[K] * (1.0f + (float)kReps * qv0c[K]) / (1.0f + qv0c[K]);
  float a, b;
  theta_rho_grad(th[c], qv[c], a, b);
  const float gB = W * (float)kG / tr0;
  lth[c] += gB * a;
  lqv[c] += gB * b;
}

// ------------------------------------------------------------- S1 acoustic substep
// Reverse of k_acoustic in two passes, without atomics (the step driver and
// lex_acoustic_substep_vjp).
// k_acoustic_vjp_faces3: per face, the total adjoint g of the new velocity = its folded
// output adjoint + the terms of the (one or two) cells whose pressure tendency reads it;
// then old velocity (+= when acc, else =), rhs (+= g dts, one owner per face), and
// gdt = g dts into gu/gv/gw (interior face layout) for the cell pass.
__global__ void k_acoustic_vjp_faces3(Dims d, const float* __restrict__ rc, const float* __restrict__ pi0c,
                                      const float* __restrict__ ps, const float* __restrict__ lpo,
                                      const float* __restrict__ luo, const float* __restrict__ lvo,
                                      const float* __restrict__ lwo, float dts, int acc, float* __restrict__ lu,
                                      float* __restrict__ lv, float* __restrict__ lw, float* __restrict__ lru,
                                      float* __restrict__ lrv, float* __restrict__ lrw, float* __restrict__ gu,
                                      float* __restrict__ gv, float* __restrict__ gw) {
  // levels 0..nz-1: the lid level nz carries no face of this pass (w levels 1..nz-1
  // only), so the grid's z extent is nz
  int k = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nz-1
  int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = blockIdx.z;                             // 0..nx
  if (k >= d.nz || j > d.ny) return;
  const float cr = (float)(-kRd / kCv);
  const float g11 = 1.1f * dts;
  const bool du = k < d.nz && j < d.ny;                      // u face (i, j, k)
  const bool dv = k < d.nz && i < d.nx;                      // v face (i, j, k)
  const bool dw = i < d.nx && j < d.ny && k > 0 && k < d.nz;  // w level k of column (i, j)
  // clamped indices: every load below is in bounds and issues before the first use
  const int kc = min(k, d.nz - 1), jc = min(j, d.ny - 1), icl = min(i, d.nx - 1);
  const int kw = min(max(k, 1), d.nz - 1);
  const long sxc = d.sxc(), syc = d.syc();
  // u face (i+3, jc+3, kc+1)
  const int Ku = kc + 1;
  const long ou = ((long)i * d.ny + jc) * d.nz + kc, au = iu(d, i + 3, jc + 3, Ku);
  const long cRu = ic(d, i + 3, jc + 3, Ku), cLu = cRu - sxc;
  const float psRu = ps[cRu], psLu = ps[cLu];
  const float lpRu = ld_or0(lpo, cRu, i < d.nx), lpLu = ld_or0(lpo, cLu, i > 0);
  const float gou = luo[au], lu0 = acc ? lu[au] : 0.f, lru0 = lru[ou], piu = pi0c[Ku];
  // v face (icl+3, j+3, kc+1)
  const long ov = ((long)icl * (d.ny + 1) + j) * d.nz + kc, av = iv(d, icl + 3, j + 3, Ku);
  const long cRv = ic(d, icl + 3, j + 3, Ku), cLv = cRv - syc;
  const float psRv = ps[cRv], psLv = ps[cLv];
  const float lpRv = ld_or0(lpo, cRv, j < d.ny), lpLv = ld_or0(lpo, cLv, j > 0);
  const float gov = lvo[av], lv0 = acc ? lv[av] : 0.f, lrv0 = lrv[ov];
  // w level kw of column (icl, jc): cells above (Kw) and below (Kw-1)
  const int Kw = kw + 1;
  const long ow = ((long)icl * d.ny + jc) * (d.nz + 1) + kw, aw = iw(d, icl + 3, jc + 3, Kw);
  const long ca = ic(d, icl + 3, jc + 3, Kw), cb = ca - 1;
  const float psa = ps[ca], psb = ps[cb], lpa = lpo[ca], lpb = lpo[cb];
  const float gow = lwo[aw], lw0 = acc ? lw[aw] : 0.f, lrw0 = lrw[ow];
  const float rwa = rc[Kw], rwb = rc[Kw - 1], pia = pi0c[Kw], pib = pi0c[Kw - 1];
  if (du) {
    const float dps = (psRu - psLu) * d.idx;
    float g = gou;
    if (i < d.nx) g += g11 * lpRu * (-cr * (piu + psRu) * d.idx - 0.5f * dps);
    if (i > 0) g += g11 * lpLu * (cr * (piu + psLu) * d.idx - 0.5f * dps);
    lu[au] = acc ? lu0 + g : g;
    const float gd = g * dts;
    lru[ou] = lru0 + gd;
    gu[ou] = gd;
  }
  if (dv) {
    const float dps = (psRv - psLv) * d.idy;
    float g = gov;
    if (j < d.ny) g += g11 * lpRv * (-cr * (piu + psRv) * d.idy - 0.5f * dps);
    if (j > 0) g += g11 * lpLv * (cr * (piu + psLv) * d.idy - 0.5f * dps);
    lv[av] = acc ? lv0 + g : g;
    const float gd = g * dts;
    lrv[ov] = lrv0 + gd;
    gv[ov] = gd;
  }
  if (dw) {
    const float r8 = 0.5f * (rwb + rwa);
    const float dzz = (pia - pib) * d.idz + (psa - psb) * d.idz;
    float g = gow;
    g += g11 * lpa * (-cr * (pia + psa) * d.idz + (-0.5f * r8 / rwa) * dzz);
    g += g11 * lpb * (cr * (pib + psb) * d.idz + (-0.5f * r8 / rwb) * dzz);
    lw[aw] = acc ? lw0 + g : g;
    const float gd = g * dts;
    lrw[ow] = lrw0 + gd;
    gw[ow] = gd;
  }
}

// k_acoustic_vjp_cells3: thread per padded cell of the interior and its one-cell x/y
// ring (levels 1..nz).  Gathers (a) the new pi' adjoint (identity) and the PGF terms of
// the cell's six faces (weights gdt from the face pass) into lp and ltr, and (b) the
// advection_pi terms of the cell and of its six neighbours into lps; the new face
// velocities are recomputed from the substep's inputs as in the forward.
// lp and lps may be the same array (the first substep of the first stage).
__global__ void k_acoustic_vjp_cells3(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                                      const float* __restrict__ v, const float* __restrict__ w,
                                      const float* __restrict__ p, const float* __restrict__ ru,
                                      const float* __restrict__ rv, const float* __restrict__ rw,
                                      const float* __restrict__ tr, const float* __restrict__ lpo,
                                      const float* __restrict__ gu, const float* __restrict__ gv,
                                      const float* __restrict__ gw, float dts, float* lp, float* __restrict__ ltr,
                                      float* lps) {
  const int Z = blockIdx.x * blockDim.x + threadIdx.x + 1;  // 1..nz
  const int Y = blockIdx.y * blockDim.y + threadIdx.y + 2;  // 2..ny+3
  const int X = blockIdx.z + 2;                             // 2..nx+3
  if (Z > d.nz || Y > d.ny + 3) return;
  const long sx = d.sxc(), sy = d.syc();
  const long c = ic(d, X, Y, Z);
  const int i = X - 3, j = Y - 3, k = Z - 1;
  const bool in_x = i >= 0 && i < d.nx, in_y = j >= 0 && j < d.ny, inner = in_x && in_y;
  // which of the six faces exist (u faces 0..nx, v faces 0..ny, w levels 1..nz-1)
  const bool fu0 = in_y && i >= 0 && i <= d.nx, fu1 = in_y && i + 1 >= 0 && i + 1 <= d.nx;
  const bool fv0 = in_x && j >= 0 && j <= d.ny, fv1 = in_x && j + 1 >= 0 && j + 1 <= d.ny;
  const bool fw0 = inner && k >= 1, fw1 = inner && k + 1 <= d.nz - 1;
  const long ou0 = fu0 ? ((long)i * d.ny + j) * d.nz + k : 0, ou1 = fu1 ? ((long)(i + 1) * d.ny + j) * d.nz + k : 0;
  const long ov0 = fv0 ? ((long)i * (d.ny + 1) + j) * d.nz + k : 0;
  const long ov1 = fv1 ? ((long)i * (d.ny + 1) + j + 1) * d.nz + k : 0;
  const long ow0 = inner ? ((long)i * d.ny + j) * (d.nz + 1) + k : 0;
  // every load issues up front (all indices in bounds)
  const float pc = p[c], pxm = p[c - sx], pxp = p[c + sx], pym = p[c - sy], pyp = p[c + sy], pzm = p[c - 1],
              pzp = p[c + 1];
  const float tc = tr[c], txm = tr[c - sx], txp = tr[c + sx], tym = tr[c - sy], typ = tr[c + sy],
              tzm = tr[c - 1], tzp = tr[c + 1];
  const long cu = iu(d, X, Y, Z), cv = iv(d, X, Y, Z), cw = iw(d, X, Y, Z);
  const float u0 = u[cu], u1 = u[cu + d.sxu()], v0 = v[cv], v1 = v[cv + d.syc()], w0 = w[cw], w1 = w[cw + 1];
  const float ru0 = ld_or0(ru, ou0, fu0), ru1 = ld_or0(ru, ou1, fu1), rv0 = ld_or0(rv, ov0, fv0),
              rv1 = ld_or0(rv, ov1, fv1), rw0 = ld_or0(rw, ow0, fw0), rw1 = ld_or0(rw, ow0 + 1, fw1);
  const float gu0 = ld_or0(gu, ou0, fu0), gu1 = ld_or0(gu, ou1, fu1), gv0 = ld_or0(gv, ov0, fv0),
              gv1 = ld_or0(gv, ov1, fv1), gw0 = ld_or0(gw, ow0, fw0), gw1 = ld_or0(gw, ow0 + 1, fw1);
  const float Lc = ld_or0(lpo, c, inner);
  const float Lxp = ld_or0(lpo, c + sx, i + 1 >= 0 && i + 1 < d.nx && in_y);
  const float Lxm = ld_or0(lpo, c - sx, i - 1 >= 0 && i - 1 < d.nx && in_y);
  const float Lyp = ld_or0(lpo, c + sy, j + 1 >= 0 && j + 1 < d.ny && in_x);
  const float Lym = ld_or0(lpo, c - sy, j - 1 >= 0 && j - 1 < d.ny && in_x);
  const float Lzp = ld_or0(lpo, c + 1, inner && k + 1 <= d.nz - 1);
  const float Lzm = ld_or0(lpo, c - 1, inner && k >= 1);
  const float rcm = rc[Z - 1], rcc = rc[Z], rcp = rc[Z + 1];
  const float h = (float)(-kCp * 0.5);
  float U0 = 0.f, U1 = 0.f, V0 = 0.f, V1 = 0.f, W0 = 0.f, W1 = 0.f;
  float alp = 0.f, atr = 0.f;  // lp and ltr sums
  // faces on this cell's - side (X, Y, level Z) and + side (X+1, Y+1, level Z+1)
  if (fu0) {
    const float dp = (pc - pxm) * d.idx, ts = tc + txm;
    U0 = u0 + (h * ts * dp + ru0) * dts;
    atr += gu0 * h * dp;
    alp += gu0 * h * ts * d.idx;
  }
  if (fu1) {
    const float dp = (pxp - pc) * d.idx, ts = txp + tc;
    U1 = u1 + (h * ts * dp + ru1) * dts;
    atr += gu1 * h * dp;
    alp -= gu1 * h * ts * d.idx;
  }
  if (fv0) {
    const float dp = (pc - pym) * d.idy, ts = tc + tym;
    V0 = v0 + (h * ts * dp + rv0) * dts;
    atr += gv0 * h * dp;
    alp += gv0 * h * ts * d.idy;
  }
  if (fv1) {
    const float dp = (pyp - pc) * d.idy, ts = typ + tc;
    V1 = v1 + (h * ts * dp + rv1) * dts;
    atr += gv1 * h * dp;
    alp -= gv1 * h * ts * d.idy;
  }
  // w levels Z (below the cell) and Z+1 (above); zero at the surface and the lid
  if (fw0) {
    const float dp = (pc - pzm) * d.idz, ts = tc + tzm;
    W0 = w0 + (h * ts * dp + rw0) * dts;
    atr += gw0 * h * dp;
    alp += gw0 * h * ts * d.idz;
  }
  if (fw1) {
    const float dp = (pzp - pc) * d.idz, ts = tzp + tc;
    W1 = w1 + (h * ts * dp + rw1) * dts;
    atr += gw1 * h * dp;
    alp -= gw1 * h * ts * d.idz;
  }
  // advection_pi reverse (one_step_integration.py:543-546): G = 1.1 dts lam(new pi')
  const float g11 = 1.1f * dts;
  const float cr = (float)(-kRd / kCv);
  float aps = 0.f;
  if (inner) {
    alp += Lc;
    const float G = g11 * Lc;
    const float div = (U1 - U0) * d.idx + (V1 - V0) * d.idy + (W1 - W0) * d.idz;
    const float gx = -0.5f * G * d.idx, gy = -0.5f * G * d.idy;
    aps += cr * div * G + (gx * U0 - gx * U1) + (gy * V0 - gy * V1);
    if (k > 0) aps += (-0.5f * (0.5f * (rcm + rcc)) / rcc * G) * W0 * d.idz;
    if (k < d.nz - 1) aps -= (-0.5f * (0.5f * (rcc + rcp)) / rcc * G) * W1 * d.idz;
    // the cells above and below read the shared w level
    if (k + 1 <= d.nz - 1) aps -= (-0.5f * (0.5f * (rcc + rcp)) / rcp * (g11 * Lzp)) * W1 * d.idz;
    if (k >= 1) aps += (-0.5f * (0.5f * (rcm + rcc)) / rcm * (g11 * Lzm)) * W0 * d.idz;
  }
  // x and y neighbours read the shared face velocities
  aps += 0.5f * g11 * d.idx * (Lxp * U1 - Lxm * U0);
  aps += 0.5f * g11 * d.idy * (Lyp * V1 - Lym * V0);
  if (lp == lps) {
    const float t = alp + aps;
    if (t != 0.f) lp[c] += t;
  } else {
    if (alp != 0.f) lp[c] += alp;
    if (aps != 0.f) lps[c] += aps;
  }
  if (atr != 0.f) ltr[c] += atr;
}

// ------------------------------------------------------------- G1 Smagorinsky
// Gather-form reverses: one thread per padded cell (or momentum point) sums every
// contribution of the flux and stress stencils that read it, so each adjoint element has
// a single writer and no atomics are needed.  Ghost cells receive the contributions of
// the boundary stencils; the caller folds them as for the other padded adjoints.

// -lam/rho0 of an interior cell, 0 elsewhere (padded indices)
__device__ __forceinline__ float sgs_ldiv(const Dims& d, const float* __restrict__ lam, const float* __restrict__ rc,
                                          int X, int Y, int Z) {
  const int i = X - 3, j = Y - 3, k = Z - 1;
  const bool ok = i >= 0 && i < d.nx && j >= 0 && j < d.ny && k >= 0 && k < d.nz;
  const float L = ld_or0(lam, ((long)i * d.ny + j) * d.nz + k, ok);
  return ok ? -L / rc[ok ? Z : 1] : 0.f;
}

// SGS scalar diffusion (turbulence.py:80-109): lam (nx, ny, nz) -> kh, s adjoints.  The
// flux through the face below padded cell cc along axis sd,
//   t = -r8 0.5 (kh[cc] + kh[cc-sd]) (s[cc] - s[cc-sd]) id,
// enters the two cells' tendencies with total weight g = (ldiv[cc-sd] - ldiv[cc]) id.
// Thread per padded cell X in [2, nx+3], Y in [2, ny+3], Z in [1, nz].
__global__ void k_smag_scalar_vjp(Dims d, const float* __restrict__ rc, const float* __restrict__ kh,
                                  const float* __restrict__ s, const float* __restrict__ lam, float* __restrict__ lkh,
                                  float* __restrict__ ls) {
  const int Z = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int X = blockIdx.z;
  if (Z < 1 || Z > d.nz || Y < 2 || Y > d.ny + 3 || X < 2 || X > d.nx + 3) return;
  const long sx = d.sxc(), sy = d.syc();
  const long c = ic(d, X, Y, Z);
  const float lc = sgs_ldiv(d, lam, rc, X, Y, Z);
  const float r = rc[Z];
  float gkh = 0.f, gs = 0.f;
  // this cell is cc (side +1) or cc - sd (side -1) of the face
  auto face = [&](long cc, long sd, float id, float r8, float g, float side) {
    if (g == 0.f) return;
    const float ds = (s[cc] - s[cc - sd]) * id;
    const float kk = 0.5f * (kh[cc] + kh[cc - sd]);
    gkh += -g * r8 * 0.5f * ds;
    gs += side * (-g * r8 * kk * id);
  };
  face(c, sx, d.idx, r, (sgs_ldiv(d, lam, rc, X - 1, Y, Z) - lc) * d.idx, 1.f);
  face(c + sx, sx, d.idx, r, (lc - sgs_ldiv(d, lam, rc, X + 1, Y, Z)) * d.idx, -1.f);
  face(c, sy, d.idy, r, (sgs_ldiv(d, lam, rc, X, Y - 1, Z) - lc) * d.idy, 1.f);
  face(c + sy, sy, d.idy, r, (lc - sgs_ldiv(d, lam, rc, X, Y + 1, Z)) * d.idy, -1.f);
  // vertical faces between levels 1..nz only (zero flux at the surface and the lid)
  if (Z >= 2) face(c, 1, d.idz, 0.5f * (rc[Z - 1] + rc[Z]), (sgs_ldiv(d, lam, rc, X, Y, Z - 1) - lc) * d.idz, 1.f);
  if (Z + 1 <= d.nz)
    face(c + 1, 1, d.idz, 0.5f * (rc[Z] + rc[Z + 1]), (lc - sgs_ldiv(d, lam, rc, X, Y, Z + 1)) * d.idz, -1.f);
  if (gkh != 0.f) lkh[c] += gkh;
  if (gs != 0.f) ls[c] += gs;
}

// -lam/rho0 weights of the SGS momentum tendencies at padded u, v points and w levels
__device__ __forceinline__ float sgs_gu(const Dims& d, const float* __restrict__ lsu, const float* __restrict__ rc,
                                        int X, int Y, int Z) {
  const int f = X - 3, j = Y - 3, k = Z - 1;
  const bool ok = f >= 0 && f <= d.nx && j >= 0 && j < d.ny && k >= 0 && k < d.nz;
  const float L = ld_or0(lsu, ((long)f * d.ny + j) * d.nz + k, ok);
  const float r = rc[ok ? Z : 1];
  return ok ? -L / (0.5f * (r + r)) : 0.f;
}
__device__ __forceinline__ float sgs_gv(const Dims& d, const float* __restrict__ lsv, const float* __restrict__ rc,
                                        int X, int Y, int Z) {
  const int i = X - 3, f = Y - 3, k = Z - 1;
  const bool ok = i >= 0 && i < d.nx && f >= 0 && f <= d.ny && k >= 0 && k < d.nz;
  const float L = ld_or0(lsv, ((long)i * (d.ny + 1) + f) * d.nz + k, ok);
  const float r = rc[ok ? Z : 1];
  return ok ? -L / (0.5f * (r + r)) : 0.f;
}
__device__ __forceinline__ float sgs_gw(const Dims& d, const float* __restrict__ lsw, const float* __restrict__ rc,
                                        int X, int Y, int l) {
  const int i = X - 3, j = Y - 3;
  const bool ok = i >= 0 && i < d.nx && j >= 0 && j < d.ny && l >= 1 && l <= d.nz - 1;
  const float L = ld_or0(lsw, ((long)i * d.ny + j) * (d.nz + 1) + l, ok);
  const int lc = ok ? l : 1;
  return ok ? -L / (0.5f * (rc[lc] + rc[lc + 1])) : 0.f;
}

// SGS momentum tendencies (turbulence.py:9-79): lam on u, v, w points -> km, s_ij
// adjoints.  tau_ij = -2 km s_ij, so with G_ij = dL/dtau_ij at a cell,
// lkm += -2 sum G_ij s_ij and ls_ij += -2 G_ij km.  G_ij gathers:
//   d11 (u): -d(tau11)/dx at faces;  d22 (v), d33 (w) likewise;
//   d12: tau12 averaged to the four cells around an edge, differenced across it (the two
//        cells shared by both edges cancel);
//   d13, d23 at u/v points: tau interpolated to w levels 1..nz-1, differenced vertically;
//   at w levels: centred differences over two columns of both adjacent levels.
// Thread per padded cell X in [2, nx+3], Y in [2, ny+3], Z in [1, nz].
__global__ void k_smag_uvw_vjp(Dims d, const float* __restrict__ rc, const float* __restrict__ km,
                               const float* __restrict__ s11, const float* __restrict__ s22,
                               const float* __restrict__ s33, const float* __restrict__ s12,
                               const float* __restrict__ s13, const float* __restrict__ s23,
                               const float* __restrict__ lsu, const float* __restrict__ lsv,
                               const float* __restrict__ lsw, float* __restrict__ lkm, float* __restrict__ l11,
                               float* __restrict__ l22, float* __restrict__ l33, float* __restrict__ l12,
                               float* __restrict__ l13, float* __restrict__ l23) {
  const int Z = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int X = blockIdx.z;
  if (Z < 1 || Z > d.nz || Y < 2 || Y > d.ny + 3 || X < 2 || X > d.nx + 3) return;
  const int nz = d.nz;
  auto U = [&](int x, int y, int z) { return sgs_gu(d, lsu, rc, x, y, z); };
  auto V = [&](int x, int y, int z) { return sgs_gv(d, lsv, rc, x, y, z); };
  auto W = [&](int x, int y, int l) { return sgs_gw(d, lsw, rc, x, y, l); };
  const float G11 = d.idx * (U(X, Y, Z) - U(X + 1, Y, Z));
  const float G22 = d.idy * (V(X, Y, Z) - V(X, Y + 1, Z));
  const float G33 = d.idz * (W(X, Y, Z - 1) - W(X, Y, Z));
  const float G12 = 0.25f * d.idy * ((U(X + 1, Y - 1, Z) + U(X, Y - 1, Z)) - (U(X + 1, Y + 1, Z) + U(X, Y + 1, Z))) +
                    0.25f * d.idx * ((V(X - 1, Y, Z) + V(X - 1, Y + 1, Z)) - (V(X + 1, Y + 1, Z) + V(X + 1, Y, Z)));
  // vertical pattern of a u (v) column: level Z gathers +g(Z-1) (Z <= nz), -g(Z+1), and
  // g(Z) where only one of the two w levels around Z is inside 1..nz-1
  const float c0 = (Z <= nz - 1 ? 1.f : 0.f) - (Z >= 2 ? 1.f : 0.f);
  const float cm = Z <= nz ? 1.f : 0.f;
  auto zp = [&](float gm, float g0, float gp) { return c0 * g0 + cm * gm - gp; };
  float G13 = 0.25f * d.idz *
              (zp(U(X, Y, Z - 1), U(X, Y, Z), U(X, Y, Z + 1)) + zp(U(X + 1, Y, Z - 1), U(X + 1, Y, Z), U(X + 1, Y, Z + 1)));
  float G23 = 0.25f * d.idz *
              (zp(V(X, Y, Z - 1), V(X, Y, Z), V(X, Y, Z + 1)) + zp(V(X, Y + 1, Z - 1), V(X, Y + 1, Z), V(X, Y + 1, Z + 1)));
  const float hx = 0.5f / (2.0f * d.dx), hy = 0.5f / (2.0f * d.dy);
  G13 += hx * ((W(X - 1, Y, Z - 1) + W(X - 1, Y, Z)) - (W(X + 1, Y, Z - 1) + W(X + 1, Y, Z)));
  G23 += hy * ((W(X, Y - 1, Z - 1) + W(X, Y - 1, Z)) - (W(X, Y + 1, Z - 1) + W(X, Y + 1, Z)));
  const long c = ic(d, X, Y, Z);
  const float k2 = -2.0f * km[c];
  float a = 0.f;
  if (G11 != 0.f) { a += G11 * s11[c]; l11[c] += k2 * G11; }
  if (G22 != 0.f) { a += G22 * s22[c]; l22[c] += k2 * G22; }
  if (G33 != 0.f) { a += G33 * s33[c]; l33[c] += k2 * G33; }
  if (G12 != 0.f) { a += G12 * s12[c]; l12[c] += k2 * G12; }
  if (G13 != 0.f) { a += G13 * s13[c]; l13[c] += k2 * G13; }
  if (G23 != 0.f) { a += G23 * s23[c]; l23[c] += k2 * G23; }
  if (a != 0.f) lkm[c] += -2.0f * a;
}

// compute_k (turbulence.py:430-446), elementwise over the padded array:
// km = cg2 sqrt(max(s2 - n2/Pr, 0)), kh = km/Pr.  Writes ls2, ln2 (not accumulated).
// dL/dcg2 = sum(dL/dkm_total * sqrt(x)) is reduced in FP64 without atomics: a grid of
// at most kSmagKBlocks blocks strides over the array, block b writes part[b], and
// k_smag_k_sum adds the partials in a fixed order to *lcg2 (deterministic).
constexpr int kSmagKBlocks = 512;

inline unsigned smag_k_blocks(long n) {
  const long b = (n + 255) / 256;
  return (unsigned)(b < kSmagKBlocks ? b : kSmagKBlocks);
}

__device__ __forceinline__ double block_sum_256(double v) {
  __shared__ double wsum[4];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  return (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

__global__ void __launch_bounds__(256) k_smag_k_vjp(long n, const float* __restrict__ s2, const float* __restrict__ n2,
                                                    const float* __restrict__ lkm, const float* __restrict__ lkh,
                                                    float csgs2, float prandtl, float* __restrict__ ls2,
                                                    float* __restrict__ ln2, double* __restrict__ part) {
  double acc = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    float x = s2[t] - n2[t] / prandtl;
    float G = lkm[t] + lkh[t] / prandtl;
    float a = 0.f, b = 0.f;
    if (x > 0.0f && G != 0.f) {
      float sq = sqrtf(x);
      float gx = G * csgs2 * 0.5f / sq;
      a = gx;
      b = -gx / prandtl;
      acc += (double)(G * sq);
    }
    ls2[t] = a;
    ln2[t] = b;
  }
  const double tot = block_sum_256(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// one block of 256 threads: *lcg2 += sum of part[0..nb)
__global__ void __launch_bounds__(256) k_smag_k_sum(const double* __restrict__ part, int nb, double* __restrict__ lcg2) {
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) a += part[b];
  const double tot = block_sum_256(a);
  if (threadIdx.x == 0) lcg2[0] += tot;
}

// sum over the cells whose N^2 reads w level l of (adjoint of that N^2) x its weight:
// cells 1..nz-2 average levels k and k+1, cell 0 extrapolates levels 1, 2, 3 and cell nz-1
// levels nz-1, nz-2, nz-3 (1.875, -1.25, 0.375).  Lc[k+1] = adjoint of cell k's N^2.
__device__ __forceinline__ float nm_level_weight(const float* __restrict__ Lc, int nz, int l) {
  float G = 0.f;
  if (l >= 1 && l <= 3) G += Lc[1] * (l == 1 ? 1.875f : (l == 2 ? -1.25f : 0.375f));
  if (l >= 1 && l <= nz - 2) G += 0.5f * Lc[l + 1];
  if (l >= 2 && l <= nz - 1) G += 0.5f * Lc[l];
  if (l >= nz - 3 && l <= nz - 1) G += Lc[nz] * (l == nz - 1 ? 1.875f : (l == nz - 2 ? -1.25f : 0.375f));
  return G;
}

// compute_nm (turbulence.py:391-427), reverse: folded adjoint at interior cells -> theta,
// qv.  N^2 at w level l = g/dz log(a/b), a = theta (1 + (Rv/Rd - 1) qv) at padded level
// l+1, b the same at level l.  Gather form, one thread per cell: the cell is the upper
// operand of level P-1 and the lower one of level P.
__global__ void k_smag_nm_vjp(Dims d, const float* __restrict__ th, const float* __restrict__ qv,
                              const float* __restrict__ ln2, float* __restrict__ lth, float* __restrict__ lqv) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const long c = ic(d, i + 3, j + 3, 0);
  const float* Lc = ln2 + c;
  const int nz = d.nz, P = k + 1;
  const float em = (float)kRepsm1, gs = (float)kG / d.dz;
  const float t = th[c + P], f = 1.0f + em * qv[c + P];
  float dth = 0.f, dqv = 0.f;
  if (P - 1 >= 1) {  // upper operand of w level P-1
    const float ga = nm_level_weight(Lc, nz, P - 1) * gs / (t * f);
    dth += ga * f;
    dqv += ga * t * em;
  }
  if (P <= nz - 1) {  // lower operand of w level P
    const float gb = -(nm_level_weight(Lc, nz, P) * gs) / (t * f);
    dth += gb * f;
    dqv += gb * t * em;
  }
  lth[c + P] += dth;
  lqv[c + P] += dqv;
}

// compute_deformation (turbulence.py:298-388), reverse in two passes.  Pass 1 (interior
// cells, in place): the folded adjoints of rho*S_ij (l_ij) and of the deformation (ldf)
// become the weights H of the velocity differences each cell's strain reads:
//   H11 = r g11, H22, H33 likewise; H12 = r g12 / 4; H13 = r g13, H23 = r g23, with
//   g_ii = l_ii + 4 q s_ii, g_ij = l_ij + 8 q s_ij, q = ldf / r^2
// (deformation = (2 (a11^2 + a22^2 + a33^2) + 4 (a12^2 + a13^2 + a23^2)) / r^2).
__global__ void k_smag_deform_h(Dims d, const float* __restrict__ rc, const float* __restrict__ s11,
                                const float* __restrict__ s22, const float* __restrict__ s33,
                                const float* __restrict__ s12, const float* __restrict__ s13,
                                const float* __restrict__ s23, const float* __restrict__ ldf, float* __restrict__ h11,
                                float* __restrict__ h22, float* __restrict__ h33, float* __restrict__ h12,
                                float* __restrict__ h13, float* __restrict__ h23) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const int K = k + 1;
  const long c = ic(d, i + 3, j + 3, K);
  const float r = rc[K];
  const float q = ldf[c] / (r * r);
  h11[c] = (h11[c] + 4.0f * q * s11[c]) * r;
  h22[c] = (h22[c] + 4.0f * q * s22[c]) * r;
  h33[c] = (h33[c] + 4.0f * q * s33[c]) * r;
  h12[c] = 0.25f * r * (h12[c] + 8.0f * q * s12[c]);
  h13[c] = r * (h13[c] + 8.0f * q * s13[c]);
  h23[c] = r * (h23[c] + 8.0f * q * s23[c]);
}

// H of an interior cell, 0 elsewhere (padded indices)
__device__ __forceinline__ float hcell(const Dims& d, const float* __restrict__ h, int X, int Y, int Z) {
  const bool ok = X >= 3 && X < d.nx + 3 && Y >= 3 && Y < d.ny + 3 && Z >= 1 && Z <= d.nz;
  return ld_or0(h, ic(d, X, Y, Z), ok);
}

// Weight of the xz (yz) edge at horizontal face position F between columns F-1 and F (of
// the other axis' column Y), w level l+1: the strain a13 of an interior cell K in
// [2, nz-1] averages the four edges around it (1/4 each); the bottom (top) cell
// extrapolates from the three edges above (below) it with weights 1.875, -1.25, 0.375
// (extrapolate_bottom_top), each halved by the two-face average.
template <int AX>
__device__ __forceinline__ float strain_edge(const Dims& d, const float* __restrict__ h, int F, int Y, int l) {
  auto H2 = [&](int Z) {
    return AX == 0 ? hcell(d, h, F, Y, Z) + hcell(d, h, F - 1, Y, Z) : hcell(d, h, Y, F, Z) + hcell(d, h, Y, F - 1, Z);
  };
  const int nz = d.nz;
  float e = 0.f;
  if (l + 1 >= 2 && l + 1 <= nz - 1) e += 0.25f * H2(l + 1);
  if (l >= 2 && l <= nz - 1) e += 0.25f * H2(l);
  if (l >= 1 && l <= 3) e += 0.5f * (l == 1 ? 1.875f : (l == 2 ? -1.25f : 0.375f)) * H2(1);
  const int t = nz - 1 - l;
  if (t >= 0 && t <= 2) e += 0.5f * (t == 0 ? 1.875f : (t == 1 ? -1.25f : 0.375f)) * H2(nz);
  return e;
}

// weight of the xy edge (corner) at (F, G): the four cells around it
__device__ __forceinline__ float strain_corner(const Dims& d, const float* __restrict__ h12, int F, int G, int Z) {
  return (hcell(d, h12, F, G, Z) + hcell(d, h12, F - 1, G - 1, Z)) +
         (hcell(d, h12, F, G - 1, Z) + hcell(d, h12, F - 1, G, Z));
}

// Pass 2: thread per padded velocity point (u, v, w arrays share the thread grid
// (nx+7, ny+7, nz+3)); every velocity adjoint gathers its strain stencils.
__global__ void k_smag_deform_vjp(Dims d, const float* __restrict__ h11, const float* __restrict__ h22,
                                  const float* __restrict__ h33, const float* __restrict__ h12,
                                  const float* __restrict__ h13, const float* __restrict__ h23,
                                  float* __restrict__ lu, float* __restrict__ lv, float* __restrict__ lw) {
  const int Z = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int X = blockIdx.z;
  if (Z > d.nz + 2 || Y > d.ny + 6) return;
  if (X < d.nx + 7 && Y < d.ny + 6 && Z < d.nz + 2) {  // u point (X, Y, Z)
    float g = d.idx * (hcell(d, h11, X - 1, Y, Z) - hcell(d, h11, X, Y, Z)) +
              0.5f * d.idy * (strain_corner(d, h12, X, Y, Z) - strain_corner(d, h12, X, Y + 1, Z)) +
              0.5f * d.idz * (strain_edge<0>(d, h13, X, Y, Z - 1) - strain_edge<0>(d, h13, X, Y, Z));
    if (g != 0.f) lu[iu(d, X, Y, Z)] += g;
  }
  if (X < d.nx + 6 && Y < d.ny + 7 && Z < d.nz + 2) {  // v point (X, Y, Z)
    float g = d.idy * (hcell(d, h22, X, Y - 1, Z) - hcell(d, h22, X, Y, Z)) +
              0.5f * d.idx * (strain_corner(d, h12, X, Y, Z) - strain_corner(d, h12, X + 1, Y, Z)) +
              0.5f * d.idz * (strain_edge<1>(d, h23, Y, X, Z - 1) - strain_edge<1>(d, h23, Y, X, Z));
    if (g != 0.f) lv[iv(d, X, Y, Z)] += g;
  }
  if (X < d.nx + 6 && Y < d.ny + 6) {  // w point (X, Y, Z), Z < nz+3
    float g = d.idz * (hcell(d, h33, X, Y, Z - 1) - hcell(d, h33, X, Y, Z)) +
              0.5f * d.idx * (strain_edge<0>(d, h13, X, Y, Z - 1) - strain_edge<0>(d, h13, X + 1, Y, Z - 1)) +
              0.5f * d.idy * (strain_edge<1>(d, h23, Y, X, Z - 1) - strain_edge<1>(d, h23, Y + 1, X, Z - 1));
    if (g != 0.f) lw[iw(d, X, Y, Z)] += g;
  }
}

// ------------------------------------------------------------------ H1 halo folds
// Reverse of the padded stores: each ghost's adjoint is added to its interior source,
// then cleared.  Work is per interior point (every ghost has exactly one source), mode 0
// scalar, 1 u, 2 v, 3 w.  Only points with ghost images do anything: those within four
// points of an x or y edge (all levels) and the bottom and top levels of every column.
__device__ __forceinline__ long fold_idx(const Dims& d, int mode, int X, int Y, int Z) {
  if (mode == 0) return ic(d, X, Y, Z);
  if (mode == 1) return iu(d, X, Y, Z);
  if (mode == 2) return iv(d, X, Y, Z);
  return iw(d, X, Y, Z);
}

// All images are loaded before anything is stored, so a point costs one memory round
// trip (the loops have fixed bounds and stay in registers).
__device__ __forceinline__ void fold_point(const Dims& d, int mode, float* __restrict__ a, int i, int j, int k,
                                           int nk) {
  int xs[3], ys[3], zs[3];
  int nxs, nys, nzs = 1;
  if (mode == 1) ghost_list_face(i, d.nx, xs, nxs); else ghost_list_x(i, d.nx, xs, nxs);
  if (mode == 2) ghost_list_face(j, d.ny, ys, nys); else ghost_list_x(j, d.ny, ys, nys);
  zs[0] = k + 1;
  if (k == 0) zs[nzs++] = 0;
  if (k == nk - 1) zs[nzs++] = nk + 1;
  if (nxs == 1 && nys == 1 && nzs == 1) return;  // no ghost image: nothing to move
  float val[3][3][3];
#pragma unroll
  for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
    for (int b1 = 0; b1 < 3; ++b1)
#pragma unroll
      for (int c1 = 0; c1 < 3; ++c1)
        val[a1][b1][c1] = (a1 < nxs && b1 < nys && c1 < nzs) ? a[fold_idx(d, mode, xs[a1], ys[b1], zs[c1])] : 0.f;
  float sum = 0.f;
#pragma unroll
  for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
    for (int b1 = 0; b1 < 3; ++b1)
#pragma unroll
      for (int c1 = 0; c1 < 3; ++c1)
        if ((a1 | b1 | c1) != 0) sum += val[a1][b1][c1];
#pragma unroll
  for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
    for (int b1 = 0; b1 < 3; ++b1)
#pragma unroll
      for (int c1 = 0; c1 < 3; ++c1)
        if ((a1 | b1 | c1) != 0 && a1 < nxs && b1 < nys && c1 < nzs) a[fold_idx(d, mode, xs[a1], ys[b1], zs[c1])] = 0.f;
  a[fold_idx(d, mode, xs[0], ys[0], zs[0])] = val[0][0][0] + sum;
}

struct FoldExt {
  int ni, nj, nk;
  __host__ __device__ FoldExt(const Dims& d, int mode)
      : ni(d.nx + (mode == 1 ? 1 : 0)), nj(d.ny + (mode == 2 ? 1 : 0)), nk(d.nz + (mode == 3 ? 1 : 0)) {}
  // columns within 4 of an x or y edge: 8 full rows of nj, then 8 per remaining row
  __host__ __device__ int nband() const { return 8 * nj + 8 * (ni - 8); }
  __host__ __device__ int ninner() const { return (ni - 8) * (nj - 8); }
  __host__ __device__ long nwork() const { return (long)nband() * nk + ninner(); }
};

// every interior point (small grids)
__global__ void k_fold(Dims d, int mode, float* __restrict__ a) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  const FoldExt e(d, mode);
  if (k >= e.nk || j >= e.nj || i >= e.ni) return;
  fold_point(d, mode, a, i, j, k, e.nk);
}

// One launch: work items [0, nband*nk) are the points of the edge-band columns (level
// fastest), the rest are the inner columns, whose only images are the vertical ghosts
// below the bottom level and above the top level.
__global__ void k_fold_edges(Dims d, int mode, float* __restrict__ a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const FoldExt e(d, mode);
  const long nb = (long)e.nband() * e.nk;
  if (t < nb) {
    const int b = (int)(t / e.nk), k = (int)(t - (long)b * e.nk);
    int i, j;
    if (b < 8 * e.nj) {
      const int r = b / e.nj;
      i = r < 4 ? r : e.ni - 8 + r;
      j = b - r * e.nj;
    } else {
      const int q = b - 8 * e.nj, r = q >> 3, c = q & 7;
      i = 4 + r;
      j = c < 4 ? c : e.nj - 8 + c;
    }
    fold_point(d, mode, a, i, j, k, e.nk);
    return;
  }
  const long q = t - nb;
  if (q >= e.ninner()) return;
  const int i = 4 + (int)(q / (e.nj - 8)), j = 4 + (int)(q % (e.nj - 8));
  float* col = a + fold_idx(d, mode, i + 3, j + 3, 0);
  const float g0 = col[0], s1 = col[1], sn = col[e.nk], gn = col[e.nk + 1];
  col[0] = 0.f;
  col[e.nk + 1] = 0.f;
  if (e.nk == 1) {
    col[1] = s1 + (g0 + gn);
  } else {
    col[1] = s1 + g0;
    col[e.nk] = sn + gn;
  }
}

inline void launch_fold(const Dims& d, int mode, float* a, hipStream_t st) {
  const FoldExt e(d, mode);
  if (e.ni < 9 || e.nj < 9) {
    hipLaunchKernelGGL(k_fold, dim3((e.nk + 63) / 64, (e.nj + 3) / 4, e.ni), dim3(64, 4, 1), 0, st, d, mode, a);
    return;
  }
  hipLaunchKernelGGL(k_fold_edges, dim3((unsigned)((e.nwork() + 255) / 256)), dim3(256), 0, st, d, mode, a);
}

// Several folds in one launch (blockIdx.y selects the field): the same work items as
// k_fold_edges for each field.  Fields with mode < 0 are skipped.
constexpr int kFoldBatch = 8;
struct FoldBatch {
  float* a[kFoldBatch];
  int mode[kFoldBatch];
};
__global__ void k_fold_edges_batch(Dims d, FoldBatch fb) {
  const int f = blockIdx.y;
  const int mode = fb.mode[f];
  if (mode < 0) return;
  float* __restrict__ a = fb.a[f];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const FoldExt e(d, mode);
  const long nb = (long)e.nband() * e.nk;
  if (t < nb) {
    const int b = (int)(t / e.nk), k = (int)(t - (long)b * e.nk);
    int i, j;
    if (b < 8 * e.nj) {
      const int r = b / e.nj;
      i = r < 4 ? r : e.ni - 8 + r;
      j = b - r * e.nj;
    } else {
      const int q = b - 8 * e.nj, r = q >> 3, c = q & 7;
      i = 4 + r;
      j = c < 4 ? c : e.nj - 8 + c;
    }
    fold_point(d, mode, a, i, j, k, e.nk);
    return;
  }
  const long q = t - nb;
  if (q >= e.ninner()) return;
  const int i = 4 + (int)(q / (e.nj - 8)), j = 4 + (int)(q % (e.nj - 8));
  float* col = a + fold_idx(d, mode, i + 3, j + 3, 0);
  const float g0 = col[0], s1 = col[1], sn = col[e.nk], gn = col[e.nk + 1];
  col[0] = 0.f;
  col[e.nk + 1] = 0.f;
  if (e.nk == 1) {
    col[1] = s1 + (g0 + gn);
  } else {
    col[1] = s1 + g0;
    col[e.nk] = sn + gn;
  }
}

// folds of n fields (modes[q], a[q]), batched by kFoldBatch (small grids: one by one)
inline void launch_fold_many(const Dims& d, int n, const int* modes, float* const* a, hipStream_t st) {
  if (d.nx + 1 < 9 || d.ny + 1 < 9 || d.nx < 9 || d.ny < 9) {
    for (int q = 0; q < n; ++q) launch_fold(d, modes[q], a[q], st);
    return;
  }
  for (int q0 = 0; q0 < n; q0 += kFoldBatch) {
    FoldBatch fb;
    long work = 0;
    const int m = (n - q0 < kFoldBatch) ? n - q0 : kFoldBatch;
    for (int q = 0; q < kFoldBatch; ++q) {
      fb.a[q] = q < m ? a[q0 + q] : nullptr;
      fb.mode[q] = q < m ? modes[q0 + q] : -1;
      if (q < m) {
        const long wq = FoldExt(d, modes[q0 + q]).nwork();
        work = wq > work ? wq : work;
      }
    }
    hipLaunchKernelGGL(k_fold_edges_batch, dim3((unsigned)((work + 255) / 256), (unsigned)m), dim3(256), 0, st, d, fb);
  }
}

// zero several buffers in one launch (blockIdx.y selects the buffer)
struct FillBatch {
  float* a[kFoldBatch];
  long n[kFoldBatch];
};
__global__ void k_fill_batch(FillBatch fb) {
  const int f = blockIdx.y;
  float* __restrict__ a = fb.a[f];
  const long n = fb.n[f];
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) a[t] = 0.0f;
}
inline void launch_zero_many(int n, float* const* a, const size_t* counts, hipStream_t st) {
  for (int q0 = 0; q0 < n; q0 += kFoldBatch) {
    FillBatch fb;
    const int m = (n - q0 < kFoldBatch) ? n - q0 : kFoldBatch;
    long mx = 0;
    for (int q = 0; q < kFoldBatch; ++q) {
      fb.a[q] = q < m ? a[q0 + q] : nullptr;
      fb.n[q] = q < m ? (long)counts[q0 + q] : 0;
      mx = fb.n[q] > mx ? fb.n[q] : mx;
    }
    long blocks = (mx + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_batch, dim3((unsigned)blocks, (unsigned)m), dim3(256), 0, st, fb);
  }
}

// ------------------------------------------------------------- E1 scalar update
// next = pad(s0 + dt T + dt sgs) (one_step_integration.py:450-465, :158-162):
// ls0 += lnext, lT = dt lnext, lsgs += dt lnext (lnext folded; interior cells).
__global__ void k_scalar_update_vjp(Dims d, const float* __restrict__ lnext, float dt, float* __restrict__ ls0,
                                    float* __restrict__ lT, float* __restrict__ lsgs) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  long c = ic(d, i + 3, j + 3, k + 1);
  long o = ((long)i * d.ny + j) * d.nz + k;
  float g = lnext[c];
  ls0[c] += g;
  lT[o] = g * dt;
  if (lsgs) lsgs[o] += g * dt;
}

__global__ void k_add_to(long n, const float* __restrict__ a, float* __restrict__ b) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] += a[t];
}

__global__ void k_fill(long n, float* __restrict__ a, float v) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

}  // namespace lex
