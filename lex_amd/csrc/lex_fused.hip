// Fused stage kernels for the device WS-RK3 driver.
//
// k_stage_scalars: the scalars' WENO advection + Euler update + frozen-SGS add-in of
// one stage (update_theta_euler / update_qv_euler, one_step_integration.py:450-465,
// :158-162), plus the flow divergence (advection.py:200-226) and theta_rho
// (one_step_integration.py:135) inline, in one launch (one scalar per workgroup).
//
// Work decomposition (each face flux computed once):
//   block  = full z column (64*ceil(nz/64) lanes) x TY rows of y of one scalar; marches
//            TX cells in x
//   x-face : from a 6-value register window along x; the lower face is the previous
//            step's upper face
//   y-face : each row computes its lower face from an LDS plane (rows j0-3..j0+TY+2);
//            the last row also computes the tile's upper face
//   z-face : each lane its lower face from the LDS column; the lid face is zero
#pragma once
#include "lex_common.hpp"

namespace lex {

constexpr int kFTY = 4;    // rows of y per block
constexpr int kFTX = 16;   // cells marched in x per block
constexpr int kMaxScalars = 8;
constexpr int kSlots = 3;  // ceil((kFTY + 6) * L / (64 * ceil(nz / 64) * kFTY)) <= 3 for L <= nz + 3

// Per-thread plane-load slots of a (kFTY + 6) x L plane (rows j0-3.. of the padded
// field, clamped to the last padded row), fixed for the whole x-march: offsets within
// one x-plane of a field with row stride `rs`.
// Slot q of thread tid is plane element tid + q * nt.
struct PlaneSlots {
  int off[kSlots];
  int n;
  int tid, nt;
  __device__ __forceinline__ int dst(int q) const { return tid + q * nt; }
};
__device__ __forceinline__ PlaneSlots plane_slots(int tid, int nt, int L, int j0, int row_max, int rs) {
  PlaneSlots p;
  p.n = 0;
  p.tid = tid;
  p.nt = nt;
#pragma unroll
  for (int q = 0; q < kSlots; ++q) {
    int e = tid + q * nt;
    int row = e / L, col = e - row * L;
    p.off[q] = min(j0 + row, row_max) * rs + col;
    if (e < (kFTY + 6) * L) p.n = q + 1;
  }
  return p;
}
__device__ __forceinline__ void plane_load(float* __restrict__ P, const float* __restrict__ plane, const PlaneSlots& ps) {
#pragma unroll
  for (int q = 0; q < kSlots; ++q)
    if (q < ps.n) P[ps.dst(q)] = plane[ps.off[q]];
}
// Split form for software pipelining: fetch the next plane into registers while the
// current one is being used, store it to LDS at the top of the next step.
struct PlaneRegs {
  float v[kSlots];
};
__device__ __forceinline__ void plane_fetch(PlaneRegs& r, const float* __restrict__ plane, const PlaneSlots& ps) {
#pragma unroll
  for (int q = 0; q < kSlots; ++q) r.v[q] = (q < ps.n) ? plane[ps.off[q]] : 0.0f;
}
__device__ __forceinline__ void plane_fetch32(PlaneRegs& r, const float* __restrict__ field, uint32_t plane_off,
                                              const PlaneSlots& ps) {
#pragma unroll
  for (int q = 0; q < kSlots; ++q) r.v[q] = (q < ps.n) ? ld32(field, plane_off + (uint32_t)ps.off[q]) : 0.0f;
}
__device__ __forceinline__ void plane_store(float* __restrict__ P, const PlaneRegs& r, const PlaneSlots& ps) {
#pragma unroll
  for (int q = 0; q < kSlots; ++q)
    if (q < ps.n) P[ps.dst(q)] = r.v[q];
}

struct ScalarSet {
  const float* s0[kMaxScalars];   // xi0 fields (padded)
  const float* s[kMaxScalars];    // stage fields (padded)
  const float* sgs[kMaxScalars];  // frozen SGS tendencies (interior) or null
  float* out[kMaxScalars];        // next-stage fields (padded, halos filled)
  float weps[kMaxScalars];
  const float* heat = nullptr;    // diabatic heating added to scalar 0's tendency (interior) or null
  // bottom-face fluxes (nx, ny) of scalar 0 (sen/Cp) and 1 (evaporation) or null: the
  // lower boundary condition of the vertical flux (advection.py:15)
  const float* sfc[2] = {nullptr, nullptr};
  int n;
};

// the surface fluxes of scalars 0 and 1 at column p, read by the surface lanes only
// (zero elsewhere and without surface fluxes), once per column before the scalar loop
struct BottomFlux {
  float f0, f1;
  __device__ __forceinline__ float of(int q) const { return q == 0 ? f0 : (q == 1 ? f1 : 0.0f); }
};
__device__ __forceinline__ BottomFlux bottom_fluxes(const ScalarSet& S, bool surface, long p) {
  BottomFlux b{0.0f, 0.0f};
  if (surface && S.sfc[0]) b.f0 = S.sfc[0][p];
  if (surface && S.sfc[1]) b.f1 = S.sfc[1][p];
  return b;
}

// Flow divergence div(rho0 u) of a cell (get_divergence, advection.py:200-226) for a
// horizontally uniform rho column, from the cell's face velocities already in registers:
// r = rho at the cell, r8lo / r8hi at its lower / upper w level (zero at the ghost levels).
__device__ __forceinline__ float fd_ops(const Dims& d, float r, float r8lo, float r8hi, float u1, float u0, float v1,
                                        float v0, float w1, float w0) {
  return (r * u1 - r * u0) * d.idx + (r * v1 - r * v0) * d.idy + (r8hi * w1 - r8lo * w0) * d.idz;
}

// The momentum kernels address every field through 32-bit element offsets of the
// thread's point, advanced by one x-plane per step of the march (one integer add per
// field instead of a 64-bit index product per load); an operand shared by a mass flux
// and the flow divergence is loaded once, and the x-1 column's operands are carried
// from the previous step.

// ---------------------------------------------------------------------------------
// k_mom_u: rhs_u = advection_u + sgs_u at u faces 0..nx-1 (face nx is the periodic
// duplicate of face 0 and is written from it).  advection.py:37-74, :561-771.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_mom_u(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                                                 const float* __restrict__ v, const float* __restrict__ w,
                                                 const float* __restrict__ sgs, const float* __restrict__ tau_x,
                                                 CorOpt cor, float* __restrict__ ru) {
  extern __shared__ float lds[];
  const float weps = 1.0e-16f;
  const int nzl = blockDim.x, tz = threadIdx.x, ty = threadIdx.y, nt = nzl * kFTY, tid = ty * nzl + tz;
  const int Lz = d.nz + 2;
  float* P = lds;
  float* FY = P + (kFTY + 6) * Lz;
  float* FZ = FY + (kFTY + 1) * nzl;
  const TileId tb = xcd_tile();
  const int j0 = tb.y * kFTY, i0 = tb.x * kFTX;
  const int k = lane_level(tz, nzl, d.nz), j = j0 + ty;
  const bool act = (k < d.nz) && (j < d.ny);
  const int K = k + 1;
  const int Kc = act ? K : 1, Jc = (j < d.ny) ? j + 3 : 3;
  const bool last_row = (ty == kFTY - 1) || (j == d.ny - 1);
  const float r = rc[Kc];
  const float ir = 1.0f / r;   // one division per thread instead of one per cell and field
  const float r8lo = (Kc <= 0 || Kc >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc - 1] + rc[Kc]);
  const float r8hi = (Kc + 1 >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc] + rc[Kc + 1]);
  const float r8w = 0.5f * (rc[Kc - 1] + rc[Kc]);   // rho at the w level below (z mass flux)
  const int iend = min(i0 + kFTX, d.nx);
  const PlaneSlots psl = plane_slots(tid, nt, Lz, j0, d.ny + 5, (int)d.syc());
  const uint32_t usx = (uint32_t)d.sxu(), usy = (uint32_t)d.syc();
  const uint32_t usxv = (uint32_t)d.sxv(), usxw = (uint32_t)d.sxw(), usyw = (uint32_t)d.syw();
  const uint32_t usi = (uint32_t)d.ny * (uint32_t)d.nz;
  // offsets of the point at x = I: ou (u), ov (v), ow (w), oi the interior (nx+1, ny, nz)
  // arrays (sgs_u, ru), op the first element of u's x-plane
  uint32_t ou = (uint32_t)(i0 + 3) * usx + (uint32_t)Jc * usy + (uint32_t)Kc;
  uint32_t ov = (uint32_t)(i0 + 3) * usxv + (uint32_t)Jc * usy + (uint32_t)Kc;
  uint32_t ow = (uint32_t)(i0 + 3) * usxw + (uint32_t)Jc * usyw + (uint32_t)Kc;
  uint32_t oi = (uint32_t)i0 * usi + (uint32_t)(Jc - 3) * (uint32_t)d.nz + (uint32_t)(Kc - 1);
  uint32_t op = (uint32_t)(i0 + 3) * usx;
  float win[6];
  float fxlo, fdlo;
  float vA, vC, wA;   // v at (I-1, J), (I-1, J+1) and w at (I-1, J): carried from the previous step
  {
    float m3 = ld32(u, ou - 3 * usx);
#pragma unroll
    for (int t = 0; t < 6; ++t) win[t] = ld32(u, ou + (uint32_t)(t - 2) * usx);
    float mass = 0.5f * (win[1] + win[2]) * r;   // cell I-1
    fxlo = mass * hrecon(m3, win[0], win[1], win[2], win[3], win[4], upwind_pos(mass), weps);
    vA = ld32(v, ov - usxv);
    vC = ld32(v, ov - usxv + usy);
    wA = ld32(w, ow - usxw);
    const float wA1 = ld32(w, ow - usxw + 1);
    fdlo = fd_ops(d, r, r8lo, r8hi, win[2], win[1], vC, vA, wA1, wA);   // fd_cell at I-1
  }
  PlaneRegs pr;
  plane_fetch32(pr, u, op, psl);
  for (int i = i0; i < iend; ++i) {
    const float uc = win[2];
    const float mx = 0.5f * (win[2] + win[3]) * r;   // x-flux at cell I
    const float fxhi = mx * hrecon(win[0], win[1], win[2], win[3], win[4], win[5], upwind_pos(mx), weps);
    // operands at column I, issued before the barrier
    const float vB = ld32(v, ov), vD = ld32(v, ov + usy);
    const float wB = ld32(w, ow), wB1 = ld32(w, ow + 1);
    const float sgv = sgs ? ld32(sgs, oi) : 0.0f;
    const float fdhi = fd_ops(d, r, r8lo, r8hi, win[3], win[2], vD, vB, wB1, wB);   // fd_cell at I
    plane_store(P, pr, psl);  // fetched one plane ahead
    if (i + 1 < iend) plane_fetch32(pr, u, op + usx, psl);
    __syncthreads();
    const float* Pc = P + (ty + 3) * Lz;
    if (act) {
      // z flux at w level K (face K-1): mass averaged over the two columns of the u point;
      // at the surface the stress averaged onto the face (periodic in x, advection.py:43-48)
      float fz = 0.0f;
      if (k > 0) {
        float m = 0.5f * (r8w * wA + r8w * wB);
        fz = m * vrecon(Pc, Lz, K - 1, upwind_pos(m), weps);
      } else if (tau_x) {
        const int il = (i == 0) ? d.nx - 1 : i - 1;
        fz = 0.5f * (tau_x[(long)il * d.ny + j] + tau_x[(long)i * d.ny + j]);
      }
      FZ[ty * nzl + k] = fz;
      float my = 0.5f * (r * vA + r * vB);
      FY[ty * nzl + k] =
          my * hrecon(Pc[K - 3 * Lz], Pc[K - 2 * Lz], Pc[K - Lz], Pc[K], Pc[K + Lz], Pc[K + 2 * Lz], upwind_pos(my), weps);
      if (last_row) {
        const float* Pn = Pc + Lz;
        float m2 = 0.5f * (r * vC + r * vD);
        FY[(ty + 1) * nzl + k] =
            m2 * hrecon(Pn[K - 3 * Lz], Pn[K - 2 * Lz], Pn[K - Lz], Pn[K], Pn[K + Lz], Pn[K + 2 * Lz], upwind_pos(m2), weps);
      }
    }
    __syncthreads();
    if (act) {
      float fzhi = (k == d.nz - 1) ? 0.0f : FZ[ty * nzl + k + 1];
      float conv = -((fxhi - fxlo) * d.idx + (FY[(ty + 1) * nzl + k] - FY[ty * nzl + k]) * d.idy +
                     (fzhi - FZ[ty * nzl + k]) * d.idz);
      float a = (conv + uc * (0.5f * (fdlo + fdhi))) * ir;   // rho8u = 0.5 (r + r) = r
      // + fv: f times the mean of the four v around the u point (v at I-1, I and J, J+1)
      if (cor.mode) a = a + cor.of(((vA + vC) + vB) + vD, cor.vg);
      if (sgs) a = a + sgv;
      st32(ru, oi, a);                                  // ((i ny + j) nz + k)
      if (i == 0) st32(ru, oi + (uint32_t)d.nx * usi, a);
    }
    fxlo = fxhi;
    fdlo = fdhi;
    vA = vB;
    vC = vD;
    wA = wB;
#pragma unroll
    for (int t = 0; t < 5; ++t) win[t] = win[t + 1];
    win[5] = (i + 1 < iend) ? ld32(u, ou + 4 * usx) : 0.0f;
    ou += usx;
    ov += usxv;
    ow += usxw;
    oi += usi;
    op += usx;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------
// k_mom_v: rhs_v = advection_v + sgs_v at v faces 0..ny-1 (face ny written from 0).
// advection.py:77-114, :816-1025.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_mom_v(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                                                 const float* __restrict__ v, const float* __restrict__ w,
                                                 const float* __restrict__ sgs, const float* __restrict__ tau_y,
                                                 CorOpt cor, float* __restrict__ rv) {
  extern __shared__ float lds[];
  const float weps = 1.0e-16f;
  const int nzl = blockDim.x, tz = threadIdx.x, ty = threadIdx.y, nt = nzl * kFTY, tid = ty * nzl + tz;
  const int Lz = d.nz + 2;
  float* P = lds;
  float* FY = P + (kFTY + 6) * Lz;
  float* FZ = FY + (kFTY + 1) * nzl;
  const TileId tb = xcd_tile();
  const int j0 = tb.y * kFTY, i0 = tb.x * kFTX;
  const int k = lane_level(tz, nzl, d.nz), j = j0 + ty;
  const bool act = (k < d.nz) && (j < d.ny);
  const int K = k + 1;
  const int Kc = act ? K : 1, Jc = (j < d.ny) ? j + 3 : 3;
  const bool last_row = (ty == kFTY - 1) || (j == d.ny - 1);
  const float r = rc[Kc];
  const float ir = 1.0f / r;   // one division per thread instead of one per cell and field
  const float r8lo = (Kc <= 0 || Kc >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc - 1] + rc[Kc]);
  const float r8hi = (Kc + 1 >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc] + rc[Kc + 1]);
  const float r8w = 0.5f * (rc[Kc - 1] + rc[Kc]);
  const int iend = min(i0 + kFTX, d.nx);
  const PlaneSlots psl = plane_slots(tid, nt, Lz, j0, d.ny + 6, (int)d.syc());
  const uint32_t usx = (uint32_t)d.sxv(), usy = (uint32_t)d.syc();
  const uint32_t usxu = (uint32_t)d.sxu(), usxw = (uint32_t)d.sxw(), usyw = (uint32_t)d.syw();
  const uint32_t usi = (uint32_t)(d.ny + 1) * (uint32_t)d.nz;
  // offsets of the point at x = I: ov (v), ou (u), ow (w), oi the interior (nx, ny+1, nz)
  // arrays (sgs_v, rv), op the first element of v's x-plane
  uint32_t ov = (uint32_t)(i0 + 3) * usx + (uint32_t)Jc * usy + (uint32_t)Kc;
  uint32_t ou = (uint32_t)(i0 + 3) * usxu + (uint32_t)Jc * usy + (uint32_t)Kc;
  uint32_t ow = (uint32_t)(i0 + 3) * usxw + (uint32_t)Jc * usyw + (uint32_t)Kc;
  uint32_t oi = (uint32_t)i0 * usi + (uint32_t)(Jc - 3) * (uint32_t)d.nz + (uint32_t)(Kc - 1);
  uint32_t op = (uint32_t)(i0 + 3) * usx;
  float win[6];
  float fxlo;
  float uw0, uw1;   // u at (I, J-1) and (I, J): the west corners of the v point
  {
    float m3 = ld32(v, ov - 3 * usx);
#pragma unroll
    for (int t = 0; t < 6; ++t) win[t] = ld32(v, ov + (uint32_t)(t - 2) * usx);
    uw0 = ld32(u, ou - usy);
    uw1 = ld32(u, ou);
    float mass = 0.5f * (r * uw0 + r * uw1);
    fxlo = mass * hrecon(m3, win[0], win[1], win[2], win[3], win[4], upwind_pos(mass), weps);
  }
  PlaneRegs pr;
  plane_fetch32(pr, v, op, psl);
  for (int i = i0; i < iend; ++i) {
    const float vc = win[2];
    const float ue0 = ld32(u, ou + usxu - usy), ue1 = ld32(u, ou + usxu);
    const float mx = 0.5f * (r * ue0 + r * ue1);
    const float fxhi = mx * hrecon(win[0], win[1], win[2], win[3], win[4], win[5], upwind_pos(mx), weps);
    // operands of the flow divergence at rows J-1 and J, the z mass flux and the SGS
    // add-in, issued before the barrier
    const float vS = ld32(v, ov - usy), vN = ld32(v, ov + usy);
    const float wA = ld32(w, ow - usyw), wA1 = ld32(w, ow - usyw + 1);
    const float wB = ld32(w, ow), wB1 = ld32(w, ow + 1);
    const float sgv = sgs ? ld32(sgs, oi) : 0.0f;
    const float fd0 = fd_ops(d, r, r8lo, r8hi, ue0, uw0, vc, vS, wA1, wA);   // fd_cell at (I, J-1)
    const float fd1 = fd_ops(d, r, r8lo, r8hi, ue1, uw1, vN, vc, wB1, wB);   // fd_cell at (I, J)
    plane_store(P, pr, psl);  // fetched one plane ahead
    if (i + 1 < iend) plane_fetch32(pr, v, op + usx, psl);
    __syncthreads();
    const float* Pc = P + (ty + 3) * Lz;
    if (act) {
      float fz = 0.0f;
      if (k > 0) {
        float m = 0.5f * (r8w * wA + r8w * wB);
        fz = m * vrecon(Pc, Lz, K - 1, upwind_pos(m), weps);
      } else if (tau_y) {   // advection.py:83-88
        const int jl = (j == 0) ? d.ny - 1 : j - 1;
        fz = 0.5f * (tau_y[(long)i * d.ny + jl] + tau_y[(long)i * d.ny + j]);
      }
      FZ[ty * nzl + k] = fz;
      float my = 0.5f * (Pc[K - Lz] + Pc[K]) * r;   // y-flux at cell J-1
      FY[ty * nzl + k] =
          my * hrecon(Pc[K - 3 * Lz], Pc[K - 2 * Lz], Pc[K - Lz], Pc[K], Pc[K + Lz], Pc[K + 2 * Lz], upwind_pos(my), weps);
      if (last_row) {
        const float* Pn = Pc + Lz;
        float m2 = 0.5f * (Pn[K - Lz] + Pn[K]) * r;
        FY[(ty + 1) * nzl + k] =
            m2 * hrecon(Pn[K - 3 * Lz], Pn[K - 2 * Lz], Pn[K - Lz], Pn[K], Pn[K + Lz], Pn[K + 2 * Lz], upwind_pos(m2), weps);
      }
    }
    __syncthreads();
    if (act) {
      float fzhi = (k == d.nz - 1) ? 0.0f : FZ[ty * nzl + k + 1];
      float conv = -((fxhi - fxlo) * d.idx + (FY[(ty + 1) * nzl + k] - FY[ty * nzl + k]) * d.idy +
                     (fzhi - FZ[ty * nzl + k]) * d.idz);
      float a = (conv + vc * (0.5f * (fd0 + fd1))) * ir;     // rho8v = 0.5 (r + r) = r
      // - fu: f times the mean of the four u around the v point (u at I, I+1 and J-1, J)
      if (cor.mode) a = a - cor.of(((uw0 + uw1) + ue0) + ue1, cor.ug);
      if (sgs) a = a + sgv;
      st32(rv, oi, a);                                  // ((i (ny+1) + j) nz + k)
      if (j == 0) st32(rv, oi + (uint32_t)d.ny * (uint32_t)d.nz, a);
    }
    fxlo = fxhi;
    uw0 = ue0;
    uw1 = ue1;
#pragma unroll
    for (int t = 0; t < 5; ++t) win[t] = win[t + 1];
    win[5] = (i + 1 < iend) ? ld32(v, ov + 4 * usx) : 0.0f;
    ov += usx;
    ou += usxu;
    ow += usxw;
    oi += usi;
    op += usx;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------
// k_mom_w: rhs_w = advection_w + b8w + sgs_w at w levels 0..nz (advection.py:117-149,
// :1070-1281; pressure_equations.py:162-172).  Lane l owns level l and computes the
// vertical flux at scalar level l+1 (its upper face); the lid level is written by the
// top lane.
// ---------------------------------------------------------------------------------
// 80 VGPRs at most: six waves per SIMD, three 512-thread workgroups per CU (at 93 the
// compiler's choice, two)
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) k_mom_w(Dims d, const float* __restrict__ rc, const float* __restrict__ th0c,
                                                 const float* __restrict__ qv0c, const float* __restrict__ u,
                                                 const float* __restrict__ v, const float* __restrict__ w,
                                                 const float* __restrict__ th, const float* __restrict__ qv,
                                                 const float* __restrict__ sgs, float* __restrict__ rw) {
  extern __shared__ float lds[];
  const float weps = 1.0e-16f;
  const int nzl = blockDim.x, tz = threadIdx.x, ty = threadIdx.y, nt = nzl * kFTY, tid = ty * nzl + tz;
  const int Lw = d.nz + 3;
  float* P = lds;                                   // (kFTY+6) x Lw plane of w
  float* FY = P + (kFTY + 6) * Lw;                  // (kFTY+1) x nzl
  float* FZ = FY + (kFTY + 1) * nzl;                // kFTY x nzl: flux at scalar level l+1
  float* FD = FZ + kFTY * nzl;                      // kFTY x nzl: flow divergence at cell l
  float* BB = FD + kFTY * nzl;                      // kFTY x nzl: buoyancy at cell l
  const TileId tb = xcd_tile();
  const int j0 = tb.y * kFTY, i0 = tb.x * kFTX;
  const int l = lane_level(tz, nzl, d.nz), j = j0 + ty;
  const bool row_ok = j < d.ny;
  const bool lane_ok = l < d.nz;                   // owns level l and cell l
  const bool act = row_ok && lane_ok;
  const int K = l + 1;                              // w padded level (cell l is scalar K = l+1)
  const int Kc = act ? K : 1, Jc = row_ok ? j + 3 : 3;
  const bool last_row = (ty == kFTY - 1) || (j == d.ny - 1);
  const bool interior = act && l >= 1;              // levels 1..nz-1 carry advection
  const float irho8w = 1.0f / (0.5f * (rc[Kc - 1] + rc[Kc]));
  // fd_cell's weights at cell Kc
  const float rK = rc[Kc];
  const float r8lo = (Kc <= 0 || Kc >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc - 1] + rc[Kc]);
  const float r8hi = (Kc + 1 >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc] + rc[Kc + 1]);
  const int iend = min(i0 + kFTX, d.nx);
  const PlaneSlots psl = plane_slots(tid, nt, Lw, j0, d.ny + 5, (int)d.syw());
  const uint32_t usx = (uint32_t)d.sxw(), usyw = (uint32_t)d.syw();
  const uint32_t usxu = (uint32_t)d.sxu(), usy = (uint32_t)d.syc(), usxv = (uint32_t)d.sxv();
  const uint32_t usi = (uint32_t)d.ny * (uint32_t)(d.nz + 1);
  // offsets of the point at x = I: ow (w), ou (u and the scalars, which share u's
  // strides), ov (v), oi the row of the interior (nx, ny, nz+1) arrays (sgs_w, rw), op the
  // first element of w's x-plane
  uint32_t ow = (uint32_t)(i0 + 3) * usx + (uint32_t)Jc * usyw + (uint32_t)Kc;
  uint32_t ou = (uint32_t)(i0 + 3) * usxu + (uint32_t)Jc * usy + (uint32_t)Kc;
  uint32_t ov = (uint32_t)(i0 + 3) * usxv + (uint32_t)Jc * usy + (uint32_t)Kc;
  uint32_t oi = (uint32_t)i0 * usi + (uint32_t)(Jc - 3) * (uint32_t)(d.nz + 1);
  uint32_t op = (uint32_t)(i0 + 3) * usx;
  float win[6];
  float fxlo;
  float uI;   // u at (I, J, K): the west face of the cell, carried from the previous step
  {
    float m3 = ld32(w, ow - 3 * usx);
#pragma unroll
    for (int t = 0; t < 6; ++t) win[t] = ld32(w, ow + (uint32_t)(t - 2) * usx);
    uI = ld32(u, ou);
    float mass = 0.5f * (rc[Kc - 1] * ld32(u, ou - 1) + rc[Kc] * uI);
    fxlo = mass * hrecon(m3, win[0], win[1], win[2], win[3], win[4], upwind_pos(mass), weps);
  }
  PlaneRegs pr;
  plane_fetch32(pr, w, op, psl);
  for (int i = i0; i < iend; ++i) {
    const float wc = win[2];
    const float uE0 = ld32(u, ou + usxu - 1), uE = ld32(u, ou + usxu);   // u at (I+1, J, K-1), (I+1, J, K)
    const float mx = 0.5f * (rc[Kc - 1] * uE0 + rc[Kc] * uE);
    const float fxhi = mx * hrecon(win[0], win[1], win[2], win[3], win[4], win[5], upwind_pos(mx), weps);
    // operands of the y mass fluxes, the flow divergence, the buoyancy and the SGS add-in,
    // issued before the barrier
    const float vA = ld32(v, ov - 1), vB = ld32(v, ov);
    const float vC = ld32(v, ov + usy - 1), vD = ld32(v, ov + usy);
    const float wT = ld32(w, ow + 1);
    const float thc = ld32(th, ou), qvc = ld32(qv, ou);
    const float sgl = sgs ? ld32(sgs, oi + (uint32_t)(Kc - 1)) : 0.0f;
    const float sgt = sgs ? ld32(sgs, oi + (uint32_t)d.nz) : 0.0f;
    plane_store(P, pr, psl);  // fetched one plane ahead
    if (i + 1 < iend) plane_fetch32(pr, w, op + usx, psl);
    if (act) {
      FD[ty * nzl + l] = fd_ops(d, rK, r8lo, r8hi, uE, uI, vD, vB, wT, wc);   // fd_cell at cell l
      BB[ty * nzl + l] = buoy_at(th0c, qv0c, thc, qvc, K);
    }
    __syncthreads();
    const float* Pc = P + (ty + 3) * Lw;
    if (act) {
      // vertical flux at scalar level c = l+1 (face between w[c] and w[c+1])
      const int cz = l + 1;
      float mz = 0.5f * (Pc[cz] + Pc[cz + 1]) * rc[cz];
      FZ[ty * nzl + l] = mz * vrecon(Pc, Lw, cz, upwind_pos(mz), weps);
      if (interior) {
        float my = 0.5f * (rc[K - 1] * vA + rc[K] * vB);
        FY[ty * nzl + l] = my * hrecon(Pc[K - 3 * Lw], Pc[K - 2 * Lw], Pc[K - Lw], Pc[K], Pc[K + Lw], Pc[K + 2 * Lw],
                                       upwind_pos(my), weps);
        if (last_row) {
          const float* Pn = Pc + Lw;
          float m2 = 0.5f * (rc[K - 1] * vC + rc[K] * vD);
          FY[(ty + 1) * nzl + l] = m2 * hrecon(Pn[K - 3 * Lw], Pn[K - 2 * Lw], Pn[K - Lw], Pn[K], Pn[K + Lw],
                                               Pn[K + 2 * Lw], upwind_pos(m2), weps);
        }
      }
    }
    __syncthreads();
    if (act) {
      const float* Bb = BB + ty * nzl;
      const int nz = d.nz;
      float b8;
      float a = 0.0f;
      if (l == 0) {
        b8 = 1.875f * Bb[0] + (-1.25f) * Bb[1] + 0.375f * Bb[2];
      } else {
        b8 = 0.5f * (Bb[l - 1] + Bb[l]);
        float conv = -((fxhi - fxlo) * d.idx + (FY[(ty + 1) * nzl + l] - FY[ty * nzl + l]) * d.idy +
                       (FZ[ty * nzl + l] - FZ[ty * nzl + l - 1]) * d.idz);
        float div8w = 0.5f * (FD[ty * nzl + l - 1] + FD[ty * nzl + l]);
        a = (conv + wc * div8w) * irho8w;
      }
      a = a + b8;
      if (sgs) a = a + sgl;
      st32(rw, oi + (uint32_t)l, a);                    // ((i ny + j) (nz+1) + l)
      if (l == nz - 1) {   // lid level: zero advection, extrapolated buoyancy
        float bt = 1.875f * Bb[nz - 1] + (-1.25f) * Bb[nz - 2] + 0.375f * Bb[nz - 3];
        float at = 0.0f + bt;
        if (sgs) at = at + sgt;
        st32(rw, oi + (uint32_t)nz, at);
      }
    }
    fxlo = fxhi;
    uI = uE;
#pragma unroll
    for (int t = 0; t < 5; ++t) win[t] = win[t + 1];
    win[5] = (i + 1 < iend) ? ld32(w, ow + 4 * usx) : 0.0f;
    ow += usx;
    ou += usxu;
    ov += usxv;
    oi += usi;
    op += usx;
    __syncthreads();
  }
}

inline size_t mom_uv_lds(const Dims& d) {
  int nzl = 64 * ((d.nz + 63) / 64);
  return sizeof(float) * ((size_t)(kFTY + 6) * (d.nz + 2) + (size_t)(2 * kFTY + 1) * nzl);
}
inline size_t mom_w_lds(const Dims& d) {
  int nzl = 64 * ((d.nz + 63) / 64);
  return sizeof(float) * ((size_t)(kFTY + 6) * (d.nz + 3) + (size_t)(4 * kFTY + 1) * nzl);
}

// The entry of a per-scalar array of ScalarSet for a wave-uniform scalar index q
// (a select chain on scalar registers, not an indexed copy of the argument struct).
template <class T>
__device__ __forceinline__ T pick_q(const T (&a)[kMaxScalars], int q) {
  T r = a[0];
#pragma unroll
  for (int t = 1; t < kMaxScalars; ++t) r = (q == t) ? a[t] : r;
  return r;
}

// ---------------------------------------------------------------------------------
// k_stage_scalars: one scalar's WENO advection + Euler update + frozen-SGS add-in per
// workgroup.  The grid's fastest dimension is the scalar, so the workgroups of the S.n
// scalars of one tile are consecutive, run on the same XCD and share the tile's u, v, w
// lines in its L2; each recomputes the tile's mass fluxes and flow divergence.  Holding
// one scalar's x window (not all S.n) keeps the kernel at 80 VGPRs: three workgroups per
// CU, whose barriers and loads overlap (the all-scalars-per-workgroup form needed 119 and
// ran one fewer, 12 % slower).  theta_rho is written by the workgroups of scalar 0
// (theta), which read qv at their cells.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_stage_scalars(Dims d, const float* __restrict__ rc,
                                                        const float* __restrict__ u, const float* __restrict__ v,
                                                        const float* __restrict__ w, ScalarSet S,
                                                        float* __restrict__ trho, float dt) {
  extern __shared__ float lds[];
  const int nzl = blockDim.x;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int nt = nzl * kFTY;
  const int tid = ty * nzl + tz;
  const int Lz = d.nz + 2;
  float* P = lds;
  float* FY = P + (kFTY + 6) * Lz;
  float* FZ = FY + (kFTY + 1) * nzl;

  const TileId tb = xcd_tile();
  const int q = tb.x;
  const int i0 = tb.y * kFTX;
  const int j0 = tb.z * kFTY;
  const float* __restrict__ s = pick_q(S.s, q);
  const float* __restrict__ s0 = pick_q(S.s0, q);
  const float* __restrict__ sgs = pick_q(S.sgs, q);
  float* __restrict__ out = pick_q(S.out, q);
  const float weps = pick_q(S.weps, q);
  const bool heat = (q == 0) && S.heat;
  const bool wtr = (q == 0) && trho;
  const float* __restrict__ sfc = (q == 0) ? S.sfc[0] : ((q == 1) ? S.sfc[1] : nullptr);
  const int k = lane_level(tz, nzl, d.nz), j = j0 + ty;
  const bool act = (k < d.nz) && (j < d.ny);
  const int K = k + 1, J = j + 3;
  const int Kc = act ? K : 1, Jc = (j < d.ny) ? J : 3;
  const bool last_row = (ty == kFTY - 1) || (j == d.ny - 1);
  const long sx = d.sxc(), sy = d.syc();
  const float r = rc[Kc];
  const float ir = 1.0f / r;
  const float r8lo = (Kc <= 0 || Kc >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc - 1] + rc[Kc]);
  const float r8hi = (Kc + 1 >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc] + rc[Kc + 1]);

  const PlaneSlots psl = plane_slots(tid, nt, Lz, j0, d.ny + 5, (int)sy);
  const uint32_t usx = (uint32_t)sx, usy = (uint32_t)sy;
  const uint32_t usxv = (uint32_t)d.sxv(), usxw = (uint32_t)d.sxw(), usyw = (uint32_t)d.syw();
  const uint32_t usi = (uint32_t)d.ny * (uint32_t)d.nz;
  uint32_t oc = (uint32_t)(i0 + 3) * usx + (uint32_t)Jc * usy + (uint32_t)Kc;
  uint32_t ov = (uint32_t)(i0 + 3) * usxv + (uint32_t)Jc * usy + (uint32_t)Kc;
  uint32_t ow = (uint32_t)(i0 + 3) * usxw + (uint32_t)Jc * usyw + (uint32_t)Kc;
  uint32_t oi = (uint32_t)i0 * usi + (uint32_t)(Jc - 3) * (uint32_t)d.nz + (uint32_t)(Kc - 1);
  uint32_t op = (uint32_t)(i0 + 3) * usx;
  float win[6];
  float fxlo;
  const int iend = min(i0 + kFTX, d.nx);
  {
    const float mlo = r * ld32(u, oc);
    const bool pos = upwind_pos(mlo);
    const float m3 = ld32(s, oc - 3 * usx);
#pragma unroll
    for (int t = 0; t < 6; ++t) win[t] = ld32(s, oc + (uint32_t)(t - 2) * usx);
    fxlo = mlo * hrecon(m3, win[0], win[1], win[2], win[3], win[4], pos, weps);
  }
  float ulo = ld32(u, oc);
  PlaneRegs pr;
  plane_fetch32(pr, s, op, psl);

  for (int i = i0; i < iend; ++i) {
    const float uhi = ld32(u, oc + usx);
    const float vlo = ld32(v, ov);
    const float vhi = ld32(v, ov + usy);
    const float wlo = ld32(w, ow);
    const float whi = ld32(w, ow + 1);
    const float fd = (r * uhi - r * ulo) * d.idx + (r * vhi - r * vlo) * d.idy + (r8hi * whi - r8lo * wlo) * d.idz;
    const float mx = r * uhi;
    const float my_lo = r * vlo;
    const float my_hi = r * vhi;
    const float mz_lo = 0.5f * (rc[Kc - 1] + rc[Kc]) * wlo;
    const bool px = upwind_pos(mx), py0 = upwind_pos(my_lo), py1 = upwind_pos(my_hi), pz = upwind_pos(mz_lo);
    const float s0v = ld32(s0, oc);
    const float sgv = sgs ? ld32(sgs, oi) : 0.0f;
    const float hv = heat ? ld32(S.heat, oi) : 0.0f;
    const float qvc = wtr ? ld32(S.s[1], oc) : 0.0f;
    const float bfv = (sfc && act && k == 0) ? sfc[(long)i * d.ny + j] : 0.0f;
    plane_store(P, pr, psl);
    if (i + 1 < iend) plane_fetch32(pr, s, op + usx, psl);
    __syncthreads();
    const float* Pc = P + (ty + 3) * Lz;
    const float qc = Pc[Kc];
    if (act) {
      float fz = (k == 0) ? bfv : mz_lo * vrecon(Pc, Lz, K - 1, pz, weps);
      FZ[ty * nzl + k] = fz;
      float fy = my_lo * hrecon(Pc[Kc - 3 * Lz], Pc[Kc - 2 * Lz], Pc[Kc - Lz], Pc[Kc], Pc[Kc + Lz], Pc[Kc + 2 * Lz],
                                py0, weps);
      FY[ty * nzl + k] = fy;
      if (last_row) {
        const float* Pn = Pc + Lz;
        FY[(ty + 1) * nzl + k] =
            my_hi * hrecon(Pn[Kc - 3 * Lz], Pn[Kc - 2 * Lz], Pn[Kc - Lz], Pn[Kc], Pn[Kc + Lz], Pn[Kc + 2 * Lz], py1, weps);
      }
    }
    float fxhi = mx * hrecon(win[0], win[1], win[2], win[3], win[4], win[5], px, weps);
    __syncthreads();
    if (act) {
      float fzhi = (k == d.nz - 1) ? 0.0f : FZ[ty * nzl + k + 1];
      float conv = -((fxhi - fxlo) * d.idx + (FY[(ty + 1) * nzl + k] - FY[ty * nzl + k]) * d.idy +
                     (fzhi - FZ[ty * nzl + k]) * d.idz);
      float tend = (conv + qc * fd) * ir;
      if (heat) tend = tend + hv;
      float nxt = s0v + tend * dt;
      if (sgs) nxt = nxt + sgv * dt;
      store_c_fast32(d, out, i, j, k, oc, nxt);
      if (wtr) store_c_fast32(d, trho, i, j, k, oc, theta_rho(qc, qvc));
    }
    fxlo = fxhi;
#pragma unroll
    for (int t = 0; t < 5; ++t) win[t] = win[t + 1];
    win[5] = (i + 1 < iend) ? ld32(s, oc + 4 * usx) : 0.0f;
    ulo = uhi;
    oc += usx;
    ov += usxv;
    ow += usxw;
    oi += usi;
    op += usx;
  }
}

inline size_t stage_scalars_lds(const Dims& d) {
  int nzl = 64 * ((d.nz + 63) / 64);
  return sizeof(float) * ((size_t)(kFTY + 6) * (d.nz + 2) + (size_t)(kFTY + 1) * nzl + (size_t)kFTY * nzl);
}

// ---------------------------------------------------------------------------------
// k_stage_scalars_db: the same arithmetic as k_stage_scalars, all scalars per workgroup:
// the planes of all NS scalars at x = I+1 are fetched into registers at the start of
// step I and written to the second of two LDS plane buffers after the step's face
// fluxes, so one x-step takes two barriers for all scalars (k_stage_scalars: two per
// scalar) and the global loads overlap the WENO arithmetic.
// LDS: 2 x NS planes + NS x (2 kFTY + 1) face rows.
// ---------------------------------------------------------------------------------
inline size_t stage_scalars_db_lds(const Dims& d, int ns) {
  int nzl = 64 * ((d.nz + 63) / 64);
  return sizeof(float) * ((size_t)2 * ns * (kFTY + 6) * (d.nz + 2) + (size_t)ns * (2 * kFTY + 1) * nzl);
}

template <int NS>
__global__ void __launch_bounds__(512) k_stage_scalars_db(Dims d, const float* __restrict__ rc,
                                                            const float* __restrict__ u, const float* __restrict__ v,
                                                            const float* __restrict__ w, ScalarSet S,
                                                            float* __restrict__ trho, float dt) {
  extern __shared__ float lds[];
  const int nzl = blockDim.x;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int nt = nzl * kFTY;
  const int tid = ty * nzl + tz;
  const int Lz = d.nz + 2;
  const int PL = (kFTY + 6) * Lz;
  float* Pb = lds;                              // [2][NS][PL]
  float* FYb = Pb + 2 * NS * PL;                // [NS][(kFTY+1) * nzl]
  float* FZb = FYb + NS * (kFTY + 1) * nzl;     // [NS][kFTY * nzl]

  const TileId tb = xcd_tile();
  const int j0 = tb.y * kFTY;
  const int i0 = tb.x * kFTX;
  const int k = lane_level(tz, nzl, d.nz), j = j0 + ty;
  const bool act = (k < d.nz) && (j < d.ny);
  const int K = k + 1;
  const int Kc = act ? K : 1, Jc = (j < d.ny) ? j + 3 : 3;
  const bool last_row = (ty == kFTY - 1) || (j == d.ny - 1);
  const long sx = d.sxc(), sy = d.syc();
  const float r = rc[Kc];
  const float ir = 1.0f / r;   // one division per thread instead of one per cell and field
  const float r8lo = (Kc <= 0 || Kc >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc - 1] + rc[Kc]);
  const float r8hi = (Kc + 1 >= d.nz + 2) ? 0.0f : 0.5f * (rc[Kc] + rc[Kc + 1]);
  const float rz = 0.5f * (rc[Kc - 1] + rc[Kc]);
  const PlaneSlots psl = plane_slots(tid, nt, Lz, j0, d.ny + 5, (int)sy);
  const int iend = min(i0 + kFTX, d.nx);

  float win[NS][6];
  float fxlo[NS];
  {
    const int I = i0 + 3;
    const float mlo = r * u[iu(d, I, Jc, Kc)];
    const bool pos = upwind_pos(mlo);
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const float* s = S.s[q];
      long c = ic(d, I, Jc, Kc);
      float m3 = s[c - 3 * sx];
#pragma unroll
      for (int t = 0; t < 6; ++t) win[q][t] = s[c + (t - 2) * sx];
      fxlo[q] = mlo * hrecon(m3, win[q][0], win[q][1], win[q][2], win[q][3], win[q][4], pos, S.weps[q]);
      plane_load(Pb + q * PL, s + (long)I * sx, psl);
    }
  }
  __syncthreads();
  float ulo = u[iu(d, i0 + 3, Jc, Kc)];
  int b = 0;
  for (int i = i0; i < iend; ++i) {
    const int I = i + 3;
    const long c = ic(d, I, Jc, Kc);
    const bool more = (i + 1 < iend);
    // prefetch the next x-plane of every scalar
    float pre[NS][kSlots];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const float* nxt = S.s[q] + (long)(I + 1) * sx;
#pragma unroll
      for (int t = 0; t < kSlots; ++t) pre[q][t] = (more && t < psl.n) ? nxt[psl.off[t]] : 0.0f;
    }
    const float uhi = u[iu(d, I + 1, Jc, Kc)];
    const float vlo = v[iv(d, I, Jc, Kc)];
    const float vhi = v[iv(d, I, Jc + 1, Kc)];
    const float wlo = w[iw(d, I, Jc, Kc)];
    const float whi = w[iw(d, I, Jc, Kc + 1)];
    const float fd = (r * uhi - r * ulo) * d.idx + (r * vhi - r * vlo) * d.idy + (r8hi * whi - r8lo * wlo) * d.idz;
    const float mx = r * uhi, my_lo = r * vlo, my_hi = r * vhi, mz_lo = rz * wlo;
    const bool px = upwind_pos(mx), py0 = upwind_pos(my_lo), py1 = upwind_pos(my_hi), pz = upwind_pos(mz_lo);
    float qcs[NS], fxhi[NS];
    const BottomFlux bf = bottom_fluxes(S, act && k == 0, (long)i * d.ny + j);
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const float weps = S.weps[q];
      const float* Pc = Pb + (b * NS + q) * PL + (ty + 3) * Lz;
      qcs[q] = Pc[Kc];
      if (act) {
        FZb[(q * kFTY + ty) * nzl + k] =
            (k == 0) ? bf.of(q) : mz_lo * vrecon(Pc, Lz, K - 1, pz, weps);
        FYb[(q * (kFTY + 1) + ty) * nzl + k] =
            my_lo * hrecon(Pc[Kc - 3 * Lz], Pc[Kc - 2 * Lz], Pc[Kc - Lz], Pc[Kc], Pc[Kc + Lz], Pc[Kc + 2 * Lz], py0,
                           weps);
        if (last_row) {
          const float* Pn = Pc + Lz;
          FYb[(q * (kFTY + 1) + ty + 1) * nzl + k] =
              my_hi * hrecon(Pn[Kc - 3 * Lz], Pn[Kc - 2 * Lz], Pn[Kc - Lz], Pn[Kc], Pn[Kc + Lz], Pn[Kc + 2 * Lz], py1,
                             weps);
        }
      }
      fxhi[q] = mx * hrecon(win[q][0], win[q][1], win[q][2], win[q][3], win[q][4], win[q][5], px, weps);
    }
    // next planes into the other buffer (its last readers finished before the previous barrier)
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      float* P = Pb + ((b ^ 1) * NS + q) * PL;
#pragma unroll
      for (int t = 0; t < kSlots; ++t)
        if (more && t < psl.n) P[psl.dst(t)] = pre[q][t];
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const float* FY = FYb + q * (kFTY + 1) * nzl;
        const float* FZ = FZb + q * kFTY * nzl;
        float fzhi = (k == d.nz - 1) ? 0.0f : FZ[ty * nzl + k + 1];
        float conv = -((fxhi[q] - fxlo[q]) * d.idx + (FY[(ty + 1) * nzl + k] - FY[ty * nzl + k]) * d.idy +
                       (fzhi - FZ[ty * nzl + k]) * d.idz);
        float tend = (conv + qcs[q] * fd) * ir;
        if (q == 0 && S.heat) tend = tend + S.heat[((long)i * d.ny + j) * d.nz + k];
        float nxt = S.s0[q][c] + tend * dt;
        if (S.sgs[q]) nxt = nxt + S.sgs[q][((long)i * d.ny + j) * d.nz + k] * dt;
        store_c_fast(d, S.out[q], i, j, k, nxt);
      }
      if (trho) store_c_fast(d, trho, i, j, k, theta_rho(qcs[0], qcs[1]));
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      fxlo[q] = fxhi[q];
#pragma unroll
      for (int t = 0; t < 5; ++t) win[q][t] = win[q][t + 1];
      win[q][5] = more ? S.s[q][c + 4 * sx] : 0.0f;
    }
    ulo = uhi;
    b ^= 1;
    __syncthreads();   // the face rows are rewritten in the next step
  }
}

}  // namespace lex
