// G1: Smagorinsky-Lilly SGS with the N^2 correction (turbulence.py).
//   k_smag_deform : compute_deformation (turbulence.py:298-388)  -> rho*S_ij, deformation
//   k_smag_nm     : compute_nm          (turbulence.py:391-427)  -> N^2
//   k_smag_k      : compute_k           (turbulence.py:430-446)  -> Km, Kh (turb_opt 1)
//   k_smag_k2     : compute_k2          (turbulence.py:449-471)  -> Kmh, Kmv (turb_opt 2)
//   k_smag_uvw / k_smag_scalar(s) : compute_smag (turbulence.py:9-111) and compute_smag2
//                   (:114-295), which differ only in the coefficients: horizontal (kmh, khh)
//                   for tau11, tau22, tau33, tau12 and the x / y scalar fluxes, vertical
//                   (kmv, khv) for tau13, tau23 and the z scalar fluxes.  turb_opt 1 passes
//                   the same array for both.
#pragma once
#include "lex_common.hpp"

namespace lex {

__device__ __forceinline__ float ex3(float a0, float a1, float a2) {
#pragma clang fp contract(off)
  return 1.875f * a0 + (-1.25f) * a1 + 0.375f * a2;
}

// The strain rates from their four operands (shared by every Smagorinsky path, so the
// one-pass kernel and the routine chain round identically).
// e12 = (du/dy + dv/dx) / 2 from u at (F, G), (F, G-1) and v at (F, G), (F-1, G)
__device__ __forceinline__ float e12_val(const Dims& d, float u1, float u0, float v1, float v0) {
#pragma clang fp contract(off)
  float dudy = (u1 - u0) * d.idy;
  float dvdx = (v1 - v0) * d.idx;
  return 0.5f * (dudy + dvdx);
}
// e13 = (du/dz + dw/dx) / 2 from u at levels l+1, l and w at (F, l+1), (F-1, l+1)
__device__ __forceinline__ float e13_val(const Dims& d, float u1, float u0, float w1, float w0) {
#pragma clang fp contract(off)
  float dudz = (u1 - u0) * d.idz;
  float dwdx = (w1 - w0) * d.idx;
  return 0.5f * (dudz + dwdx);
}
// e23 = (dv/dz + dw/dy) / 2 from v at levels l+1, l and w at (G, l+1), (G-1, l+1)
__device__ __forceinline__ float e23_val(const Dims& d, float v1, float v0, float w1, float w0) {
#pragma clang fp contract(off)
  float dvdz = (v1 - v0) * d.idz;
  float dwdy = (w1 - w0) * d.idy;
  return 0.5f * (dvdz + dwdy);
}
// e12 at corner (u-face F, v-face G), scalar level K
__device__ __forceinline__ float e12_at(const Dims& d, const float* __restrict__ u, const float* __restrict__ v,
                                        int F, int G, int K) {
  return e12_val(d, u[iu(d, F, G, K)], u[iu(d, F, G - 1, K)], v[iv(d, F, G, K)], v[iv(d, F - 1, G, K)]);
}
// e13 at (u-face F, row J, interior w level l in 1..nz-1)
__device__ __forceinline__ float e13_at(const Dims& d, const float* __restrict__ u, const float* __restrict__ w,
                                        int F, int J, int l) {
  return e13_val(d, u[iu(d, F, J, l + 1)], u[iu(d, F, J, l)], w[iw(d, F, J, l + 1)], w[iw(d, F - 1, J, l + 1)]);
}
// e23 at (column I, v-face G, interior w level l)
__device__ __forceinline__ float e23_at(const Dims& d, const float* __restrict__ v, const float* __restrict__ w,
                                        int I, int G, int l) {
  return e23_val(d, v[iv(d, I, G, l + 1)], v[iv(d, I, G, l)], w[iw(d, I, G, l + 1)], w[iw(d, I, G - 1, l + 1)]);
}

// Vertical stencil of N^2 at a cell: an interior cell averages the w levels k and k+1;
// the bottom (top) cell extrapolates from levels 1, 2, 3 (nz-1, nz-2, nz-3) with 1.875,
// -1.25, 0.375 (extrapolate_bottom_top).  Written as three weighted levels for every
// lane, so the boundary lanes take no separate path.
struct ZStencil {
  int l[3];
  float c[3];
};
__device__ __forceinline__ ZStencil z_stencil(int k, int nz) {
  ZStencil z;
  if (k == 0) {
    z.l[0] = 1; z.l[1] = 2; z.l[2] = 3;
  } else if (k == nz - 1) {
    z.l[0] = nz - 1; z.l[1] = nz - 2; z.l[2] = nz - 3;
  } else {
    z.l[0] = k; z.l[1] = k + 1; z.l[2] = k;
  }
  const bool edge = (k == 0) || (k == nz - 1);
  z.c[0] = edge ? 1.875f : 0.5f;
  z.c[1] = edge ? -1.25f : 0.5f;
  z.c[2] = edge ? 0.375f : 0.0f;
  return z;
}

struct SmagCell {
  float a11, a22, a33, a12, a13, a23, df;
};

// rho S_ij and the deformation at interior cell (i, j, k) (compute_deformation,
// turbulence.py:298-388)
__device__ __forceinline__ SmagCell smag_deform_cell(const Dims& d, const float* __restrict__ rc,
                                                     const float* __restrict__ u, const float* __restrict__ v,
                                                     const float* __restrict__ w, int i, int j, int k) {
#pragma clang fp contract(off)
  int I = i + 3, J = j + 3, K = k + 1, nz = d.nz;
  float r = rc[K];
  float a11 = r * (u[iu(d, I + 1, J, K)] - u[iu(d, I, J, K)]) * d.idx;
  float a22 = r * (v[iv(d, I, J + 1, K)] - v[iv(d, I, J, K)]) * d.idy;
  float a33 = r * (w[iw(d, I, J, K + 1)] - w[iw(d, I, J, K)]) * d.idz;
  float a12 = r * (0.25f * (e12_at(d, u, v, I, J, K) + e12_at(d, u, v, I + 1, J + 1, K) + e12_at(d, u, v, I, J + 1, K) +
                            e12_at(d, u, v, I + 1, J, K)));
  float a13, a23;
  if (k == 0) {
    float b1 = ex3(e13_at(d, u, w, I + 1, J, 1), e13_at(d, u, w, I + 1, J, 2), e13_at(d, u, w, I + 1, J, 3));
    float b0 = ex3(e13_at(d, u, w, I, J, 1), e13_at(d, u, w, I, J, 2), e13_at(d, u, w, I, J, 3));
    a13 = 0.5f * (b1 + b0);
    float c1 = ex3(e23_at(d, v, w, I, J + 1, 1), e23_at(d, v, w, I, J + 1, 2), e23_at(d, v, w, I, J + 1, 3));
    float c0 = ex3(e23_at(d, v, w, I, J, 1), e23_at(d, v, w, I, J, 2), e23_at(d, v, w, I, J, 3));
    a23 = 0.5f * (c1 + c0);
  } else if (k == nz - 1) {
    float b1 = ex3(e13_at(d, u, w, I + 1, J, nz - 1), e13_at(d, u, w, I + 1, J, nz - 2), e13_at(d, u, w, I + 1, J, nz - 3));
    float b0 = ex3(e13_at(d, u, w, I, J, nz - 1), e13_at(d, u, w, I, J, nz - 2), e13_at(d, u, w, I, J, nz - 3));
    a13 = 0.5f * (b1 + b0);
    float c1 = ex3(e23_at(d, v, w, I, J + 1, nz - 1), e23_at(d, v, w, I, J + 1, nz - 2), e23_at(d, v, w, I, J + 1, nz - 3));
    float c0 = ex3(e23_at(d, v, w, I, J, nz - 1), e23_at(d, v, w, I, J, nz - 2), e23_at(d, v, w, I, J, nz - 3));
    a23 = 0.5f * (c1 + c0);
  } else {
    a13 = 0.25f * (e13_at(d, u, w, I, J, k) + e13_at(d, u, w, I + 1, J, k + 1) + e13_at(d, u, w, I, J, k + 1) +
                   e13_at(d, u, w, I + 1, J, k));
    a23 = 0.25f * (e23_at(d, v, w, I, J, k) + e23_at(d, v, w, I, J + 1, k + 1) + e23_at(d, v, w, I, J, k + 1) +
                   e23_at(d, v, w, I, J + 1, k));
  }
  a13 = r * a13;
  a23 = r * a23;
  float df = (2.0f * (a11 * a11 + a22 * a22 + a33 * a33) + 4.0f * (a12 * a12 + a13 * a13 + a23 * a23)) * (1.0f / (r * r));
  return SmagCell{a11, a22, a33, a12, a13, a23, df};
}

__global__ void k_smag_deform(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                              const float* __restrict__ v, const float* __restrict__ w, float* __restrict__ s11,
                              float* __restrict__ s22, float* __restrict__ s33, float* __restrict__ s12,
                              float* __restrict__ s13, float* __restrict__ s23, float* __restrict__ deform) {
  const TileId tb = xcd_tile();
  int k = tb.x * blockDim.x + threadIdx.x;
  int j = tb.y * blockDim.y + threadIdx.y;
  int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  const SmagCell c = smag_deform_cell(d, rc, u, v, w, i, j, k);
  store_c_fast(d, s11, i, j, k, c.a11);
  store_c_fast(d, s22, i, j, k, c.a22);
  store_c_fast(d, s33, i, j, k, c.a33);
  store_c_fast(d, s12, i, j, k, c.a12);
  store_c_fast(d, s13, i, j, k, c.a13);
  store_c_fast(d, s23, i, j, k, c.a23);
  store_c_fast(d, deform, i, j, k, c.df);
}

// N^2 at a w level from theta and qv of the cells above (1) and below (0)
__device__ __forceinline__ float n28w_val(const Dims& d, float th1, float qv1, float th0, float qv0) {
#pragma clang fp contract(off)
  float a = th1 * (1.0f + (float)kRepsm1 * qv1);
  float b = th0 * (1.0f + (float)kRepsm1 * qv0);
  return logf(a / b) * ((float)kG / d.dz);
}
// N^2 at interior w level l (1..nz-1) of column (I,J)
__device__ __forceinline__ float n28w_at(const Dims& d, const float* __restrict__ th, const float* __restrict__ qv,
                                         long c, int l) {
  return n28w_val(d, th[c + l + 1], qv[c + l + 1], th[c + l], qv[c + l]);
}

// N^2 at interior cell (i, j, k) (compute_nm, turbulence.py:391-427)
__device__ __forceinline__ float smag_n2_cell(const Dims& d, const float* __restrict__ th,
                                              const float* __restrict__ qv, int i, int j, int k) {
#pragma clang fp contract(off)
  long c = ic(d, i + 3, j + 3, 0);
  const ZStencil z = z_stencil(k, d.nz);
  float val = 0.f;
#pragma unroll
  for (int t = 0; t < 3; ++t) val += z.c[t] * n28w_at(d, th, qv, c, z.l[t]);
  return val;
}

__global__ void k_smag_nm(Dims d, const float* __restrict__ th, const float* __restrict__ qv, float* __restrict__ n2) {
  const TileId tb = xcd_tile();
  int k = tb.x * blockDim.x + threadIdx.x;
  int j = tb.y * blockDim.y + threadIdx.y;
  int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  store_c_fast(d, n2, i, j, k, smag_n2_cell(d, th, qv, i, j, k));
}

// compute_k's eddy viscosity from the deformation and N^2 (turbulence.py:430-446)
__device__ __forceinline__ float smag_km(float s2, float n2, float csgs2, float prandtl) {
#pragma clang fp contract(off)
  float x = s2 - n2 * (1.0f / prandtl);
  x = (x > 0.0f) ? x : 0.0f;
  return csgs2 * sqrtf(x);
}

// elementwise over the padded array
__global__ void k_smag_k(long n, const float* __restrict__ s2, const float* __restrict__ n2, float* __restrict__ km,
                         float* __restrict__ kh, float csgs2, float prandtl) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float m = smag_km(s2[t], n2[t], csgs2, prandtl);
  km[t] = m;
  kh[t] = m * (1.0f / prandtl);
}

// compute_k2 (turbulence.py:449-471), elementwise over the padded arrays with prandtl 1:
// Kmh = Cs^2 dx dy Dh with Dh = (2 (s11^2 + s22^2 + s33^2) + 4 s12^2) / rho^2 (no square
// root, as the reference has it); Kmv = Cs^2 Lv^2 sqrt(max(Dv - N^2, 0)) with
// Dv = 4 (s13^2 + s23^2) / rho^2 and 1/Lv^2 = (0.4 z)^-2 + dz^-2, z the cell-centre height
// of the padded level ((K - 1/2) dz, negative in the ghost level below the surface).
// Khh = Kmh and Khv = Kmv (prandtl 1).
// compute_k2's (Kmh, Kmv) at padded level K with density r (turbulence.py:449-471)
__device__ __forceinline__ float2 smag_k2(const Dims& d, float r, int K, float a11, float a22, float a33, float a12,
                                          float a13, float a23, float n2, float csq) {
#pragma clang fp contract(off)
  const float z = ((float)K - 0.5f) * d.dz;
  const float zk = 0.4f * z;
  const float lv2 = 1.0f / (1.0f / (zk * zk) + 1.0f / (d.dz * d.dz));
  const float ir2 = 1.0f / (r * r);
  const float dh = (2.0f * (a11 * a11 + a22 * a22 + a33 * a33) + 4.0f * (a12 * a12)) * ir2;
  const float dv = 4.0f * (a13 * a13 + a23 * a23) * ir2;
  return make_float2(csq * (d.dx * d.dy) * dh, csq * lv2 * sqrtf(fmaxf(dv - n2 / 1.0f, 0.0f)));
}

__global__ void k_smag_k2(Dims d, const float* __restrict__ rc, const float* __restrict__ s11,
                          const float* __restrict__ s22, const float* __restrict__ s33, const float* __restrict__ s12,
                          const float* __restrict__ s13, const float* __restrict__ s23, const float* __restrict__ n2,
                          float* __restrict__ kmh, float* __restrict__ kmv, float csq) {
  const long n = d.nc();
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int K = (int)(t % (d.nz + 2));
  const float2 k2 = smag_k2(d, rc[K], K, s11[t], s22[t], s33[t], s12[t], s13[t], s23[t], n2[t], csq);
  kmh[t] = k2.x;
  kmv[t] = k2.y;
}

// compute_deformation -> compute_nm -> compute_k (aniso 0) or compute_k2 (aniso 1) in one
// pass per interior cell, writing the stresses tau_ij = -2 K s_ij (horizontal K for 11,
// 22, 33, 12, vertical for 13, 23) and the scalar diffusivities (khh; khv when aniso) with
// their ghost images: the same values, operation for operation, as k_smag_deform,
// k_smag_nm, k_smag_k / k_smag_k2 and the products inside k_smag_uvw.
struct SmagTau {
  float* t[6];
  float* khh;
  float* khv;
};
__global__ void k_smag_tau(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                           const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ th,
                           const float* __restrict__ qv, float csgs2, float prandtl, int aniso, float csq, SmagTau o) {
#pragma clang fp contract(off)
  const TileId tb = xcd_tile();
  int k = tb.x * blockDim.x + threadIdx.x;
  int j = tb.y * blockDim.y + threadIdx.y;
  int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  const SmagCell c = smag_deform_cell(d, rc, u, v, w, i, j, k);
  const float n2 = smag_n2_cell(d, th, qv, i, j, k);
  float kmh, kmv, khh, khv;
  if (aniso) {
    const float2 k2 = smag_k2(d, rc[k + 1], k + 1, c.a11, c.a22, c.a33, c.a12, c.a13, c.a23, n2, csq);
    kmh = k2.x;
    kmv = k2.y;
    khh = kmh;
    khv = kmv;
  } else {
    kmh = smag_km(c.df, n2, csgs2, prandtl);
    kmv = kmh;
    khh = kmh * (1.0f / prandtl);
    khv = khh;
  }
  store_c_fast(d, o.t[0], i, j, k, -2.0f * kmh * c.a11);
  store_c_fast(d, o.t[1], i, j, k, -2.0f * kmh * c.a22);
  store_c_fast(d, o.t[2], i, j, k, -2.0f * kmh * c.a33);
  store_c_fast(d, o.t[3], i, j, k, -2.0f * kmh * c.a12);
  store_c_fast(d, o.t[4], i, j, k, -2.0f * kmv * c.a13);
  store_c_fast(d, o.t[5], i, j, k, -2.0f * kmv * c.a23);
  store_c_fast(d, o.khh, i, j, k, khh);
  if (aniso) store_c_fast(d, o.khv, i, j, k, khv);
}

// tau_ij = -2 km s_ij at padded scalar index c: horizontal coefficient for 11, 22, 33, 12,
// vertical for 13, 23; with PRE the arrays hold tau_ij already (k_smag_tau)
#define TAU(arr, c) (PRE ? arr[c] : -2.0f * kmh[c] * arr[c])
#define TAUV(arr, c) (PRE ? arr[c] : -2.0f * kmv[c] * arr[c])

template <bool PRE>
__global__ void k_smag_uvw(Dims d, const float* __restrict__ rc, const float* __restrict__ kmh,
                           const float* __restrict__ kmv, const float* __restrict__ s11, const float* __restrict__ s22,
                           const float* __restrict__ s33, const float* __restrict__ s12, const float* __restrict__ s13,
                           const float* __restrict__ s23, float* __restrict__ su, float* __restrict__ sv,
                           float* __restrict__ sw) {
#pragma clang fp contract(off)   // the stresses are rounded products whether precomputed or not
  const TileId tb = xcd_tile();
  // level k of the three outputs; the lid level nz of sgs_w (zero) is written by the
  // thread of k = nz - 1, so the grid's z extent is nz, not nz + 1 (a workgroup row of
  // 64 lanes for the one extra level would be idle but for one lane)
  int k = tb.x * blockDim.x + threadIdx.x;  // 0..nz-1
  int j = tb.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = tb.z;                             // 0..nx
  if (k >= d.nz || j > d.ny) return;
  const long sx = d.sxc(), sy = d.syc();
  int I = i + 3, J = j + 3, K = k + 1, nz = d.nz;
  // interpolate_scalar2w_0 of tau at w level l: zero at l=0, nz
  auto t13w = [&](long col, int l) -> float {
    if (l <= 0 || l >= nz) return 0.0f;
    return 0.5f * (TAUV(s13, col + l) + TAUV(s13, col + l + 1));
  };
  auto t23w = [&](long col, int l) -> float {
    if (l <= 0 || l >= nz) return 0.0f;
    return 0.5f * (TAUV(s23, col + l) + TAUV(s23, col + l + 1));
  };
  if (k < nz && j < d.ny) {  // sgs_u at u face i (0..nx)
    long c = ic(d, I, J, K);
    float r = rc[K];
    float d11 = (TAU(s11, c) - TAU(s11, c - sx)) * d.idx;
    auto d12y = [&](long cc) { return (TAU(s12, cc) - TAU(s12, cc - sy)) * d.idy; };
    float d12 = 0.25f * (d12y(c - sx) + d12y(c + sy) + d12y(c - sx + sy) + d12y(c));
    auto d13s = [&](long cc) {
      long col = cc - K;
      return (t13w(col, k + 1) - t13w(col, k)) * d.idz;
    };
    float d13 = 0.5f * (d13s(c - sx) + d13s(c));
    su[((long)i * d.ny + j) * nz + k] = -(d11 + d12 + d13) * (1.0f / (0.5f * (r + r)));
  }
  if (k < nz && i < d.nx) {  // sgs_v at v face j (0..ny)
    long c = ic(d, I, J, K);
    float r = rc[K];
    float d22 = (TAU(s22, c) - TAU(s22, c - sy)) * d.idy;
    auto d12x = [&](long cc) { return (TAU(s12, cc) - TAU(s12, cc - sx)) * d.idx; };
    float d21 = 0.25f * (d12x(c - sy) + d12x(c + sx) + d12x(c) + d12x(c + sx - sy));
    auto d23s = [&](long cc) {
      long col = cc - K;
      return (t23w(col, k + 1) - t23w(col, k)) * d.idz;
    };
    float d23 = 0.5f * (d23s(c - sy) + d23s(c));
    sv[((long)i * (d.ny + 1) + j) * nz + k] = -(d21 + d22 + d23) * (1.0f / (0.5f * (r + r)));
  }
  if (i < d.nx && j < d.ny) {  // sgs_w at level k (0..nz)
    float val = 0.0f;
    if (k > 0 && k < nz) {
      int l = k;
      long cu = ic(d, I, J, l + 1), cl = ic(d, I, J, l);  // cells above and below the level
      float d33 = (TAU(s33, cu) - TAU(s33, cl)) * d.idz;
      const float i2dx = 1.0f / (2.0f * d.dx), i2dy = 1.0f / (2.0f * d.dy);
      auto d13x = [&](long cc) { return (TAUV(s13, cc + sx) - TAUV(s13, cc - sx)) * i2dx; };
      auto d23y = [&](long cc) { return (TAUV(s23, cc + sy) - TAUV(s23, cc - sy)) * i2dy; };
      float a = 0.5f * (d13x(cl) + d13x(cu));
      float b = 0.5f * (d23y(cl) + d23y(cu));
      float r8w = 0.5f * (rc[l] + rc[l + 1]);
      val = -(a + b + d33) * (1.0f / r8w);
    }
    sw[((long)i * d.ny + j) * (nz + 1) + k] = val;
    if (k == nz - 1) sw[((long)i * d.ny + j) * (nz + 1) + nz] = 0.0f;   // the lid level
  }
}

// SGS diffusion of a scalar (turbulence.py:80-109, :207-234): khh on the x / y faces,
// khv on the z faces
__device__ float smag_scalar_point(const Dims& d, const float* __restrict__ rc, const float* __restrict__ kh,
                                   const float* __restrict__ khv, const float* __restrict__ s, int i, int j, int k) {
#pragma clang fp contract(off)   // the same rounding in k_smag_scalars and k_smag_fused
  const long sx = d.sxc(), sy = d.syc();
  int I = i + 3, J = j + 3, K = k + 1, nz = d.nz;
  long c = ic(d, I, J, K);
  float r = rc[K];
  float r8u = 0.5f * (r + r);
  auto tx = [&](long cc) { return -r8u * (0.5f * (kh[cc] + kh[cc - sx])) * ((s[cc] - s[cc - sx]) * d.idx); };
  auto ty = [&](long cc) { return -r8u * (0.5f * (kh[cc] + kh[cc - sy])) * ((s[cc] - s[cc - sy]) * d.idy); };
  // w level l between padded levels l and l+1; zero at l=0, nz
  auto tz = [&](int l) -> float {
    if (l <= 0 || l >= nz) return 0.0f;
    long cl = c - K + l;
    float r8w = 0.5f * (rc[l] + rc[l + 1]);
    return -r8w * (0.5f * (khv[cl] + khv[cl + 1])) * ((s[cl + 1] - s[cl]) * d.idz);
  };
  float div = (tx(c + sx) - tx(c)) * d.idx + (ty(c + sy) - ty(c)) * d.idy + (tz(k + 1) - tz(k)) * d.idz;
  return -div * (1.0f / r);
}

__global__ void k_smag_scalar(Dims d, const float* __restrict__ rc, const float* __restrict__ kh,
                              const float* __restrict__ khv, const float* __restrict__ s, float* __restrict__ out) {
  const TileId tb = xcd_tile();
  int k = tb.x * blockDim.x + threadIdx.x;
  int j = tb.y * blockDim.y + threadIdx.y;
  int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  out[((long)i * d.ny + j) * d.nz + k] = smag_scalar_point(d, rc, kh, khv, s, i, j, k);
}

// All scalars' SGS diffusion in one pass: the diffusivity faces (kh averaged to the six
// faces of the cell, times rho at the face) are computed once and applied to every scalar.
struct SgsScalars {
  const float* s[10];
  float* out[10];
  int n;
};

__global__ void k_smag_scalars(Dims d, const float* __restrict__ rc, const float* __restrict__ kh,
                               const float* __restrict__ khv, SgsScalars S) {
#pragma clang fp contract(off)
  const TileId tb = xcd_tile();
  int k = tb.x * blockDim.x + threadIdx.x;
  int j = tb.y * blockDim.y + threadIdx.y;
  int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  const long sx = d.sxc(), sy = d.syc();
  const int I = i + 3, J = j + 3, K = k + 1, nz = d.nz;
  const long c = ic(d, I, J, K);
  const float r = rc[K];
  const float r8u = 0.5f * (r + r);
  // -rho_face * kh_face per face (same association as smag_scalar_point)
  const float kx0 = -r8u * (0.5f * (kh[c] + kh[c - sx])), kx1 = -r8u * (0.5f * (kh[c + sx] + kh[c]));
  const float ky0 = -r8u * (0.5f * (kh[c] + kh[c - sy])), ky1 = -r8u * (0.5f * (kh[c + sy] + kh[c]));
  // vertical faces: zero flux at the surface and the lid (levels c-1 and c+1 exist in the
  // padded column, so the loads need no branch)
  const bool zlo = k > 0, zhi = k + 1 < nz;
  const float kz0 = zlo ? -(0.5f * (rc[K - 1] + rc[K])) * (0.5f * (khv[c - 1] + khv[c])) : 0.0f;
  const float kz1 = zhi ? -(0.5f * (rc[K] + rc[K + 1])) * (0.5f * (khv[c] + khv[c + 1])) : 0.0f;
  const long o = ((long)i * d.ny + j) * d.nz + k;
  for (int q = 0; q < S.n; ++q) {
    const float* sc = S.s[q];
    const float s0 = sc[c], sxp = sc[c + sx], sxm = sc[c - sx], syp = sc[c + sy], sym = sc[c - sy];
    const float szp = sc[c + 1], szm = sc[c - 1];
    float tx1 = kx1 * ((sxp - s0) * d.idx), tx0 = kx0 * ((s0 - sxm) * d.idx);
    float ty1 = ky1 * ((syp - s0) * d.idy), ty0 = ky0 * ((s0 - sym) * d.idy);
    float tz1 = zhi ? kz1 * ((szp - s0) * d.idz) : 0.0f;
    float tz0 = zlo ? kz0 * ((s0 - szm) * d.idz) : 0.0f;
    float div = (tx1 - tx0) * d.idx + (ty1 - ty0) * d.idy + (tz1 - tz0) * d.idz;
    S.out[q][o] = -div * (1.0f / r);
  }
}

// ---------------------------------------------------------------------------------
// k_smag_fused: the whole frozen-SGS evaluation of a step in one pass, with the stresses
// and diffusivities kept on chip: compute_deformation -> compute_nm -> compute_k /
// compute_k2 -> compute_smag (compute_smag2) for u, v, w and every scalar.  The same
// operations in the same order as k_smag_tau + k_smag_uvw<true> + k_smag_scalars (and so
// as the routine-by-routine chain), bit for bit; the stresses are never written to HBM.
//
// A workgroup owns whole z columns of TYT = blockDim.y rows j0-1 .. j0+TYT-2 and marches
// x from x0-1 to x1.  At plane p it evaluates the stresses of its TYT rows (phase 1):
//   - the e12 corner values and the e13 edge values of the upper x face are carried to
//     the next plane as its lower face (two evaluations of each per cell instead of four);
//   - N^2 at the cell's lower and upper w levels (two logarithms instead of three); the
//     surface and lid cells take the extrapolation levels from the lanes above / below;
//   - t11 and the diffusivity stay in registers for the x differences; t12 and t13 go to
//     a four-plane LDS ring (x, y and z neighbours), t22, t23, t33 and the diffusivities
//     to a three-plane ring.
// Then (phase 2) its TYT-2 inner rows form the stress divergence on the u, v, w points
// of plane p-1 and the scalars' SGS diffusion there.  One barrier per plane.  The lanes
// above / below the surface and lid cells take values from within their wave, so nz mod
// 64 is not 1 or 2 (the host checks, and otherwise launches the three-kernel form).
//
// Workgroups of at most 768 threads held to 80 VGPRs (five spilled, a few dwords of
// private memory): two workgroups and 24 waves per CU, where 1024-thread workgroups at
// 87 VGPRs ran one and 16; the pass waits on memory and the extra waves cover it (11 %
// faster at 256x256x128 despite 12 % more recomputed halo rows).
// ---------------------------------------------------------------------------------
inline int smag_fused_rows(const Dims& d) {
  const int nzl = 64 * ((d.nz + 63) / 64);
  const int r = 768 / nzl;
  return r > 8 ? 8 : r;
}
inline bool smag_fused_ok(const Dims& d) {
  const int m = d.nz % 64;
  return d.nz >= 4 && m != 1 && m != 2 && smag_fused_rows(d) >= 3;
}
inline size_t smag_fused_lds(const Dims& d, bool aniso) {
  const int nzl = 64 * ((d.nz + 63) / 64);
  return sizeof(float) * (size_t)smag_fused_rows(d) * nzl * (aniso ? 23 : 20);
}
constexpr int kSmagAx = 16;   // x cells per workgroup
inline dim3 smag_fused_grid(const Dims& d) {
  const int ty = smag_fused_rows(d) - 2;
  return dim3((d.ny + ty - 1) / ty, (d.nx + kSmagAx - 1) / kSmagAx, 1);
}

template <bool ANISO>
__global__ void __launch_bounds__(768, 6) k_smag_fused(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                                                      const float* __restrict__ v, const float* __restrict__ w,
                                                      const float* __restrict__ th, const float* __restrict__ qv,
                                                      float csgs2, float prandtl, float csq, SgsScalars S,
                                                      float* __restrict__ su, float* __restrict__ sv,
                                                      float* __restrict__ sw) {
#pragma clang fp contract(off)
  extern __shared__ float lds[];
  const int nzl = blockDim.x, TYT = blockDim.y;
  const int PL = TYT * nzl;
  // rings deep enough for one barrier per plane: while the slowest thread forms the
  // divergences of plane i (t12, t13 of planes i-1..i+1; the rest of plane i), the
  // fastest may already store plane i+2
  float* L12 = lds;             // [4][TYT][nzl]
  float* L13 = L12 + 4 * PL;    // [4]
  float* L22 = L13 + 4 * PL;    // [3]
  float* L23 = L22 + 3 * PL;    // [3]
  float* L33 = L23 + 3 * PL;    // [3]
  float* LKH = L33 + 3 * PL;    // [3] khh
  float* LKV = LKH + 3 * PL;    // [3] khv (ANISO)
  const int lane = threadIdx.x, row = threadIdx.y;
  const TileId tb = xcd_tile();
  const int j0 = tb.x * (TYT - 2);
  const int x0 = tb.y * kSmagAx, x1 = min(x0 + kSmagAx, d.nx);
  const int nz = d.nz;
  const int j = j0 - 1 + row;                 // -1 .. j0+TYT-2 (rows past ny are never read)
  // halo rows and planes are evaluated at their periodic images (the interior cell's own
  // inputs), so their stresses equal the images the routine chain copies into the ghosts
  // whatever the input ghosts hold
  const int jw = wrap_cell(min(j, d.ny), d.ny);
  const int kc = min(lane, nz - 1);           // idle lanes compute level nz-1
  const int J = jw + 3, K = kc + 1;
  const float r = rc[K];
  const float ir2 = 1.0f / (r * r), ipr = 1.0f / prandtl;
  const int me = row * nzl + lane;            // LDS index of this thread's (row, level)
  const bool surf = (kc == 0), lid = (kc == nz - 1);
  // Addressing: the x plane is the same for the whole workgroup, so every field is read
  // through a wave-uniform pointer to its plane (scalar registers) plus a 32-bit lane
  // offset: lc for the scalar, u and v layouts (row stride nz + 2), lw for w (nz + 3).
  const long sxc = d.sxc(), sxu = d.sxu(), sxv = d.sxv(), sxw = d.sxw(), syc = d.syc(), syw = d.syw();
  const uint32_t lc0 = (uint32_t)(J * (nz + 2) + K), lw0 = (uint32_t)(J * (nz + 3) + K);

  // carried lower-face values of the next plane: e12 at (I, J) and (I, J+1), e13 at
  // (I, J) on w levels k and k+1
  float c00 = 0.f, c01 = 0.f, e13lo_k = 0.f, e13lo_k1 = 0.f;
  // t11 and khh at planes p-2, p-1 (the output plane), p
  float t11_a = 0.f, t11_b = 0.f, kh_a = 0.f, kh_b = 0.f, kh_n = 0.f, t11_n = 0.f;

  // ---- phase 1: the stresses and diffusivities of cell (p, jw, kc)
  // first: the previous plane was not the x-neighbour in memory (start of the march, or
  // a wrap to the periodic image), so the lower-face values are evaluated, not carried
  auto phase1 = [&](int p, bool first) {
#pragma clang fp contract(off)
    const int I = wrap_cell(p, d.nx) + 3;
    const uint32_t oc = lane_off(lc0), ow = lane_off(lw0);
    const float* U0 = u + I * sxu;            // u planes I, I+1
    const float* U1 = U0 + sxu;
    const float* V0 = v + I * sxv;            // v planes I-1, I, I+1
    const float* Vm = V0 - sxv;
    const float* V1 = V0 + sxv;
    const float* W0 = w + I * sxw;            // w planes I-1, I, I+1
    const float* Wm = W0 - sxw;
    const float* W1 = W0 + sxw;
    const float* TH = th + I * sxc;
    const float* QV = qv + I * sxc;
    const float a11 = r * (ld32(U1, oc) - ld32(U0, oc)) * d.idx;
    const float a22 = r * (ld32(V0 + syc, oc) - ld32(V0, oc)) * d.idy;
    const float a33 = r * (ld32(W0 + 1, ow) - ld32(W0, ow)) * d.idz;
    // e12 at corner (F, G) = e12_val(u(F, G), u(F, G-1), v(F, G), v(F-1, G))
    if (first) {
      c00 = e12_val(d, ld32(U0, oc), ld32(U0 - syc, oc), ld32(V0, oc), ld32(Vm, oc));
      c01 = e12_val(d, ld32(U0 + syc, oc), ld32(U0, oc), ld32(V0 + syc, oc), ld32(Vm + syc, oc));
      e13lo_k = e13_val(d, ld32(U0, oc), ld32(U0 - 1, oc), ld32(W0, ow), ld32(Wm, ow));
      e13lo_k1 = e13_val(d, ld32(U0 + 1, oc), ld32(U0, oc), ld32(W0 + 1, ow), ld32(Wm + 1, ow));
    }
    const float c11 = e12_val(d, ld32(U1 + syc, oc), ld32(U1, oc), ld32(V1 + syc, oc), ld32(V0 + syc, oc));
    const float c10 = e12_val(d, ld32(U1, oc), ld32(U1 - syc, oc), ld32(V1, oc), ld32(V0, oc));
    const float a12 = r * (0.25f * (c00 + c11 + c01 + c10));
    // e13 on the upper x face (u plane I+1, w planes I+1 and I), levels k and k+1
    const float e13hi_k = e13_val(d, ld32(U1, oc), ld32(U1 - 1, oc), ld32(W1, ow), ld32(W0, ow));
    const float e13hi_k1 = e13_val(d, ld32(U1 + 1, oc), ld32(U1, oc), ld32(W1 + 1, ow), ld32(W0 + 1, ow));
    // e23 on the y faces J (f0) and J+1 (f1), levels k and k+1:
    // e23_val(v(G, l+1), v(G, l), w(G, l+1), w(G-1, l+1))
    const float f0k = e23_val(d, ld32(V0, oc), ld32(V0 - 1, oc), ld32(W0, ow), ld32(W0 - syw, ow));
    const float f1k1 = e23_val(d, ld32(V0 + syc + 1, oc), ld32(V0 + syc, oc), ld32(W0 + syw + 1, ow), ld32(W0 + 1, ow));
    const float f0k1 = e23_val(d, ld32(V0 + 1, oc), ld32(V0, oc), ld32(W0 + 1, ow), ld32(W0 - syw + 1, ow));
    const float f1k = e23_val(d, ld32(V0 + syc, oc), ld32(V0 + syc - 1, oc), ld32(W0 + syw, ow), ld32(W0, ow));
    // surface cell: levels 1, 2, 3 = (k+1) of lanes 0, 1, 2; lid cell: levels nz-1,
    // nz-2, nz-3 = k of lanes nz-1, nz-2, nz-3 (all within the wave)
    const float b1s2 = __shfl_down(e13hi_k1, 1, 64), b1s3 = __shfl_down(e13hi_k1, 2, 64);
    const float b0s2 = __shfl_down(e13lo_k1, 1, 64), b0s3 = __shfl_down(e13lo_k1, 2, 64);
    const float c1s2 = __shfl_down(f1k1, 1, 64), c1s3 = __shfl_down(f1k1, 2, 64);
    const float c0s2 = __shfl_down(f0k1, 1, 64), c0s3 = __shfl_down(f0k1, 2, 64);
    const float b1l2 = __shfl_up(e13hi_k, 1, 64), b1l3 = __shfl_up(e13hi_k, 2, 64);
    const float b0l2 = __shfl_up(e13lo_k, 1, 64), b0l3 = __shfl_up(e13lo_k, 2, 64);
    const float c1l2 = __shfl_up(f1k, 1, 64), c1l3 = __shfl_up(f1k, 2, 64);
    const float c0l2 = __shfl_up(f0k, 1, 64), c0l3 = __shfl_up(f0k, 2, 64);
    float a13, a23;
    if (surf) {
      const float b1 = ex3(e13hi_k1, b1s2, b1s3), b0 = ex3(e13lo_k1, b0s2, b0s3);
      a13 = 0.5f * (b1 + b0);
      const float c1 = ex3(f1k1, c1s2, c1s3), c0 = ex3(f0k1, c0s2, c0s3);
      a23 = 0.5f * (c1 + c0);
    } else if (lid) {
      const float b1 = ex3(e13hi_k, b1l2, b1l3), b0 = ex3(e13lo_k, b0l2, b0l3);
      a13 = 0.5f * (b1 + b0);
      const float c1 = ex3(f1k, c1l2, c1l3), c0 = ex3(f0k, c0l2, c0l3);
      a23 = 0.5f * (c1 + c0);
    } else {
      a13 = 0.25f * (e13lo_k + e13hi_k1 + e13lo_k1 + e13hi_k);
      a23 = 0.25f * (f0k + f1k1 + f0k1 + f1k);
    }
    a13 = r * a13;
    a23 = r * a23;
    // N^2 at w levels k (cells k-1, k) and k+1 (cells k, k+1) of this column
    const float thc = ld32(TH, oc), qvc = ld32(QV, oc);
    const float nA = n28w_val(d, thc, qvc, ld32(TH - 1, oc), ld32(QV - 1, oc));
    const float nB = n28w_val(d, ld32(TH + 1, oc), ld32(QV + 1, oc), thc, qvc);
    const float nBs2 = __shfl_down(nB, 1, 64), nBs3 = __shfl_down(nB, 2, 64);
    const float nAl2 = __shfl_up(nA, 1, 64), nAl3 = __shfl_up(nA, 2, 64);
    float n2 = 0.f;
    if (surf) {
      n2 = n2 + 1.875f * nB;
      n2 = n2 + (-1.25f) * nBs2;
      n2 = n2 + 0.375f * nBs3;
    } else if (lid) {
      n2 = n2 + 1.875f * nA;
      n2 = n2 + (-1.25f) * nAl2;
      n2 = n2 + 0.375f * nAl3;
    } else {
      n2 = n2 + 0.5f * nA;
      n2 = n2 + 0.5f * nB;
      n2 = n2 + 0.0f * nA;
    }
    float kmh, kmv, khh, khv;
    if (ANISO) {
      const float2 k2 = smag_k2(d, r, K, a11, a22, a33, a12, a13, a23, n2, csq);
      kmh = k2.x;
      kmv = k2.y;
      khh = kmh;
      khv = kmv;
    } else {
      const float df = (2.0f * (a11 * a11 + a22 * a22 + a33 * a33) + 4.0f * (a12 * a12 + a13 * a13 + a23 * a23)) * ir2;
      kmh = smag_km(df, n2, csgs2, prandtl);
      kmv = kmh;
      khh = kmh * ipr;
      khv = khh;
    }
    const int s4 = (p + 4) % 4, s3 = (p + 3) % 3;   // p >= -1
    L12[s4 * PL + me] = -2.0f * kmh * a12;
    L13[s4 * PL + me] = -2.0f * kmv * a13;
    L22[s3 * PL + me] = -2.0f * kmh * a22;
    L23[s3 * PL + me] = -2.0f * kmv * a23;
    L33[s3 * PL + me] = -2.0f * kmh * a33;
    LKH[s3 * PL + me] = khh;
    if (ANISO) LKV[s3 * PL + me] = khv;
    t11_n = -2.0f * kmh * a11;
    kh_n = khh;
    // the upper face becomes the next plane's lower face
    c00 = c10;
    c01 = c11;
    e13lo_k = e13hi_k;
    e13lo_k1 = e13hi_k1;
  };

  phase1(x0 - 1, true);
  t11_a = t11_n;
  kh_a = kh_n;
  phase1(x0, x0 == 0);
  t11_b = t11_n;
  kh_b = kh_n;
  const bool out_row = row >= 1 && row <= TYT - 2 && j < d.ny && lane < nz;
  const int k = kc;
  const float irr = 1.0f / (0.5f * (r + r)), ir = 1.0f / r;
  const float ir8w = 1.0f / (0.5f * (rc[k] + rc[k + 1]));
  const float i2dx = 1.0f / (2.0f * d.dx), i2dy = 1.0f / (2.0f * d.dy);
  const uint32_t ui = (uint32_t)(jw * nz + k), vi = ui, wi = (uint32_t)(jw * (nz + 1) + k);
  const long psu = (long)d.ny * nz, psv = (long)(d.ny + 1) * nz, psw = (long)d.ny * (nz + 1);
  for (int i = x0; i < x1; ++i) {
    phase1(i + 1, i + 1 == d.nx);
    __syncthreads();
    if (out_row) {
      const int sm = (i + 3) % 4, sc = (i + 4) % 4, sp = (i + 5) % 4, c2 = (i + 3) % 3;
      const float* T12m = L12 + sm * PL;
      const float* T12c = L12 + sc * PL;
      const float* T12p = L12 + sp * PL;
      const float* T13m = L13 + sm * PL;
      const float* T13c = L13 + sc * PL;
      const float* T13p = L13 + sp * PL;
      const float* T22 = L22 + c2 * PL;
      const float* T23 = L23 + c2 * PL;
      const float* T33 = L33 + c2 * PL;
      const float* KH = LKH + c2 * PL;
      const float* KV = ANISO ? LKV + c2 * PL : KH;
      const int up = me + nzl, dn = me - nzl;                     // rows j+1, j-1
      const int zm = (k > 0) ? me - 1 : me, zp = (k + 1 < nz) ? me + 1 : me;   // levels k-1, k+1 (clamped)
      // sgs_u at u face i (cells i-1, i of row j), k_smag_uvw
      {
        const float d11 = (t11_b - t11_a) * d.idx;
        auto d12y = [&](const float* T, int a, int b) {
#pragma clang fp contract(off)
          return (T[a] - T[b]) * d.idy;
        };
        const float d12 = 0.25f * (d12y(T12m, me, dn) + d12y(T12c, up, me) + d12y(T12m, up, me) + d12y(T12c, me, dn));
        auto t13w_hi = [&](const float* T) { return (k + 1 >= nz) ? 0.0f : 0.5f * (T[me] + T[zp]); };
        auto t13w_lo = [&](const float* T) { return (k <= 0) ? 0.0f : 0.5f * (T[zm] + T[me]); };
        const float d13 = 0.5f * ((t13w_hi(T13m) - t13w_lo(T13m)) * d.idz + (t13w_hi(T13c) - t13w_lo(T13c)) * d.idz);
        const float val = -(d11 + d12 + d13) * irr;
        st32(su + i * psu, ui, val);
        if (i == 0) st32(su + d.nx * psu, ui, val);   // face nx, the periodic duplicate of face 0
      }
      // sgs_v at v face j (rows j-1, j of plane i)
      {
        const float d22 = (T22[me] - T22[dn]) * d.idy;
        auto d12x = [&](const float* A, const float* B, int a) {
#pragma clang fp contract(off)
          return (A[a] - B[a]) * d.idx;
        };
        const float d21 = 0.25f * (d12x(T12c, T12m, dn) + d12x(T12p, T12c, me) + d12x(T12c, T12m, me) + d12x(T12p, T12c, dn));
        auto t23w_hi = [&](int a) { return (k + 1 >= nz) ? 0.0f : 0.5f * (T23[a] + T23[a + (zp - me)]); };
        auto t23w_lo = [&](int a) { return (k <= 0) ? 0.0f : 0.5f * (T23[a + (zm - me)] + T23[a]); };
        const float d23 = 0.5f * ((t23w_hi(dn) - t23w_lo(dn)) * d.idz + (t23w_hi(me) - t23w_lo(me)) * d.idz);
        const float val = -(d21 + d22 + d23) * irr;
        st32(sv + i * psv, vi, val);
        if (j == 0) st32(sv + i * psv + (long)d.ny * nz, vi, val);   // face ny
      }
      // sgs_w at level k of cell column (i, j); zero at the surface and the lid
      {
        float val = 0.0f;
        if (k > 0) {
          const float d33 = (T33[me] - T33[zm]) * d.idz;
          const float a = 0.5f * ((T13p[zm] - T13m[zm]) * i2dx + (T13p[me] - T13m[me]) * i2dx);
          const float b = 0.5f * ((T23[zm + nzl] - T23[zm - nzl]) * i2dy + (T23[up] - T23[dn]) * i2dy);
          val = -(a + b + d33) * ir8w;
        }
        float* W = sw + i * psw;
        st32(W, wi, val);
        if (k == nz - 1) st32(W + 1, wi, 0.0f);
      }
      // SGS diffusion of every scalar at cell (i, j, k), k_smag_scalars
      {
        const uint32_t oc = lc0;
        const float r8u = 0.5f * (r + r);
        const float kx0 = -r8u * (0.5f * (kh_b + kh_a)), kx1 = -r8u * (0.5f * (kh_n + kh_b));
        const float ky0 = -r8u * (0.5f * (KH[me] + KH[dn])), ky1 = -r8u * (0.5f * (KH[up] + KH[me]));
        const bool zlo = k > 0, zhi = k + 1 < nz;
        const float kz0 = zlo ? -(0.5f * (rc[K - 1] + rc[K])) * (0.5f * (KV[zm] + KV[me])) : 0.0f;
        const float kz1 = zhi ? -(0.5f * (rc[K] + rc[K + 1])) * (0.5f * (KV[me] + KV[zp])) : 0.0f;
        for (int q = 0; q < S.n; ++q) {
          const float* P = S.s[q] + (i + 3) * sxc;
          const float s0 = ld32(P, oc), sxp = ld32(P + sxc, oc), sxm = ld32(P - sxc, oc);
          const float syp = ld32(P + syc, oc), sym = ld32(P - syc, oc);
          const float szp = ld32(P + 1, oc), szm = ld32(P - 1, oc);
          float tx1 = kx1 * ((sxp - s0) * d.idx), tx0 = kx0 * ((s0 - sxm) * d.idx);
          float ty1 = ky1 * ((syp - s0) * d.idy), ty0 = ky0 * ((s0 - sym) * d.idy);
          float tz1 = zhi ? kz1 * ((szp - s0) * d.idz) : 0.0f;
          float tz0 = zlo ? kz0 * ((s0 - szm) * d.idz) : 0.0f;
          float div = (tx1 - tx0) * d.idx + (ty1 - ty0) * d.idy + (tz1 - tz0) * d.idz;
          st32(S.out[q] + i * psu, ui, -div * ir);
        }
      }
    }
    t11_a = t11_b;
    t11_b = t11_n;
    kh_a = kh_b;
    kh_b = kh_n;
  }
}

#undef TAU
#undef TAUV
}  // namespace lex
