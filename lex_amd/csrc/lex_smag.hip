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

// e12 at corner (u-face F, v-face G), scalar level K
__device__ __forceinline__ float e12_at(const Dims& d, const float* __restrict__ u, const float* __restrict__ v,
                                        int F, int G, int K) {
#pragma clang fp contract(off)
  float dudy = (u[iu(d, F, G, K)] - u[iu(d, F, G - 1, K)]) * d.idy;
  float dvdx = (v[iv(d, F, G, K)] - v[iv(d, F - 1, G, K)]) * d.idx;
  return 0.5f * (dudy + dvdx);
}
// e13 at (u-face F, row J, interior w level l in 1..nz-1)
__device__ __forceinline__ float e13_at(const Dims& d, const float* __restrict__ u, const float* __restrict__ w,
                                        int F, int J, int l) {
#pragma clang fp contract(off)
  float dudz = (u[iu(d, F, J, l + 1)] - u[iu(d, F, J, l)]) * d.idz;
  float dwdx = (w[iw(d, F, J, l + 1)] - w[iw(d, F - 1, J, l + 1)]) * d.idx;
  return 0.5f * (dudz + dwdx);
}
// e23 at (column I, v-face G, interior w level l)
__device__ __forceinline__ float e23_at(const Dims& d, const float* __restrict__ v, const float* __restrict__ w,
                                        int I, int G, int l) {
#pragma clang fp contract(off)
  float dvdz = (v[iv(d, I, G, l + 1)] - v[iv(d, I, G, l)]) * d.idz;
  float dwdy = (w[iw(d, I, G, l + 1)] - w[iw(d, I, G - 1, l + 1)]) * d.idy;
  return 0.5f * (dvdz + dwdy);
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
  float df = (2.0f * (a11 * a11 + a22 * a22 + a33 * a33) + 4.0f * (a12 * a12 + a13 * a13 + a23 * a23)) / (r * r);
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

// N^2 at interior w level l (1..nz-1) of column (I,J)
__device__ __forceinline__ float n28w_at(const Dims& d, const float* __restrict__ th, const float* __restrict__ qv,
                                         long c, int l) {
#pragma clang fp contract(off)
  float a = th[c + l + 1] * (1.0f + (float)kRepsm1 * qv[c + l + 1]);
  float b = th[c + l] * (1.0f + (float)kRepsm1 * qv[c + l]);
  return logf(a / b) / d.dz * (float)kG;
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
  float x = s2 - n2 / prandtl;
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
  kh[t] = m / prandtl;
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
  const float dh = (2.0f * (a11 * a11 + a22 * a22 + a33 * a33) + 4.0f * (a12 * a12)) / (r * r);
  const float dv = 4.0f * (a13 * a13 + a23 * a23) / (r * r);
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
    khh = kmh / prandtl;
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
    su[((long)i * d.ny + j) * nz + k] = -(d11 + d12 + d13) / (0.5f * (r + r));
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
    sv[((long)i * (d.ny + 1) + j) * nz + k] = -(d21 + d22 + d23) / (0.5f * (r + r));
  }
  if (i < d.nx && j < d.ny) {  // sgs_w at level k (0..nz)
    float val = 0.0f;
    if (k > 0 && k < nz) {
      int l = k;
      long cu = ic(d, I, J, l + 1), cl = ic(d, I, J, l);  // cells above and below the level
      float d33 = (TAU(s33, cu) - TAU(s33, cl)) * d.idz;
      auto d13x = [&](long cc) { return (TAUV(s13, cc + sx) - TAUV(s13, cc - sx)) / (2.0f * d.dx); };
      auto d23y = [&](long cc) { return (TAUV(s23, cc + sy) - TAUV(s23, cc - sy)) / (2.0f * d.dy); };
      float a = 0.5f * (d13x(cl) + d13x(cu));
      float b = 0.5f * (d23y(cl) + d23y(cu));
      float r8w = 0.5f * (rc[l] + rc[l + 1]);
      val = -(a + b + d33) / r8w;
    }
    sw[((long)i * d.ny + j) * (nz + 1) + k] = val;
    if (k == nz - 1) sw[((long)i * d.ny + j) * (nz + 1) + nz] = 0.0f;   // the lid level
  }
}

// SGS diffusion of a scalar (turbulence.py:80-109, :207-234): khh on the x / y faces,
// khv on the z faces
__device__ float smag_scalar_point(const Dims& d, const float* __restrict__ rc, const float* __restrict__ kh,
                                   const float* __restrict__ khv, const float* __restrict__ s, int i, int j, int k) {
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
  return -div / r;
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
    S.out[q][o] = -div / r;
  }
}

#undef TAU
#undef TAUV
}  // namespace lex
