// Per-operator kernels: one kernel per reference routine on the LEX hot path, one
// thread per output point.  They are the drop-in entry points behind the Python
// mirror modules; the fused stage kernels in lex_stage.hip reuse the device functions
// defined here.
#pragma once
#include "lex_common.hpp"

namespace lex {

// ----------------------------------------------------------------------------------
// A1 pieces: mass-weighted face velocities for a horizontally uniform rho0 column.
// get_divergence (advection.py:200-226): rho at x/y faces is 0.5*(rho+rho) = rho, at
// w levels 0.5*(rho[K-1]+rho[K]) with zero at the padded end levels.
// ----------------------------------------------------------------------------------
__device__ __forceinline__ float rho8w_col(const float* __restrict__ r, int Kw, int nz) {
  if (Kw <= 0 || Kw >= nz + 2) return 0.0f;
  return 0.5f * (r[Kw - 1] + r[Kw]);
}

// flow divergence at padded scalar cell (I,J,K), any cell of the padded array
__device__ __forceinline__ float flow_div_at(const Dims& d, const float* __restrict__ r, const float* __restrict__ u,
                                            const float* __restrict__ v, const float* __restrict__ w, int I, int J, int K) {
  float rk = r[K];
  float dxp = (rk * u[iu(d, I + 1, J, K)] - rk * u[iu(d, I, J, K)]) * d.idx;
  float dyp = (rk * v[iv(d, I, J + 1, K)] - rk * v[iv(d, I, J, K)]) * d.idy;
  float dzp = (rho8w_col(r, K + 1, d.nz) * w[iw(d, I, J, K + 1)] - rho8w_col(r, K, d.nz) * w[iw(d, I, J, K)]) * d.idz;
  return dxp + dyp + dzp;
}

__global__ void k_div_rho(Dims d, const float* __restrict__ r, const float* __restrict__ u, const float* __restrict__ v,
                          const float* __restrict__ w, float* __restrict__ out) {
  int K = blockIdx.x * blockDim.x + threadIdx.x;
  int J = blockIdx.y * blockDim.y + threadIdx.y;
  int I = blockIdx.z;
  if (K >= d.nz + 2 || J >= d.ny + 6) return;
  out[ic(d, I, J, K)] = flow_div_at(d, r, u, v, w, I, J, K);
}

// get_divergence2 (advection.py:229-236)
__global__ void k_div(Dims d, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                      float* __restrict__ out) {
  int K = blockIdx.x * blockDim.x + threadIdx.x;
  int J = blockIdx.y * blockDim.y + threadIdx.y;
  int I = blockIdx.z;
  if (K >= d.nz + 2 || J >= d.ny + 6) return;
  out[ic(d, I, J, K)] = (u[iu(d, I + 1, J, K)] - u[iu(d, I, J, K)]) * d.idx +
                        (v[iv(d, I, J + 1, K)] - v[iv(d, I, J, K)]) * d.idy +
                        (w[iw(d, I, J, K + 1)] - w[iw(d, I, J, K)]) * d.idz;
}

// ----------------------------------------------------------------------------------
// A2: advection_scalar (advection.py:9-34) with vertical_flux_scalar2 (:305-416) and
// horizontal_flux_scalar (:419-516).
// ----------------------------------------------------------------------------------
__device__ __forceinline__ float scalar_xflux(const Dims& d, float r, const float* __restrict__ s,
                                              const float* __restrict__ u, int m, int J, int K, float weps) {
  float mass = r * u[iu(d, m, J, K)];
  long b = ic(d, m, J, K);
  long sx = d.sxc();
  float val = hrecon(s[b - 3 * sx], s[b - 2 * sx], s[b - sx], s[b], s[b + sx], s[b + 2 * sx], upwind_pos(mass), weps);
  return mass * val;
}

__device__ __forceinline__ float scalar_yflux(const Dims& d, float r, const float* __restrict__ s,
                                              const float* __restrict__ v, int I, int m, int K, float weps) {
  float mass = r * v[iv(d, I, m, K)];
  long b = ic(d, I, m, K);
  long sy = d.syc();
  float val = hrecon(s[b - 3 * sy], s[b - 2 * sy], s[b - sy], s[b], s[b + sy], s[b + 2 * sy], upwind_pos(mass), weps);
  return mass * val;
}

// vertical flux at face f (between padded levels f and f+1) of a scalar column
__device__ __forceinline__ float scalar_zflux(const Dims& d, const float* __restrict__ rc, const float* __restrict__ s,
                                              const float* __restrict__ w, int I, int J, int f, float weps) {
  float mass = 0.5f * (rc[f] + rc[f + 1]) * w[iw(d, I, J, f + 1)];
  float val = vrecon(s + ic(d, I, J, 0), d.nz + 2, f, upwind_pos(mass), weps);
  return mass * val;
}

__device__ float adv_scalar_point(const Dims& d, const float* __restrict__ rc, const float* __restrict__ s,
                                  const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                  float fdc, float sfc, int i, int j, int k, float weps) {
  int I = i + 3, J = j + 3, K = k + 1;
  float r = rc[K];
  float fx0 = scalar_xflux(d, r, s, u, I, J, K, weps);
  float fx1 = scalar_xflux(d, r, s, u, I + 1, J, K, weps);
  float fy0 = scalar_yflux(d, r, s, v, I, J, K, weps);
  float fy1 = scalar_yflux(d, r, s, v, I, J + 1, K, weps);
  float fz0 = (k == 0) ? sfc : scalar_zflux(d, rc, s, w, I, J, K - 1, weps);
  float fz1 = scalar_zflux(d, rc, s, w, I, J, K, weps);
  float conv = -((fx1 - fx0) * d.idx + (fy1 - fy0) * d.idy + (fz1 - fz0) * d.idz);
  return (conv + s[ic(d, I, J, K)] * fdc) / r;
}

__global__ void k_adv_scalar(Dims d, const float* __restrict__ rc, const float* __restrict__ s,
                             const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                             const float* __restrict__ fd, const float* __restrict__ sfc, float* __restrict__ out,
                             float weps) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  float sf = sfc ? sfc[(long)i * d.ny + j] : 0.0f;
  out[((long)i * d.ny + j) * d.nz + k] =
      adv_scalar_point(d, rc, s, u, v, w, fd[ic(d, i + 3, j + 3, k + 1)], sf, i, j, k, weps);
}

// ----------------------------------------------------------------------------------
// A3: advection_u/v/w (advection.py:37-149) with their vertical (:561-675, :816-930,
// :1070-1183) and horizontal (:678-771, :933-1025, :1186-1281) fluxes.
// ----------------------------------------------------------------------------------
// w-level mass flux 0.5*(rho[Kw-1]+rho[Kw])*w at (Ic,Jc,Kw) (vertical_flux_*2: rho0w)
__device__ __forceinline__ float mz_at(const Dims& d, const float* __restrict__ rc, const float* __restrict__ w,
                                       int Ic, int Jc, int Kw) {
  return 0.5f * (rc[Kw - 1] + rc[Kw]) * w[iw(d, Ic, Jc, Kw)];
}

__device__ float adv_u_point(const Dims& d, const float* __restrict__ rc, const float* __restrict__ u,
                             const float* __restrict__ v, const float* __restrict__ w, float fdl, float fdr,
                             float sfc_face, int f, int j, int k, float weps) {
  int I = f + 3, J = j + 3, K = k + 1;
  float r = rc[K];
  long sx = d.sxu(), sy = d.syc();
  // x flux at scalar cells I-1 and I (flux point m between u[m-1] and u[m])
  float fx[2], fy[2], fz[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int m = I + t;
    long b = iu(d, m, J, K);
    float mass = 0.5f * (u[b - sx] + u[b]) * r;
    fx[t] = mass * hrecon(u[b - 3 * sx], u[b - 2 * sx], u[b - sx], u[b], u[b + sx], u[b + 2 * sx], upwind_pos(mass), weps);
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int m = J + t;
    float mass = 0.5f * (r * v[iv(d, I - 1, m, K)] + r * v[iv(d, I, m, K)]);
    long b = iu(d, I, m, K);
    fy[t] = mass * hrecon(u[b - 3 * sy], u[b - 2 * sy], u[b - sy], u[b], u[b + sy], u[b + 2 * sy], upwind_pos(mass), weps);
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int fz_ = K - 1 + t;  // face between padded levels fz_ and fz_+1 (w level fz_+1)
    float mass = 0.5f * (mz_at(d, rc, w, I - 1, J, fz_ + 1) + mz_at(d, rc, w, I, J, fz_ + 1));
    fz[t] = mass * vrecon(u + iu(d, I, J, 0), d.nz + 2, fz_, upwind_pos(mass), weps);
  }
  if (k == 0) fz[0] = sfc_face;
  float conv = -((fx[1] - fx[0]) * d.idx + (fy[1] - fy[0]) * d.idy + (fz[1] - fz[0]) * d.idz);
  float div8u = 0.5f * (fdl + fdr);
  float rho8u = 0.5f * (r + r);
  return (conv + u[iu(d, I, J, K)] * div8u) / rho8u;
}

__device__ float adv_v_point(const Dims& d, const float* __restrict__ rc, const float* __restrict__ u,
                             const float* __restrict__ v, const float* __restrict__ w, float fdl, float fdr,
                             float sfc_face, int i, int f, int k, float weps) {
  int I = i + 3, J = f + 3, K = k + 1;
  float r = rc[K];
  long sx = d.sxv(), sy = d.syc();
  float fx[2], fy[2], fz[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int m = I + t;
    float mass = 0.5f * (r * u[iu(d, m, J - 1, K)] + r * u[iu(d, m, J, K)]);
    long b = iv(d, m, J, K);
    fx[t] = mass * hrecon(v[b - 3 * sx], v[b - 2 * sx], v[b - sx], v[b], v[b + sx], v[b + 2 * sx], upwind_pos(mass), weps);
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int m = J + t;
    long b = iv(d, I, m, K);
    float mass = 0.5f * (v[b - sy] + v[b]) * r;
    fy[t] = mass * hrecon(v[b - 3 * sy], v[b - 2 * sy], v[b - sy], v[b], v[b + sy], v[b + 2 * sy], upwind_pos(mass), weps);
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int fz_ = K - 1 + t;
    float mass = 0.5f * (mz_at(d, rc, w, I, J - 1, fz_ + 1) + mz_at(d, rc, w, I, J, fz_ + 1));
    fz[t] = mass * vrecon(v + iv(d, I, J, 0), d.nz + 2, fz_, upwind_pos(mass), weps);
  }
  if (k == 0) fz[0] = sfc_face;
  float conv = -((fx[1] - fx[0]) * d.idx + (fy[1] - fy[0]) * d.idy + (fz[1] - fz[0]) * d.idz);
  float div8v = 0.5f * (fdl + fdr);
  float rho8v = 0.5f * (r + r);
  return (conv + v[iv(d, I, J, K)] * div8v) / rho8v;
}

// w at interior level l (1..nz-1; the caller zeroes l=0 and l=nz)
__device__ float adv_w_point(const Dims& d, const float* __restrict__ rc, const float* __restrict__ u,
                             const float* __restrict__ v, const float* __restrict__ w, float fdb, float fdt,
                             int i, int j, int l, float weps) {
  int I = i + 3, J = j + 3, K = l + 1;
  long sx = d.sxw(), sy = d.syw();
  float fx[2], fy[2], fz[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int m = I + t;
    float mass = 0.5f * (rc[K - 1] * u[iu(d, m, J, K - 1)] + rc[K] * u[iu(d, m, J, K)]);
    long b = iw(d, m, J, K);
    fx[t] = mass * hrecon(w[b - 3 * sx], w[b - 2 * sx], w[b - sx], w[b], w[b + sx], w[b + 2 * sx], upwind_pos(mass), weps);
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int m = J + t;
    float mass = 0.5f * (rc[K - 1] * v[iv(d, I, m, K - 1)] + rc[K] * v[iv(d, I, m, K)]);
    long b = iw(d, I, m, K);
    fy[t] = mass * hrecon(w[b - 3 * sy], w[b - 2 * sy], w[b - sy], w[b], w[b + sy], w[b + 2 * sy], upwind_pos(mass), weps);
  }
  const float* col = w + iw(d, I, J, 0);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int c = K - 1 + t;  // scalar level c: face between w[c] and w[c+1]
    float mass = 0.5f * (col[c] + col[c + 1]) * rc[c];
    fz[t] = mass * vrecon(col, d.nz + 3, c, upwind_pos(mass), weps);
  }
  float conv = -((fx[1] - fx[0]) * d.idx + (fy[1] - fy[0]) * d.idy + (fz[1] - fz[0]) * d.idz);
  float div8w = 0.5f * (fdb + fdt);
  float rho8w = 0.5f * (rc[K - 1] + rc[K]);
  return (conv + col[K] * div8w) / rho8w;
}

__global__ void k_adv_u(Dims d, const float* __restrict__ rc, const float* __restrict__ u, const float* __restrict__ v,
                        const float* __restrict__ w, const float* __restrict__ fd, const float* __restrict__ sfc,
                        float* __restrict__ out, float weps) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int f = blockIdx.z;  // 0..nx
  if (k >= d.nz || j >= d.ny) return;
  float sface = 0.0f;
  if (sfc) {
    int il = (f - 1 + d.nx) % d.nx, ir = f % d.nx;
    sface = 0.5f * (sfc[(long)il * d.ny + j] + sfc[(long)ir * d.ny + j]);
  }
  int I = f + 3, J = j + 3, K = k + 1;
  out[((long)f * d.ny + j) * d.nz + k] =
      adv_u_point(d, rc, u, v, w, fd[ic(d, I - 1, J, K)], fd[ic(d, I, J, K)], sface, f, j, k, weps);
}

__global__ void k_adv_v(Dims d, const float* __restrict__ rc, const float* __restrict__ u, const float* __restrict__ v,
                        const float* __restrict__ w, const float* __restrict__ fd, const float* __restrict__ sfc,
                        float* __restrict__ out, float weps) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int f = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = blockIdx.z;
  if (k >= d.nz || f > d.ny) return;
  float sface = 0.0f;
  if (sfc) {
    int jl = (f - 1 + d.ny) % d.ny, jr = f % d.ny;
    sface = 0.5f * (sfc[(long)i * d.ny + jl] + sfc[(long)i * d.ny + jr]);
  }
  int I = i + 3, J = f + 3, K = k + 1;
  out[((long)i * (d.ny + 1) + f) * d.nz + k] =
      adv_v_point(d, rc, u, v, w, fd[ic(d, I, J - 1, K)], fd[ic(d, I, J, K)], sface, i, f, k, weps);
}

__global__ void k_adv_w(Dims d, const float* __restrict__ rc, const float* __restrict__ u, const float* __restrict__ v,
                        const float* __restrict__ w, const float* __restrict__ fd, float* __restrict__ out, float weps) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (l > d.nz || j >= d.ny) return;
  float val = 0.0f;
  if (l > 0 && l < d.nz) {
    int I = i + 3, J = j + 3, K = l + 1;
    val = adv_w_point(d, rc, u, v, w, fd[ic(d, I, J, K - 1)], fd[ic(d, I, J, K)], i, j, l, weps);
  }
  out[((long)i * d.ny + j) * (d.nz + 1) + l] = val;
}

// ----------------------------------------------------------------------------------
// A4: advection_pi (advection.py:152-197), cc1 = cc2 = 1/2 on the uniform grid
// ----------------------------------------------------------------------------------
__device__ float adv_pi_point(const Dims& d, const float* __restrict__ rc, const float* __restrict__ pi0c,
                              const float* __restrict__ ps, const float* __restrict__ u, const float* __restrict__ v,
                              const float* __restrict__ w, int i, int j, int k) {
  int I = i + 3, J = j + 3, K = k + 1;
  long c = ic(d, I, J, K);
  long sx = d.sxc(), sy = d.syc();
  float pc = ps[c];
  float udx0 = u[iu(d, I, J, K)] * ((pc - ps[c - sx]) * d.idx);
  float udx1 = u[iu(d, I + 1, J, K)] * ((ps[c + sx] - pc) * d.idx);
  float vdy0 = v[iv(d, I, J, K)] * ((pc - ps[c - sy]) * d.idy);
  float vdy1 = v[iv(d, I, J + 1, K)] * ((ps[c + sy] - pc) * d.idy);
  // w levels K (lower) and K+1 (upper); zero at surface (l=0) and lid (l=nz)
  float dz0 = (pi0c[K] - pi0c[K - 1]) * d.idz + (pc - ps[c - 1]) * d.idz;
  float dz1 = (pi0c[K + 1] - pi0c[K]) * d.idz + (ps[c + 1] - pc) * d.idz;
  float wdz0 = (k == 0) ? 0.0f : w[iw(d, I, J, K)] * dz0;
  float wdz1 = (k == d.nz - 1) ? 0.0f : w[iw(d, I, J, K + 1)] * dz1;
  float r8w0 = 0.5f * (rc[K - 1] + rc[K]);
  float r8w1 = 0.5f * (rc[K] + rc[K + 1]);
  return -(0.5f * (udx1 + udx0) + 0.5f * (vdy1 + vdy0) + (0.5f * wdz0 * r8w0 + 0.5f * wdz1 * r8w1) / rc[K]);
}

__global__ void k_adv_pi(Dims d, const float* __restrict__ rc, const float* __restrict__ pi0c,
                         const float* __restrict__ ps, const float* __restrict__ u, const float* __restrict__ v,
                         const float* __restrict__ w, float* __restrict__ out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  out[((long)i * d.ny + j) * d.nz + k] = adv_pi_point(d, rc, pi0c, ps, u, v, w, i, j, k);
}

// ----------------------------------------------------------------------------------
// B1: calculate_buoyancy (pressure_equations.py:162-172) with interpolate_scalar2w and
// extrapolate_bottom_top (:123-159)
// ----------------------------------------------------------------------------------
__device__ __forceinline__ float theta_rho(float th, float qv) {
  return th * (1.0f + (float)kReps * qv) / (1.0f + qv);
}

__device__ __forceinline__ float buoy_at(const float* __restrict__ th0c, const float* __restrict__ qv0c, float th,
                                         float qv, int K) {
  float tr0 = th0c[K] * (1.0f + (float)kReps * qv0c[K]) / (1.0f + qv0c[K]);
  float tr = theta_rho(th, qv);
  return (float)kG * (tr - tr0) / tr0;
}

__device__ __forceinline__ float extrap3(float a0, float a1, float a2) {
  return 1.875f * a0 + (-1.25f) * a1 + 0.375f * a2;
}

// B at interior w level l (0..nz)
__device__ float b8w_point(const Dims& d, const float* __restrict__ th0c, const float* __restrict__ qv0c,
                           const float* __restrict__ th, const float* __restrict__ qv, int I, int J, int l) {
  long c = ic(d, I, J, 0);
  auto B = [&](int k) { return buoy_at(th0c, qv0c, th[c + k + 1], qv[c + k + 1], k + 1); };
  if (l == 0) return extrap3(B(0), B(1), B(2));
  if (l == d.nz) return extrap3(B(d.nz - 1), B(d.nz - 2), B(d.nz - 3));
  return 0.5f * (B(l - 1) + B(l));
}

__global__ void k_buoyancy(Dims d, const float* __restrict__ th0c, const float* __restrict__ qv0c,
                           const float* __restrict__ th, const float* __restrict__ qv, float* __restrict__ b,
                           float* __restrict__ b8w) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (l > d.nz || j >= d.ny) return;
  int I = i + 3, J = j + 3;
  if (l < d.nz) b[((long)i * d.ny + j) * d.nz + l] = buoy_at(th0c, qv0c, th[ic(d, I, J, l + 1)], qv[ic(d, I, J, l + 1)], l + 1);
  b8w[((long)i * d.ny + j) * (d.nz + 1) + l] = b8w_point(d, th0c, qv0c, th, qv, I, J, l);
}

// theta_rho over the whole padded array (one_step_integration.py:135)
__global__ void k_theta_rho(long n, const float* __restrict__ th, const float* __restrict__ qv, float* __restrict__ out) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = theta_rho(th[t], qv[t]);
}

// ----------------------------------------------------------------------------------
// P1: pressure_gradient_force (pressure_gradient_coriolis.py:7-37)
// ----------------------------------------------------------------------------------
__device__ __forceinline__ float pgf_x(const Dims& d, const float* __restrict__ p, const float* __restrict__ tr,
                                       int I, int J, int K) {
  long c = ic(d, I, J, K);
  float dp = (p[c] - p[c - d.sxc()]) * d.idx;
  return (float)(-kCp * 0.5) * (tr[c] + tr[c - d.sxc()]) * dp;
}
__device__ __forceinline__ float pgf_y(const Dims& d, const float* __restrict__ p, const float* __restrict__ tr,
                                       int I, int J, int K) {
  long c = ic(d, I, J, K);
  float dp = (p[c] - p[c - d.syc()]) * d.idy;
  return (float)(-kCp * 0.5) * (tr[c] + tr[c - d.syc()]) * dp;
}
// w level Kw (1..nz+1); zero at l=0 and l=nz
__device__ __forceinline__ float pgf_z(const Dims& d, const float* __restrict__ p, const float* __restrict__ tr,
                                       int I, int J, int Kw) {
  if (Kw <= 1 || Kw >= d.nz + 1) return 0.0f;
  long c = ic(d, I, J, Kw);
  float dp = (p[c] - p[c - 1]) * d.idz;
  return (float)(-kCp * 0.5) * (tr[c] + tr[c - 1]) * dp;
}

__global__ void k_pgf(Dims d, const float* __restrict__ p, const float* __restrict__ tr, float* __restrict__ pu,
                      float* __restrict__ pv, float* __restrict__ pw) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nz
  int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = blockIdx.z;                             // 0..nx
  if (k > d.nz || j > d.ny) return;
  if (k < d.nz && j < d.ny) pu[((long)i * d.ny + j) * d.nz + k] = pgf_x(d, p, tr, i + 3, j + 3, k + 1);
  if (k < d.nz && i < d.nx) pv[((long)i * (d.ny + 1) + j) * d.nz + k] = pgf_y(d, p, tr, i + 3, j + 3, k + 1);
  if (i < d.nx && j < d.ny) pw[((long)i * d.ny + j) * (d.nz + 1) + k] = pgf_z(d, p, tr, i + 3, j + 3, k + 1);
}

// ----------------------------------------------------------------------------------
// P2: calculate_coriolis_force (pressure_gradient_coriolis.py:80-95)
// ----------------------------------------------------------------------------------
// fv at u point (I, J, K) and fu at v point (I, J, K) (pressure_gradient_coriolis.py:80-113)
__device__ __forceinline__ float cor_fv(const Dims& d, const float* __restrict__ v, int I, int J, int K,
                                        const CorOpt& c) {
  return c.of(v[iv(d, I - 1, J, K)] + v[iv(d, I - 1, J + 1, K)] + v[iv(d, I, J, K)] + v[iv(d, I, J + 1, K)], c.vg);
}
__device__ __forceinline__ float cor_fu(const Dims& d, const float* __restrict__ u, int I, int J, int K,
                                        const CorOpt& c) {
  return c.of(u[iu(d, I, J - 1, K)] + u[iu(d, I, J, K)] + u[iu(d, I + 1, J - 1, K)] + u[iu(d, I + 1, J, K)], c.ug);
}

__global__ void k_coriolis(Dims d, const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ fu,
                           float* __restrict__ fv, CorOpt c) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = blockIdx.z;                             // 0..nx
  if (k >= d.nz || j > d.ny) return;
  if (j < d.ny) fv[((long)i * d.ny + j) * d.nz + k] = cor_fv(d, v, i + 3, j + 3, k + 1, c);
  if (i < d.nx) fu[((long)i * (d.ny + 1) + j) * d.nz + k] = cor_fu(d, u, i + 3, j + 3, k + 1, c);
}

// momentum right-hand sides of the staged stage (one_step_integration.py:126-133):
// rhs_u = adv4u + fv + sgs_u, rhs_v = adv4v - fu + sgs_v, rhs_w = adv4w + b8w + sgs_w,
// each optional term NULL when absent.  Thread (k, j, i) over (nz+1, ny+1, nx+1).
__global__ void k_momentum_rhs(Dims d, const float* __restrict__ au, const float* __restrict__ av,
                               const float* __restrict__ aw, const float* __restrict__ fu,
                               const float* __restrict__ fv, const float* __restrict__ b8w,
                               const float* __restrict__ su, const float* __restrict__ sv,
                               const float* __restrict__ sw, float* __restrict__ ru, float* __restrict__ rv,
                               float* __restrict__ rw) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nz
  int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = blockIdx.z;                             // 0..nx
  if (k > d.nz || j > d.ny) return;
  if (k < d.nz && j < d.ny) {   // u points (nx+1, ny, nz)
    const long o = ((long)i * d.ny + j) * d.nz + k;
    float a = au[o];
    if (fv) a = a + fv[o];
    if (su) a = a + su[o];
    ru[o] = a;
  }
  if (k < d.nz && i < d.nx) {   // v points (nx, ny+1, nz)
    const long o = ((long)i * (d.ny + 1) + j) * d.nz + k;
    float a = av[o];
    if (fu) a = a - fu[o];
    if (sv) a = a + sv[o];
    rv[o] = a;
  }
  if (i < d.nx && j < d.ny) {   // w points (nx, ny, nz+1)
    const long o = ((long)i * d.ny + j) * (d.nz + 1) + k;
    float a = aw[o];
    if (b8w) a = a + b8w[o];
    if (sw) a = a + sw[o];
    rw[o] = a;
  }
}

// ----------------------------------------------------------------------------------
// H1: padding (one_step_integration.py:950-985, pressure_equations.py:111-120)
// mode 0 = scalar/w (padding3_array), 1 = u (padding3u_array), 2 = v (padding3v_array)
// Input is an interior array (ni, nj, nk); output padded (ni+6 or +6/+... , nk+2).
// ----------------------------------------------------------------------------------
__device__ __forceinline__ int src_index(int P, int n, int g, bool face) {
  // padded index P along an axis padded by g on each side; returns the interior index
  if (!face) {
    int t = P - g;
    if (t < 0) t += n;
    if (t >= n) t -= n;
    return t;
  }
  // face axis: interior count n+1 (faces 0..n); ghosts skip the duplicate face
  int t = P - g;
  if (t < 0) return t + n;        // padded 0..g-1 -> faces n-g..n-1
  if (t > n) return t - n;        // padded n+g+1.. -> faces 1..g
  return t;
}

__global__ void k_pad(int mode, int ni, int nj, int nk, int g, const float* __restrict__ in, float* __restrict__ out) {
  // ni, nj, nk: interior extents of `in` (for mode 1 ni = nx+1, for mode 2 nj = ny+1)
  int K = blockIdx.x * blockDim.x + threadIdx.x;
  int J = blockIdx.y * blockDim.y + threadIdx.y;
  int I = blockIdx.z;
  int PI = ni + 2 * g, PJ = nj + 2 * g, PK = nk + 2;
  if (K >= PK || J >= PJ || I >= PI) return;
  int si = (mode == 1) ? src_index(I, ni - 1, g, true) : src_index(I, ni, g, false);
  int sj = (mode == 2) ? src_index(J, nj - 1, g, true) : src_index(J, nj, g, false);
  int sk = K - 1;
  if (sk < 0) sk = 0;
  if (sk >= nk) sk = nk - 1;
  out[((long)I * PJ + J) * PK + K] = in[((long)si * nj + sj) * nk + sk];
}

__global__ void k_set_w_bc(Dims d, float* __restrict__ w) {
  int J = blockIdx.x * blockDim.x + threadIdx.x;
  int I = blockIdx.y;
  if (J >= d.ny + 6) return;
  long b = iw(d, I, J, 0);
  w[b] = 0.0f;
  w[b + 1] = 0.0f;
  w[b + d.nz + 1] = 0.0f;
  w[b + d.nz + 2] = 0.0f;
}

// ----------------------------------------------------------------------------------
// S1 arithmetic of one cell (one_step_integration.py:528-553): the six face velocities
// after pressure_gradient_force + rhs, set_w_bc at the surface and lid, the divergence,
// advection_pi with the fixed pi* and the damped pi' update.  Operands: the cell's faces
// (u0/u1 lower/upper x face, ...), the rhs on those faces, the 7-point stencils of pi',
// theta_rho and pi*, and the rho0 / pi0 column values at levels K-1, K, K+1.
// ----------------------------------------------------------------------------------
struct AcousticOut {
  float un0, un1, vn0, vn1, wn0, wn1, pn;
};

__device__ __forceinline__ AcousticOut acoustic_cell(
    const Dims& d, int k, float u0, float u1, float v0, float v1, float w0, float w1, float ru0, float ru1, float rv0,
    float rv1, float rw0, float rw1, float pc, float pxm, float pxp, float pym, float pyp, float pzm, float pzp, float tc,
    float txm, float txp, float tym, float typ, float tzm, float tzp, float sc, float sxm, float sxp, float sym,
    float syp, float szm, float szp, float rcm, float rcc, float rcp, float pim, float pic, float pip, float dts) {
  // Contraction is spelled out (fmaf) and otherwise off, so the value of every output
  // does not depend on the surrounding kernel's code generation: k_acoustic and both
  // substeps of k_acoustic2 round identically.
#pragma clang fp contract(off)
  const int K = k + 1;
  const float h = (float)(-kCp * 0.5);
  // new velocities on the six faces of the cell (pgf_x/y/z with the same operation order)
  const float un0 = fmaf(fmaf(h * (tc + txm), (pc - pxm) * d.idx, ru0), dts, u0);
  const float un1 = fmaf(fmaf(h * (txp + tc), (pxp - pc) * d.idx, ru1), dts, u1);
  const float vn0 = fmaf(fmaf(h * (tc + tym), (pc - pym) * d.idy, rv0), dts, v0);
  const float vn1 = fmaf(fmaf(h * (typ + tc), (pyp - pc) * d.idy, rv1), dts, v1);
  // set_w_bc: surface (l=0) and lid (l=nz) are zero after the update; pgf_z vanishes at
  // levels 1 and nz+1
  const float pz0 = (K <= 1) ? 0.0f : h * (tc + tzm) * ((pc - pzm) * d.idz);
  const float pz1 = (K + 1 >= d.nz + 1) ? 0.0f : h * (tzp + tc) * ((pzp - pc) * d.idz);
  const float wn0 = (k == 0) ? 0.0f : fmaf(pz0 + rw0, dts, w0);
  const float wn1 = (k == d.nz - 1) ? 0.0f : fmaf(pz1 + rw1, dts, w1);
  const float div = fmaf(wn1 - wn0, d.idz, fmaf(vn1 - vn0, d.idy, (un1 - un0) * d.idx));
  // advection_pi with the new velocities and the fixed pi*
  const float udx0 = un0 * ((sc - sxm) * d.idx);
  const float udx1 = un1 * ((sxp - sc) * d.idx);
  const float vdy0 = vn0 * ((sc - sym) * d.idy);
  const float vdy1 = vn1 * ((syp - sc) * d.idy);
  const float dz0 = fmaf(sc - szm, d.idz, (pic - pim) * d.idz);
  const float dz1 = fmaf(szp - sc, d.idz, (pip - pic) * d.idz);
  const float wdz0 = (k == 0) ? 0.0f : wn0 * dz0;
  const float wdz1 = (k == d.nz - 1) ? 0.0f : wn1 * dz1;
  const float r8w0 = 0.5f * (rcm + rcc);
  const float r8w1 = 0.5f * (rcc + rcp);
  const float advp =
      -(fmaf(0.5f, udx1 + udx0, 0.5f * (vdy1 + vdy0)) + fmaf(0.5f * wdz0, r8w0, 0.5f * wdz1 * r8w1) / rcc);
  const float piadv = pic + sc;
  const float dpdt = fmaf((float)(-kRd / kCv), piadv * div, advp);
  const float pn = fmaf(1.1f * dpdt, dts, pc);
  AcousticOut o;
  o.un0 = un0;
  o.un1 = un1;
  o.vn0 = vn0;
  o.vn1 = vn1;
  o.wn0 = wn0;
  o.wn1 = wn1;
  o.pn = pn;
  return o;
}

// ----------------------------------------------------------------------------------
// S1: one forward-backward acoustic substep (one_step_integration.py:528-553).
// Reads u,v,w,pi' of the current substep, writes the next (halos filled, w BC set).
// ----------------------------------------------------------------------------------
__global__ void k_acoustic(Dims d, const float* __restrict__ rc, const float* __restrict__ pi0c,
                           const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                           const float* __restrict__ p, const float* __restrict__ ps, const float* __restrict__ ru,
                           const float* __restrict__ rv, const float* __restrict__ rw, const float* __restrict__ tr,
                           float* __restrict__ uo, float* __restrict__ vo, float* __restrict__ wo,
                           float* __restrict__ po, float dts) {
  const TileId tb = xcd_tile();
  int k = tb.x * blockDim.x + threadIdx.x;
  int j = tb.y * blockDim.y + threadIdx.y;
  int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  int I = i + 3, J = j + 3, K = k + 1;
  const long c = ic(d, I, J, K);
  const long sx = d.sxc(), sy = d.syc();
  // Every operand is loaded unconditionally (all indices are inside the padded arrays,
  // the w levels K and K+1 exist for K in 1..nz) so the loads issue back to back; the
  // surface and lid cases are selected afterwards.
  const float rcm = rc[K - 1], rcc = rc[K], rcp = rc[K + 1];
  const float pim = pi0c[K - 1], pic = pi0c[K], pip = pi0c[K + 1];
  const long cu = iu(d, I, J, K), cv = iv(d, I, J, K), cw = iw(d, I, J, K);
  const long ou = ((long)i * d.ny + j) * d.nz + k, ov = ((long)i * (d.ny + 1) + j) * d.nz + k;
  const long ow = ((long)i * d.ny + j) * (d.nz + 1) + k;
  const float u0 = u[cu], u1 = u[cu + d.sxu()], v0 = v[cv], v1 = v[cv + d.syc()], w0 = w[cw], w1 = w[cw + 1];
  const float ru0 = ru[ou], ru1 = ru[ou + (long)d.ny * d.nz], rv0 = rv[ov], rv1 = rv[ov + d.nz];
  const float rw0 = rw[ow], rw1 = rw[ow + 1];
  const float pc = p[c], pxm = p[c - sx], pxp = p[c + sx], pym = p[c - sy], pyp = p[c + sy], pzm = p[c - 1],
              pzp = p[c + 1];
  const float tc = tr[c], txm = tr[c - sx], txp = tr[c + sx], tym = tr[c - sy], typ = tr[c + sy],
              tzm = tr[c - 1], tzp = tr[c + 1];
  const float sc = ps[c], sxm = ps[c - sx], sxp = ps[c + sx], sym = ps[c - sy], syp = ps[c + sy],
              szm = ps[c - 1], szp = ps[c + 1];
  const AcousticOut o = acoustic_cell(d, k, u0, u1, v0, v1, w0, w1, ru0, ru1, rv0, rv1, rw0, rw1, pc, pxm, pxp, pym,
                                      pyp, pzm, pzp, tc, txm, txp, tym, typ, tzm, tzp, sc, sxm, sxp, sym, syp, szm, szp,
                                      rcm, rcc, rcp, pim, pic, pip, dts);
  const float pn = o.pn, un0 = o.un0, un1 = o.un1, vn0 = o.vn0, vn1 = o.vn1, wn0 = o.wn0, wn1 = o.wn1;
  store_c_fast(d, po, i, j, k, pn);
  store_u_fast(d, uo, i, j, k, un0);
  if (i == d.nx - 1) store_u(d, uo, d.nx, j, k, un1);
  store_v_fast(d, vo, i, j, k, vn0);
  if (j == d.ny - 1) store_v(d, vo, i, d.ny, k, vn1);
  store_w_fast(d, wo, i, j, k, wn0);
  if (k == d.nz - 1) store_w(d, wo, i, j, d.nz, wn1);
}

// ----------------------------------------------------------------------------------
// k_acoustic2: two consecutive S1 substeps in one launch (the same arithmetic as two
// k_acoustic launches, acoustic_cell for each).  A workgroup owns whole z-columns of a
// strip of R-2 rows in y and marches `ax` cells along x.  Each thread takes one
// (row, level) of the strip plus one halo row on each side and, at x position s,
//   substep A: advances cell s from the launch's input (global loads, as k_acoustic);
//   its new pi' goes to a three-slab LDS ring;
//   substep B: advances cell s-1 from the substep-A values: pi' of s-2, s-1, s from its
//   own registers, the y and z neighbours from the ring, its faces' velocities and the
//   operands of the rhs, theta_rho and pi* terms carried from its substep-A step of s-1.
// Halo rows and the x positions x0-1 and x1 are computed at their periodic images
// (wrapped indices), so every value substep B reads equals what the second of two
// k_acoustic launches would read through the ghost cells.  Only substep B writes.
// ----------------------------------------------------------------------------------
constexpr int kA2Rows = 8;  // rows per workgroup (two of them halo rows), at most

inline int acoustic2_lanes(const Dims& d) { return 64 * ((d.nz + 63) / 64); }
inline int acoustic2_rows(const Dims& d) {
  const int r = 1024 / acoustic2_lanes(d);
  return r > kA2Rows ? kA2Rows : r;
}
inline bool acoustic2_ok(const Dims& d) { return acoustic2_rows(d) >= 4; }
inline size_t acoustic2_lds(const Dims& d) {
  return (size_t)3 * acoustic2_rows(d) * (acoustic2_lanes(d) + 2) * sizeof(float);
}
// cells along x per workgroup: 16 (2 halo positions per 16), fewer when that would leave
// the 256 CUs with under two workgroups each
inline int acoustic2_ax(const Dims& d) {
  const int ty = acoustic2_rows(d) - 2, strips = (d.ny + ty - 1) / ty;
  int ax = 16;
  while (ax > 4 && (long)strips * ((d.nx + ax - 1) / ax) < 512) ax >>= 1;
  return ax;
}
inline dim3 acoustic2_grid(const Dims& d) {
  const int ty = acoustic2_rows(d) - 2, ax = acoustic2_ax(d);
  return dim3((d.ny + ty - 1) / ty, (d.nx + ax - 1) / ax, 1);
}


__global__ void __launch_bounds__(1024) k_acoustic2(Dims d, const float* __restrict__ rc,
                                                    const float* __restrict__ pi0c, const float* __restrict__ u,
                                                    const float* __restrict__ v, const float* __restrict__ w,
                                                    const float* __restrict__ p, const float* __restrict__ ps,
                                                    const float* __restrict__ ru, const float* __restrict__ rv,
                                                    const float* __restrict__ rw, const float* __restrict__ tr,
                                                    float* __restrict__ uo, float* __restrict__ vo,
                                                    float* __restrict__ wo, float* __restrict__ po, float dts,
                                                    int ax) {
  extern __shared__ float ring[];  // [3 slabs][R rows][NZL + 2 levels]
  const int NZL = blockDim.x, R = blockDim.y, LR = NZL + 2;
  const int k = threadIdx.x, r = threadIdx.y;
  const TileId tb = xcd_tile();
  const int j = tb.x * (R - 2) - 1 + r;  // -1 .. ny (rows past ny+... are clamped, never stored)
  const int jw = wrap_cell(min(j, d.ny), d.ny);
  const int x0 = tb.y * ax, x1 = min(x0 + ax, d.nx);
  const bool act = k < d.nz;
  const int kc = act ? k : d.nz - 1;
  const bool bcell = act && r >= 1 && r <= R - 2 && j < d.ny;
  const int J = jw + 3, K = kc + 1;
  const long sx = d.sxc(), sy = d.syc();
  const float rcm = rc[K - 1], rcc = rc[K], rcp = rc[K + 1];
  const float pim = pi0c[K - 1], pic = pi0c[K], pip = pi0c[K + 1];
  // substep-A results and operands of the cell that substep B advances next
  float bu0 = 0.f, bu1 = 0.f, bv0 = 0.f, bv1 = 0.f, bw0 = 0.f, bw1 = 0.f, bp = 0.f, bpm = 0.f;
  float bru0 = 0.f, bru1 = 0.f, brv0 = 0.f, brv1 = 0.f, brw0 = 0.f, brw1 = 0.f;
  float btc = 0.f, btxm = 0.f, btxp = 0.f, btym = 0.f, btyp = 0.f, btzm = 0.f, btzp = 0.f;
  float bsc = 0.f, bsxm = 0.f, bsxp = 0.f, bsym = 0.f, bsyp = 0.f, bszm = 0.f, bszp = 0.f;
  // x-neighbour values carried along the march: p and u at the previous position (the
  // next cell's pxm / pc and u0), valid when the previous position is the x-neighbour in
  // memory (not across the wrap to a periodic image); theta_rho, pi* and ru from the
  // substep-B operands (btc, btxp, bsc, bsxp, bru1)
  float cpc = 0.f, cpxp = 0.f, cu1 = 0.f;
  const bool zlo_ld = (k & 63) == 0 || kc == 0;               // lanes whose z neighbours are
  const bool zhi_ld = (k & 63) == 63 || kc >= d.nz - 1;       // not held by an adjacent lane
  for (int s = x0 - 1; s <= x1; ++s) {
    // ---- substep A at cell (s, j, k), taken at its periodic image (sw, jw, kc)
    const int sw = wrap_cell(s, d.nx), I = sw + 3;
    const bool carry = (s > x0 - 1) && sw == wrap_cell(s - 1, d.nx) + 1;
    const long c = ic(d, I, J, K);
    const long cu = iu(d, I, J, K), cv = iv(d, I, J, K), cw = iw(d, I, J, K);
    const long ou = ((long)sw * d.ny + jw) * d.nz + kc, ov = ((long)sw * (d.ny + 1) + jw) * d.nz + kc;
    const long ow = ((long)sw * d.ny + jw) * (d.nz + 1) + kc;
    const float u1 = u[cu + d.sxu()], v0 = v[cv], v1 = v[cv + d.syc()], w0 = w[cw], w1 = w[cw + 1];
    const float ru1 = ru[ou + (long)d.ny * d.nz], rv0 = rv[ov], rv1 = rv[ov + d.nz];
    const float rw0 = rw[ow], rw1 = rw[ow + 1];
    const float pxp = p[c + sx], pym = p[c - sy], pyp = p[c + sy];
    const float txp = tr[c + sx], tym = tr[c - sy], typ = tr[c + sy];
    const float sxp = ps[c + sx], sym = ps[c - sy], syp = ps[c + sy];
    float u0, ru0, pc, pxm, tc, txm, sc, sxm;
    if (carry) {
      u0 = cu1; ru0 = bru1; pc = cpxp; pxm = cpc; tc = btxp; txm = btc; sc = bsxp; sxm = bsc;
    } else {
      u0 = u[cu]; ru0 = ru[ou]; pc = p[c]; pxm = p[c - sx]; tc = tr[c]; txm = tr[c - sx]; sc = ps[c]; sxm = ps[c - sx];
    }
    // z neighbours from the adjacent lanes (level k-1 is lane k-1); the surface, lid,
    // idle and wave-edge lanes load theirs
    float pzm = __shfl_up(pc, 1, 64), pzp = __shfl_down(pc, 1, 64);
    float tzm = __shfl_up(tc, 1, 64), tzp = __shfl_down(tc, 1, 64);
    float szm = __shfl_up(sc, 1, 64), szp = __shfl_down(sc, 1, 64);
    if (zlo_ld) {
      pzm = p[c - 1];
      tzm = tr[c - 1];
      szm = ps[c - 1];
    }
    if (zhi_ld) {
      pzp = p[c + 1];
      tzp = tr[c + 1];
      szp = ps[c + 1];
    }
    cpc = pc;
    cpxp = pxp;
    cu1 = u1;
    const AcousticOut a = acoustic_cell(d, kc, u0, u1, v0, v1, w0, w1, ru0, ru1, rv0, rv1, rw0, rw1, pc, pxm, pxp,
                                        pym, pyp, pzm, pzp, tc, txm, txp, tym, typ, tzm, tzp, sc, sxm, sxp, sym, syp,
                                        szm, szp, rcm, rcc, rcp, pim, pic, pip, dts);
    ring[((s - x0 + 3) % 3) * R * LR + r * LR + k + 1] = a.pn;
    __syncthreads();
    // ---- substep B at cell (s-1, j, k); the ring slab of s-1 was completed one step ago
    if (s > x0 && bcell) {
      const int i = s - 1;
      const float* L = ring + ((i - x0 + 3) % 3) * R * LR + r * LR + k + 1;
      const AcousticOut b = acoustic_cell(d, kc, bu0, bu1, bv0, bv1, bw0, bw1, bru0, bru1, brv0, brv1, brw0, brw1, bp,
                                          bpm, a.pn, L[-LR], L[LR], L[-1], L[1], btc, btxm, btxp, btym, btyp, btzm,
                                          btzp, bsc, bsxm, bsxp, bsym, bsyp, bszm, bszp, rcm, rcc, rcp, pim, pic, pip,
                                          dts);
      store_c_fast(d, po, i, j, k, b.pn);
      store_u_fast(d, uo, i, j, k, b.un0);
      if (i == d.nx - 1) store_u(d, uo, d.nx, j, k, b.un1);
      store_v_fast(d, vo, i, j, k, b.vn0);
      if (j == d.ny - 1) store_v(d, vo, i, d.ny, k, b.vn1);
      store_w_fast(d, wo, i, j, k, b.wn0);
      if (k == d.nz - 1) store_w(d, wo, i, j, d.nz, b.wn1);
    }
    bpm = bp;
    bp = a.pn;
    bu0 = a.un0; bu1 = a.un1; bv0 = a.vn0; bv1 = a.vn1; bw0 = a.wn0; bw1 = a.wn1;
    bru0 = ru0; bru1 = ru1; brv0 = rv0; brv1 = rv1; brw0 = rw0; brw1 = rw1;
    btc = tc; btxm = txm; btxp = txp; btym = tym; btyp = typ; btzm = tzm; btzp = tzp;
    bsc = sc; bsxm = sxm; bsxp = sxp; bsym = sym; bsyp = syp; bszm = szm; bszp = szp;
  }
}

// k_acoustic_m: the same substep with each thread marching kAX cells along x: the x-face
// velocity, pi', theta_rho and pi* of the upper neighbour are carried to the next cell
// (one x-face update per cell instead of two), and the z neighbours of pi', theta_rho
// and pi* come from the adjacent lanes of the wave.  Block 64 (z) x 4 (y).
constexpr int kAX = 8;

__global__ void __launch_bounds__(256) k_acoustic_m(Dims d, const float* __restrict__ rc, const float* __restrict__ pi0c,
                                                    const float* __restrict__ u, const float* __restrict__ v,
                                                    const float* __restrict__ w, const float* __restrict__ p,
                                                    const float* __restrict__ ps, const float* __restrict__ ru,
                                                    const float* __restrict__ rv, const float* __restrict__ rw,
                                                    const float* __restrict__ tr, float* __restrict__ uo,
                                                    float* __restrict__ vo, float* __restrict__ wo,
                                                    float* __restrict__ po, float dts) {
  const int lane = threadIdx.x;
  const int k = blockIdx.z * 64 + lane;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int i0 = blockIdx.x * kAX;
  const bool act = (k < d.nz) && (j < d.ny);
  const int kc = min(k, d.nz - 1), jc = min(j, d.ny - 1);   // clamped (idle lanes load valid data)
  const int J = jc + 3, K = kc + 1;
  const long sx = d.sxc(), sy = d.syc();
  const float h = (float)(-kCp * 0.5);
  const float cr = (float)(-kRd / kCv);
  const float r = rc[K];
  const float r8w0 = 0.5f * (rc[K - 1] + rc[K]);
  const float r8w1 = 0.5f * (rc[K] + rc[K + 1]);
  const float dpi0lo = (pi0c[K] - pi0c[K - 1]) * d.idz;
  const float dpi0hi = (pi0c[K + 1] - pi0c[K]) * d.idz;
  const float pi0k = pi0c[K];
  const int iend = min(i0 + kAX, d.nx);
  // carried x-direction values at cell I (and ps at I-1) and the new velocity of face I
  long c = ic(d, i0 + 3, J, K);
  float p_c = p[c], tr_c = tr[c], ps_c = ps[c], ps_m = ps[c - sx];
  float un0 = u[iu(d, i0 + 3, J, K)] +
              (h * (tr_c + tr[c - sx]) * ((p_c - p[c - sx]) * d.idx) + ru[((long)i0 * d.ny + jc) * d.nz + kc]) * dts;
  for (int i = i0; i < iend; ++i) {
    const int I = i + 3;
    c = ic(d, I, J, K);
    // x: upper face I+1
    const float p_p = p[c + sx], tr_p = tr[c + sx], ps_p = ps[c + sx];
    const float un1 = u[iu(d, I + 1, J, K)] +
                      (h * (tr_p + tr_c) * ((p_p - p_c) * d.idx) + ru[((long)(i + 1) * d.ny + jc) * d.nz + kc]) * dts;
    // y: faces J and J+1
    const float p_s = p[c - sy], p_n = p[c + sy], tr_s = tr[c - sy], tr_n = tr[c + sy];
    const float ps_s = ps[c - sy], ps_n = ps[c + sy];
    const float vn0 = v[iv(d, I, J, K)] +
                      (h * (tr_c + tr_s) * ((p_c - p_s) * d.idy) + rv[((long)i * (d.ny + 1) + jc) * d.nz + kc]) * dts;
    const float vn1 = v[iv(d, I, J + 1, K)] +
                      (h * (tr_n + tr_c) * ((p_n - p_c) * d.idy) + rv[((long)i * (d.ny + 1) + jc + 1) * d.nz + kc]) * dts;
    // z: neighbours from the adjacent lanes (lane 0 / 63 of the wave load their own)
    float p_b = __shfl_up(p_c, 1, 64), tr_b = __shfl_up(tr_c, 1, 64), ps_b = __shfl_up(ps_c, 1, 64);
    float p_t = __shfl_down(p_c, 1, 64), tr_t = __shfl_down(tr_c, 1, 64), ps_t = __shfl_down(ps_c, 1, 64);
    if (lane == 0) {
      p_b = p[c - 1];
      tr_b = tr[c - 1];
      ps_b = ps[c - 1];
    }
    if (lane == 63 || k >= d.nz - 1) {
      p_t = p[c + 1];
      tr_t = tr[c + 1];
      ps_t = ps[c + 1];
    }
    // pgf_z is zero at the surface and lid levels, and set_w_bc zeroes the new w there
    const float wn0 = (kc == 0) ? 0.0f
                                : w[iw(d, I, J, K)] + (h * (tr_c + tr_b) * ((p_c - p_b) * d.idz) +
                                                       rw[((long)i * d.ny + jc) * (d.nz + 1) + kc]) * dts;
    const float wn1 = (kc == d.nz - 1)
                          ? 0.0f
                          : w[iw(d, I, J, K + 1)] + (h * (tr_t + tr_c) * ((p_t - p_c) * d.idz) +
                                                     rw[((long)i * d.ny + jc) * (d.nz + 1) + kc + 1]) * dts;
    const float div = (un1 - un0) * d.idx + (vn1 - vn0) * d.idy + (wn1 - wn0) * d.idz;
    const float udx0 = un0 * ((ps_c - ps_m) * d.idx);
    const float udx1 = un1 * ((ps_p - ps_c) * d.idx);
    const float vdy0 = vn0 * ((ps_c - ps_s) * d.idy);
    const float vdy1 = vn1 * ((ps_n - ps_c) * d.idy);
    const float dz0 = dpi0lo + (ps_c - ps_b) * d.idz;
    const float dz1 = dpi0hi + (ps_t - ps_c) * d.idz;
    const float wdz0 = (kc == 0) ? 0.0f : wn0 * dz0;
    const float wdz1 = (kc == d.nz - 1) ? 0.0f : wn1 * dz1;
    const float advp = -(0.5f * (udx1 + udx0) + 0.5f * (vdy1 + vdy0) + (0.5f * wdz0 * r8w0 + 0.5f * wdz1 * r8w1) / r);
    const float dpdt = cr * ((pi0k + ps_c) * div) + advp;
    const float pn = p_c + 1.1f * dpdt * dts;
    if (act) {
      store_c_fast(d, po, i, j, k, pn);
      store_u_fast(d, uo, i, j, k, un0);
      if (i == d.nx - 1) store_u(d, uo, d.nx, j, k, un1);
      store_v_fast(d, vo, i, j, k, vn0);
      if (j == d.ny - 1) store_v(d, vo, i, d.ny, k, vn1);
      store_w_fast(d, wo, i, j, k, wn0);
      if (k == d.nz - 1) store_w(d, wo, i, j, d.nz, wn1);
    }
    un0 = un1;
    ps_m = ps_c;
    p_c = p_p;
    tr_c = tr_p;
    ps_c = ps_p;
  }
}

// ----------------------------------------------------------------------------------
// E1 / M1: Euler updates (one_step_integration.py:450-465, :876-891) and SGS add-in
// (:113-118, :158-162): next = pad(base + tend*dt) [+ pad(sgs*dt)]
// ----------------------------------------------------------------------------------
__global__ void k_scalar_update(Dims d, const float* __restrict__ rc, const float* __restrict__ s0,
                                const float* __restrict__ s, const float* __restrict__ u, const float* __restrict__ v,
                                const float* __restrict__ w, const float* __restrict__ fd, const float* __restrict__ sfc,
                                const float* __restrict__ heat, const float* __restrict__ sgs, float* __restrict__ out,
                                float* __restrict__ tend_out, float dt, float weps) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  int I = i + 3, J = j + 3, K = k + 1;
  long o = ((long)i * d.ny + j) * d.nz + k;
  float sf = sfc ? sfc[(long)i * d.ny + j] : 0.0f;
  float t = adv_scalar_point(d, rc, s, u, v, w, fd[ic(d, I, J, K)], sf, i, j, k, weps);
  if (heat) t = t + heat[o];
  float nxt = s0[ic(d, I, J, K)] + t * dt;
  if (sgs) {
    nxt = nxt + sgs[o] * dt;
    t = t + sgs[o];
  }
  if (tend_out) tend_out[o] = t;
  store_c_fast(d, out, i, j, k, nxt);
}

__global__ void k_momentum_update(Dims d, const float* __restrict__ u0, const float* __restrict__ v0,
                                  const float* __restrict__ w0, const float* __restrict__ p, const float* __restrict__ tr,
                                  const float* __restrict__ au, const float* __restrict__ av, const float* __restrict__ aw,
                                  const float* __restrict__ fu, const float* __restrict__ fv, const float* __restrict__ b8w,
                                  const float* __restrict__ su, const float* __restrict__ sv, const float* __restrict__ sw,
                                  float* __restrict__ uo, float* __restrict__ vo, float* __restrict__ wo,
                                  float* __restrict__ duo, float* __restrict__ dvo, float* __restrict__ dwo, float dt) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nz
  int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = blockIdx.z;                             // 0..nx
  if (k > d.nz || j > d.ny) return;
  int I = i + 3, J = j + 3, K = k + 1;
  if (k < d.nz && j < d.ny) {
    long o = ((long)i * d.ny + j) * d.nz + k;
    float du = au[o] + pgf_x(d, p, tr, I, J, K) + (fv ? fv[o] : 0.0f);
    float un = u0[iu(d, I, J, K)] + du * dt;
    if (su) un = un + su[o] * dt;
    if (duo) duo[o] = su ? du + su[o] : du;
    store_u(d, uo, i, j, k, un);
  }
  if (k < d.nz && i < d.nx) {
    long o = ((long)i * (d.ny + 1) + j) * d.nz + k;
    float dv = av[o] + pgf_y(d, p, tr, I, J, K) - (fu ? fu[o] : 0.0f);
    float vn = v0[iv(d, I, J, K)] + dv * dt;
    if (sv) vn = vn + sv[o] * dt;
    if (dvo) dvo[o] = sv ? dv + sv[o] : dv;
    store_v(d, vo, i, j, k, vn);
  }
  if (i < d.nx && j < d.ny) {
    long o = ((long)i * d.ny + j) * (d.nz + 1) + k;
    float dw = aw[o] + pgf_z(d, p, tr, I, J, K) + b8w[o];
    float wn = w0[iw(d, I, J, K)] + dw * dt;
    // set_w_bc, then the SGS add (sgs_w is zero at surface and lid)
    if (k == 0 || k == d.nz) wn = 0.0f;
    if (sw) wn = wn + sw[o] * dt;
    if (dwo) dwo[o] = sw ? dw + sw[o] : dw;
    store_w(d, wo, i, j, k, wn);
  }
}

}  // namespace lex
