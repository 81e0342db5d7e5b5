// Bulk atmosphere-ocean fluxes at the lowest model level: the CAM scheme of
// atm_ocn_flux (boundary_conditions.py:7-75) with q_sat_sfc, cdn and psi_m_s
// (:78-103), a fixed eight passes of the Monin-Obukhov iteration, one thread per column.
//   k_atm_ocn_flux      : inputs sliced from the padded state as rk_sub_step0 does
//                         (one_step_integration.py:37-47): u, v averaged to the cell
//                         centre, theta, qv, rho0 of level 1, z = dz/2; also writes
//                         sen/Cp, the bottom theta flux of the scalar update (:151)
//   k_atm_ocn_flux_cols : the reference's own signature on (n,) arrays
// Outputs keep the reference's final sign flip (upward fluxes positive).
#pragma once
#include "lex_common.hpp"

namespace lex {

constexpr int kAtmOcnIters = 8;          // namelist_n_constants.py:43
constexpr float kAtmOcnMinWind = 0.25f;  // :44
constexpr float kSfcZRef = 10.0f;        // :45
constexpr float kSfcZTRef = 2.0f;        // :46
constexpr float kKarman = 0.4f;          // :47
constexpr float kLatVap = 2.501e6f;      // :60

struct SfcOut {
  float* tau_x;
  float* tau_y;
  float* sen;
  float* evap;
  float* t_ref;
  float* q_ref;
  float* u10n;
  float* sen_cp;   // sen / Cp (may be null)
};

struct PsiMS {
  float m, s;
};

__device__ __forceinline__ PsiMS psi_m_s(float hol) {
  const float xsq = fmaxf(sqrtf(fabsf(1.0f - 16.0f * hol)), 1.0f);
  const float xqq = sqrtf(xsq);
  PsiMS p;
  if (hol > 0.0f) {
    p.m = -5.0f * hol;
    p.s = p.m;
  } else {
    p.m = logf((1.0f + xqq * (2.0f + xqq)) * (1.0f + xqq * xqq) / 8.0f) - 2.0f * atanf(xqq) + 1.571f;
    p.s = 2.0f * logf((1.0f + xqq * xqq) / 2.0f);
  }
  return p;
}

__device__ __forceinline__ float cdn_neutral(float u) { return 0.0027f / u + 0.000142f + 0.0000764f * u; }

__device__ __forceinline__ void atm_ocn_column(float z, float ub, float vb, float th, float q, float rho, float ts,
                                               const SfcOut& o, long p) {
  const float cpf = (float)kCp;
  const float cpvir = (float)(1810.0 / kCp - 1.0);
  const float v_mag = fmaxf(sqrtf(ub * ub + vb * vb), kAtmOcnMinWind);
  const float ssq = 0.98f * (640380.0f / expf(5107.4f / ts)) / rho;
  const float dth = th - ts;
  const float dq = q - ssq;
  const float alz = logf(z / kSfcZRef);
  const float cp = cpf * (1.0f + cpvir * ssq);
  float rdn = sqrtf(cdn_neutral(v_mag));
  float rhn = (dth > 0.0f) ? 0.018f : 0.0327f;
  const float ren = 0.0346f;
  float u_star = rdn * v_mag, t_star = rhn * dth, q_star = ren * dq;
  const float qfac = 1.0f / (float)kRepsm1;
  const float kg = kKarman * (float)kG;
  float hol = 0.0f, rh = 0.0f, re = 0.0f, psi_xh = 0.0f, u10n = v_mag;
  for (int it = 0; it < kAtmOcnIters; ++it) {
    hol = kg * z * (t_star / th + q_star / (qfac + q)) / (u_star * u_star);
    hol = fminf(fmaxf(hol, -10.0f), 10.0f);
    const PsiMS ps = psi_m_s(hol);
    psi_xh = ps.s;
    float rd = rdn / (1.0f + rdn / kKarman * (alz - ps.m));
    u10n = v_mag * rd / rdn;
    rdn = sqrtf(cdn_neutral(u10n));
    rhn = (hol > 0.0f) ? 0.018f : 0.0327f;
    rd = rdn / (1.0f + rdn / kKarman * (alz - ps.m));
    rh = rhn / (1.0f + rhn / kKarman * (alz - ps.s));
    re = ren / (1.0f + ren / kKarman * (alz - ps.s));
    u_star = rd * v_mag;
    t_star = rh * dth;
    q_star = re * dq;
  }
  const float tau = rho * u_star * u_star;
  const float sen = cp * tau * t_star / u_star;
  const float evap = kLatVap * tau * q_star / u_star / kLatVap;
  const float psi_x2 = psi_m_s(hol * kSfcZTRef / z).s;
  const float al2 = logf(kSfcZRef / kSfcZTRef);
  const float prof = alz + al2 - psi_xh + psi_x2;
  o.tau_x[p] = -(tau * ub / v_mag);
  o.tau_y[p] = -(tau * vb / v_mag);
  o.sen[p] = -sen;
  o.evap[p] = -evap;
  o.t_ref[p] = th - dth * ((rh / kKarman) * prof) - 0.01f * kSfcZTRef;
  o.q_ref[p] = q - dq * ((re / kKarman) * prof);
  o.u10n[p] = u10n;
  if (o.sen_cp) o.sen_cp[p] = -sen / cpf;
}

// one thread per column (i, j): lanes along y, blocks along x
__global__ void k_atm_ocn_flux(Dims d, const float* __restrict__ rc, const float* __restrict__ th,
                               const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ qv,
                               const float* __restrict__ surface_t, SfcOut o) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= d.ny || i >= d.nx) return;
  const int I = i + 3, J = j + 3, K = 1;
  const float ub = 0.5f * (u[iu(d, I, J, K)] + u[iu(d, I + 1, J, K)]);
  const float vb = 0.5f * (v[iv(d, I, J, K)] + v[iv(d, I, J + 1, K)]);
  const long p = (long)i * d.ny + j;
  atm_ocn_column(0.5f * d.dz, ub, vb, th[ic(d, I, J, K)], qv[ic(d, I, J, K)], rc[K], surface_t[p], o, p);
}

__global__ void k_atm_ocn_flux_cols(long n, const float* __restrict__ z, const float* __restrict__ ub,
                                    const float* __restrict__ vb, const float* __restrict__ th,
                                    const float* __restrict__ q, const float* __restrict__ rho,
                                    const float* __restrict__ ts, SfcOut o) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  atm_ocn_column(z[p], ub[p], vb[p], th[p], q[p], rho[p], ts[p], o, p);
}

inline void launch_atm_ocn_flux(const Dims& d, const float* rc, const float* th, const float* u, const float* v,
                                const float* qv, const float* surface_t, const SfcOut& o, hipStream_t st) {
  hipLaunchKernelGGL(k_atm_ocn_flux, dim3((d.ny + 63) / 64, d.nx, 1), dim3(64, 1, 1), 0, st, d, rc, th, u, v, qv,
                     surface_t, o);
}

}  // namespace lex
