// solver_opt 3: vertically implicit acoustic substep (sound_implicit_rk_step0/1/2,
// one_step_integration.py:559-635, :678-754, :797-873).
//
// u and v are stepped forward as in the explicit substep; w and pi' are coupled
// implicitly in the vertical with off-centring alpha (namelist_n_constants.py:14), which
// leaves one tridiagonal system per column for the new w at levels 1..nz-1.  Columns are
// independent within a substep (the horizontal terms are explicit), so one workgroup owns
// one column: lane k holds cell k (its lower w level k), the column's pressure tendency
// and the system live in LDS, and the system is solved by parallel cyclic reduction
// (ceil(log2 nz) rounds of neighbour elimination at doubling distance, every lane one
// equation; the reference calls jax.lax.linalg.tridiagonal_solve).  Everything else -
// PGF, divergence, advection_pi, the 1.1 divergence damping, halos - is the explicit
// substep's arithmetic (k_acoustic).
#pragma once
#include "lex_common.hpp"
#include "lex_ops.hip"

namespace lex {

constexpr float kAlphaImplicit = 0.6f;  // off-centring of the implicit step (namelist_n_constants.py:14)

// column workgroup: nzl = 64 * ceil(nz / 64) lanes; LDS 6 * nzl + 1 floats
inline int acoustic_vi_lanes(const Dims& d) { return 64 * ((d.nz + 63) / 64); }
inline size_t acoustic_vi_lds(const Dims& d) { return (6 * (size_t)acoustic_vi_lanes(d) + 1) * sizeof(float); }

__device__ __forceinline__ float advpi_cell(float un0, float un1, float vn0, float vn1, float wn0, float wn1, float sc,
                                            float sxm, float sxp, float sym, float syp, float szm, float szp, float pim,
                                            float pic, float pip, float rcm, float rcc, float rcp, bool bottom, bool top,
                                            const Dims& d) {
  // advection_pi (advection.py:152-197) as in k_acoustic; the w term is zero at the lids
  float udx0 = un0 * ((sc - sxm) * d.idx);
  float udx1 = un1 * ((sxp - sc) * d.idx);
  float vdy0 = vn0 * ((sc - sym) * d.idy);
  float vdy1 = vn1 * ((syp - sc) * d.idy);
  float dz0 = (pic - pim) * d.idz + (sc - szm) * d.idz;
  float dz1 = (pip - pic) * d.idz + (szp - sc) * d.idz;
  float wdz0 = bottom ? 0.0f : wn0 * dz0;
  float wdz1 = top ? 0.0f : wn1 * dz1;
  float r8w0 = 0.5f * (rcm + rcc);
  float r8w1 = 0.5f * (rcc + rcp);
  return -(0.5f * (udx1 + udx0) + 0.5f * (vdy1 + vdy0) + (0.5f * wdz0 * r8w0 + 0.5f * wdz1 * r8w1) / rcc);
}

__global__ void __launch_bounds__(256) k_acoustic_vi(Dims d, const float* __restrict__ rc,
                                                     const float* __restrict__ pi0c, const float* __restrict__ u,
                                                     const float* __restrict__ v, const float* __restrict__ w,
                                                     const float* __restrict__ p, const float* __restrict__ ps,
                                                     const float* __restrict__ ru, const float* __restrict__ rv,
                                                     const float* __restrict__ rw, const float* __restrict__ tr,
                                                     float* __restrict__ uo, float* __restrict__ vo,
                                                     float* __restrict__ wo, float* __restrict__ po, float dts,
                                                     float alpha) {
  extern __shared__ float lds[];
  const int nzl = blockDim.x, nz = d.nz;
  float* A = lds;
  float* B = A + nzl;
  float* C = B + nzl;
  float* D = C + nzl;
  float* PT = D + nzl;   // pressure tendency with beta * w_now, per cell
  float* WN = PT + nzl;  // new w at levels 0..nz
  const TileId tb = xcd_tile();
  const int j = tb.x, i = tb.y;
  const int k = threadIdx.x;
  const bool act = k < nz;
  const int kc = act ? k : nz - 1;
  const int I = i + 3, J = j + 3, K = kc + 1;
  const long c = ic(d, I, J, K);
  const long sx = d.sxc(), sy = d.syc();
  const float beta = 1.0f - alpha;
  const float h = (float)(-kCp * 0.5);
  const float cr = (float)(-kRd / kCv);
  // ---- loads (all in bounds: levels K-2 .. K+1 clamped at 0)
  const float rcm = rc[K - 1], rcc = rc[K], rcp = rc[K + 1];
  const float pim2 = pi0c[max(K - 2, 0)], pim = pi0c[K - 1], pic = pi0c[K], pip = pi0c[K + 1];
  const long cu = iu(d, I, J, K), cv = iv(d, I, J, K), cw = iw(d, I, J, K);
  const long ou = ((long)i * d.ny + j) * nz + kc, ov = ((long)i * (d.ny + 1) + j) * nz + kc;
  const long ow = ((long)i * d.ny + j) * (nz + 1) + kc;
  const float u0 = u[cu], u1 = u[cu + d.sxu()], v0 = v[cv], v1 = v[cv + d.syc()], w0 = w[cw], w1 = w[cw + 1];
  const float ru0 = ru[ou], ru1 = ru[ou + (long)d.ny * nz], rv0 = rv[ov], rv1 = rv[ov + nz], rw0 = rw[ow];
  const float pc = p[c], pxm = p[c - sx], pxp = p[c + sx], pym = p[c - sy], pyp = p[c + sy], pzm = p[c - 1];
  const float tc = tr[c], txm = tr[c - sx], txp = tr[c + sx], tym = tr[c - sy], typ = tr[c + sy], tzm = tr[c - 1];
  const float sc = ps[c], sxm = ps[c - sx], sxp = ps[c + sx], sym = ps[c - sy], syp = ps[c + sy], szm = ps[c - 1],
              szp = ps[c + 1], szm2 = ps[c - (K >= 2 ? 2 : 1)];
  // ---- explicit u, v (pressure_gradient_force_horizontal with pi' now)
  const float un0 = u0 + (h * (tc + txm) * ((pc - pxm) * d.idx) + ru0) * dts;
  const float un1 = u1 + (h * (txp + tc) * ((pxp - pc) * d.idx) + ru1) * dts;
  const float vn0 = v0 + (h * (tc + tym) * ((pc - pym) * d.idy) + rv0) * dts;
  const float vn1 = v1 + (h * (typ + tc) * ((pyp - pc) * d.idy) + rv1) * dts;
  const bool bottom = kc == 0, top = kc == nz - 1;
  const float piadv = pic + sc;
  // ---- pressure tendency with beta * w_now (conv_mix + adv4pi_mix)
  {
    const float wb0 = beta * w0, wb1 = beta * w1;
    const float div = (un1 - un0) * d.idx + (vn1 - vn0) * d.idy + (wb1 - wb0) * d.idz;
    const float adv = advpi_cell(un0, un1, vn0, vn1, wb0, wb1, sc, sxm, sxp, sym, syp, szm, szp, pim, pic, pip, rcm,
                                 rcc, rcp, bottom, top, d);
    PT[k] = act ? cr * (piadv * div) + adv : 0.0f;
  }
  __syncthreads();
  // ---- the system at level l = k (1..nz-1): cells below (padded K-1) and above (K)
  float a = 0.f, bb = 1.f, cc = 0.f, dd = 0.f;
  if (act && k >= 1) {
    const float thr8w = 0.5f * (tc + tzm);
    const float lhs = (alpha * dts) * (alpha * dts) * (float)kCp * thr8w * d.idz;
    const float pa_m2 = pim2 + szm2, pa_m = pim + szm, pa_c = pic + sc, pa_p = pip + szp;
    const float rdcv = (float)(kRd / kCv);
    a = (k == 1) ? 0.0f : -lhs * (rdcv * pa_m * d.idz - (pa_m - pa_m2) * d.idz * 0.5f);
    cc = (k == nz - 1) ? 0.0f : -lhs * (rdcv * pa_c * d.idz + (pa_p - pa_c) * d.idz * 0.5f);
    bb = 1.0f + lhs * rdcv * (pa_c * d.idz + pa_m * d.idz);
    const float rhs3 = alpha * dts * dts * (float)kCp * thr8w;
    const float pw = h * (tc + tzm) * ((pc - pzm) * d.idz);
    dd = w0 + dts * (rw0 + pw) - rhs3 * ((PT[k] - PT[k - 1]) * d.idz);
  }
  // ---- parallel cyclic reduction over the column (rows 0 and >= nz are identity rows)
  for (int s = 1; s < nz; s <<= 1) {
    const bool hm = k - s >= 0, hp = k + s < nzl;
    __syncthreads();
    A[k] = a;
    B[k] = bb;
    C[k] = cc;
    D[k] = dd;
    __syncthreads();
    const float am = hm ? A[k - s] : 0.f, bm = hm ? B[k - s] : 1.f, cm = hm ? C[k - s] : 0.f, dm = hm ? D[k - s] : 0.f;
    const float ap = hp ? A[k + s] : 0.f, bp = hp ? B[k + s] : 1.f, cp = hp ? C[k + s] : 0.f, dp = hp ? D[k + s] : 0.f;
    const float k1 = a / bm, k2 = cc / bp;
    a = -am * k1;
    cc = -cp * k2;
    bb = bb - cm * k1 - ap * k2;
    dd = dd - dm * k1 - dp * k2;
  }
  WN[k] = (act && k >= 1) ? dd / bb : 0.0f;
  if (k == 0) WN[nz] = 0.0f;
  __syncthreads();
  if (!act) return;
  // ---- pi' with w_mix = alpha w_new + beta w_now
  const float wn0 = WN[k], wn1 = WN[k + 1];
  const float wm0 = alpha * wn0 + beta * w0, wm1 = alpha * wn1 + beta * w1;
  const float div = (un1 - un0) * d.idx + (vn1 - vn0) * d.idy + (wm1 - wm0) * d.idz;
  const float adv = advpi_cell(un0, un1, vn0, vn1, wm0, wm1, sc, sxm, sxp, sym, syp, szm, szp, pim, pic, pip, rcm,
                               rcc, rcp, bottom, top, d);
  const float pn = pc + 1.1f * (cr * (piadv * div) + adv) * dts;
  store_c_fast(d, po, i, j, k, pn);
  store_u_fast(d, uo, i, j, k, un0);
  if (i == d.nx - 1) store_u(d, uo, d.nx, j, k, un1);
  store_v_fast(d, vo, i, j, k, vn0);
  if (j == d.ny - 1) store_v(d, vo, i, d.ny, k, vn1);
  store_w_fast(d, wo, i, j, k, wn0);
  if (k == nz - 1) store_w(d, wo, i, j, nz, 0.0f);
}

}  // namespace lex
