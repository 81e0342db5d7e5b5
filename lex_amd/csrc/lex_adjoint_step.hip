// Reverse-mode driver of one compressible WS-RK3 step (solver_opt 2).
//
// rk3_step_vjp re-runs the step with the per-operator kernels, recording a tape of every
// stage state, the frozen SGS fields, the per-stage momentum rhs and theta_rho, and the
// input of every acoustic substep; it then walks the stages backwards (2, 1, 0) and the
// Smagorinsky chain last, accumulating the adjoints of the step's input state and of Cs.
// Forward structure: rk_sub_step0/1/2 (one_step_integration.py:16-447) and
// ws_rk3_big_step (one_sprint_functions.py:10-30), as in rk3_step (lex_stage.hip).
#pragma once
#include <algorithm>
#include "lex_adjoint.hip"
#include "lex_stage.hip"

namespace lex {

struct VjpLayout {
  size_t nc, nu, nv, nw, iu_, iv_, iw_, ic_;
  size_t state;                       // bytes of one state block
  size_t acou;                        // bytes of one (u, v, w, pip) block
  int nsub[3];
  int np;
  size_t off_XA, off_XB, off_Xn, off_sm, off_su, off_sv, off_sw, off_ss, off_fd;
  size_t off_tr[3], off_ru[3], off_rv[3], off_rw[3];
  size_t off_ac;                      // acoustic substep inputs t >= 1, stage by stage
  size_t off_lXA, off_lXB, off_lP[2], off_lU, off_lV, off_lW, off_lru, off_lrv, off_lrw, off_ltr;
  size_t off_lsu, off_lsv, off_lsw, off_lss, off_lT, off_lsm, off_lcg2;
  size_t off_slots;                   // slot planes of the advection reverses (never cleared)
  size_t total;
};

inline void n_substeps(double dt, int n_sound, int* nsub) {
  const double dts_small = dt / (double)n_sound;
  nsub[0] = 1;
  nsub[1] = (int)nearbyint((dt / 2.0) / dts_small);
  nsub[2] = (int)nearbyint(dt / dts_small);
}

inline VjpLayout vjp_layout(const Dims& d, double dt, int n_sound, int np) {
  VjpLayout L;
  L.np = np;
  L.nc = d.nc(); L.nu = d.nu(); L.nv = d.nv(); L.nw = d.nw();
  L.iu_ = (size_t)(d.nx + 1) * d.ny * d.nz;
  L.iv_ = (size_t)d.nx * (d.ny + 1) * d.nz;
  L.iw_ = (size_t)d.nx * d.ny * (d.nz + 1);
  L.ic_ = (size_t)d.nx * d.ny * d.nz;
  n_substeps(dt, n_sound, L.nsub);
  L.state = al256(L.nu * 4) + al256(L.nv * 4) + al256(L.nw * 4) + (3 + np) * al256(L.nc * 4);
  L.acou = al256(L.nu * 4) + al256(L.nv * 4) + al256(L.nw * 4) + al256(L.nc * 4);
  const size_t C = al256(L.nc * 4), U = al256(L.iu_ * 4), V = al256(L.iv_ * 4), W = al256(L.iw_ * 4),
               S = al256(L.ic_ * 4);
  size_t o = 0;
  L.off_XA = o; o += L.state;
  L.off_XB = o; o += L.state;
  L.off_Xn = o; o += L.state;
  L.off_sm = o; o += 10 * C;
  L.off_su = o; o += U;
  L.off_sv = o; o += V;
  L.off_sw = o; o += W;
  L.off_ss = o; o += (2 + np) * S;
  L.off_fd = o; o += C;
  for (int s = 0; s < 3; ++s) {
    L.off_tr[s] = o; o += C;
    L.off_ru[s] = o; o += U;
    L.off_rv[s] = o; o += V;
    L.off_rw[s] = o; o += W;
  }
  L.off_ac = o; o += (size_t)(L.nsub[0] + L.nsub[1] + L.nsub[2] - 3) * L.acou;
  L.off_lXA = o; o += L.state;
  L.off_lXB = o; o += L.state;
  L.off_lP[0] = o; o += L.acou;
  L.off_lP[1] = o; o += L.acou;
  L.off_lU = o; o += U;
  L.off_lV = o; o += V;
  L.off_lW = o; o += W;
  L.off_lru = o; o += U;
  L.off_lrv = o; o += V;
  L.off_lrw = o; o += W;
  L.off_ltr = o; o += C;
  L.off_lsu = o; o += U;
  L.off_lsv = o; o += V;
  L.off_lsw = o; o += W;
  L.off_lss = o; o += (2 + np) * S;
  L.off_lT = o; o += S;
  L.off_lsm = o; o += 10 * C;
  L.off_lcg2 = o; o += 256 + al256((size_t)kSmagKBlocks * sizeof(double));  // + reduction partials
  // the scalar reverse and the momentum reverses use the planes in turn; every launch
  // rewrites all of its planes, so they need no clearing
  L.off_slots = o; o += al256(std::max(slot_bytes_scalar(d), slot_bytes_mom(d)));
  L.total = o;
  return L;
}

// index into the acoustic tape: stage s, substep t >= 1
inline size_t ac_off(const VjpLayout& L, int s, int t) {
  size_t k = 0;
  for (int q = 0; q < s; ++q) k += (size_t)(L.nsub[q] - 1);
  return L.off_ac + (k + (size_t)(t - 1)) * L.acou;
}

__global__ void k_cs_grad(const double* __restrict__ lcg2, double factor, double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] += lcg2[0] * factor;
}

__global__ void k_zero_d(double* __restrict__ a) {
  if (blockIdx.x == 0 && threadIdx.x == 0) a[0] = 0.0;
}

struct VjpCtx {
  const Dims& d;
  hipStream_t st;
  dim3 bk;
  dim3 G(int a, int b, int c) const { return dim3((a + 63) / 64, (b + 3) / 4, c); }
  void zero(float* p, size_t n) const {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (long)n, p, 0.0f);
  }
  void fold(int mode, float* a) const { launch_fold(d, mode, a, st); }
};

// padded field sizes of a state block in state_ptrs order (theta, u, v, w, pip, qv, ...)
inline size_t field_n(const VjpLayout& L, int q) {
  return q == 1 ? L.nu : (q == 2 ? L.nv : (q == 3 ? L.nw : L.nc));
}
inline int field_mode(int q) { return q == 1 ? 1 : (q == 2 ? 2 : (q == 3 ? 3 : 0)); }

// Forward pass with tape.  X0: the step's input state (not modified).
template <class O, class B>
void rk3_forward_tape(const Dims& d, const O& o, const B& b, float* const* X0, char* ws, const VjpLayout& L,
                      hipStream_t st, float* const* out = nullptr) {
  VjpCtx cx{d, st, dim3(64, 4, 1)};
  const int nf = 6 + o.n_passive, nsc = 2 + o.n_passive;
  Rk3Layout RL;
  RL.np = o.n_passive; RL.nc = L.nc; RL.nu = L.nu; RL.nv = L.nv; RL.nw = L.nw;
  float *XA[16], *XB[16], *Xn[16];
  state_ptrs(ws + L.off_XA, RL, XA);
  state_ptrs(ws + L.off_XB, RL, XB);
  state_ptrs(ws + L.off_Xn, RL, Xn);
  (void)nf;
  float* sm[10];
  for (int q = 0; q < 10; ++q) sm[q] = (float*)(ws + L.off_sm + q * al256(L.nc * 4));
  float* su = (float*)(ws + L.off_su);
  float* sv = (float*)(ws + L.off_sv);
  float* sw = (float*)(ws + L.off_sw);
  float* ss[10];
  for (int q = 0; q < nsc; ++q) ss[q] = (float*)(ws + L.off_ss + q * al256(L.ic_ * 4));
  float* fd = (float*)(ws + L.off_fd);
  const float* rc = b.rho0;
  const bool turb = o.turb != 0;
  const double dtd = (double)o.dt;
  const double dts_small = dtd / (double)o.n_sound;
  for (int s = 0; s < 3; ++s) {
    const double dt_s = (s == 0) ? dtd / 3.0 : (s == 1 ? dtd / 2.0 : dtd);
    float* const* src = (s == 0) ? X0 : (s == 1 ? XA : XB);
    float* const* dst = (s == 0) ? XA : (s == 1 ? XB : (out ? out : (float* const*)Xn));
    float* tr = (float*)(ws + L.off_tr[s]);
    float* ru = (float*)(ws + L.off_ru[s]);
    float* rv = (float*)(ws + L.off_rv[s]);
    float* rw = (float*)(ws + L.off_rw[s]);
    // turb 1: Smagorinsky recorded here; kTurbExternal: the caller's SGS tendencies are
    // already in the tape's SGS slots
    if (s == 0 && o.turb == 1) {
      hipLaunchKernelGGL(k_smag_deform, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, rc, src[1], src[2], src[3], sm[0],
                         sm[1], sm[2], sm[3], sm[4], sm[5], sm[6]);
      hipLaunchKernelGGL(k_smag_nm, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, src[0], src[5], sm[7]);
      float gs = cbrtf(d.dx * d.dy * d.dz);
      float cg = o.cs * gs;
      long n = (long)L.nc;
      hipLaunchKernelGGL(k_smag_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, sm[6], sm[7], sm[8], sm[9],
                         cg * cg, o.prandtl);
      hipLaunchKernelGGL(k_smag_uvw<false>, cx.G(d.nz, d.ny + 1, d.nx + 1), cx.bk, 0, st, d, rc, sm[8], sm[8], sm[0],
                         sm[1], sm[2], sm[3], sm[4], sm[5], su, sv, sw);
      for (int q = 0; q < nsc; ++q) {
        int fi = (q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2));
        hipLaunchKernelGGL(k_smag_scalar, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, rc, sm[9], sm[9], src[fi], ss[q]);
      }
    }
    // the fused stage kernels of the forward step where they apply (same trajectory as
    // lex_rk3_step impl 0), else one kernel per reference routine
    const int nzl = 64 * ((d.nz + 63) / 64);
    const bool fused = nzl * kFTY <= 1024 && nsc >= 2 && nsc <= 6;
    if (fused) {
      dim3 fb(nzl, kFTY, 1), fg((d.nx + kFTX - 1) / kFTX, (d.ny + kFTY - 1) / kFTY, 1);
      // the taped forward covers cor_opt 0 (its reverse has no Coriolis term; the entry
      // points reject cor != 0)
      const CorOpt nocor{0, 0.0f, 0.0f, 0.0f};
      hipLaunchKernelGGL(k_mom_u, fg, fb, mom_uv_lds(d), st, d, rc, src[1], src[2], src[3],
                         turb ? (const float*)su : nullptr, (const float*)nullptr, nocor, ru);
      hipLaunchKernelGGL(k_mom_v, fg, fb, mom_uv_lds(d), st, d, rc, src[1], src[2], src[3],
                         turb ? (const float*)sv : nullptr, (const float*)nullptr, nocor, rv);
      hipLaunchKernelGGL(k_mom_w, fg, fb, mom_w_lds(d), st, d, rc, b.theta0, b.qv0, src[1], src[2], src[3], src[0],
                         src[5], turb ? (const float*)sw : nullptr, rw);
      ScalarSet S;
      S.heat = nullptr;
      S.n = nsc;
      for (int q = 0; q < nsc; ++q) {
        int fi = (q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2));
        S.s0[q] = X0[fi];
        S.s[q] = src[fi];
        S.sgs[q] = turb ? ss[q] : nullptr;
        S.out[q] = dst[fi];
        S.weps[q] = (q == 0) ? 1.0e-17f : 1.0e-18f;
      }
      size_t lds = stage_scalars_lds(d);
      float dts_f = (float)dt_s;
      hipLaunchKernelGGL(k_stage_scalars, dim3(nsc, fg.x, fg.y), fb, lds, st, d, rc, src[1], src[2], src[3], S, tr,
                         dts_f);
    } else {
      hipLaunchKernelGGL(k_div_rho, cx.G(d.nz + 2, d.ny + 6, d.nx + 6), cx.bk, 0, st, d, rc, src[1], src[2], src[3],
                         fd);
      hipLaunchKernelGGL(k_rhs_uv, cx.G(d.nz, d.ny + 1, d.nx + 1), cx.bk, 0, st, d, rc, src[1], src[2], src[3], fd,
                         turb ? (const float*)su : nullptr, turb ? (const float*)sv : nullptr, (const float*)nullptr,
                         (const float*)nullptr, CorOpt{0, 0.0f, 0.0f, 0.0f}, ru, rv);
      hipLaunchKernelGGL(k_rhs_w, cx.G(d.nz + 1, d.ny, d.nx), cx.bk, 0, st, d, rc, b.theta0, b.qv0, src[1], src[2],
                         src[3], src[0], src[5], fd, turb ? (const float*)sw : nullptr, rw);
      long n = (long)L.nc;
      hipLaunchKernelGGL(k_theta_rho, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, src[0], src[5], tr);
    }
    const int nsub = L.nsub[s];
    const float dts = (s == 0) ? (float)dt_s : (float)dts_small;
    float* in[4] = {X0[1], X0[2], X0[3], X0[4]};
    for (int t = 0; t < nsub; ++t) {
      float* out[4];
      if (t == nsub - 1) {
        out[0] = dst[1]; out[1] = dst[2]; out[2] = dst[3]; out[3] = dst[4];
      } else {
        acou_ptrs(ws + ac_off(L, s, t + 1), RL, out);
      }
      hipLaunchKernelGGL(k_acoustic, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, rc, b.pi0, in[0], in[1], in[2], in[3],
                         src[4], ru, rv, rw, tr, out[0], out[1], out[2], out[3], dts);
      for (int c = 0; c < 4; ++c) in[c] = out[c];
    }
    if (!fused) {
      for (int q = 0; q < nsc; ++q) {
        int fi = (q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2));
        float weps = (q == 0) ? 1.0e-17f : 1.0e-18f;
        hipLaunchKernelGGL(k_scalar_update, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, rc, X0[fi], src[fi], src[1],
                           src[2], src[3], fd, (const float*)nullptr, (const float*)nullptr,
                           turb ? (const float*)ss[q] : nullptr, dst[fi], (float*)nullptr, (float)dt_s, weps);
      }
    }
  }
}

// Adjoint of one step.  lam_out: adjoint of the step's output state (padded; folded in
// place).  lam_in: receives the adjoint of the input state (padded, folded: ghosts zero).
// lam_cs: accumulates dL/dCs.  X0 is the step's input state (not modified).
template <class O, class B>
int rk3_step_vjp(const Dims& d, const O& o, const B& b, float* const* X0, float* const* lam_out,
                 float* const* lam_in, double* lam_cs, char* ws, char* adj, const VjpLayout& L, hipStream_t st,
                 bool run_tape = true) {
  VjpCtx cx{d, st, dim3(64, 4, 1)};
  const int nf = 6 + o.n_passive, nsc = 2 + o.n_passive;
  if (run_tape) rk3_forward_tape(d, o, b, X0, ws, L, st);
  // adjoint buffers live in `adj` (offsets from off_lXA on); the tape in `ws`
  auto A = [&](size_t off) { return adj + (off - L.off_lXA); };
  Rk3Layout RL;
  RL.np = o.n_passive; RL.nc = L.nc; RL.nu = L.nu; RL.nv = L.nv; RL.nw = L.nw;
  float *XA[16], *XB[16], *lXA[16], *lXB[16];
  state_ptrs(ws + L.off_XA, RL, XA);
  state_ptrs(ws + L.off_XB, RL, XB);
  state_ptrs(A(L.off_lXA), RL, lXA);
  state_ptrs(A(L.off_lXB), RL, lXB);
  float* lP[2][4];
  acou_ptrs(A(L.off_lP[0]), RL, lP[0]);
  acou_ptrs(A(L.off_lP[1]), RL, lP[1]);
  float* sm[10];
  float* lsm[10];
  for (int q = 0; q < 10; ++q) {
    sm[q] = (float*)(ws + L.off_sm + q * al256(L.nc * 4));
    lsm[q] = (float*)(A(L.off_lsm) + q * al256(L.nc * 4));
  }
  // face weights g dts of the acoustic reverse, interior face layouts
  float* gU = (float*)(A(L.off_lU));
  float* gV = (float*)(A(L.off_lV));
  float* gW = (float*)(A(L.off_lW));
  float* lru = (float*)(A(L.off_lru));
  float* lrv = (float*)(A(L.off_lrv));
  float* lrw = (float*)(A(L.off_lrw));
  float* ltr = (float*)(A(L.off_ltr));
  float* lsu = (float*)(A(L.off_lsu));
  float* lsv = (float*)(A(L.off_lsv));
  float* lsw = (float*)(A(L.off_lsw));
  float* lss[10];
  for (int q = 0; q < nsc; ++q) lss[q] = (float*)(A(L.off_lss) + q * al256(L.ic_ * 4));
  float* lT = (float*)(A(L.off_lT));
  double* lcg2 = (double*)(A(L.off_lcg2));
  float* slots = (float*)(A(L.off_slots));
  const SlotPlanes spS = slots_scalar(d, slots);
  const SlotPlanes spU = slots_u(d, slots);
  const SlotPlanes spV = slots_v(d, slots + kSlotsMom * spU.plane);
  const SlotPlanes spW = slots_w(d, slots + kSlotsMom * (spU.plane + spV.plane));
  const float* rc = b.rho0;
  const bool turb = o.turb != 0;
  const double dtd = (double)o.dt;
  const double dts_small = dtd / (double)o.n_sound;

  {
    size_t cnt[16];
    for (int q = 0; q < nf; ++q) cnt[q] = field_n(L, q);
    launch_zero_many(nf, lam_in, cnt, st);
  }
  int fmodes[16];
  for (int q = 0; q < nf; ++q) fmodes[q] = field_mode(q);
  // the whole adjoint scratch region (stage adjoints, rhs, SGS and Smagorinsky adjoints)
  // is contiguous: one fill
  cx.zero((float*)A(L.off_lXA), (L.off_lcg2 - L.off_lXA) / sizeof(float));
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, st, lcg2);

  for (int s = 2; s >= 0; --s) {
    const double dt_s = (s == 0) ? dtd / 3.0 : (s == 1 ? dtd / 2.0 : dtd);
    float* const* src = (s == 0) ? X0 : (s == 1 ? (float* const*)XA : (float* const*)XB);
    float* const* lsrc = (s == 0) ? lam_in : (s == 1 ? (float* const*)lXA : (float* const*)lXB);
    float* const* ldst = (s == 2) ? lam_out : (s == 1 ? (float* const*)lXB : (float* const*)lXA);
    const float* tr = (const float*)(ws + L.off_tr[s]);
    const float* ru = (const float*)(ws + L.off_ru[s]);
    const float* rv = (const float*)(ws + L.off_rv[s]);
    const float* rw = (const float*)(ws + L.off_rw[s]);
    launch_fold_many(d, nf, fmodes, ldst, st);
    cx.zero(lru, (L.off_lsu - L.off_lru) / sizeof(float));  // lru, lrv, lrw, ltr are contiguous
    // scalars: next = pad(s0 + dt_s adv(s; u, v, w) + dt_s sgs)
    for (int q = 0; q < nsc; ++q) {
      int fi = (q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2));
      float weps = (q == 0) ? 1.0e-17f : 1.0e-18f;
      hipLaunchKernelGGL(k_scalar_update_vjp, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, (const float*)ldst[fi],
                         (float)dt_s, lam_in[fi], lT, turb ? lss[q] : nullptr);
      hipLaunchKernelGGL(k_adv_scalar_vjp, cx.G(d.nz, d.ny + 1, d.nx + 1), cx.bk, 0, st, d, rc, (const float*)src[fi],
                         (const float*)src[1], (const float*)src[2], (const float*)src[3], (const float*)lT, weps,
                         spS, lsrc[fi], lsrc[1], lsrc[2], lsrc[3]);
      hipLaunchKernelGGL(k_gather_scalar, cx.G(d.nz + 2, d.ny + 6, d.nx + 6), cx.bk, 0, st, d, spS, lsrc[fi]);
    }
    // acoustic substeps, last to first
    const int nsub = L.nsub[s];
    const float dts = (s == 0) ? (float)dt_s : (float)dts_small;
    float* lo[4] = {ldst[1], ldst[2], ldst[3], ldst[4]};
    for (int t = nsub - 1; t >= 0; --t) {
      float* in[4];
      float* li[4];
      const int acc = (t == 0);   // the step input's adjoint accumulates; ping-pong buffers are overwritten
      if (t == 0) {
        for (int c = 0; c < 4; ++c) { in[c] = X0[1 + c]; li[c] = lam_in[1 + c]; }
      } else {
        acou_ptrs(ws + ac_off(L, s, t), RL, in);
        for (int c = 0; c < 4; ++c) li[c] = lP[t & 1][c];
        cx.zero(li[3], L.nc);
      }
      hipLaunchKernelGGL(k_acoustic_vjp_faces3, cx.G(d.nz, d.ny + 1, d.nx + 1), cx.bk, 0, st, d, rc, b.pi0,
                         (const float*)src[4], (const float*)lo[3], (const float*)lo[0], (const float*)lo[1],
                         (const float*)lo[2], dts, acc, li[0], li[1], li[2], lru, lrv, lrw, gU, gV, gW);
      hipLaunchKernelGGL(k_acoustic_vjp_cells3, cx.G(d.nz, d.ny + 2, d.nx + 2), cx.bk, 0, st, d, rc,
                         (const float*)in[0], (const float*)in[1], (const float*)in[2], (const float*)in[3], ru, rv, rw,
                         tr, (const float*)lo[3], (const float*)gU, (const float*)gV, (const float*)gW, dts, li[3], ltr,
                         lsrc[4]);
      // the velocity adjoints are written at interior points only (their ghosts are never
      // read); the pi' adjoint also has ghost contributions from the PGF stencil
      if (t > 0) cx.fold(0, li[3]);
      for (int c = 0; c < 4; ++c) lo[c] = li[c];
    }
    // momentum rhs = advection (+ b8w for w) + frozen SGS
    if (turb) {
      hipLaunchKernelGGL(k_add_to, dim3((unsigned)((L.iu_ + 255) / 256)), dim3(256), 0, st, (long)L.iu_,
                         (const float*)lru, lsu);
      hipLaunchKernelGGL(k_add_to, dim3((unsigned)((L.iv_ + 255) / 256)), dim3(256), 0, st, (long)L.iv_,
                         (const float*)lrv, lsv);
      hipLaunchKernelGGL(k_add_to, dim3((unsigned)((L.iw_ + 255) / 256)), dim3(256), 0, st, (long)L.iw_,
                         (const float*)lrw, lsw);
    }
    hipLaunchKernelGGL(k_adv_u_vjp, cx.G(d.nz, d.ny + 1, d.nx + 2), cx.bk, 0, st, d, rc, (const float*)src[1],
                       (const float*)src[2], (const float*)src[3], (const float*)lru, spU, lsrc[1]);
    hipLaunchKernelGGL(k_adv_v_vjp, cx.G(d.nz, d.ny + 2, d.nx + 1), cx.bk, 0, st, d, rc, (const float*)src[1],
                       (const float*)src[2], (const float*)src[3], (const float*)lrv, spV, lsrc[2]);
    hipLaunchKernelGGL(k_adv_w_vjp, cx.G(d.nz + 1, d.ny + 1, d.nx + 1), cx.bk, 0, st, d, rc, (const float*)src[1],
                       (const float*)src[2], (const float*)src[3], (const float*)lrw, spW);
    hipLaunchKernelGGL(k_gather_lu, cx.G(d.nz + 2, d.ny + 6, d.nx + 7), cx.bk, 0, st, d, spU, spV, spW, lsrc[1]);
    hipLaunchKernelGGL(k_gather_lv, cx.G(d.nz + 2, d.ny + 7, d.nx + 6), cx.bk, 0, st, d, spU, spV, spW, lsrc[2]);
    hipLaunchKernelGGL(k_gather_lw, cx.G(d.nz + 3, d.ny + 6, d.nx + 6), cx.bk, 0, st, d, spU, spV, spW, lsrc[3]);
    hipLaunchKernelGGL(k_b8w_vjp, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, b.theta0, b.qv0,
                       (const float*)src[0], (const float*)src[5], (const float*)lrw, lsrc[0], lsrc[5]);
    long n = (long)L.nc;
    hipLaunchKernelGGL(k_theta_rho_vjp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                       (const float*)src[0], (const float*)src[5], (const float*)ltr, lsrc[0], lsrc[5]);
  }

  // turb 2: the adjoints of the caller's SGS tendencies (lsu, lsv, lsw, lss in the adjoint
  // workspace) are the result; turb 1 takes them through the Smagorinsky reverse
  if (o.turb == 1) {
    // lsm was cleared with the scratch region and is not written before this point
    for (int q = 0; q < nsc; ++q) {
      int fi = (q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2));
      hipLaunchKernelGGL(k_smag_scalar_vjp, cx.G(d.nz + 2, d.ny + 6, d.nx + 6), cx.bk, 0, st, d, rc, (const float*)sm[9],
                         (const float*)X0[fi], (const float*)lss[q], lsm[9], lam_in[fi]);
    }
    hipLaunchKernelGGL(k_smag_uvw_vjp, cx.G(d.nz + 2, d.ny + 6, d.nx + 6), cx.bk, 0, st, d, rc, (const float*)sm[8],
                       (const float*)sm[0], (const float*)sm[1], (const float*)sm[2], (const float*)sm[3],
                       (const float*)sm[4], (const float*)sm[5], (const float*)lsu, (const float*)lsv,
                       (const float*)lsw, lsm[8], lsm[0], lsm[1], lsm[2], lsm[3], lsm[4], lsm[5]);
    float gs = cbrtf(d.dx * d.dy * d.dz);
    float cg = o.cs * gs;
    long n = (long)L.nc;
    const unsigned nbk = smag_k_blocks(n);
    double* kpart = (double*)(A(L.off_lcg2 + 256));
    hipLaunchKernelGGL(k_smag_k_vjp, dim3(nbk), dim3(256), 0, st, n, (const float*)sm[6], (const float*)sm[7],
                       (const float*)lsm[8], (const float*)lsm[9], cg * cg, o.prandtl, lsm[6], lsm[7], kpart);
    hipLaunchKernelGGL(k_smag_k_sum, dim3(1), dim3(256), 0, st, (const double*)kpart, (int)nbk, lcg2);
    {
      const int zm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      launch_fold_many(d, 8, zm, lsm, st);
    }
    hipLaunchKernelGGL(k_smag_nm_vjp, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, (const float*)X0[0],
                       (const float*)X0[5], (const float*)lsm[7], lam_in[0], lam_in[5]);
    hipLaunchKernelGGL(k_smag_deform_h, cx.G(d.nz, d.ny, d.nx), cx.bk, 0, st, d, rc, (const float*)sm[0],
                       (const float*)sm[1], (const float*)sm[2], (const float*)sm[3], (const float*)sm[4],
                       (const float*)sm[5], (const float*)lsm[6], lsm[0], lsm[1], lsm[2], lsm[3], lsm[4], lsm[5]);
    hipLaunchKernelGGL(k_smag_deform_vjp, cx.G(d.nz + 3, d.ny + 7, d.nx + 7), cx.bk, 0, st, d, (const float*)lsm[0],
                       (const float*)lsm[1], (const float*)lsm[2], (const float*)lsm[3], (const float*)lsm[4],
                       (const float*)lsm[5], lam_in[1], lam_in[2], lam_in[3]);
    // d(cg^2)/dCs = 2 Cs gs^2
    if (lam_cs) {
      double factor = 2.0 * (double)o.cs * (double)gs * (double)gs;
      hipLaunchKernelGGL(k_cs_grad, dim3(1), dim3(64), 0, st, (const double*)lcg2, factor, lam_cs);
    }
  }
  launch_fold_many(d, nf, fmodes, lam_in, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return (int)e;
  return 0;
}

}  // namespace lex
