// R1/R2: the WS-RK3 step for the compressible core (solver_opt 2) on the device.
// Stage structure follows rk_sub_step0/1/2 (one_step_integration.py:16-447) and
// ws_rk3_big_step (one_sprint_functions.py:10-30):
//   stage s (dt_s = dt/3, dt/2, dt) restarts from xi0;
//   Smagorinsky tendencies are computed in stage 0 and frozen (:57-76, :245);
//   momentum rhs = advection + buoyancy(w) + SGS; acoustic substeps 1, round(n/2), n;
//   scalars: s_next = pad(s0 + adv*dt_s) + pad(sgs*dt_s).
#pragma once
#include "lex_common.hpp"
#include "lex_prof.hpp"
#include "lex_fused.hip"
#include "lex_acoustic_vi.hip"
#include "lex_surface.hip"

namespace lex {

// momentum right-hand sides rhs_u = adv4u + fv + sgs_u, rhs_v = adv4v - fu + sgs_v,
// rhs_w = adv4w + b8w + sgs_w (one_step_integration.py:126-133); fu, fv from the stage's
// u, v (cor_opt 1 or 2, off with cor.mode 0), sgs may be NULL.
// tau_x, tau_y: surface stresses (nx, ny) averaged onto the bottom u / v faces, or null.
__global__ void k_rhs_uv(Dims d, const float* __restrict__ rc, const float* __restrict__ u,
                         const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ fd,
                         const float* __restrict__ su, const float* __restrict__ sv,
                         const float* __restrict__ tau_x, const float* __restrict__ tau_y, CorOpt cor,
                         float* __restrict__ ru, float* __restrict__ rv) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = blockIdx.z;                             // 0..nx
  if (k >= d.nz || j > d.ny) return;
  const float weps = 1.0e-16f;  // one_step_integration.py:896
  int I = i + 3, J = j + 3, K = k + 1;
  if (j < d.ny) {
    long o = ((long)i * d.ny + j) * d.nz + k;
    float tf = 0.0f;
    if (tau_x) tf = 0.5f * (tau_x[(long)((i - 1 + d.nx) % d.nx) * d.ny + j] + tau_x[(long)(i % d.nx) * d.ny + j]);
    float a = adv_u_point(d, rc, u, v, w, fd[ic(d, I - 1, J, K)], fd[ic(d, I, J, K)], tf, i, j, k, weps);
    if (cor.mode) a = a + cor_fv(d, v, I, J, K, cor);
    if (su) a = a + su[o];
    ru[o] = a;
  }
  if (i < d.nx) {
    long o = ((long)i * (d.ny + 1) + j) * d.nz + k;
    float tf = 0.0f;
    if (tau_y) tf = 0.5f * (tau_y[(long)i * d.ny + (j - 1 + d.ny) % d.ny] + tau_y[(long)i * d.ny + j % d.ny]);
    float a = adv_v_point(d, rc, u, v, w, fd[ic(d, I, J - 1, K)], fd[ic(d, I, J, K)], tf, i, j, k, weps);
    if (cor.mode) a = a - cor_fu(d, u, I, J, K, cor);
    if (sv) a = a + sv[o];
    rv[o] = a;
  }
}

__global__ void k_rhs_w(Dims d, const float* __restrict__ rc, const float* __restrict__ th0c,
                        const float* __restrict__ qv0c, const float* __restrict__ u, const float* __restrict__ v,
                        const float* __restrict__ w, const float* __restrict__ th, const float* __restrict__ qv,
                        const float* __restrict__ fd, const float* __restrict__ sw, float* __restrict__ rw) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (l > d.nz || j >= d.ny) return;
  const float weps = 1.0e-16f;
  int I = i + 3, J = j + 3, K = l + 1;
  float a = 0.0f;
  if (l > 0 && l < d.nz) a = adv_w_point(d, rc, u, v, w, fd[ic(d, I, J, K - 1)], fd[ic(d, I, J, K)], i, j, l, weps);
  a = a + b8w_point(d, th0c, qv0c, th, qv, I, J, l);
  long o = ((long)i * d.ny + j) * (d.nz + 1) + l;
  if (sw) a = a + sw[o];
  rw[o] = a;
}

__global__ void k_copy(long n, const float* __restrict__ a, float* __restrict__ b) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

// Newtonian heating (newton_radiation, radiation.py:7-16): relaxation of theta to the
// base-state theta0 over 12 h, the tendency capped at +-1 K / 12 h beyond +-1 K.
__global__ void k_newton_heating(Dims d, const float* __restrict__ th, const float* __restrict__ th0c,
                                 float* __restrict__ heat) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const float tau = (float)(12.0 * 3600.0);
  const float dth = th[ic(d, i + 3, j + 3, k + 1)] - th0c[k + 1];
  float t = -dth / tau;
  t = (dth > 1.0f) ? -1.0f / tau : t;
  t = (dth < -1.0f) ? 1.0f / tau : t;
  heat[((long)i * d.ny + j) * d.nz + k] = t;
}

// Horizontal mean of theta per level (rayleigh_damping's theta_avg,
// boundary_conditions.py:136), deterministic in FP64: block (i, c) sums rows
// c*cy .. (c+1)*cy-1 of x-plane i for every level (lanes along z, kMeanRows rows of
// threads, combined in a fixed order); then one wave per level adds the nx*kMeanChunks
// partials (lane l takes l, l+64, ... in order, then a fixed butterfly).
constexpr int kMeanRows = 4;
constexpr int kMeanChunks = 4;
__global__ void __launch_bounds__(1024) k_level_partial(Dims d, const float* __restrict__ th,
                                                        double* __restrict__ part) {
  __shared__ double red[kMeanRows][256];
  const int k = threadIdx.x, r = threadIdx.y;
  const int i = blockIdx.x, c = blockIdx.y;
  const int cy = (d.ny + kMeanChunks - 1) / kMeanChunks;
  const int jend = min((c + 1) * cy, d.ny);
  const int kc = k < d.nz ? k : d.nz - 1;
  double s0 = 0.0, s1 = 0.0;
  int j = c * cy + r;
  for (; j + kMeanRows < jend; j += 2 * kMeanRows) {
    s0 += (double)th[ic(d, i + 3, j + 3, kc + 1)];
    s1 += (double)th[ic(d, i + 3, j + kMeanRows + 3, kc + 1)];
  }
  if (j < jend) s0 += (double)th[ic(d, i + 3, j + 3, kc + 1)];
  red[r][k] = s0 + s1;
  __syncthreads();
  if (r == 0 && k < d.nz) {
    double sum = red[0][k];
    for (int q = 1; q < kMeanRows; ++q) sum += red[q][k];
    part[((long)i * kMeanChunks + c) * d.nz + k] = sum;
  }
}

// Ramp of the damping layer at height z (setup_damping_tau, setup_lex.py:149-156): zero
// up to z_damping, 0.5 (1 - cos(pi (z - z_damping) / depth)) above, depth measured to the
// top w level of the padded grid, (nz + 1) dz.
__device__ __forceinline__ float damp_tau(double z, double zd, double depth) {
  return (z <= zd) ? 0.0f : (float)(0.5 * (1.0 - cos(M_PI * (z - zd) / depth)));
}

// Per level k: the theta mean, the ramp at the cell centre (tauh) and at w level k (tauf;
// the last wave also writes w level nz).  Blocks of 4 waves, one level per wave.
__global__ void __launch_bounds__(256) k_level_mean(Dims d, double zd, const double* __restrict__ part,
                                                    float* __restrict__ mean, float* __restrict__ tauh,
                                                    float* __restrict__ tauf) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= d.nz) return;   // whole waves
  const int np = d.nx * kMeanChunks;
  double s = 0.0;
  for (int p = lane; p < np; p += 64) s += part[(long)p * d.nz + k];
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
  if (lane == 0) {
    const double dz = (double)d.dz, depth = (d.nz + 1) * dz - zd;
    mean[k] = (float)(s / ((double)d.nx * (double)d.ny));
    tauh[k] = damp_tau((k + 0.5) * dz, zd, depth);
    tauf[k] = damp_tau((double)k * dz, zd, depth);
    if (k == d.nz - 1) tauf[d.nz] = damp_tau((double)d.nz * dz, zd, depth);
  }
}

// Upper-level Rayleigh damping (rayleigh_damping, boundary_conditions.py:119-141) applied
// to the stage result in place (one_step_integration.py:164-175): u, v, w relax to zero and
// theta to its horizontal mean with rate alpha x ramp, ghosts refilled.
// Levels below k0 have a zero ramp, where the update leaves every value unchanged, so the
// launch covers levels k0..nz-1 only: thread t of x-plane blockIdx.y takes level
// k0 + t % nl of row t / nl (nl = nz - k0), so a wave runs along whole layer columns.
__global__ void __launch_bounds__(256) k_rayleigh(Dims d, int k0, float alpha, float dt,
                                                  const float* __restrict__ mean, const float* __restrict__ tauh,
                                                  const float* __restrict__ tauf, float* __restrict__ th,
                                                  float* __restrict__ u, float* __restrict__ v,
                                                  float* __restrict__ w) {
  const int nl = d.nz - k0;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (t >= nl * d.ny) return;
  const int j = t / nl, k = k0 + t % nl;
  const int I = i + 3, J = j + 3, K = k + 1;
  const float th_k = tauh[k];
  const float tf_k = tauf[k];
  const float hu = 0.5f * (th_k + th_k);
  const float un = u[iu(d, I, J, K)], vn = v[iv(d, I, J, K)], wn = w[iw(d, I, J, K)], tn = th[ic(d, I, J, K)];
  store_u_fast(d, u, i, j, k, un + (-alpha * (hu * un)) * dt);
  if (i == d.nx - 1) {
    const float ue = u[iu(d, d.nx + 3, J, K)];
    store_u(d, u, d.nx, j, k, ue + (-alpha * (hu * ue)) * dt);
  }
  store_v_fast(d, v, i, j, k, vn + (-alpha * (hu * vn)) * dt);
  if (j == d.ny - 1) {
    const float ve = v[iv(d, I, d.ny + 3, K)];
    store_v(d, v, i, d.ny, k, ve + (-alpha * (hu * ve)) * dt);
  }
  store_w_fast(d, w, i, j, k, wn + (-alpha * (tf_k * wn)) * dt);
  if (k == d.nz - 1) {
    const float wt = w[iw(d, I, J, d.nz + 1)];
    store_w(d, w, i, j, d.nz, wt + (-alpha * (tauf[d.nz] * wt)) * dt);
  }
  store_c_fast(d, th, i, j, k, tn + (-alpha * th_k * (tn - mean[k])) * dt);
}

struct Rk3Layout {
  size_t nc, nu, nv, nw;        // padded sizes (elements)
  size_t iu_, iv_, iw_, ic_;    // interior sizes (u, v, w points, scalar)
  size_t off_stA, off_stB;      // two stage states (6 + np padded fields each)
  size_t off_P[2];              // acoustic ping-pong (u, v, w, pip)
  size_t off_fd, off_tr;        // flow divergence, theta_rho
  size_t off_ru, off_rv, off_rw;
  size_t off_su, off_sv, off_sw, off_ss;   // frozen SGS tendencies (ss: 2+np scalars)
  size_t off_smag;              // 10 padded scratch fields
  size_t off_heat;              // Newtonian heating of theta (interior), rad
  size_t off_lmean;             // damp: per-level partial sums (nx x kMeanChunks x nz doubles), then
                                // the theta mean (nz floats) and the ramps tauh (nz), tauf (nz+1)
  size_t off_sfc;               // surface fluxes: 8 (nx, ny) fields (SfcOut order), sfc
  size_t total;
  int np;
};

inline size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

inline Rk3Layout rk3_layout(const Dims& d, int np) {
  Rk3Layout L;
  L.np = np;
  L.nc = d.nc();
  L.nu = d.nu();
  L.nv = d.nv();
  L.nw = d.nw();
  L.iu_ = (size_t)(d.nx + 1) * d.ny * d.nz;
  L.iv_ = (size_t)d.nx * (d.ny + 1) * d.nz;
  L.iw_ = (size_t)d.nx * d.ny * (d.nz + 1);
  L.ic_ = (size_t)d.nx * d.ny * d.nz;
  size_t state = al256(L.nu * 4) + al256(L.nv * 4) + al256(L.nw * 4) + (3 + np) * al256(L.nc * 4);
  size_t acou = al256(L.nu * 4) + al256(L.nv * 4) + al256(L.nw * 4) + al256(L.nc * 4);
  size_t o = 0;
  L.off_stA = o; o += state;
  L.off_stB = o; o += state;
  L.off_P[0] = o; o += acou;
  L.off_P[1] = o; o += acou;
  L.off_fd = o; o += al256(L.nc * 4);
  L.off_tr = o; o += al256(L.nc * 4);
  L.off_ru = o; o += al256(L.iu_ * 4);
  L.off_rv = o; o += al256(L.iv_ * 4);
  L.off_rw = o; o += al256(L.iw_ * 4);
  L.off_su = o; o += al256(L.iu_ * 4);
  L.off_sv = o; o += al256(L.iv_ * 4);
  L.off_sw = o; o += al256(L.iw_ * 4);
  L.off_ss = o; o += (2 + np) * al256(L.ic_ * 4);
  L.off_smag = o; o += 10 * al256(L.nc * 4);
  L.off_heat = o; o += al256(L.ic_ * 4);
  L.off_lmean = o; o += al256((size_t)d.nx * kMeanChunks * d.nz * 8) + 3 * al256((size_t)(d.nz + 1) * 4);
  L.off_sfc = o; o += 8 * al256((size_t)d.nx * d.ny * 4);
  L.total = o;
  return L;
}

// pointers of a state block laid out as theta, u, v, w, pip, qv, passive...
inline void state_ptrs(char* base, const Rk3Layout& L, float** f) {
  size_t o = 0;
  f[1] = (float*)(base + o); o += al256(L.nu * 4);
  f[2] = (float*)(base + o); o += al256(L.nv * 4);
  f[3] = (float*)(base + o); o += al256(L.nw * 4);
  f[0] = (float*)(base + o); o += al256(L.nc * 4);
  f[4] = (float*)(base + o); o += al256(L.nc * 4);
  f[5] = (float*)(base + o); o += al256(L.nc * 4);
  for (int q = 0; q < L.np; ++q) {
    f[6 + q] = (float*)(base + o);
    o += al256(L.nc * 4);
  }
}

inline void acou_ptrs(char* base, const Rk3Layout& L, float** f) {  // u, v, w, pip
  size_t o = 0;
  f[0] = (float*)(base + o); o += al256(L.nu * 4);
  f[1] = (float*)(base + o); o += al256(L.nv * 4);
  f[2] = (float*)(base + o); o += al256(L.nw * 4);
  f[3] = (float*)(base + o);
}

// turb code of a step whose SGS tendencies the caller has placed in the workspace (or
// tape) slots, frozen over the step: the learned SGS net
constexpr int kTurbExternal = 9;

// Host function called at the start of every stage with the stage's input state (6 +
// n_passive padded fields in state order) and the SGS slots of the workspace; it fills
// the slots for that stage by work enqueued on the step's stream (the learned SGS net
// run inside each stage).  Returns 0 or an error code.
typedef int (*StageSgsFn)(int stage, float* const* stage_state, float* su, float* sv, float* sw, float* const* ss,
                          void* user);

struct StepOpts {
  float dt;
  int n_sound;
  int turb;
  float cs;
  float prandtl;
  int n_passive;
  int damp = 0;
  int rad = 0;
  float z_damping = 9600.0f;
  float rd_alpha = 1.0f / 120.0f;
  int sfc = 0;
};

// the surface-flux fields of the workspace
inline SfcOut sfc_fields(char* ws, const Rk3Layout& L, const Dims& d) {
  float* f[8];
  for (int q = 0; q < 8; ++q) f[q] = (float*)(ws + L.off_sfc + q * al256((size_t)d.nx * d.ny * 4));
  return SfcOut{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]};
}

template <class O, class B>
int rk3_step(const Dims& d, const O& o, const B& b, float* const* state, char* ws, const Rk3Layout& L,
             hipStream_t st, StageSgsFn stage_sgs = nullptr, void* stage_user = nullptr) {
  const int nf = 6 + o.n_passive;
  const CorOpt cor{o.cor, o.fcor, o.u_geo, o.v_geo};
  float* X0[16];
  float* XA[16];
  float* XB[16];
  for (int q = 0; q < nf; ++q) X0[q] = state[q];
  state_ptrs(ws + L.off_stA, L, XA);
  state_ptrs(ws + L.off_stB, L, XB);
  float* P[2][4];
  acou_ptrs(ws + L.off_P[0], L, P[0]);
  acou_ptrs(ws + L.off_P[1], L, P[1]);
  float* fd = (float*)(ws + L.off_fd);
  float* tr = (float*)(ws + L.off_tr);
  float* ru = (float*)(ws + L.off_ru);
  float* rv = (float*)(ws + L.off_rv);
  float* rw = (float*)(ws + L.off_rw);
  float* su = (float*)(ws + L.off_su);
  float* sv = (float*)(ws + L.off_sv);
  float* sw = (float*)(ws + L.off_sw);
  float* ss[10];
  for (int q = 0; q < 2 + o.n_passive; ++q) ss[q] = (float*)(ws + L.off_ss + q * al256(L.ic_ * 4));
  float* sm[10];
  for (int q = 0; q < 10; ++q) sm[q] = (float*)(ws + L.off_smag + q * al256(L.nc * 4));
  float* heat = o.rad ? (float*)(ws + L.off_heat) : nullptr;
  double* lpart = (double*)(ws + L.off_lmean);
  float* lmean = (float*)(ws + L.off_lmean + al256((size_t)d.nx * kMeanChunks * d.nz * 8));
  float* ltauh = lmean + al256((size_t)(d.nz + 1) * 4) / 4;
  float* ltauf = ltauh + al256((size_t)(d.nz + 1) * 4) / 4;
  const SfcOut sf = sfc_fields(ws, L, d);
  const float* tau_x = o.sfc ? sf.tau_x : nullptr;
  const float* tau_y = o.sfc ? sf.tau_y : nullptr;
  const float* sen_cp = o.sfc ? sf.sen_cp : nullptr;
  const float* evap = o.sfc ? sf.evap : nullptr;
  const dim3 bk(64, 4, 1);
  auto G = [&](int a, int bb, int c) { return dim3((a + 63) / 64, (bb + 3) / 4, c); };
  const float* rc = b.rho0;
  const double dtd = (double)o.dt;
  const double dts_small = dtd / (double)o.n_sound;

  for (int s = 0; s < 3; ++s) {
    const double dt_s = (s == 0) ? dtd / 3.0 : (s == 1 ? dtd / 2.0 : dtd);
    float** src = (s == 0) ? X0 : (s == 1 ? XA : XB);
    float** dst = (s == 0) ? XA : (s == 1 ? XB : X0);
    // diabatic heating from the stage-0 theta, kept for the step (one_step_integration.py:31-35)
    if (s == 0 && o.rad)
      LEX_LAUNCH(KID_HEATING, k_newton_heating, G(d.nz, d.ny, d.nx), bk, 0, st, d, X0[0], b.theta0, heat);
    // surface fluxes from the stage-0 state, kept for the step (one_step_integration.py:37-47, :209)
    if (s == 0 && o.sfc)
      LEX_LAUNCH(KID_SFC, k_atm_ocn_flux, dim3((d.ny + 63) / 64, d.nx, 1), dim3(64, 1, 1), 0, st, d, rc, X0[0],
                 X0[1], X0[2], X0[5], b.surface_t, sf);
    // turb 1 / 2: Smagorinsky (isotropic / horizontal-vertical) computed here in stage 0;
    // kTurbExternal: SGS tendencies supplied by the caller in the workspace slots (the
    // learned SGS net); all frozen over the step
    const bool turb = o.turb != 0;
    if (stage_sgs) {   // SGS tendencies of this stage from its input state (caller's function)
      const int rc = stage_sgs(s, src, su, sv, sw, ss, stage_user);
      if (rc) return -5;
    }
    if (s == 0 && (o.turb == 1 || o.turb == 2)) {
      long n = (long)L.nc;
      const float gs = cbrtf(d.dx * d.dy * d.dz);
      const float cg = o.cs * gs;
      const float* khh;
      const float* khv;
      if (o.impl != 1 && smag_fused_ok(d)) {
        // the fused paths: the whole evaluation in one pass, stresses on chip
        SgsScalars SS;
        SS.n = 2 + o.n_passive;
        for (int q = 0; q < SS.n; ++q) {
          SS.s[q] = src[(q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2))];
          SS.out[q] = ss[q];
        }
        const bool an = (o.turb == 2);
        const dim3 fbk(64 * ((d.nz + 63) / 64), smag_fused_rows(d), 1);
        if (an)
          LEX_LAUNCH(KID_SMAG_ONE, k_smag_fused<true>, smag_fused_grid(d), fbk, smag_fused_lds(d, true), st, d, rc,
                     src[1], src[2], src[3], src[0], src[5], cg * cg, o.prandtl, o.cs * o.cs, SS, su, sv, sw);
        else
          LEX_LAUNCH(KID_SMAG_ONE, k_smag_fused<false>, smag_fused_grid(d), fbk, smag_fused_lds(d, false), st, d, rc,
                     src[1], src[2], src[3], src[0], src[5], cg * cg, o.prandtl, o.cs * o.cs, SS, su, sv, sw);
      } else if (o.impl != 1) {
        // the fused paths when the one-pass kernel does not apply (nz mod 64 of 1 or 2):
        // deformation, N^2 and the coefficients in one pass writing the stresses and
        // diffusivities (k_smag_tau), then the stress divergence
        SmagTau T;
        for (int q = 0; q < 6; ++q) T.t[q] = sm[q];
        T.khh = sm[6];
        T.khv = (o.turb == 2) ? sm[7] : nullptr;
        LEX_LAUNCH(KID_SMAG_FUSED, k_smag_tau, G(d.nz, d.ny, d.nx), bk, 0, st, d, rc, src[1], src[2], src[3], src[0],
                   src[5], cg * cg, o.prandtl, (int)(o.turb == 2), o.cs * o.cs, T);
        LEX_LAUNCH(KID_SMAG_UVW, k_smag_uvw<true>, G(d.nz, d.ny + 1, d.nx + 1), bk, 0, st, d, rc,
                   (const float*)nullptr, (const float*)nullptr, sm[0], sm[1], sm[2], sm[3], sm[4], sm[5], su, sv, sw);
        khh = sm[6];
        khv = (o.turb == 2) ? sm[7] : sm[6];
        SgsScalars SS;
        SS.n = 2 + o.n_passive;
        for (int q = 0; q < SS.n; ++q) {
          SS.s[q] = src[(q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2))];
          SS.out[q] = ss[q];
        }
        LEX_LAUNCH(KID_SMAG_SCALAR, k_smag_scalars, G(d.nz, d.ny, d.nx), bk, 0, st, d, rc, khh, khv, SS);
      } else {
        // one kernel per reference routine: compute_deformation -> compute_nm -> compute_k
        // (compute_k2) -> compute_smag (compute_smag2), one call per scalar
        LEX_LAUNCH(KID_SMAG_DEFORM, k_smag_deform, G(d.nz, d.ny, d.nx), bk, 0, st, d, rc, src[1], src[2], src[3], sm[0],
                   sm[1], sm[2], sm[3], sm[4], sm[5], sm[6]);
        LEX_LAUNCH(KID_SMAG_NM, k_smag_nm, G(d.nz, d.ny, d.nx), bk, 0, st, d, src[0], src[5], sm[7]);
        if (o.turb == 1) {
          LEX_LAUNCH(KID_SMAG_K, k_smag_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, sm[6], sm[7], sm[8],
                     sm[9], cg * cg, o.prandtl);
        } else {  // sm[8] = Kmh = Khh, sm[9] = Kmv = Khv
          LEX_LAUNCH(KID_SMAG_K, k_smag_k2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, rc, sm[0], sm[1],
                     sm[2], sm[3], sm[4], sm[5], sm[7], sm[8], sm[9], o.cs * o.cs);
        }
        // coefficients: turb 1 Km (sm[8]) for every stress, Kh (sm[9]) for every scalar
        // face; turb 2 Kmh / Kmv (sm[8] / sm[9]) for both
        const float* kmh = sm[8];
        const float* kmv = (o.turb == 1) ? sm[8] : sm[9];
        khh = (o.turb == 1) ? sm[9] : sm[8];
        khv = sm[9];
        LEX_LAUNCH(KID_SMAG_UVW, k_smag_uvw<false>, G(d.nz, d.ny + 1, d.nx + 1), bk, 0, st, d, rc, kmh, kmv, sm[0],
                   sm[1], sm[2], sm[3], sm[4], sm[5], su, sv, sw);
        LEX_LAUNCH(KID_SMAG_SCALAR, k_smag_scalar, G(d.nz, d.ny, d.nx), bk, 0, st, d, rc, khh, khv, src[0], ss[0]);
        LEX_LAUNCH(KID_SMAG_SCALAR, k_smag_scalar, G(d.nz, d.ny, d.nx), bk, 0, st, d, rc, khh, khv, src[5], ss[1]);
        for (int q = 0; q < o.n_passive; ++q)
          LEX_LAUNCH(KID_SMAG_SCALAR, k_smag_scalar, G(d.nz, d.ny, d.nx), bk, 0, st, d, rc, khh, khv, src[6 + q],
                     ss[2 + q]);
      }
    }
    const int nsc = 2 + o.n_passive;
    const bool fused_ok = (64 * ((d.nz + 63) / 64)) * kFTY <= 1024 && nsc >= 2 && nsc <= 6;
    // impl 0: fused stage kernels; impl 2: the same with the double-buffered scalar kernel;
    // impl 1: one kernel per reference routine
    const bool fused = (o.impl == 0 || o.impl == 2);
    if (fused && !fused_ok) return -4;  // the caller falls back to impl 1
    if (o.impl == 2 && (stage_scalars_db_lds(d, nsc) > 65536 || 64 * ((d.nz + 63) / 64) * kFTY > 512)) return -4;
    if (fused) {
      const int nzl = 64 * ((d.nz + 63) / 64);
      dim3 fb(nzl, kFTY, 1), fg((d.nx + kFTX - 1) / kFTX, (d.ny + kFTY - 1) / kFTY, 1);
      LEX_LAUNCH(KID_STAGE_MOMENTUM, k_mom_u, fg, fb, mom_uv_lds(d), st, d, rc, src[1], src[2], src[3],
                 turb ? (const float*)su : nullptr, tau_x, cor, ru);
      LEX_LAUNCH(KID_MOM_V, k_mom_v, fg, fb, mom_uv_lds(d), st, d, rc, src[1], src[2], src[3],
                 turb ? (const float*)sv : nullptr, tau_y, cor, rv);
      LEX_LAUNCH(KID_MOM_W, k_mom_w, fg, fb, mom_w_lds(d), st, d, rc, b.theta0, b.qv0, src[1], src[2], src[3],
                 src[0], src[5], turb ? (const float*)sw : nullptr, rw);
    } else {
      LEX_LAUNCH(KID_DIV_RHO, k_div_rho, G(d.nz + 2, d.ny + 6, d.nx + 6), bk, 0, st, d, rc, src[1], src[2], src[3], fd);
      LEX_LAUNCH(KID_RHS_UV, k_rhs_uv, G(d.nz, d.ny + 1, d.nx + 1), bk, 0, st, d, rc, src[1], src[2], src[3], fd,
                 turb ? su : nullptr, turb ? sv : nullptr, tau_x, tau_y, cor, ru, rv);
      LEX_LAUNCH(KID_RHS_W, k_rhs_w, G(d.nz + 1, d.ny, d.nx), bk, 0, st, d, rc, b.theta0, b.qv0, src[1], src[2], src[3],
                 src[0], src[5], fd, turb ? sw : nullptr, rw);
    }
    if (fused) {
      // all scalars of the stage + theta_rho in one launch (k_stage_scalars)
      ScalarSet S;
      S.n = nsc;
      for (int q = 0; q < nsc; ++q) {
        int fi = (q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2));
        S.s0[q] = X0[fi];
        S.s[q] = src[fi];
        S.sgs[q] = turb ? ss[q] : nullptr;
        S.out[q] = dst[fi];
        S.weps[q] = (q == 0) ? 1.0e-17f : 1.0e-18f;
      }
      S.heat = heat;
      S.sfc[0] = sen_cp;
      S.sfc[1] = evap;
      const int nzl = 64 * ((d.nz + 63) / 64);
      dim3 fb(nzl, kFTY, 1), fg((d.nx + kFTX - 1) / kFTX, (d.ny + kFTY - 1) / kFTY, 1);
      size_t lds = stage_scalars_lds(d);
      float dts_f = (float)dt_s;
      if (o.impl == 2) {
        size_t ldb = stage_scalars_db_lds(d, nsc);
        switch (nsc) {
          case 2: LEX_LAUNCH(KID_STAGE_SCALARS, k_stage_scalars_db<2>, fg, fb, ldb, st, d, rc, src[1], src[2], src[3], S, tr, dts_f); break;
          case 3: LEX_LAUNCH(KID_STAGE_SCALARS, k_stage_scalars_db<3>, fg, fb, ldb, st, d, rc, src[1], src[2], src[3], S, tr, dts_f); break;
          case 4: LEX_LAUNCH(KID_STAGE_SCALARS, k_stage_scalars_db<4>, fg, fb, ldb, st, d, rc, src[1], src[2], src[3], S, tr, dts_f); break;
          default: return -4;
        }
      } else {
        dim3 g1(nsc, (d.nx + kFTX - 1) / kFTX, (d.ny + kFTY - 1) / kFTY);   // scalar fastest
        LEX_LAUNCH(KID_STAGE_SCALARS, k_stage_scalars, g1, fb, lds, st, d, rc, src[1], src[2], src[3], S, tr, dts_f);
      }
    } else {
      long n = (long)L.nc;
      LEX_LAUNCH(KID_THETA_RHO, k_theta_rho, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, src[0], src[5], tr);
    }
    // acoustic substeps
    int nsub;
    float dts;
    if (s == 0) {
      nsub = 1;
      dts = (float)dt_s;
    } else {
      nsub = (int)nearbyint(dt_s / dts_small);
      dts = (float)dts_small;
    }
    float* in[4] = {X0[1], X0[2], X0[3], X0[4]};
    // the fused path advances two substeps per launch while two remain
    const bool pairs = (o.solver != 3) && (o.impl == 0) && acoustic2_ok(d);
    for (int t = 0; t < nsub;) {
      const int tn = t + ((pairs && t + 1 < nsub) ? 2 : 1);
      float* out[4];
      bool last = (tn == nsub);
      bool alias = last && (dst == X0) && (in[0] == X0[1]);
      if (last && !alias) {
        out[0] = dst[1]; out[1] = dst[2]; out[2] = dst[3]; out[3] = dst[4];
      } else {
        int pp = (in[0] == P[0][0]) ? 1 : 0;
        for (int c = 0; c < 4; ++c) out[c] = P[pp][c];
      }
      if (o.solver == 3) {
        LEX_LAUNCH(KID_ACOUSTIC_VI, k_acoustic_vi, dim3(d.ny, d.nx, 1), dim3(acoustic_vi_lanes(d)), acoustic_vi_lds(d),
                   st, d, rc, b.pi0, in[0], in[1], in[2], in[3], src[4], ru, rv, rw, tr, out[0], out[1], out[2], out[3],
                   dts, kAlphaImplicit);
      } else if (o.impl == 2) {
        dim3 ag((d.nx + kAX - 1) / kAX, (d.ny + 3) / 4, (d.nz + 63) / 64);
        LEX_LAUNCH(KID_ACOUSTIC, k_acoustic_m, ag, dim3(64, 4, 1), 0, st, d, rc, b.pi0, in[0], in[1], in[2], in[3],
                   src[4], ru, rv, rw, tr, out[0], out[1], out[2], out[3], dts);
      } else if (tn == t + 2) {
        LEX_LAUNCH(KID_ACOUSTIC2, k_acoustic2, acoustic2_grid(d), dim3(acoustic2_lanes(d), acoustic2_rows(d), 1),
                   acoustic2_lds(d), st, d, rc, b.pi0, in[0], in[1], in[2], in[3], src[4], ru, rv, rw, tr, out[0],
                   out[1], out[2], out[3], dts, acoustic2_ax(d));
      } else {
        LEX_LAUNCH(KID_ACOUSTIC, k_acoustic, G(d.nz, d.ny, d.nx), bk, 0, st, d, rc, b.pi0, in[0], in[1], in[2], in[3],
                   src[4], ru, rv, rw, tr, out[0], out[1], out[2], out[3], dts);
      }
      if (alias) {
        const size_t ns[4] = {L.nu, L.nv, L.nw, L.nc};
        for (int c = 0; c < 4; ++c)
          LEX_LAUNCH(KID_COPY, k_copy, dim3((unsigned)((ns[c] + 255) / 256)), dim3(256), 0, st, (long)ns[c], out[c],
                             dst[1 + c]);
      }
      for (int c = 0; c < 4; ++c) in[c] = out[c];
      t = tn;
    }
    // scalars: theta (weps 1e-17), qv and passive (1e-18) (one_step_integration.py:905, :913)
    if (!fused) {
      for (int q = 0; q < nsc; ++q) {
        int fi = (q == 0) ? 0 : (q == 1 ? 5 : 6 + (q - 2));
        float weps = (q == 0) ? 1.0e-17f : 1.0e-18f;
        LEX_LAUNCH(KID_SCALAR_UPDATE, k_scalar_update, G(d.nz, d.ny, d.nx), bk, 0, st, d, rc, X0[fi], src[fi], src[1],
                   src[2], src[3], fd, (q == 0) ? sen_cp : (q == 1 ? evap : nullptr),
                   (q == 0) ? (const float*)heat : nullptr,
                   turb ? ss[q] : nullptr, dst[fi], (float*)nullptr, (float)dt_s, weps);
      }
    }
    // upper-level Rayleigh damping of the stage result (one_step_integration.py:164-175)
    if (o.damp) {
      LEX_LAUNCH(KID_DAMPING, k_level_partial, dim3(d.nx, kMeanChunks), dim3(64 * ((d.nz + 63) / 64), kMeanRows), 0,
                 st, d, (const float*)dst[0], lpart);
      LEX_LAUNCH(KID_DAMPING, k_level_mean, dim3((d.nz + 3) / 4), dim3(256), 0, st, d, (double)o.z_damping,
                 (const double*)lpart, lmean, ltauh, ltauf);
      // first level whose cell centre or lower w level can lie above z_damping
      int k0 = (int)floor((double)o.z_damping / (double)d.dz - 0.5);
      k0 = k0 < 0 ? 0 : (k0 > d.nz - 1 ? d.nz - 1 : k0);
      const int nl = d.nz - k0;
      LEX_LAUNCH(KID_DAMPING, k_rayleigh, dim3((unsigned)((nl * d.ny + 255) / 256), d.nx), dim3(256), 0, st, d, k0,
                 o.rd_alpha, (float)dt_s, (const float*)lmean, (const float*)ltauh, (const float*)ltauf, dst[0],
                 dst[1], dst[2], dst[3]);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return (int)e;
  return 0;
}

}  // namespace lex
