// extern "C" entry points of liblexhip.so (declared in include/lex_capi.h).
// Unity build: the kernel sources are included here so one hipcc invocation produces
// the whole library.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <cmath>

#include "lex_common.hpp"
#include "lex_ops.hip"
#include "lex_smag.hip"
#include "lex_pressure.hip"
#include "lex_stage.hip"
#include "../../include/lex_capi.h"

using namespace lex;

namespace {
thread_local std::string g_err;

int fail(int code, const char* fmt, const char* a = "") {
  char buf[512];
  snprintf(buf, sizeof(buf), fmt, a);
  g_err = buf;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return (int)e;
  }
  return 0;
}

bool grid_ok(const lex_grid_t* g) {
  return g && g->nx >= 6 && g->ny >= 6 && g->nz >= 6 && g->dx > 0.f && g->dy > 0.f && g->dz > 0.f;
}

Dims mkdims(const lex_grid_t* g) {
  Dims d;
  d.nx = g->nx;
  d.ny = g->ny;
  d.nz = g->nz;
  d.dx = g->dx;
  d.dy = g->dy;
  d.dz = g->dz;
  d.idx = 1.0f / g->dx;
  d.idy = 1.0f / g->dy;
  d.idz = 1.0f / g->dz;
  return d;
}

inline hipStream_t S(void* s) { return (hipStream_t)s; }

// (z-lanes, rows) block over an (nkz, nj, ni) index space
inline dim3 blk() { return dim3(64, 4, 1); }
inline dim3 grd(int nkz, int nj, int ni) { return dim3((nkz + 63) / 64, (nj + 3) / 4, ni); }

#define REQ(ptr) \
  if (!(ptr)) return fail(-2, "null pointer argument: %s", #ptr)
#define REQGRID(g) \
  if (!grid_ok(g)) return fail(-1, "invalid grid (need nx,ny,nz >= 6 and positive spacings)%s")
}  // namespace

extern "C" {

const char* lex_last_error(void) { return g_err.c_str(); }
int lex_abi_version(void) { return 1; }

// ---------------------------------- H1 -------------------------------------------
int lex_pad(int mode, int ni, int nj, int nk, int pad, const float* in, float* out, void* stream) {
  REQ(in);
  REQ(out);
  if (mode < 0 || mode > 2 || (pad != 1 && pad != 3)) return fail(-1, "lex_pad: bad mode/pad%s");
  if (ni < pad + 1 || nj < pad + 1 || nk < 1) return fail(-1, "lex_pad: extents too small%s");
  int PI = ni + 2 * pad, PJ = nj + 2 * pad, PK = nk + 2;
  hipLaunchKernelGGL(k_pad, grd(PK, PJ, PI), blk(), 0, S(stream), mode, ni, nj, nk, pad, in, out);
  return check_launch("lex_pad");
}

int lex_set_w_bc(const lex_grid_t* g, float* w, void* stream) {
  REQGRID(g);
  REQ(w);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_set_w_bc, dim3((d.ny + 6 + 63) / 64, d.nx + 6), dim3(64), 0, S(stream), d, w);
  return check_launch("lex_set_w_bc");
}

// ---------------------------------- A1 -------------------------------------------
int lex_div_rho(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w, float* out,
                void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_div_rho, grd(d.nz + 2, d.ny + 6, d.nx + 6), blk(), 0, S(stream), d, r, u, v, w, out);
  return check_launch("lex_div_rho");
}

int lex_div(const lex_grid_t* g, const float* u, const float* v, const float* w, float* out, void* stream) {
  REQGRID(g);
  REQ(u); REQ(v); REQ(w); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_div, grd(d.nz + 2, d.ny + 6, d.nx + 6), blk(), 0, S(stream), d, u, v, w, out);
  return check_launch("lex_div");
}

// ------------------------------- A2-A4 -------------------------------------------
int lex_adv_scalar(const lex_grid_t* g, const float* r, const float* s, const float* u, const float* v, const float* w,
                   float weps, const float* fd, const float* sfc, float* out, void* stream) {
  REQGRID(g);
  REQ(r); REQ(s); REQ(u); REQ(v); REQ(w); REQ(fd); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_adv_scalar, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, r, s, u, v, w, fd, sfc, out, weps);
  return check_launch("lex_adv_scalar");
}

int lex_adv_u(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w, float weps,
              const float* fd, const float* sfc, float* out, void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(fd); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_adv_u, grd(d.nz, d.ny, d.nx + 1), blk(), 0, S(stream), d, r, u, v, w, fd, sfc, out, weps);
  return check_launch("lex_adv_u");
}

int lex_adv_v(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w, float weps,
              const float* fd, const float* sfc, float* out, void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(fd); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_adv_v, grd(d.nz, d.ny + 1, d.nx), blk(), 0, S(stream), d, r, u, v, w, fd, sfc, out, weps);
  return check_launch("lex_adv_v");
}

int lex_adv_w(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w, float weps,
              const float* fd, float* out, void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(fd); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_adv_w, grd(d.nz + 1, d.ny, d.nx), blk(), 0, S(stream), d, r, u, v, w, fd, out, weps);
  return check_launch("lex_adv_w");
}

int lex_adv_pi(const lex_grid_t* g, const float* r, const float* pi0, const float* pip, const float* u,
               const float* v, const float* w, float* out, void* stream) {
  REQGRID(g);
  REQ(r); REQ(pi0); REQ(pip); REQ(u); REQ(v); REQ(w); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_adv_pi, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, r, pi0, pip, u, v, w, out);
  return check_launch("lex_adv_pi");
}

// ---------------------------------- B1 -------------------------------------------
int lex_buoyancy(const lex_grid_t* g, const float* th0, const float* qv0, const float* th, const float* qv, float* b,
                 float* b8w, void* stream) {
  REQGRID(g);
  REQ(th0); REQ(qv0); REQ(th); REQ(qv); REQ(b); REQ(b8w);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_buoyancy, grd(d.nz + 1, d.ny, d.nx), blk(), 0, S(stream), d, th0, qv0, th, qv, b, b8w);
  return check_launch("lex_buoyancy");
}

int lex_theta_rho(long n, const float* th, const float* qv, float* out, void* stream) {
  REQ(th); REQ(qv); REQ(out);
  if (n <= 0) return fail(-1, "lex_theta_rho: n <= 0%s");
  hipLaunchKernelGGL(k_theta_rho, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), n, th, qv, out);
  return check_launch("lex_theta_rho");
}

// ------------------------------- P1/P2 -------------------------------------------
int lex_pgf(const lex_grid_t* g, const float* p, const float* tr, float* pu, float* pv, float* pw, void* stream) {
  REQGRID(g);
  REQ(p); REQ(tr); REQ(pu); REQ(pv); REQ(pw);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_pgf, grd(d.nz + 1, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, p, tr, pu, pv, pw);
  return check_launch("lex_pgf");
}

int lex_coriolis(const lex_grid_t* g, const float* u, const float* v, float fcor, float* fu, float* fv, void* stream) {
  REQGRID(g);
  REQ(u); REQ(v); REQ(fu); REQ(fv);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_coriolis, grd(d.nz, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, u, v, fu, fv, fcor);
  return check_launch("lex_coriolis");
}

// ------------------------------- P3/P4 -------------------------------------------
int lex_rtt(const lex_grid_t* g, const float* rtc, const float* th, const float* qv, const float* b, float* rtt,
            void* stream) {
  REQGRID(g);
  REQ(rtc); REQ(th); REQ(qv); REQ(b); REQ(rtt);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_rtt, grd(d.nz + 2, d.ny + 6, d.nx + 6), blk(), 0, S(stream), d, rtc, th, qv, b, rtt);
  return check_launch("lex_rtt");
}

int lex_p_rhs(const lex_grid_t* g, const float* rtc, const float* au, const float* av, const float* aw,
              const float* fu, const float* fv, const float* b, float* rhs, void* stream) {
  REQGRID(g);
  REQ(rtc); REQ(au); REQ(av); REQ(aw); REQ(b); REQ(rhs);
  if ((fu == nullptr) != (fv == nullptr)) return fail(-2, "lex_p_rhs: fu and fv must both be given or both NULL%s");
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_p_rhs, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, rtc, au, av, aw, fu, fv, b, rhs);
  return check_launch("lex_p_rhs");
}

int lex_p_matvec(const lex_grid_t* g, const float* rtt, const float* x, float* out, void* stream) {
  REQGRID(g);
  REQ(rtt); REQ(x); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_p_matvec, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, rtt, x, out);
  return check_launch("lex_p_matvec");
}

size_t lex_bicgstab_workspace_bytes(const lex_grid_t* g) {
  if (!grid_ok(g)) return 0;
  size_t n = (size_t)g->nx * g->ny * g->nz;
  size_t vec = ((n * sizeof(float) + 255) / 256) * 256;
  return 256 + (size_t)kBicgBlocks * 4 * sizeof(double) + 8 * vec;
}

int lex_bicgstab(const lex_grid_t* g, const float* rtt, const float* rhs, const float* x0, float* x, double tol,
                 double atol, int maxiter, void* ws, size_t ws_bytes, int* k_out, void* stream) {
  REQGRID(g);
  REQ(rtt); REQ(rhs); REQ(x0); REQ(x); REQ(ws);
  if (ws_bytes < lex_bicgstab_workspace_bytes(g)) return fail(-3, "lex_bicgstab: workspace too small%s");
  if (maxiter < 0) return fail(-1, "lex_bicgstab: maxiter < 0%s");
  if (x == x0) return fail(-1, "lex_bicgstab: x must not alias x0%s");
  Dims d = mkdims(g);
  long n = (long)g->nx * g->ny * g->nz;
  size_t vec = (((size_t)n * sizeof(float) + 255) / 256) * 256;
  char* base = (char*)ws;
  BicgState* st = (BicgState*)base;
  double* partials = (double*)(base + 256);
  char* v0 = base + 256 + (size_t)kBicgBlocks * 4 * sizeof(double);
  float* r = (float*)(v0 + 0 * vec);
  float* rhat = (float*)(v0 + 1 * vec);
  float* p[2] = {(float*)(v0 + 2 * vec), (float*)(v0 + 3 * vec)};
  float* q[2] = {(float*)(v0 + 4 * vec), (float*)(v0 + 5 * vec)};
  float* s = (float*)(v0 + 6 * vec);
  float* t = (float*)(v0 + 7 * vec);
  hipStream_t hs = S(stream);
  int nb = (int)std::min<long>(kBicgBlocks, (n + kBicgThreads - 1) / kBicgThreads);
  hipLaunchKernelGGL(k_bicg_reset, dim3(1), dim3(64), 0, hs, st, maxiter);
  hipLaunchKernelGGL(k_bicg_init, dim3(nb), dim3(kBicgThreads), 0, hs, d, rtt, rhs, x0, x, r, rhat, p[0], q[0],
                     partials, st, tol, atol);
  int rc = check_launch("lex_bicgstab(init)");
  if (rc) return rc;
  int cur = 0;
  const int poll = 8;
  for (int it = 0; it < maxiter; ++it) {
    hipLaunchKernelGGL(k_bicg_pq, dim3(nb), dim3(kBicgThreads), 0, hs, d, rtt, r, rhat, p[cur], q[cur], p[1 - cur],
                       q[1 - cur], partials, st);
    hipLaunchKernelGGL(k_bicg_st, dim3(nb), dim3(kBicgThreads), 0, hs, d, rtt, r, q[1 - cur], s, t, partials, st);
    hipLaunchKernelGGL(k_bicg_xr, dim3(nb), dim3(kBicgThreads), 0, hs, n, p[1 - cur], s, t, rhat, x, r, partials, st);
    cur = 1 - cur;
    rc = check_launch("lex_bicgstab(iter)");
    if (rc) return rc;
    if ((it % poll) == poll - 1) {
      int done = 0;
      hipError_t e = hipMemcpyAsync(&done, &st->done, sizeof(int), hipMemcpyDeviceToHost, hs);
      if (e == hipSuccess) e = hipStreamSynchronize(hs);
      if (e != hipSuccess) return fail((int)e, "lex_bicgstab: %s", hipGetErrorString(e));
      if (done) break;
    }
  }
  if (k_out) {
    int kk = 0;
    hipError_t e = hipMemcpyAsync(&kk, &st->k, sizeof(int), hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    if (e != hipSuccess) return fail((int)e, "lex_bicgstab: %s", hipGetErrorString(e));
    *k_out = kk;
  }
  return 0;
}

// ------------------------------- S1/M1/E1 ----------------------------------------
int lex_acoustic_substep(const lex_grid_t* g, const float* r, const float* pi0, const float* u, const float* v,
                         const float* w, const float* p, const float* ps, const float* ru, const float* rv,
                         const float* rw, const float* tr, float dts, float* uo, float* vo, float* wo, float* po,
                         void* stream) {
  REQGRID(g);
  REQ(r); REQ(pi0); REQ(u); REQ(v); REQ(w); REQ(p); REQ(ps); REQ(ru); REQ(rv); REQ(rw); REQ(tr);
  REQ(uo); REQ(vo); REQ(wo); REQ(po);
  if (uo == u || vo == v || wo == w || po == p) return fail(-1, "lex_acoustic_substep: outputs must not alias inputs%s");
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_acoustic, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, r, pi0, u, v, w, p, ps, ru, rv, rw, tr,
                     uo, vo, wo, po, dts);
  return check_launch("lex_acoustic_substep");
}

int lex_scalar_update(const lex_grid_t* g, const float* r, const float* s0, const float* s, const float* u,
                      const float* v, const float* w, const float* fd, const float* sfc, const float* heat,
                      const float* sgs, float dt, float weps, float* out, float* tend, void* stream) {
  REQGRID(g);
  REQ(r); REQ(s0); REQ(s); REQ(u); REQ(v); REQ(w); REQ(fd); REQ(out);
  if (out == s) return fail(-1, "lex_scalar_update: out must not alias s%s");
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_scalar_update, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, r, s0, s, u, v, w, fd, sfc, heat,
                     sgs, out, tend, dt, weps);
  return check_launch("lex_scalar_update");
}

int lex_momentum_update(const lex_grid_t* g, const float* u0, const float* v0, const float* w0, const float* p,
                        const float* tr, const float* au, const float* av, const float* aw, const float* fu,
                        const float* fv, const float* b8w, const float* su, const float* sv, const float* sw, float dt,
                        float* uo, float* vo, float* wo, float* duo, float* dvo, float* dwo, void* stream) {
  REQGRID(g);
  REQ(u0); REQ(v0); REQ(w0); REQ(p); REQ(tr); REQ(au); REQ(av); REQ(aw); REQ(b8w); REQ(uo); REQ(vo); REQ(wo);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_momentum_update, grd(d.nz + 1, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, u0, v0, w0, p, tr,
                     au, av, aw, fu, fv, b8w, su, sv, sw, uo, vo, wo, duo, dvo, dwo, dt);
  return check_launch("lex_momentum_update");
}

// ---------------------------------- G1 -------------------------------------------
int lex_smag_deformation(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w,
                         float* s11, float* s22, float* s33, float* s12, float* s13, float* s23, float* df,
                         void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(s11); REQ(s22); REQ(s33); REQ(s12); REQ(s13); REQ(s23); REQ(df);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_smag_deform, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, r, u, v, w, s11, s22, s33, s12,
                     s13, s23, df);
  return check_launch("lex_smag_deformation");
}

int lex_smag_nm(const lex_grid_t* g, const float* th, const float* qv, float* n2, void* stream) {
  REQGRID(g);
  REQ(th); REQ(qv); REQ(n2);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_smag_nm, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, th, qv, n2);
  return check_launch("lex_smag_nm");
}

int lex_smag_k(const lex_grid_t* g, const float* df, const float* n2, float cs, float prandtl, float* km, float* kh,
               void* stream) {
  REQGRID(g);
  REQ(df); REQ(n2); REQ(km); REQ(kh);
  Dims d = mkdims(g);
  // grid scale cbrt(dx dy dz) and (cs*gs)^2 in FP32 (turbulence.py:433-444)
  float gs = cbrtf(g->dx * g->dy * g->dz);
  float cg = cs * gs;
  long n = d.nc();
  hipLaunchKernelGGL(k_smag_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), n, df, n2, km, kh, cg * cg,
                     prandtl);
  return check_launch("lex_smag_k");
}

int lex_smag_momentum(const lex_grid_t* g, const float* r, const float* km, const float* s11, const float* s22,
                      const float* s33, const float* s12, const float* s13, const float* s23, float* su, float* sv,
                      float* sw, void* stream) {
  REQGRID(g);
  REQ(r); REQ(km); REQ(s11); REQ(s22); REQ(s33); REQ(s12); REQ(s13); REQ(s23); REQ(su); REQ(sv); REQ(sw);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_smag_uvw, grd(d.nz + 1, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, r, km, s11, s22, s33, s12,
                     s13, s23, su, sv, sw);
  return check_launch("lex_smag_momentum");
}

int lex_smag_scalar(const lex_grid_t* g, const float* r, const float* kh, const float* s, float* out, void* stream) {
  REQGRID(g);
  REQ(r); REQ(kh); REQ(s); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_smag_scalar, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, r, kh, s, out);
  return check_launch("lex_smag_scalar");
}

// ---------------------------------- I1 -------------------------------------------
int lex_base_state(int nz, double dz, double theta0, const double* rh, double* out, int max_iter, double tol,
                   int* iters, void* stream) {
  REQ(rh); REQ(out);
  if (nz < 3 || dz <= 0.0) return fail(-1, "lex_base_state: bad nz/dz%s");
  // out[5*(nz+2)] receives the iteration count (as a double)
  hipLaunchKernelGGL(k_base_state, dim3(1), dim3(64), 0, S(stream), nz, dz, theta0, rh, out, max_iter, tol,
                     out + 5 * (nz + 2));
  int rc = check_launch("lex_base_state");
  if (!rc && iters) {
    double hi = 0.0;
    hipError_t e = hipMemcpyAsync(&hi, out + 5 * (nz + 2), sizeof(double), hipMemcpyDeviceToHost, S(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
    if (e != hipSuccess) return fail((int)e, "lex_base_state: %s", hipGetErrorString(e));
    *iters = (int)hi;
  }
  return rc;
}

// ------------------------------ profiling ----------------------------------------
int lex_profile_enable(int on, int capacity_pairs) {
  Prof& p = prof();
  if (on) {
    int cap = capacity_pairs > 0 ? capacity_pairs : 8192;
    if ((int)p.ev.size() < 2 * cap) {
      for (hipEvent_t e : p.ev) (void)hipEventDestroy(e);
      p.ev.assign(2 * cap, nullptr);
      for (auto& e : p.ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return fail((int)r, "lex_profile_enable: %s", hipGetErrorString(r));
      }
      p.ids.assign(cap, -1);
    }
    p.used = 0;
  }
  p.on = on != 0;
  return 0;
}

int lex_profile_read(double* ms, int* counts, int n) {
  REQ(ms); REQ(counts);
  Prof& p = prof();
  for (int i = 0; i < n; ++i) {
    ms[i] = 0.0;
    counts[i] = 0;
  }
  if (p.used > 0) {
    hipError_t r = hipEventSynchronize(p.ev[p.used - 1]);
    if (r != hipSuccess) return fail((int)r, "lex_profile_read: %s", hipGetErrorString(r));
    for (int s = 0; s < p.used; s += 2) {
      float t = 0.f;
      r = hipEventElapsedTime(&t, p.ev[s], p.ev[s + 1]);
      if (r != hipSuccess) return fail((int)r, "lex_profile_read: %s", hipGetErrorString(r));
      int id = p.ids[s / 2];
      if (id >= 0 && id < n) {
        ms[id] += t;
        counts[id] += 1;
      }
    }
  }
  p.used = 0;
  return KID_COUNT;
}

const char* lex_profile_kernel_name(int id) { return kernel_name(id); }

// ---------------------------------- R2 -------------------------------------------
size_t lex_rk3_workspace_bytes(const lex_grid_t* g, int n_passive) {
  if (!grid_ok(g) || n_passive < 0) return 0;
  return rk3_layout(mkdims(g), n_passive).total;
}

int lex_rk3_step(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                 void* ws, size_t ws_bytes, void* stream) {
  REQGRID(g);
  REQ(o); REQ(b); REQ(state); REQ(ws);
  REQ(b->rho0); REQ(b->theta0); REQ(b->qv0); REQ(b->pi0);
  if (o->n_passive < 0 || o->n_passive > 8) return fail(-1, "lex_rk3_step: n_passive must be 0..8%s");
  if (o->n_sound < 1) return fail(-1, "lex_rk3_step: n_sound must be >= 1%s");
  for (int q = 0; q < 6 + o->n_passive; ++q)
    if (!state[q]) return fail(-2, "lex_rk3_step: null state pointer%s");
  Dims d = mkdims(g);
  Rk3Layout L = rk3_layout(d, o->n_passive);
  if (ws_bytes < L.total) return fail(-3, "lex_rk3_step: workspace too small%s");
  if (o->impl != 0 && o->impl != 1) return fail(-1, "lex_rk3_step: impl must be 0 or 1%s");
  if (o->impl == 0 && (64 * ((g->nz + 63) / 64)) * kFTY > 1024)
    return fail(-1, "lex_rk3_step: the fused kernels need nz <= 256 (use impl 1)%s");
  int rc = rk3_step(d, *o, *b, state, (char*)ws, L, S(stream));
  if (rc < 0) return fail(rc, "lex_rk3_step: unsupported configuration%s");
  if (rc > 0) return fail(rc, "lex_rk3_step: HIP error %s", hipGetErrorString((hipError_t)rc));
  return 0;
}

}  // extern "C"
