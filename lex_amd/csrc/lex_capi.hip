// extern "C" entry points of liblexhip.so (declared in include/lex_capi.h).
// Unity build: the kernel sources are included here so one hipcc invocation produces
// the whole library.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <cmath>
#include <mutex>

#include "lex_common.hpp"
#include "lex_ops.hip"
#include "lex_smag.hip"
#include "lex_pressure.hip"
#include "lex_stage.hip"
#include "lex_adjoint_step.hip"
#include "lex_conv.hip"
#include "lex_conv_f32.hip"
#include "lex_sgs_io.hip"
#include "lex_dl_post.hip"
#include "lex_dl_train.hip"
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
  if (!(g && g->nx >= 6 && g->ny >= 6 && g->nz >= 6 && g->dx > 0.f && g->dy > 0.f && g->dz > 0.f)) return false;
  // the kernels address fields with 32-bit element offsets: the largest padded field
  // (u, v or w) must stay below 2^30 elements
  const long px = g->nx + 7L, py = g->ny + 7L, pz = g->nz + 3L;
  return px * py * pz < kMaxOffset32;
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
  if (!grid_ok(g)) return fail(-1, "invalid grid (need nx,ny,nz >= 6, positive spacings and fewer than 2^30 "   \
                                  "elements per padded field)%s")
}  // namespace

extern "C" {

const char* lex_last_error(void) { return g_err.c_str(); }
int lex_abi_version(void) { return 2; }

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

int lex_coriolis(const lex_grid_t* g, const float* u, const float* v, float fcor, float* fu, float* fv,
                 void* stream) {
  REQGRID(g);
  REQ(u); REQ(v); REQ(fu); REQ(fv);
  Dims d = mkdims(g);
  const CorOpt c{1, fcor, 0.0f, 0.0f};
  hipLaunchKernelGGL(k_coriolis, grd(d.nz, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, u, v, fu, fv, c);
  return check_launch("lex_coriolis");
}

int lex_coriolis_lspgrad(const lex_grid_t* g, const float* u, const float* v, float fcor, float u_geo, float v_geo,
                         float* fu, float* fv, void* stream) {
  REQGRID(g);
  REQ(u); REQ(v); REQ(fu); REQ(fv);
  Dims d = mkdims(g);
  const CorOpt c{2, fcor, u_geo, v_geo};
  hipLaunchKernelGGL(k_coriolis, grd(d.nz, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, u, v, fu, fv, c);
  return check_launch("lex_coriolis_lspgrad");
}

int lex_momentum_rhs(const lex_grid_t* g, const float* au, const float* av, const float* aw, const float* fu,
                     const float* fv, const float* b8w, const float* su, const float* sv, const float* sw,
                     float* ru, float* rv, float* rw, void* stream) {
  REQGRID(g);
  REQ(au); REQ(av); REQ(aw); REQ(ru); REQ(rv); REQ(rw);
  if ((fu == nullptr) != (fv == nullptr)) return fail(-2, "lex_momentum_rhs: fu and fv must both be given or both NULL%s");
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_momentum_rhs, grd(d.nz + 1, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, au, av, aw, fu, fv,
                     b8w, su, sv, sw, ru, rv, rw);
  return check_launch("lex_momentum_rhs");
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
  return 256 + kBicgPartialBytes + 8 * vec;
}

}  // extern "C"

// Scratch of lex_bicgstab_noalloc: one copy per GPU, held until the process ends.
// Sized for LEX_BICGSTAB_NOALLOC_MAX_CELLS cells in the BiCGSTAB workspace layout
// (loop state, partial sums, eight vectors).
constexpr size_t kNoAllocVec = (size_t)LEX_BICGSTAB_NOALLOC_MAX_CELLS * sizeof(float);
constexpr size_t kNoAllocBytes = 256 + kBicgPartialBytes + 8 * kNoAllocVec;
static_assert(kNoAllocVec % 256 == 0, "vectors stay 256-byte aligned");
static __device__ __align__(256) char scratch[kNoAllocBytes];

namespace {
// The BiCGSTAB iteration over a workspace laid out as: loop state (256 B), the
// per-block FP64 partial sums, then eight 256-byte-aligned vectors of n floats
// (r, r-hat, p x2, q x2, s, t).  Shared by the caller-workspace and the
// allocation-free entry points.
int bicgstab_run(const lex_grid_t* g, const float* rtt, const float* rhs, const float* x0, float* x, double tol,
                 double atol, int maxiter, char* base, int* k_out, hipStream_t hs, const char* who) {
  Dims d = mkdims(g);
  long n = (long)g->nx * g->ny * g->nz;
  size_t vec = (((size_t)n * sizeof(float) + 255) / 256) * 256;
  BicgState* st = (BicgState*)base;
  double* partials = (double*)(base + 256);
  char* v0 = base + 256 + kBicgPartialBytes;
  float* r = (float*)(v0 + 0 * vec);
  float* rhat = (float*)(v0 + 1 * vec);
  float* p[2] = {(float*)(v0 + 2 * vec), (float*)(v0 + 3 * vec)};
  float* q[2] = {(float*)(v0 + 4 * vec), (float*)(v0 + 5 * vec)};
  float* s = (float*)(v0 + 6 * vec);
  float* t = (float*)(v0 + 7 * vec);
  // kBicgCellsPerThread cells per thread, at most kBicgBlocks blocks (the partial-sum capacity)
  const long cpb = (long)kBicgCellsPerThread * kBicgThreads;
  const int nb = (int)std::min<long>(kBicgBlocks, std::max<long>(1, (n + cpb - 1) / cpb));
  // the loop state (flags, counters, scalars) is cleared by a kernel, not a memset
  hipLaunchKernelGGL(k_bicg_reset, dim3(1), dim3(64), 0, hs, st, maxiter);
  LEX_LAUNCH(KID_BICG_INIT, k_bicg_init, dim3(nb), dim3(kBicgThreads), 0, hs, d, rtt, rhs, x0, x, r, rhat, p[0],
             q[0], partials);
  int rc = check_launch(who);
  if (rc) return rc;
  int cur = 0;
  const int poll = 8;
  int it = 0;
  for (; it < maxiter; ++it) {
    LEX_LAUNCH(KID_BICG_PQ, k_bicg_pq, dim3(nb), dim3(kBicgThreads), 0, hs, d, rtt, r, rhat, p[cur], q[cur],
               p[1 - cur], q[1 - cur], partials, st, it, nb, tol, atol);
    LEX_LAUNCH(KID_BICG_ST, k_bicg_st, dim3(nb), dim3(kBicgThreads), 0, hs, d, rtt, r, q[1 - cur], s, t, partials, st,
               it, nb, tol, atol);
    LEX_LAUNCH(KID_BICG_XR, k_bicg_xr, dim3(nb), dim3(kBicgThreads), 0, hs, n, p[1 - cur], s, t, rhat, x, r, partials,
               st, it, nb, tol, atol);
    cur = 1 - cur;
    rc = check_launch(who);
    if (rc) return rc;
    if ((it % poll) == poll - 1) {
      int done = 0;
      hipError_t e = hipMemcpyAsync(&done, &st->done, sizeof(int), hipMemcpyDeviceToHost, hs);
      if (e == hipSuccess) e = hipStreamSynchronize(hs);
      if (e != hipSuccess) return fail((int)e, "lex_bicgstab: %s", hipGetErrorString(e));
      if (done) {
        ++it;
        break;
      }
    }
  }
  // maxiter reached without the loop test failing: the exit code is the last count
  hipLaunchKernelGGL(k_bicg_finish, dim3(1), dim3(64), 0, hs, st, it);
  rc = check_launch(who);
  if (rc) return rc;
  if (k_out) {
    int kk = 0;
    hipError_t e = hipMemcpyAsync(&kk, &st->k_final, sizeof(int), hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    if (e != hipSuccess) return fail((int)e, "lex_bicgstab: %s", hipGetErrorString(e));
    *k_out = kk;
  }
  return 0;
}

// Device address of `scratch` per device, looked up once per device (with that device
// current) on the first call there.  In host code `scratch` is only a host-side stand-in
// for the symbol; kernels and copies get the looked-up address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(char** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail((int)e, "lex_bicgstab_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(-1, "lex_bicgstab_noalloc: device index out of range%s");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(e != hipSuccess ? (int)e : -4, "lex_bicgstab_noalloc: hipGetSymbolAddress(scratch): %s",
                  e != hipSuccess ? hipGetErrorString(e) : "null address");
    g_scratch_ptr[dev] = (char*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" {

int lex_bicgstab(const lex_grid_t* g, const float* rtt, const float* rhs, const float* x0, float* x, double tol,
                 double atol, int maxiter, void* ws, size_t ws_bytes, int* k_out, void* stream) {
  REQGRID(g);
  REQ(rtt); REQ(rhs); REQ(x0); REQ(x); REQ(ws);
  if (ws_bytes < lex_bicgstab_workspace_bytes(g)) return fail(-3, "lex_bicgstab: workspace too small%s");
  if (maxiter < 0) return fail(-1, "lex_bicgstab: maxiter < 0%s");
  if (x == x0) return fail(-1, "lex_bicgstab: x must not alias x0%s");
  return bicgstab_run(g, rtt, rhs, x0, x, tol, atol, maxiter, (char*)ws, k_out, S(stream), "lex_bicgstab");
}

int lex_bicgstab_noalloc(const lex_grid_t* g, const float* rtt, const float* rhs, const float* x0, float* x,
                         double tol, double atol, int maxiter, int* k_out, void* stream) {
  REQGRID(g);
  REQ(rtt); REQ(rhs); REQ(x0); REQ(x);
  if (maxiter < 0) return fail(-1, "lex_bicgstab_noalloc: maxiter < 0%s");
  if (x == x0) return fail(-1, "lex_bicgstab_noalloc: x must not alias x0%s");
  if ((long long)g->nx * g->ny * g->nz > (long long)LEX_BICGSTAB_NOALLOC_MAX_CELLS)
    return fail(-3, "lex_bicgstab_noalloc: nx*ny*nz exceeds LEX_BICGSTAB_NOALLOC_MAX_CELLS (the static "
                    "scratch's capacity)%s");
  char* base = nullptr;
  int rc = scratch_for_current_device(&base);
  if (rc) return rc;
  return bicgstab_run(g, rtt, rhs, x0, x, tol, atol, maxiter, base, k_out, S(stream), "lex_bicgstab_noalloc");
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

int lex_acoustic_substep_pair(const lex_grid_t* g, const float* r, const float* pi0, const float* u, const float* v,
                              const float* w, const float* p, const float* ps, const float* ru, const float* rv,
                              const float* rw, const float* tr, float dts, float* uo, float* vo, float* wo, float* po,
                              void* stream) {
  REQGRID(g);
  REQ(r); REQ(pi0); REQ(u); REQ(v); REQ(w); REQ(p); REQ(ps); REQ(ru); REQ(rv); REQ(rw); REQ(tr);
  REQ(uo); REQ(vo); REQ(wo); REQ(po);
  if (uo == u || vo == v || wo == w || po == p)
    return fail(-1, "lex_acoustic_substep_pair: outputs must not alias inputs%s");
  Dims d = mkdims(g);
  if (!acoustic2_ok(d)) return fail(-1, "lex_acoustic_substep_pair: nz must be <= 256%s");
  hipLaunchKernelGGL(k_acoustic2, acoustic2_grid(d), dim3(acoustic2_lanes(d), acoustic2_rows(d), 1), acoustic2_lds(d),
                     S(stream), d, r, pi0, u, v, w, p, ps, ru, rv, rw, tr, uo, vo, wo, po, dts, acoustic2_ax(d));
  return check_launch("lex_acoustic_substep_pair");
}

int lex_acoustic_substep_implicit(const lex_grid_t* g, const float* r, const float* pi0, const float* u,
                                  const float* v, const float* w, const float* p, const float* ps, const float* ru,
                                  const float* rv, const float* rw, const float* tr, float dts, float alpha, float* uo,
                                  float* vo, float* wo, float* po, void* stream) {
  REQGRID(g);
  REQ(r); REQ(pi0); REQ(u); REQ(v); REQ(w); REQ(p); REQ(ps); REQ(ru); REQ(rv); REQ(rw); REQ(tr);
  REQ(uo); REQ(vo); REQ(wo); REQ(po);
  if (uo == u || vo == v || wo == w || po == p)
    return fail(-1, "lex_acoustic_substep_implicit: outputs must not alias inputs%s");
  if (g->nz < 2 || g->nz > 256) return fail(-1, "lex_acoustic_substep_implicit: nz must be 2..256%s");
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_acoustic_vi, dim3(d.ny, d.nx, 1), dim3(acoustic_vi_lanes(d)), acoustic_vi_lds(d), S(stream),
                     d, r, pi0, u, v, w, p, ps, ru, rv, rw, tr, uo, vo, wo, po, dts, alpha);
  return check_launch("lex_acoustic_substep_implicit");
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
  hipLaunchKernelGGL(k_smag_uvw<false>, grd(d.nz, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, r, km, km, s11, s22, s33,
                     s12, s13, s23, su, sv, sw);
  return check_launch("lex_smag_momentum");
}

int lex_smag_scalar(const lex_grid_t* g, const float* r, const float* kh, const float* s, float* out, void* stream) {
  REQGRID(g);
  REQ(r); REQ(kh); REQ(s); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_smag_scalar, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, r, kh, kh, s, out);
  return check_launch("lex_smag_scalar");
}

int lex_smag_k2(const lex_grid_t* g, const float* r, const float* s11, const float* s22, const float* s33,
                const float* s12, const float* s13, const float* s23, const float* n2, float cs, float* kmh, float* kmv,
                void* stream) {
  REQGRID(g);
  REQ(r); REQ(s11); REQ(s22); REQ(s33); REQ(s12); REQ(s13); REQ(s23); REQ(n2); REQ(kmh); REQ(kmv);
  Dims d = mkdims(g);
  long n = d.nc();
  hipLaunchKernelGGL(k_smag_k2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), d, r, s11, s22, s33, s12,
                     s13, s23, n2, kmh, kmv, cs * cs);
  return check_launch("lex_smag_k2");
}

int lex_smag2_momentum(const lex_grid_t* g, const float* r, const float* kmh, const float* kmv, const float* s11,
                       const float* s22, const float* s33, const float* s12, const float* s13, const float* s23,
                       float* su, float* sv, float* sw, void* stream) {
  REQGRID(g);
  REQ(r); REQ(kmh); REQ(kmv); REQ(s11); REQ(s22); REQ(s33); REQ(s12); REQ(s13); REQ(s23); REQ(su); REQ(sv); REQ(sw);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_smag_uvw<false>, grd(d.nz, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, r, kmh, kmv, s11, s22,
                     s33, s12, s13, s23, su, sv, sw);
  return check_launch("lex_smag2_momentum");
}

int lex_smag2_scalar(const lex_grid_t* g, const float* r, const float* khh, const float* khv, const float* s,
                     float* out, void* stream) {
  REQGRID(g);
  REQ(r); REQ(khh); REQ(khv); REQ(s); REQ(out);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_smag_scalar, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, r, khh, khv, s, out);
  return check_launch("lex_smag2_scalar");
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

int lex_rk3_workspace_sgs_offsets(const lex_grid_t* g, int n_passive, size_t* offsets) {
  REQGRID(g);
  REQ(offsets);
  if (n_passive < 0 || n_passive > 8) return fail(-1, "lex_rk3_workspace_sgs_offsets: n_passive must be 0..8%s");
  const Rk3Layout L = rk3_layout(mkdims(g), n_passive);
  offsets[0] = L.off_su;
  offsets[1] = L.off_sv;
  offsets[2] = L.off_sw;
  for (int q = 0; q < 2 + n_passive; ++q) offsets[3 + q] = L.off_ss + q * al256(L.ic_ * 4);
  return 0;
}

static int rk3_step_entry(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                          float* const* state, void* ws, size_t ws_bytes, void* stream,
                          StageSgsFn stage_sgs = nullptr, void* stage_user = nullptr);

int lex_rk3_step(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                 void* ws, size_t ws_bytes, void* stream) {
  REQ(o);
  if (o->turb < 0 || o->turb > 2) return fail(-1, "lex_rk3_step: turb must be 0, 1 or 2%s");
  if (o->cor < 0 || o->cor > 2) return fail(-1, "lex_rk3_step: cor must be 0, 1 or 2%s");
  return rk3_step_entry(g, o, b, state, ws, ws_bytes, stream);
}

static int rk3_step_entry(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                          float* const* state, void* ws, size_t ws_bytes, void* stream, StageSgsFn stage_sgs,
                          void* stage_user) {
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
  if (o->impl < 0 || o->impl > 2) return fail(-1, "lex_rk3_step: impl must be 0, 1 or 2%s");
  if (o->solver != 0 && o->solver != 2 && o->solver != 3) return fail(-1, "lex_rk3_step: solver must be 0, 2 or 3%s");
  if (o->solver == 3 && (g->nz < 2 || g->nz > 256)) return fail(-1, "lex_rk3_step: solver 3 needs nz 2..256%s");
  if (o->damp && g->nz > 256) return fail(-1, "lex_rk3_step: damping needs nz <= 256%s");
  if (o->sfc && !b->surface_t) return fail(-2, "lex_rk3_step: sfc needs the surface temperature (surface_t)%s");
  if (o->impl != 1 && (64 * ((g->nz + 63) / 64)) * kFTY > 1024)
    return fail(-1, "lex_rk3_step: the fused kernels need nz <= 256 (use impl 1)%s");
  int rc = rk3_step(d, *o, *b, state, (char*)ws, L, S(stream), stage_sgs, stage_user);
  if (rc == -5 && stage_sgs) return fail(-5, "lex_rk3_step_stage_sgs: the stage SGS function failed%s");
  if (rc < 0) return fail(rc, "lex_rk3_step: unsupported configuration%s");
  if (rc > 0) return fail(rc, "lex_rk3_step: HIP error %s", hipGetErrorString((hipError_t)rc));
  return 0;
}

// ------------------------------------------------------------------ adjoint (VJP)
int lex_fold(const lex_grid_t* g, int mode, float* a, void* stream) {
  REQGRID(g);
  REQ(a);
  if (mode < 0 || mode > 3) return fail(-1, "lex_fold: mode must be 0..3%s");
  Dims d = mkdims(g);
  launch_fold(d, mode, a, S(stream));
  return check_launch("lex_fold");
}

// Stream-ordered scratch for the slot planes of a stand-alone advection reverse: taken
// from the stream's memory pool before the kernels and returned after them on the same
// stream.
struct SlotScratch {
  float* p = nullptr;
  hipStream_t st = nullptr;
  hipError_t take(size_t bytes, hipStream_t s) {
    st = s;
    keep_pool_memory();
    return hipMallocAsync((void**)&p, bytes, s);
  }
  // The default pool returns freed memory to the driver at every synchronisation
  // (release threshold 0), so each call would map its scratch afresh; keep it pooled.
  static void keep_pool_memory() {
    static std::once_flag once;
    std::call_once(once, [] {
      int dev = 0;
      hipMemPool_t pool;
      if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      }
    });
  }
  ~SlotScratch() {
    if (p) (void)hipFreeAsync(p, st);
  }
};

int lex_adv_scalar_vjp(const lex_grid_t* g, const float* r, const float* s, const float* u, const float* v,
                       const float* w, const float* lam, float weps, float* ls, float* lu, float* lv, float* lw,
                       void* stream) {
  REQGRID(g);
  REQ(r); REQ(s); REQ(u); REQ(v); REQ(w); REQ(lam); REQ(ls); REQ(lu); REQ(lv); REQ(lw);
  Dims d = mkdims(g);
  SlotScratch sc;
  if (sc.take(slot_bytes_scalar(d), S(stream)) != hipSuccess)
    return fail(-4, "%s: no device memory for the slot planes", "lex_adv_scalar_vjp");
  const SlotPlanes sp = slots_scalar(d, sc.p);
  hipLaunchKernelGGL(k_adv_scalar_vjp, grd(d.nz, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, r, s, u, v, w, lam, weps, sp,
                     ls, lu, lv, lw);
  hipLaunchKernelGGL(k_gather_scalar, grd(d.nz + 2, d.ny + 6, d.nx + 6), blk(), 0, S(stream), d, sp, ls);
  return check_launch("lex_adv_scalar_vjp");
}

// one momentum reverse (which = 0, 1, 2 for u, v, w) and the three gathers of its slots
static int adv_mom_vjp(const char* name, int which, const Dims& d, const float* r, const float* u, const float* v,
                       const float* w, const float* lam, float* lu, float* lv, float* lw, hipStream_t st) {
  SlotScratch sc;
  const size_t bytes = (size_t)kSlotsMom * 4 *
                       (which == 0 ? slots_u(d, nullptr).plane : (which == 1 ? slots_v(d, nullptr).plane
                                                                               : slots_w(d, nullptr).plane));
  if (sc.take(bytes, st) != hipSuccess) return fail(-4, "%s: no device memory for the slot planes", name);
  SlotPlanes none = slot_planes(nullptr, 0, 0, 0);
  SlotPlanes U = none, V = none, W = none;
  if (which == 0) {
    U = slots_u(d, sc.p);
    hipLaunchKernelGGL(k_adv_u_vjp, grd(d.nz, d.ny + 1, d.nx + 2), blk(), 0, st, d, r, u, v, w, lam, U, lu);
  } else if (which == 1) {
    V = slots_v(d, sc.p);
    hipLaunchKernelGGL(k_adv_v_vjp, grd(d.nz, d.ny + 2, d.nx + 1), blk(), 0, st, d, r, u, v, w, lam, V, lv);
  } else {
    W = slots_w(d, sc.p);
    hipLaunchKernelGGL(k_adv_w_vjp, grd(d.nz + 1, d.ny + 1, d.nx + 1), blk(), 0, st, d, r, u, v, w, lam, W);
  }
  hipLaunchKernelGGL(k_gather_lu, grd(d.nz + 2, d.ny + 6, d.nx + 7), blk(), 0, st, d, U, V, W, lu);
  hipLaunchKernelGGL(k_gather_lv, grd(d.nz + 2, d.ny + 7, d.nx + 6), blk(), 0, st, d, U, V, W, lv);
  hipLaunchKernelGGL(k_gather_lw, grd(d.nz + 3, d.ny + 6, d.nx + 6), blk(), 0, st, d, U, V, W, lw);
  return check_launch(name);
}

int lex_adv_u_vjp(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w,
                  const float* lam, float* lu, float* lv, float* lw, void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(lam); REQ(lu); REQ(lv); REQ(lw);
  return adv_mom_vjp("lex_adv_u_vjp", 0, mkdims(g), r, u, v, w, lam, lu, lv, lw, S(stream));
}

int lex_adv_v_vjp(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w,
                  const float* lam, float* lu, float* lv, float* lw, void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(lam); REQ(lu); REQ(lv); REQ(lw);
  return adv_mom_vjp("lex_adv_v_vjp", 1, mkdims(g), r, u, v, w, lam, lu, lv, lw, S(stream));
}

int lex_adv_w_vjp(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w,
                  const float* lam, float* lu, float* lv, float* lw, void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(lam); REQ(lu); REQ(lv); REQ(lw);
  return adv_mom_vjp("lex_adv_w_vjp", 2, mkdims(g), r, u, v, w, lam, lu, lv, lw, S(stream));
}

int lex_buoyancy_vjp(const lex_grid_t* g, const float* th0, const float* qv0, const float* th, const float* qv,
                     const float* lam_b8w, float* lth, float* lqv, void* stream) {
  REQGRID(g);
  REQ(th0); REQ(qv0); REQ(th); REQ(qv); REQ(lam_b8w); REQ(lth); REQ(lqv);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_b8w_vjp, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, th0, qv0, th, qv, lam_b8w, lth, lqv);
  return check_launch("lex_buoyancy_vjp");
}

int lex_theta_rho_vjp(long n, const float* th, const float* qv, const float* ltr, float* lth, float* lqv,
                      void* stream) {
  REQ(th); REQ(qv); REQ(ltr); REQ(lth); REQ(lqv);
  if (n <= 0) return fail(-1, "lex_theta_rho_vjp: n <= 0%s");
  hipLaunchKernelGGL(k_theta_rho_vjp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), n, th, qv, ltr, lth,
                     lqv);
  return check_launch("lex_theta_rho_vjp");
}

int lex_acoustic_substep_vjp(const lex_grid_t* g, const float* r, const float* pi0, const float* u, const float* v,
                             const float* w, const float* p, const float* ps, const float* ru, const float* rv,
                             const float* rw, const float* tr, float dts, float* luo, float* lvo, float* lwo,
                             float* lpo, float* lU, float* lV, float* lW, float* lu, float* lv, float* lw, float* lp,
                             float* lps, float* ltr, float* lru, float* lrv, float* lrw, void* stream) {
  REQGRID(g);
  REQ(r); REQ(pi0); REQ(u); REQ(v); REQ(w); REQ(p); REQ(ps); REQ(ru); REQ(rv); REQ(rw); REQ(tr);
  REQ(luo); REQ(lvo); REQ(lwo); REQ(lpo); REQ(lU); REQ(lV); REQ(lW);
  REQ(lu); REQ(lv); REQ(lw); REQ(lp); REQ(lps); REQ(ltr); REQ(lru); REQ(lrv); REQ(lrw);
  if (lu == luo || lv == lvo || lw == lwo || lp == lpo)
    return fail(-1, "lex_acoustic_substep_vjp: input adjoints must not alias output adjoints%s");
  Dims d = mkdims(g);
  hipStream_t st = S(stream);
  launch_fold(d, 1, luo, st);
  launch_fold(d, 2, lvo, st);
  launch_fold(d, 3, lwo, st);
  launch_fold(d, 0, lpo, st);
  const size_t nU = (size_t)(d.nx + 1) * d.ny * d.nz, nV = (size_t)d.nx * (d.ny + 1) * d.nz,
               nW = (size_t)d.nx * d.ny * (d.nz + 1);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((nU + 255) / 256)), dim3(256), 0, st, (long)nU, lU, 0.0f);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((nV + 255) / 256)), dim3(256), 0, st, (long)nV, lV, 0.0f);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((nW + 255) / 256)), dim3(256), 0, st, (long)nW, lW, 0.0f);
  // the step's atomic-free two-pass reverse: faces (velocity adjoints accumulated, rhs,
  // face weights g dts into lU, lV, lW), then cells (pi', pi*, theta_rho adjoints)
  hipLaunchKernelGGL(k_acoustic_vjp_faces3, grd(d.nz, d.ny + 1, d.nx + 1), blk(), 0, st, d, r, pi0, ps,
                     (const float*)lpo, (const float*)luo, (const float*)lvo, (const float*)lwo, dts, 1, lu, lv, lw, lru,
                     lrv, lrw, lU, lV, lW);
  hipLaunchKernelGGL(k_acoustic_vjp_cells3, grd(d.nz, d.ny + 2, d.nx + 2), blk(), 0, st, d, r, u, v, w, p, ru, rv, rw,
                     tr, (const float*)lpo, (const float*)lU, (const float*)lV, (const float*)lW, dts, lp, ltr, lps);
  return check_launch("lex_acoustic_substep_vjp");
}

size_t lex_smag_vjp_workspace_bytes(const lex_grid_t* g) {
  if (!grid_ok(g)) return 0;
  Dims d = mkdims(g);
  return 20 * al256((size_t)d.nc() * 4) + 256 + al256((size_t)kSmagKBlocks * sizeof(double));
}

int lex_smag_vjp(const lex_grid_t* g, const float* r, const float* u, const float* v, const float* w, const float* th,
                 const float* qv, int n_scalars, const float* const* scalars, float cs, float prandtl,
                 const float* lsu, const float* lsv, const float* lsw, const float* const* lss, float* lu, float* lv,
                 float* lw, float* lth, float* lqv, float* const* lscalars, double* lcs, void* ws, size_t ws_bytes,
                 void* stream) {
  REQGRID(g);
  REQ(r); REQ(u); REQ(v); REQ(w); REQ(th); REQ(qv); REQ(lsu); REQ(lsv); REQ(lsw); REQ(lu); REQ(lv); REQ(lw);
  REQ(lth); REQ(lqv); REQ(ws);
  if (n_scalars < 0 || n_scalars > 10) return fail(-1, "lex_smag_vjp: n_scalars must be 0..10%s");
  if (n_scalars > 0 && (!scalars || !lss || !lscalars)) return fail(-2, "lex_smag_vjp: null scalar lists%s");
  if (ws_bytes < lex_smag_vjp_workspace_bytes(g)) return fail(-3, "lex_smag_vjp: workspace too small%s");
  Dims d = mkdims(g);
  hipStream_t st = S(stream);
  const size_t C = al256((size_t)d.nc() * 4);
  float* sm[10];
  float* lsm[10];
  for (int q = 0; q < 10; ++q) {
    sm[q] = (float*)((char*)ws + q * C);
    lsm[q] = (float*)((char*)ws + (10 + q) * C);
  }
  double* lcg2 = (double*)((char*)ws + 20 * C);
  long n = (long)d.nc();
  dim3 e1((unsigned)((n + 255) / 256)), e2(256);
  hipLaunchKernelGGL(k_smag_deform, grd(d.nz, d.ny, d.nx), blk(), 0, st, d, r, u, v, w, sm[0], sm[1], sm[2], sm[3],
                     sm[4], sm[5], sm[6]);
  hipLaunchKernelGGL(k_smag_nm, grd(d.nz, d.ny, d.nx), blk(), 0, st, d, th, qv, sm[7]);
  float gs = cbrtf(d.dx * d.dy * d.dz);
  float cg = cs * gs;
  hipLaunchKernelGGL(k_smag_k, e1, e2, 0, st, n, sm[6], sm[7], sm[8], sm[9], cg * cg, prandtl);
  for (int q = 0; q < 10; ++q) hipLaunchKernelGGL(k_fill, e1, e2, 0, st, n, lsm[q], 0.0f);
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, st, lcg2);
  for (int q = 0; q < n_scalars; ++q) {
    if (!scalars[q] || !lss[q] || !lscalars[q]) return fail(-2, "lex_smag_vjp: null scalar pointer%s");
    hipLaunchKernelGGL(k_smag_scalar_vjp, grd(d.nz + 2, d.ny + 6, d.nx + 6), blk(), 0, st, d, r, (const float*)sm[9], scalars[q],
                       lss[q], lsm[9], lscalars[q]);
  }
  hipLaunchKernelGGL(k_smag_uvw_vjp, grd(d.nz + 2, d.ny + 6, d.nx + 6), blk(), 0, st, d, r, (const float*)sm[8],
                     (const float*)sm[0], (const float*)sm[1], (const float*)sm[2], (const float*)sm[3],
                     (const float*)sm[4], (const float*)sm[5], lsu, lsv, lsw, lsm[8], lsm[0], lsm[1], lsm[2], lsm[3],
                     lsm[4], lsm[5]);
  const unsigned nbk = smag_k_blocks(n);
  double* kpart = (double*)((char*)ws + 20 * C + 256);
  hipLaunchKernelGGL(k_smag_k_vjp, dim3(nbk), dim3(256), 0, st, n, (const float*)sm[6], (const float*)sm[7],
                     (const float*)lsm[8], (const float*)lsm[9], cg * cg, prandtl, lsm[6], lsm[7], kpart);
  hipLaunchKernelGGL(k_smag_k_sum, dim3(1), dim3(256), 0, st, (const double*)kpart, (int)nbk, lcg2);
  for (int q = 0; q < 8; ++q) launch_fold(d, 0, lsm[q], st);
  hipLaunchKernelGGL(k_smag_nm_vjp, grd(d.nz, d.ny, d.nx), blk(), 0, st, d, th, qv, (const float*)lsm[7], lth, lqv);
  hipLaunchKernelGGL(k_smag_deform_h, grd(d.nz, d.ny, d.nx), blk(), 0, st, d, r, (const float*)sm[0],
                     (const float*)sm[1], (const float*)sm[2], (const float*)sm[3], (const float*)sm[4],
                     (const float*)sm[5], (const float*)lsm[6], lsm[0], lsm[1], lsm[2], lsm[3], lsm[4], lsm[5]);
  hipLaunchKernelGGL(k_smag_deform_vjp, grd(d.nz + 3, d.ny + 7, d.nx + 7), blk(), 0, st, d, (const float*)lsm[0],
                     (const float*)lsm[1], (const float*)lsm[2], (const float*)lsm[3], (const float*)lsm[4],
                     (const float*)lsm[5], lu, lv, lw);
  if (lcs)
    hipLaunchKernelGGL(k_cs_grad, dim3(1), dim3(64), 0, st, (const double*)lcg2,
                       2.0 * (double)cs * (double)gs * (double)gs, lcs);
  return check_launch("lex_smag_vjp");
}

size_t lex_rk3_vjp_workspace_bytes(const lex_grid_t* g, float dt, int n_sound, int n_passive) {
  if (!grid_ok(g) || n_passive < 0 || n_passive > 8 || n_sound < 1 || !(dt > 0.f)) return 0;
  Dims d = mkdims(g);
  return vjp_layout(d, (double)dt, n_sound, n_passive).total;
}

int lex_rk3_step_vjp(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                     float* const* lam_out, float* const* lam_in, double* lam_cs, void* ws, size_t ws_bytes,
                     void* stream) {
  REQGRID(g);
  REQ(o); REQ(b); REQ(state); REQ(lam_out); REQ(lam_in); REQ(ws);
  REQ(b->rho0); REQ(b->theta0); REQ(b->qv0); REQ(b->pi0);
  if (o->n_passive < 0 || o->n_passive > 8) return fail(-1, "lex_rk3_step_vjp: n_passive must be 0..8%s");
  if (o->n_sound < 1 || !(o->dt > 0.f)) return fail(-1, "lex_rk3_step_vjp: need n_sound >= 1 and dt > 0%s");
  for (int q = 0; q < 6 + o->n_passive; ++q) {
    if (!state[q] || !lam_out[q] || !lam_in[q]) return fail(-2, "lex_rk3_step_vjp: null field pointer%s");
    if (lam_in[q] == lam_out[q] || lam_in[q] == state[q])
      return fail(-1, "lex_rk3_step_vjp: lam_in must not alias lam_out or the state%s");
  }
  if (o->solver == 3) return fail(-1, "lex_rk3_step_vjp: the adjoint covers solver_opt 2 (explicit acoustic substeps)%s");
  if (o->damp || o->rad || o->sfc || o->turb == 2 || o->cor) return fail(-1, "lex_rk3_step_vjp: the adjoint covers damp_opt, rad_opt, sfc_opt, cor_opt and turb_opt 2 off%s");
  Dims d = mkdims(g);
  VjpLayout L = vjp_layout(d, (double)o->dt, o->n_sound, o->n_passive);
  if (ws_bytes < L.total) return fail(-3, "lex_rk3_step_vjp: workspace too small%s");
  int rc = rk3_step_vjp(d, *o, *b, state, lam_out, lam_in, lam_cs, (char*)ws, (char*)ws + L.off_lXA, L, S(stream));
  if (rc > 0) return fail(rc, "lex_rk3_step_vjp: HIP error %s", hipGetErrorString((hipError_t)rc));
  return 0;
}

size_t lex_rk3_tape_bytes(const lex_grid_t* g, float dt, int n_sound, int n_passive) {
  if (!grid_ok(g) || n_passive < 0 || n_passive > 8 || n_sound < 1 || !(dt > 0.f)) return 0;
  Dims d = mkdims(g);
  return vjp_layout(d, (double)dt, n_sound, n_passive).off_lXA;
}

size_t lex_rk3_adjoint_bytes(const lex_grid_t* g, float dt, int n_sound, int n_passive) {
  if (!grid_ok(g) || n_passive < 0 || n_passive > 8 || n_sound < 1 || !(dt > 0.f)) return 0;
  Dims d = mkdims(g);
  VjpLayout L = vjp_layout(d, (double)dt, n_sound, n_passive);
  return L.total - L.off_lXA;
}

int lex_rk3_step_tape(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                      float* const* state_out, void* tape, size_t tape_bytes, void* stream) {
  REQGRID(g);
  REQ(o); REQ(b); REQ(state); REQ(state_out); REQ(tape);
  REQ(b->rho0); REQ(b->theta0); REQ(b->qv0); REQ(b->pi0);
  if (o->n_passive < 0 || o->n_passive > 8) return fail(-1, "lex_rk3_step_tape: n_passive must be 0..8%s");
  if (o->n_sound < 1 || !(o->dt > 0.f)) return fail(-1, "lex_rk3_step_tape: need n_sound >= 1 and dt > 0%s");
  for (int q = 0; q < 6 + o->n_passive; ++q) {
    if (!state[q] || !state_out[q]) return fail(-2, "lex_rk3_step_tape: null field pointer%s");
    if (state[q] == state_out[q]) return fail(-1, "lex_rk3_step_tape: state_out must not alias state%s");
  }
  if (o->solver == 3) return fail(-1, "lex_rk3_step_tape: the adjoint covers solver_opt 2 (explicit acoustic substeps)%s");
  if (o->damp || o->rad || o->sfc || o->turb == 2 || o->cor) return fail(-1, "lex_rk3_step_tape: the adjoint covers damp_opt, rad_opt, sfc_opt, cor_opt and turb_opt 2 off%s");
  Dims d = mkdims(g);
  VjpLayout L = vjp_layout(d, (double)o->dt, o->n_sound, o->n_passive);
  if (tape_bytes < L.off_lXA) return fail(-3, "lex_rk3_step_tape: tape too small%s");
  rk3_forward_tape(d, *o, *b, state, (char*)tape, L, S(stream), state_out);
  return check_launch("lex_rk3_step_tape");
}

int lex_rk3_step_vjp_taped(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                           float* const* state, const void* tape, size_t tape_bytes, float* const* lam_out,
                           float* const* lam_in, double* lam_cs, void* adj, size_t adj_bytes, void* stream) {
  REQGRID(g);
  REQ(o); REQ(b); REQ(state); REQ(tape); REQ(lam_out); REQ(lam_in); REQ(adj);
  REQ(b->rho0); REQ(b->theta0); REQ(b->qv0); REQ(b->pi0);
  if (o->n_passive < 0 || o->n_passive > 8) return fail(-1, "lex_rk3_step_vjp_taped: n_passive must be 0..8%s");
  if (o->n_sound < 1 || !(o->dt > 0.f)) return fail(-1, "lex_rk3_step_vjp_taped: need n_sound >= 1 and dt > 0%s");
  for (int q = 0; q < 6 + o->n_passive; ++q) {
    if (!state[q] || !lam_out[q] || !lam_in[q]) return fail(-2, "lex_rk3_step_vjp_taped: null field pointer%s");
    if (lam_in[q] == lam_out[q] || lam_in[q] == state[q])
      return fail(-1, "lex_rk3_step_vjp_taped: lam_in must not alias lam_out or the state%s");
  }
  if (o->solver == 3) return fail(-1, "lex_rk3_step_vjp_taped: the adjoint covers solver_opt 2 (explicit acoustic substeps)%s");
  if (o->damp || o->rad || o->sfc || o->turb == 2 || o->cor) return fail(-1, "lex_rk3_step_vjp_taped: the adjoint covers damp_opt, rad_opt, sfc_opt, cor_opt and turb_opt 2 off%s");
  Dims d = mkdims(g);
  VjpLayout L = vjp_layout(d, (double)o->dt, o->n_sound, o->n_passive);
  if (tape_bytes < L.off_lXA) return fail(-3, "lex_rk3_step_vjp_taped: tape too small%s");
  if (adj_bytes < L.total - L.off_lXA) return fail(-3, "lex_rk3_step_vjp_taped: adjoint workspace too small%s");
  int rc = rk3_step_vjp(d, *o, *b, state, lam_out, lam_in, lam_cs, (char*)tape, (char*)adj, L, S(stream), false);
  if (rc > 0) return fail(rc, "lex_rk3_step_vjp_taped: HIP error %s", hipGetErrorString((hipError_t)rc));
  return 0;
}

// recorded step and reverse with caller-supplied SGS tendencies (turb 2: the learned net)
int lex_rk3_step_tape_ext_sgs(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                              float* const* state, const float* su, const float* sv, const float* sw,
                              const float* const* ss, float* const* state_out, void* tape, size_t tape_bytes,
                              void* stream) {
  REQGRID(g);
  REQ(o); REQ(b); REQ(state); REQ(state_out); REQ(tape); REQ(su); REQ(sv); REQ(sw); REQ(ss);
  REQ(b->rho0); REQ(b->theta0); REQ(b->qv0); REQ(b->pi0);
  if (o->n_passive < 0 || o->n_passive > 8) return fail(-1, "lex_rk3_step_tape_ext_sgs: n_passive must be 0..8%s");
  if (o->n_sound < 1 || !(o->dt > 0.f)) return fail(-1, "lex_rk3_step_tape_ext_sgs: need n_sound >= 1 and dt > 0%s");
  for (int q = 0; q < 6 + o->n_passive; ++q) {
    if (!state[q] || !state_out[q]) return fail(-2, "lex_rk3_step_tape_ext_sgs: null field pointer%s");
    if (state[q] == state_out[q]) return fail(-1, "lex_rk3_step_tape_ext_sgs: state_out must not alias state%s");
  }
  for (int q = 0; q < 2 + o->n_passive; ++q)
    if (!ss[q]) return fail(-2, "lex_rk3_step_tape_ext_sgs: null scalar SGS pointer%s");
  if (o->solver == 3) return fail(-1, "lex_rk3_step_tape_ext_sgs: the adjoint covers solver_opt 2 (explicit acoustic substeps)%s");
  if (o->damp || o->rad || o->sfc) return fail(-1, "lex_rk3_step_tape_ext_sgs: the adjoint covers damp_opt, rad_opt and sfc_opt off%s");
  Dims d = mkdims(g);
  VjpLayout L = vjp_layout(d, (double)o->dt, o->n_sound, o->n_passive);
  if (tape_bytes < L.off_lXA) return fail(-3, "lex_rk3_step_tape_ext_sgs: tape too small%s");
  hipStream_t st = S(stream);
  char* t = (char*)tape;
  auto cp = [&](const float* src, size_t off, size_t n) {
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (long)n, src, (float*)(t + off));
  };
  cp(su, L.off_su, L.iu_);
  cp(sv, L.off_sv, L.iv_);
  cp(sw, L.off_sw, L.iw_);
  for (int q = 0; q < 2 + o->n_passive; ++q) cp(ss[q], L.off_ss + q * al256(L.ic_ * 4), L.ic_);
  lex_step_opts_t o2 = *o;
  o2.turb = kTurbExternal;
  rk3_forward_tape(d, o2, *b, state, t, L, st, state_out);
  return check_launch("lex_rk3_step_tape_ext_sgs");
}

int lex_rk3_step_vjp_taped_ext_sgs(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                                   float* const* state, const void* tape, size_t tape_bytes, float* const* lam_out,
                                   float* const* lam_in, float* lsu, float* lsv, float* lsw, float* const* lss,
                                   void* adj, size_t adj_bytes, void* stream) {
  REQGRID(g);
  REQ(o); REQ(b); REQ(state); REQ(tape); REQ(lam_out); REQ(lam_in); REQ(adj); REQ(lsu); REQ(lsv); REQ(lsw); REQ(lss);
  REQ(b->rho0); REQ(b->theta0); REQ(b->qv0); REQ(b->pi0);
  if (o->n_passive < 0 || o->n_passive > 8) return fail(-1, "lex_rk3_step_vjp_taped_ext_sgs: n_passive must be 0..8%s");
  if (o->n_sound < 1 || !(o->dt > 0.f))
    return fail(-1, "lex_rk3_step_vjp_taped_ext_sgs: need n_sound >= 1 and dt > 0%s");
  for (int q = 0; q < 6 + o->n_passive; ++q) {
    if (!state[q] || !lam_out[q] || !lam_in[q]) return fail(-2, "lex_rk3_step_vjp_taped_ext_sgs: null field pointer%s");
    if (lam_in[q] == lam_out[q] || lam_in[q] == state[q])
      return fail(-1, "lex_rk3_step_vjp_taped_ext_sgs: lam_in must not alias lam_out or the state%s");
  }
  for (int q = 0; q < 2 + o->n_passive; ++q)
    if (!lss[q]) return fail(-2, "lex_rk3_step_vjp_taped_ext_sgs: null scalar SGS adjoint pointer%s");
  if (o->solver == 3) return fail(-1, "lex_rk3_step_vjp_taped_ext_sgs: the adjoint covers solver_opt 2 (explicit acoustic substeps)%s");
  if (o->damp || o->rad || o->sfc) return fail(-1, "lex_rk3_step_vjp_taped_ext_sgs: the adjoint covers damp_opt, rad_opt and sfc_opt off%s");
  Dims d = mkdims(g);
  VjpLayout L = vjp_layout(d, (double)o->dt, o->n_sound, o->n_passive);
  if (tape_bytes < L.off_lXA) return fail(-3, "lex_rk3_step_vjp_taped_ext_sgs: tape too small%s");
  if (adj_bytes < L.total - L.off_lXA) return fail(-3, "lex_rk3_step_vjp_taped_ext_sgs: adjoint workspace too small%s");
  lex_step_opts_t o2 = *o;
  o2.turb = kTurbExternal;
  hipStream_t st = S(stream);
  int rc = rk3_step_vjp(d, o2, *b, state, lam_out, lam_in, nullptr, (char*)tape, (char*)adj, L, st, false);
  if (rc > 0) return fail(rc, "lex_rk3_step_vjp_taped_ext_sgs: HIP error %s", hipGetErrorString((hipError_t)rc));
  // the adjoints of the frozen SGS tendencies sit in the adjoint workspace
  const char* a = (const char*)adj;
  auto cp = [&](size_t off, float* dst, size_t n) {
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (long)n,
                       (const float*)(a + (off - L.off_lXA)), dst);
  };
  cp(L.off_lsu, lsu, L.iu_);
  cp(L.off_lsv, lsv, L.iv_);
  cp(L.off_lsw, lsw, L.iw_);
  for (int q = 0; q < 2 + o->n_passive; ++q) cp(L.off_lss + q * al256(L.ic_ * 4), lss[q], L.ic_);
  return check_launch("lex_rk3_step_vjp_taped_ext_sgs");
}

// ------------------------------------------------------------------ X4 conv3d (BF16 MFMA)
}  // extern "C"
namespace {
inline int padc(int c) { return c <= 16 ? 16 : (c <= 32 ? 32 : 64); }
// forward and data-gradient convolutions take up to 64 channels; the weight gradient 32
inline bool conv_ok(int nx, int ny, int nz, int cin, int cout, int cmax = 32) {
  return nx > 0 && ny > 0 && nz > 0 && cin >= 1 && cin <= cmax && cout >= 1 && cout <= cmax;
}
inline int conv_nstep(int cinp) { return cinp <= 32 ? (27 + 32 / cinp - 1) / (32 / cinp) : 27 * (cinp / 32); }
}  // namespace
extern "C" {

size_t lex_conv3d_packed_weight_bytes(int cin, int cout, int flip) {
  int ci = flip ? cout : cin, co = flip ? cin : cout;
  if (ci < 1 || ci > 64 || co < 1 || co > 64) return 0;
  int cip = padc(ci), cop = padc(co);
  return (size_t)cop * conv_nstep(cip) * 32 * 2;
}

int lex_conv3d_pack_weights(const float* w, int cin, int cout, int flip, void* wp, void* stream) {
  REQ(w); REQ(wp);
  if (cin < 1 || cin > 64 || cout < 1 || cout > 64) return fail(-1, "lex_conv3d_pack_weights: channels must be 1..64%s");
  int ci = flip ? cout : cin, co = flip ? cin : cout;
  int cip = padc(ci), cop = padc(co);
  int nstep = conv_nstep(cip);
  long n = (long)cop * nstep * 32;
  hipLaunchKernelGGL(k_conv3_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), w, cin, cout, cip, cop,
                     nstep, flip ? 1 : 0, (__bf16*)wp);
  return check_launch("lex_conv3d_pack_weights");
}

}  // extern "C"
namespace {
template <int CI, int CO>
void launch_fwd(int nx, int ny, int nz, const void* x, const void* wp, const float* bias, int mode, int cout_real,
                void* y, float* yf, float* pre, hipStream_t st) {
  if (CI <= 32 && CO == 32) {   // persistent, weights in registers
    constexpr int CIP = CI <= 32 ? CI : 32;
#define LEX_FWD32P(MD)                                                                                  \
  if ((mode & 15) == MD) {                                                                               \
    hipLaunchKernelGGL((k_conv3_fwd32p<CIP, MD>), dim3(conv_fwd32p_grid<CIP, MD>(nx, ny, nz)), dim3(256), 0, st, \
                       nx, ny, nz, (const __bf16*)x, (const __bf16*)wp, bias, cout_real, (__bf16*)y, yf, pre); \
    return;                                                                                             \
  }
    LEX_FWD32P(0) LEX_FWD32P(1) LEX_FWD32P(2) LEX_FWD32P(3) LEX_FWD32P(4) LEX_FWD32P(5) LEX_FWD32P(6) LEX_FWD32P(7)
    LEX_FWD32P(8)
#undef LEX_FWD32P
  }
  if (CO % 32 == 0) {   // 32 x 32 MFMA tiles: one A fragment per 32 output channels
    dim3 grid((nz + CT2Z - 1) / CT2Z, (ny + CT2Y - 1) / CT2Y, (nx + CT2X - 1) / CT2X);
    hipLaunchKernelGGL((k_conv3_fwd32<CI, (CO % 32 == 0 ? CO : 32)>), grid, dim3(256), 0, st, nx, ny, nz,
                       (const __bf16*)x, (const __bf16*)wp, bias, mode, cout_real, (__bf16*)y, yf, pre);
    return;
  }
  if (CI <= 32 && CO == 16) {   // persistent, weights in registers
    constexpr int CIP = CI <= 32 ? CI : 32;
#define LEX_FWDP(MD)                                                                                   \
  if ((mode & 7) == MD) {                                                                              \
    hipLaunchKernelGGL((k_conv3_fwdp<CIP, MD>), dim3(conv_fwdp_grid<CIP, MD>(nx, ny, nz)), dim3(256), 0, st, \
                       nx, ny, nz, (const __bf16*)x, (const __bf16*)wp, bias, cout_real, (__bf16*)y, yf, pre); \
    return;                                                                                            \
  }
    LEX_FWDP(0) LEX_FWDP(1) LEX_FWDP(2) LEX_FWDP(3) LEX_FWDP(4) LEX_FWDP(5) LEX_FWDP(6) LEX_FWDP(7)
#undef LEX_FWDP
  }
  dim3 grid((nz + CTZ - 1) / CTZ, (ny + CTY - 1) / CTY, (nx + CTX - 1) / CTX);
  hipLaunchKernelGGL((k_conv3_fwd<CI, CO>), grid, dim3(256), 0, st, nx, ny, nz, (const __bf16*)x, (const __bf16*)wp,
                     bias, mode, cout_real, (__bf16*)y, yf, pre);
}
// workgroups of k_conv3_wgrad: one per CU (the accumulators fill the register file)
inline unsigned wgrad_blocks(int nx, int ny, int nz) {
  long nt = (long)((nx + CTX - 1) / CTX) * ((ny + CTY - 1) / CTY) * ((nz + CTZ - 1) / CTZ);
  return (unsigned)(nt < 256 ? nt : 256);
}
template <int CI, int CO>
// part: wgrad_blocks() x (27 CI CO + CO) floats; one k_partial_sum adds the weight and
// (db not null) the bias partial sums
void launch_wgrad(int nx, int ny, int nz, const void* x, const void* g, float* dw, float* part, float* db,
                  int cout, hipStream_t st) {
  const unsigned nb = wgrad_blocks(nx, ny, nz);
  hipLaunchKernelGGL((k_conv3_wgrad<CI, CO>), dim3(nb), dim3(256), 0, st, nx, ny, nz, (const __bf16*)x,
                     (const __bf16*)g, part, db ? 1 : 0);
  const int n = 27 * CI * CO, n2 = db ? cout : 0;
  hipLaunchKernelGGL(k_partial_sum, dim3((n + n2 + 63) / 64), dim3(64), 0, st, (int)nb, n, n + CO,
                     (const float*)part, dw, n2, db);
}
}  // namespace
extern "C" {

static int conv_fwd(const char* name, int nx, int ny, int nz, int cin, int cout, const void* x, const void* wp,
                    const float* bias, int mode, void* y, float* yf, float* pre, hipStream_t st) {
  if (!conv_ok(nx, ny, nz, cin, cout, 64)) return fail(-1, "%s: bad sizes (channels 1..64)", name);
  if ((mode & 2) ? !yf : !y) return fail(-2, "%s: missing output", name);
  if ((mode & 4) && (nx < 2 || ny < 2)) return fail(-1, "%s: periodic padding needs nx, ny >= 2", name);
  const int cip = padc(cin), cop = padc(cout);
  // the GELU-gradient epilogue (mode 8) exists in the persistent 32-output-channel kernel only
  if ((mode & 8) && (mode != 8 || cip > 32 || cop != 32 || !pre))
    return fail(-1, "%s: the fused GELU gradient needs <= 32 input and 17..32 output channels", name);
#define LEX_FWD(CI, CO) \
  if (cip == CI && cop == CO) { launch_fwd<CI, CO>(nx, ny, nz, x, wp, bias, mode, cout, y, yf, pre, st); return check_launch(name); }
  LEX_FWD(16, 16) LEX_FWD(16, 32) LEX_FWD(16, 64)
  LEX_FWD(32, 16) LEX_FWD(32, 32) LEX_FWD(32, 64)
  LEX_FWD(64, 16) LEX_FWD(64, 32) LEX_FWD(64, 64)
#undef LEX_FWD
  return fail(-1, "%s: unsupported channel padding", name);
}

int lex_conv3d_bf16_fwd(int nx, int ny, int nz, int cin, int cout, const void* x, const void* wp, const float* bias,
                        int gelu, int out_fp32, void* y, float* yf, float* pre, void* stream) {
  REQ(x); REQ(wp); REQ(bias);
  const int mode = (gelu ? 1 : 0) | (out_fp32 ? 2 : 0);
  return conv_fwd("lex_conv3d_bf16_fwd", nx, ny, nz, cin, cout, x, wp, bias, mode, y, yf, pre, S(stream));
}

int lex_conv3d_bf16_fwd_circular(int nx, int ny, int nz, int cin, int cout, const void* x, const void* wp,
                                 const float* bias, int gelu, int out_fp32, void* y, float* yf, float* pre,
                                 void* stream) {
  REQ(x); REQ(wp); REQ(bias);
  const int mode = (gelu ? 1 : 0) | (out_fp32 ? 2 : 0) | 4;
  return conv_fwd("lex_conv3d_bf16_fwd_circular", nx, ny, nz, cin, cout, x, wp, bias, mode, y, yf, pre, S(stream));
}

int lex_conv3d_bf16_bwd_data(int nx, int ny, int nz, int cin, int cout, const void* g_pre, const void* wp_flip,
                             const float* zero_bias, float* dx, void* stream) {
  REQ(g_pre); REQ(wp_flip); REQ(zero_bias); REQ(dx);
  if (!conv_ok(nx, ny, nz, cin, cout, 64)) return fail(-1, "lex_conv3d_bf16_bwd_data: bad sizes (channels 1..64)%s");
  // transposed convolution: input channels = cout, output channels = cin
  return lex_conv3d_bf16_fwd(nx, ny, nz, cout, cin, g_pre, wp_flip, zero_bias, 0, 1, nullptr, dx, nullptr, stream);
}

int lex_conv3d_bf16_bwd_data_gelu(int nx, int ny, int nz, int cin, int cout, const void* g_pre, const void* wp_flip,
                                  const float* zero_bias, const float* pre_in, void* g_pre_in, void* stream) {
  REQ(g_pre); REQ(wp_flip); REQ(zero_bias); REQ(pre_in); REQ(g_pre_in);
  if (!conv_ok(nx, ny, nz, cin, cout, 64))
    return fail(-1, "lex_conv3d_bf16_bwd_data_gelu: bad sizes (channels 1..64)%s");
  // transposed convolution (input channels = cout, output channels = cin), its epilogue
  // multiplying by gelu'(pre_in) and writing BF16; pre_in is read, never written
  return conv_fwd("lex_conv3d_bf16_bwd_data_gelu", nx, ny, nz, cout, cin, g_pre, wp_flip, zero_bias, 8, g_pre_in,
                  nullptr, const_cast<float*>(pre_in), S(stream));
}

int lex_conv3d_bf16_bwd_weight(int nx, int ny, int nz, int cin, int cout, const void* x, const void* g_pre, float* dw,
                               float* db, void* stream) {
  REQ(x); REQ(g_pre); REQ(dw);
  if (!conv_ok(nx, ny, nz, cin, cout)) return fail(-1, "lex_conv3d_bf16_bwd_weight: bad sizes (channels 1..32)%s");
  int cip = padc(cin), cop = padc(cout);
  hipStream_t st = S(stream);
  // per-workgroup partial sums, added in workgroup order (bit-reproducible gradients):
  // stream-ordered scratch, returned on the stream
  // (the bias gradient's partial sums come from the same workgroups, after the weights')
  const size_t nb = wgrad_blocks(nx, ny, nz), npart = nb * (27 * cip * cop + cop);
  SlotScratch sc;
  if (sc.take(npart * sizeof(float), st) != hipSuccess)
    return fail(-4, "%s: no device memory for the partial sums", "lex_conv3d_bf16_bwd_weight");
  if (cip == 16 && cop == 16) launch_wgrad<16, 16>(nx, ny, nz, x, g_pre, dw, sc.p, db, cout, st);
  else if (cip == 16 && cop == 32) launch_wgrad<16, 32>(nx, ny, nz, x, g_pre, dw, sc.p, db, cout, st);
  else if (cip == 32 && cop == 16) launch_wgrad<32, 16>(nx, ny, nz, x, g_pre, dw, sc.p, db, cout, st);
  else launch_wgrad<32, 32>(nx, ny, nz, x, g_pre, dw, sc.p, db, cout, st);
  return check_launch("lex_conv3d_bf16_bwd_weight");
}

int lex_conv3d_f32(const float* x, int nix, int niy, int niz, int cin, const float* w, const float* bias, int k,
                   int stride, int lhs_dilation, int pad_lo_x, int pad_lo_y, int pad_lo_z, int relu, float* y, int nox,
                   int noy, int noz, int cout, void* stream) {
  REQ(x); REQ(w); REQ(y);
  if (nix < 1 || niy < 1 || niz < 1 || nox < 1 || noy < 1 || noz < 1 || cin < 1 || cout < 1)
    return fail(-1, "lex_conv3d_f32: extents must be positive%s");
  if ((k != 1 && k != 3) || stride < 1 || lhs_dilation < 1)
    return fail(-1, "lex_conv3d_f32: k must be 1 or 3, stride and dilation >= 1%s");
  const long nout = (long)nox * noy * noz * cout;
  if (nout >= (1L << 31) || (long)nix * niy * niz * cin >= (1L << 31))
    return fail(-1, "lex_conv3d_f32: tensor too large%s");
  ConvGeom g{nix, niy, niz, cin, nox, noy, noz, cout, k, stride, lhs_dilation, pad_lo_x, pad_lo_y, pad_lo_z};
  hipLaunchKernelGGL(k_conv3_f32, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, S(stream), g, x, w, bias, relu, y);
  return check_launch("lex_conv3d_f32");
}

// the 8-channel-chunk forms of the BF16 channels-last writers: whole 16-byte chunks,
// aligned output, 32-bit chunk indices
static bool chunk8_ok(long ncell, int cp, const void* out) {
  return cp % 8 == 0 && ((uintptr_t)out & 15) == 0 && ncell * (cp / 8) < (1L << 31);
}
// every row of an FP32 array with this base and row stride starts on 16 bytes
static bool rows16(const float* base, int stride) { return ((uintptr_t)base & 15) == 0 && stride % 4 == 0; }

int lex_gelu_bwd(long ncell, int cp, int c_real, int cpre, const float* gpost, int gstride, const float* pre, int gelu,
                 void* gpre, void* stream) {
  REQ(gpost); REQ(gpre);
  if (gelu) REQ(pre);
  if (ncell <= 0 || cp < c_real || c_real < 1) return fail(-1, "lex_gelu_bwd: bad sizes%s");
  long n = ncell * cp;
  if (chunk8_ok(ncell, cp, gpre)) {
    const unsigned nch = (unsigned)(cp / 8), nchunk = (unsigned)(ncell * nch);
    const int vec = (rows16(gpost, gstride) ? 1 : 0) | (gelu && rows16(pre, cpre) ? 2 : 0);
    hipLaunchKernelGGL(k_gelu_bwd8, dim3((nchunk + 255) / 256), dim3(256), 0, S(stream), nchunk, nch, c_real, cpre,
                       gpost, gstride, pre, gelu, vec, (__bf16*)gpre);
  } else {
    hipLaunchKernelGGL(k_gelu_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), ncell, cp, c_real, cpre,
                       gpost, gstride, pre, gelu, (__bf16*)gpre);
  }
  return check_launch("lex_gelu_bwd");
}

int lex_to_bf16_channels_last(long ncell, int c_real, int cstride, int cp, const float* in, void* out, void* stream) {
  REQ(in); REQ(out);
  if (ncell <= 0 || cp < c_real || c_real < 1) return fail(-1, "lex_to_bf16_channels_last: bad sizes%s");
  long n = ncell * cp;
  if (chunk8_ok(ncell, cp, out)) {
    const unsigned nch = (unsigned)(cp / 8), nchunk = (unsigned)(ncell * nch);
    hipLaunchKernelGGL(k_to_bf16_cl8, dim3((nchunk + 255) / 256), dim3(256), 0, S(stream), nchunk, nch, c_real,
                       cstride, rows16(in, cstride) ? 1 : 0, in, (__bf16*)out);
  } else {
    hipLaunchKernelGGL(k_to_bf16_cl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), ncell, c_real,
                       cstride, cp, in, (__bf16*)out);
  }
  return check_launch("lex_to_bf16_channels_last");
}

// ------------------------------------------------------------------ learned-SGS data paths
int lex_sgs_features(const lex_grid_t* g, const float* theta0, const float* qv0, const float* u, const float* v,
                     const float* w, const float* th, const float* qv, const float* shift, const float* scale,
                     float* out, void* stream) {
  REQGRID(g);
  REQ(theta0); REQ(qv0); REQ(u); REQ(v); REQ(w); REQ(th); REQ(qv); REQ(shift); REQ(scale); REQ(out);
  Dims d = mkdims(g);
  FeatNorm nrm;
  for (int c = 0; c < 11; ++c) {
    nrm.shift[c] = shift[c];
    nrm.scale[c] = scale[c];
  }
  hipLaunchKernelGGL(k_sgs_features, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, theta0, qv0, u, v, w, th, qv, nrm,
                     out);
  return check_launch("lex_sgs_features");
}

// ------------------------------------------------- post-step deep-learning correction
int lex_dl_pack(const lex_grid_t* g, const float* th, const float* u, const float* v, const float* w, const float* qv,
                const float* lo, const float* den, float* out, void* stream) {
  REQGRID(g);
  REQ(th); REQ(u); REQ(v); REQ(w); REQ(qv); REQ(lo); REQ(den); REQ(out);
  Dims d = mkdims(g);
  DlNorm nrm;
  nrm.lo = lo;
  nrm.den = den;
  hipLaunchKernelGGL(k_dl_pack, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, th, u, v, w, qv, nrm, out);
  return check_launch("lex_dl_pack");
}

int lex_dl_stagger(const lex_grid_t* g, const float* c, float* tc, float* uc, float* vc, float* wc, float* qc,
                   void* stream) {
  REQGRID(g);
  REQ(c); REQ(tc); REQ(uc); REQ(vc); REQ(wc); REQ(qc);
  Dims d = mkdims(g);
  if (d.nx < 2 || d.ny < 2 || d.nz < 2) return fail(-1, "lex_dl_stagger: nx, ny, nz must be >= 2%s");
  hipLaunchKernelGGL(k_dl_stagger, grd(d.nz + 1, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, c, tc, uc, vc, wc, qc);
  return check_launch("lex_dl_stagger");
}

int lex_dl_apply(const lex_grid_t* g, const float* c, const float* r_th, const float* r_u, const float* r_v,
                 const float* r_w, const float* r_qv, float* th, float* u, float* v, float* w, float* qv,
                 void* stream) {
  REQGRID(g);
  REQ(c); REQ(r_th); REQ(r_u); REQ(r_v); REQ(r_w); REQ(r_qv); REQ(th); REQ(u); REQ(v); REQ(w); REQ(qv);
  Dims d = mkdims(g);
  if (d.nx < 2 || d.ny < 2 || d.nz < 2) return fail(-1, "lex_dl_apply: nx, ny, nz must be >= 2%s");
  DlRanges rg;
  rg.th = r_th;
  rg.u = r_u;
  rg.v = r_v;
  rg.w = r_w;
  rg.qv = r_qv;
  hipLaunchKernelGGL(k_dl_apply, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, c, rg, th, u, v, w, qv);
  return check_launch("lex_dl_apply");
}

int lex_laplace(const lex_grid_t* g, long nfields, const float* s, float* out, void* stream) {
  REQGRID(g);
  REQ(s); REQ(out);
  if (nfields < 0) return fail(-1, "lex_laplace: nfields must be >= 0%s");
  const long per = (long)g->nx * g->ny * g->nz;
  const long chunk = 65535 / g->nx;   // fields per launch (grid z = fields x nx)
  for (long f0 = 0; f0 < nfields; f0 += chunk) {
    const long nf = (nfields - f0 < chunk) ? nfields - f0 : chunk;
    hipLaunchKernelGGL(k_laplace, dim3((g->nz + 63) / 64, g->ny, (unsigned)(nf * g->nx)), dim3(64), 0, S(stream),
                       g->nx, g->ny, g->nz, g->dx, g->dy, g->dz, s + f0 * per, out + f0 * per);
  }
  return check_launch("lex_laplace");
}

int lex_weighted_sum_grad(long n, int op, const float* a, const float* b, const float* w, long period, float* out,
                          void* stream) {
  REQ(a); REQ(w); REQ(out);
  if (op != 0 && op != 1) return fail(-1, "lex_weighted_sum_grad: op must be 0 or 1%s");
  if (op == 0 && !b) return fail(-2, "lex_weighted_sum_grad: op 0 needs b%s");
  if (n < 0 || period < 1) return fail(-1, "lex_weighted_sum_grad: n >= 0 and period >= 1 required%s");
  if (n == 0) return 0;
  const long blocks = (period + 255) / 256 < 4096 ? (period + 255) / 256 : 4096;
  hipLaunchKernelGGL(k_wgrad, dim3((unsigned)blocks), dim3(256), 0, S(stream), n, op, a, b, w, period, out);
  return check_launch("lex_weighted_sum_grad");
}

int lex_dl_apply_vjp(const lex_grid_t* g, const float* l_th, const float* l_u, const float* l_v, const float* l_w,
                     const float* l_qv, const float* r_th, const float* r_u, const float* r_v, const float* r_w,
                     const float* r_qv, float* dc, void* stream) {
  REQGRID(g);
  REQ(l_th); REQ(l_u); REQ(l_v); REQ(l_w); REQ(l_qv); REQ(r_th); REQ(r_u); REQ(r_v); REQ(r_w); REQ(r_qv); REQ(dc);
  Dims d = mkdims(g);
  DlRanges rg;
  rg.th = r_th;
  rg.u = r_u;
  rg.v = r_v;
  rg.w = r_w;
  rg.qv = r_qv;
  hipLaunchKernelGGL(k_dl_apply_vjp, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, l_th, l_u, l_v, l_w, l_qv, rg, dc);
  return check_launch("lex_dl_apply_vjp");
}

int lex_dl_pack_vjp(const lex_grid_t* g, const float* dp, const float* lo, const float* den, float* l_th, float* l_u,
                    float* l_v, float* l_w, float* l_qv, void* stream) {
  REQGRID(g);
  REQ(dp); REQ(den); REQ(l_th); REQ(l_u); REQ(l_v); REQ(l_w); REQ(l_qv);
  Dims d = mkdims(g);
  DlNorm nrm{lo, den};
  hipLaunchKernelGGL(k_dl_pack_vjp, dim3((d.nz + 1 + 63) / 64, (d.ny + 1 + 3) / 4, d.nx + 1), dim3(64, 4, 1), 0,
                     S(stream), d, dp, nrm, l_th, l_u, l_v, l_w, l_qv);
  return check_launch("lex_dl_pack_vjp");
}

size_t lex_weighted_sum_workspace_bytes(void) { return (size_t)kWsumBlocks * sizeof(double); }

int lex_weighted_sum(long n, int op, const float* a, const float* b, const float* w, long period, double* out,
                     void* workspace, size_t workspace_bytes, void* stream) {
  REQ(a); REQ(w); REQ(out); REQ(workspace);
  if (op != 0 && op != 1) return fail(-1, "lex_weighted_sum: op must be 0 or 1%s");
  if (op == 0 && !b) return fail(-2, "lex_weighted_sum: op 0 needs b%s");
  if (n < 0 || period < 1) return fail(-1, "lex_weighted_sum: n >= 0 and period >= 1 required%s");
  if (workspace_bytes < (size_t)kWsumBlocks * sizeof(double))
    return fail(-3, "lex_weighted_sum: workspace too small%s");
  double* part = (double*)workspace;
  hipLaunchKernelGGL(k_wsum_partial, dim3(kWsumBlocks), dim3(kWsumThreads), 0, S(stream), n, op, a, b, w, period, part);
  hipLaunchKernelGGL(k_wsum_final, dim3(1), dim3(64), 0, S(stream), (const double*)part, out);
  return check_launch("lex_weighted_sum");
}

int lex_atm_ocn_flux(long n, const float* z, const float* ub, const float* vb, const float* th, const float* q,
                     const float* rho, const float* ts, float* const* out, void* stream) {
  REQ(z); REQ(ub); REQ(vb); REQ(th); REQ(q); REQ(rho); REQ(ts); REQ(out);
  if (n < 0) return fail(-1, "lex_atm_ocn_flux: n must be >= 0%s");
  for (int c = 0; c < 7; ++c)
    if (!out[c]) return fail(-2, "lex_atm_ocn_flux: null output pointer%s");
  if (n == 0) return 0;
  SfcOut o{out[0], out[1], out[2], out[3], out[4], out[5], out[6], nullptr};
  hipLaunchKernelGGL(k_atm_ocn_flux_cols, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), n, z, ub, vb,
                     th, q, rho, ts, o);
  return check_launch("lex_atm_ocn_flux");
}

int lex_surface_fluxes(const lex_grid_t* g, const float* rho0, const float* th, const float* u, const float* v,
                       const float* qv, const float* ts, float* const* out, void* stream) {
  REQGRID(g);
  REQ(rho0); REQ(th); REQ(u); REQ(v); REQ(qv); REQ(ts); REQ(out);
  for (int c = 0; c < 7; ++c)
    if (!out[c]) return fail(-2, "lex_surface_fluxes: null output pointer%s");
  Dims d = mkdims(g);
  SfcOut o{out[0], out[1], out[2], out[3], out[4], out[5], out[6], nullptr};
  launch_atm_ocn_flux(d, rho0, th, u, v, qv, ts, o, S(stream));
  return check_launch("lex_surface_fluxes");
}

int lex_sgs_apply(const lex_grid_t* g, const float* y, const float* out_scale, float* su, float* sv, float* sw,
                  float* sth, float* sqv, void* stream) {
  REQGRID(g);
  REQ(y); REQ(out_scale); REQ(su); REQ(sv); REQ(sw); REQ(sth); REQ(sqv);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_sgs_apply, grd(d.nz + 1, d.ny + 1, d.nx + 1), blk(), 0, S(stream), d, y, out_scale[0],
                     out_scale[1], out_scale[2], out_scale[3], out_scale[4], su, sv, sw, sth, sqv);
  return check_launch("lex_sgs_apply");
}

int lex_sgs_apply_vjp(const lex_grid_t* g, const float* lsu, const float* lsv, const float* lsw, const float* lsth,
                      const float* lsqv, const float* out_scale, float* ly, void* stream) {
  REQGRID(g);
  REQ(lsu); REQ(lsv); REQ(lsw); REQ(lsth); REQ(lsqv); REQ(out_scale); REQ(ly);
  Dims d = mkdims(g);
  hipLaunchKernelGGL(k_sgs_apply_vjp, grd(d.nz, d.ny, d.nx), blk(), 0, S(stream), d, lsu, lsv, lsw, lsth, lsqv,
                     out_scale[0], out_scale[1], out_scale[2], out_scale[3], out_scale[4], ly);
  return check_launch("lex_sgs_apply_vjp");
}

int lex_sgs_features_vjp(const lex_grid_t* g, const float* lfeat, const float* scale, float* lu, float* lv, float* lw,
                         float* lth, float* lqv, void* stream) {
  REQGRID(g);
  REQ(lfeat); REQ(scale); REQ(lu); REQ(lv); REQ(lw); REQ(lth); REQ(lqv);
  Dims d = mkdims(g);
  FeatNorm nrm;
  for (int c = 0; c < 11; ++c) {
    nrm.shift[c] = 0.f;
    nrm.scale[c] = scale[c];
  }
  const SgsFeatAdj A{d, lfeat, nrm};
  hipStream_t st = S(stream);
  hipLaunchKernelGGL(k_sgs_features_vjp_u, grd(d.nz + 2, d.ny + 6, d.nx + 7), blk(), 0, st, A, lu, lth, lqv);
  hipLaunchKernelGGL(k_sgs_features_vjp_v, grd(d.nz + 2, d.ny + 7, d.nx + 6), blk(), 0, st, A, lv);
  hipLaunchKernelGGL(k_sgs_features_vjp_w, grd(d.nz + 3, d.ny + 6, d.nx + 6), blk(), 0, st, A, lw);
  return check_launch("lex_sgs_features_vjp");
}

int lex_coarsen(const lex_grid_t* g, int mode, int fx, int fy, int fz, const float* a, float* out, void* stream) {
  REQGRID(g);
  REQ(a); REQ(out);
  if (mode < 0 || mode > 3 || fx < 1 || fy < 1 || fz < 1) return fail(-1, "lex_coarsen: bad mode or factors%s");
  if (g->nx % fx || g->ny % fy || g->nz % fz) return fail(-1, "lex_coarsen: the grid must divide by the factors%s");
  Dims d = mkdims(g);
  int ni = d.nx / fx + (mode == 1), nj = d.ny / fy + (mode == 2), nk = d.nz / fz + (mode == 3);
  hipLaunchKernelGGL(k_coarsen, grd(nk, nj, ni), blk(), 0, S(stream), d, mode, fx, fy, fz, a, out);
  return check_launch("lex_coarsen");
}

int lex_rk3_step_stage_sgs(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                           float* const* state, lex_stage_sgs_fn fn, void* user, void* ws, size_t ws_bytes,
                           void* stream) {
  REQ(o); REQ(fn);
  if (o->cor < 0 || o->cor > 2) return fail(-1, "lex_rk3_step_stage_sgs: cor must be 0, 1 or 2%s");
  lex_step_opts_t o2 = *o;
  o2.turb = kTurbExternal;
  return rk3_step_entry(g, &o2, b, state, ws, ws_bytes, stream, (StageSgsFn)fn, user);
}

int lex_rk3_step_ext_sgs(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                         const float* su, const float* sv, const float* sw, const float* const* ss, void* ws,
                         size_t ws_bytes, void* stream) {
  REQGRID(g);
  REQ(o); REQ(b); REQ(state); REQ(ws); REQ(su); REQ(sv); REQ(sw); REQ(ss);
  if (o->n_passive < 0 || o->n_passive > 8) return fail(-1, "lex_rk3_step_ext_sgs: n_passive must be 0..8%s");
  Dims d = mkdims(g);
  Rk3Layout L = rk3_layout(d, o->n_passive);
  if (ws_bytes < L.total) return fail(-3, "lex_rk3_step_ext_sgs: workspace too small%s");
  hipStream_t st = S(stream);
  char* w = (char*)ws;
  auto cp = [&](const float* src, size_t off, size_t n) {
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (long)n, src, (float*)(w + off));
  };
  cp(su, L.off_su, L.iu_);
  cp(sv, L.off_sv, L.iv_);
  cp(sw, L.off_sw, L.iw_);
  for (int q = 0; q < 2 + o->n_passive; ++q) {
    if (!ss[q]) return fail(-2, "lex_rk3_step_ext_sgs: null scalar SGS pointer%s");
    cp(ss[q], L.off_ss + q * al256(L.ic_ * 4), L.ic_);
  }
  lex_step_opts_t o2 = *o;
  o2.turb = kTurbExternal;
  return rk3_step_entry(g, &o2, b, state, ws, ws_bytes, stream);
}

}  // extern "C"
