// P3/P4: pseudo-incompressible pressure equation (pressure_equations.py) and the
// unpreconditioned BiCGSTAB of jax.scipy.sparse.linalg.bicgstab that the reference
// calls (one_step_integration.py:468-487).  Vectors are FP32 (nx, ny, nz); every
// reduction (dot product) and the scalar recurrence run in FP64.
//
// One iteration = 3 kernels, each a grid-stride pass whose dot products leave FP64
// per-block partials that the following kernels sum in a fixed order (deterministic).
// All kernels check a device-side `done` flag, so the host can enqueue iterations ahead
// and poll the flag only every few iterations.
#pragma once
#include "lex_common.hpp"

namespace lex {

// --------------------------- rtt, RHS, matvec ------------------------------------
// calculate_rtt (pressure_equations.py:175-195)
__global__ void k_rtt(Dims d, const float* __restrict__ rtc, const float* __restrict__ th,
                      const float* __restrict__ qv, const float* __restrict__ buoy, float* __restrict__ rtt) {
  int K = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nz+1
  int J = blockIdx.y * blockDim.y + threadIdx.y;  // padded
  int I = blockIdx.z;
  if (K >= d.nz + 2 || J >= d.ny + 6) return;
  long c = ic(d, I, J, K);
  float val;
  if (K >= 1 && K <= d.nz) {
    val = rtc[K] * th[c] * (1.0f + (float)kReps * qv[c]) / (1.0f + qv[c]);
  } else {
    int i = (I - 3 + d.nx) % d.nx, j = (J - 3 + d.ny) % d.ny;
    const float* bcol = buoy + ((long)i * d.ny + j) * d.nz;
    auto rb = [&](int k) { return rtc[k + 1] * bcol[k] / (float)kCp; };
    int nz = d.nz;
    val = (K == 0) ? (1.875f * rb(0) + (-1.25f) * rb(1) + 0.375f * rb(2))
                   : (1.875f * rb(nz - 1) + (-1.25f) * rb(nz - 2) + 0.375f * rb(nz - 3));
  }
  rtt[c] = val;
}

// rhs_of_pressure_equation (pressure_equations.py:51-77) on interior (nx,ny,nz) arrays
__global__ void k_p_rhs(Dims d, const float* __restrict__ rtc, const float* __restrict__ au,
                        const float* __restrict__ av, const float* __restrict__ aw, const float* __restrict__ fu,
                        const float* __restrict__ fv, const float* __restrict__ buoy, float* __restrict__ rhs) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  float rt = rtc[k + 1];
  auto U = [&](const float* a, int f) { return a[((long)f * ny + j) * nz + k]; };
  auto V = [&](const float* a, int f) { return a[((long)i * (ny + 1) + f) * nz + k]; };
  auto r8w = [&](int l) -> float { return (l <= 0 || l >= nz) ? 0.0f : 0.5f * (rtc[l] + rtc[l + 1]); };
  auto W = [&](int l) { return aw[((long)i * ny + j) * (nz + 1) + l]; };
  float adv = (rt * U(au, i + 1) - rt * U(au, i)) * d.idx + (rt * V(av, j + 1) - rt * V(av, j)) * d.idy +
              (r8w(k + 1) * W(k + 1) - r8w(k) * W(k)) * d.idz;
  float cor = 0.0f;
  if (fu && fv) cor = (rt * U(fv, i + 1) - rt * U(fv, i)) * d.idx + (rt * (-V(fu, j + 1)) - rt * (-V(fu, j))) * d.idy;
  const float* bcol = buoy + ((long)i * ny + j) * nz;
  auto rb = [&](int kk) { return rtc[kk + 1] * bcol[kk]; };
  auto rbw = [&](int l) -> float {
    if (l == 0) return 1.875f * rb(0) + (-1.25f) * rb(1) + 0.375f * rb(2);
    if (l == nz) return 1.875f * rb(nz - 1) + (-1.25f) * rb(nz - 2) + 0.375f * rb(nz - 3);
    return 0.5f * (rb(l - 1) + rb(l));
  };
  float bz = (rbw(k + 1) - rbw(k)) * d.idz;
  rhs[((long)i * ny + j) * nz + k] = adv + cor + bz;
  (void)nx;
}

// affine matvec laplace_of_pressure (pressure_equations.py:8-48) at interior (i,j,k)
__device__ __forceinline__ float matvec_point(const Dims& d, const float* __restrict__ rtt,
                                              const float* __restrict__ x, int i, int j, int k) {
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  auto X = [&](int a, int b, int c) {
    a = (a + nx) % nx;
    b = (b + ny) % ny;
    c = c < 0 ? 0 : (c >= nz ? nz - 1 : c);
    return x[((long)a * ny + b) * nz + c];
  };
  long cc = ic(d, i + 3, j + 3, k + 1);
  const long sx = d.sxc(), sy = d.syc();
  float xc = X(i, j, k);
  float px0 = 0.5f * (rtt[cc] + rtt[cc - sx]) * ((xc - X(i - 1, j, k)) * d.idx);
  float px1 = 0.5f * (rtt[cc + sx] + rtt[cc]) * ((X(i + 1, j, k) - xc) * d.idx);
  float py0 = 0.5f * (rtt[cc] + rtt[cc - sy]) * ((xc - X(i, j - 1, k)) * d.idy);
  float py1 = 0.5f * (rtt[cc + sy] + rtt[cc]) * ((X(i, j + 1, k) - xc) * d.idy);
  float pz0 = (k == 0) ? rtt[cc - 1] : 0.5f * (rtt[cc] + rtt[cc - 1]) * ((xc - X(i, j, k - 1)) * d.idz);
  float pz1 = (k == nz - 1) ? rtt[cc + 1] : 0.5f * (rtt[cc + 1] + rtt[cc]) * ((X(i, j, k + 1) - xc) * d.idz);
  return (float)kCp * ((px1 - px0) * d.idx + (py1 - py0) * d.idy + (pz1 - pz0) * d.idz);
}

__global__ void k_p_matvec(Dims d, const float* __restrict__ rtt, const float* __restrict__ x, float* __restrict__ out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  out[((long)i * d.ny + j) * d.nz + k] = matvec_point(d, rtt, x, i, j, k);
}

// ------------------------------- BiCGSTAB ----------------------------------------
// Every dot product is reduced in two phases across a kernel boundary: a kernel writes
// one FP64 partial sum per block, and each block of the next kernel sums all partials
// itself, in the same fixed order, so every block derives bit-identical scalars.  No
// kernel needs a grid-wide fence or a last-block ticket (on a multi-XCD part a
// device-scope release writes back the XCD's L2, which made those tails dominate).
//
// Scalars entering iteration `it` (rho, alpha, omega, k) are double-buffered by parity:
// iteration it reads s[it & 1] and its last kernel's block 0 writes s[(it + 1) & 1].
// The loop test (JAX: |r|^2 > atol^2 and 0 <= k < maxiter) runs at the top of each
// iteration; the first kernel that finds it false sets the sticky `done` and `k_final`.
struct BicgScalars {
  double rho, alpha, omega;
  int k, pad;
};
struct BicgState {
  BicgScalars s[2];
  int done;      // sticky: the loop has ended
  int k_final;   // exit code (iterations, or -10 / -11 on breakdown)
  int maxiter;
  int pad;
};
static_assert(sizeof(BicgState) <= 256, "BiCGSTAB loop state fits its 256-byte slot");

constexpr int kBicgThreads = 256;
constexpr int kBicgBlocks = 1024;   // upper bound on blocks (partials per sum)
constexpr int kBicgCellsPerThread = 2;   // grid size (1-16 measured within 15 % on 64x64x40)
// partial-sum slots (each kBicgBlocks doubles): b.b; (r.r, rhat.r) x 2 parities;
// rhat.q; (t.s, t.t, s.s)
constexpr int kBicgSlots = 1 + 4 + 1 + 3;
constexpr size_t kBicgPartialBytes = (size_t)kBicgSlots * kBicgBlocks * sizeof(double);
struct BicgPartials {
  double* bb;      // [nb]
  double* rr[2];   // [2][nb] by parity: r.r, then rhat.r
  double* rq;      // [nb]
  double* ts;      // [3][nb]: t.s, t.t, s.s
};
__host__ __device__ inline BicgPartials bicg_partials(double* base) {
  BicgPartials P;
  P.bb = base;
  P.rr[0] = base + kBicgBlocks;
  P.rr[1] = base + 3 * kBicgBlocks;
  P.rq = base + 5 * kBicgBlocks;
  P.ts = base + 6 * kBicgBlocks;
  return P;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block-level sum of NV per-thread values; the result is returned to every thread.
template <int NV>
__device__ __forceinline__ void block_sum_all(const double (&v)[NV], double (&out)[NV]) {
  __shared__ double sh[NV][kBicgThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double s = wave_sum(v[q]);
    if (lane == 0) sh[q][wv] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double t = 0.0;
    for (int w = 0; w < kBicgThreads / 64; ++w) t += sh[q][w];
    out[q] = t;
  }
  __syncthreads();
}

// Sum of NV partial arrays (stride kBicgBlocks, nb entries each) in a fixed order,
// returned to every thread of the block.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ P, int nb, double (&tot)[NV]) {
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBicgThreads) s += P[q * kBicgBlocks + b];
    v[q] = s;
  }
  block_sum_all<NV>(v, tot);
}

// Write this block's sums of NV per-thread values to P[q * kBicgBlocks + blockIdx.x].
template <int NV>
__device__ __forceinline__ void put_partials(const double (&acc)[NV], double* __restrict__ P) {
  double tot[NV];
  block_sum_all<NV>(acc, tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) P[q * kBicgBlocks + blockIdx.x] = tot[q];
  }
}

// Interior cell (i, j, k) of the unpadded linear index t (x slowest, z fastest).
__device__ __forceinline__ void cell_of(const Dims& d, long t, int& i, int& j, int& k) {
  const long jt = t / d.nz;
  k = (int)(t - jt * d.nz);
  i = (int)(jt / d.ny);
  j = (int)(jt - (long)i * d.ny);
}

// laplace_of_pressure (pressure_equations.py:8-48) at (i, j, k) of a vector given by
// F(offset) on the unpadded (nx, ny, nz) layout: periodic in x and y, z clamped to the
// end levels; arithmetic in the order of matvec_point.
template <class F>
__device__ __forceinline__ float matvec_at(const Dims& d, const float* __restrict__ rtt, int i, int j, int k, long t,
                                           const F& X, float& xc) {
  const long syz = (long)d.ny * d.nz;
  const long oxm = (i == 0) ? t + (long)(d.nx - 1) * syz : t - syz;
  const long oxp = (i == d.nx - 1) ? t - (long)(d.nx - 1) * syz : t + syz;
  const long oym = (j == 0) ? t + (long)(d.ny - 1) * d.nz : t - d.nz;
  const long oyp = (j == d.ny - 1) ? t - (long)(d.ny - 1) * d.nz : t + d.nz;
  const long ozm = (k == 0) ? t : t - 1;
  const long ozp = (k == d.nz - 1) ? t : t + 1;
  const long cc = ic(d, i + 3, j + 3, k + 1);
  const long sx = d.sxc(), sy = d.syc();
  xc = X(t);
  float px0 = 0.5f * (rtt[cc] + rtt[cc - sx]) * ((xc - X(oxm)) * d.idx);
  float px1 = 0.5f * (rtt[cc + sx] + rtt[cc]) * ((X(oxp) - xc) * d.idx);
  float py0 = 0.5f * (rtt[cc] + rtt[cc - sy]) * ((xc - X(oym)) * d.idy);
  float py1 = 0.5f * (rtt[cc + sy] + rtt[cc]) * ((X(oyp) - xc) * d.idy);
  float pz0 = (k == 0) ? rtt[cc - 1] : 0.5f * (rtt[cc] + rtt[cc - 1]) * ((xc - X(ozm)) * d.idz);
  float pz1 = (k == d.nz - 1) ? rtt[cc + 1] : 0.5f * (rtt[cc + 1] + rtt[cc]) * ((X(ozp) - xc) * d.idz);
  return (float)kCp * ((px1 - px0) * d.idx + (py1 - py0) * d.idy + (pz1 - pz0) * d.idz);
}

// Scalars every kernel of iteration `it` derives from the partials: |r|^2 and rhat.r
// of the iteration's starting residual, and atol^2.
struct BicgTop {
  double rr, rho_new, atol2;
};
__device__ __forceinline__ BicgTop bicg_top(const BicgPartials& P, int nb, int it, double tol, double atol) {
  double a[2], bb[1];
  sum_partials<2>(P.rr[(it + 1) & 1], nb, a);
  sum_partials<1>(P.bb, nb, bb);
  BicgTop t;
  t.rr = a[0];
  t.rho_new = a[1];
  t.atol2 = fmax(tol * tol * bb[0], atol * atol);
  return t;
}

// clear the loop state (a kernel, not a memset, so the solve can be graph-captured)
__global__ void k_bicg_reset(BicgState* st, int maxiter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int q = 0; q < 2; ++q) {
      st->s[q].rho = st->s[q].alpha = st->s[q].omega = 1.0;
      st->s[q].k = 0;
    }
    st->maxiter = maxiter;
    st->done = 0;
    st->k_final = 0;
  }
}

// init: r = b - A(x0), rhat = p = q = r, x = x0; partials b.b and (r.r, rhat.r = r.r)
// in the parity slot iteration 0 reads
__global__ void __launch_bounds__(kBicgThreads) k_bicg_init(Dims d, const float* __restrict__ rtt,
                                                           const float* __restrict__ b, const float* __restrict__ x0,
                                                           float* __restrict__ x, float* __restrict__ r,
                                                           float* __restrict__ rhat, float* __restrict__ p,
                                                           float* __restrict__ q, double* __restrict__ partials) {
  const BicgPartials P = bicg_partials(partials);
  const long n = (long)d.nx * d.ny * d.nz;
  double acc[2] = {0.0, 0.0};
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(d, t, i, j, k);
    float ri = b[t] - matvec_point(d, rtt, x0, i, j, k);
    x[t] = x0[t];
    r[t] = ri;
    rhat[t] = ri;
    p[t] = ri;
    q[t] = ri;
    acc[0] += (double)b[t] * b[t];
    acc[1] += (double)ri * ri;
  }
  double bb[1] = {acc[0]};
  double rr[2] = {acc[1], acc[1]};
  put_partials<1>(bb, P.bb);
  put_partials<2>(rr, P.rr[1]);
}

// K1 of iteration it: the loop test; beta; p' = r + beta (p - omega q) (recomputed on
// the stencil's points); q' = A(p'); partials rhat.q'
__global__ void __launch_bounds__(kBicgThreads) k_bicg_pq(Dims d, const float* __restrict__ rtt,
                                                         const float* __restrict__ r, const float* __restrict__ rhat,
                                                         const float* __restrict__ p, const float* __restrict__ q,
                                                         float* __restrict__ pn, float* __restrict__ qn,
                                                         double* __restrict__ partials, BicgState* st, int it, int nb,
                                                         double tol, double atol) {
  if (st->done) return;
  const BicgPartials P = bicg_partials(partials);
  const BicgTop top = bicg_top(P, nb, it, tol, atol);
  const BicgScalars sc = st->s[it & 1];
  if (!(top.rr > top.atol2 && sc.k < st->maxiter && sc.k >= 0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->k_final = sc.k;
    }
    return;
  }
  const double beta_d = top.rho_new / sc.rho * sc.alpha / sc.omega;
  const float beta = (float)beta_d, om = (float)sc.omega;
  const long n = (long)d.nx * d.ny * d.nz;
  double acc[1] = {0.0};
  auto Pv = [&](long o) { return r[o] + beta * (p[o] - om * q[o]); };
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(d, t, i, j, k);
    float xc;
    const float qv = matvec_at(d, rtt, i, j, k, t, Pv, xc);
    pn[t] = xc;
    qn[t] = qv;
    acc[0] += (double)rhat[t] * qv;
  }
  put_partials<1>(acc, P.rq);
}

// K2: alpha = rho_new / rhat.q; s = r - alpha q (recomputed on the stencil), t = A(s);
// partials t.s, t.t, s.s
__global__ void __launch_bounds__(kBicgThreads) k_bicg_st(Dims d, const float* __restrict__ rtt,
                                                         const float* __restrict__ r, const float* __restrict__ q,
                                                         float* __restrict__ s, float* __restrict__ tv,
                                                         double* __restrict__ partials, BicgState* st, int it, int nb,
                                                         double tol, double atol) {
  if (st->done) return;
  const BicgPartials P = bicg_partials(partials);
  const BicgTop top = bicg_top(P, nb, it, tol, atol);
  double rq[1];
  sum_partials<1>(P.rq, nb, rq);
  const float al = (float)(top.rho_new / rq[0]);
  const long n = (long)d.nx * d.ny * d.nz;
  double acc[3] = {0.0, 0.0, 0.0};
  auto Sv = [&](long o) { return r[o] - al * q[o]; };
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(d, t, i, j, k);
    float xc;
    const float tt = matvec_at(d, rtt, i, j, k, t, Sv, xc);
    s[t] = xc;
    tv[t] = tt;
    acc[0] += (double)tt * xc;
    acc[1] += (double)tt * tt;
    acc[2] += (double)xc * xc;
  }
  put_partials<3>(acc, P.ts);
}

// K3: omega = t.s / t.t; x, r updates (early exit when |s|^2 < atol^2); partials r.r,
// rhat.r for the next iteration; block 0 advances the recurrence scalars
__global__ void __launch_bounds__(kBicgThreads) k_bicg_xr(long n, const float* __restrict__ p,
                                                         const float* __restrict__ s, const float* __restrict__ tv,
                                                         const float* __restrict__ rhat, float* __restrict__ x,
                                                         float* __restrict__ r, double* __restrict__ partials,
                                                         BicgState* st, int it, int nb, double tol, double atol) {
  if (st->done) return;
  const BicgPartials P = bicg_partials(partials);
  const BicgTop top = bicg_top(P, nb, it, tol, atol);
  double rq[1], ts[3];
  sum_partials<1>(P.rq, nb, rq);
  sum_partials<3>(P.ts, nb, ts);
  const double alpha_new = top.rho_new / rq[0];
  const double omega_new = ts[0] / ts[1];
  const bool early = ts[2] < top.atol2;
  const float al = (float)alpha_new, om = (float)omega_new;
  double acc[2] = {0.0, 0.0};
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    float rn;
    if (early) {
      x[t] = x[t] + al * p[t];
      rn = s[t];
    } else {
      x[t] = x[t] + (al * p[t] + om * s[t]);
      rn = s[t] - om * tv[t];
    }
    r[t] = rn;
    acc[0] += (double)rn * rn;
    acc[1] += (double)rhat[t] * rn;
  }
  put_partials<2>(acc, P.rr[it & 1]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const BicgScalars sc = st->s[it & 1];
    int kn = (omega_new == 0.0 || alpha_new == 0.0) ? -11 : sc.k + 1;
    if (top.rho_new == 0.0) kn = -10;
    BicgScalars nx;
    nx.rho = top.rho_new;
    nx.alpha = alpha_new;
    nx.omega = omega_new;
    nx.k = kn;
    nx.pad = 0;
    st->s[(it + 1) & 1] = nx;
  }
}

// after the host's loop: the exit code when no kernel ended the loop (maxiter reached)
__global__ void k_bicg_finish(BicgState* st, int iters) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && !st->done) {
    st->done = 1;
    st->k_final = st->s[iters & 1].k;
  }
}

// ------------------------------ I1: base state -----------------------------------
// Hydrostatic base state in FP64 on one column (setup_lex.py:103-137): pi0 by the
// cumulative sum from pi=1 at the surface, qv0 = RH * rslf iterated to 1e-9.
__device__ double rslf_d(double p, double t) {
  double esl = 611.2 * exp(17.67 * (t - 273.15) / (t - 29.65));
  if (esl > p * 0.5) esl = p * 0.5;
  return (kRd / kRv) * esl / (p - esl);
}

__global__ void k_base_state(int nz, double dz, double theta0v, const double* __restrict__ rh, double* __restrict__ out,
                             int max_iter, double tol, double* iters) {
  // single thread: out holds 5 columns of nz+2 (pi0, rho0_theta0, rho0, theta0, qv0)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* pi0 = out;
  double* rt0 = out + (nz + 2);
  double* rho0 = out + 2 * (nz + 2);
  double* th0 = out + 3 * (nz + 2);
  double* qv0 = out + 4 * (nz + 2);
  for (int k = 0; k < nz + 2; ++k) {
    th0[k] = theta0v;
    qv0[k] = 0.0;
  }
  auto integrate = [&]() {
    double acc = 0.0, prev = 1.0;
    for (int k = 1; k <= nz; ++k) {
      double trho = th0[k] * (1.0 + kReps * qv0[k]) / (1.0 + qv0[k]);
      acc += -kG / kCp / trho;
      double cur = acc * dz + 1.0;
      pi0[k] = 0.5 * (prev + cur);
      prev = cur;
    }
    pi0[0] = pi0[1];
    pi0[nz + 1] = pi0[nz];
  };
  auto update_qv = [&](double* dst) {
    for (int k = 0; k < nz + 2; ++k)
      dst[k] = rslf_d(pow(pi0[k], kCp / kRd) * kP00, th0[k] * pi0[k]) * rh[k];
  };
  integrate();
  update_qv(qv0);
  // prev = 1.5 qv0; iterate while max|qv0 - prev| > tol
  double diff = 0.0;
  for (int k = 0; k < nz + 2; ++k) diff = fmax(diff, fabs(qv0[k] - 1.5 * qv0[k]));
  int it = 0;
  while (diff > tol && it < max_iter) {
    // keep the previous profile in rt0 as scratch
    for (int k = 0; k < nz + 2; ++k) rt0[k] = qv0[k];
    integrate();
    update_qv(qv0);
    diff = 0.0;
    for (int k = 0; k < nz + 2; ++k) diff = fmax(diff, fabs(qv0[k] - rt0[k]));
    ++it;
  }
  for (int k = 0; k < nz + 2; ++k) {
    double rt = pow(pi0[k], kCv / kRd) * kP00 / kRd;
    rho0[k] = rt / th0[k] / ((1.0 + kReps * qv0[k]) / (1.0 + qv0[k]));
    rt0[k] = rho0[k] * th0[k];  // setup_lex.py:64
  }
  *iters = (double)it;
}

}  // namespace lex
