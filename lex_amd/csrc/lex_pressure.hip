// P3/P4: pseudo-incompressible pressure equation (pressure_equations.py) and the
// unpreconditioned BiCGSTAB of jax.scipy.sparse.linalg.bicgstab that the reference
// calls (one_step_integration.py:468-487).  Vectors are FP32 (nx, ny, nz); every
// reduction (dot product) and the scalar recurrence run in FP64.
//
// One iteration = 3 kernels, each a grid-stride pass with a deterministic two-level
// reduction (per-block FP64 partials, the last block to finish sums them in a fixed
// order and advances the scalar state).  All kernels check a device-side `done` flag,
// so the host can enqueue iterations ahead and poll the flag only every few iterations.
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
struct BicgState {
  double bs, atol2;
  double rho, alpha, omega;   // recurrence scalars (rho = previous rho_)
  double rho_new;             // rhat.r of the current residual
  double rr;                  // r.r of the current residual
  double alpha_new, omega_new, ss;
  int k;                      // iteration counter / exit code
  int done;
  int early;
  int maxiter;
  unsigned int counter[4];    // last-block tickets, one per kernel
  int pad;
};

constexpr int kBicgThreads = 256;
constexpr int kBicgBlocks = 1024;  // grid-stride; partials buffer holds kBicgBlocks*4 doubles

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// block-sum NV values per thread into lane 0 of the block; returns in `out` on thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (&out)[NV]) {
  __shared__ double sh[NV][kBicgThreads / 64];
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = wave_sum(v[q]);
    if (lane == 0) sh[q][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int t = 0; t < kBicgThreads / 64; ++t) s += sh[q][t];
      out[q] = s;
    }
  }
}

// Returns true on the one block that finished last; partial sums of this block are in
// `part` (NV values), written to partials[q*gridDim.x + blockIdx.x].
template <int NV>
__device__ bool publish_partials(double (&part)[NV], double* __restrict__ partials, unsigned int* ticket) {
  __shared__ int last;
  if (threadIdx.x == 0) {
    for (int q = 0; q < NV; ++q) partials[q * gridDim.x + blockIdx.x] = part[q];
    __threadfence();
    unsigned int t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  return last != 0;
}

template <int NV>
__device__ void final_sums(const double* __restrict__ partials, double (&tot)[NV]) {
  // executed by the whole last block; fixed summation order -> deterministic
  __shared__ double sh[NV][kBicgThreads];
  __threadfence();
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) s += ((volatile const double*)partials)[q * gridDim.x + b];
    sh[q][threadIdx.x] = s;
  }
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < NV; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
    __syncthreads();
  }
  for (int q = 0; q < NV; ++q) tot[q] = sh[q][0];
}

// zero the tickets and set the iteration cap (runs before k_bicg_init)
__global__ void k_bicg_reset(BicgState* st, int maxiter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int c = 0; c < 4; ++c) st->counter[c] = 0;
    st->maxiter = maxiter;
    st->done = 0;
    st->k = 0;
  }
}

// init: r = b - A(x0), rhat = p = q = r, x = x0; sums: b.b, r.r (rhat.r = r.r)
__global__ void __launch_bounds__(kBicgThreads) k_bicg_init(Dims d, const float* __restrict__ rtt,
                                                           const float* __restrict__ b, const float* __restrict__ x0,
                                                           float* __restrict__ x, float* __restrict__ r,
                                                           float* __restrict__ rhat, float* __restrict__ p,
                                                           float* __restrict__ q, double* __restrict__ partials,
                                                           BicgState* st, double tol, double atol) {
  long n = (long)d.nx * d.ny * d.nz;
  double acc[2] = {0.0, 0.0};
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int k = t % d.nz;
    long jt = t / d.nz;
    int j = jt % d.ny, i = jt / d.ny;
    float ri = b[t] - matvec_point(d, rtt, x0, i, j, k);
    x[t] = x0[t];
    r[t] = ri;
    rhat[t] = ri;
    p[t] = ri;
    q[t] = ri;
    acc[0] += (double)b[t] * b[t];
    acc[1] += (double)ri * ri;
  }
  double part[2];
  block_sum<2>(acc, part);
  if (publish_partials<2>(part, partials, &st->counter[0])) {
    double tot[2];
    final_sums<2>(partials, tot);
    if (threadIdx.x == 0) {
      st->bs = tot[0];
      st->atol2 = fmax(tol * tol * tot[0], atol * atol);
      st->rr = tot[1];
      st->rho_new = tot[1];
      st->rho = st->alpha = st->omega = 1.0;
      st->k = 0;
      st->done = !(st->rr > st->atol2 && 0 < st->maxiter);
      for (int c = 0; c < 4; ++c) st->counter[c] = 0;
    }
  }
}

// K1: (loop test) beta; p' = r + beta (p - omega q) recomputed on the 7-point stencil;
// q' = A(p'); sum rhat.q'  -> alpha
__global__ void __launch_bounds__(kBicgThreads) k_bicg_pq(Dims d, const float* __restrict__ rtt,
                                                         const float* __restrict__ r, const float* __restrict__ rhat,
                                                         const float* __restrict__ p, const float* __restrict__ q,
                                                         float* __restrict__ pn, float* __restrict__ qn,
                                                         double* __restrict__ partials, BicgState* st) {
  if (st->done) return;
  long n = (long)d.nx * d.ny * d.nz;
  double beta_d = st->rho_new / st->rho * st->alpha / st->omega;
  float beta = (float)beta_d, om = (float)st->omega;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  double acc[1] = {0.0};
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int k = t % nz;
    long jt = t / nz;
    int j = jt % ny, i = jt / ny;
    auto P = [&](int a, int b2, int c) {
      a = (a + nx) % nx;
      b2 = (b2 + ny) % ny;
      c = c < 0 ? 0 : (c >= nz ? nz - 1 : c);
      long o = ((long)a * ny + b2) * nz + c;
      return r[o] + beta * (p[o] - om * q[o]);
    };
    long cc = ic(d, i + 3, j + 3, k + 1);
    const long sx = d.sxc(), sy = d.syc();
    float xc = P(i, j, k);
    float px0 = 0.5f * (rtt[cc] + rtt[cc - sx]) * ((xc - P(i - 1, j, k)) * d.idx);
    float px1 = 0.5f * (rtt[cc + sx] + rtt[cc]) * ((P(i + 1, j, k) - xc) * d.idx);
    float py0 = 0.5f * (rtt[cc] + rtt[cc - sy]) * ((xc - P(i, j - 1, k)) * d.idy);
    float py1 = 0.5f * (rtt[cc + sy] + rtt[cc]) * ((P(i, j + 1, k) - xc) * d.idy);
    float pz0 = (k == 0) ? rtt[cc - 1] : 0.5f * (rtt[cc] + rtt[cc - 1]) * ((xc - P(i, j, k - 1)) * d.idz);
    float pz1 = (k == nz - 1) ? rtt[cc + 1] : 0.5f * (rtt[cc + 1] + rtt[cc]) * ((P(i, j, k + 1) - xc) * d.idz);
    float qv = (float)kCp * ((px1 - px0) * d.idx + (py1 - py0) * d.idy + (pz1 - pz0) * d.idz);
    pn[t] = xc;
    qn[t] = qv;
    acc[0] += (double)rhat[t] * qv;
  }
  double part[1];
  block_sum<1>(acc, part);
  if (publish_partials<1>(part, partials, &st->counter[1])) {
    double tot[1];
    final_sums<1>(partials, tot);
    if (threadIdx.x == 0) {
      st->alpha_new = st->rho_new / tot[0];
      st->counter[1] = 0;
    }
  }
}

// K2: s = r - alpha q (recomputed on the stencil), t = A(s); sums t.s, t.t, s.s
__global__ void __launch_bounds__(kBicgThreads) k_bicg_st(Dims d, const float* __restrict__ rtt,
                                                         const float* __restrict__ r, const float* __restrict__ q,
                                                         float* __restrict__ s, float* __restrict__ tv,
                                                         double* __restrict__ partials, BicgState* st) {
  if (st->done) return;
  long n = (long)d.nx * d.ny * d.nz;
  float al = (float)st->alpha_new;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int k = t % nz;
    long jt = t / nz;
    int j = jt % ny, i = jt / ny;
    auto S = [&](int a, int b2, int c) {
      a = (a + nx) % nx;
      b2 = (b2 + ny) % ny;
      c = c < 0 ? 0 : (c >= nz ? nz - 1 : c);
      long o = ((long)a * ny + b2) * nz + c;
      return r[o] - al * q[o];
    };
    long cc = ic(d, i + 3, j + 3, k + 1);
    const long sx = d.sxc(), sy = d.syc();
    float xc = S(i, j, k);
    float px0 = 0.5f * (rtt[cc] + rtt[cc - sx]) * ((xc - S(i - 1, j, k)) * d.idx);
    float px1 = 0.5f * (rtt[cc + sx] + rtt[cc]) * ((S(i + 1, j, k) - xc) * d.idx);
    float py0 = 0.5f * (rtt[cc] + rtt[cc - sy]) * ((xc - S(i, j - 1, k)) * d.idy);
    float py1 = 0.5f * (rtt[cc + sy] + rtt[cc]) * ((S(i, j + 1, k) - xc) * d.idy);
    float pz0 = (k == 0) ? rtt[cc - 1] : 0.5f * (rtt[cc] + rtt[cc - 1]) * ((xc - S(i, j, k - 1)) * d.idz);
    float pz1 = (k == nz - 1) ? rtt[cc + 1] : 0.5f * (rtt[cc + 1] + rtt[cc]) * ((S(i, j, k + 1) - xc) * d.idz);
    float tt = (float)kCp * ((px1 - px0) * d.idx + (py1 - py0) * d.idy + (pz1 - pz0) * d.idz);
    s[t] = xc;
    tv[t] = tt;
    acc[0] += (double)tt * xc;
    acc[1] += (double)tt * tt;
    acc[2] += (double)xc * xc;
  }
  double part[3];
  block_sum<3>(acc, part);
  if (publish_partials<3>(part, partials, &st->counter[2])) {
    double tot[3];
    final_sums<3>(partials, tot);
    if (threadIdx.x == 0) {
      st->omega_new = tot[0] / tot[1];
      st->ss = tot[2];
      st->early = tot[2] < st->atol2;
      st->counter[2] = 0;
    }
  }
}

// K3: x, r updates; sums r.r, rhat.r; scalar recurrence and the loop test
__global__ void __launch_bounds__(kBicgThreads) k_bicg_xr(long n, const float* __restrict__ p,
                                                         const float* __restrict__ s, const float* __restrict__ tv,
                                                         const float* __restrict__ rhat, float* __restrict__ x,
                                                         float* __restrict__ r, double* __restrict__ partials,
                                                         BicgState* st) {
  if (st->done) return;
  float al = (float)st->alpha_new, om = (float)st->omega_new;
  int early = st->early;
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
  double part[2];
  block_sum<2>(acc, part);
  if (publish_partials<2>(part, partials, &st->counter[3])) {
    double tot[2];
    final_sums<2>(partials, tot);
    if (threadIdx.x == 0) {
      double a = st->alpha_new, o = st->omega_new, rn = st->rho_new;
      int kn = (o == 0.0 || a == 0.0) ? -11 : st->k + 1;
      if (rn == 0.0) kn = -10;
      st->k = kn;
      st->rho = rn;
      st->alpha = a;
      st->omega = o;
      st->rr = tot[0];
      st->rho_new = tot[1];
      st->done = !(st->rr > st->atol2 && kn < st->maxiter && kn >= 0);
      st->counter[3] = 0;
    }
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
