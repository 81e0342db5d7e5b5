// The loss of the reference's post-step-correction training (lex_dl_train.py:49-99):
//   k_laplace          : laplace_of_tensor (dataset_generation.py:271-320) on a stack of
//                        interior scalar fields: periodic in x and y, the z neighbours
//                        beyond the ends copied from the end levels (so the z derivative
//                        is zero on the bottom and top faces).  The reference's
//                        s_z[:, :, (0, -1)] = 0 on its 5-D stacks indexes the x axis: it
//                        zeroes the z derivative of the first and last x planes, so their
//                        Laplacian has no vertical term; kept as the reference has it.
//   k_wsum_partial /
//   k_wsum_final       : deterministic FP64 sums of weighted terms over n elements with
//                        a weight array repeating every `period` elements:
//                          op 0: w * 0.5 (a - b)^2      (optax.l2_loss times the weight)
//                          op 1: w * max(-a, 0)         (the negative-qv penalty)
//                        fixed block partials added in a fixed order, so the result does
//                        not depend on scheduling
#pragma once
#include "lex_common.hpp"

namespace lex {

// lanes along z, one block row per (j, x-plane of a field): blockIdx.y = j,
// blockIdx.z = f * nx + i (the caller launches field chunks with nx * chunk <= 65535)
__global__ void k_laplace(int nx, int ny, int nz, float dx, float dy, float dz, const float* __restrict__ s,
                          float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  const int j = blockIdx.y;
  const int f = blockIdx.z / nx, i = blockIdx.z - f * nx;
  const float* a = s + (long)f * nx * ny * nz;
  auto at = [&](int ii, int jj, int kk) { return a[((long)ii * ny + jj) * nz + kk]; };
  const float c = at(i, j, k);
  const int im = (i == 0) ? nx - 1 : i - 1, ip = (i == nx - 1) ? 0 : i + 1;
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  // first differences on the faces, then their differences
  const float sx_lo = (c - at(im, j, k)) / dx, sx_hi = (at(ip, j, k) - c) / dx;
  const float sy_lo = (c - at(i, jm, k)) / dy, sy_hi = (at(i, jp, k) - c) / dy;
  const bool zt = (i != 0) && (i != nx - 1);
  const float sz_lo = (zt && k > 0) ? (c - at(i, j, k - 1)) / dz : 0.0f;
  const float sz_hi = (zt && k < nz - 1) ? (at(i, j, k + 1) - c) / dz : 0.0f;
  out[(long)f * nx * ny * nz + ((long)i * ny + j) * nz + k] =
      (sx_hi - sx_lo) / dx + (sy_hi - sy_lo) / dy + (sz_hi - sz_lo) / dz;
}

constexpr int kWsumBlocks = 1024;
constexpr int kWsumThreads = 256;

__device__ __forceinline__ double wsum_term(int op, float a, float b, float w) {
  if (op == 0) {
    const double d = (double)a - (double)b;
    return (double)w * 0.5 * d * d;
  }
  return (double)w * (a < 0.0f ? -(double)a : 0.0);
}

__global__ void __launch_bounds__(kWsumThreads) k_wsum_partial(long n, int op, const float* __restrict__ a,
                                                               const float* __restrict__ b,
                                                               const float* __restrict__ w, long period,
                                                               double* __restrict__ part) {
  __shared__ double red[kWsumThreads];
  double acc = 0.0;
  const long stride = (long)kWsumBlocks * kWsumThreads;
  // thread p takes weight index p (grid-stride over the period) in every repetition
  for (long p = (long)blockIdx.x * kWsumThreads + threadIdx.x; p < period; p += stride) {
    const float wp = w[p];
    for (long t = p; t < n; t += period) acc += wsum_term(op, a[t], b ? b[t] : 0.0f, wp);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int m = kWsumThreads / 2; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(64) k_wsum_final(const double* __restrict__ part, double* __restrict__ out) {
  double s = 0.0;
  for (int p = threadIdx.x; p < kWsumBlocks; p += 64) s += part[p];
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
  if (threadIdx.x == 0) out[0] = s;
}

// The gradient of those terms with respect to a, accumulated: out += w (a - b) (op 0) or
// out += -w where a < 0 (op 1), w repeating every `period` elements.
__global__ void k_wgrad(long n, int op, const float* __restrict__ a, const float* __restrict__ b,
                        const float* __restrict__ w, long period, float* __restrict__ out) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < period; p += stride) {
    const float wp = w[p];
    for (long t = p; t < n; t += period) {
      const float g = (op == 0) ? wp * (a[t] - b[t]) : (a[t] < 0.0f ? -wp : 0.0f);
      out[t] += g;
    }
  }
}

}  // namespace lex
