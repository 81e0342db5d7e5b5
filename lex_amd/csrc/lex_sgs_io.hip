// X4/X5 data paths around the learned SGS net.
//   k_sgs_features : padded state -> 11 cell-centred input channels (channels-last FP32):
//                    u, v, w at cell centres, theta', qv' and the six rate-of-strain
//                    components (the staggering of compute_deformation,
//                    turbulence.py:298-388, without the rho factor)
//   k_sgs_apply    : 5 cell-centred net outputs -> sgs_u, sgs_v, sgs_w, sgs_theta, sgs_qv
//                    on the points and shapes of compute_smag's outputs (turbulence.py:9-111)
//   k_coarsen      : block-mean coarsening of a fine run (X5,
//                    initial_comparison/plot_theta_pert_coarsened.m:7-8): cell averages
//                    for scalars, face-area averages for u, v, w (divergence-preserving
//                    restriction on the C grid)
#pragma once
#include "lex_common.hpp"
#include "lex_smag.hip"

namespace lex {

// per-channel affine normalisation applied to the features: out = (x - shift) * scale
struct FeatNorm {
  float shift[11];
  float scale[11];
};

__global__ void k_sgs_features(Dims d, const float* __restrict__ th0c, const float* __restrict__ qv0c,
                               const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                               const float* __restrict__ th, const float* __restrict__ qv, FeatNorm nrm,
                               float* __restrict__ out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  int I = i + 3, J = j + 3, K = k + 1, nz = d.nz;
  float f[11];
  f[0] = 0.5f * (u[iu(d, I, J, K)] + u[iu(d, I + 1, J, K)]);
  f[1] = 0.5f * (v[iv(d, I, J, K)] + v[iv(d, I, J + 1, K)]);
  f[2] = 0.5f * (w[iw(d, I, J, K)] + w[iw(d, I, J, K + 1)]);
  f[3] = th[ic(d, I, J, K)] - th0c[K];
  f[4] = qv[ic(d, I, J, K)] - qv0c[K];
  f[5] = (u[iu(d, I + 1, J, K)] - u[iu(d, I, J, K)]) * d.idx;
  f[6] = (v[iv(d, I, J + 1, K)] - v[iv(d, I, J, K)]) * d.idy;
  f[7] = (w[iw(d, I, J, K + 1)] - w[iw(d, I, J, K)]) * d.idz;
  f[8] = 0.25f * (e12_at(d, u, v, I, J, K) + e12_at(d, u, v, I + 1, J + 1, K) + e12_at(d, u, v, I, J + 1, K) +
                  e12_at(d, u, v, I + 1, J, K));
  if (k == 0 || k == nz - 1) {
    int l0 = (k == 0) ? 1 : nz - 1, dl = (k == 0) ? 1 : -1;
    f[9] = 0.5f * (ex3(e13_at(d, u, w, I + 1, J, l0), e13_at(d, u, w, I + 1, J, l0 + dl),
                       e13_at(d, u, w, I + 1, J, l0 + 2 * dl)) +
                   ex3(e13_at(d, u, w, I, J, l0), e13_at(d, u, w, I, J, l0 + dl), e13_at(d, u, w, I, J, l0 + 2 * dl)));
    f[10] = 0.5f * (ex3(e23_at(d, v, w, I, J + 1, l0), e23_at(d, v, w, I, J + 1, l0 + dl),
                        e23_at(d, v, w, I, J + 1, l0 + 2 * dl)) +
                    ex3(e23_at(d, v, w, I, J, l0), e23_at(d, v, w, I, J, l0 + dl), e23_at(d, v, w, I, J, l0 + 2 * dl)));
  } else {
    f[9] = 0.25f * (e13_at(d, u, w, I, J, k) + e13_at(d, u, w, I + 1, J, k + 1) + e13_at(d, u, w, I, J, k + 1) +
                    e13_at(d, u, w, I + 1, J, k));
    f[10] = 0.25f * (e23_at(d, v, w, I, J, k) + e23_at(d, v, w, I, J + 1, k + 1) + e23_at(d, v, w, I, J, k + 1) +
                     e23_at(d, v, w, I, J + 1, k));
  }
  long o = (((long)i * d.ny + j) * d.nz + k) * 11;
#pragma unroll
  for (int c = 0; c < 11; ++c) out[o + c] = (f[c] - nrm.shift[c]) * nrm.scale[c];
}

// net output y (nx, ny, nz, 5) cell-centred, scaled by oscale[c] -> SGS tendencies:
// u faces average the two adjacent cells (periodic), w levels the cells below and above
// (zero at the surface and lid, as compute_smag's sgs_w), scalars at the cells.
__global__ void k_sgs_apply(Dims d, const float* __restrict__ y, float os0, float os1, float os2, float os3,
                            float os4, float* __restrict__ su, float* __restrict__ sv, float* __restrict__ sw,
                            float* __restrict__ sth, float* __restrict__ sqv) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nz
  int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny
  int i = blockIdx.z;                             // 0..nx
  if (k > d.nz || j > d.ny) return;
  auto Y = [&](int ii, int jj, int kk, int c) {
    ii = (ii + d.nx) % d.nx;
    jj = (jj + d.ny) % d.ny;
    return y[(((long)ii * d.ny + jj) * d.nz + kk) * 5 + c];
  };
  if (k < d.nz && j < d.ny) su[((long)i * d.ny + j) * d.nz + k] = os0 * 0.5f * (Y(i - 1, j, k, 0) + Y(i, j, k, 0));
  if (k < d.nz && i < d.nx) sv[((long)i * (d.ny + 1) + j) * d.nz + k] = os1 * 0.5f * (Y(i, j - 1, k, 1) + Y(i, j, k, 1));
  if (i < d.nx && j < d.ny) {
    float val = (k == 0 || k == d.nz) ? 0.0f : os2 * 0.5f * (Y(i, j, k - 1, 2) + Y(i, j, k, 2));
    sw[((long)i * d.ny + j) * (d.nz + 1) + k] = val;
    if (k < d.nz) {
      long o = ((long)i * d.ny + j) * d.nz + k;
      sth[o] = os3 * Y(i, j, k, 3);
      sqv[o] = os4 * Y(i, j, k, 4);
    }
  }
}

// ---------------------------------------------------------------- reverses (training
// the net through the solver)
// k_sgs_apply_vjp: adjoints of sgs_u, sgs_v, sgs_w, sgs_theta, sgs_qv -> adjoint of the
// net output ly (nx, ny, nz, 5), written (not accumulated).  Gather per cell: cell i is
// read by u faces i and i+1, and (periodic) by face 0 when i = nx-1 and face nx when
// i = 0; cell k by w levels k and k+1 inside 1..nz-1.
__global__ void k_sgs_apply_vjp(Dims d, const float* __restrict__ lsu, const float* __restrict__ lsv,
                                const float* __restrict__ lsw, const float* __restrict__ lsth,
                                const float* __restrict__ lsqv, float os0, float os1, float os2, float os3, float os4,
                                float* __restrict__ ly) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  auto U = [&](int f) { return lsu[((long)f * ny + j) * nz + k]; };
  auto V = [&](int f) { return lsv[((long)i * (ny + 1) + f) * nz + k]; };
  float gu = U(i) + U(i + 1);
  if (i == nx - 1) gu += U(0);
  if (i == 0) gu += U(nx);
  float gv = V(j) + V(j + 1);
  if (j == ny - 1) gv += V(0);
  if (j == 0) gv += V(ny);
  const long ow = ((long)i * ny + j) * (nz + 1);
  float gw = 0.f;
  if (k >= 1) gw += lsw[ow + k];
  if (k + 1 <= nz - 1) gw += lsw[ow + k + 1];
  const long o = ((long)i * ny + j) * nz + k;
  float* out = ly + o * 5;
  out[0] = os0 * 0.5f * gu;
  out[1] = os1 * 0.5f * gv;
  out[2] = os2 * 0.5f * gw;
  out[3] = os3 * lsth[o];
  out[4] = os4 * lsqv[o];
}

// reverses of the strain-rate stencils of k_sgs_features (scatter with atomics)
__device__ __forceinline__ void sgs_e12_vjp(const Dims& d, float* lu, float* lv, int F, int G, int K, float g) {
  unsafeAtomicAdd(lu + iu(d, F, G, K), 0.5f * g * d.idy);
  unsafeAtomicAdd(lu + iu(d, F, G - 1, K), -0.5f * g * d.idy);
  unsafeAtomicAdd(lv + iv(d, F, G, K), 0.5f * g * d.idx);
  unsafeAtomicAdd(lv + iv(d, F - 1, G, K), -0.5f * g * d.idx);
}
__device__ __forceinline__ void sgs_e13_vjp(const Dims& d, float* lu, float* lw, int F, int J, int l, float g) {
  unsafeAtomicAdd(lu + iu(d, F, J, l + 1), 0.5f * g * d.idz);
  unsafeAtomicAdd(lu + iu(d, F, J, l), -0.5f * g * d.idz);
  unsafeAtomicAdd(lw + iw(d, F, J, l + 1), 0.5f * g * d.idx);
  unsafeAtomicAdd(lw + iw(d, F - 1, J, l + 1), -0.5f * g * d.idx);
}
__device__ __forceinline__ void sgs_e23_vjp(const Dims& d, float* lv, float* lw, int I, int G, int l, float g) {
  unsafeAtomicAdd(lv + iv(d, I, G, l + 1), 0.5f * g * d.idz);
  unsafeAtomicAdd(lv + iv(d, I, G, l), -0.5f * g * d.idz);
  unsafeAtomicAdd(lw + iw(d, I, G, l + 1), 0.5f * g * d.idy);
  unsafeAtomicAdd(lw + iw(d, I, G - 1, l + 1), -0.5f * g * d.idy);
}

// k_sgs_features_vjp: adjoint of the normalised features lf (nx, ny, nz, 11) -> padded
// adjoints of u, v, w, theta, qv (accumulated, ghost lines included: fold afterwards).
__global__ void k_sgs_features_vjp(Dims d, const float* __restrict__ lf, FeatNorm nrm, float* __restrict__ lu,
                                   float* __restrict__ lv, float* __restrict__ lw, float* __restrict__ lth,
                                   float* __restrict__ lqv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.nz || j >= d.ny) return;
  const int I = i + 3, J = j + 3, K = k + 1, nz = d.nz;
  const long o = (((long)i * d.ny + j) * d.nz + k) * 11;
  float g[11];
#pragma unroll
  for (int c = 0; c < 11; ++c) g[c] = lf[o + c] * nrm.scale[c];
  unsafeAtomicAdd(lu + iu(d, I, J, K), 0.5f * g[0] - g[5] * d.idx);
  unsafeAtomicAdd(lu + iu(d, I + 1, J, K), 0.5f * g[0] + g[5] * d.idx);
  unsafeAtomicAdd(lv + iv(d, I, J, K), 0.5f * g[1] - g[6] * d.idy);
  unsafeAtomicAdd(lv + iv(d, I, J + 1, K), 0.5f * g[1] + g[6] * d.idy);
  unsafeAtomicAdd(lw + iw(d, I, J, K), 0.5f * g[2] - g[7] * d.idz);
  unsafeAtomicAdd(lw + iw(d, I, J, K + 1), 0.5f * g[2] + g[7] * d.idz);
  unsafeAtomicAdd(lth + ic(d, I, J, K), g[3]);
  unsafeAtomicAdd(lqv + ic(d, I, J, K), g[4]);
  const float h12 = 0.25f * g[8];
  sgs_e12_vjp(d, lu, lv, I, J, K, h12);
  sgs_e12_vjp(d, lu, lv, I + 1, J + 1, K, h12);
  sgs_e12_vjp(d, lu, lv, I, J + 1, K, h12);
  sgs_e12_vjp(d, lu, lv, I + 1, J, K, h12);
  if (k == 0 || k == nz - 1) {
    const int l0 = (k == 0) ? 1 : nz - 1, dl = (k == 0) ? 1 : -1;
    const float cf[3] = {1.875f, -1.25f, 0.375f};
    for (int t = 0; t < 3; ++t) {
      const int l = l0 + t * dl;
      sgs_e13_vjp(d, lu, lw, I + 1, J, l, 0.5f * cf[t] * g[9]);
      sgs_e13_vjp(d, lu, lw, I, J, l, 0.5f * cf[t] * g[9]);
      sgs_e23_vjp(d, lv, lw, I, J + 1, l, 0.5f * cf[t] * g[10]);
      sgs_e23_vjp(d, lv, lw, I, J, l, 0.5f * cf[t] * g[10]);
    }
  } else {
    const float a = 0.25f * g[9], b = 0.25f * g[10];
    sgs_e13_vjp(d, lu, lw, I, J, k, a);
    sgs_e13_vjp(d, lu, lw, I + 1, J, k + 1, a);
    sgs_e13_vjp(d, lu, lw, I, J, k + 1, a);
    sgs_e13_vjp(d, lu, lw, I + 1, J, k, a);
    sgs_e23_vjp(d, lv, lw, I, J, k, b);
    sgs_e23_vjp(d, lv, lw, I, J + 1, k + 1, b);
    sgs_e23_vjp(d, lv, lw, I, J, k + 1, b);
    sgs_e23_vjp(d, lv, lw, I, J + 1, k, b);
  }
}

// Coarsening of a padded fine field by (fx, fy, fz) into an interior coarse field.
// mode 0: cell block mean of a scalar -> (nxc, nyc, nzc);
// mode 1: u, mean over the fy*fz fine faces of each coarse face -> (nxc+1, nyc, nzc);
// mode 2: v -> (nxc, nyc+1, nzc); mode 3: w -> (nxc, nyc, nzc+1).
__global__ void k_coarsen(Dims d, int mode, int fx, int fy, int fz, const float* __restrict__ a,
                          float* __restrict__ out) {
  const int nxc = d.nx / fx, nyc = d.ny / fy, nzc = d.nz / fz;
  const int ni = nxc + (mode == 1), nj = nyc + (mode == 2), nk = nzc + (mode == 3);
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int i = blockIdx.z;
  if (k >= nk || j >= nj || i >= ni) return;
  float s = 0.0f;
  int n = 0;
  if (mode == 0) {
    for (int a1 = 0; a1 < fx; ++a1)
      for (int b1 = 0; b1 < fy; ++b1)
        for (int c1 = 0; c1 < fz; ++c1) {
          s += a[ic(d, i * fx + a1 + 3, j * fy + b1 + 3, k * fz + c1 + 1)];
          ++n;
        }
  } else if (mode == 1) {
    for (int b1 = 0; b1 < fy; ++b1)
      for (int c1 = 0; c1 < fz; ++c1) {
        s += a[iu(d, i * fx + 3, j * fy + b1 + 3, k * fz + c1 + 1)];
        ++n;
      }
  } else if (mode == 2) {
    for (int a1 = 0; a1 < fx; ++a1)
      for (int c1 = 0; c1 < fz; ++c1) {
        s += a[iv(d, i * fx + a1 + 3, j * fy + 3, k * fz + c1 + 1)];
        ++n;
      }
  } else {
    for (int a1 = 0; a1 < fx; ++a1)
      for (int b1 = 0; b1 < fy; ++b1) {
        s += a[iw(d, i * fx + a1 + 3, j * fy + b1 + 3, k * fz + 1)];
        ++n;
      }
  }
  out[((long)i * nj + j) * nk + k] = s / (float)n;
}

}  // namespace lex
