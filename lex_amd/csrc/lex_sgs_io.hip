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

// Reverse of k_sgs_features in gather form (no atomics, bit-reproducible): one thread per
// padded u, v or w point adds, in a fixed order, every term of the cells whose features
// read that point.  The features of cell (i, j, k) read: u at its faces I, I+1 (the centre
// mean and du/dx), v and w likewise; e12 as the mean of its four (x, y) corners, each a
// difference of u across y and of v across x; e13 (e23) as the mean of its four x-z (y-z)
// edges at its lower and upper w levels, or, in the surface and top cells, the
// 1.875 / -1.25 / 0.375 extrapolation of the edges at levels 1..3 (nz-1..nz-3).
struct SgsFeatAdj {
  Dims d;
  const float* lf;
  FeatNorm nrm;
  // scaled adjoint of feature c of interior cell (i, j, k), zero outside the interior
  __device__ __forceinline__ float g(int c, int i, int j, int k) const {
    const bool ok = (unsigned)i < (unsigned)d.nx && (unsigned)j < (unsigned)d.ny && (unsigned)k < (unsigned)d.nz;
    return ok ? lf[(((long)i * d.ny + j) * d.nz + k) * 11 + c] * nrm.scale[c] : 0.f;
  }
  // weight of e12 corner (F, G) (padded) at cell level k: 1/4 of the e12 adjoint of each
  // of the four cells around it
  __device__ __forceinline__ float h12(int F, int G, int k) const {
    return 0.25f * g(8, F - 4, G - 4, k) + 0.25f * g(8, F - 3, G - 4, k) + 0.25f * g(8, F - 4, G - 3, k) +
           0.25f * g(8, F - 3, G - 3, k);
  }
  // weight of the e13 (c = 9) or e23 (c = 10) edge of cell (i, j) at padded w level l + 1
  __device__ __forceinline__ float hedge(int c, int i, int j, int l) const {
    const int nz = d.nz;
    float h = 0.f;
    if (l - 1 >= 1 && l - 1 <= nz - 2) h += 0.25f * g(c, i, j, l - 1);   // interior cell k = l-1 (its upper edge)
    if (l >= 1 && l <= nz - 2) h += 0.25f * g(c, i, j, l);               // interior cell k = l (its lower edge)
    if (l >= 1 && l <= 3) h += 0.5f * (l == 1 ? 1.875f : (l == 2 ? -1.25f : 0.375f)) * g(c, i, j, 0);
    if (nz - 1 > 0 && l >= nz - 3 && l <= nz - 1)
      h += 0.5f * (l == nz - 1 ? 1.875f : (l == nz - 2 ? -1.25f : 0.375f)) * g(c, i, j, nz - 1);
    return h;
  }
  // e13 edge (F, J, l): cells i = F-4, F-3 of row j; e23 edge (I, G, l): rows j = G-4, G-3
  __device__ __forceinline__ float h13(int F, int j, int l) const { return hedge(9, F - 4, j, l) + hedge(9, F - 3, j, l); }
  __device__ __forceinline__ float h23(int i, int G, int l) const { return hedge(10, i, G - 4, l) + hedge(10, i, G - 3, l); }
};

// lu at padded (F, G, L), and lth, lqv of cell (F-3, G-3, L-1) when it is interior
__global__ void k_sgs_features_vjp_u(SgsFeatAdj A, float* __restrict__ lu, float* __restrict__ lth,
                                     float* __restrict__ lqv) {
  const Dims& d = A.d;
  const int L = blockIdx.x * blockDim.x + threadIdx.x;
  const int G = blockIdx.y * blockDim.y + threadIdx.y;
  const int F = blockIdx.z;
  if (L >= d.nz + 2 || G >= d.ny + 6) return;
  const int i = F - 3, j = G - 3, k = L - 1;
  float a = 0.5f * A.g(0, i, j, k) - A.g(5, i, j, k) * d.idx;           // face I of cell i
  a += 0.5f * A.g(0, i - 1, j, k) + A.g(5, i - 1, j, k) * d.idx;        // face I+1 of cell i-1
  a += 0.5f * d.idy * (A.h12(F, G, k) - A.h12(F, G + 1, k));
  a += 0.5f * d.idz * (A.h13(F, j, L - 1) - A.h13(F, j, L));
  lu[iu(d, F, G, L)] += a;
  if ((unsigned)i < (unsigned)d.nx && (unsigned)j < (unsigned)d.ny && (unsigned)k < (unsigned)d.nz) {
    lth[ic(d, F, G, L)] += A.g(3, i, j, k);
    lqv[ic(d, F, G, L)] += A.g(4, i, j, k);
  }
}

__global__ void k_sgs_features_vjp_v(SgsFeatAdj A, float* __restrict__ lv) {
  const Dims& d = A.d;
  const int L = blockIdx.x * blockDim.x + threadIdx.x;
  const int G = blockIdx.y * blockDim.y + threadIdx.y;
  const int F = blockIdx.z;
  if (L >= d.nz + 2 || G >= d.ny + 7) return;
  const int i = F - 3, j = G - 3, k = L - 1;
  float a = 0.5f * A.g(1, i, j, k) - A.g(6, i, j, k) * d.idy;           // face J of cell j
  a += 0.5f * A.g(1, i, j - 1, k) + A.g(6, i, j - 1, k) * d.idy;        // face J+1 of cell j-1
  a += 0.5f * d.idx * (A.h12(F, G, k) - A.h12(F + 1, G, k));
  a += 0.5f * d.idz * (A.h23(i, G, L - 1) - A.h23(i, G, L));
  lv[iv(d, F, G, L)] += a;
}

__global__ void k_sgs_features_vjp_w(SgsFeatAdj A, float* __restrict__ lw) {
  const Dims& d = A.d;
  const int L = blockIdx.x * blockDim.x + threadIdx.x;
  const int G = blockIdx.y * blockDim.y + threadIdx.y;
  const int F = blockIdx.z;
  if (L >= d.nz + 3 || G >= d.ny + 6) return;
  const int i = F - 3, j = G - 3;
  float a = 0.5f * A.g(2, i, j, L - 1) - A.g(7, i, j, L - 1) * d.idz;   // lower level of cell L-1
  a += 0.5f * A.g(2, i, j, L - 2) + A.g(7, i, j, L - 2) * d.idz;        // upper level of cell L-2
  a += 0.5f * d.idx * (A.h13(F, j, L - 1) - A.h13(F + 1, j, L - 1));
  a += 0.5f * d.idy * (A.h23(i, G, L - 1) - A.h23(i, G + 1, L - 1));
  lw[iw(d, F, G, L)] += a;
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
