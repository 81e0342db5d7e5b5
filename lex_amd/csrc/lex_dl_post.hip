// The reference's post-step deep-learning correction (ws_rk3_sprint_dl,
// one_sprint_functions.py:118-165) around the step:
//   k_dl_pack    : phys_state_tuple2array (dl_models.py:125-155) - theta, u, v, w, qv
//                  normalised per level by the scaling extremes, u, v, w averaged to
//                  cell centres (unstagger_uvw, :166-171), stacked channels-last (x,y,z,5)
//   k_dl_stagger : correct_array2tuple (:158-163) with stagger_uvw (:174-182) - the net's
//                  five centred channels back to theta, u, v, w, qv points
//   k_dl_apply   : the correction step of ws_rk3_sprint_dl (:137-160) - the staggered
//                  corrections times the scaling range added to the state, ghosts refilled
//                  (padding3 / padding3u / padding3v), pi' untouched
// stagger_uvw keeps the reference's end faces: u face 0 takes the average of the last
// two centres and u face nx the average of the first two (likewise for v); w is zero at
// the surface and the lid.
#pragma once
#include "lex_common.hpp"

namespace lex {

// per-level offsets and denominators of the normalisation, concatenated per field:
// theta, u, v (nz levels each), w (nz+1 levels), qv (nz); lo = min over x, y of the
// field's scaling minimum, den = max over x, y of its maximum - lo + c_f
struct DlNorm {
  const float* lo;
  const float* den;
};

__global__ void k_dl_pack(Dims d, const float* __restrict__ th, const float* __restrict__ u,
                          const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ qv,
                          DlNorm nrm, float* __restrict__ out) {
  const TileId tb = xcd_tile();
  const int k = tb.x * blockDim.x + threadIdx.x;
  const int j = tb.y * blockDim.y + threadIdx.y;
  const int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  const int I = i + 3, J = j + 3, K = k + 1, nz = d.nz;
  const float* lo = nrm.lo;
  const float* den = nrm.den;
  const float tn = (th[ic(d, I, J, K)] - lo[k]) / den[k];
  const float u0 = (u[iu(d, I, J, K)] - lo[nz + k]) / den[nz + k];
  const float u1 = (u[iu(d, I + 1, J, K)] - lo[nz + k]) / den[nz + k];
  const float v0 = (v[iv(d, I, J, K)] - lo[2 * nz + k]) / den[2 * nz + k];
  const float v1 = (v[iv(d, I, J + 1, K)] - lo[2 * nz + k]) / den[2 * nz + k];
  const float w0 = (w[iw(d, I, J, K)] - lo[3 * nz + k]) / den[3 * nz + k];
  const float w1 = (w[iw(d, I, J, K + 1)] - lo[3 * nz + k + 1]) / den[3 * nz + k + 1];
  const float qn = (qv[ic(d, I, J, K)] - lo[4 * nz + 1 + k]) / den[4 * nz + 1 + k];
  float* o = out + (((long)i * d.ny + j) * nz + k) * 5;
  o[0] = tn;
  o[1] = 0.5f * (u0 + u1);
  o[2] = 0.5f * (v0 + v1);
  o[3] = 0.5f * (w0 + w1);
  o[4] = qn;
}

// staggered corrections from the centred channels c (x, y, z, 5)
__device__ __forceinline__ float dl_c(const Dims& d, const float* __restrict__ c, int i, int j, int k, int ch) {
  return c[(((long)i * d.ny + j) * d.nz + k) * 5 + ch];
}
__device__ __forceinline__ float dl_u_face(const Dims& d, const float* __restrict__ c, int f, int j, int k) {
  // faces 1..nx-1 average their two centres; face 0 = face nx-1's value, face nx = face 1's
  const int g = (f == 0) ? d.nx - 1 : (f == d.nx ? 1 : f);
  return 0.5f * (dl_c(d, c, g - 1, j, k, 1) + dl_c(d, c, g, j, k, 1));
}
__device__ __forceinline__ float dl_v_face(const Dims& d, const float* __restrict__ c, int i, int f, int k) {
  const int g = (f == 0) ? d.ny - 1 : (f == d.ny ? 1 : f);
  return 0.5f * (dl_c(d, c, i, g - 1, k, 2) + dl_c(d, c, i, g, k, 2));
}
__device__ __forceinline__ float dl_w_level(const Dims& d, const float* __restrict__ c, int i, int j, int l) {
  return (l == 0 || l == d.nz) ? 0.0f : 0.5f * (dl_c(d, c, i, j, l - 1, 3) + dl_c(d, c, i, j, l, 3));
}

// corrections as interior arrays: theta, qv (nx, ny, nz), u (nx+1, ny, nz),
// v (nx, ny+1, nz), w (nx, ny, nz+1).  One thread per (i, j, k) of the (nx+1, ny+1, nz+1)
// box.
__global__ void k_dl_stagger(Dims d, const float* __restrict__ c, float* __restrict__ tc, float* __restrict__ uc,
                             float* __restrict__ vc, float* __restrict__ wc, float* __restrict__ qc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  if (k > d.nz || j > d.ny || i > d.nx) return;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  if (i < nx && j < ny && k < nz) {
    const long o = ((long)i * ny + j) * nz + k;
    tc[o] = dl_c(d, c, i, j, k, 0);
    qc[o] = dl_c(d, c, i, j, k, 4);
  }
  if (j < ny && k < nz) uc[((long)i * ny + j) * nz + k] = dl_u_face(d, c, i, j, k);
  if (i < nx && k < nz) vc[((long)i * (ny + 1) + j) * nz + k] = dl_v_face(d, c, i, j, k);
  if (i < nx && j < ny) wc[((long)i * ny + j) * (nz + 1) + k] = dl_w_level(d, c, i, j, k);
}

// state += staggered correction x range (ranges: the scaling maximum minus minimum on
// each field's own points), in place, ghosts refilled.  One thread per cell; the cells
// of the last column / row / level also update face nx, face ny and level nz.
struct DlRanges {
  const float* th;
  const float* u;
  const float* v;
  const float* w;
  const float* qv;
};

__global__ void k_dl_apply(Dims d, const float* __restrict__ c, DlRanges rg, float* __restrict__ th,
                           float* __restrict__ u, float* __restrict__ v, float* __restrict__ w,
                           float* __restrict__ qv) {
  const TileId tb = xcd_tile();
  const int k = tb.x * blockDim.x + threadIdx.x;
  const int j = tb.y * blockDim.y + threadIdx.y;
  const int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  const int I = i + 3, J = j + 3, K = k + 1, nx = d.nx, ny = d.ny, nz = d.nz;
  const long o = ((long)i * ny + j) * nz + k;
  store_c_fast(d, th, i, j, k, th[ic(d, I, J, K)] + dl_c(d, c, i, j, k, 0) * rg.th[o]);
  store_c_fast(d, qv, i, j, k, qv[ic(d, I, J, K)] + dl_c(d, c, i, j, k, 4) * rg.qv[o]);
  store_u_fast(d, u, i, j, k, u[iu(d, I, J, K)] + dl_u_face(d, c, i, j, k) * rg.u[o]);
  if (i == nx - 1)
    store_u(d, u, nx, j, k, u[iu(d, nx + 3, J, K)] + dl_u_face(d, c, nx, j, k) * rg.u[((long)nx * ny + j) * nz + k]);
  store_v_fast(d, v, i, j, k, v[iv(d, I, J, K)] + dl_v_face(d, c, i, j, k) * rg.v[((long)i * (ny + 1) + j) * nz + k]);
  if (j == ny - 1)
    store_v(d, v, i, ny, k,
            v[iv(d, I, ny + 3, K)] + dl_v_face(d, c, i, ny, k) * rg.v[((long)i * (ny + 1) + ny) * nz + k]);
  const long ow = ((long)i * ny + j) * (nz + 1) + k;
  store_w_fast(d, w, i, j, k, w[iw(d, I, J, K)] + dl_w_level(d, c, i, j, k) * rg.w[ow]);
  if (k == nz - 1) store_w(d, w, i, j, nz, w[iw(d, I, J, nz + 1)] + dl_w_level(d, c, i, j, nz) * rg.w[ow + 1]);
}

}  // namespace lex
