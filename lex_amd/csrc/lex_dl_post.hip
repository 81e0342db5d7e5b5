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

// Reverse of k_dl_apply with respect to the correction (the training's gradient through
// the post-step correction, one_sprint_functions.py:54-83): dc (nx, ny, nz, 5) = the
// transposed stagger of the state adjoint (padded, ghosts folded) times the ranges.  The u
// correction of cell i enters faces i and i+1 (those in 1..nx-1), face 0 when i is nx-2
// or nx-1 and face nx when i is 0 or 1 (the end faces of stagger_uvw); v alike; w levels
// k and k+1 (those in 1..nz-1).  One thread per cell, written once.
__global__ void k_dl_apply_vjp(Dims d, const float* __restrict__ lth, const float* __restrict__ lu,
                               const float* __restrict__ lv, const float* __restrict__ lw,
                               const float* __restrict__ lq, DlRanges rg, float* __restrict__ dc) {
  const TileId tb = xcd_tile();
  const int k = tb.x * blockDim.x + threadIdx.x;
  const int j = tb.y * blockDim.y + threadIdx.y;
  const int i = tb.z;
  if (k >= d.nz || j >= d.ny) return;
  const int I = i + 3, J = j + 3, K = k + 1, nx = d.nx, ny = d.ny, nz = d.nz;
  const long o = ((long)i * ny + j) * nz + k;
  auto fu = [&](int f) { return lu[iu(d, f + 3, J, K)] * rg.u[((long)f * ny + j) * nz + k]; };
  auto fv = [&](int f) { return lv[iv(d, I, f + 3, K)] * rg.v[((long)i * (ny + 1) + f) * nz + k]; };
  auto fw = [&](int l) { return lw[iw(d, I, J, l + 1)] * rg.w[((long)i * ny + j) * (nz + 1) + l]; };
  float su = 0.0f, sv = 0.0f, sw = 0.0f;
  if (i >= 1) su += fu(i);
  if (i + 1 <= nx - 1) su += fu(i + 1);
  if (i >= nx - 2) su += fu(0);
  if (i <= 1) su += fu(nx);
  if (j >= 1) sv += fv(j);
  if (j + 1 <= ny - 1) sv += fv(j + 1);
  if (j >= ny - 2) sv += fv(0);
  if (j <= 1) sv += fv(ny);
  if (k >= 1) sw += fw(k);
  if (k + 1 <= nz - 1) sw += fw(k + 1);
  float* c = dc + o * 5;
  c[0] = lth[ic(d, I, J, K)] * rg.th[o];
  c[1] = 0.5f * su;
  c[2] = 0.5f * sv;
  c[3] = 0.5f * sw;
  c[4] = lq[ic(d, I, J, K)] * rg.qv[o];
}

// Reverse of k_dl_pack: adds the adjoint of the state's interior values (theta, qv at
// cells, u faces 0..nx, v faces 0..ny, w levels 0..nz) for the input adjoint dp
// (nx, ny, nz, 5).  One thread per point of the (nx+1, ny+1, nz+1) box; every adjoint
// element is written by one thread.
__global__ void k_dl_pack_vjp(Dims d, const float* __restrict__ dp, DlNorm nrm, float* __restrict__ lth,
                              float* __restrict__ lu, float* __restrict__ lv, float* __restrict__ lw,
                              float* __restrict__ lq) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = blockIdx.z;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  if (k > nz || j > ny || i > nx) return;
  auto g = [&](int ii, int jj, int kk, int ch) { return dp[(((long)ii * ny + jj) * nz + kk) * 5 + ch]; };
  const float* den = nrm.den;
  if (i < nx && j < ny && k < nz) {
    lth[ic(d, i + 3, j + 3, k + 1)] += g(i, j, k, 0) / den[k];
    lq[ic(d, i + 3, j + 3, k + 1)] += g(i, j, k, 4) / den[4 * nz + 1 + k];
  }
  if (j < ny && k < nz) {   // u face i
    float a = 0.0f;
    if (i >= 1) a += g(i - 1, j, k, 1);
    if (i < nx) a += g(i, j, k, 1);
    lu[iu(d, i + 3, j + 3, k + 1)] += 0.5f * a / den[nz + k];
  }
  if (i < nx && k < nz) {   // v face j
    float a = 0.0f;
    if (j >= 1) a += g(i, j - 1, k, 2);
    if (j < ny) a += g(i, j, k, 2);
    lv[iv(d, i + 3, j + 3, k + 1)] += 0.5f * a / den[2 * nz + k];
  }
  if (i < nx && j < ny) {   // w level k
    float a = 0.0f;
    if (k >= 1) a += g(i, j, k - 1, 3);
    if (k < nz) a += g(i, j, k, 3);
    lw[iw(d, i + 3, j + 3, k + 1)] += 0.5f * a / den[3 * nz + k];
  }
}

}  // namespace lex
