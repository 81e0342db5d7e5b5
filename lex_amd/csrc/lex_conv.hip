// X4: the 3x3x3 'SAME' convolutions of the SGS correction net on BF16 MFMA with FP32
// accumulation.  The architecture follows simpleCNN (dl_models.py:19-30): four Conv
// layers (kernel 3x3x3, stride 1, zero 'SAME' padding, bias) with the tanh-approximate
// GELU (flax nn.gelu's default) between them.
//
// Layout: activations channels-last (x, y, z, c) with the channel count padded to a
// multiple of 16 (CIN, COUT template arguments, 16 or 32).  Weights are kept in FP32 as
// W[tap][cin][cout] (flax's (3,3,3,in,out) kernel, tap = kx*9 + ky*3 + kz) and packed
// for the kernels into BF16 rows Wp[cout][step*32 + k] (one MFMA K-step per 32 k).
//
// MFMA: v_mfma_f32_16x16x32_bf16.  Lane l supplies A[row l%16][k 8(l/16)..+7],
// B[k 8(l/16)..+7][col l%16] and receives D[row 4(l/16)+i][col l%16], i = 0..3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "lex_common.hpp"

namespace lex {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CTX = 4, CTY = 4, CTZ = 16;  // output tile of one workgroup (4 waves)
constexpr int HX = CTX + 2, HY = CTY + 2, HZ = CTZ + 2;

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// tanh for the GELU of BF16 activations: 1 - 2 / (1 + e^(2y)) from the hardware
// exponential and reciprocal (a few instructions instead of the library's tens).  It
// saturates to +-1 with e^(2y) at inf / 0; near 0 its absolute error is ~1e-7, which the
// GELU's factor x makes a relative error of the same size, far below BF16's 4e-3.
__device__ __forceinline__ float tanh_fast(float y) {
  const float e = __expf(2.0f * y);
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}
// tanh-approximate GELU and its derivative.  0.5 x (1 + tanh(y)) = x / (1 + e^(-2y)) with
// y = k0 (x + k1 x^3): seven instructions, e^(-2y) = 2^(x (c0 + c1 x^2)) with
// c0 = -2 k0 log2(e), c1 = -2 k0 k1 log2(e); it saturates to x and to -0 like the tanh form.
__device__ __forceinline__ float gelu_t(float x) {
  const float c0 = (float)(-2.0 * 0.7978845608028654 * 1.4426950408889634);
  const float c1 = (float)(-2.0 * 0.7978845608028654 * 0.044715 * 1.4426950408889634);
  const float e = __builtin_amdgcn_exp2f(x * fmaf(c1, x * x, c0));
  return x * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_t_grad(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  float t = tanh_fast(k0 * (x + k1 * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k0 * (1.0f + 3.0f * k1 * x * x);
}

// Element stride of one cell in an LDS halo tile: the CIN channels plus 8 (16 bytes) of
// padding.  An MFMA operand read takes 16 bytes from each of 16 consecutive cells (one z
// run); with a stride of CIN elements (32 or 64 bytes for 16 or 32 channels) those reads
// fall on 2 or 4 distinct bank groups of the 256-byte bank row, with CIN + 8 (48 or 80
// bytes, an odd multiple of 16) on 16 distinct ones.
__host__ __device__ constexpr int conv_cell_stride(int cin) { return cin + 8; }

// K-steps of 32 over (tap, channel): CIN <= 32 packs 32 / CIN taps per step; CIN = 64
// spends two steps per tap (channels 0..31, then 32..63).
template <int CIN>
struct ConvSteps {
  static constexpr int TPS = CIN <= 32 ? 32 / CIN : 1;              // taps per K-step
  static constexpr int SPT = CIN <= 32 ? 1 : CIN / 32;              // K-steps per tap
  static constexpr int NSTEP = CIN <= 32 ? (27 + TPS - 1) / TPS : 27 * SPT;
};
// tap and first channel of a lane's 8 k values (kl = 8 * (lane / 16)) in step s
__host__ __device__ __forceinline__ void conv_step_k(int cinp, int s, int kl, int& tap, int& ch) {
  if (cinp <= 32) {
    tap = s * (32 / cinp) + kl / cinp;
    ch = kl % cinp;
  } else {
    tap = s / (cinp / 32);
    ch = (s % (cinp / 32)) * 32 + kl;
  }
}

// Forward convolution.  mode bit 0: GELU after the bias; bit 1: FP32 output Yf with
// cout_real channels (else BF16 Y with COUT channels); bit 2: periodic x and y.  pre (optional): FP32
// pre-activation, COUT channels.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_fwd(int nx, int ny, int nz, const __bf16* __restrict__ X,
                                                   const __bf16* __restrict__ Wp, const float* __restrict__ bias,
                                                   int mode, int cout_real, __bf16* __restrict__ Y,
                                                   float* __restrict__ Yf, float* __restrict__ pre) {
  constexpr int NSTEP = ConvSteps<CIN>::NSTEP, NT = COUT / 16, CS = conv_cell_stride(CIN);
  __shared__ __attribute__((aligned(16))) __bf16 tile[HX * HY * HZ * CS];
  // neighbouring tiles (which share halo cells) on one XCD: z tiles fastest, then y, x
  const TileId tb = xcd_tile();
  const int z0 = tb.x * CTZ, y0 = tb.y * CTY, x0 = tb.z * CTX;
  constexpr int nchunk = HX * HY * HZ * CIN / 8;
  const bool circ = (mode & 4) != 0;  // periodic x and y (flax padding 'CIRCULAR' there), zero z
  for (int c = threadIdx.x; c < nchunk; c += 256) {
    int cell = c / (CIN / 8), g = c % (CIN / 8);
    int hz = cell % HZ, hy = (cell / HZ) % HY, hx = cell / (HZ * HY);
    int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
    if (circ) {
      gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
      gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
    }
    bf16x8 v;
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) {
      v = *(const bf16x8*)(X + (((long)gx * ny + gy) * nz + gz) * CIN + 8 * g);
    } else {
      for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
    }
    *(bf16x8*)(tile + cell * CS + 8 * g) = v;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 15, g = l >> 4;
  f32x4 acc[CTY][NT];
  for (int a = 0; a < CTY; ++a)
    for (int b = 0; b < NT; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int kl = 8 * g;  // this lane's first k within a step
  // the weight fragments of step s+1 are fetched while step s runs
  bf16x8 bw[NT], bwn[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) bw[nt] = *(const bf16x8*)(Wp + (long)(nt * 16 + m) * (NSTEP * 32) + kl);
  for (int s = 0; s < NSTEP; ++s) {
    int tap, ch0;
    conv_step_k(CIN, s, kl, tap, ch0);
    if (tap > 26) tap = 26;              // padded K-step: weights are zero there
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    const int sn = (s + 1 < NSTEP) ? s + 1 : s;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      bwn[nt] = *(const bf16x8*)(Wp + (long)(nt * 16 + m) * (NSTEP * 32) + sn * 32 + kl);
    bf16x8 a[CTY];
#pragma unroll
    for (int mt = 0; mt < CTY; ++mt) {
      const int cell = ((w + 1 + dx) * HY + (mt + 1 + dy)) * HZ + (m + 1 + dz);
      a[mt] = *(const bf16x8*)(tile + cell * CS + ch0);
    }
#pragma unroll
    for (int mt = 0; mt < CTY; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma16(a[mt], bw[nt], acc[mt][nt]);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) bw[nt] = bwn[nt];
  }
  const int xc = x0 + w;
#pragma unroll
  for (int mt = 0; mt < CTY; ++mt) {
    const int yc = y0 + mt;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int n = nt * 16 + m;
      const float bn = bias[n];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int zc = z0 + 4 * g + i;
        if (xc < nx && yc < ny && zc < nz) {
          const long cell = ((long)xc * ny + yc) * nz + zc;
          float v = acc[mt][nt][i] + bn;
          if (pre) pre[cell * COUT + n] = v;
          if (mode & 1) v = gelu_t(v);
          if (mode & 2) {
            if (n < cout_real) Yf[cell * cout_real + n] = v;
          } else {
            Y[cell * COUT + n] = (__bf16)v;
          }
        }
      }
    }
  }
}

// The forward convolution for COUT a multiple of 32 on v_mfma_f32_32x32x16_bf16: one
// MFMA covers 32 cells x 32 output channels, so each A fragment read from LDS (32 cells x
// 16 k) feeds 32 (or 64) output channels, half the LDS bytes per flop of the 16x16x32
// form, whose A fragments feed 16.  Output tile 4 x 4 x 16 (x, y, z) per 4-wave
// workgroup, as k_conv3_fwd (the same 52 KB halo tile for 32 channels): wave w owns
// x = x0 + w and two accumulator blocks, each 2 y rows x 16 z cells as the 32 MFMA rows
// (row m: y offset m / 16, z offset m % 16).  Operand layout: lane l supplies A[row l%32][k 8(l/32)..+7]
// and B[k 8(l/32)..+7][col l%32] and receives D[row 8(i/4) + 4(l/32) + i%4][col l%32]
// in accumulator i = 0..15.  Same arguments and results as k_conv3_fwd.
constexpr int CT2X = 4, CT2Y = 4, CT2Z = 16;
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_fwd32(int nx, int ny, int nz, const __bf16* __restrict__ X,
                                                     const __bf16* __restrict__ Wp, const float* __restrict__ bias,
                                                     int mode, int cout_real, __bf16* __restrict__ Y,
                                                     float* __restrict__ Yf, float* __restrict__ pre) {
  constexpr int NSTEP = ConvSteps<CIN>::NSTEP, NT = COUT / 32, NK = 2 * NSTEP;   // 16-wide K steps
  constexpr int GX = CT2X + 2, GY = CT2Y + 2, GZ = CT2Z + 2, CS = conv_cell_stride(CIN);
  __shared__ __attribute__((aligned(16))) __bf16 tile[GX * GY * GZ * CS];
  const TileId tb = xcd_tile();  // neighbouring tiles on one XCD (shared halo cells in its L2)
  const int z0 = tb.x * CT2Z, y0 = tb.y * CT2Y, x0 = tb.z * CT2X;
  constexpr int nchunk = GX * GY * GZ * CIN / 8;
  const bool circ = (mode & 4) != 0;  // periodic x and y (flax padding 'CIRCULAR' there), zero z
  for (int c = threadIdx.x; c < nchunk; c += 256) {
    int cell = c / (CIN / 8), g = c % (CIN / 8);
    int hz = cell % GZ, hy = (cell / GZ) % GY, hx = cell / (GZ * GY);
    int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
    if (circ) {
      gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
      gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
    }
    bf16x8 v;
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) {
      v = *(const bf16x8*)(X + (((long)gx * ny + gy) * nz + gz) * CIN + 8 * g);
    } else {
      for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
    }
    *(bf16x8*)(tile + cell * CS + 8 * g) = v;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 31, h = l >> 5;
  f32x16 acc[CT2Y / 2][NT];
#pragma unroll
  for (int a = 0; a < CT2Y / 2; ++a)
#pragma unroll
    for (int b = 0; b < NT; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.f;
  // fully unrolled: the tap of a K step does not depend on the lane (a step of 16 k is
  // one tap's channels 16 s' .. 16 s' + 15), so every tap offset is a constant and only
  // the channel offset 8 (l / 32) varies by lane
  const __bf16* tw = tile + ((w + 1) * GY + 1) * GZ * CS + (((m >> 4) * GZ + (m & 15) + 1) * CS) + 8 * h;
  const __bf16* wl = Wp + (long)m * (NSTEP * 32) + 8 * h;
#pragma unroll
  for (int s = 0; s < NK; ++s) {
    const int kk0 = s * 16;          // the step's first k (lane half 0)
    int tap, ch0;
    conv_step_k(CIN, kk0 / 32, kk0 % 32, tap, ch0);
    if (tap > 26) tap = 26;          // padded K: weights are zero there
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    bf16x8 b[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) b[nt] = *(const bf16x8*)(wl + (long)nt * 32 * (NSTEP * 32) + kk0);
#pragma unroll
    for (int mt = 0; mt < CT2Y / 2; ++mt) {
      // cell ((w + 1 + dx), (2 mt + m/16 + 1 + dy), (m%16 + 1 + dz)), channel ch0 + 8 h
      const int off = ((dx * GY + 2 * mt + dy) * GZ + dz) * CS + ch0;
      const bf16x8 a = *(const bf16x8*)(tw + off);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[nt], acc[mt][nt], 0, 0, 0);
    }
  }
  const int xc = x0 + w;
#pragma unroll
  for (int mt = 0; mt < CT2Y / 2; ++mt) {
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int n = nt * 32 + m;
      const float bn = bias[n];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int row = 8 * (i / 4) + 4 * h + (i % 4);   // MFMA row of accumulator i
        const int yc = y0 + 2 * mt + (row >> 4), zc = z0 + (row & 15);
        if (xc < nx && yc < ny && zc < nz) {
          const long cell = ((long)xc * ny + yc) * nz + zc;
          float v = acc[mt][nt][i] + bn;
          if (pre) pre[cell * COUT + n] = v;
          if (mode & 1) v = gelu_t(v);
          if (mode & 2) {
            if (n < cout_real) Yf[cell * cout_real + n] = v;
          } else {
            Y[cell * COUT + n] = (__bf16)v;
          }
        }
      }
    }
  }
}

// Persistent form for CIN <= 32 and 32 output channels (the SGS net's 32-channel layers).
// Workgroup b owns the contiguous tile range [b T / G, (b+1) T / G) (z tiles fastest, then
// y, then x): consecutive tiles of one workgroup share halo cells in its L2.  The weight
// fragments of every K step stay in registers for the whole launch (NK x 4 VGPRs: 216 for
// 32 input channels, so one wave per SIMD; 108 for 16, which fit two), the halo of tile
// t+1 is fetched into registers while tile t's MFMAs and epilogue run, and the K loop
// reads only LDS.  MODE: bit 0 GELU after the bias, bit 1 FP32 output Yf with cout_real
// channels (else BF16 Y), bit 2 periodic x and y, bit 3 GELU gradient: `pre` is read (the
// FP32 pre-activation of the layer below, 32 channels) and Y = BF16((acc + bias) *
// gelu'(pre)), the data gradient and the k_gelu_bwd pass after it in one (bits 0-2
// clear).  Same results as k_conv3_fwd32 (and, bit 3, as that pair).
template <int CIN, int MODE>
__global__ __launch_bounds__(256, CIN <= 16 ? 2 : 1) void k_conv3_fwd32p(int nx, int ny, int nz, const __bf16* __restrict__ X,
                                                      const __bf16* __restrict__ Wp, const float* __restrict__ bias,
                                                      int cout_real, __bf16* __restrict__ Y, float* __restrict__ Yf,
                                                      float* __restrict__ pre) {
  constexpr int COUT = 32;
  static_assert(CIN <= 32, "persistent form: the weights of CIN <= 32 fit the register file");
  constexpr int NSTEP = ConvSteps<CIN>::NSTEP, NK = 2 * NSTEP;   // 16-wide K steps
  constexpr int GX = CT2X + 2, GY = CT2Y + 2, GZ = CT2Z + 2, CS = conv_cell_stride(CIN);
  constexpr int NCH = CIN / 8, NTOT = GX * GY * GZ * NCH, NPF = (NTOT + 255) / 256;  // 16-byte chunks
  constexpr bool GELU = (MODE & 1) != 0, F32 = (MODE & 2) != 0, CIRC = (MODE & 4) != 0, GGRAD = (MODE & 8) != 0;
  static_assert(!GGRAD || (MODE & 7) == 0, "the GELU-gradient epilogue writes BF16 without GELU");
  constexpr bool PF = GGRAD && CIN > 16;   // pre-activations prefetched for the whole tile
  // LDS row (one z run of GZ cells) stride RS elements, padded to 256 bytes modulo 512: an
  // operand read takes two z runs of 16 cells in adjacent rows (lanes 0-15, 16-31), whose
  // 16-byte slots then fall on the two complementary halves of the 512-byte bank row
  constexpr int RS = (GZ * CS * 2 + (((256 - GZ * CS * 2) % 512) + 512) % 512) / 2;
  static_assert((RS * 2) % 512 == 256, "row stride");
  __shared__ __attribute__((aligned(16))) __bf16 tile[GX * GY * RS];
  const int ntz = (nz + CT2Z - 1) / CT2Z, nty = (ny + CT2Y - 1) / CT2Y, ntx = (nx + CT2X - 1) / CT2X;
  const int ntiles = ntz * nty * ntx;
  const int t_begin = (int)((long)blockIdx.x * ntiles / gridDim.x);
  const int t_end = (int)((long)(blockIdx.x + 1) * ntiles / gridDim.x);
  if (t_begin >= t_end) return;  // uniform over the workgroup
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 31, h = l >> 5;
  const float bn = bias[m];
  bf16x8 wr[NK];
  {
    const __bf16* wl = Wp + (long)m * (NSTEP * 32) + 8 * h;
#pragma unroll
    for (int s = 0; s < NK; ++s) wr[s] = *(const bf16x8*)(wl + s * 16);
  }
  bf16x8 pf[NPF];
  // thread t loads chunk t % NCH of halo cells t / NCH + STEP j; the cell coordinates
  // advance by a constant per j (no divisions per tile)
  constexpr int STEP = 256 / NCH, SZ = STEP % GZ, SY = (STEP / GZ) % GY, SX = STEP / (GZ * GY);
  const int g = threadIdx.x % NCH, q0 = threadIdx.x / NCH;
  const int hz0 = q0 % GZ, hy0 = (q0 / GZ) % GY, hx0 = q0 / (GZ * GY);
  auto fetch = [&](int t) {
    const int z0 = (t % ntz) * CT2Z, y0 = ((t / ntz) % nty) * CT2Y, x0 = (t / (ntz * nty)) * CT2X;
    int hz = hz0, hy = hy0, hx = hx0;
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int c = threadIdx.x + 256 * j;
      bf16x8 v;
      for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.0f;
      if (c < NTOT) {
        int gx = x0 + hx - 1, gy = y0 + hy - 1;
        const int gz = z0 + hz - 1;
        if (CIRC) {
          gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
          gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        }
        if ((unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz)
          v = *(const bf16x8*)(X + (((long)gx * ny + gy) * nz + gz) * CIN + 8 * g);
      }
      pf[j] = v;
      hz += SZ;
      const int cz = hz >= GZ;
      hz -= cz ? GZ : 0;
      hy += SY + cz;
      const int cy = hy >= GY;
      hy -= cy ? GY : 0;
      hx += SX + cy;
    }
  };
  // LDS element offset of this thread's chunk j: cell (hx, hy, hz) at (hx GY + hy) RS + hz CS
  auto lds_off = [&](int j) {
    const int cell = q0 + STEP * j;
    return ((cell / GZ) * RS + (cell % GZ) * CS) + 8 * g;   // constant-folded per j
  };
  fetch(t_begin);
  // LDS operand of K step s for accumulator block mt: cell ((w + 1 + dx), (2 mt + m/16 +
  // 1 + dy), (m%16 + 1 + dz)), channel ch0 + 8 h; the tap of a step is lane-independent
  // base at the tap (-1, -1, -1) corner, so every offset below is a non-negative constant
  // (an LDS instruction's immediate offset is unsigned)
  const __bf16* tw = tile + (w * GY + (m >> 4)) * RS + (m & 15) * CS + 8 * h;
  auto a_at = [&](int s, int mt) {
    int tap, ch0;
    conv_step_k(CIN, (s * 16) / 32, (s * 16) % 32, tap, ch0);
    if (tap > 26) tap = 26;          // padded K: weights are zero there
    const int ox = tap / 9, oy = (tap / 3) % 3, oz = tap % 3;
    return *(const bf16x8*)(tw + (ox * GY + 2 * mt + oy) * RS + oz * CS + ch0);
  };
  constexpr int AD = 2;   // LDS operands fetched AD steps ahead
  for (int t = t_begin; t < t_end; ++t) {
    __syncthreads();  // every wave has finished reading the previous tile
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int c = threadIdx.x + 256 * j;
      if (c < NTOT) *(bf16x8*)(tile + lds_off(j)) = pf[j];
    }
    __syncthreads();
    // GELU gradient: the tile's pre-activations, loaded ahead of the next tile's halo so
    // that they arrive during the MFMAs and waiting for them does not wait for the halo
    // (CIN <= 16, two waves per SIMD: no registers to spare, loaded per row below)
    float pv[PF ? CT2Y / 2 : 1][2][8];
    if constexpr (PF) {
      const int z0 = (t % ntz) * CT2Z, y0 = ((t / ntz) % nty) * CT2Y, x0 = (t / (ntz * nty)) * CT2X;
      const int xc = x0 + w;
#pragma unroll
      for (int mt = 0; mt < CT2Y / 2; ++mt)
#pragma unroll
        for (int yq = 0; yq < 2; ++yq)
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int yc = y0 + 2 * mt + yq, zc = z0 + 4 * h + 8 * (j / 4) + (j % 4);
            pv[mt][yq][j] = (xc < nx && yc < ny && zc < nz) ? pre[(((long)xc * ny + yc) * nz + zc) * COUT + m] : 0.f;
          }
    }
    if (t + 1 < t_end) fetch(t + 1);  // in flight during this tile's MFMAs and epilogue
    f32x16 acc[CT2Y / 2];
#pragma unroll
    for (int a = 0; a < CT2Y / 2; ++a)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][i] = 0.f;
    bf16x8 aq[AD][CT2Y / 2];
#pragma unroll
    for (int s = 0; s < AD; ++s)
#pragma unroll
      for (int mt = 0; mt < CT2Y / 2; ++mt) aq[s][mt] = a_at(s < NK ? s : NK - 1, mt);
#pragma unroll
    for (int s = 0; s < NK; ++s) {
      bf16x8 a[CT2Y / 2];
#pragma unroll
      for (int mt = 0; mt < CT2Y / 2; ++mt) a[mt] = aq[s % AD][mt];
      if (s + AD < NK) {
#pragma unroll
        for (int mt = 0; mt < CT2Y / 2; ++mt) aq[s % AD][mt] = a_at(s + AD, mt);
      }
#pragma unroll
      for (int mt = 0; mt < CT2Y / 2; ++mt)
        acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mt], wr[s], acc[mt], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);  // keep the LDS prefetch distance written here
    }
    // epilogue: accumulator i of block mt is cell (x0 + w, y0 + 2 mt + i / 8, z0 + 4 h +
    // zo), zo = 8 ((i / 4) % 2) + i % 4, output channel m; each (x, y) row is addressed
    // from one base with constant offsets
    const int z0 = (t % ntz) * CT2Z, y0 = ((t / ntz) % nty) * CT2Y, x0 = (t / (ntz * nty)) * CT2X;
    const int xc = x0 + w;
    const bool zfull = z0 + CT2Z <= nz;
    if (xc < nx) {
#pragma unroll
      for (int mt = 0; mt < CT2Y / 2; ++mt) {
#pragma unroll
        for (int yq = 0; yq < 2; ++yq) {
          const int yc = y0 + 2 * mt + yq;
          if (yc >= ny) continue;  // uniform over the wave
          const long c0 = ((long)xc * ny + yc) * nz + z0 + 4 * h;
          float* pr = pre ? pre + c0 * COUT + m : nullptr;
          float prow[8];   // GELU gradient without the tile prefetch: this row's 8 values first
          if constexpr (GGRAD && !PF) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
              prow[j] = (z0 + 4 * h + 8 * (j / 4) + (j % 4) < nz) ? pr[(8 * (j / 4) + (j % 4)) * COUT] : 0.f;
          }
          auto out = [&](int j) {
            const int zo = 8 * (j / 4) + (j % 4);
            float v = acc[mt][8 * yq + j] + bn;
            if constexpr (GGRAD) {
              const float pa = PF ? pv[PF ? mt : 0][yq][j] : prow[j];
              v = m < cout_real ? v * gelu_t_grad(pa) : 0.f;   // padded channels: +0
            } else if (pr) {
              pr[zo * COUT] = v;
            }
            if (GELU) v = gelu_t(v);
            if (F32) {
              if (m < cout_real) Yf[(c0 + zo) * cout_real + m] = v;
            } else {
              Y[(c0 + zo) * COUT + m] = (__bf16)v;
            }
          };
          if (zfull) {
#pragma unroll
            for (int j = 0; j < 8; ++j) out(j);
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if (z0 + 4 * h + 8 * (j / 4) + (j % 4) < nz) out(j);
          }
        }
      }
    }
  }
}

// Persistent form of k_conv3_fwd for CIN <= 32 and 16 output channels (the net's output
// layer and the data gradient of its input layer), organised as k_conv3_fwd32p: a
// contiguous tile range per workgroup, every K step's weight fragment in registers
// (NSTEP x 4 VGPRs), the next tile's halo fetched during the current tile's MFMAs and
// epilogue.  Same MFMA layout and results as k_conv3_fwd.
template <int CIN, int MODE>
__global__ __launch_bounds__(256, 2) void k_conv3_fwdp(int nx, int ny, int nz, const __bf16* __restrict__ X,
                                                    const __bf16* __restrict__ Wp, const float* __restrict__ bias,
                                                    int cout_real, __bf16* __restrict__ Y, float* __restrict__ Yf,
                                                    float* __restrict__ pre) {
  constexpr int COUT = 16;
  static_assert(CIN <= 32, "persistent form: the weights of CIN <= 32 fit the register file");
  constexpr int NSTEP = ConvSteps<CIN>::NSTEP, CS = conv_cell_stride(CIN);
  constexpr int NCH = CIN / 8, NTOT = HX * HY * HZ * NCH, NPF = (NTOT + 255) / 256;  // 16-byte chunks
  constexpr bool GELU = (MODE & 1) != 0, F32 = (MODE & 2) != 0, CIRC = (MODE & 4) != 0;
  __shared__ __attribute__((aligned(16))) __bf16 tile[HX * HY * HZ * CS];
  const int ntz = (nz + CTZ - 1) / CTZ, nty = (ny + CTY - 1) / CTY, ntx = (nx + CTX - 1) / CTX;
  const int ntiles = ntz * nty * ntx;
  const int t_begin = (int)((long)blockIdx.x * ntiles / gridDim.x);
  const int t_end = (int)((long)(blockIdx.x + 1) * ntiles / gridDim.x);
  if (t_begin >= t_end) return;  // uniform over the workgroup
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 15, g = l >> 4;
  const int kl = 8 * g;  // this lane's first k within a step
  const float bn = bias[m];
  bf16x8 wr[NSTEP];
#pragma unroll
  for (int s = 0; s < NSTEP; ++s) wr[s] = *(const bf16x8*)(Wp + (long)m * (NSTEP * 32) + s * 32 + kl);
  constexpr int STEP = 256 / NCH, SZ = STEP % HZ, SY = (STEP / HZ) % HY, SX = STEP / (HZ * HY);
  const int gc = threadIdx.x % NCH, q0 = threadIdx.x / NCH;
  const int hz0 = q0 % HZ, hy0 = (q0 / HZ) % HY, hx0 = q0 / (HZ * HY);
  bf16x8 pf[NPF];
  auto fetch = [&](int t) {
    const int z0 = (t % ntz) * CTZ, y0 = ((t / ntz) % nty) * CTY, x0 = (t / (ntz * nty)) * CTX;
    int hz = hz0, hy = hy0, hx = hx0;
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int c = threadIdx.x + 256 * j;
      bf16x8 v;
      for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.0f;
      if (c < NTOT) {
        int gx = x0 + hx - 1, gy = y0 + hy - 1;
        const int gz = z0 + hz - 1;
        if (CIRC) {
          gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
          gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        }
        if ((unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz)
          v = *(const bf16x8*)(X + (((long)gx * ny + gy) * nz + gz) * CIN + 8 * gc);
      }
      pf[j] = v;
      hz += SZ;
      const int cz = hz >= HZ;
      hz -= cz ? HZ : 0;
      hy += SY + cz;
      const int cy = hy >= HY;
      hy -= cy ? HY : 0;
      hx += SX + cy;
    }
  };
  fetch(t_begin);
  // LDS operand of step s for y row mt: cell (w + 1 + dx, mt + 1 + dy, m + 1 + dz), channel
  // ch0 (tap and channel of the lane's k group, conv_step_k)
  // base at the tap (-1, -1, -1) corner: non-negative constant offsets when the tap of a
  // step is lane-independent (CIN = 32)
  const __bf16* tc = tile + ((w * HY) * HZ + m) * CS;
  auto a_at = [&](int s, int mt) {
    int tap, ch0;
    conv_step_k(CIN, s, kl, tap, ch0);
    if (tap > 26) tap = 26;              // padded K step: weights are zero there
    const int ox = tap / 9, oy = (tap / 3) % 3, oz = tap % 3;
    return *(const bf16x8*)(tc + ((ox * HY + mt + oy) * HZ + oz) * CS + ch0);
  };
  for (int t = t_begin; t < t_end; ++t) {
    __syncthreads();  // every wave has finished reading the previous tile
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int c = threadIdx.x + 256 * j;
      if (c < NTOT) *(bf16x8*)(tile + (q0 + STEP * j) * CS + 8 * gc) = pf[j];
    }
    __syncthreads();
    if (t + 1 < t_end) fetch(t + 1);  // in flight during this tile's MFMAs and epilogue
    f32x4 acc[CTY];
#pragma unroll
    for (int a = 0; a < CTY; ++a) acc[a] = (f32x4){0.f, 0.f, 0.f, 0.f};
    bf16x8 an[CTY];
#pragma unroll
    for (int mt = 0; mt < CTY; ++mt) an[mt] = a_at(0, mt);
#pragma unroll
    for (int s = 0; s < NSTEP; ++s) {
      bf16x8 a[CTY];
#pragma unroll
      for (int mt = 0; mt < CTY; ++mt) a[mt] = an[mt];
      if (s + 1 < NSTEP) {
#pragma unroll
        for (int mt = 0; mt < CTY; ++mt) an[mt] = a_at(s + 1, mt);
      }
#pragma unroll
      for (int mt = 0; mt < CTY; ++mt) acc[mt] = mfma16(a[mt], wr[s], acc[mt]);
      __builtin_amdgcn_sched_barrier(0);  // keep the LDS prefetch distance written here
    }
    // epilogue: accumulator i of row mt is cell (x0 + w, y0 + mt, z0 + 4 g + i), channel m
    const int z0 = (t % ntz) * CTZ, y0 = ((t / ntz) % nty) * CTY, x0 = (t / (ntz * nty)) * CTX;
    const int xc = x0 + w;
    if (xc < nx) {
#pragma unroll
      for (int mt = 0; mt < CTY; ++mt) {
        const int yc = y0 + mt;
        if (yc >= ny) continue;  // uniform over the wave
        const long c0 = ((long)xc * ny + yc) * nz + z0 + 4 * g;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (z0 + 4 * g + i < nz) {
            float v = acc[mt][i] + bn;
            if (pre) pre[(c0 + i) * COUT + m] = v;
            if (GELU) v = gelu_t(v);
            if (F32) {
              if (m < cout_real) Yf[(c0 + i) * cout_real + m] = v;
            } else {
              Y[(c0 + i) * COUT + m] = (__bf16)v;
            }
          }
        }
      }
    }
  }
}

template <int CIN, int MODE>
inline unsigned conv_fwdp_grid(int nx, int ny, int nz) {
  static int resident = 0;
  if (resident == 0) {
    int dev = 0, ncu = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_conv3_fwdp<CIN, MODE>, 256, 0);
    resident = (ncu > 0 ? ncu : 256) * (per > 0 ? per : 1);
  }
  const long ntiles = (long)((nz + CTZ - 1) / CTZ) * ((ny + CTY - 1) / CTY) * ((nx + CTX - 1) / CTX);
  return (unsigned)(ntiles < resident ? ntiles : resident);
}

// persistent grid of k_conv3_fwd32p: as many workgroups as the device holds at once
// (occupancy from the runtime), at most one per tile
template <int CIN, int MODE>
inline unsigned conv_fwd32p_grid(int nx, int ny, int nz) {
  static int resident = 0;
  if (resident == 0) {
    int dev = 0, ncu = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_conv3_fwd32p<CIN, MODE>, 256, 0);
    resident = (ncu > 0 ? ncu : 256) * (per > 0 ? per : 1);
  }
  const long ntiles = (long)((nz + CT2Z - 1) / CT2Z) * ((ny + CT2Y - 1) / CT2Y) * ((nx + CT2X - 1) / CT2X);
  return (unsigned)(ntiles < resident ? ntiles : resident);
}

// Weight gradient: dW[tap][c][n] += sum_p X[p + o(tap)][c] * G[p][n] over the cells of
// the grid (G: BF16 gradient of the pre-activation, COUT channels).  Workgroups loop
// over output tiles and store their FP32 partial sums (all 27 CIN COUT of them) in the
// slot of part that belongs to the workgroup; k_wgrad_reduce then adds the slots in
// workgroup order, so the result does not depend on the order the workgroups ran in.
//
// The reduction runs over cells, so an MFMA operand lane needs 8 consecutive cells of
// one channel: both tiles are transposed into channel rows in LDS, the X rows as z-runs
// of XR = 24 elements per (c, hx, hy): the CTZ interior cells at 8..8+CTZ-1, the lower
// halo at 7 and the upper halo at 0.  A tap's z offset dz then shifts the 8-cell window
// by one element, which is rebuilt from one aligned 16-byte read and one neighbouring
// dword with byte-align (the upper halo read for dz = +1 at the end of the run wraps to
// offset 0, where it is stored).
constexpr int XR = 24;

__device__ __forceinline__ bf16x8 xrun_window(const __bf16* __restrict__ row, int hz, int dz) {
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 c = *(const u32x4*)(row + 8 + hz);
  if (dz == 0) return __builtin_bit_cast(bf16x8, c);
  u32x4 r;
  if (dz < 0) {
    const unsigned lo = *(const unsigned*)(row + 6 + hz);  // elements 6+hz, 7+hz
    r[0] = __builtin_amdgcn_alignbyte(c[0], lo, 2);
    r[1] = __builtin_amdgcn_alignbyte(c[1], c[0], 2);
    r[2] = __builtin_amdgcn_alignbyte(c[2], c[1], 2);
    r[3] = __builtin_amdgcn_alignbyte(c[3], c[2], 2);
  } else {
    const unsigned hi = *(const unsigned*)(row + (16 + hz) % XR);  // element 16+hz (or the upper halo)
    r[0] = __builtin_amdgcn_alignbyte(c[1], c[0], 2);
    r[1] = __builtin_amdgcn_alignbyte(c[2], c[1], 2);
    r[2] = __builtin_amdgcn_alignbyte(c[3], c[2], 2);
    r[3] = __builtin_amdgcn_alignbyte(hi, c[3], 2);
  }
  return __builtin_bit_cast(bf16x8, r);
}

// Workgroup b stores its partial sums at part + b * (27 CIN COUT + COUT): the weights', then
// (with_bias) the bias gradient's, which rides along: once a G tile is in LDS, the 256 /
// COUT threads of each channel row add COUT consecutive cells of it, in order, to one
// running FP32 sum each; at the end the sums of a channel are added in thread order.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_wgrad(int nx, int ny, int nz, const __bf16* __restrict__ X,
                                                     const __bf16* __restrict__ G, float* __restrict__ part,
                                                     int with_bias) {
  static_assert(CTZ == 16 && XR == 24, "z-run layout assumes a 16-cell z tile");
  constexpr int MT = COUT / 16, NT = CIN / 16, TPW = 7;  // taps per wave: 7, 7, 7, 6
  constexpr int NCELL = CTX * CTY * CTZ;
  constexpr int TPC = 256 / COUT;   // bias: threads per channel row, COUT cells each
  static_assert(NCELL == 256 && NCELL % TPC == 0, "bias sums: one 256-cell tile per 256 threads");
  float bsum = 0.f;
  // channel rows padded to an odd multiple of 16 bytes (modulo 256): the 16 lanes of an
  // operand read take 16 different channel rows at one offset, which then fall on 16
  // distinct bank groups instead of 1 (G rows of 512 bytes) or 4 (X rows of 1728)
  constexpr int XS = HX * HY * XR + 8, GS = NCELL + 8;   // row strides in elements
  static_assert((XS * 2) % 32 == 16 && (GS * 2) % 32 == 16, "odd multiples of 16 bytes");
  __shared__ __attribute__((aligned(16))) __bf16 xt[CIN * XS];  // [c][hx][hy][z-run]
  __shared__ __attribute__((aligned(16))) __bf16 gt[COUT * GS];  // [n][cell]
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 15, g = l >> 4;
  f32x4 acc[TPW][MT][NT];
#pragma unroll
  for (int a = 0; a < TPW; ++a)
    for (int b = 0; b < MT; ++b)
      for (int c = 0; c < NT; ++c) acc[a][b][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int tz = (nz + CTZ - 1) / CTZ, ty = (ny + CTY - 1) / CTY, tx = (nx + CTX - 1) / CTX;
  const int ntiles = tx * ty * tz;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int z0 = (tile % tz) * CTZ, y0 = ((tile / tz) % ty) * CTY, x0 = (tile / (tz * ty)) * CTX;
    __syncthreads();
    // X halo tile: one 16-byte load = 8 channels of one cell, scattered to 8 channel rows
    // (neighbouring lanes take neighbouring cells, so their 2-byte stores share banks
    // only pairwise)
    for (int e = threadIdx.x; e < HX * HY * HZ * (CIN / 8); e += 256) {
      const int cell = e % (HX * HY * HZ), cg = e / (HX * HY * HZ);
      const int hz = cell % HZ, hy = (cell / HZ) % HY, hx = cell / (HZ * HY);
      const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
      bf16x8 v;
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) {
        v = *(const bf16x8*)(X + (((long)gx * ny + gy) * nz + gz) * CIN + 8 * cg);
      } else {
        for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
      }
      __bf16* dst = xt + (8 * cg) * XS + (hx * HY + hy) * XR + (hz == HZ - 1 ? 0 : 7 + hz);
#pragma unroll
      for (int j = 0; j < 8; ++j) dst[j * XS] = v[j];
    }
    // G tile: 8 channels of one cell per load, scattered to 8 channel rows
    for (int e = threadIdx.x; e < NCELL * (COUT / 8); e += 256) {
      const int cell = e % NCELL, ng = e / NCELL;
      const int hz = cell % CTZ, hy = (cell / CTZ) % CTY, hx = cell / (CTZ * CTY);
      const int gx = x0 + hx, gy = y0 + hy, gz = z0 + hz;
      bf16x8 v;
      if (gx < nx && gy < ny && gz < nz) {
        v = *(const bf16x8*)(G + (((long)gx * ny + gy) * nz + gz) * COUT + 8 * ng);
      } else {
        for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) gt[(8 * ng + j) * GS + cell] = v[j];
    }
    __syncthreads();
    if (with_bias) {   // uniform
      const __bf16* row = gt + (threadIdx.x / TPC) * GS + (threadIdx.x % TPC) * COUT;
#pragma unroll
      for (int r = 0; r < COUT / 8; ++r) {
        const bf16x8 v = *(const bf16x8*)(row + 8 * r);
#pragma unroll
        for (int j = 0; j < 8; ++j) bsum += (float)v[j];
      }
    }
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int tap = w + 4 * q;
      if (tap > 26) continue;
      const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
      for (int ks = 0; ks < NCELL / 32; ++ks) {
        const int k0 = ks * 32 + 8 * g;  // 8 consecutive cells along z of one (hx, hy) line
        const int hx = k0 / (CTY * CTZ), hy = (k0 / CTZ) % CTY, hz = k0 % CTZ;
        bf16x8 a[MT], b[NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) a[mt] = *(const bf16x8*)(gt + (mt * 16 + m) * GS + k0);
        const int rb = (hx + 1 + dx) * HY + (hy + 1 + dy);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) b[nt] = xrun_window(xt + (nt * 16 + m) * XS + rb * XR, hz, dz);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) acc[q][mt][nt] = mfma16(a[mt], b[nt], acc[q][mt][nt]);
      }
    }
  }
  float* dW = part + (long)blockIdx.x * (27 * CIN * COUT + COUT);
  if (with_bias) {
    float* red = reinterpret_cast<float*>(xt);   // the X tile is no longer read
    __syncthreads();
    red[threadIdx.x] = bsum;
    __syncthreads();
    if (threadIdx.x < COUT) {
      float s = 0.f;
      for (int i = 0; i < TPC; ++i) s += red[threadIdx.x * TPC + i];
      dW[27 * CIN * COUT + threadIdx.x] = s;
    }
  }
  // D[row = n][col = c]: lane holds rows 4g+i of the M-tile, column m of the N-tile; the
  // four waves' taps (w, w+4, ..) cover all 27, so every element of the slot is stored
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int tap = w + 4 * q;
    if (tap > 26) continue;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
      for (int nt = 0; nt < NT; ++nt)
        for (int i = 0; i < 4; ++i) {
          const int n = mt * 16 + 4 * g + i, c = nt * 16 + m;
          dW[((long)tap * CIN + c) * COUT + n] = acc[q][mt][nt][i];
        }
  }
}

// out[e] += sum over b < nb of part[b * stride + e], in order of b (e < n); the loads are
// issued 16 at a time ahead of their additions, so a thread waits for one memory round
// trip per 16 partial sums, not per partial sum
__global__ void k_partial_sum(int nb, int n, int stride, const float* __restrict__ part, float* __restrict__ out,
                              int n2, float* __restrict__ out2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n + n2) return;
  float s = 0.f;
  int b = 0;
  for (; b + 16 <= nb; b += 16) {
    float v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = part[(long)(b + q) * stride + e];
#pragma unroll
    for (int q = 0; q < 16; ++q) s += v[q];
  }
  for (; b < nb; ++b) s += part[(long)b * stride + e];
  if (e < n) {
    out[e] += s;
  } else {
    out2[e - n] += s;
  }
}

// channels n0 .. n0+7 of one FP32 row (zero from c_real on); vec: the row is 16-byte
// aligned (two 16-byte loads when all 8 channels are real)
__device__ __forceinline__ void load_row8(const float* __restrict__ src, int n0, int c_real, bool vec, float (&v)[8]) {
  if (vec && n0 + 8 <= c_real) {
    const float4 a = *(const float4*)(src + n0), b = *(const float4*)(src + n0 + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = n0 + j < c_real ? src[n0 + j] : 0.f;
  }
}

// k_gelu_bwd with one thread per 8 output channels of a cell (one 16-byte store), as
// k_to_bf16_cl8: nch = cp / 8, nchunk = ncell * nch < 2^31, gpre 16-byte aligned
// (vec bit 0: gpost rows 16-byte aligned, bit 1: pre rows)
__global__ void k_gelu_bwd8(unsigned nchunk, unsigned nch, int c_real, int cpre, const float* __restrict__ gpost,
                            int gstride, const float* __restrict__ pre, int gelu, int vec, __bf16* __restrict__ gpre) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nchunk) return;
  const unsigned p = t / nch;
  const int n0 = 8 * (int)(t - p * nch);
  float g[8], x[8];
  load_row8(gpost + (long)p * gstride, n0, c_real, (vec & 1) != 0, g);
  if (gelu) load_row8(pre + (long)p * cpre, n0, c_real, (vec & 2) != 0, x);
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float v = 0.f;
    if (n0 + j < c_real) {
      v = g[j];
      if (gelu) v *= gelu_t_grad(x[j]);
    }
    o[j] = (__bf16)v;
  }
  *(bf16x8*)(gpre + 8L * t) = o;
}

// g_pre = g_post * gelu'(pre) (or g_post when gelu is 0) -> BF16, cp channels (the
// channels >= c_real are zero)
__global__ void k_gelu_bwd(long ncell, int cp, int c_real, int cpre, const float* __restrict__ gpost, int gstride,
                           const float* __restrict__ pre, int gelu, __bf16* __restrict__ gpre) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * cp) return;
  long p = t / cp;
  int n = (int)(t % cp);
  float v = 0.f;
  if (n < c_real) {
    v = gpost[p * gstride + n];
    if (gelu) v *= gelu_t_grad(pre[p * cpre + n]);
  }
  gpre[t] = (__bf16)v;
}

// Pack FP32 W[tap][cin][cout] into BF16 rows Wp[n][s*32 + k] of a CINP -> COUTP
// convolution.  flip = 1 packs the transposed convolution of the data gradient:
// input channels = cout, output channels = cin, taps mirrored (tap -> 26 - tap).
__global__ void k_conv3_pack(const float* __restrict__ W, int cin, int cout, int cinp, int coutp, int nstep,
                             int flip, __bf16* __restrict__ Wp) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)coutp * nstep * 32) return;
  int n = (int)(t / (nstep * 32));
  int k = (int)(t % (nstep * 32));
  int s = k / 32, kl = k % 32;
  int tap, ch;
  conv_step_k(cinp, s, kl, tap, ch);
  float v = 0.f;
  if (tap < 27) {
    if (!flip) {
      if (ch < cin && n < cout) v = W[((long)tap * cin + ch) * cout + n];
    } else {  // input channel ch is a forward output channel, output n a forward input channel
      if (ch < cout && n < cin) v = W[((long)(26 - tap) * cin + n) * cout + ch];
    }
  }
  Wp[t] = (__bf16)v;
}

// The same with one thread per 8 output channels of a cell (one 16-byte store) and 32-bit
// indexing: nch = cp / 8 chunks per cell, nchunk = ncell * nch < 2^31, out 16-byte aligned;
// vec: in and cs keep every row 16-byte aligned.
__global__ void k_to_bf16_cl8(unsigned nchunk, unsigned nch, int c_real, int cs, int vec, const float* __restrict__ in,
                              __bf16* __restrict__ out) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nchunk) return;
  const unsigned p = t / nch;
  const int n0 = 8 * (int)(t - p * nch);
  float x[8];
  load_row8(in + (long)p * cs, n0, c_real, vec != 0, x);
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (__bf16)x[j];
  *(bf16x8*)(out + 8L * t) = v;
}

// FP32 channels-last (c_real channels, stride cs) -> BF16 padded to cp channels
__global__ void k_to_bf16_cl(long ncell, int c_real, int cs, int cp, const float* __restrict__ in,
                             __bf16* __restrict__ out) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * cp) return;
  long p = t / cp;
  int n = (int)(t % cp);
  out[t] = (__bf16)(n < c_real ? in[p * cs + n] : 0.f);
}

}  // namespace lex
