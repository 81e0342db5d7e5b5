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

namespace lex {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CTX = 4, CTY = 4, CTZ = 16;  // output tile of one workgroup (4 waves)
constexpr int HX = CTX + 2, HY = CTY + 2, HZ = CTZ + 2;

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// tanh-approximate GELU and its derivative
__device__ __forceinline__ float gelu_t(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  return 0.5f * x * (1.0f + tanhf(k0 * (x + k1 * x * x * x)));
}
__device__ __forceinline__ float gelu_t_grad(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  float t = tanhf(k0 * (x + k1 * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k0 * (1.0f + 3.0f * k1 * x * x);
}

template <int CIN>
struct ConvSteps {
  static constexpr int TPS = 32 / CIN;                  // taps per K-step
  static constexpr int NSTEP = (27 + TPS - 1) / TPS;   // K-steps
};

// Forward convolution.  mode bit 0: GELU after the bias; bit 1: FP32 output Yf with
// cout_real channels (else BF16 Y with COUT channels).  pre (optional): FP32
// pre-activation, COUT channels.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_fwd(int nx, int ny, int nz, const __bf16* __restrict__ X,
                                                   const __bf16* __restrict__ Wp, const float* __restrict__ bias,
                                                   int mode, int cout_real, __bf16* __restrict__ Y,
                                                   float* __restrict__ Yf, float* __restrict__ pre) {
  constexpr int TPS = ConvSteps<CIN>::TPS, NSTEP = ConvSteps<CIN>::NSTEP, NT = COUT / 16;
  __shared__ __attribute__((aligned(16))) __bf16 tile[HX * HY * HZ * CIN];
  const int z0 = blockIdx.x * CTZ, y0 = blockIdx.y * CTY, x0 = blockIdx.z * CTX;
  constexpr int nchunk = HX * HY * HZ * CIN / 8;
  for (int c = threadIdx.x; c < nchunk; c += 256) {
    int cell = c / (CIN / 8), g = c % (CIN / 8);
    int hz = cell % HZ, hy = (cell / HZ) % HY, hx = cell / (HZ * HY);
    int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
    bf16x8 v;
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) {
      v = *(const bf16x8*)(X + (((long)gx * ny + gy) * nz + gz) * CIN + 8 * g);
    } else {
      for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
    }
    *(bf16x8*)(tile + cell * CIN + 8 * g) = v;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 15, g = l >> 4;
  f32x4 acc[CTY][NT];
  for (int a = 0; a < CTY; ++a)
    for (int b = 0; b < NT; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < NSTEP; ++s) {
    const int kl = 8 * g;                // this lane's first k within the step
    int tap = s * TPS + kl / CIN;
    const int ch0 = kl % CIN;
    if (tap > 26) tap = 26;              // padded K-step: weights are zero there
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    bf16x8 bw[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      bw[nt] = *(const bf16x8*)(Wp + (long)(nt * 16 + m) * (NSTEP * 32) + s * 32 + kl);
#pragma unroll
    for (int mt = 0; mt < CTY; ++mt) {
      const int cell = ((w + 1 + dx) * HY + (mt + 1 + dy)) * HZ + (m + 1 + dz);
      bf16x8 a = *(const bf16x8*)(tile + cell * CIN + ch0);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma16(a, bw[nt], acc[mt][nt]);
    }
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

// Weight gradient: dW[tap][c][n] += sum_p X[p + o(tap)][c] * G[p][n] over the cells of
// the grid (G: BF16 gradient of the pre-activation, COUT channels).  Workgroups loop
// over output tiles and flush their FP32 partial sums with one atomic add per element.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_wgrad(int nx, int ny, int nz, const __bf16* __restrict__ X,
                                                     const __bf16* __restrict__ G, float* __restrict__ dW) {
  constexpr int MT = COUT / 16, NT = CIN / 16, TPW = 7;  // taps per wave: 7, 7, 7, 6
  __shared__ __attribute__((aligned(16))) __bf16 xt[CIN * HX * HY * HZ];        // [c][hx][hy][hz]
  __shared__ __attribute__((aligned(16))) __bf16 gt[COUT * CTX * CTY * CTZ];    // [n][cell]
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
    for (int e = threadIdx.x; e < HX * HY * HZ * CIN; e += 256) {
      int c = e % CIN, cell = e / CIN;
      int hz = cell % HZ, hy = (cell / HZ) % HY, hx = cell / (HZ * HY);
      int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
      __bf16 v = (__bf16)0.0f;
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) v = X[(((long)gx * ny + gy) * nz + gz) * CIN + c];
      xt[c * (HX * HY * HZ) + cell] = v;
    }
    for (int e = threadIdx.x; e < CTX * CTY * CTZ * COUT; e += 256) {
      int n = e % COUT, cell = e / COUT;
      int hz = cell % CTZ, hy = (cell / CTZ) % CTY, hx = cell / (CTZ * CTY);
      int gx = x0 + hx, gy = y0 + hy, gz = z0 + hz;
      __bf16 v = (__bf16)0.0f;
      if (gx < nx && gy < ny && gz < nz) v = G[(((long)gx * ny + gy) * nz + gz) * COUT + n];
      gt[n * (CTX * CTY * CTZ) + cell] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int tap = w + 4 * q;
      if (tap > 26) continue;
      const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
      for (int ks = 0; ks < (CTX * CTY * CTZ) / 32; ++ks) {
        const int k0 = ks * 32 + 8 * g;  // 8 consecutive cells along z of one (hx, hy) line
        const int hx = k0 / (CTY * CTZ), hy = (k0 / CTZ) % CTY, hz = k0 % CTZ;
        bf16x8 a[MT], b[NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) a[mt] = *(const bf16x8*)(gt + (mt * 16 + m) * (CTX * CTY * CTZ) + k0);
        const int hb = ((hx + 1 + dx) * HY + (hy + 1 + dy)) * HZ + (hz + 1 + dz);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const __bf16* src = xt + (nt * 16 + m) * (HX * HY * HZ) + hb;
#pragma unroll
          for (int j = 0; j < 8; ++j) b[nt][j] = src[j];
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) acc[q][mt][nt] = mfma16(a[mt], b[nt], acc[q][mt][nt]);
      }
    }
  }
  // D[row = n][col = c]: lane holds rows 4g+i of the M-tile, column m of the N-tile
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int tap = w + 4 * q;
    if (tap > 26) continue;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
      for (int nt = 0; nt < NT; ++nt)
        for (int i = 0; i < 4; ++i) {
          const int n = mt * 16 + 4 * g + i, c = nt * 16 + m;
          unsafeAtomicAdd(dW + ((long)tap * CIN + c) * COUT + n, acc[q][mt][nt][i]);
        }
  }
}

// bias gradient: db[n] += sum_p G[p][n]
__global__ void k_conv3_bgrad(long ncell, int cout, const __bf16* __restrict__ G, float* __restrict__ db) {
  const int n = blockIdx.y;
  float s = 0.f;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < ncell; p += (long)gridDim.x * blockDim.x)
    s += (float)G[p * cout + n];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(db + n, s);
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
  int tps = 32 / cinp;
  int tap = s * tps + kl / cinp, ch = kl % cinp;
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
