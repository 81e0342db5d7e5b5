// General FP32 3-D convolution for the reference's AutoEncoder (dl_models.py:33-68):
// strided 'SAME' convolutions, transposed convolutions and its dense layer, all as one
// correlation over an input that may be dilated and is zero-padded:
//
//   y[o, co] = act(bias[co] + sum_{t in k^3} sum_ci W[t, ci, co] x[p(o, t), ci])
//   p_d = o * stride - pad_lo + t      (per axis, in dilated-input coordinates)
//   the tap reads x[p_d / dil] when p_d >= 0, p_d % dil == 0 and p_d / dil < n_in,
//   and zero otherwise (padding, or a hole of the dilated input).
//
// flax's Conv with stride s and 'SAME' padding is dil = 1 with the TF/flax split of the
// total padding (pad_lo = total / 2); flax's ConvTranspose (transpose_kernel=False) is
// stride 1 over the input dilated by s with lax.conv_transpose's padding; the dense
// layer is k = 1 on the flattened (1, 1, 1, n) input.  Layout: channels last,
// x[(ix * ny + iy) * nz + iz][c]; W[t][ci][co] with t = (tx * k + ty) * k + tz (flax's
// (kx, ky, kz, ci, co) kernel).  One thread per output value, co fastest, so the
// threads of a wave read consecutive weights and share the input value.
#pragma once
#include "lex_common.hpp"

namespace lex {

struct ConvGeom {
  int nix, niy, niz, ci;
  int nox, noy, noz, co;
  int k, stride, dil;
  int plx, ply, plz;
};

__device__ __forceinline__ int conv_src(int o, int t, const ConvGeom& g, int pl, int n_in) {
  const int pd = o * g.stride - pl + t;
  if (pd < 0 || pd % g.dil != 0) return -1;
  const int p = pd / g.dil;
  return p < n_in ? p : -1;
}

__global__ void __launch_bounds__(256) k_conv3_f32(ConvGeom g, const float* __restrict__ x, const float* __restrict__ w,
                                                   const float* __restrict__ bias, int relu, float* __restrict__ y) {
  const long nout = (long)g.nox * g.noy * g.noz;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nout * g.co) return;
  const int c = (int)(idx % g.co);
  const long o = idx / g.co;
  const int oz = (int)(o % g.noz);
  const long r = o / g.noz;
  const int oy = (int)(r % g.noy);
  const int ox = (int)(r / g.noy);
  float acc = bias ? bias[c] : 0.0f;
  for (int tx = 0; tx < g.k; ++tx) {
    const int px = conv_src(ox, tx, g, g.plx, g.nix);
    if (px < 0) continue;
    for (int ty = 0; ty < g.k; ++ty) {
      const int py = conv_src(oy, ty, g, g.ply, g.niy);
      if (py < 0) continue;
      for (int tz = 0; tz < g.k; ++tz) {
        const int pz = conv_src(oz, tz, g, g.plz, g.niz);
        if (pz < 0) continue;
        const float* __restrict__ xp = x + (((long)px * g.niy + py) * g.niz + pz) * g.ci;
        const float* __restrict__ wp = w + ((long)((tx * g.k + ty) * g.k + tz) * g.ci) * g.co + c;
        for (int i = 0; i < g.ci; ++i) acc = fmaf(xp[i], wp[(long)i * g.co], acc);
      }
    }
  }
  y[o * g.co + c] = relu ? fmaxf(acc, 0.0f) : acc;
}

}  // namespace lex
