"""Learned SGS correction net on hand-written BF16 MFMA convolutions.

Architecture of simpleCNN (dl_models.py:19-30): four 3x3x3 'SAME' convolutions with bias,
in -> 32 -> 32 -> 32 -> out, the tanh-approximate GELU (flax nn.gelu) after the first three.
Here in = 11 (resolved u, v, w, theta', qv' and the six rate-of-strain components at cell
centres) and out = 5 (SGS tendencies of u, v, w, theta, qv at cell centres).

Activations are channels-last BF16 on the device; every convolution, its data and
weight gradients and the GELU backward run in liblexhip (lex_conv3d_*); weights are
FP32 masters packed to BF16 per call.  torch provides the tensors, autograd plumbing
and the optimiser only.
"""
from __future__ import annotations

import math

import torch

from . import _lib

CHANNELS = (11, 32, 32, 32, 5)


def _padc(c):
    return 16 if c <= 16 else (32 if c <= 32 else 64)


def _pack(w, flip):
    cin, cout = int(w.shape[1]), int(w.shape[2])
    nbytes = _lib.load().lex_conv3d_packed_weight_bytes(cin, cout, int(flip))
    wp = torch.empty(nbytes // 2, dtype=torch.bfloat16, device=w.device)
    _lib.call("lex_conv3d_pack_weights", _lib.ptr(w.contiguous()), cin, cout, int(flip), _lib.ptr(wp), _lib.stream())
    return wp


def to_bf16_channels_last(x, cp=None):
    """FP32 (..., c) channels-last -> BF16 (ncell, cp) with zero padding channels."""
    c = int(x.shape[-1])
    cp = _padc(c) if cp is None else cp
    xf = x.reshape(-1, c).contiguous().float()
    out = torch.empty((xf.shape[0], cp), dtype=torch.bfloat16, device=x.device)
    _lib.call("lex_to_bf16_channels_last", xf.shape[0], c, c, cp, _lib.ptr(xf), _lib.ptr(out), _lib.stream())
    return out


def conv3d_fwd(x_bf, shape, w, b, gelu, out_fp32, want_pre=False, circular=False):
    """One convolution: x_bf (ncell, padded cin) BF16 -> BF16 (ncell, padded cout) or FP32 (ncell, cout).
    circular: periodic x, y and zero z padding (else zero padding everywhere)."""
    entry = "lex_conv3d_bf16_fwd_circular" if circular else "lex_conv3d_bf16_fwd"
    nx, ny, nz = shape
    cin, cout = int(w.shape[1]), int(w.shape[2])
    cop = _padc(cout)
    ncell = nx * ny * nz
    wp = _pack(w, 0)
    bp = torch.zeros(cop, dtype=torch.float32, device=w.device)
    bp[:cout] = b
    pre = torch.empty((ncell, cop), dtype=torch.float32, device=w.device) if want_pre else None
    if out_fp32:
        yf = torch.empty((ncell, cout), dtype=torch.float32, device=w.device)
        _lib.call(entry, nx, ny, nz, cin, cout, _lib.ptr(x_bf), _lib.ptr(wp), _lib.ptr(bp), int(gelu), 1,
                  None, _lib.ptr(yf), _lib.ptr(pre), _lib.stream())
        return yf, pre
    y = torch.empty((ncell, cop), dtype=torch.bfloat16, device=w.device)
    _lib.call(entry, nx, ny, nz, cin, cout, _lib.ptr(x_bf), _lib.ptr(wp), _lib.ptr(bp), int(gelu), 0,
              _lib.ptr(y), None, _lib.ptr(pre), _lib.stream())
    return y, pre


def conv3d_bwd_data(g_pre_bf, shape, w):
    """dx (ncell, cin) FP32 from the BF16 pre-activation gradient (ncell, padded cout)."""
    nx, ny, nz = shape
    cin, cout = int(w.shape[1]), int(w.shape[2])
    wpf = _pack(w, 1)
    zb = torch.zeros(_padc(cin), dtype=torch.float32, device=w.device)
    dx = torch.empty((nx * ny * nz, cin), dtype=torch.float32, device=w.device)
    _lib.call("lex_conv3d_bf16_bwd_data", nx, ny, nz, cin, cout, _lib.ptr(g_pre_bf), _lib.ptr(wpf), _lib.ptr(zb),
              _lib.ptr(dx), _lib.stream())
    return dx


def fused_gelu_grad_ok(cin, cout):
    """Shapes the one-pass data gradient + GELU backward covers (17..32 output, <= 32 input channels)."""
    return _padc(cin) == 32 and cout <= 32


def conv3d_bwd_data_gelu(g_pre_bf, shape, w, pre_in):
    """BF16 (ncell, padded cin) = conv3d_bwd_data(g_pre_bf, shape, w) * gelu'(pre_in), in one
    pass; pre_in: the FP32 pre-activation (ncell, padded cin) of the layer below."""
    nx, ny, nz = shape
    cin, cout = int(w.shape[1]), int(w.shape[2])
    wpf = _pack(w, 1)
    zb = torch.zeros(_padc(cin), dtype=torch.float32, device=w.device)
    out = torch.empty((nx * ny * nz, _padc(cin)), dtype=torch.bfloat16, device=w.device)
    _lib.call("lex_conv3d_bf16_bwd_data_gelu", nx, ny, nz, cin, cout, _lib.ptr(g_pre_bf), _lib.ptr(wpf), _lib.ptr(zb),
              _lib.ptr(pre_in.contiguous()), _lib.ptr(out), _lib.stream())
    return out


def conv3d_bwd_weight(x_bf, g_pre_bf, shape, cin, cout):
    """(dW [27, cin, cout], db [cout]) FP32."""
    nx, ny, nz = shape
    dwp = torch.zeros((27, _padc(cin), _padc(cout)), dtype=torch.float32, device=x_bf.device)
    db = torch.zeros(cout, dtype=torch.float32, device=x_bf.device)
    _lib.call("lex_conv3d_bf16_bwd_weight", nx, ny, nz, cin, cout, _lib.ptr(x_bf), _lib.ptr(g_pre_bf), _lib.ptr(dwp),
              _lib.ptr(db), _lib.stream())
    return dwp[:, :cin, :cout].contiguous(), db


def gelu_bwd(g_post, gstride, pre, cp, c_real, gelu):
    ncell = g_post.shape[0]
    out = torch.empty((ncell, cp), dtype=torch.bfloat16, device=g_post.device)
    _lib.call("lex_gelu_bwd", ncell, cp, c_real, cp, _lib.ptr(g_post.contiguous()), gstride,
              _lib.ptr(pre) if gelu else None, int(gelu), _lib.ptr(out), _lib.stream())
    return out


def _net_forward(ctx, x, params, circular):
    nx, ny, nz, cin = x.shape
    shape = (nx, ny, nz)
    a = to_bf16_channels_last(x)
    acts, pres = [a], []
    nl = len(params) // 2
    out = None
    for li in range(nl):
        w, b = params[2 * li], params[2 * li + 1]
        last = li == nl - 1
        y, pre = conv3d_fwd(acts[-1], shape, w, b, gelu=not last, out_fp32=last, want_pre=not last,
                            circular=circular)
        if last:
            out = y
        else:
            acts.append(y)
            pres.append(pre)
    ctx.shape = shape
    ctx.save_for_backward(*params, *acts, *pres)
    ctx.nl = nl
    ctx.need_x = ctx.needs_input_grad[0]
    ctx.circular = circular
    return out.view(nx, ny, nz, -1)


class SgsNetFn(torch.autograd.Function):
    """x (nx, ny, nz, cin) FP32 -> (nx, ny, nz, cout) FP32 through the layers (zero padding)."""

    @staticmethod
    def forward(ctx, x, *params):
        return _net_forward(ctx, x, params, False)

    @staticmethod
    def backward(ctx, gy):
        if ctx.circular:
            raise NotImplementedError("the backward of the periodic-padding (Zcnn) layers is not built")
        nl = ctx.nl
        saved = ctx.saved_tensors
        params, acts, pres = saved[:2 * nl], saved[2 * nl:3 * nl], saved[3 * nl:]
        shape = ctx.shape
        g = gy.reshape(-1, gy.shape[-1]).contiguous().float()
        grads = [None] * (2 * nl)
        dx = None
        gp = None   # BF16 pre-activation gradient of layer li, when the layer above produced it
        for li in reversed(range(nl)):
            w = params[2 * li]
            cin, cout = int(w.shape[1]), int(w.shape[2])
            last = li == nl - 1
            if gp is None:
                gp = gelu_bwd(g, cout, None if last else pres[li], _padc(cout), cout, gelu=not last)
            dw, db = conv3d_bwd_weight(acts[li], gp, shape, cin, cout)
            grads[2 * li], grads[2 * li + 1] = dw, db
            if li > 0 and fused_gelu_grad_ok(cin, cout):
                # the data gradient and the GELU backward of layer li - 1 in one pass
                gp = conv3d_bwd_data_gelu(gp, shape, w, pres[li - 1])
            elif li > 0 or ctx.need_x:
                g = conv3d_bwd_data(gp, shape, w)
                gp = None
                if li == 0:
                    dx = g.view(*shape, cin)
        return (dx, *grads)


class CircularNetFn(SgsNetFn):
    """The same layers with periodic x, y and zero z padding (inference only)."""

    @staticmethod
    def forward(ctx, x, *params):
        return _net_forward(ctx, x, params, True)


class SgsNet(torch.nn.Module):
    """simpleCNN-shaped net with FP32 master weights W[27, cin, cout] and biases."""

    def __init__(self, channels=CHANNELS, seed=2031, device="cuda", circular=False):
        super().__init__()
        self.circular = bool(circular)
        gen = torch.Generator().manual_seed(seed)
        self.channels = tuple(channels)
        ws, bs = [], []
        for cin, cout in zip(channels[:-1], channels[1:]):
            # flax's default kernel init: lecun_normal (variance 1/fan_in), zero bias
            std = 1.0 / math.sqrt(27 * cin)
            ws.append(torch.nn.Parameter((torch.randn(27, cin, cout, generator=gen) * std).to(device)))
            bs.append(torch.nn.Parameter(torch.zeros(cout, device=device)))
        self.weights = torch.nn.ParameterList(ws)
        self.biases = torch.nn.ParameterList(bs)

    def params(self):
        out = []
        for w, b in zip(self.weights, self.biases):
            out += [w, b]
        return out

    def forward(self, x):
        fn = CircularNetFn if self.circular else SgsNetFn
        return fn.apply(x, *self.params())
