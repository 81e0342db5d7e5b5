"""The reference's post-step deep-learning correction: its network and data maps.

Mirrors dl_models.py of the reference for the pieces ws_rk3_sprint_dl
(one_sprint_functions.py:118-165) calls around the step:

* ``simpleCNN`` (dl_models.py:19-30): four 3x3x3 'SAME' convolutions with bias,
  5 -> 32 -> 32 -> 32 -> 5 channels, tanh-GELU between them, on the BF16 MFMA kernels
  of liblexhip (``lex_amd.sgs_net.SgsNet`` with five input channels).
* ``phys_state_tuple2array`` (:125-155) and ``correct_array2tuple`` (:158-163, with
  ``stagger_uvw``): ``lex_dl_pack`` and ``lex_dl_stagger``.

The flax parameters of the reference become the net object itself (``dl_params``).
``Zcnn`` (:71-122) runs on the same kernels with periodic x/y padding (inference).
``AutoEncoder`` (:33-68, the reference's default model, namelist_n_constants.py:78) runs
in FP32 on ``lex_conv3d_f32``: strided 'SAME' convolutions, the dense layer and the
transposed convolutions (inference).
"""
from __future__ import annotations

import torch

from . import _lib
from .sgs_net import SgsNet


def simpleCNN(seed=2031, device="cuda"):  # noqa: N802 - the reference's class name
    """The reference's simpleCNN shape on the MFMA kernels: 5 -> 32 -> 32 -> 32 -> 5."""
    return SgsNet(channels=(5, 32, 32, 32, 5), seed=seed, device=device)


ZCNN_CHANNELS = (5, 64, 32, 32, 32, 64, 32, 32, 32, 64, 32, 32, 32, 5)


def Zcnn(seed=2031, device="cuda"):  # noqa: N802 - the reference's class name
    """Zcnn (dl_models.py:71-122): thirteen 3x3x3 convolutions, 5 -> 64 -> 32 -> 32 -> 32 ->
    64 -> 32 -> 32 -> 32 -> 64 -> 32 -> 32 -> 32 -> 5, GELU between them, periodic in x and y;
    the zero planes the reference concatenates above and below the column and resets
    after every layer are zero padding in z.  Inference only (the 64-channel weight
    gradient is not built)."""
    return SgsNet(channels=ZCNN_CHANNELS, seed=seed, device=device, circular=True)


# (kind, features, stride or padding) of dl_models.py:36-66; ReLU after all but the last
AUTOENCODER_LAYERS = (("conv", 64, 1), ("conv", 32, 2), ("conv", 32, 2), ("conv", 64, 2), ("conv", 128, 2),
                      ("conv", 256, 2), ("dense", 512, None), ("convT", 256, "SAME"), ("convT", 128, "VALID"),
                      ("convT", 64, "SAME"), ("convT", 32, "SAME"), ("convT", 32, "SAME"), ("conv", 64, 1),
                      ("conv", 5, 1))


def _same_pads(n, k, s):
    """'SAME' padding (lo, hi) of a strided window: ceil(n / s) outputs, the smaller half
    of the total padding low (lax.padtype_to_pads)."""
    out = -(-n // s)
    total = max((out - 1) * s + k - n, 0)
    return total // 2, total - total // 2


def _transpose_pads(k, s, padding):
    """Padding of a transposed convolution's input-dilated correlation (lax.conv_transpose)."""
    pad_len = k + s - 2 + (max(k - s, 0) if padding == "VALID" else 0)
    if padding == "VALID" or s > k - 1:
        pad_a = k - 1
    else:
        pad_a = -(-pad_len // 2)
    return pad_a, pad_len - pad_a


class AutoEncoder(torch.nn.Module):  # noqa: N801 - the reference's class name
    """AutoEncoder (dl_models.py:33-68) on one (40, 40, 40, 5) sample: a stride-1 and five
    stride-2 'SAME' convolutions (5 -> 64 -> 32 -> 32 -> 64 -> 128 -> 256, 40^3 -> 2^3),
    flatten, Dense 2048 -> 512, five stride-2 transposed convolutions (512 -> 256 -> 128
    -> 64 -> 32 -> 32, 1 -> 2 -> 5 -> 10 -> 20 -> 40; the second 'VALID'), two stride-1
    convolutions (32 -> 64 -> 5); ReLU after every layer but the last.  FP32 on
    lex_conv3d_f32, inference only.  Kernels in flax's layout (3, 3, 3, ci, co), the
    dense kernel (2048, 512); random normal(0, 1/sqrt(fan_in)) weights and zero biases
    stand in for trained parameters (flax's initialisers are not restated)."""

    SIZE = 40   # the dense layer fixes the sample to 40^3 (dl_models.py:49-52)

    def __init__(self, seed=2031, device="cuda"):
        super().__init__()
        gen = torch.Generator(device="cpu").manual_seed(seed)
        ws, bs = [], []
        ci = 5
        for kind, f, _ in AUTOENCODER_LAYERS:
            shape = (2 * 2 * 2 * ci, f) if kind == "dense" else (3, 3, 3, ci, f)
            fan_in = shape[0] if kind == "dense" else 27 * ci
            w = torch.randn(shape, generator=gen) * fan_in ** -0.5
            ws.append(torch.nn.Parameter(w.to(device).contiguous(), requires_grad=False))
            bs.append(torch.nn.Parameter(torch.zeros(f, device=device), requires_grad=False))
            ci = f
        self.weights = torch.nn.ParameterList(ws)
        self.biases = torch.nn.ParameterList(bs)

    def forward(self, x):
        n = self.SIZE
        h = x.reshape(x.shape[-4:]).contiguous().float()
        if tuple(h.shape) != (n, n, n, 5):
            raise ValueError(f"AutoEncoder takes one (40, 40, 40, 5) sample, got {tuple(x.shape)}")
        last = len(AUTOENCODER_LAYERS) - 1
        for q, ((kind, f, arg), w, b) in enumerate(zip(AUTOENCODER_LAYERS, self.weights, self.biases)):
            ni = tuple(h.shape[:3])
            ci = h.shape[3]
            if kind == "conv":
                pads = [_same_pads(m, 3, arg) for m in ni]
                no = [(m + lo + hi - 3) // arg + 1 for m, (lo, hi) in zip(ni, pads)]
                k, stride, dil = 3, arg, 1
            elif kind == "convT":
                pads = [_transpose_pads(3, 2, arg)] * 3
                no = [(m - 1) * 2 + 1 + lo + hi - 2 for m, (lo, hi) in zip(ni, pads)]
                k, stride, dil = 3, 1, 2
            else:   # flatten (channels last, as the reference's reshape) and Dense
                ci = h.numel()
                ni, pads, no, k, stride, dil = (1, 1, 1), [(0, 0)] * 3, [1, 1, 1], 1, 1, 1
            y = torch.empty((*no, f), dtype=torch.float32, device=h.device)
            _lib.call("lex_conv3d_f32", _lib.ptr(h), ni[0], ni[1], ni[2], ci, _lib.ptr(w), _lib.ptr(b), k, stride,
                      dil, pads[0][0], pads[1][0], pads[2][0], int(q != last), _lib.ptr(y), no[0], no[1], no[2], f,
                      _lib.stream())
            h = y
        return h


def level_norm(scaling_params):
    """(lo, den), each (5 nz + 1,) FP32: per-level offsets and denominators of the
    normalisation in phys_state_tuple2array (dl_models.py:127-151), concatenated for
    theta, u, v (nz levels), w (nz + 1) and qv (nz)."""
    los, dens = [], []
    for q in range(5):
        mn, mx = scaling_params[2 * q], scaling_params[2 * q + 1]
        mn, mx = torch.as_tensor(mn).float().squeeze(), torch.as_tensor(mx).float().squeeze()
        lo = torch.amin(mn, dim=(0, 1))
        den = torch.amax(mx, dim=(0, 1)) - lo + (1.0e-5 if q == 4 else 0.01)
        los.append(lo)
        dens.append(den)
    return torch.cat(los).contiguous(), torch.cat(dens).contiguous()


def scaling_ranges(scaling_params, device):
    """The scaling maximum minus minimum of theta, u, v, w, qv on their own points."""
    out = []
    for q in range(5):
        mn, mx = scaling_params[2 * q], scaling_params[2 * q + 1]
        r = (torch.as_tensor(mx).float() - torch.as_tensor(mn).float()).squeeze()
        out.append(r.to(device).contiguous())
    return tuple(out)


def _grid(phys_state):
    th = phys_state[0]
    return _lib.grid(th.shape[0] - 6, th.shape[1] - 6, th.shape[2] - 2, 1.0, 1.0, 1.0)


def phys_state_tuple2array(phys_state, scaling_params, norm=None):
    """(1, nx, ny, nz, 5) FP32 network input (dl_models.py:125-155)."""
    th, u, v, w, _, qv = phys_state[:6]
    g = _grid(phys_state)
    lo, den = norm if norm is not None else level_norm(scaling_params)
    lo, den = lo.to(th.device), den.to(th.device)
    out = torch.empty((g.nx, g.ny, g.nz, 5), dtype=torch.float32, device=th.device)
    _lib.call("lex_dl_pack", g, *[_lib.ptr(a) for a in (th, u, v, w, qv, lo, den, out)], _lib.stream())
    return out.view(1, g.nx, g.ny, g.nz, 5)


def correct_array2tuple(correct):
    """(nx, ny, nz, 5) correction -> (theta, u, v, w, qv) corrections on their own points
    (dl_models.py:158-163 with stagger_uvw, :174-182)."""
    c = correct.reshape(correct.shape[-4:]).contiguous().float()
    nx, ny, nz = c.shape[:3]
    g = _lib.grid(nx, ny, nz, 1.0, 1.0, 1.0)
    dev = c.device
    tc = torch.empty((nx, ny, nz), dtype=torch.float32, device=dev)
    uc = torch.empty((nx + 1, ny, nz), dtype=torch.float32, device=dev)
    vc = torch.empty((nx, ny + 1, nz), dtype=torch.float32, device=dev)
    wc = torch.empty((nx, ny, nz + 1), dtype=torch.float32, device=dev)
    qc = torch.empty_like(tc)
    _lib.call("lex_dl_stagger", g, *[_lib.ptr(a) for a in (c, tc, uc, vc, wc, qc)], _lib.stream())
    return tc, uc, vc, wc, qc


def apply_correction(phys_state, correct, ranges):
    """In place: theta, u, v, w, qv += staggered correction x scaling range, ghosts
    refilled (one_sprint_functions.py:137-160)."""
    th, u, v, w, _, qv = phys_state[:6]
    g = _grid(phys_state)
    c = correct.reshape(g.nx, g.ny, g.nz, 5).contiguous().float()
    _lib.call("lex_dl_apply", g, _lib.ptr(c), *[_lib.ptr(r) for r in ranges],
              *[_lib.ptr(a) for a in (th, u, v, w, qv)], _lib.stream())
    return phys_state
