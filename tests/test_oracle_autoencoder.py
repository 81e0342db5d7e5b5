"""Pins of the oracle's AutoEncoder restatement (dl_models.py:33-68), on the CPU.

flax is absent, so its layer semantics are pinned by independent implementations:
  * strided 'SAME' convolution: a per-output loop, and torch's conv3d on the input padded
    with lax's split (smaller half low);
  * ConvTranspose (flax default transpose_kernel=False): a per-output loop over the
    input-dilated correlation, and torch's conv_transpose3d with the kernel flipped
    (its adjoint-of-convolution definition flips it), cropped to lax's padding;
  * the layer chain maps a 40^3 sample to 2^3 x 256 (the dense layer's 2048 inputs) and
    back to 40^3 x 5 ('VALID' on the second transposed layer: 2 -> 5).
Each pin is shown to fail under a planted mistake.
"""
from __future__ import annotations

import inspect
import textwrap

import numpy as np
import pytest
import torch

from oracle import lex_oracle as o


def _mutant(name, old, new):
    src = textwrap.dedent(inspect.getsource(getattr(o, name)))
    assert old in src, (name, old)
    ns = dict(vars(o))
    exec(compile(src.replace(old, new), f"<mutant {name}>", "exec"), ns)
    return ns[name]


def _loop_conv(x, W, b, stride, pads, dil):
    """Per-output loop of the dilated, padded, strided correlation (independent of the
    oracle's slicing)."""
    k = W.shape[0]
    n_in = x.shape[:3]
    n_dil = tuple((n - 1) * dil + 1 for n in n_in)
    no = tuple((n_dil[a] + pads[a][0] + pads[a][1] - k) // stride + 1 for a in range(3))
    out = np.zeros(no + (W.shape[4],))
    for o0 in range(no[0]):
        for o1 in range(no[1]):
            for o2 in range(no[2]):
                acc = b.copy()
                for t0 in range(k):
                    for t1 in range(k):
                        for t2 in range(k):
                            p = [oo * stride - pads[a][0] + t for a, (oo, t) in enumerate(((o0, t0), (o1, t1), (o2, t2)))]
                            if any(pp < 0 or pp % dil or pp // dil >= n_in[a] for a, pp in enumerate(p)):
                                continue
                            acc = acc + x[p[0] // dil, p[1] // dil, p[2] // dil] @ W[t0, t1, t2]
                out[o0, o1, o2] = acc
    return out


def _case(shape=(5, 6, 4), ci=2, co=3, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape + (ci,)), rng.normal(size=(3, 3, 3, ci, co)), rng.normal(size=co)


@pytest.mark.parametrize("shape", [(5, 6, 4), (3, 3, 3), (4, 7, 5)])
def test_strided_same_conv_vs_loop_and_torch(shape):
    x, W, b = _case(shape)
    pads = tuple(o.same_pads(n, 3, 2) for n in shape)
    got = o.conv3d_general(x, W, b, stride=2, pads=pads)
    assert got.shape[:3] == tuple(-(-n // 2) for n in shape)
    np.testing.assert_allclose(got, _loop_conv(x, W, b, 2, pads, 1), rtol=1e-12, atol=1e-12)
    xt = torch.from_numpy(x).permute(3, 0, 1, 2)[None]
    xt = torch.nn.functional.pad(xt, (pads[2][0], pads[2][1], pads[1][0], pads[1][1], pads[0][0], pads[0][1]))
    wt = torch.from_numpy(W).permute(4, 3, 0, 1, 2)
    ref = torch.nn.functional.conv3d(xt, wt, torch.from_numpy(b), stride=2)[0].permute(1, 2, 3, 0).numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("padding", ["SAME", "VALID"])
@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 4, 2), (5, 5, 5)])
def test_conv_transpose_vs_loop_and_torch(shape, padding):
    x, W, b = _case(shape, seed=3)
    pads = (o.transpose_pads(3, 2, padding),) * 3
    got = o.conv3d_general(x, W, b, stride=1, pads=pads, dil=2)
    n_out = tuple(2 * n if padding == "SAME" else 2 * n + 1 for n in shape)
    assert got.shape[:3] == n_out
    np.testing.assert_allclose(got, _loop_conv(x, W, b, 1, pads, 2), rtol=1e-12, atol=1e-12)
    # torch: conv_transpose3d(x, flip(W)) with stride 2 and no padding is the correlation
    # over the dilated input padded (2, 2); lax's 'SAME' is that minus the last output
    xt = torch.from_numpy(x).permute(3, 0, 1, 2)[None]
    wt = torch.from_numpy(np.ascontiguousarray(W[::-1, ::-1, ::-1])).permute(3, 4, 0, 1, 2)
    full = torch.nn.functional.conv_transpose3d(xt, wt, torch.from_numpy(b), stride=2)[0].permute(1, 2, 3, 0).numpy()
    ref = full[:n_out[0], :n_out[1], :n_out[2]]
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_layer_chain_shapes():
    sizes = [40]
    for kind, _, arg in o.AUTOENCODER_LAYERS[1:6]:
        lo, hi = o.same_pads(sizes[-1], 3, arg)
        sizes.append((sizes[-1] + lo + hi - 3) // 2 + 1)
    assert sizes == [40, 20, 10, 5, 3, 2]
    up = [1]
    for kind, _, arg in o.AUTOENCODER_LAYERS[7:12]:
        lo, hi = o.transpose_pads(3, 2, arg)
        up.append((up[-1] - 1) * 2 + 1 + lo + hi - 2)
    assert up == [1, 2, 5, 10, 20, 40]
    rng = np.random.default_rng(1)
    params, ci = [], 5
    for kind, f, _ in o.AUTOENCODER_LAYERS:
        fan = 2 * 2 * 2 * 256 if kind == "dense" else 27 * ci
        W = rng.normal(0, fan ** -0.5, (fan, f) if kind == "dense" else (3, 3, 3, ci, f))
        params.append((W, rng.normal(0, 0.01, f)))
        ci = f
    y = o.autoencoder_forward(params, rng.normal(size=(40, 40, 40, 5)))
    assert y.shape == (40, 40, 40, 5) and np.isfinite(y).all() and np.abs(y).max() > 0


def test_autoencoder_pins_bite():
    x, W, b = _case((5, 6, 4))
    # padding split the other way round (larger half low)
    mut = _mutant("same_pads", "return total // 2, total - total // 2", "return total - total // 2, total // 2")
    pads = tuple(mut(n, 3, 2) for n in (5, 6, 4))
    good = tuple(o.same_pads(n, 3, 2) for n in (5, 6, 4))
    assert pads != good
    assert not np.allclose(o.conv3d_general(x, W, b, stride=2, pads=pads),
                           _loop_conv(x, W, b, 2, good, 1), rtol=1e-6)
    # the transposed layer with the kernel flipped (transpose_kernel=True semantics)
    x, W, b = _case((3, 4, 2), seed=3)
    pads = (o.transpose_pads(3, 2, "SAME"),) * 3
    flipped = o.conv3d_general(x, np.ascontiguousarray(W[::-1, ::-1, ::-1]), b, stride=1, pads=pads, dil=2)
    assert not np.allclose(flipped, _loop_conv(x, W, b, 1, pads, 2), rtol=1e-6)
    # 'SAME' padding of the transposed layer taken as symmetric
    mut = _mutant("transpose_pads", "pad_a = k - 1 if s > k - 1 else int(math.ceil(pad_len / 2))",
                  "pad_a = pad_len // 2")
    assert mut(3, 2, "SAME") != o.transpose_pads(3, 2, "SAME")
