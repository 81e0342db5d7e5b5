"""BF16 MFMA convolutions of the SGS net against torch's conv3d on the CPU.

Kernel checks use a float64 CPU reference fed with the same BF16-rounded operands, so
the only difference is FP32 accumulation order: rel-L2 < 1e-5.  The whole net (BF16
activations and gradients) is checked against plain FP32 torch autograd with the
tolerance BF16 rounding allows: rel-L2 < 3e-2.  Grid sizes are ragged on purpose
(not multiples of the 4 x 4 x 16 tile).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (10, 6, 20)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _torch_w(W):
    """W[27, cin, cout] -> torch conv3d weight (cout, cin, 3, 3, 3)."""
    cin, cout = W.shape[1], W.shape[2]
    return W.reshape(3, 3, 3, cin, cout).permute(4, 3, 0, 1, 2).contiguous()


def _ncdhw(x):
    return x.permute(3, 0, 1, 2).unsqueeze(0).contiguous()


def _gelu(x):
    import torch
    return torch.nn.functional.gelu(x, approximate="tanh")


@pytest.mark.parametrize("cin,cout", [(11, 32), (32, 32), (32, 5), (5, 16)])
def test_conv_forward(cin, cout):
    import torch
    from lex_amd import sgs_net as sn
    g = torch.Generator().manual_seed(cin * 100 + cout)
    x = torch.randn(*SHAPE, cin, generator=g)
    W = torch.randn(27, cin, cout, generator=g) / np.sqrt(27 * cin)
    b = torch.randn(cout, generator=g) * 0.1
    xb, Wb = x.bfloat16().double(), W.bfloat16().double()
    ref_pre = torch.nn.functional.conv3d(_ncdhw(xb), _torch_w(Wb), padding=1)[0].permute(1, 2, 3, 0) + b.double()
    xd = sn.to_bf16_channels_last(x.cuda())
    for gelu in (0, 1):
        ref = _gelu(ref_pre) if gelu else ref_pre
        yf, pre = sn.conv3d_fwd(xd, SHAPE, W.cuda(), b.cuda(), gelu=gelu, out_fp32=True, want_pre=True)
        got = yf.cpu().double().reshape(*SHAPE, cout)
        assert _rel(got, ref) < 1e-5, (gelu, _rel(got, ref))
        assert _rel(pre.cpu().double()[:, :cout].reshape(*SHAPE, cout), ref_pre) < 1e-5
        yb, _ = sn.conv3d_fwd(xd, SHAPE, W.cuda(), b.cuda(), gelu=gelu, out_fp32=False)
        gotb = yb.cpu().double()[:, :cout].reshape(*SHAPE, cout)
        assert _rel(gotb, ref) < 8e-3  # one BF16 rounding of the output
        assert not bool(yb.cpu()[:, cout:].float().abs().max() > 0) if yb.shape[1] > cout else True


@pytest.mark.parametrize("cin,cout", [(11, 32), (32, 32), (32, 5)])
def test_conv_backward_kernels(cin, cout):
    import torch
    from lex_amd import sgs_net as sn
    g = torch.Generator().manual_seed(7 + cin + cout)
    x = torch.randn(*SHAPE, cin, generator=g)
    W = torch.randn(27, cin, cout, generator=g) / np.sqrt(27 * cin)
    gy = torch.randn(*SHAPE, cout, generator=g)
    xb, Wb, gb = x.bfloat16().double(), W.bfloat16().double(), gy.bfloat16().double()
    # references: input and weight gradients of conv3d (float64)
    dx_ref = torch.nn.grad.conv3d_input((1, cin) + SHAPE, _torch_w(Wb), _ncdhw(gb), padding=1)[0].permute(1, 2, 3, 0)
    dw_ref = torch.nn.grad.conv3d_weight(_ncdhw(xb), (cout, cin, 3, 3, 3), _ncdhw(gb), padding=1)
    dw_ref = dw_ref.permute(2, 3, 4, 1, 0).reshape(27, cin, cout)
    db_ref = gb.sum(dim=(0, 1, 2))
    xd = sn.to_bf16_channels_last(x.cuda())
    gd = sn.to_bf16_channels_last(gy.cuda())
    dx = sn.conv3d_bwd_data(gd, SHAPE, W.cuda()).cpu().double().reshape(*SHAPE, cin)
    assert _rel(dx, dx_ref) < 1e-5, _rel(dx, dx_ref)
    dw, db = sn.conv3d_bwd_weight(xd, gd, SHAPE, cin, cout)
    assert _rel(dw.cpu(), dw_ref) < 1e-5, _rel(dw.cpu(), dw_ref)
    assert _rel(db.cpu(), db_ref) < 1e-5, _rel(db.cpu(), db_ref)


@pytest.mark.parametrize("cin,cout,shape", [(32, 32, (13, 9, 37)), (32, 5, (10, 6, 20)), (20, 32, (24, 24, 16)),
                                            (32, 32, (40, 40, 48))], ids=["32-32", "32-5", "20-32", "32-32-big"])
def test_fused_gelu_gradient_equals_two_pass(cin, cout, shape):
    """conv3d_bwd_data_gelu (the data gradient with the GELU backward of the layer below in
    its epilogue) equals conv3d_bwd_data followed by gelu_bwd bit for bit, signed zeros
    and the zero padded channels (cin 20) included; 40x40x48 gives every workgroup of the
    persistent grid several tiles."""
    import torch
    from lex_amd import sgs_net as sn
    g = torch.Generator().manual_seed(11 * cin + cout)
    ncell = int(np.prod(shape))
    W = torch.randn(27, cin, cout, generator=g) / np.sqrt(27 * cin)
    gy = torch.randn(*shape, cout, generator=g)
    pre = torch.full((ncell, sn._padc(cin)), -1.5)   # gelu'(-1.5) < 0: an unmasked 0 * gelu' would be -0
    pre[:, :cin] = torch.randn(ncell, cin, generator=g) * 2.0
    gd = sn.to_bf16_channels_last(gy.cuda())
    pre_d, Wd = pre.cuda(), W.cuda()
    assert sn.fused_gelu_grad_ok(cin, cout)
    two = sn.gelu_bwd(sn.conv3d_bwd_data(gd, shape, Wd), cin, pre_d, sn._padc(cin), cin, gelu=True)
    one = sn.conv3d_bwd_data_gelu(gd, shape, Wd, pre_d)
    torch.cuda.synchronize()
    assert one.shape == two.shape == (ncell, sn._padc(cin))
    assert torch.equal(one.view(torch.int16), two.view(torch.int16))
    assert float(one.float().abs().max()) > 0


@pytest.mark.parametrize("c,cs,cp", [(11, 11, 16), (32, 32, 32), (5, 7, 16), (3, 3, 12), (20, 20, 64)])
def test_bf16_channels_last_conversion(c, cs, cp):
    """lex_to_bf16_channels_last equals torch's round-to-nearest-even BF16 conversion bit for
    bit, with zero padding channels; cp 12 takes the element-wise form, the others the
    8-channel-chunk form."""
    import torch
    from lex_amd import _lib
    g = torch.Generator().manual_seed(c * 7 + cp)
    ncell = 1237
    x = torch.randn(ncell, cs, generator=g) * 3.0
    out = torch.empty((ncell, cp), dtype=torch.bfloat16, device="cuda")
    _lib.call("lex_to_bf16_channels_last", ncell, c, cs, cp, _lib.ptr(x.cuda()), _lib.ptr(out), _lib.stream())
    ref = torch.zeros((ncell, cp), dtype=torch.bfloat16)
    ref[:, :c] = x[:, :c].bfloat16()
    assert torch.equal(out.cpu().view(torch.int16), ref.view(torch.int16))


@pytest.mark.parametrize("gelu", [0, 1])
def test_gelu_bwd_chunked_equals_elementwise(gelu):
    """lex_gelu_bwd's 8-channel-chunk form (cp 16) and its element-wise form (cp 12) give the
    same bits in the real channels and zeros in the padding."""
    import torch
    from lex_amd import _lib
    g = torch.Generator().manual_seed(5 + gelu)
    ncell, c = 1237, 11
    gpost = torch.randn(ncell, 13, generator=g).cuda()
    pre = (torch.randn(ncell, 16, generator=g) * 2.0).cuda()
    outs = {}
    for cp in (16, 12):
        out = torch.empty((ncell, cp), dtype=torch.bfloat16, device="cuda")
        _lib.call("lex_gelu_bwd", ncell, cp, c, 16, _lib.ptr(gpost), 13, _lib.ptr(pre) if gelu else None, gelu,
                  _lib.ptr(out), _lib.stream())
        outs[cp] = out.cpu()
    assert torch.equal(outs[16][:, :c].view(torch.int16), outs[12][:, :c].view(torch.int16))
    assert not bool(outs[16][:, c:].float().abs().max() > 0) and not bool(outs[12][:, c:].float().abs().max() > 0)
    assert float(outs[16].float().abs().max()) > 0


def test_net_forward_backward_vs_fp32_autograd():
    import torch
    from lex_amd.sgs_net import SgsNet
    net = SgsNet(seed=5)
    g = torch.Generator().manual_seed(9)
    x = torch.randn(*SHAPE, 11, generator=g)
    tgt = torch.randn(*SHAPE, 5, generator=g) * 0.1
    xd = x.cuda().requires_grad_(True)
    y = net(xd)
    loss = 0.5 * ((y - tgt.cuda()) ** 2).sum()
    loss.backward()
    # FP32 CPU reference with the same weights
    ws = [w.detach().cpu().clone().requires_grad_(True) for w in net.weights]
    bs = [b.detach().cpu().clone().requires_grad_(True) for b in net.biases]
    xr = x.clone().requires_grad_(True)
    h = _ncdhw(xr)
    for i, (w, b) in enumerate(zip(ws, bs)):
        h = torch.nn.functional.conv3d(h, _torch_w(w), b, padding=1)
        if i < 3:
            h = _gelu(h)
    yr = h[0].permute(1, 2, 3, 0)
    lr = 0.5 * ((yr - tgt) ** 2).sum()
    lr.backward()
    assert _rel(y.detach().cpu(), yr.detach()) < 3e-2
    for a, r in zip(net.weights, ws):
        assert _rel(a.grad.cpu(), r.grad) < 3e-2, _rel(a.grad.cpu(), r.grad)
    for a, r in zip(net.biases, bs):
        assert _rel(a.grad.cpu(), r.grad) < 3e-2, _rel(a.grad.cpu(), r.grad)
    assert _rel(xd.grad.cpu(), xr.grad) < 3e-2


def _ref_conv(xb, Wb, circular):
    import torch
    xin = _ncdhw(xb)
    if circular:  # periodic x and y, zero z (Zcnn's padding)
        xin = torch.nn.functional.pad(xin, (0, 0, 1, 1, 1, 1), mode="circular")
        return torch.nn.functional.conv3d(xin, _torch_w(Wb), padding=(0, 0, 1))[0].permute(1, 2, 3, 0)
    return torch.nn.functional.conv3d(xin, _torch_w(Wb), padding=1)[0].permute(1, 2, 3, 0)


@pytest.mark.parametrize("cin,cout,circular", [(32, 64, False), (64, 32, False), (64, 64, False), (5, 64, True),
                                               (64, 32, True), (32, 32, True), (32, 5, True)])
def test_conv_forward_64_channels_and_periodic_padding(cin, cout, circular):
    """64-channel layers (two K-steps per tap) and the periodic x/y padding of Zcnn's
    layers against torch conv3d in float64 on the same BF16-rounded operands."""
    import torch
    from lex_amd import sgs_net as sn
    g = torch.Generator().manual_seed(cin * 1000 + cout * 10 + int(circular))
    x = torch.randn(*SHAPE, cin, generator=g)
    W = torch.randn(27, cin, cout, generator=g) / np.sqrt(27 * cin)
    b = torch.randn(cout, generator=g) * 0.1
    ref = _gelu(_ref_conv(x.bfloat16().double(), W.bfloat16().double(), circular) + b.double())
    xd = sn.to_bf16_channels_last(x.cuda())
    yf, _ = sn.conv3d_fwd(xd, SHAPE, W.cuda(), b.cuda(), gelu=1, out_fp32=True, circular=circular)
    got = yf.cpu().double().reshape(*SHAPE, cout)
    assert _rel(got, ref) < 1e-5, _rel(got, ref)


@pytest.mark.parametrize("cin,cout,circular,out_fp32", [(32, 32, False, False), (11, 32, False, True),
                                                        (32, 5, True, True), (32, 32, True, True)])
def test_conv_forward_several_tiles_per_workgroup(cin, cout, circular, out_fp32):
    """The persistent forward kernels (16 or 32 output channels, at most 32 input
    channels) on a 70x36x40 grid: 486 output tiles of 4x4x16, more than the device's
    resident workgroups, so workgroups run several tiles each and fetch the next tile's
    halo while finishing the current one; the grid is ragged in x and z.  Against torch
    conv3d in float64 on the same BF16-rounded operands (BF16 output: within its
    rounding)."""
    import torch
    from lex_amd import sgs_net as sn
    shape = (70, 36, 40)
    g = torch.Generator().manual_seed(cin * 100 + cout + 7 * int(circular))
    x = torch.randn(*shape, cin, generator=g)
    W = torch.randn(27, cin, cout, generator=g) / np.sqrt(27 * cin)
    b = torch.randn(cout, generator=g) * 0.1
    ref = _gelu(_ref_conv(x.bfloat16().double(), W.bfloat16().double(), circular) + b.double())
    xd = sn.to_bf16_channels_last(x.cuda())
    if out_fp32:
        yf, _ = sn.conv3d_fwd(xd, shape, W.cuda(), b.cuda(), gelu=1, out_fp32=True, circular=circular)
        got = yf.cpu().double().reshape(*shape, cout)
        assert _rel(got, ref) < 1e-5, _rel(got, ref)
    else:
        yb, _ = sn.conv3d_fwd(xd, shape, W.cuda(), b.cuda(), gelu=1, out_fp32=False, circular=circular)
        got = yb.cpu().double().reshape(*shape, -1)[..., :cout]
        assert _rel(got, ref) < 4e-3, _rel(got, ref)


def test_zcnn_forward_vs_fp32():
    """The reference's Zcnn shape (13 layers, 64/32 channels, periodic x/y) against FP32
    torch with the same weights: the BF16 budget."""
    import torch
    from lex_amd import dl_models as dlm
    net = dlm.Zcnn(seed=3)
    g = torch.Generator().manual_seed(4)
    x = torch.randn(*SHAPE, 5, generator=g)
    with torch.no_grad():
        y = net(x.cuda()).cpu().double()
        h = x.double()
        nl = len(net.weights)
        for i, (w, b) in enumerate(zip(net.weights, net.biases)):
            h = _ref_conv(h, w.detach().cpu().double(), True) + b.detach().cpu().double()
            if i < nl - 1:
                h = _gelu(h)
    assert y.shape == h.shape
    assert _rel(y, h) < 3e-2, _rel(y, h)
