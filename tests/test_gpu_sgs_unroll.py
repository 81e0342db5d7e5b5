"""Training the SGS net through the solver: parity of the hand-written reverse pieces.

* lex_sgs_apply_vjp and lex_sgs_features_vjp (+ lex_fold) are the transposes of linear
  maps: dot-product identities <J v, w> = <v, J^T w> on random vectors (FP32 rounding
  tolerance written in each test).
* lex_rk3_step_vjp_taped_ext_sgs: adjoints of the state and of the caller's frozen SGS
  tendencies against FP64 central differences of the oracle step with the same SGS
  (ws_rk3_big_step_ext_sgs).
* The assembled unroll gradient with respect to the net's last-layer bias against
  central differences of the GPU unroll itself (the bias enters in FP32; the last layer's
  weights are zero, so the BF16 inputs of later steps do not feed back).
"""
import ctypes

import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import f32, f64, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _dot(a, b):
    return float(np.sum(np.asarray(a, np.float64) * np.asarray(b, np.float64)))


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-30)


def _step_setup(nx=16, ny=16, nz=16, seed=3):
    from lex_amd.setup_lex import build_case, config2
    case, opts = config2()
    case = case.scaled(nx, ny, nz)
    phys, base, grids, _ = build_case(case)
    og = o.Grid(nx, ny, nz, case.dx, case.dy, case.dz)
    oo = o.Options(dt=opts["dt"], solver_opt=2, n_sound=opts["n_sound"], turb_opt=0)
    rng = np.random.default_rng(seed)
    st = [from_dev(a) for a in phys]
    st[1] = o.padding3u_array(np.concatenate([ui := rng.normal(0, 1.0, (nx, ny, nz)), ui[:1]], 0))
    st[2] = o.padding3v_array(np.concatenate([vi := rng.normal(0, 1.0, (nx, ny, nz)), vi[:, :1]], 1))
    st[3] = o.set_w_bc(o.padding3_array(rng.normal(0, 0.5, (nx, ny, nz + 1))))
    st = [f64(f32(a)) for a in st]
    return case, opts, og, oo, st, base, grids


def _sgs_shapes(nx, ny, nz):
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz), (nx, ny, nz)]


def test_sgs_apply_vjp_dot():
    import torch
    from lex_amd import _lib, adjoint as adj
    nx, ny, nz = 12, 10, 8
    g = _lib.grid(nx, ny, nz, 200.0, 200.0, 200.0)
    rng = np.random.default_rng(1)
    y = rng.normal(size=(nx, ny, nz, 5))
    lam = [rng.normal(size=s) for s in _sgs_shapes(nx, ny, nz)]
    osc = (ctypes.c_float * 5)(1.5, 0.5, 2.0, 0.25, 3.0)
    outs = [torch.empty(s, dtype=torch.float32, device="cuda") for s in _sgs_shapes(nx, ny, nz)]
    y_d = to_dev(y)   # held: a pointer to a temporary would dangle
    _lib.call("lex_sgs_apply", g, _lib.ptr(y_d), osc, *[_lib.ptr(t) for t in outs], _lib.stream())
    ly = adj.sgs_apply_vjp(g, tuple(to_dev(a) for a in lam), osc)
    lhs = sum(_dot(from_dev(t), l) for t, l in zip(outs, lam))
    rhs = _dot(y, from_dev(ly))
    assert _rel(rhs, lhs) < 1e-5, (lhs, rhs)


def test_sgs_features_vjp_dot():
    import torch
    from lex_amd import _lib, adjoint as adj
    nx, ny, nz = 12, 10, 8
    g = _lib.grid(nx, ny, nz, 200.0, 250.0, 150.0)
    rng = np.random.default_rng(2)
    # halo-consistent direction: padded from interior fields as the forward stores pad
    di = [rng.normal(size=(nx, ny, nz)), rng.normal(size=(nx + 1, ny, nz)), rng.normal(size=(nx, ny + 1, nz)),
          rng.normal(size=(nx, ny, nz + 1)), np.zeros((nx, ny, nz)), rng.normal(size=(nx, ny, nz))]
    dp = [o.padding3_array(di[0]), o.padding3u_array(di[1]), o.padding3v_array(di[2]), o.padding3_array(di[3]),
          o.padding3_array(di[4]), o.padding3_array(di[5])]
    th0 = to_dev(300.0 + np.linspace(0, 5, nz + 2))
    qv0 = to_dev(np.linspace(0.01, 0.002, nz + 2))
    shift = (ctypes.c_float * 11)(*([0.0] * 11))
    scale = (ctypes.c_float * 11)(*[0.5 + 0.1 * c for c in range(11)])

    def feats(state):
        out = torch.empty((nx, ny, nz, 11), dtype=torch.float32, device="cuda")
        _lib.call("lex_sgs_features", g, _lib.ptr(th0), _lib.ptr(qv0), _lib.ptr(state[1]), _lib.ptr(state[2]),
                  _lib.ptr(state[3]), _lib.ptr(state[0]), _lib.ptr(state[5]), shift, scale, _lib.ptr(out),
                  _lib.stream())
        return from_dev(out)

    d_dev = [to_dev(a) for a in dp]
    z_dev = [to_dev(np.zeros_like(a)) for a in dp]
    jd = feats(d_dev) - feats(z_dev)  # affine map: the constant part cancels
    lf = rng.normal(size=(nx, ny, nz, 11))
    lam = tuple(torch.zeros_like(t) for t in z_dev)
    adj.sgs_features_vjp(g, to_dev(lf), scale, lam)
    gi = [from_dev(lam[0])[3:-3, 3:-3, 1:-1], from_dev(lam[1])[3:-3, 3:-3, 1:-1],
          from_dev(lam[2])[3:-3, 3:-3, 1:-1], from_dev(lam[3])[3:-3, 3:-3, 1:-1], None,
          from_dev(lam[5])[3:-3, 3:-3, 1:-1]]
    lhs = _dot(jd, lf)
    rhs = sum(_dot(a, b) for a, b in zip(di, gi) if b is not None)
    # theta enters as theta - theta0 (theta0 ~ 300): FP32 rounding of that difference
    assert _rel(rhs, lhs) < 2e-4, (lhs, rhs)


def test_network_and_feature_gradients_are_bit_reproducible():
    """The learned-SGS gradients use no atomics: the weight and bias gradients of a
    convolution are per-workgroup partial sums added in workgroup order, and the feature
    reverse gathers per u, v, w point.  Two calls on the same inputs agree bit for bit
    (on a grid large enough that float atomics would reorder their additions)."""
    import torch
    from lex_amd import _lib, adjoint as adj, sgs_net as sn
    nx, ny, nz = 48, 40, 32
    gen = torch.Generator(device="cuda").manual_seed(17)
    x = torch.randn(nx * ny * nz, 32, device="cuda", generator=gen).to(torch.bfloat16)
    gp = torch.randn(nx * ny * nz, 32, device="cuda", generator=gen).to(torch.bfloat16)
    r1 = sn.conv3d_bwd_weight(x, gp, (nx, ny, nz), 32, 32)
    r2 = sn.conv3d_bwd_weight(x, gp, (nx, ny, nz), 32, 32)
    assert torch.equal(r1[0], r2[0]) and torch.equal(r1[1], r2[1])
    assert float(r1[0].abs().max()) > 0 and float(r1[1].abs().max()) > 0
    g = _lib.grid(nx, ny, nz, 200.0, 250.0, 150.0)
    scale = (ctypes.c_float * 11)(*[0.5 + 0.1 * c for c in range(11)])
    lf = torch.randn(nx, ny, nz, 11, device="cuda", generator=gen)
    shapes = [(nx + 6, ny + 6, nz + 2), (nx + 7, ny + 6, nz + 2), (nx + 6, ny + 7, nz + 2), (nx + 6, ny + 6, nz + 3),
              (nx + 6, ny + 6, nz + 2), (nx + 6, ny + 6, nz + 2)]
    runs = []
    for _ in range(2):
        lam = tuple(torch.zeros(sh, dtype=torch.float32, device="cuda") for sh in shapes)
        runs.append(adj.sgs_features_vjp(g, lf, scale, lam))
    for q in (0, 1, 2, 3, 5):
        assert torch.equal(runs[0][q], runs[1][q]), q
        assert float(runs[0][q].abs().max()) > 0, q


def test_ext_sgs_step_vjp_vs_fd():
    from lex_amd import adjoint as adj
    from lex_amd._core import grid_from
    nx = ny = nz = 16
    case, opts, og, oo, st, base, grids = _step_setup(nx, ny, nz)
    bs = tuple(from_dev(a) for a in base[:5]) + (None,)
    rng = np.random.default_rng(9)
    sgs = [rng.normal(0, 1e-3, s) for s in _sgs_shapes(nx, ny, nz)]
    sgs[2][:, :, 0] = 0.0
    sgs[2][:, :, -1] = 0.0
    sgs = [f64(f32(a)) for a in sgs]

    def F(xs, ss):
        return o.ws_rk3_big_step_ext_sgs(tuple(xs), bs, og, oo, tuple(ss))[0]

    out = F(st, sgs)
    lam = [rng.normal(size=a.shape) for a in out]
    lam[4] = lam[4] * 1e3
    g = grid_from(base[1], *grids[3:6])
    dev = tuple(to_dev(a) for a in st)
    tape, y = adj.step_tape_ext_sgs(dev, tuple(to_dev(a) for a in sgs), base, g, opts["dt"], opts["n_sound"])
    # the recorded step is the forward step with these SGS tendencies
    for q in (0, 1, 3):
        d_gpu, d_ref = from_dev(y[q]) - st[q], out[q] - st[q]
        assert np.linalg.norm(d_gpu - d_ref) < 1e-3 * np.linalg.norm(d_ref) + 1e-12
    li, lsgs = adj.step_vjp_taped_ext_sgs(dev, tape, tuple(to_dev(a) for a in lam), base, g, opts["dt"],
                                          opts["n_sound"])
    eps = 1e-3
    for c, name in enumerate(["su", "sv", "sw", "s_theta", "s_qv"]):
        v = [np.zeros_like(a) for a in sgs]
        v[c] = rng.normal(0, 1e-3, sgs[c].shape)
        fp = F(st, [a + eps * b for a, b in zip(sgs, v)])
        fm = F(st, [a - eps * b for a, b in zip(sgs, v)])
        fd = sum(_dot(l, (p - m) / (2 * eps)) for l, p, m in zip(lam, fp, fm))
        got = _dot(from_dev(lsgs[c]), v[c])
        assert _rel(got, fd) < 2e-3, (name, got, fd)
    # and the state adjoint with these SGS tendencies (theta direction)
    vi = rng.normal(0, 0.3, (nx, ny, nz))
    vq = [np.zeros_like(a) for a in st]
    vq[0] = o.padding3_array(vi)
    fp = F([a + 1e-6 * b for a, b in zip(st, vq)], sgs)
    fm = F([a - 1e-6 * b for a, b in zip(st, vq)], sgs)
    fd = sum(_dot(l, (p - m) / 2e-6) for l, p, m in zip(lam, fp, fm))
    assert _rel(_dot(from_dev(li[0]), vq[0]), fd) < 2e-3


def test_unroll_gradient_last_bias_vs_fd():
    import torch
    from lex_amd.sgs_net import SgsNet
    from lex_amd.sgs_train import CoarseModel
    from lex_amd.sgs_unroll import OUT_SCALE, NetSolverUnroll, field_loss
    from lex_amd.setup_lex import config2
    case, opts = config2()
    case = case.scaled(16, 16, 16)
    model = CoarseModel(case, opts["dt"], opts["n_sound"])
    for c in range(5):
        model.out_scale[c] = OUT_SCALE[c]
    net = SgsNet(seed=5)
    with torch.no_grad():
        net.weights[-1].zero_()
        net.biases[-1].copy_(torch.tensor([0.3, -0.2, 0.1, 0.5, 0.4]))
    from lex_amd.setup_lex import build_case
    phys, _, _, _ = build_case(case)
    rng = np.random.default_rng(4)
    tgt = tuple(a + to_dev(rng.normal(0, 0.05, a.shape)) for a in phys)
    un = NetSolverUnroll(model, net, checkpoint_every=1)
    nsteps = 2
    st, ck = un.forward(phys, nsteps)
    L, lam = field_loss(st, tgt)
    net.zero_grad(set_to_none=True)
    un.backward(ck, nsteps, lam)
    gb = net.biases[-1].grad.detach().cpu().numpy()

    def loss_at(b):
        with torch.no_grad():
            net.biases[-1].copy_(torch.tensor(b, device="cuda"))
        s, _ = un.forward(phys, nsteps)
        return float(field_loss(s, tgt)[0])

    b0 = np.array([0.3, -0.2, 0.1, 0.5, 0.4])
    # the w and qv biases carry O(1) gradients; the u, v and theta ones are 1e-3 of those
    # (a uniform forcing on the periodic domain), below the FP32 rounding noise of the
    # GPU forward that the differences are taken from
    for c in (2, 4):
        e = np.zeros(5)
        e[c] = 0.5
        fd = (loss_at(b0 + e) - loss_at(b0 - e)) / 1.0
        assert _rel(gb[c], fd) < 3e-2, (c, gb[c], fd)
