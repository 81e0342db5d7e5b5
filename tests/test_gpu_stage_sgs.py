"""The step with the SGS tendencies recomputed at the start of every RK3 stage
(lex_rk3_step_stage_sgs, the learned SGS net run inside each stage), through the C ABI.

- a uniform wind damped by sgs_u = -k u (built on the device from the net's input
  features by lex_sgs_features / lex_sgs_apply) over a dry rest state advances by the
  WS-RK3 polynomial of -k dt, for every implementation of the device step; the frozen
  value 1 - k dt is rejected;
- a moist bubble on the coarse learned-SGS grid with a fixed pointwise "net"
  (tanh of an affine map of the 11 features) against the FP64 / FP32 oracle
  (ws_rk3_big_step_stage_sgs with sgs_features / sgs_apply), on the whole state and on
  the SGS contribution (the step with the net minus the step without SGS);
- an exception raised by the stage function aborts the step and reaches the caller;
- the coarse model's "net_stage" mode with the BF16 SgsNet runs and stays finite.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import from_dev, rel_l2

pytestmark = pytest.mark.gpu


def _opts(dt, n_sound, impl=0):
    from lex_amd import _lib
    return _lib.StepOpts(float(dt), int(n_sound), 0, 0.18, 1.0 / 3.0, 0, int(impl), 2)


def _cols(base):
    from lex_amd import _lib
    from lex_amd._core import column
    cols = [column(a) for a in (base[1], base[2], base[4], base[3], base[0])]
    return cols, _lib.BaseCols(*[c.data_ptr() for c in cols])


def _ws(g, dev):
    import torch
    from lex_amd import _lib
    return torch.empty(_lib.load().lex_rk3_workspace_bytes(g, 0), dtype=torch.uint8, device=dev)


class _PointwiseNet:
    """Stage function: features -> y = tanh(f W + c) (or a caller rule) -> lex_sgs_apply."""

    def __init__(self, g, base, shift, scale, out_scale, rule):
        import torch
        self.g = g
        self.th0c, self.qv0c = _cols(base)[0][1], _cols(base)[0][2]
        self.shift = (ctypes.c_float * 11)(*shift)
        self.scale = (ctypes.c_float * 11)(*scale)
        self.out_scale = (ctypes.c_float * 5)(*out_scale)
        self.feat = torch.empty((g.nx, g.ny, g.nz, 11), dtype=torch.float32, device=base[0].device)
        self.rule = rule
        self.stages = []

    def __call__(self, stage, sp, su, sv, sw, ss):
        from lex_amd import _lib
        self.stages.append(stage)
        g = self.g
        _lib.call("lex_sgs_features", g, _lib.ptr(self.th0c), _lib.ptr(self.qv0c), sp[1], sp[2], sp[3], sp[0], sp[5],
                  self.shift, self.scale, _lib.ptr(self.feat), _lib.stream())
        y = self.rule(self.feat).contiguous()
        _lib.call("lex_sgs_apply", g, _lib.ptr(y), self.out_scale, su, sv, sw, ss[0], ss[1], _lib.stream())


@pytest.mark.parametrize("impl", [0, 1, 2])
def test_linear_damping_on_the_device(impl):
    import torch
    from lex_amd.setup_lex import BubbleCase, build_case
    from lex_amd.sgs_train import stage_sgs_step
    from lex_amd._core import grid_from
    case = BubbleCase(16, 12, 12, 1000.0, 1000.0, 1000.0, amplitude=0.0)
    phys, base, grids, _ = build_case(case)
    g = grid_from(base[1], *grids[3:6])
    u0, k, dt = 1.7, 1e-2, 20.0
    st = tuple(t.clone() for t in phys)
    st[1].fill_(u0)

    def rule(feat):
        y = torch.zeros(feat.shape[:3] + (5,), dtype=torch.float32, device=feat.device)
        y[..., 0] = -k * feat[..., 0]
        return y
    net = _PointwiseNet(g, base, [0.0] * 11, [1.0] * 11, [1.0] * 5, rule)
    cols, bc = _cols(base)
    ws = _ws(g, st[0].device)
    stage_sgs_step(g, _opts(dt, 10, impl), bc, st, net, ws)
    torch.cuda.synchronize()
    assert net.stages == [0, 1, 2]
    z = -k * dt
    np.testing.assert_allclose(from_dev(st[1]), u0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=2e-6)
    assert np.abs(from_dev(st[1]) - u0 * (1 + z)).min() > 1e-2
    assert float(st[3].abs().max()) < 1e-6


def _coarse_case(nx=24, ny=24, nz=16):
    from lex_amd.setup_lex import build_case
    from lex_amd.sgs_train import member_case
    case = member_case(0, nx, ny, nz, 600.0, 300.0)
    return case, build_case(case)


def _oracle_run(init, base, og, oo, nsteps, dtype, net=None):
    """FP64 / FP32 oracle steps: with net = (W, c, shift, scale, out_scale) the pointwise
    net runs in every stage; without, the step has no SGS."""
    xi = tuple(a.astype(dtype) for a in init)
    bs = tuple(b.astype(dtype) for b in base) + (None,)
    if net is not None:
        W, c, shift, scale, osc = (np.asarray(a, dtype) for a in net)

        def fn(stage, x):
            f = (o.sgs_features(x, bs, og) - shift) * scale
            return o.sgs_apply(np.tanh(f @ W + c), osc)
    for _ in range(nsteps):
        if net is None:
            xi, _ = o.ws_rk3_big_step(xi, bs, og, oo)
        else:
            xi, _ = o.ws_rk3_big_step_stage_sgs(xi, bs, og, oo, fn)
    return xi


@pytest.mark.parametrize("impl", [0, 1])
def test_stage_net_bubble_vs_oracle(impl):
    import torch
    from lex_amd.sgs_train import stage_sgs_step
    from lex_amd._core import grid_from
    case, (phys, base, grids, _) = _coarse_case()
    g = grid_from(base[1], *grids[3:6])
    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    dt, n_sound, nsteps = 4.0, 10, 2
    init = tuple(from_dev(a) for a in phys)
    b = tuple(from_dev(a) for a in base[:5])
    f0 = o.sgs_features(init, tuple(b) + (None,), og).reshape(-1, 11)
    shift = np.zeros(11, np.float32)
    scale = np.float32(1.0) / np.maximum(f0.std(axis=0), 1e-6).astype(np.float32)
    rng = np.random.default_rng(7)
    W = (0.5 * rng.normal(size=(11, 5))).astype(np.float32)
    c = (0.1 * rng.normal(size=5)).astype(np.float32)
    out_scale = np.array([2e-3, 2e-3, 2e-3, 2e-3, 2e-7], np.float32)
    Wd, cd = torch.from_numpy(W).cuda(), torch.from_numpy(c).cuda()

    def rule(feat):
        acc = cd.expand(feat.shape[:3] + (5,)).clone()
        for ch in range(11):   # elementwise FP32 (no library GEMM)
            acc = acc + feat[..., ch:ch + 1] * Wd[ch]
        return torch.tanh(acc)
    net = _PointwiseNet(g, base, shift.tolist(), scale.tolist(), out_scale.tolist(), rule)
    cols, bc = _cols(base)
    ws = _ws(g, phys[0].device)
    st = tuple(t.clone() for t in phys)
    for _ in range(nsteps):
        stage_sgs_step(g, _opts(dt, n_sound, impl), bc, st, net, ws)
    st0 = tuple(t.clone() for t in phys)
    from lex_amd.one_sprint_functions import fused_step
    for _ in range(nsteps):
        fused_step(st0, base, g, (2, False, False, 0, False, False, 0), dt=dt, n_sound=n_sound, impl=impl)
    torch.cuda.synchronize()
    assert net.stages == [0, 1, 2] * nsteps
    got = tuple(from_dev(a) for a in st)
    none = tuple(from_dev(a) for a in st0)
    oo = o.Options(dt=dt, solver_opt=2, n_sound=n_sound, turb_opt=0)
    args = (W, c, shift, scale, out_scale)
    r64 = _oracle_run(init, b, og, oo, nsteps, np.float64, args)
    r32 = _oracle_run(init, b, og, oo, nsteps, np.float32, args)
    n64 = _oracle_run(init, b, og, oo, nsteps, np.float64)
    n32 = _oracle_run(init, b, og, oo, nsteps, np.float32)
    names = ["theta", "u", "v", "w", "pip", "qv"]
    for n, a, x64, x32, a0 in zip(names, got, r64, r32, init):
        da, db, dc = a - a0, x64 - a0, np.asarray(x32, np.float64) - a0
        e, e32 = rel_l2(da, db), rel_l2(dc, db)
        print(f"impl {impl} {n}: increment error {e:.2e} (FP32 oracle {e32:.2e})")
        assert e <= max(3 * e32, 1e-6), (n, e, e32)
    for q, n in enumerate(names):
        if n == "pip":
            continue
        cg = got[q] - none[q]
        co = r64[q] - n64[q]
        c32 = np.asarray(r32[q], np.float64) - n32[q]
        assert np.linalg.norm(co) > 0, n
        e, e32 = rel_l2(cg, co), rel_l2(c32, co)
        print(f"impl {impl} {n}: SGS contribution error {e:.2e} (FP32 oracle {e32:.2e})")
        assert e <= max(3 * e32, 1e-4), (n, e, e32)


def test_stage_function_error_reaches_the_caller():
    import torch
    from lex_amd.sgs_train import stage_sgs_step
    from lex_amd._core import grid_from
    case, (phys, base, grids, _) = _coarse_case(16, 16, 8)
    g = grid_from(base[1], *grids[3:6])
    cols, bc = _cols(base)
    ws = _ws(g, phys[0].device)
    net = _PointwiseNet(g, base, [0.0] * 11, [1.0] * 11, [1.0] * 5,
                        lambda f: torch.zeros(f.shape[:3] + (5,), dtype=torch.float32, device=f.device))

    def fn(stage, sp, su, sv, sw, ss):
        if stage == 1:
            raise ValueError("stage function failed on purpose")
        net(stage, sp, su, sv, sw, ss)
    st = tuple(t.clone() for t in phys)
    with pytest.raises(ValueError, match="on purpose"):
        stage_sgs_step(g, _opts(4.0, 10), bc, st, fn, ws)
    torch.cuda.synchronize()
    assert net.stages == [0]
    # the library is usable afterwards: a full step with the same function minus the failure
    stage_sgs_step(g, _opts(4.0, 10), bc, st, net, ws)
    torch.cuda.synchronize()
    assert net.stages == [0, 0, 1, 2]


def test_coarse_model_net_stage_mode():
    import time

    import torch
    from lex_amd.sgs_net import SgsNet
    from lex_amd.sgs_train import CoarseModel, member_case
    case = member_case(0, 96, 96, 48, 600.0, 300.0)
    model = CoarseModel(case, 4.0, 10)
    from lex_amd.setup_lex import build_case
    phys = build_case(case)[0]
    net = SgsNet(seed=2031, device="cuda")
    # an untrained net: features normalised to unit spread, outputs scaled to a small
    # tendency (1e-4 per second), as the fitted normalisation would
    f = model.features(phys).reshape(-1, 11)
    sd = f.std(dim=0).clamp_min(1e-6)
    for ch in range(11):
        model.shift[ch] = float(f[:, ch].mean())
        model.scale[ch] = float(1.0 / sd[ch])
    for ch in range(5):
        model.out_scale[ch] = 1e-4 if ch < 4 else 1e-8
    out = {}
    for mode in ("smag", "net", "net_stage"):
        st = tuple(a.clone() for a in phys)
        model.step(st, mode, net)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            model.step(st, mode, net)
        torch.cuda.synchronize()
        out[mode] = (time.perf_counter() - t0) / 5 * 1e3
        assert all(bool(torch.isfinite(a).all()) for a in st), mode
    print("96x96x48 coarse step ms: " + ", ".join(f"{k} {v:.3f}" for k, v in out.items()))
