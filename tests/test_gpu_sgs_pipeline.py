"""Learned-SGS data paths through the C ABI: coarsening against the oracle, the net's
input features against the oracle's deformation, the tendency staggering, the step with
caller-supplied SGS tendencies (bit-identical to the Smagorinsky step when fed the
Smagorinsky tendencies), and the small end-to-end train/score pipeline."""
import ctypes

import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import f32, f64, from_dev, random_state, to_dev

pytestmark = pytest.mark.gpu


def test_coarsen_matches_oracle():
    import torch
    from lex_amd import _lib
    nx, ny, nz, f = 24, 18, 12, (6, 3, 3)
    gr, phys, base = random_state(nx, ny, nz, seed=3)
    g = _lib.grid(nx, ny, nz, gr.dx, gr.dy, gr.dz)
    p32 = [f32(a) for a in phys]
    nxc, nyc, nzc = nx // 6, ny // 3, nz // 3
    cases = [(0, p32[0], (nxc, nyc, nzc), o.coarsen_block_mean(f64(p32[0][3:-3, 3:-3, 1:-1]), *f)),
             (1, p32[1], (nxc + 1, nyc, nzc), o.coarsen_face_mean(f64(p32[1][3:-3, 3:-3, 1:-1]), 0, *f)),
             (2, p32[2], (nxc, nyc + 1, nzc), o.coarsen_face_mean(f64(p32[2][3:-3, 3:-3, 1:-1]), 1, *f)),
             (3, p32[3], (nxc, nyc, nzc + 1), o.coarsen_face_mean(f64(p32[3][3:-3, 3:-3, 1:-1]), 2, *f))]
    for mode, a, shp, ref in cases:
        out = torch.empty(shp, dtype=torch.float32, device="cuda")
        a_d = to_dev(a)   # held: a pointer to a temporary would dangle
        _lib.call("lex_coarsen", g, mode, *f, _lib.ptr(a_d), _lib.ptr(out), _lib.stream())
        got = from_dev(out)
        np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-6 * float(np.abs(ref).max()))


def test_features_match_oracle_deformation():
    import torch
    from lex_amd import _lib
    nx, ny, nz = 16, 12, 12
    gr, phys, base = random_state(nx, ny, nz, seed=4)
    g = _lib.grid(nx, ny, nz, gr.dx, gr.dy, gr.dz)
    p = [f64(f32(a)) for a in phys]
    b = [f64(f32(a)) for a in base[:5]]
    s = o.compute_deformation(np.ones_like(b[1]), p[1], p[2], p[3], gr)
    th0c, qv0c = to_dev(np.ascontiguousarray(b[2][3, 3, :])), to_dev(np.ascontiguousarray(b[4][3, 3, :]))
    out = torch.empty((nx, ny, nz, 11), dtype=torch.float32, device="cuda")
    shift = (ctypes.c_float * 11)(*([0.0] * 11))
    scale = (ctypes.c_float * 11)(*([1.0] * 11))
    dv = [to_dev(a) for a in p]
    _lib.call("lex_sgs_features", g, _lib.ptr(th0c), _lib.ptr(qv0c), _lib.ptr(dv[1]), _lib.ptr(dv[2]), _lib.ptr(dv[3]),
              _lib.ptr(dv[0]), _lib.ptr(dv[5]), shift, scale, _lib.ptr(out), _lib.stream())
    got = from_dev(out)
    I = (slice(3, -3), slice(3, -3), slice(1, -1))
    ref = [0.5 * (p[1][3:-4, 3:-3, 1:-1] + p[1][4:-3, 3:-3, 1:-1]),
           0.5 * (p[2][3:-3, 3:-4, 1:-1] + p[2][3:-3, 4:-3, 1:-1]),
           0.5 * (p[3][3:-3, 3:-3, 1:-2] + p[3][3:-3, 3:-3, 2:-1]),
           p[0][I] - b[2][I], p[5][I] - b[4][I]] + [a[I] for a in s[:6]]
    for c, r in enumerate(ref):
        np.testing.assert_allclose(got[..., c], r, rtol=1e-5, atol=1e-6 * max(1.0, float(np.abs(r).max())),
                                   err_msg=f"channel {c}")


def test_sgs_apply_staggering():
    import torch
    from lex_amd import _lib
    nx, ny, nz = 12, 10, 8
    g = _lib.grid(nx, ny, nz, 100.0, 100.0, 100.0)
    x = np.arange(nx, dtype=np.float32)[:, None, None, None] * np.ones((1, ny, nz, 5), np.float32)
    y = to_dev(x + np.arange(5, dtype=np.float32))
    osc = (ctypes.c_float * 5)(1.0, 2.0, 3.0, 4.0, 5.0)
    su = torch.empty((nx + 1, ny, nz), device="cuda")
    sv = torch.empty((nx, ny + 1, nz), device="cuda")
    sw = torch.empty((nx, ny, nz + 1), device="cuda")
    st = torch.empty((nx, ny, nz), device="cuda")
    sq = torch.empty_like(st)
    _lib.call("lex_sgs_apply", g, _lib.ptr(y), osc, _lib.ptr(su), _lib.ptr(sv), _lib.ptr(sw), _lib.ptr(st), _lib.ptr(sq),
              _lib.stream())
    su, sv, sw, st = from_dev(su), from_dev(sv), from_dev(sw), from_dev(st)
    # u faces average cells i-1 and i (periodic): face i in 1..nx-1 -> i - 0.5; faces 0, nx -> (nx-1)/2
    assert np.allclose(su[1:nx, 0, 0], np.arange(1, nx) - 0.5)
    assert np.allclose(su[0], (nx - 1) / 2.0) and np.allclose(su[nx], (nx - 1) / 2.0)
    assert np.allclose(sv, 2.0 * (np.arange(nx)[:, None, None] + 1.0))
    assert np.all(sw[:, :, 0] == 0) and np.all(sw[:, :, -1] == 0)
    assert np.allclose(sw[:, :, 1:-1], 3.0 * (np.arange(nx)[:, None, None] + 2.0))
    assert np.allclose(st, 4.0 * (np.arange(nx)[:, None, None] + 3.0))


def test_ext_sgs_step_equals_smagorinsky_step():
    """Feeding lex_rk3_step_ext_sgs the Smagorinsky tendencies gives the turb-1 step bit for bit."""
    import torch
    from lex_amd import _lib, turbulence as tb
    from lex_amd.one_sprint_functions import fused_step
    from lex_amd._core import column, grid_from
    from lex_amd.setup_lex import build_case, config2
    case, opts = config2()
    case = case.scaled(24, 20, 16)
    phys, base, grids, _ = build_case(case)
    g = grid_from(base[1], *grids[3:6])
    a = tuple(t.clone() for t in phys)
    fused_step(a, base, g, (2, False, False, 0, False, False, 1), dt=2.0, n_sound=10)
    x3d, y3d, z3d, x4u, y4v, z4w = grids[:6]
    rho = base[1]
    s11, s22, s33, s12, s13, s23, df = tb.compute_deformation(rho, phys[1], phys[2], phys[3], x3d, y3d, z3d, x4u, y4v,
                                                              z4w)
    n2 = tb.compute_nm(phys[0], base[3], phys[5], z3d)
    km, kh = tb.compute_k(df, n2, x4u, y4v, z4w, cs=0.18, prandtl=1.0 / 3.0)
    sg = tb.compute_smag(rho, km, kh, s11, s22, s33, s12, s13, s23, phys[0], phys[5], x3d, y3d, z3d, x4u, y4v, z4w,
                         extra_scalars=tuple(phys[6:]))
    b = tuple(t.clone() for t in phys)
    cols = [column(x) for x in (base[1], base[2], base[4], base[3], base[0])]
    bc = _lib.BaseCols(*[c.data_ptr() for c in cols])
    o2 = _lib.StepOpts(2.0, 10, 2, 0.18, 1.0 / 3.0, len(b) - 6, 0)
    ws = torch.empty(_lib.load().lex_rk3_workspace_bytes(g, len(b) - 6), dtype=torch.uint8, device="cuda")
    ptrs = (ctypes.c_void_p * len(b))(*[_lib.ptr(x) for x in b])
    ss = (ctypes.c_void_p * len(sg[3:]))(*[_lib.ptr(x) for x in sg[3:]])
    _lib.call("lex_rk3_step_ext_sgs", g, o2, bc, ptrs, _lib.ptr(sg[0]), _lib.ptr(sg[1]), _lib.ptr(sg[2]), ss,
              _lib.ptr(ws), ws.numel(), _lib.stream())
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_small_pipeline_runs():
    """Fine runs -> coarsened pairs -> training -> coarse runs scored against the truth."""
    from lex_amd import sgs_train
    cfg = dict(sgs_train.SMALL)
    res = sgs_train.run(cfg, log=lambda s: print(s, flush=True))
    assert res["pairs"] == cfg["members"] * cfg["snaps"]
    assert all(np.isfinite(res["train_mse"]))
    for k in ("none", "smag", "net", "net_stage"):
        assert np.isfinite(res["theta_rmse_final"][k]), k
    assert res["chosen"] in ("net", "net_stage", "smagorinsky")
    print(res["theta_rmse_mean"], res["chosen"])
