"""The horizontal / vertical Smagorinsky closure (turb_opt 2: compute_k2, compute_smag2,
turbulence.py:114-295, 449-471) on the device vs the oracle: the coefficients and the
tendencies per operator (FP32 error within a multiple of the FP32 oracle's own), and
whole steps with turb_opt 2 (fused kernels, one kernel per routine, double-buffered
scalars) vs the FP64/FP32 oracle.  The device writes sgs_w = 0 at the surface, where
compute_smag2 keeps its stencil value; the step sets w to zero there after every
acoustic substep, so the state does not depend on it."""
import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import check_fp32, f32, f64, from_dev, random_state, rel_l2, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(16, 12, 12), (24, 20, 40)], ids=lambda s: "x".join(map(str, s)))
def test_k2_and_smag2_operators(shape):
    from lex_amd import turbulence as tb
    from lex_amd.setup_lex import make_grids
    nx, ny, nz = shape
    gr, phys, base = random_state(nx, ny, nz, seed=nx + 1, passive=1)
    p32 = tuple(f32(a) for a in phys)
    b32 = tuple(f32(a) for a in base[:5])
    p64, b64 = tuple(f64(a) for a in p32), tuple(f64(a) for a in b32)
    dv, bd = tuple(to_dev(a) for a in p32), tuple(to_dev(a) for a in b32)
    g = make_grids(gr.nx, gr.ny, gr.nz, gr.dx, gr.dy, gr.dz)
    d = tb.compute_deformation(bd[1], dv[1], dv[2], dv[3], *g[:6])
    n2 = tb.compute_nm(dv[0], None, dv[5], g[2])
    # the coefficients on shared FP32 inputs (the device's own s_ij and n2)
    ins = [f32(from_dev(a)) for a in d[:6] + (n2,)]
    kmh, kmv, khh, khv = tb.compute_k2(bd[1], *[to_dev(a) for a in ins], g[3], g[4], g[5], g[2])
    k64 = o.compute_k2(b64[1], *[f64(a) for a in ins], gr)
    k32 = o.compute_k2(b32[1], *ins, gr)
    check_fp32(from_dev(kmh), k64[0], k32[0], "kmh", factor=16.0)
    check_fp32(from_dev(kmv), k64[1], k32[1], "kmv", factor=16.0)
    assert khh.data_ptr() == kmh.data_ptr() and khv.data_ptr() == kmv.data_ptr()
    # the tendencies on shared inputs
    ks = [f32(from_dev(a)) for a in (kmh, kmv)]
    ss = [f32(from_dev(a)) for a in d[:6]]
    kd = [to_dev(a) for a in ks]
    sg = tb.compute_smag2(bd[1], kd[0], kd[1], kd[0], kd[1], *[to_dev(a) for a in ss], dv[0], dv[5], *g[:6],
                          extra_scalars=(dv[6],))
    s64 = o.compute_smag2(b64[1], f64(ks[0]), f64(ks[1]), f64(ks[0]), f64(ks[1]), *[f64(a) for a in ss],
                          (p64[0], p64[5], p64[6]), gr)
    s32 = o.compute_smag2(b32[1], ks[0], ks[1], ks[0], ks[1], *ss, (p32[0], p32[5], p32[6]), gr)
    for n, a, b, c in zip(["sgs_u", "sgs_v", "sgs_w", "sgs_theta", "sgs_qv", "sgs_passive"], sg, s64, s32):
        a = from_dev(a)
        if n == "sgs_w":
            assert not a[:, :, 0].any()
            a, b, c = a[:, :, 1:], b[:, :, 1:], c[:, :, 1:]
        check_fp32(a, b, c, n)


def _setup(nx, ny, nz):
    from lex_amd import namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    case = case.scaled(nx, ny, nz)
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], opts["n_sound"], opts["solver_opt"], 2
    mo = (opts["solver_opt"], False, False, 0, False, False, 2)
    og = o.Grid(nx, ny, nz, case.dx, case.dy, case.dz)
    oo = o.Options(dt=opts["dt"], solver_opt=opts["solver_opt"], n_sound=opts["n_sound"], turb_opt=2)
    return phys, base, grids, mo, og, oo


def _oracle_run(phys, base, og, oo, nsteps, dtype):
    xi = tuple(from_dev(a).astype(dtype) for a in phys)
    bs = tuple(from_dev(a).astype(dtype) for a in base[:5]) + (None,)
    for _ in range(nsteps):
        xi, _ = o.ws_rk3_big_step(xi, bs, og, oo)
    return xi


@pytest.mark.parametrize("impl,nsteps", [(0, 1), (0, 3), (1, 2), (2, 1)])
def test_turb_opt_2_step_vs_oracle(impl, nsteps):
    from lex_amd import one_sprint_functions as sp
    from lex_amd._core import grid_from
    phys, base, grids, mo, og, oo = _setup(24, 20, 24)
    init = tuple(from_dev(a) for a in phys)
    st = tuple(a.clone() for a in phys)
    g = grid_from(base[1], *grids[3:6])
    for _ in range(nsteps):
        sp.fused_step(st, base, g, mo, impl=impl)
    got = tuple(from_dev(a) for a in st)
    r64 = _oracle_run(phys, base, og, oo, nsteps, np.float64)
    r32 = _oracle_run(phys, base, og, oo, nsteps, np.float32)
    for n, a, b, c, a0 in zip(["theta", "u", "v", "w", "pip", "qv"], got, r64, r32, init):
        da, db, dc = a - a0, b - a0, c - a0
        e_g, e_32 = rel_l2(da, db), rel_l2(dc, db)
        print(f"{n}: GPU {e_g:.2e} FP32 oracle {e_32:.2e}")
        assert e_g <= max(10.0 * e_32, 2e-6), n
    # the closure changes the result against turb_opt 1
    oo.turb_opt = 1
    iso = _oracle_run(phys, base, og, oo, nsteps, np.float64)
    assert rel_l2(r64[0] - init[0], iso[0] - init[0]) > 1e-4
