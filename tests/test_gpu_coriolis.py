"""Coriolis options on the GPU, through the C ABI (pressure_gradient_coriolis.py:80-113,
one_step_integration.py:83-89, :126-133).

- lex_coriolis / lex_coriolis_lspgrad against the FP64 oracle;
- lex_momentum_rhs against the same FP32 arithmetic in NumPy (bit-exact);
- the device step (fused kernels, per-routine kernels, double-buffered scalars) on a
  uniform wind over a dry rest state: the WS-RK3 polynomial of the f-plane rotation;
- bubble steps with cor_opt 1 and 2 (device step and the staged pseudo-incompressible
  path) against the FP64 / FP32 oracle, on the whole state and on the Coriolis
  contribution itself (the step with Coriolis minus the same step without).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import f32, f64, from_dev, random_state, rel_l2, to_dev

pytestmark = pytest.mark.gpu

NAMES = ["theta", "u", "v", "w", "pip", "qv", "qc", "qr"]


def test_coriolis_kernels_vs_oracle():
    from lex_amd import pressure_gradient_coriolis as pgc
    gr, phys, base = random_state(17, 11, 9, seed=21, vel=4.0)
    u, v = f32(phys[1]), f32(phys[2])
    for fn_g, fn_o in ((pgc.calculate_coriolis_force, o.calculate_coriolis_force),
                       (pgc.calculate_coriolis_force_with_lspgrad, o.calculate_coriolis_force_with_lspgrad)):
        fu, fv = fn_g(to_dev(u), to_dev(v))
        ru, rv = fn_o(f64(u), f64(v))
        ru32, rv32 = fn_o(u, v)
        for name, a, b, c in (("fu", fu, ru, ru32), ("fv", fv, rv, rv32)):
            e = rel_l2(from_dev(a), b)
            assert e <= max(4 * rel_l2(c, b), 1e-7), (fn_o.__name__, name, e)
    # geostrophic balance on the device: the (1, 0) m/s wind feels no force
    fu, fv = pgc.calculate_coriolis_force_with_lspgrad(to_dev(np.ones_like(u)), to_dev(np.zeros_like(v)))
    assert float(fu.abs().max()) == 0.0 and float(fv.abs().max()) == 0.0


def test_momentum_rhs_is_the_reference_sum():
    import torch
    from lex_amd import _lib
    from lex_amd.one_step_integration import momentum_rhs
    nx, ny, nz = 13, 10, 9
    g = _lib.grid(nx, ny, nz, 100.0, 100.0, 100.0)
    rng = np.random.default_rng(5)
    shp_u, shp_v, shp_w = (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)
    a = [f32(rng.normal(size=s)) for s in (shp_u, shp_v, shp_w)]
    fu, fv = f32(rng.normal(size=shp_v)), f32(rng.normal(size=shp_u))
    b8w = f32(rng.normal(size=shp_w))
    sgs = [f32(rng.normal(size=s)) for s in (shp_u, shp_v, shp_w)]
    got = momentum_rhs(g, *[to_dev(x) for x in a], to_dev(fu), to_dev(fv), to_dev(b8w),
                       tuple(to_dev(x) for x in sgs))
    ref = ((a[0] + fv) + sgs[0], (a[1] - fu) + sgs[1], (a[2] + b8w) + sgs[2])
    for x, y in zip(got, ref):
        assert np.array_equal(x.cpu().numpy(), y)
    got = momentum_rhs(g, *[to_dev(x) for x in a], None, None, to_dev(b8w), None)
    for x, y in zip(got, (a[0], a[1], a[2] + b8w)):
        assert np.array_equal(x.cpu().numpy(), y)
    torch.cuda.synchronize()


def _rest_case(nx=16, ny=12, nz=12, dx=1000.0):
    from lex_amd.setup_lex import BubbleCase, build_case
    case = BubbleCase(nx, ny, nz, dx, dx, dx, amplitude=0.0)
    return case, build_case(case)


@pytest.mark.parametrize("impl", [0, 1, 2])
@pytest.mark.parametrize("cor_opt,ug", [(1, 0.0), (2, 1.0)])
def test_inertial_oscillation_on_the_device_step(impl, cor_opt, ug):
    import torch
    from lex_amd import namelist_n_constants as nl, one_sprint_functions as sp
    from lex_amd._core import grid_from
    case, (phys, base, grids, _) = _rest_case()
    # f dt = 1e-2: the second-order term of the WS-RK3 polynomial ((f dt)^2 / 2 u ~ 8.5e-5)
    # stands far above the FP32 rounding of accumulating 1 + 10 + 20 substep increments
    # into u ~ 1.7 (ulp 1.2e-7).  The state stays exactly uniform, so no acoustic mode is
    # excited whatever the acoustic Courant number.
    u0, v0, dt = 1.7, -0.6, 100.0
    st = list(t.clone() for t in phys)
    st[1].fill_(u0)
    st[2].fill_(v0)
    st = tuple(st)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = dt, 20, 2, 1
    g = grid_from(base[1], *grids[3:6])
    sp.fused_step(st, base, g, (2, False, False, cor_opt, False, False, 1), impl=impl)
    torch.cuda.synchronize()
    A = o.fCor * dt * np.array([[0.0, 1.0], [-1.0, 0.0]])
    y = (np.eye(2) + A + A @ A / 2 + A @ A @ A / 6) @ np.array([u0 - ug, v0])
    du, dv = from_dev(st[1]) - u0, from_dev(st[2]) - v0
    np.testing.assert_allclose(du, (y[0] + ug) - u0, rtol=1e-4, atol=3e-6)
    np.testing.assert_allclose(dv, y[1] - v0, rtol=1e-4, atol=3e-6)
    # the second-order term is resolved: a first-order (forward Euler) result is rejected
    y1 = (np.eye(2) + A) @ np.array([u0 - ug, v0])
    assert np.abs(du - ((y1[0] + ug) - u0)).max() > 3e-5
    assert float(st[3].abs().max()) < 1e-6


def _bubble(cor_opt, shape=(24, 20, 24)):
    from lex_amd import namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    case = case.scaled(*shape)
    phys, base, grids, _ = build_case(case)
    # a 5 m/s mean wind so the f-plane force is visible over a few steps
    phys[1].add_(5.0)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    og = o.Grid(*shape, case.dx, case.dy, case.dz)
    oo = o.Options(dt=opts["dt"], solver_opt=2, n_sound=7, turb_opt=1, cor_opt=cor_opt)
    return phys, base, grids, og, oo


def _oracle(init, base, og, oo, nsteps, dtype):
    xi = tuple(a.astype(dtype) for a in init)
    bs = tuple(b.astype(dtype) for b in base) + (None,)
    for _ in range(nsteps):
        xi, _ = o.ws_rk3_big_step(xi, bs, og, oo)
    return xi


@pytest.mark.parametrize("impl", [0, 1])
@pytest.mark.parametrize("cor_opt", [1, 2])
def test_bubble_steps_with_coriolis_vs_oracle(cor_opt, impl):
    import dataclasses
    from lex_amd import one_sprint_functions as sp
    from lex_amd._core import grid_from
    nsteps = 3
    phys, base, grids, og, oo = _bubble(cor_opt)
    g = grid_from(base[1], *grids[3:6])
    init = tuple(from_dev(a) for a in phys)
    b = tuple(from_dev(a) for a in base[:5])
    runs = {}
    for c in (0, cor_opt):
        st = tuple(t.clone() for t in phys)
        for _ in range(nsteps):
            sp.fused_step(st, base, g, (2, False, False, c, False, False, 1), impl=impl)
        runs[c] = tuple(from_dev(a) for a in st)
    r64 = {c: _oracle(init, b, og, dataclasses.replace(oo, cor_opt=c), nsteps, np.float64) for c in (0, cor_opt)}
    r32 = {c: _oracle(init, b, og, dataclasses.replace(oo, cor_opt=c), nsteps, np.float32) for c in (0, cor_opt)}
    got = runs[cor_opt]
    for n, a, x64, x32, a0 in zip(NAMES, got, r64[cor_opt], r32[cor_opt], init):
        da, db, dc = a - a0, x64 - a0, np.asarray(x32, np.float64) - a0
        if np.linalg.norm(db) == 0:
            continue
        assert rel_l2(da, db) <= max(3 * rel_l2(dc, db), 1e-6), (n, rel_l2(da, db), rel_l2(dc, db))
    # the Coriolis contribution (with minus without) on u and v
    for q in (1, 2):
        cg = runs[cor_opt][q] - runs[0][q]
        co = r64[cor_opt][q] - r64[0][q]
        c32 = np.asarray(r32[cor_opt][q], np.float64) - r32[0][q]
        assert np.linalg.norm(co) > 0
        e, e32 = rel_l2(cg, co), rel_l2(c32, co)
        print(f"cor_opt {cor_opt} impl {impl} field {q}: contribution error {e:.2e} (FP32 oracle {e32:.2e})")
        assert e <= max(3 * e32, 1e-4), (q, e, e32)


@pytest.mark.parametrize("cor_opt", [1, 2])
def test_pseudo_incompressible_step_with_coriolis_vs_oracle(cor_opt):
    """The staged solver_opt 1 path: Coriolis in the pressure RHS and the momentum update."""
    from lex_amd import namelist_n_constants as nl, one_sprint_functions as sp
    from lex_amd.setup_lex import build_case, config0
    case, opts = config0()
    case = case.scaled(20, 16, 12)
    phys, base, grids, _ = build_case(case)
    phys[1].add_(5.0)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], opts["n_sound"], 1, 1
    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    init = tuple(from_dev(a) for a in phys)
    b = tuple(from_dev(a) for a in base[:5])
    got = {}
    for c in (0, cor_opt):
        out, _ = sp.ws_rk3_big_step(phys, base, grids, (1, False, False, c, False, False, 1))
        got[c] = tuple(from_dev(a) for a in out)
    mk = lambda c: o.Options(dt=opts["dt"], solver_opt=1, n_sound=opts["n_sound"], turb_opt=1, cor_opt=c)  # noqa: E731
    ref = {c: _oracle(init, b, og, mk(c), 1, np.float64) for c in (0, cor_opt)}
    r32 = {c: _oracle(init, b, og, mk(c), 1, np.float32) for c in (0, cor_opt)}
    for q in (1, 2):   # the Coriolis contribution to u and v after one step
        cg = got[cor_opt][q] - got[0][q]
        co = ref[cor_opt][q] - ref[0][q]
        c32 = np.asarray(r32[cor_opt][q], np.float64) - r32[0][q]
        assert np.linalg.norm(co) > 0
        e, e32 = rel_l2(cg, co), rel_l2(c32, co)
        print(f"solver_opt 1 cor_opt {cor_opt} field {q}: contribution error {e:.2e} (FP32 oracle {e32:.2e})")
        assert e <= max(3 * e32, 1e-3), (q, e, e32)
