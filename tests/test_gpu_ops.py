"""Per-operator parity: HIP kernels (through the C ABI via the mirror modules) vs the oracle.

Inputs are rounded to float32 once; the oracle runs on them in float64 (reference
values) and in float32 (the reference's own precision).  A kernel passes when its error
against float64 is within a small multiple of the float32 oracle's error, so the
tolerance is the intrinsic FP32 drift of the reference algorithm (stated per test).
Halo and index maps are checked bit-exactly.
"""
import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import check_fp32, f32, f64, from_dev, random_state, to_dev, rel_l2

pytestmark = pytest.mark.gpu

SIZES = [(16, 12, 12), (24, 20, 16)]


@pytest.fixture(scope="module", params=SIZES, ids=lambda s: "x".join(map(str, s)))
def case(request):
    import torch  # noqa: F401
    nx, ny, nz = request.param
    gr, phys, base = random_state(nx, ny, nz, seed=nx, passive=1)
    p32 = tuple(f32(a) for a in phys)
    b32 = tuple(f32(a) for a in base[:5])
    grids = make_grids(gr)
    return dict(gr=gr, p32=p32, b32=b32, p64=tuple(f64(a) for a in p32), b64=tuple(f64(a) for a in b32),
                dev=tuple(to_dev(a) for a in p32), bdev=tuple(to_dev(a) for a in b32), grids=grids)


def make_grids(gr):
    from lex_amd.setup_lex import make_grids as mg
    return mg(gr.nx, gr.ny, gr.nz, gr.dx, gr.dy, gr.dz)


def test_padding_bit_exact(case):
    from lex_amd import one_step_integration as one, pressure_equations as pe
    gr = case["gr"]
    rng = np.random.default_rng(3)
    for fn_o, fn_g, shape in [(o.padding3_array, one.padding3_array, (gr.nx, gr.ny, gr.nz)),
                              (o.padding3u_array, one.padding3u_array, (gr.nx + 1, gr.ny, gr.nz)),
                              (o.padding3v_array, one.padding3v_array, (gr.nx, gr.ny + 1, gr.nz)),
                              (o.padding3_array, one.padding3_array, (gr.nx, gr.ny, gr.nz + 1)),
                              (o.padding_array, pe.padding_array, (gr.nx, gr.ny, gr.nz))]:
        a = f32(rng.normal(size=shape))
        assert np.array_equal(from_dev(fn_g(to_dev(a))), f64(fn_o(a)))


def test_set_w_bc_bit_exact(case):
    from lex_amd.boundary_conditions import set_w_bc
    w = case["p32"][3] + 1.0
    assert np.array_equal(from_dev(set_w_bc(to_dev(w))), f64(o.set_w_bc(w)))


def test_divergences(case):
    from lex_amd import advection as adv
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    got = from_dev(adv.get_divergence(bd[1], dv[1], dv[2], dv[3], g[3], g[4], g[5]))
    check_fp32(got, o.get_divergence(b64[1], p64[1], p64[2], p64[3], gr),
               o.get_divergence(b32[1], p32[1], p32[2], p32[3], gr), "get_divergence")
    got = from_dev(adv.get_divergence2(dv[1], dv[2], dv[3], g[3], g[4], g[5]))
    check_fp32(got, o.get_divergence2(p64[1], p64[2], p64[3], gr), o.get_divergence2(p32[1], p32[2], p32[3], gr),
               "get_divergence2")


@pytest.mark.parametrize("field", ["theta", "qv", "passive"])
def test_advection_scalar(case, field):
    from lex_amd import advection as adv
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    idx, weps = {"theta": (0, o.WEPS_THETA), "qv": (5, o.WEPS_QV), "passive": (6, o.WEPS_QV)}[field]
    sfc = np.linspace(-0.1, 0.1, gr.nx * gr.ny).reshape(gr.nx, gr.ny).astype(np.float32)
    fd64 = o.get_divergence(b64[1], p64[1], p64[2], p64[3], gr)
    fd32 = o.get_divergence(b32[1], p32[1], p32[2], p32[3], gr)
    ref64 = o.advection_scalar(b64[1], p64[idx], p64[1], p64[2], p64[3], weps, fd64, f64(sfc), gr)
    ref32 = o.advection_scalar(b32[1], p32[idx], p32[1], p32[2], p32[3], weps, fd32, sfc, gr)
    fdg = adv.get_divergence(bd[1], dv[1], dv[2], dv[3], g[3], g[4], g[5])
    got = from_dev(adv.advection_scalar(bd[1], dv[idx], dv[1], dv[2], dv[3], weps, fdg, to_dev(sfc), g[3], g[4], g[5]))
    check_fp32(got, ref64, ref32, f"advection_scalar[{field}]")


@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_advection_momentum(case, comp):
    from lex_amd import advection as adv
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    fd64 = o.get_divergence(b64[1], p64[1], p64[2], p64[3], gr)
    fd32 = o.get_divergence(b32[1], p32[1], p32[2], p32[3], gr)
    fdg = adv.get_divergence(bd[1], dv[1], dv[2], dv[3], g[3], g[4], g[5])
    if comp == "u":
        s = np.cos(np.arange(gr.nx * gr.ny)).reshape(gr.nx, gr.ny).astype(np.float32) * 0.05
        ref64 = o.advection_u(b64[1], p64[1], p64[2], p64[3], 1e-16, fd64, f64(s), gr)
        ref32 = o.advection_u(b32[1], p32[1], p32[2], p32[3], 1e-16, fd32, s, gr)
        got = adv.advection_u(bd[1], dv[1], dv[2], dv[3], 1e-16, fdg, to_dev(s), g[0], g[4], g[5])
    elif comp == "v":
        s = np.sin(np.arange(gr.nx * gr.ny)).reshape(gr.nx, gr.ny).astype(np.float32) * 0.05
        ref64 = o.advection_v(b64[1], p64[1], p64[2], p64[3], 1e-16, fd64, f64(s), gr)
        ref32 = o.advection_v(b32[1], p32[1], p32[2], p32[3], 1e-16, fd32, s, gr)
        got = adv.advection_v(bd[1], dv[1], dv[2], dv[3], 1e-16, fdg, to_dev(s), g[3], g[1], g[5])
    else:
        ref64 = o.advection_w(b64[1], p64[1], p64[2], p64[3], 1e-16, fd64, gr)
        ref32 = o.advection_w(b32[1], p32[1], p32[2], p32[3], 1e-16, fd32, gr)
        got = adv.advection_w(bd[1], dv[1], dv[2], dv[3], 1e-16, fdg, g[3], g[4], g[2])
    check_fp32(from_dev(got), ref64, ref32, f"advection_{comp}")


def test_advection_pi(case):
    from lex_amd import advection as adv
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    ref64 = o.advection_pi(b64[1], b64[3], p64[4], p64[1], p64[2], p64[3], gr)
    ref32 = o.advection_pi(b32[1], b32[3], p32[4], p32[1], p32[2], p32[3], gr)
    got = adv.advection_pi(bd[1], bd[3], dv[4], dv[1], dv[2], dv[3], g[0], g[1], g[2])
    check_fp32(from_dev(got), ref64, ref32, "advection_pi")


def test_buoyancy_rtt(case):
    from lex_amd import pressure_equations as pe
    gr = case["gr"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    b, b8w = pe.calculate_buoyancy(bd[2], dv[0], bd[4], dv[5])
    r64 = o.calculate_buoyancy(b64[2], p64[0], b64[4], p64[5])
    r32 = o.calculate_buoyancy(b32[2], p32[0], b32[4], p32[5])
    check_fp32(from_dev(b), r64[0], r32[0], "buoyancy")
    check_fp32(from_dev(b8w), r64[1], r32[1], "buoyancy@w")
    rtt = pe.calculate_rtt(bd[0], dv[0], dv[5], b)
    check_fp32(from_dev(rtt), o.calculate_rtt(b64[0], p64[0], p64[5], r64[0]),
               o.calculate_rtt(b32[0], p32[0], p32[5], r32[0]), "calculate_rtt")


def test_pgf_coriolis(case):
    from lex_amd import pressure_gradient_coriolis as pgc
    gr, g = case["gr"], case["grids"]
    p32, p64, dv = case["p32"], case["p64"], case["dev"]
    tr32 = o.theta_rho_of(p32[0], p32[5])
    tr64 = o.theta_rho_of(p64[0], p64[5])
    got = pgc.pressure_gradient_force(None, dv[4], to_dev(tr32), g[0], g[1], g[2])
    r64 = o.pressure_gradient_force(p64[4], f64(tr32), gr)
    r32 = o.pressure_gradient_force(p32[4], tr32, gr)
    for name, a, b, c in zip("uvw", got, r64, r32):
        check_fp32(from_dev(a), b, c, f"pgf_{name}")
    del tr64
    fu, fv = pgc.calculate_coriolis_force(dv[1], dv[2])
    r64 = o.calculate_coriolis_force(p64[1], p64[2])
    r32 = o.calculate_coriolis_force(p32[1], p32[2])
    check_fp32(from_dev(fu), r64[0], r32[0], "coriolis fu")
    check_fp32(from_dev(fv), r64[1], r32[1], "coriolis fv")


def test_pressure_rhs_matvec(case):
    from lex_amd import pressure_equations as pe
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    rng = np.random.default_rng(7)
    au = f32(rng.normal(size=(gr.nx + 1, gr.ny, gr.nz)))
    av = f32(rng.normal(size=(gr.nx, gr.ny + 1, gr.nz)))
    aw = f32(rng.normal(size=(gr.nx, gr.ny, gr.nz + 1)))
    fu = f32(rng.normal(size=(gr.nx, gr.ny + 1, gr.nz)))
    fv = f32(rng.normal(size=(gr.nx + 1, gr.ny, gr.nz)))
    b = f32(o.calculate_buoyancy(b32[2], p32[0], b32[4], p32[5])[0])
    rhs = pe.rhs_of_pressure_equation(bd[0], bd[3], None, to_dev(au), to_dev(av), to_dev(aw), to_dev(fu), to_dev(fv),
                                      to_dev(b), g[0], g[3], g[1], g[4], g[5])
    r64 = o.rhs_of_pressure_equation(b64[0], f64(au), f64(av), f64(aw), f64(fu), f64(fv), f64(b), gr)
    r32 = o.rhs_of_pressure_equation(b32[0], au, av, aw, fu, fv, b, gr)
    check_fp32(from_dev(rhs[0]), r64, r32, "rhs_of_pressure_equation")
    rtt32 = f32(o.calculate_rtt(b32[0], p32[0], p32[5], b))
    x = f32(rng.normal(size=(gr.nx, gr.ny, gr.nz)))
    got = pe.laplace_of_pressure(g[0], g[3], g[1], g[4], g[2], g[5], to_dev(rtt32), to_dev(x))
    check_fp32(from_dev(got), o.laplace_of_pressure(f64(rtt32), f64(x), gr), o.laplace_of_pressure(rtt32, x, gr),
               "laplace_of_pressure")


def test_smagorinsky(case):
    from lex_amd import turbulence as tb
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    d = tb.compute_deformation(bd[1], dv[1], dv[2], dv[3], *g[:6])
    r64 = o.compute_deformation(b64[1], p64[1], p64[2], p64[3], gr)
    r32 = o.compute_deformation(b32[1], p32[1], p32[2], p32[3], gr)
    for n, a, b, c in zip(["s11", "s22", "s33", "s12", "s13", "s23", "deform"], d, r64, r32):
        check_fp32(from_dev(a), b, c, n)
    n2 = tb.compute_nm(dv[0], None, dv[5], g[2])
    n64 = o.compute_nm(p64[0], p64[5], gr)
    n32 = o.compute_nm(p32[0], p32[5], gr)
    check_fp32(from_dev(n2), n64, n32, "compute_nm", factor=16.0)
    km, kh = tb.compute_k(d[6], n2, g[3], g[4], g[5])
    k64 = o.compute_k(r64[6], n64, gr)
    k32 = o.compute_k(r32[6], n32, gr)
    check_fp32(from_dev(km), k64[0], k32[0], "km", factor=16.0)
    # compute_smag on shared inputs (the GPU's own km/kh/s_ij, rounded) isolates the divergence
    ins = [f32(from_dev(a)) for a in (km, kh) + d[:6]]
    sg = tb.compute_smag(bd[1], *[to_dev(a) for a in ins], dv[0], dv[5], *g[:6], extra_scalars=(dv[6],))
    s64 = o.compute_smag(b64[1], *[f64(a) for a in ins], (p64[0], p64[5], p64[6]), gr)
    s32 = o.compute_smag(b32[1], *ins, (p32[0], p32[5], p32[6]), gr)
    for n, a, b, c in zip(["sgs_u", "sgs_v", "sgs_w", "sgs_theta", "sgs_qv", "sgs_passive"], sg, s64, s32):
        check_fp32(from_dev(a), b, c, n)


def test_acoustic_substeps(case):
    from lex_amd import one_step_integration as one
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    rng = np.random.default_rng(11)
    ru = f32(rng.normal(0, 1e-2, (gr.nx + 1, gr.ny, gr.nz)))
    rv = f32(rng.normal(0, 1e-2, (gr.nx, gr.ny + 1, gr.nz)))
    rw = f32(rng.normal(0, 1e-2, (gr.nx, gr.ny, gr.nz + 1)))
    tr = f32(o.theta_rho_of(p32[0], p32[5]))
    pstar = f32(p32[4] * 0.5)
    n, dts = 3, 0.05
    got = one.sound_rk_step1(dv[1], dv[2], dv[3], bd[3], dv[4], to_dev(pstar), to_dev(ru), to_dev(rv), to_dev(rw),
                             0.0, to_dev(tr), bd[1], n, dts, *g[:8])
    r64 = o.sound_rk_step(p64[1], p64[2], p64[3], b64[3], p64[4], f64(pstar), f64(ru), f64(rv), f64(rw), f64(tr),
                          b64[1], n, dts, gr)
    r32 = o.sound_rk_step(p32[1], p32[2], p32[3], b32[3], p32[4], pstar, ru, rv, rw, tr, b32[1], n, dts, gr)
    base = (p64[1], p64[2], p64[3], p64[4])
    for name, a, b, c, a0 in zip(["u", "v", "w", "pip"], got, r64, r32, base):
        # compare increments (the state itself is dominated by its initial value)
        check_fp32(from_dev(a) - a0, b - a0, c - a0, f"acoustic {name}")


def test_implicit_acoustic_substeps(case):
    """solver_opt 3: vertically implicit substeps (PCR tridiagonal solve per column on the
    GPU, Thomas in the oracle) against the FP64/FP32 oracle."""
    from lex_amd import one_step_integration as one
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    rng = np.random.default_rng(12)
    ru = f32(rng.normal(0, 1e-2, (gr.nx + 1, gr.ny, gr.nz)))
    rv = f32(rng.normal(0, 1e-2, (gr.nx, gr.ny + 1, gr.nz)))
    rw = f32(rng.normal(0, 1e-2, (gr.nx, gr.ny, gr.nz + 1)))
    tr = f32(o.theta_rho_of(p32[0], p32[5]))
    pstar = f32(p32[4] * 0.5)
    n, dts = 3, 0.2
    got = one.sound_implicit_rk_step1(dv[1], dv[2], dv[3], bd[3], dv[4], to_dev(pstar), to_dev(ru), to_dev(rv),
                                      to_dev(rw), 0.0, to_dev(tr), bd[1], n, dts, *g[:8])
    r64 = o.sound_implicit_rk_step(p64[1], p64[2], p64[3], b64[3], p64[4], f64(pstar), f64(ru), f64(rv), f64(rw),
                                   f64(tr), b64[1], n, dts, gr)
    r32 = o.sound_implicit_rk_step(p32[1], p32[2], p32[3], b32[3], p32[4], pstar, ru, rv, rw, tr, b32[1], n, dts, gr)
    base = (p64[1], p64[2], p64[3], p64[4])
    for name, a, b, c, a0 in zip(["u", "v", "w", "pip"], got, r64, r32, base):
        check_fp32(from_dev(a) - a0, b - a0, c - a0, f"implicit acoustic {name}")


def test_scalar_and_momentum_updates(case):
    from lex_amd import one_step_integration as one, advection as adv
    gr, g = case["gr"], case["grids"]
    p32, p64, b32, b64, dv, bd = case["p32"], case["p64"], case["b32"], case["b64"], case["dev"], case["bdev"]
    dt = 0.3
    fdg = adv.get_divergence(bd[1], dv[1], dv[2], dv[3], g[3], g[4], g[5])
    th0 = f32(p32[0] + 0.1)
    nxt, tend = one.update_theta_euler(bd[1], to_dev(th0), dv[0], dv[1], dv[2], dv[3], fdg, None, None,
                                       g[3], g[4], g[5], dt)
    z2 = np.zeros((gr.nx, gr.ny))
    fd64 = o.get_divergence(b64[1], p64[1], p64[2], p64[3], gr)
    fd32 = o.get_divergence(b32[1], p32[1], p32[2], p32[3], gr)
    r64 = o.update_scalar_euler(b64[1], f64(th0), p64[0], p64[1], p64[2], p64[3], fd64, z2, 0.0, o.WEPS_THETA, dt, gr)
    r32 = o.update_scalar_euler(b32[1], th0, p32[0], p32[1], p32[2], p32[3], fd32, z2.astype(np.float32), 0.0,
                                o.WEPS_THETA, dt, gr)
    check_fp32(from_dev(nxt) - f64(th0), r64[0] - f64(th0), r32[0] - th0, "update_theta_euler")
    check_fp32(from_dev(tend), r64[1], r32[1], "theta tendency")
    # momentum (solver_opt 1 path)
    rng = np.random.default_rng(5)
    au = f32(rng.normal(0, 1e-2, (gr.nx + 1, gr.ny, gr.nz)))
    av = f32(rng.normal(0, 1e-2, (gr.nx, gr.ny + 1, gr.nz)))
    aw = f32(rng.normal(0, 1e-2, (gr.nx, gr.ny, gr.nz + 1)))
    b8w = f32(rng.normal(0, 1e-2, (gr.nx, gr.ny, gr.nz + 1)))
    tr = f32(o.theta_rho_of(p32[0], p32[5]))
    fu = f32(np.zeros((gr.nx, gr.ny + 1, gr.nz)))
    fv = f32(np.zeros((gr.nx + 1, gr.ny, gr.nz)))
    got = one.update_momentum_eqn_euler(dv[1], dv[2], dv[3], None, dv[4], to_dev(tr), to_dev(au), to_dev(av),
                                        to_dev(aw), None, None, to_dev(b8w), g[0], g[1], g[2], dt)
    r64 = o.update_momentum_eqn_euler(p64[1], p64[2], p64[3], p64[4], f64(tr), f64(au), f64(av), f64(aw), f64(fu),
                                      f64(fv), f64(b8w), dt, gr)
    r32 = o.update_momentum_eqn_euler(p32[1], p32[2], p32[3], p32[4], tr, au, av, aw, fu, fv, b8w, dt, gr)
    for name, a, b, c, a0 in zip(["u", "v", "w"], got[:3], r64[:3], r32[:3], (p64[1], p64[2], p64[3])):
        check_fp32(from_dev(a) - a0, b - a0, c - a0, f"momentum {name}")


def test_bicgstab_matches_oracle(case):
    """Converged/stagnated BiCGSTAB iterate vs the oracle's (JAX semantics), mean removed."""
    from lex_amd import one_step_integration as one, pressure_equations as pe
    gr, g = case["gr"], case["grids"]
    p32, b32, dv, bd = case["p32"], case["b32"], case["dev"], case["bdev"]
    rng = np.random.default_rng(2)
    au = f32(rng.normal(0, 1e-3, (gr.nx + 1, gr.ny, gr.nz)))
    av = f32(rng.normal(0, 1e-3, (gr.nx, gr.ny + 1, gr.nz)))
    aw = f32(rng.normal(0, 1e-3, (gr.nx, gr.ny, gr.nz + 1)))
    aw[:, :, 0] = 0.0
    aw[:, :, -1] = 0.0
    b = f32(o.calculate_buoyancy(b32[2], p32[0], b32[4], p32[5])[0])
    rtt = f32(o.calculate_rtt(b32[0], p32[0], p32[5], b))
    z_u = np.zeros((gr.nx, gr.ny + 1, gr.nz), np.float32)
    z_v = np.zeros((gr.nx + 1, gr.ny, gr.nz), np.float32)
    x_g, k_g = one.solve_pres_eqn(dv[4], bd[0], bd[3], to_dev(rtt), to_dev(au), to_dev(av), to_dev(aw),
                                  to_dev(z_u), to_dev(z_v), to_dev(b), g[0], g[3], g[1], g[4], g[2], g[5])
    x_o, k_o = o.solve_pres_eqn(f64(p32[4]), f64(b32[0]), f64(rtt), f64(au), f64(av), f64(aw), f64(z_u), f64(z_v),
                                f64(b), gr, o.Options())
    xg = from_dev(x_g)
    xg -= xg.mean()
    xo = x_o - x_o.mean()
    # the same algorithm in different precision: residuals agree, iterates agree to the
    # solver tolerance scale
    A64 = lambda x: o.laplace_of_pressure(f64(rtt), x, gr)  # noqa: E731
    rhs = o.rhs_of_pressure_equation(f64(b32[0]), f64(au), f64(av), f64(aw), f64(z_u), f64(z_v), f64(b), gr)
    res_g = rel_l2(A64(from_dev(x_g)), rhs)
    res_o = rel_l2(A64(x_o), rhs)
    assert res_g < max(3 * res_o, 1e-3), (res_g, res_o, k_g, k_o)
    assert rel_l2(xg, xo) < 0.05, (rel_l2(xg, xo), k_g, k_o)


def test_base_state_fp64(case):
    from lex_amd.setup_lex import base_state_gpu
    gr = case["gr"]
    z = (np.arange(gr.nz + 2) - 0.5) * gr.dz
    rh = 0.8 - 0.6 * np.clip(z, 0, 10000.0) / 10000.0
    got = base_state_gpu(gr.nz, gr.dz, 300.0, rh)
    ref = o.hydrostatic_base_state(gr, 300.0, lambda zz: 0.8 - 0.6 * np.clip(zz, 0, 10000.0) / 10000.0)
    for k in ("pi0", "rho0", "rho0_theta0", "qv0"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-13, atol=1e-15, err_msg=k)
    assert got["iters"] == ref["iters"]


@pytest.mark.parametrize("shape", [(16, 12, 12), (20, 13, 70), (9, 7, 130), (10, 9, 200)],
                         ids=lambda s: "x".join(map(str, s)))
def test_acoustic_substep_pair_equals_two_substeps(shape):
    """lex_acoustic_substep_pair (two substeps in one launch; the halo rows and x ends of a
    workgroup's strip are recomputed at their periodic images) against two
    lex_acoustic_substep calls on periodically padded inputs: bit-identical, ghosts
    included.  Ragged strips (ny not a multiple of the strip, nz not a multiple of 64) and
    the 4-row strips of nz > 128 are covered."""
    import torch
    from lex_amd import _lib
    nx, ny, nz = shape
    rng = np.random.default_rng(31 + nz)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()

    u = dev(o.padding3u_array(rng.normal(0, 1, (nx + 1, ny, nz))))
    v = dev(o.padding3v_array(rng.normal(0, 1, (nx, ny + 1, nz))))
    w = dev(o.set_w_bc(o.padding3_array(rng.normal(0, 0.5, (nx, ny, nz + 1)))))
    p = dev(o.padding3_array(rng.normal(0, 1e-3, (nx, ny, nz))))
    ps = dev(o.padding3_array(rng.normal(0, 1e-3, (nx, ny, nz))))
    tr = dev(o.padding3_array(300.0 + rng.normal(0, 1, (nx, ny, nz))))
    ru = dev(rng.normal(0, 1e-2, (nx + 1, ny, nz)))
    rv = dev(rng.normal(0, 1e-2, (nx, ny + 1, nz)))
    rw = dev(rng.normal(0, 1e-2, (nx, ny, nz + 1)))
    rc = dev(1.1 - 0.3 * np.linspace(0, 1, nz + 2))
    pc = dev(1.0 - 0.1 * np.linspace(0, 1, nz + 2))
    g = _lib.grid(nx, ny, nz, 100.0, 100.0, 100.0)
    dts = 0.1
    ins = (u, v, w, p)
    mid = [torch.zeros_like(a) for a in ins]
    two = [torch.zeros_like(a) for a in ins]
    one = [torch.zeros_like(a) for a in ins]
    rest = [_lib.ptr(a) for a in (ps, ru, rv, rw, tr)]

    def run(name, src, dst):
        _lib.call(name, g, _lib.ptr(rc), _lib.ptr(pc), *[_lib.ptr(a) for a in src], *rest, dts,
                  *[_lib.ptr(a) for a in dst], _lib.stream())

    run("lex_acoustic_substep", ins, mid)
    run("lex_acoustic_substep", mid, two)
    run("lex_acoustic_substep_pair", ins, one)
    torch.cuda.synchronize()
    for name, a, b in zip(("u", "v", "w", "pip"), two, one):
        assert torch.isfinite(a).all(), name
        assert torch.equal(a, b), f"{name}: max |diff| {float((a - b).abs().max()):.3e}"
