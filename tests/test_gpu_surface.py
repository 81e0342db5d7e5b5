"""Surface fluxes (sfc_opt) on the device vs the oracle: the bulk atmosphere-ocean scheme
(atm_ocn_flux, boundary_conditions.py:7-75) on columns and on a state's lowest level, and
whole steps with the fluxes on (fused kernels, one kernel per routine, double-buffered
scalars, the implicit acoustic solver) vs the FP64 oracle.

Tolerance: the scheme's own FP32 error, measured as the FP32 oracle's distance from the
FP64 oracle on the same inputs, times 20 (floor 1e-5 of the field's largest value).  The
reference switches the neutral heat-transfer coefficient with the sign of z/L, so columns
with z/L near zero are ill-conditioned in any precision; the draws keep clear of it
(a warm sea under subsaturated air, or humid air over a cooler sea).
"""
import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import from_dev, rel_l2, to_dev

pytestmark = pytest.mark.gpu

FIELDS = ["tau_x", "tau_y", "sen", "evap", "t_ref", "q_ref", "u10n"]


def _columns(side, n=20000, seed=5):
    rng = np.random.default_rng(seed)
    dth, rh, wind = {"unstable": ((-3.0, -0.2), (0.5, 0.95), 15.0),
                     "stable": ((0.2, 3.0), (0.995, 1.0), 15.0),
                     "calm": ((-3.0, -0.2), (0.5, 0.95), 0.17)}[side]
    z = rng.uniform(2.0, 60.0, n)
    u = rng.uniform(-wind, wind, n)
    v = rng.uniform(-wind, wind, n)
    ts = rng.uniform(285.0, 303.0, n)
    th = ts + rng.uniform(*dth, n)
    rho = rng.uniform(1.1, 1.2, n)
    q = 0.98 * o.q_sat_sfc(ts) / rho * rng.uniform(*rh, n)
    return [a.astype(np.float32) for a in (z, u, v, th, q, rho, ts)]


def _check(got, ins64, ins32):
    e64 = o.atm_ocn_flux(*ins64)
    e32 = o.atm_ocn_flux(*ins32)
    for name, g, a, b in zip(FIELDS, got, e64, e32):
        scale = float(np.abs(a).max())
        err_g = float(np.abs(g - a).max()) / scale
        err_32 = float(np.abs(b.astype(np.float64) - a).max()) / scale
        print(f"{name}: GPU {err_g:.2e} FP32 oracle {err_32:.2e}")
        assert err_g <= max(20.0 * err_32, 1e-5), name


@pytest.mark.parametrize("side", ["unstable", "stable", "calm"])
def test_atm_ocn_flux_columns_vs_oracle(side):
    from lex_amd import boundary_conditions as bc
    ins = _columns(side)
    got = [from_dev(a) for a in bc.atm_ocn_flux(*[to_dev(a) for a in ins])]
    _check(got, [a.astype(np.float64) for a in ins], ins)
    if side == "calm":   # every wind under the 0.25 m/s floor of the scheme
        assert (np.hypot(ins[1], ins[2]) < 0.25).all()


def _setup(nx, ny, nz, dts=(1.0, 3.0), wind=4.0, quiet=False):
    """A reduced workload-recipe state with a mean wind over a sea 1-3 K warmer than the
    lowest level; quiet: no bubble and no noise, so the surface fluxes drive all change."""
    from lex_amd import namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    import torch
    case, opts = config1()
    case = case.scaled(nx, ny, nz)
    if quiet:
        case.amplitude, case.rh_bubble = 0.0, 0.0
        case.theta_noise, case.uvw_noise = ("none", 0.0, 2025), (0.0, 2026)
    phys, base, grids, _ = build_case(case)
    phys = list(phys)
    phys[1] = phys[1] + wind          # a mean wind for the stress
    phys = tuple(phys)
    th_b = from_dev(phys[0])[3:-3, 3:-3, 1]
    rng = np.random.default_rng(2)
    ts = (th_b + rng.uniform(*dts, th_b.shape)).astype(np.float32)   # a warmer sea
    base = tuple(base[:5]) + (torch.as_tensor(ts, device=phys[0].device),)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], opts["n_sound"], opts["solver_opt"], 1
    mo = (opts["solver_opt"], False, False, 0, True, False, 1)
    og = o.Grid(nx, ny, nz, case.dx, case.dy, case.dz)
    oo = o.Options(dt=opts["dt"], solver_opt=opts["solver_opt"], n_sound=opts["n_sound"], turb_opt=1, sfc_opt=True)
    return phys, base, grids, mo, og, oo, ts


def test_surface_fluxes_of_a_state_vs_oracle():
    from lex_amd import boundary_conditions as bc
    from lex_amd._core import grid_from
    phys, base, grids, _, og, _, ts = _setup(24, 20, 24)
    g = grid_from(base[1], *grids[3:6])
    got = [from_dev(a) for a in bc.surface_fluxes(phys, base, g)]
    x64 = tuple(from_dev(a).astype(np.float64) for a in phys)
    b64 = tuple(from_dev(a).astype(np.float64) for a in base[:5]) + (ts.astype(np.float64),)
    x32 = tuple(from_dev(a).astype(np.float32) for a in phys)
    b32 = tuple(from_dev(a).astype(np.float32) for a in base[:5]) + (ts,)
    e64 = o.surface_fluxes(x64, b64, og)
    e32 = o.surface_fluxes(x32, b32, og)
    for name, gg, a, b in zip(FIELDS, got, e64, e32):
        scale = float(np.abs(a).max())
        err_g, err_32 = float(np.abs(gg - a).max()) / scale, float(np.abs(b - a).max()) / scale
        print(f"{name}: GPU {err_g:.2e} FP32 oracle {err_32:.2e}")
        assert err_g <= max(20.0 * err_32, 1e-5), name
    assert (got[2] > 0).all() and (got[0] < 0).all()


def _oracle_run(phys, base, ts, og, oo, nsteps, dtype):
    xi = tuple(from_dev(a).astype(dtype) for a in phys)
    bs = tuple(from_dev(a).astype(dtype) for a in base[:5]) + (ts.astype(dtype),)
    for _ in range(nsteps):
        xi, _ = o.ws_rk3_big_step(xi, bs, og, oo)
    return xi


@pytest.mark.parametrize("flow,impl,solver,nsteps", [("quiet", 0, 2, 1), ("quiet", 0, 2, 3), ("quiet", 1, 2, 2),
                                                      ("quiet", 2, 2, 1), ("quiet", 0, 3, 2), ("bubble", 0, 2, 2)])
def test_sfc_step_vs_oracle(flow, impl, solver, nsteps):
    """Whole steps with the surface fluxes on vs the oracle, as perturbations from the
    initial state.  In the quiet flow the fluxes are the only horizontal forcing (without
    them theta and u do not change) and their effect on theta, u and qv is ten times the
    FP32 error or more, so the comparison tests them directly."""
    from lex_amd import namelist_n_constants as nl, one_sprint_functions as sp
    from lex_amd._core import grid_from
    if flow == "quiet":   # a strong flux (sea 4-8 K warmer, 8 m/s) over a state that is otherwise at rest
        phys, base, grids, mo, og, oo, ts = _setup(24, 20, 24, dts=(4.0, 8.0), wind=8.0, quiet=True)
    else:
        phys, base, grids, mo, og, oo, ts = _setup(24, 20, 24)
    if solver == 3:
        nl.solver_opt, nl.n_sound = 3, 6
        mo = (3,) + tuple(mo[1:])
        oo.solver_opt, oo.n_sound = 3, 6
    init = tuple(from_dev(a) for a in phys)
    st = tuple(a.clone() for a in phys)
    g = grid_from(base[1], *grids[3:6])
    for _ in range(nsteps):
        sp.fused_step(st, base, g, mo, impl=impl)
    got = tuple(from_dev(a) for a in st)
    r64 = _oracle_run(phys, base, ts, og, oo, nsteps, np.float64)
    r32 = _oracle_run(phys, base, ts, og, oo, nsteps, np.float32)
    for n, a, b, c, a0 in zip(["theta", "u", "v", "w", "pip", "qv"], got, r64, r32, init):
        da, db, dc = a - a0, b - a0, c - a0
        e_g, e_32 = rel_l2(da, db), rel_l2(dc, db)
        print(f"{n}: GPU {e_g:.2e} FP32 oracle {e_32:.2e}")
        assert e_g <= max(10.0 * e_32, 2e-6), n
    oo.sfc_opt = False
    plain = _oracle_run(phys, base, ts, og, oo, nsteps, np.float64)
    for q in (0, 1, 5):
        eff = np.linalg.norm(r64[q] - plain[q])
        assert eff > 0.0
        if flow == "quiet":
            assert eff > 10.0 * np.linalg.norm(r32[q] - r64[q]), q


def test_sfc_others_of_a_step():
    """ws_rk3_big_step returns the fluxes of its input state in sfc_others (the first
    stage's, one_step_integration.py:186), and get_first_step_flux the same."""
    from lex_amd import boundary_conditions as bc, one_sprint_functions as sp
    from lex_amd._core import grid_from
    phys, base, grids, mo, _, _, _ = _setup(16, 12, 16)
    new, sfc_others = sp.ws_rk3_big_step(phys, base, grids, mo)
    ref = bc.surface_fluxes(phys, base, grid_from(base[1], *grids[3:6]))
    first = sp.get_first_step_flux(phys, base, grids, mo)
    assert len(sfc_others) == 9
    for a, b, c in zip(sfc_others[2:], ref, first[2:]):
        assert np.array_equal(from_dev(a), from_dev(b)) and np.array_equal(from_dev(c), from_dev(b))
    assert not np.array_equal(from_dev(new[0]), from_dev(phys[0]))
