"""WS-RK3 step and sprint with the reference's signatures (one_sprint_functions.py).

For the compressible core (solver_opt 2) the whole step runs in the fused device driver
`lex_rk3_step` (one C call per step); the pseudo-incompressible core (solver_opt 1)
composes the stage functions of lex_amd.one_step_integration.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from . import boundary_conditions as bc
from . import namelist_n_constants as nl
from . import one_step_integration as one
from . import pressure_gradient_coriolis as pgc
from ._core import column, grid_from

_WS: dict = {}


def _workspace(g, n_passive, device):
    key = (g.nx, g.ny, g.nz, n_passive, str(device))
    ws = _WS.get(key)
    if ws is None:
        nbytes = _lib.load().lex_rk3_workspace_bytes(g, n_passive)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        _WS.clear()
        _WS[key] = ws
    return ws


def fused_step(phys_state, base_state, grid, model_opt, dt=None, n_sound=None, cs=None, prandtl=None, impl=0):
    """Advance a compressible (solver_opt 2 or 3) state by one WS-RK3 step in place on the
    device.  impl 0 runs the fused stage kernels, impl 1 one kernel per reference routine;
    solver_opt 3 takes the vertically implicit acoustic substeps.
    """
    solver_opt, cor_opt, turb_opt = one._check_opts(model_opt, physics=True)
    if solver_opt not in (2, 3):
        raise NotImplementedError("the fused step covers solver_opt 2 and 3")
    if cor_opt not in (0, 1, 2):
        raise NotImplementedError(f"cor_opt {cor_opt} does not exist in the reference")
    rho0_theta0, rho0, theta0, pi0, qv0 = base_state[:5]
    n_passive = len(phys_state) - 6
    sfc = bool(model_opt[4])
    opts = _lib.StepOpts(float(nl.dt if dt is None else dt), int(nl.n_sound if n_sound is None else n_sound),
                         int(turb_opt), float(nl.cs if cs is None else cs),
                         float(nl.prandtl if prandtl is None else prandtl), n_passive, int(impl), int(solver_opt),
                         int(bool(model_opt[1])), int(bool(model_opt[2])), float(nl.z_damping), float(nl.rd_alpha),
                         int(sfc), int(cor_opt), float(nl.fCor), float(pgc.U_GEO), float(pgc.V_GEO))
    cols = [column(a) for a in (rho0, theta0, qv0, pi0, rho0_theta0)]
    if sfc:
        if len(base_state) < 6 or base_state[5] is None:
            raise ValueError("sfc_opt needs the surface temperature base_state[5]")
        ts = torch.as_tensor(base_state[5], dtype=torch.float32, device=phys_state[0].device).contiguous()
        cols.append(ts)
    bcols = _lib.BaseCols(*[c.data_ptr() for c in cols])
    ptrs = (ctypes.c_void_p * len(phys_state))(*[_lib.ptr(a) for a in phys_state])
    ws = _workspace(grid, n_passive, phys_state[0].device)
    _lib.call("lex_rk3_step", grid, opts, bcols, ptrs, _lib.ptr(ws), ws.numel(), _lib.stream())
    return phys_state


def ws_rk3_big_step(phys_state, base_state, grids, model_opt):
    """One Wicker-Skamarock RK3 step (one_sprint_functions.py:10-30).

    Returns (phys_state, sfc_others) with new tensors, as the reference does.
    """
    solver_opt = model_opt[0]
    if solver_opt in (2, 3):
        g = grid_from(base_state[1], *grids[3:6])
        new = tuple(a.clone() for a in phys_state)
        sfc_others = _first_stage_sfc(phys_state, base_state, g, model_opt)
        fused_step(new, base_state, g, model_opt)
        return new, sfc_others
    xi0 = phys_state
    xi1, _, sfc_others, heating, sgs = one.rk_sub_step0(xi0, xi0, base_state, grids, model_opt, nl.dt / 3.0)
    xi2, _ = one.rk_sub_step1(xi0, xi1, base_state, grids, heating, sfc_others, sgs, model_opt, nl.dt / 2.0)
    xi3, _ = one.rk_sub_step2(xi0, xi2, base_state, grids, heating, sfc_others, sgs, model_opt, nl.dt)
    return xi3, sfc_others


def ws_rk3_big_step_staged(phys_state, base_state, grids, model_opt):
    """The same step through the stage functions (used to cross-check the fused driver)."""
    xi0 = phys_state
    xi1, _, sfc_others, heating, sgs = one.rk_sub_step0(xi0, xi0, base_state, grids, model_opt, nl.dt / 3.0)
    xi2, _ = one.rk_sub_step1(xi0, xi1, base_state, grids, heating, sfc_others, sgs, model_opt, nl.dt / 2.0)
    xi3, _ = one.rk_sub_step2(xi0, xi2, base_state, grids, heating, sfc_others, sgs, model_opt, nl.dt)
    return xi3, sfc_others


def ws_rk3_sprint(phys_state, base_state, grids, model_opt):
    """sprint_n steps (one_sprint_functions.py:109-115)."""
    sfc_others = None
    for _ in range(nl.sprint_n):
        phys_state, sfc_others = ws_rk3_big_step(phys_state, base_state, grids, model_opt)
    return phys_state, sfc_others


def ws_rk3_sprint_dl(phys_state, base_state, grids, model_opt, dl_params, scaling_params):
    """sprint_n steps, each followed by the network's post-step correction
    (one_sprint_functions.py:118-165).  dl_params: the network (a module mapping the
    (nx, ny, nz, 5) input to the (nx, ny, nz, 5) correction, e.g. dl_models.simpleCNN());
    scaling_params: the reference's 14-tuple of scaling extremes (theta_min, theta_max,
    u_min, ..., qv_la_max).  Returns (phys_state, sfc_others) as new tensors."""
    from . import dl_models as dlm
    state = tuple(a.clone() for a in phys_state)
    norm = dlm.level_norm(scaling_params)
    ranges = dlm.scaling_ranges(scaling_params, state[0].device)
    sfc_others = None
    with torch.no_grad():
        for _ in range(nl.sprint_n):
            if model_opt[0] in (2, 3):
                g = grid_from(base_state[1], *grids[3:6])
                sfc_others = _first_stage_sfc(state, base_state, g, model_opt)
                fused_step(state, base_state, g, model_opt)
            else:
                state, sfc_others = ws_rk3_big_step(state, base_state, grids, model_opt)
            x = dlm.phys_state_tuple2array(state, scaling_params, norm)
            corr = dl_params(x[0])
            dlm.apply_correction(state, corr, ranges)
    return state, sfc_others


def _first_stage_sfc(phys_state, base_state, g, model_opt):
    """sfc_others of a step: the surface fluxes of its input state (the first stage's,
    one_step_integration.py:37-52, :186), zeros with sfc_opt off."""
    if not model_opt[4]:
        return one._sfc_etc(-999, phys_state[0])
    return one._sfc_etc(-999, phys_state[0], bc.surface_fluxes(phys_state, base_state, g))


def get_first_step_flux(phys_state, base_state, grids, model_opt):
    """Surface fluxes of the state for output (one_sprint_functions.py:168-174); zero with
    sfc_opt off."""
    return _first_stage_sfc(phys_state, base_state, grid_from(base_state[1], *grids[3:6]), model_opt)
