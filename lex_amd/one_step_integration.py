"""RK stages, acoustic substeps and Euler updates with the reference's signatures
(one_step_integration.py), composed from liblexhip kernels.

State tuples are (theta, u, v, w, pip, qv) as in setup_lex.py:66, optionally followed by
passive scalars (qc, qr) that are advected and SGS-diffused like qv.  Options that are
off in every LEX configuration on this path (surface fluxes, Rayleigh damping,
Newtonian heating, the pi' constant correction) raise NotImplementedError instead of
being silently ignored.
"""
from __future__ import annotations

import torch

from . import _lib
from . import namelist_n_constants as nl
from ._core import column, empty, grid_from
from .advection import get_divergence, advection_u, advection_v, advection_w
from .pressure_equations import calculate_buoyancy, calculate_rtt
from .pressure_gradient_coriolis import calculate_coriolis_force, calculate_coriolis_force_with_lspgrad
from . import turbulence as turb

WEPS_MOM = 1.0e-16    # one_step_integration.py:896
WEPS_THETA = 1.0e-17  # :905
WEPS_QV = 1.0e-18     # :913


# ------------------------------------------------------------------- H1 padding
def _pad(mode, arr):
    ni, nj, nk = arr.shape
    pi_, pj = ni + 6, nj + 6
    out = torch.empty((pi_, pj, nk + 2), dtype=torch.float32, device=arr.device)
    _lib.call("lex_pad", mode, ni, nj, nk, 3, _lib.ptr(arr.contiguous()), _lib.ptr(out), _lib.stream())
    return out


def padding3_array(arr):
    """Periodic 3-cell x/y halo, z ghost = nearest level (one_step_integration.py:950-960)."""
    return _pad(0, arr)


def padding3u_array(arr):
    """u-point halo skipping the duplicate face (one_step_integration.py:963-973)."""
    return _pad(1, arr)


def padding3v_array(arr):
    """v-point halo skipping the duplicate face (one_step_integration.py:976-985)."""
    return _pad(2, arr)


def _interior(a):
    return a[3:-3, 3:-3, 1:-1]


def _grid(grids, like):
    x3d, y3d, z3d, x3d4u, y3d4v, z3d4w = grids[:6]
    return grid_from(like, x3d4u, y3d4v, z3d4w)


# ------------------------------------------------------------------- E1 scalars
def _scalar_update(g, rho0, s0, s, u, v, w, fd, sfc, heating, sgs, dt, weps):
    out = torch.empty_like(s0)
    tend = empty(g.nx, g.ny, g.nz, like=s0)
    _lib.call("lex_scalar_update", g, _lib.ptr(column(rho0)), _lib.ptr(s0), _lib.ptr(s), _lib.ptr(u), _lib.ptr(v),
              _lib.ptr(w), _lib.ptr(fd), _lib.ptr(sfc), _lib.ptr(heating), _lib.ptr(sgs), float(dt), float(weps),
              _lib.ptr(out), _lib.ptr(tend), _lib.stream())
    return out, tend


def _dev_or_none(a, like):
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        t = a.to(device=like.device, dtype=torch.float32).contiguous()
    else:
        t = torch.as_tensor(a, dtype=torch.float32, device=like.device).contiguous()
    return None if not bool(t.any()) else t


def update_theta_euler(rho0_now, theta_now0, theta_now, u_now, v_now, w_now, flow_divergence, sfc_flux, heating,
                       x3d4u, y3d4v, z3d4w, dt):
    """theta_next = pad(theta0 + adv*dt) (one_step_integration.py:450-457)."""
    g = grid_from(rho0_now, x3d4u, y3d4v, z3d4w)
    return _scalar_update(g, rho0_now, theta_now0, theta_now, u_now, v_now, w_now, flow_divergence,
                          _dev_or_none(sfc_flux, rho0_now), _dev_or_none(heating, rho0_now), None, dt, WEPS_THETA)


def update_qv_euler(rho0_now, qv_now0, qv_now, u_now, v_now, w_now, flow_divergence, sfc_flux, x3d4u, y3d4v, z3d4w,
                    dt):
    """qv_next = pad(qv0 + adv*dt) (one_step_integration.py:460-465)."""
    g = grid_from(rho0_now, x3d4u, y3d4v, z3d4w)
    return _scalar_update(g, rho0_now, qv_now0, qv_now, u_now, v_now, w_now, flow_divergence,
                          _dev_or_none(sfc_flux, rho0_now), None, None, dt, WEPS_QV)


def compute_theta_tendency(rho0, theta, u, v, w, flow_divergence, sfc_flux, heating, x3d4u, y3d4v, z3d4w):
    """(one_step_integration.py:903-909)."""
    from .advection import advection_scalar
    t = advection_scalar(rho0, theta, u, v, w, WEPS_THETA, flow_divergence, _dev_or_none(sfc_flux, rho0),
                         x3d4u, y3d4v, z3d4w)
    h = _dev_or_none(heating, rho0)
    return t if h is None else t + h


def compute_qv_tendency(rho0, qv, u, v, w, flow_divergence, sfc_flux, x3d4u, y3d4v, z3d4w):
    """(one_step_integration.py:912-916)."""
    from .advection import advection_scalar
    return advection_scalar(rho0, qv, u, v, w, WEPS_QV, flow_divergence, _dev_or_none(sfc_flux, rho0),
                            x3d4u, y3d4v, z3d4w)


# --------------------------------------------------------------- A3 momentum prep
def prep_momentum_eqn(rho0, u, v, w, flow_divergence, u_sfc_flux, v_sfc_flux, x3d, y3d, z3d, x3d4u, y3d4v, z3d4w):
    """(adv4u, adv4v, adv4w) (one_step_integration.py:894-900)."""
    a_u = advection_u(rho0, u, v, w, WEPS_MOM, flow_divergence, _dev_or_none(u_sfc_flux, rho0), x3d, y3d4v, z3d4w)
    a_v = advection_v(rho0, u, v, w, WEPS_MOM, flow_divergence, _dev_or_none(v_sfc_flux, rho0), x3d4u, y3d, z3d4w)
    a_w = advection_w(rho0, u, v, w, WEPS_MOM, flow_divergence, x3d4u, y3d4v, z3d)
    return a_u, a_v, a_w


# ------------------------------------------------------------ M1 momentum Euler
def update_momentum_eqn_euler(u_now0, v_now0, w_now0, pi0_now, pip_now, theta_now, adv4u, adv4v, adv4w, fu, fv, b8w,
                              x3d, y3d, z3d, dt, sgs=None):
    """(one_step_integration.py:876-891); `sgs` optionally folds in the SGS add-in (:113-118)."""
    g = grid_from(pip_now, x3d, y3d, z3d)
    un, vn, wn = torch.empty_like(u_now0), torch.empty_like(v_now0), torch.empty_like(w_now0)
    du, dv, dw = torch.empty_like(adv4u), torch.empty_like(adv4v), torch.empty_like(adv4w)
    su, sv, sw = sgs if sgs is not None else (None, None, None)
    _lib.call("lex_momentum_update", g, _lib.ptr(u_now0), _lib.ptr(v_now0), _lib.ptr(w_now0), _lib.ptr(pip_now),
              _lib.ptr(theta_now), _lib.ptr(adv4u), _lib.ptr(adv4v), _lib.ptr(adv4w), _lib.ptr(fu), _lib.ptr(fv),
              _lib.ptr(b8w), _lib.ptr(su), _lib.ptr(sv), _lib.ptr(sw), float(dt), _lib.ptr(un), _lib.ptr(vn),
              _lib.ptr(wn), _lib.ptr(du), _lib.ptr(dv), _lib.ptr(dw), _lib.stream())
    return un, vn, wn, du, dv, dw


# ---------------------------------------------------------------- P4 solver
def solve_pres_eqn(pip_prev, rho0_theta0, pi0, rtt, adv4u, adv4v, adv4w, fu, fv, buoyancy,
                   x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, tol=1.0e-4, atol=1.0e-8, maxiter=100):
    """BiCGSTAB solve of the pi' equation, x0 = interior of pip_prev (one_step_integration.py:468-487).

    Returns (pip interior, info) with info the iteration count or a breakdown code
    (-10, -11) -- the reference's info is JAX's placeholder None.
    """
    g = grid_from(rho0_theta0, x3d4u, y3d4v, z3d4w)
    rt = column(rho0_theta0)
    rhs = empty(g.nx, g.ny, g.nz, like=rtt)
    _lib.call("lex_p_rhs", g, _lib.ptr(rt), _lib.ptr(adv4u), _lib.ptr(adv4v), _lib.ptr(adv4w), _lib.ptr(fu),
              _lib.ptr(fv), _lib.ptr(buoyancy), _lib.ptr(rhs), _lib.stream())
    x0 = _interior(pip_prev).contiguous()
    x = torch.empty_like(x0)
    nbytes = _lib.load().lex_bicgstab_workspace_bytes(g)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=rtt.device)
    import ctypes
    k = ctypes.c_int(0)
    _lib.call("lex_bicgstab", g, _lib.ptr(rtt), _lib.ptr(rhs), _lib.ptr(x0), _lib.ptr(x), float(tol), float(atol),
              int(maxiter), _lib.ptr(ws), nbytes, ctypes.byref(k), _lib.stream())
    return x, k.value


# --------------------------------------------------------------- S1 acoustic
ALPHA = 0.6  # off-centring of the vertically implicit acoustic step (namelist_n_constants.py:14)


def _sound(u_now, v_now, w_now, pi0, pip_now, pip_star, rhs_u, rhs_v, rhs_w, rhs_pi, theta_rho, rho0, rk_n_sound,
           dt_sound, g, implicit=False):
    if not (isinstance(rhs_pi, (int, float)) and rhs_pi == 0.0):
        raise NotImplementedError("rhs_pi is a zero placeholder in the reference (one_step_integration.py:134)")
    rc, p0c = column(rho0), column(pi0)
    cur = (u_now, v_now, w_now, pip_now)
    left = int(rk_n_sound)
    while left > 0:
        nxt = tuple(torch.empty_like(a) for a in cur)
        args = [g, _lib.ptr(rc), _lib.ptr(p0c), *[_lib.ptr(a) for a in cur], _lib.ptr(pip_star), _lib.ptr(rhs_u),
                _lib.ptr(rhs_v), _lib.ptr(rhs_w), _lib.ptr(theta_rho), float(dt_sound)]
        if implicit:
            _lib.call("lex_acoustic_substep_implicit", *args, float(ALPHA), *[_lib.ptr(a) for a in nxt], _lib.stream())
            left -= 1
        elif left >= 2 and g.nz <= 256:
            # two passes of the loop in one launch
            _lib.call("lex_acoustic_substep_pair", *args, *[_lib.ptr(a) for a in nxt], _lib.stream())
            left -= 2
        else:
            _lib.call("lex_acoustic_substep", *args, *[_lib.ptr(a) for a in nxt], _lib.stream())
            left -= 1
        cur = nxt
    return cur


def sound_rk_step0(u_now, v_now, w_now, pi0, pip_now, pip_star, rhs_u, rhs_v, rhs_w, rhs_pi, theta_rho, rho0,
                   rk_n_sound, dt_sound, x3d, y3d, z3d, x3d4u, y3d4v, z3d4w, cc1=None, cc2=None):
    """Forward-backward acoustic substeps (one_step_integration.py:519-556)."""
    g = grid_from(pip_now, x3d4u, y3d4v, z3d4w)
    return _sound(u_now, v_now, w_now, pi0, pip_now, pip_star, rhs_u, rhs_v, rhs_w, rhs_pi, theta_rho, rho0,
                  rk_n_sound, dt_sound, g)


sound_rk_step1 = sound_rk_step0   # identical bodies in the reference (:638-675)
sound_rk_step2 = sound_rk_step0   # (:757-794)


def sound_implicit_rk_step0(u_now, v_now, w_now, pi0, pip_now, pip_star, rhs_u, rhs_v, rhs_w, rhs_pi, theta_rho, rho0,
                            rk_n_sound, dt_sound, x3d, y3d, z3d, x3d4u, y3d4v, z3d4w, cc1=None, cc2=None):
    """Vertically implicit acoustic substeps, solver_opt 3 (one_step_integration.py:559-635)."""
    g = grid_from(pip_now, x3d4u, y3d4v, z3d4w)
    return _sound(u_now, v_now, w_now, pi0, pip_now, pip_star, rhs_u, rhs_v, rhs_w, rhs_pi, theta_rho, rho0,
                  rk_n_sound, dt_sound, g, implicit=True)


sound_implicit_rk_step1 = sound_implicit_rk_step0   # identical bodies in the reference (:678-754)
sound_implicit_rk_step2 = sound_implicit_rk_step0   # (:797-873)


def _theta_rho(theta, qv):
    out = torch.empty_like(theta)
    _lib.call("lex_theta_rho", theta.numel(), _lib.ptr(theta), _lib.ptr(qv), _lib.ptr(out), _lib.stream())
    return out


# ------------------------------------------------------------------ R1 stages
def _check_opts(model_opt, physics=False):
    """Validates model_opt; physics=True (the device step driver) also accepts damp_opt,
    rad_opt, sfc_opt and turb_opt 2, which the per-routine stage functions here do not
    implement."""
    solver_opt, damp_opt, rad_opt, cor_opt, sfc_opt, pic_opt, turb_opt = model_opt
    if solver_opt not in (1, 2, 3):
        raise NotImplementedError(f"solver_opt {solver_opt} does not exist in the reference")
    flags = (("pic_opt", pic_opt),)
    if not physics:
        flags = (("damp_opt", damp_opt), ("rad_opt", rad_opt), ("sfc_opt", sfc_opt)) + flags
    for name, flag in flags:
        if flag:
            raise NotImplementedError(f"{name} is not built on this path"
                                      + ("" if physics else " (damp_opt, rad_opt, sfc_opt: use the fused step)"))
    if turb_opt not in (0, 1, 2) or (turb_opt == 2 and not physics):
        raise NotImplementedError("turb_opt 2 (horizontal / vertical Smagorinsky) runs in the fused step")
    return solver_opt, cor_opt, turb_opt


def _n_sound(stage, dt):
    """Acoustic substeps and their length per stage (one_step_integration.py:121-122, :253-254, :383-384)."""
    if stage == 0:
        return 1, dt
    dt_sound = nl.dt / nl.n_sound
    return round(dt / dt_sound), dt_sound


def momentum_rhs(g, adv4u, adv4v, adv4w, fu, fv, b8w, sgs_tend=None):
    """The stage's momentum right-hand sides (one_step_integration.py:126-133): adv4u + fv
    + sgs_u, adv4v - fu + sgs_v, adv4w + b8w + sgs_w; absent terms are None."""
    ru, rv, rw = torch.empty_like(adv4u), torch.empty_like(adv4v), torch.empty_like(adv4w)
    su, sv, sw = sgs_tend[:3] if sgs_tend is not None else (None, None, None)
    _lib.call("lex_momentum_rhs", g, _lib.ptr(adv4u), _lib.ptr(adv4v), _lib.ptr(adv4w), _lib.ptr(fu), _lib.ptr(fv),
              _lib.ptr(b8w), _lib.ptr(su), _lib.ptr(sv), _lib.ptr(sw), _lib.ptr(ru), _lib.ptr(rv), _lib.ptr(rw),
              _lib.stream())
    return ru, rv, rw


def _stage(stage, phys_state_now, phys_state, base_state, grids, model_opt, dt, sgs_tend, want_tends=True):
    solver_opt, cor_opt, turb_opt = _check_opts(model_opt)
    rho0_theta0, rho0, theta0, pi0, qv0 = base_state[:5]
    x3d, y3d, z3d, x3d4u, y3d4v, z3d4w = grids[:6]
    th0, u0, v0, w0, pip0, qv0_now = phys_state_now[:6]
    th, u, v, w, pip, qv = phys_state[:6]
    pas0, pas = tuple(phys_state_now[6:]), tuple(phys_state[6:])
    g = grid_from(rho0, x3d4u, y3d4v, z3d4w)
    if stage == 0:
        if turb_opt == 1:
            s11, s22, s33, s12, s13, s23, deform = turb.compute_deformation(rho0, u, v, w, x3d, y3d, z3d, x3d4u,
                                                                            y3d4v, z3d4w)
            n2 = turb.compute_nm(th, pi0, qv, z3d)
            km, kh = turb.compute_k(deform, n2, x3d4u, y3d4v, z3d4w)
            sgs_tend = turb.compute_smag(rho0, km, kh, s11, s22, s33, s12, s13, s23, th, qv, x3d, y3d, z3d, x3d4u,
                                         y3d4v, z3d4w, extra_scalars=pas)
        else:
            sgs_tend = None
    buoyancy, b8w = calculate_buoyancy(theta0, th, qv0, qv)
    fd = get_divergence(rho0, u, v, w, x3d4u, y3d4v, z3d4w)
    adv4u, adv4v, adv4w = prep_momentum_eqn(rho0, u, v, w, fd, None, None, x3d, y3d, z3d, x3d4u, y3d4v, z3d4w)
    if cor_opt == 1:
        fu, fv = calculate_coriolis_force(u, v)
    elif cor_opt == 2:
        fu, fv = calculate_coriolis_force_with_lspgrad(u, v)
    else:
        fu = fv = None   # the reference's zero arrays (one_step_integration.py:88-89) as NULL
    tends = None
    info = -999
    trho = _theta_rho(th, qv)
    if solver_opt == 1:
        rtt = calculate_rtt(rho0_theta0, th, qv, buoyancy)
        pint, info = solve_pres_eqn(pip0, rho0_theta0, pi0, rtt, adv4u, adv4v, adv4w, fu, fv, buoyancy,
                                    x3d, x3d4u, y3d, y3d4v, z3d, z3d4w)
        pip_n = padding3_array(pint)
        sg = sgs_tend[:3] if sgs_tend is not None else None
        un, vn, wn, du, dv, dw = update_momentum_eqn_euler(u0, v0, w0, pi0, pip_n, trho, adv4u, adv4v, adv4w,
                                                           fu, fv, b8w, x3d, y3d, z3d, dt, sgs=sg)
    else:
        # rhs_u = adv4u + fv + sgs_u, rhs_v = adv4v - fu + sgs_v, rhs_w = adv4w + b8w + sgs_w
        # (one_step_integration.py:126-133), assembled on the device
        rhs_u, rhs_v, rhs_w = momentum_rhs(g, adv4u, adv4v, adv4w, fu, fv, b8w, sgs_tend)
        n, dts = _n_sound(stage, dt)
        un, vn, wn, pip_n = _sound(u0, v0, w0, pi0, pip0, pip, rhs_u, rhs_v, rhs_w, 0.0, trho, rho0, n, dts, g,
                                   implicit=solver_opt == 3)
    sg_th = sgs_tend[3] if sgs_tend is not None else None
    sg_qv = sgs_tend[4] if sgs_tend is not None else None
    thn, dth = _scalar_update(g, rho0, th0, th, u, v, w, fd, None, None, sg_th, dt, WEPS_THETA)
    qvn, dqv = _scalar_update(g, rho0, qv0_now, qv, u, v, w, fd, None, None, sg_qv, dt, WEPS_QV)
    pasn = []
    for q, (a0, a) in enumerate(zip(pas0, pas)):
        sq = sgs_tend[5 + q] if sgs_tend is not None else None
        pasn.append(_scalar_update(g, rho0, a0, a, u, v, w, fd, None, None, sq, dt, WEPS_QV)[0])
    new_state = (thn, un, vn, wn, pip_n, qvn) + tuple(pasn)
    if want_tends:
        if solver_opt == 1:
            tends = (dth, du, dv, dw, dqv)
        else:
            tends = (dth, (pip_n - pip0) / dt, (un - u0) / dt, (vn - v0) / dt, (wn - w0) / dt, dqv)
    return new_state, tends, sgs_tend, info


def _sfc_etc(info, like, fluxes=None):
    """The reference's sfc_etc tuple (one_step_integration.py:186): info, pip_const, then
    tau_x, tau_y, sen, evap, t_ref, q_ref, u10n (zeros without surface fluxes)."""
    if fluxes is None:
        nx, ny = like.shape[0] - 6, like.shape[1] - 6
        z = torch.zeros(nx, ny, dtype=torch.float32, device=like.device)
        fluxes = (z,) * 7
    return (info, -999.9) + tuple(fluxes)


def rk_sub_step0(phys_state_now, phys_state, base_state, grids, model_opt, dt):
    """First WS-RK3 stage; computes the frozen SGS tendencies (one_step_integration.py:16-188)."""
    st, tends, sgs, info = _stage(0, phys_state_now, phys_state, base_state, grids, model_opt, dt, None)
    nx, ny, nz = phys_state[0].shape[0] - 6, phys_state[0].shape[1] - 6, phys_state[0].shape[2] - 2
    heating = torch.zeros(nx, ny, nz, dtype=torch.float32, device=phys_state[0].device)
    return st, tends, _sfc_etc(info, phys_state[0]), heating, sgs


def rk_sub_step1(phys_state_now, phys_state, base_state, grids, heating, sfc_others, sgs_tend, model_opt, dt):
    """Second stage with frozen physics (one_step_integration.py:191-317)."""
    st, tends, _, _ = _stage(1, phys_state_now, phys_state, base_state, grids, model_opt, dt, sgs_tend)
    return st, tends


def rk_sub_step2(phys_state_now, phys_state, base_state, grids, heating, sfc_others, sgs_tend, model_opt, dt):
    """Third stage (one_step_integration.py:320-447)."""
    st, tends, _, _ = _stage(2, phys_state_now, phys_state, base_state, grids, model_opt, dt, sgs_tend)
    return st, tends
