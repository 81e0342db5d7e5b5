"""Hand-written adjoint of the compressible WS-RK3 step and a differentiable sprint.

The reference obtains gradients only implicitly, through ``jax.value_and_grad`` of its
training loss (lex_dl_train.py:114), without rematerialisation.  Here one step is
recorded by ``lex_rk3_step_tape`` (stage states, frozen SGS fields, rhs, theta_rho and every
acoustic substep input) and reversed by ``lex_rk3_step_vjp_taped`` (the stages backwards,
then the Smagorinsky chain); ``lex_rk3_step_vjp`` does both in one call.  ``WsRk3Sprint``
wraps N steps in a ``torch.autograd.Function`` that keeps the full state every k steps; its
backward pass records each segment once from its checkpoint and reverses it.

Gradients flow to the initial potential-temperature perturbation theta'_0 (interior) and
to the Smagorinsky constant Cs; the other initial fields are constants.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._core import column
from .one_sprint_functions import fused_step

_WSV: dict = {}
_WSS: dict = {}


def _vjp_workspace(g, dt, n_sound, n_passive, device):
    key = (g.nx, g.ny, g.nz, float(dt), int(n_sound), n_passive, str(device))
    ws = _WSV.get(key)
    if ws is None:
        nbytes = _lib.load().lex_rk3_vjp_workspace_bytes(g, float(dt), int(n_sound), int(n_passive))
        if nbytes == 0:
            raise _lib.LexError("lex_rk3_vjp_workspace_bytes: invalid grid or options")
        _WSV.clear()
        ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        _WSV[key] = ws
    return ws


def _base_cols(base_state):
    rho0_theta0, rho0, theta0, pi0, qv0 = base_state[:5]
    cols = [column(a) for a in (rho0, theta0, qv0, pi0, rho0_theta0)]
    return cols, _lib.BaseCols(*[c.data_ptr() for c in cols])


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[_lib.ptr(t) for t in ts])


def step_vjp(phys_state, lam_out, base_state, grid, dt, n_sound, cs, prandtl, turb=1):
    """Adjoint of one solver_opt-2 WS-RK3 step.

    phys_state: the step's input state (padded tensors, not modified); lam_out: adjoint of
    the step's output state (padded; not modified).  Returns (lam_in, dcs): the adjoint of
    the input state (padded, ghosts folded onto the interior) and dL/dCs as a device
    float64 tensor of shape (1,).
    """
    n_passive = len(phys_state) - 6
    opts = _lib.StepOpts(float(dt), int(n_sound), int(turb), float(cs), float(prandtl), n_passive, 1)
    cols, bc = _base_cols(base_state)
    lo = tuple(t.clone() for t in lam_out)
    li = tuple(torch.empty_like(t) for t in phys_state)
    dcs = torch.zeros(1, dtype=torch.float64, device=phys_state[0].device)
    ws = _vjp_workspace(grid, dt, n_sound, n_passive, phys_state[0].device)
    _lib.call("lex_rk3_step_vjp", grid, opts, bc, _ptrs(phys_state), _ptrs(lo), _ptrs(li), dcs.data_ptr(),
              _lib.ptr(ws), ws.numel(), _lib.stream())
    del cols
    return li, dcs


def _bytes(fn, g, dt, n_sound, n_passive):
    nb = getattr(_lib.load(), fn)(g, float(dt), int(n_sound), int(n_passive))
    if nb == 0:
        raise _lib.LexError(f"{fn}: invalid grid or options")
    return nb


def step_tape(phys_state, base_state, grid, dt, n_sound, cs, prandtl, tape=None, turb=1):
    """One step recorded for the reverse pass: returns (tape, output state); phys_state is
    not modified.  `tape` may be a preallocated uint8 tensor of lex_rk3_tape_bytes."""
    n_passive = len(phys_state) - 6
    if tape is None:
        tape = torch.empty(_bytes("lex_rk3_tape_bytes", grid, dt, n_sound, n_passive), dtype=torch.uint8,
                           device=phys_state[0].device)
    out = tuple(torch.empty_like(t) for t in phys_state)
    opts = _lib.StepOpts(float(dt), int(n_sound), int(turb), float(cs), float(prandtl), n_passive, 1)
    cols, bc = _base_cols(base_state)
    _lib.call("lex_rk3_step_tape", grid, opts, bc, _ptrs(phys_state), _ptrs(out), _lib.ptr(tape), tape.numel(),
              _lib.stream())
    del cols
    return tape, out


def step_vjp_taped(phys_state, tape, lam_out, base_state, grid, dt, n_sound, cs, prandtl, turb=1):
    """Reverse of the step recorded by step_tape (same inputs); returns (lam_in, dcs)."""
    n_passive = len(phys_state) - 6
    opts = _lib.StepOpts(float(dt), int(n_sound), int(turb), float(cs), float(prandtl), n_passive, 1)
    cols, bc = _base_cols(base_state)
    lo = tuple(t.clone() for t in lam_out)
    li = tuple(torch.empty_like(t) for t in phys_state)
    dcs = torch.zeros(1, dtype=torch.float64, device=phys_state[0].device)
    key = ("adj", grid.nx, grid.ny, grid.nz, float(dt), int(n_sound), n_passive, str(phys_state[0].device))
    ws = _WSV.get(key)
    if ws is None:
        _WSV.clear()
        ws = torch.empty(_bytes("lex_rk3_adjoint_bytes", grid, dt, n_sound, n_passive), dtype=torch.uint8,
                         device=phys_state[0].device)
        _WSV[key] = ws
    _lib.call("lex_rk3_step_vjp_taped", grid, opts, bc, _ptrs(phys_state), _lib.ptr(tape), tape.numel(), _ptrs(lo),
              _ptrs(li), dcs.data_ptr(), _lib.ptr(ws), ws.numel(), _lib.stream())
    del cols
    return li, dcs


def _adj_workspace(grid, dt, n_sound, n_passive, device):
    key = ("adj", grid.nx, grid.ny, grid.nz, float(dt), int(n_sound), n_passive, str(device))
    ws = _WSV.get(key)
    if ws is None:
        _WSV.clear()
        ws = torch.empty(_bytes("lex_rk3_adjoint_bytes", grid, dt, n_sound, n_passive), dtype=torch.uint8,
                         device=device)
        _WSV[key] = ws
    return ws


def step_tape_ext_sgs(phys_state, sgs, base_state, grid, dt, n_sound, tape=None):
    """step_tape with caller-supplied SGS tendencies sgs = (su, sv, sw, s_theta, s_qv, ...)
    frozen over the step (the learned net's); returns (tape, output state)."""
    n_passive = len(phys_state) - 6
    if tape is None:
        tape = torch.empty(_bytes("lex_rk3_tape_bytes", grid, dt, n_sound, n_passive), dtype=torch.uint8,
                           device=phys_state[0].device)
    out = tuple(torch.empty_like(t) for t in phys_state)
    opts = _lib.StepOpts(float(dt), int(n_sound), 2, 0.18, 1.0 / 3.0, n_passive, 1)
    cols, bc = _base_cols(base_state)
    _lib.call("lex_rk3_step_tape_ext_sgs", grid, opts, bc, _ptrs(phys_state), _lib.ptr(sgs[0]), _lib.ptr(sgs[1]),
              _lib.ptr(sgs[2]), _ptrs(sgs[3:]), _ptrs(out), _lib.ptr(tape), tape.numel(), _lib.stream())
    del cols
    return tape, out


def step_vjp_taped_ext_sgs(phys_state, tape, lam_out, base_state, grid, dt, n_sound):
    """Reverse of step_tape_ext_sgs: returns (lam_in, lam_sgs) with lam_sgs the adjoints of
    (su, sv, sw, s_theta, s_qv, ...)."""
    n_passive = len(phys_state) - 6
    g = grid
    dev = phys_state[0].device
    opts = _lib.StepOpts(float(dt), int(n_sound), 2, 0.18, 1.0 / 3.0, n_passive, 1)
    cols, bc = _base_cols(base_state)
    lo = tuple(t.clone() for t in lam_out)
    li = tuple(torch.empty_like(t) for t in phys_state)
    lsu = torch.empty((g.nx + 1, g.ny, g.nz), dtype=torch.float32, device=dev)
    lsv = torch.empty((g.nx, g.ny + 1, g.nz), dtype=torch.float32, device=dev)
    lsw = torch.empty((g.nx, g.ny, g.nz + 1), dtype=torch.float32, device=dev)
    lss = tuple(torch.empty((g.nx, g.ny, g.nz), dtype=torch.float32, device=dev) for _ in range(2 + n_passive))
    ws = _adj_workspace(grid, dt, n_sound, n_passive, dev)
    _lib.call("lex_rk3_step_vjp_taped_ext_sgs", grid, opts, bc, _ptrs(phys_state), _lib.ptr(tape), tape.numel(),
              _ptrs(lo), _ptrs(li), _lib.ptr(lsu), _lib.ptr(lsv), _lib.ptr(lsw), _ptrs(lss), _lib.ptr(ws), ws.numel(),
              _lib.stream())
    del cols
    return li, (lsu, lsv, lsw) + lss


def sgs_apply_vjp(grid, lsgs, out_scale):
    """Adjoint of lex_sgs_apply: (lsu, lsv, lsw, ls_theta, ls_qv) -> ly (nx, ny, nz, 5)."""
    g = grid
    ly = torch.empty((g.nx, g.ny, g.nz, 5), dtype=torch.float32, device=lsgs[0].device)
    _lib.call("lex_sgs_apply_vjp", g, _lib.ptr(lsgs[0]), _lib.ptr(lsgs[1]), _lib.ptr(lsgs[2]), _lib.ptr(lsgs[3]),
              _lib.ptr(lsgs[4]), out_scale, _lib.ptr(ly), _lib.stream())
    return ly


def sgs_features_vjp(grid, lfeat, scale, lam):
    """Adds the adjoint of lex_sgs_features (normalised features' adjoint lfeat) into the
    padded state adjoint lam = (theta, u, v, w, pip, qv, ...) and folds the ghost lines."""
    _lib.call("lex_sgs_features_vjp", grid, _lib.ptr(lfeat.contiguous()), scale, _lib.ptr(lam[1]), _lib.ptr(lam[2]),
              _lib.ptr(lam[3]), _lib.ptr(lam[0]), _lib.ptr(lam[5]), _lib.stream())
    for t, mode in ((lam[0], 0), (lam[1], 1), (lam[2], 2), (lam[3], 3), (lam[5], 0)):
        fold(t, mode, grid)
    return lam


def fold(adj, mode, grid):
    """Add ghost adjoints onto their interior sources in place (mode 0 c, 1 u, 2 v, 3 w)."""
    _lib.call("lex_fold", grid, int(mode), _lib.ptr(adj), _lib.stream())
    return adj


# ---------------------------------------------------------------- per-operator VJPs
def adv_scalar_vjp(grid, rho0c, s, u, v, w, lam, weps, ls, lu, lv, lw):
    _lib.call("lex_adv_scalar_vjp", grid, _lib.ptr(rho0c), _lib.ptr(s), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w),
              _lib.ptr(lam), float(weps), _lib.ptr(ls), _lib.ptr(lu), _lib.ptr(lv), _lib.ptr(lw), _lib.stream())


def adv_momentum_vjp(which, grid, rho0c, u, v, w, lam, lu, lv, lw):
    _lib.call(f"lex_adv_{which}_vjp", grid, _lib.ptr(rho0c), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w),
              _lib.ptr(lam), _lib.ptr(lu), _lib.ptr(lv), _lib.ptr(lw), _lib.stream())


def buoyancy_vjp(grid, theta0c, qv0c, th, qv, lam_b8w, lth, lqv):
    _lib.call("lex_buoyancy_vjp", grid, _lib.ptr(theta0c), _lib.ptr(qv0c), _lib.ptr(th), _lib.ptr(qv),
              _lib.ptr(lam_b8w), _lib.ptr(lth), _lib.ptr(lqv), _lib.stream())


def theta_rho_vjp(th, qv, ltr, lth, lqv):
    _lib.call("lex_theta_rho_vjp", th.numel(), _lib.ptr(th), _lib.ptr(qv), _lib.ptr(ltr), _lib.ptr(lth),
              _lib.ptr(lqv), _lib.stream())


def acoustic_substep_vjp(grid, rho0c, pi0c, u, v, w, p, ps, ru, rv, rw, tr, dts, luo, lvo, lwo, lpo,
                         lu, lv, lw, lp, lps, ltr, lru, lrv, lrw):
    """luo..lpo are folded in place; scratch for the face adjoints is allocated here."""
    lU, lV, lW = torch.empty_like(ru), torch.empty_like(rv), torch.empty_like(rw)
    _lib.call("lex_acoustic_substep_vjp", grid, *[_lib.ptr(t) for t in (rho0c, pi0c, u, v, w, p, ps, ru, rv, rw,
                                                                        tr)],
              float(dts), *[_lib.ptr(t) for t in (luo, lvo, lwo, lpo, lU, lV, lW, lu, lv, lw, lp, lps, ltr, lru,
                                                  lrv, lrw)], _lib.stream())


def smag_vjp(grid, rho0c, u, v, w, th, qv, scalars, cs, prandtl, lsu, lsv, lsw, lss, lu, lv, lw, lth, lqv,
             lscalars, lcs=None):
    key = (grid.nx, grid.ny, grid.nz, str(u.device))
    ws = _WSS.get(key)
    if ws is None:
        ws = torch.empty(_lib.load().lex_smag_vjp_workspace_bytes(grid), dtype=torch.uint8, device=u.device)
        _WSS.clear()
        _WSS[key] = ws
    n = len(scalars)
    _lib.call("lex_smag_vjp", grid, _lib.ptr(rho0c), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), _lib.ptr(th),
              _lib.ptr(qv), n, _ptrs(scalars), float(cs), float(prandtl), _lib.ptr(lsu), _lib.ptr(lsv),
              _lib.ptr(lsw), _ptrs(lss), _lib.ptr(lu), _lib.ptr(lv), _lib.ptr(lw), _lib.ptr(lth), _lib.ptr(lqv),
              _ptrs(lscalars), None if lcs is None else lcs.data_ptr(), _lib.ptr(ws), ws.numel(), _lib.stream())


# ---------------------------------------------------------------- differentiable sprint
def _interior(a):
    return a[3:-3, 3:-3, 1:-1]


def _pad_scalar(interior, grid):
    out = torch.empty((grid.nx + 6, grid.ny + 6, grid.nz + 2), dtype=torch.float32, device=interior.device)
    _lib.call("lex_pad", 0, grid.nx, grid.ny, grid.nz, 3, _lib.ptr(interior.contiguous()), _lib.ptr(out),
              _lib.stream())
    return out


class WsRk3Sprint(torch.autograd.Function):
    """theta'_0 (interior), Cs -> theta' after nsteps WS-RK3 steps (interior).

    forward(theta_p0, cs, state0, base_state, grid, cfg) with cfg = dict(dt, n_sound,
    prandtl, nsteps, checkpoint_every).  state0 supplies u, v, w, pi', qv and passive
    scalars (its theta is replaced by theta0 + theta'_0).  Checkpoints of the full state
    are kept every `checkpoint_every` steps; the backward pass recomputes each segment
    with the recording step (step_tape) and walks it back with step_vjp_taped.  The last
    segment is recorded during the forward pass itself (the backward pass needs one
    segment's records at a time anyway), so it is never recomputed.
    """

    @staticmethod
    def forward(ctx, theta_p0, cs, state0, base_state, grid, cfg):
        cs_f = float(cs)
        th0_int = _interior(base_state[2]).contiguous()
        th = _pad_scalar(th0_int + theta_p0.detach().to(torch.float32), grid)
        x = (th,) + tuple(t.clone() for t in state0[1:])
        k = int(cfg["checkpoint_every"])
        n = int(cfg["nsteps"])
        ckpts = {0: tuple(t.clone() for t in x)}
        s_last = ((n - 1) // k) * k if n > 0 else 0
        seg_states, seg_tapes = [], []
        with torch.no_grad():
            for i in range(n):
                if i < s_last:
                    fused_step(x, base_state, grid, (2, False, False, 0, False, False, 1), dt=cfg["dt"],
                               n_sound=cfg["n_sound"], cs=cs_f, prandtl=cfg["prandtl"])
                    if (i + 1) % k == 0:
                        ckpts[i + 1] = tuple(t.clone() for t in x)
                else:
                    tp, y = step_tape(x, base_state, grid, cfg["dt"], cfg["n_sound"], cs_f, cfg["prandtl"])
                    seg_states.append(x)
                    seg_tapes.append(tp)
                    x = y
        ctx.ckpts = ckpts
        ctx.seg = (s_last, seg_states, seg_tapes)
        ctx.base_state, ctx.grid, ctx.cfg, ctx.cs = base_state, grid, cfg, cs_f
        ctx.cs_is_tensor = torch.is_tensor(cs)
        return _interior(x[0]) - th0_int

    @staticmethod
    def backward(ctx, grad_out):
        cfg, grid, base = ctx.cfg, ctx.grid, ctx.base_state
        n, k = int(cfg["nsteps"]), int(cfg["checkpoint_every"])
        x0 = ctx.ckpts[0]
        lam = tuple(torch.zeros_like(t) for t in x0)
        _interior(lam[0]).copy_(grad_out.to(torch.float32))
        dcs = torch.zeros(1, dtype=torch.float64, device=x0[0].device)
        starts = sorted(ctx.ckpts.keys())
        s_last, seg_states, tapes = ctx.seg
        for s in reversed(starts):
            end = min(s + k, n)
            if s == s_last:
                states = seg_states  # recorded by the forward pass
            else:
                # record every step of the segment once (reusing the record buffers), then
                # reverse it from the records
                states = [ctx.ckpts[s]]
                for t in range(end - s):
                    tp = tapes[t] if t < len(tapes) else None
                    tp, y = step_tape(states[-1], base, grid, cfg["dt"], cfg["n_sound"], ctx.cs, cfg["prandtl"],
                                      tape=tp)
                    if t >= len(tapes):
                        tapes.append(tp)
                    states.append(y)
            for t in reversed(range(end - s)):
                lam, d = step_vjp_taped(states[t], tapes[t], lam, base, grid, cfg["dt"], cfg["n_sound"], ctx.cs,
                                        cfg["prandtl"])
                dcs += d
        g_theta = _interior(lam[0]).clone()
        g_cs = dcs.to(torch.float32).reshape(()) if ctx.cs_is_tensor else None
        ctx.ckpts = None
        ctx.seg = None
        return g_theta, g_cs, None, None, None, None


def sprint_theta(theta_p0, cs, state0, base_state, grid, dt, n_sound, nsteps, checkpoint_every, prandtl=1.0 / 3.0):
    """Differentiable nsteps-step run: returns theta' (interior) at the last step."""
    cfg = dict(dt=float(dt), n_sound=int(n_sound), prandtl=float(prandtl), nsteps=int(nsteps),
               checkpoint_every=int(checkpoint_every))
    return WsRk3Sprint.apply(theta_p0, cs, state0, base_state, grid, cfg)
