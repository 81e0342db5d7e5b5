"""Boundary conditions on the hot path: w at the surface and lid
(boundary_conditions.py:104-109) and the bulk atmosphere-ocean surface fluxes
(atm_ocn_flux, :7-75)."""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._core import column, grid_from


def set_w_bc(w: torch.Tensor) -> torch.Tensor:
    """Zero w on padded levels 0, 1 and nz+1, nz+2 (boundary_conditions.py:104-109).

    Returns a new tensor, as the reference does.
    """
    out = w.clone()
    nx, ny, nz = w.shape[0] - 6, w.shape[1] - 6, w.shape[2] - 3
    g = _lib.grid(nx, ny, nz, 1.0, 1.0, 1.0)
    _lib.call("lex_set_w_bc", g, _lib.ptr(out), _lib.stream())
    return out


def _f32(a, like):
    return torch.as_tensor(a, dtype=torch.float32, device=like.device).contiguous()


def atm_ocn_flux(z_bottom, u_bottom, v_bottom, theta_bottom, q_bottom, rho_bottom, surface_t):
    """Surface fluxes at the atmosphere-ocean interface (boundary_conditions.py:7-75), the
    CAM bulk scheme with eight fixed Monin-Obukhov passes, on the device.  Arguments of one
    shape (any); returns (tau_x, tau_y, sen, evap, t_ref, q_ref, u10n) of that shape with
    the reference's signs (upward fluxes positive)."""
    like = theta_bottom if torch.is_tensor(theta_bottom) else torch.as_tensor(theta_bottom)
    if not like.is_cuda:
        raise _lib.LexError("atm_ocn_flux needs device tensors (got a CPU tensor)")
    shape = like.shape
    ins = [_f32(a, like).expand(shape).contiguous()
           for a in (z_bottom, u_bottom, v_bottom, theta_bottom, q_bottom, rho_bottom, surface_t)]
    outs = [torch.empty(shape, dtype=torch.float32, device=like.device) for _ in range(7)]
    arr = (ctypes.c_void_p * 7)(*[o.data_ptr() for o in outs])
    _lib.call("lex_atm_ocn_flux", int(like.numel()), *[_lib.ptr(a) for a in ins], arr, _lib.stream())
    return tuple(outs)


def surface_fluxes(phys_state, base_state, grid):
    """atm_ocn_flux on the lowest level of a padded state as rk_sub_step0 slices it
    (one_step_integration.py:37-47); base_state[5] is the (nx, ny) surface temperature."""
    th, u, v, _, _, qv = phys_state[:6]
    ts = _f32(base_state[5], th)
    outs = [torch.empty((grid.nx, grid.ny), dtype=torch.float32, device=th.device) for _ in range(7)]
    arr = (ctypes.c_void_p * 7)(*[o.data_ptr() for o in outs])
    _lib.call("lex_surface_fluxes", grid, _lib.ptr(column(base_state[1])), _lib.ptr(th), _lib.ptr(u), _lib.ptr(v),
              _lib.ptr(qv), _lib.ptr(ts), arr, _lib.stream())
    return tuple(outs)


__all__ = ["set_w_bc", "atm_ocn_flux", "surface_fluxes", "grid_from"]
