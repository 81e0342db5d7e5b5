"""Advection operators with the reference's signatures (advection.py), on HIP kernels.

All arrays are float32 device tensors in the reference's padded layout; the grid
coordinate arguments (x3d4u, ...) only supply the uniform spacings.
"""
from __future__ import annotations

import torch

from . import _lib
from ._core import column, empty, grid_from


def get_divergence(rho0, u, v, w, x3d4u, y3d4v, z3d4w):
    """div(rho0 u) on the padded scalar array (advection.py:200-226)."""
    g = grid_from(rho0, x3d4u, y3d4v, z3d4w)
    out = torch.empty_like(rho0)
    _lib.call("lex_div_rho", g, _lib.ptr(column(rho0)), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), _lib.ptr(out),
              _lib.stream())
    return out


def get_divergence2(u, v, w, x3d4u, y3d4v, z3d4w):
    """div(u) on the padded scalar array (advection.py:229-236)."""
    nx, ny, nz = u.shape[0] - 7, u.shape[1] - 6, u.shape[2] - 2
    g = _lib.grid(nx, ny, nz, float(x3d4u[1, 0, 0] - x3d4u[0, 0, 0]), float(y3d4v[0, 1, 0] - y3d4v[0, 0, 0]),
                  float(z3d4w[0, 0, 1] - z3d4w[0, 0, 0]))
    out = torch.empty((nx + 6, ny + 6, nz + 2), dtype=torch.float32, device=u.device)
    _lib.call("lex_div", g, _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), _lib.ptr(out), _lib.stream())
    return out


def advection_scalar(rho0, scalar, u, v, w, weps, flow_divergence, scalar_sfc_flux, x3d4u, y3d4v, z3d4w):
    """WENO flux-divergence tendency of a scalar at interior cells (advection.py:9-34)."""
    g = grid_from(rho0, x3d4u, y3d4v, z3d4w)
    out = empty(g.nx, g.ny, g.nz, like=rho0)
    sfc = None if scalar_sfc_flux is None else scalar_sfc_flux.to(torch.float32).contiguous()
    _lib.call("lex_adv_scalar", g, _lib.ptr(column(rho0)), _lib.ptr(scalar), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w),
              float(weps), _lib.ptr(flow_divergence), _lib.ptr(sfc), _lib.ptr(out), _lib.stream())
    return out


def advection_u(rho0, u, v, w, weps, flow_divergence, u_sfc_flux, x3d, y3d4v, z3d4w):
    """u-momentum tendency at the nx+1 interior u faces (advection.py:37-74)."""
    g = grid_from(rho0, x3d, y3d4v, z3d4w)
    out = empty(g.nx + 1, g.ny, g.nz, like=rho0)
    sfc = None if u_sfc_flux is None else u_sfc_flux.to(torch.float32).contiguous()
    _lib.call("lex_adv_u", g, _lib.ptr(column(rho0)), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), float(weps),
              _lib.ptr(flow_divergence), _lib.ptr(sfc), _lib.ptr(out), _lib.stream())
    return out


def advection_v(rho0, u, v, w, weps, flow_divergence, v_sfc_flux, x3d4u, y3d, z3d4w):
    """v-momentum tendency at the ny+1 interior v faces (advection.py:77-114)."""
    g = grid_from(rho0, x3d4u, y3d, z3d4w)
    out = empty(g.nx, g.ny + 1, g.nz, like=rho0)
    sfc = None if v_sfc_flux is None else v_sfc_flux.to(torch.float32).contiguous()
    _lib.call("lex_adv_v", g, _lib.ptr(column(rho0)), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), float(weps),
              _lib.ptr(flow_divergence), _lib.ptr(sfc), _lib.ptr(out), _lib.stream())
    return out


def advection_w(rho0, u, v, w, weps, flow_divergence, x3d4u, y3d4v, z3d):
    """w-momentum tendency at the nz+1 interior w levels (advection.py:117-149)."""
    g = grid_from(rho0, x3d4u, y3d4v, z3d)
    out = empty(g.nx, g.ny, g.nz + 1, like=rho0)
    _lib.call("lex_adv_w", g, _lib.ptr(column(rho0)), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), float(weps),
              _lib.ptr(flow_divergence), _lib.ptr(out), _lib.stream())
    return out


def advection_pi(rho0, pi0, pip, u, v, w, x3d, y3d, z3d, cc1=None, cc2=None):
    """Low-order -(u.grad pi') tendency (advection.py:152-197); cc1 = cc2 = 1/2 on uniform grids."""
    g = grid_from(rho0, x3d, y3d, z3d)
    out = empty(g.nx, g.ny, g.nz, like=rho0)
    _lib.call("lex_adv_pi", g, _lib.ptr(column(rho0)), _lib.ptr(column(pi0)), _lib.ptr(pip), _lib.ptr(u),
              _lib.ptr(v), _lib.ptr(w), _lib.ptr(out), _lib.stream())
    return out
