"""Pressure-equation pieces with the reference's signatures (pressure_equations.py), on HIP."""
from __future__ import annotations

import torch

from . import _lib
from ._core import column, empty, grid_from


def padding_array(arr):
    """1-cell periodic x/y halo, z copy (pressure_equations.py:111-120)."""
    ni, nj, nk = arr.shape
    out = torch.empty((ni + 2, nj + 2, nk + 2), dtype=torch.float32, device=arr.device)
    _lib.call("lex_pad", 0, ni, nj, nk, 1, _lib.ptr(arr.contiguous()), _lib.ptr(out), _lib.stream())
    return out


def calculate_buoyancy(theta0, theta, qv0, qv):
    """B at interior cells and at w levels with extrapolated ends (pressure_equations.py:162-172)."""
    nx, ny, nz = theta.shape[0] - 6, theta.shape[1] - 6, theta.shape[2] - 2
    g = _lib.grid(nx, ny, nz, 1.0, 1.0, 1.0)
    b = empty(nx, ny, nz, like=theta)
    b8w = empty(nx, ny, nz + 1, like=theta)
    _lib.call("lex_buoyancy", g, _lib.ptr(column(theta0)), _lib.ptr(column(qv0)), _lib.ptr(theta), _lib.ptr(qv),
              _lib.ptr(b), _lib.ptr(b8w), _lib.stream())
    return b, b8w


def calculate_rtt(rho0_theta0, theta, qv, buoyancy):
    """rho0 theta0 theta_rho with the boundary-flux ghosts (pressure_equations.py:175-195)."""
    nx, ny, nz = theta.shape[0] - 6, theta.shape[1] - 6, theta.shape[2] - 2
    g = _lib.grid(nx, ny, nz, 1.0, 1.0, 1.0)
    out = torch.empty_like(theta)
    _lib.call("lex_rtt", g, _lib.ptr(column(rho0_theta0)), _lib.ptr(theta), _lib.ptr(qv),
              _lib.ptr(buoyancy.contiguous()), _lib.ptr(out), _lib.stream())
    return out


def laplace_of_pressure(x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, rtt, pi):
    """Affine matvec Cp div(rtt grad pi) (pressure_equations.py:8-48)."""
    g = grid_from(rtt, x3d, y3d, z3d)
    out = torch.empty_like(pi)
    _lib.call("lex_p_matvec", g, _lib.ptr(rtt), _lib.ptr(pi.contiguous()), _lib.ptr(out), _lib.stream())
    return out


def _rhs(g, rt_col, adv4u, adv4v, adv4w, fu, fv, buoyancy, like):
    out = empty(g.nx, g.ny, g.nz, like=like)
    _lib.call("lex_p_rhs", g, _lib.ptr(rt_col), _lib.ptr(adv4u), _lib.ptr(adv4v), _lib.ptr(adv4w), _lib.ptr(fu),
              _lib.ptr(fv), _lib.ptr(buoyancy), _lib.ptr(out), _lib.stream())
    return out


def rhs_of_pressure_equation(rho0_theta0, pi0, rtt, adv4u, adv4v, adv4w, fu, fv, buoyancy,
                             x3d, x3d4u, y3d, y3d4v, z3d4w):
    """(rhs, rhs_adv, rhs_cor, rhs_buoy, rhs_pres) (pressure_equations.py:51-77)."""
    g = grid_from(rho0_theta0, x3d4u, y3d4v, z3d4w)
    rt = column(rho0_theta0)
    rhs = _rhs(g, rt, adv4u, adv4v, adv4w, fu, fv, buoyancy, adv4u)
    zu, zv, zw = torch.zeros_like(adv4u), torch.zeros_like(adv4v), torch.zeros_like(adv4w)
    zb = torch.zeros_like(buoyancy)
    rhs_adv = _rhs(g, rt, adv4u, adv4v, adv4w, None, None, zb, adv4u)
    rhs_cor = _rhs(g, rt, zu, zv, zw, fu, fv, zb, adv4u)
    rhs_buoy = _rhs(g, rt, zu, zv, zw, None, None, buoyancy, adv4u)
    return rhs, rhs_adv, rhs_cor, rhs_buoy, 0.0
