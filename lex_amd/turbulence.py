"""Smagorinsky SGS with the reference's signatures (turbulence.py:9-471), on HIP: the
isotropic closure (compute_k, compute_smag; turb_opt 1) and the horizontal / vertical one
(compute_k2, compute_smag2; turb_opt 2)."""
from __future__ import annotations

import torch

from . import _lib
from . import namelist_n_constants as nl
from ._core import column, empty, grid_from


def compute_deformation(rho, u, v, w, x3d, y3d, z3d, x3d4u, y3d4v, z3d4w):
    """(s11, s22, s33, s12, s13, s23, deformation), rho-weighted, padded (turbulence.py:298-388)."""
    g = grid_from(rho, x3d4u, y3d4v, z3d4w)
    outs = [torch.empty_like(rho) for _ in range(7)]
    _lib.call("lex_smag_deformation", g, _lib.ptr(column(rho)), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w),
              *[_lib.ptr(o) for o in outs], _lib.stream())
    return tuple(outs)


def compute_nm(theta, pi, qv, z3d):
    """Squared Brunt-Vaisala frequency at cells, padded (turbulence.py:391-427)."""
    nx, ny, nz = theta.shape[0] - 6, theta.shape[1] - 6, theta.shape[2] - 2
    dz = float(z3d[0, 0, 1] - z3d[0, 0, 0])
    g = _lib.grid(nx, ny, nz, 1.0, 1.0, dz)
    out = torch.empty_like(theta)
    _lib.call("lex_smag_nm", g, _lib.ptr(theta), _lib.ptr(qv), _lib.ptr(out), _lib.stream())
    return out


def compute_k(s2, n2, x3d4u, y3d4v, z3d4w, cs=None, prandtl=None):
    """Km, Kh (turbulence.py:430-446); cs defaults to the reference's 0.18."""
    g = grid_from(s2, x3d4u, y3d4v, z3d4w)
    km = torch.empty_like(s2)
    kh = torch.empty_like(s2)
    _lib.call("lex_smag_k", g, _lib.ptr(s2), _lib.ptr(n2), float(nl.cs if cs is None else cs),
              float(nl.prandtl if prandtl is None else prandtl), _lib.ptr(km), _lib.ptr(kh), _lib.stream())
    return km, kh


def compute_scalar_sgs(rho, kh, scalar, x3d4u, y3d4v, z3d4w):
    """SGS diffusion tendency of one scalar (turbulence.py:80-109)."""
    g = grid_from(rho, x3d4u, y3d4v, z3d4w)
    out = empty(g.nx, g.ny, g.nz, like=rho)
    _lib.call("lex_smag_scalar", g, _lib.ptr(column(rho)), _lib.ptr(kh), _lib.ptr(scalar), _lib.ptr(out),
              _lib.stream())
    return out


def compute_smag(rho, km, kh, s11, s22, s33, s12, s13, s23, theta, qv, x3d, y3d, z3d, x3d4u, y3d4v, z3d4w,
                 extra_scalars=()):
    """(sgs_u, sgs_v, sgs_w, sgs_theta, sgs_qv[, sgs of extra scalars]) (turbulence.py:9-111)."""
    g = grid_from(rho, x3d4u, y3d4v, z3d4w)
    su = empty(g.nx + 1, g.ny, g.nz, like=rho)
    sv = empty(g.nx, g.ny + 1, g.nz, like=rho)
    sw = empty(g.nx, g.ny, g.nz + 1, like=rho)
    _lib.call("lex_smag_momentum", g, _lib.ptr(column(rho)), _lib.ptr(km), _lib.ptr(s11), _lib.ptr(s22),
              _lib.ptr(s33), _lib.ptr(s12), _lib.ptr(s13), _lib.ptr(s23), _lib.ptr(su), _lib.ptr(sv), _lib.ptr(sw),
              _lib.stream())
    scal = tuple(compute_scalar_sgs(rho, kh, s, x3d4u, y3d4v, z3d4w) for s in (theta, qv) + tuple(extra_scalars))
    return (su, sv, sw) + scal


def compute_k2(rho, s11, s22, s33, s12, s13, s23, n2, x3d4u, y3d4v, z3d4w, z3d, cs=None):
    """(kmh, kmv, khh, khv), padded (turbulence.py:449-471): prandtl 1, so khh is kmh and
    khv is kmv; z3d gives the cell-centre heights of the uniform grid."""
    g = grid_from(rho, x3d4u, y3d4v, z3d4w)
    kmh = torch.empty_like(s11)
    kmv = torch.empty_like(s11)
    _lib.call("lex_smag_k2", g, _lib.ptr(column(rho)), *[_lib.ptr(a) for a in (s11, s22, s33, s12, s13, s23, n2)],
              float(nl.cs if cs is None else cs), _lib.ptr(kmh), _lib.ptr(kmv), _lib.stream())
    return kmh, kmv, kmh, kmv


def compute_smag2(rho, kmh, kmv, khh, khv, s11, s22, s33, s12, s13, s23, theta, qv, x3d, y3d, z3d, x3d4u, y3d4v,
                  z3d4w, extra_scalars=()):
    """(sgs_u, sgs_v, sgs_w, sgs_theta, sgs_qv[, ...]) with horizontal and vertical
    coefficients (turbulence.py:114-295); sgs_w is zero at the surface."""
    g = grid_from(rho, x3d4u, y3d4v, z3d4w)
    su = empty(g.nx + 1, g.ny, g.nz, like=rho)
    sv = empty(g.nx, g.ny + 1, g.nz, like=rho)
    sw = empty(g.nx, g.ny, g.nz + 1, like=rho)
    _lib.call("lex_smag2_momentum", g, _lib.ptr(column(rho)), _lib.ptr(kmh), _lib.ptr(kmv),
              *[_lib.ptr(a) for a in (s11, s22, s33, s12, s13, s23, su, sv, sw)], _lib.stream())
    outs = []
    for sc in (theta, qv) + tuple(extra_scalars):
        out = empty(g.nx, g.ny, g.nz, like=rho)
        _lib.call("lex_smag2_scalar", g, _lib.ptr(column(rho)), _lib.ptr(khh), _lib.ptr(khv), _lib.ptr(sc),
                  _lib.ptr(out), _lib.stream())
        outs.append(out)
    return (su, sv, sw) + tuple(outs)
