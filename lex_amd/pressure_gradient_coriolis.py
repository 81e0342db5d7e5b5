"""Pressure gradient and Coriolis force with the reference's signatures (pressure_gradient_coriolis.py)."""
from __future__ import annotations

from . import _lib
from . import namelist_n_constants as nl
from ._core import empty, grid_from


def pressure_gradient_force(pi0, pip, theta, x3d, y3d, z3d):
    """-Cp theta_face grad(pi') on u, v, w interiors; w zero at surface/lid (pressure_gradient_coriolis.py:7-37)."""
    g = grid_from(pip, x3d, y3d, z3d)
    pu = empty(g.nx + 1, g.ny, g.nz, like=pip)
    pv = empty(g.nx, g.ny + 1, g.nz, like=pip)
    pw = empty(g.nx, g.ny, g.nz + 1, like=pip)
    _lib.call("lex_pgf", g, _lib.ptr(pip), _lib.ptr(theta), _lib.ptr(pu), _lib.ptr(pv), _lib.ptr(pw), _lib.stream())
    return pu, pv, pw


def calculate_coriolis_force(u, v):
    """f-plane Coriolis: fu at v points, fv at u points (pressure_gradient_coriolis.py:80-95)."""
    nx, ny, nz = u.shape[0] - 7, u.shape[1] - 6, u.shape[2] - 2
    g = _lib.grid(nx, ny, nz, 1.0, 1.0, 1.0)
    fu = empty(nx, ny + 1, nz, like=u)
    fv = empty(nx + 1, ny, nz, like=u)
    _lib.call("lex_coriolis", g, _lib.ptr(u), _lib.ptr(v), float(nl.fCor), _lib.ptr(fu), _lib.ptr(fv), _lib.stream())
    return fu, fv


# the geostrophic wind the reference balances with its large-scale pressure gradient
# (pressure_gradient_coriolis.py:102-111 writes the constants inline)
U_GEO, V_GEO = 1.0, 0.0


def calculate_coriolis_force_with_lspgrad(u, v):
    """f-plane Coriolis minus the large-scale pressure gradient that balances the
    geostrophic wind (1, 0) m/s: fu = f (u_bar - 1) at v points, fv = f (v_bar - 0) at u
    points (pressure_gradient_coriolis.py:98-113)."""
    nx, ny, nz = u.shape[0] - 7, u.shape[1] - 6, u.shape[2] - 2
    g = _lib.grid(nx, ny, nz, 1.0, 1.0, 1.0)
    fu = empty(nx, ny + 1, nz, like=u)
    fv = empty(nx + 1, ny, nz, like=u)
    _lib.call("lex_coriolis_lspgrad", g, _lib.ptr(u), _lib.ptr(v), float(nl.fCor), U_GEO, V_GEO, _lib.ptr(fu),
              _lib.ptr(fv), _lib.stream())
    return fu, fv
