"""CPU oracle: a NumPy restatement of LEX's per-step dynamics (test infrastructure only).

This module is the checker for the HIP path in ``lex_amd``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it; the product path never calls it.

Every function restates one reference routine and cites it as ``file:line`` of the
reference repository (MetLab-HKUST/LEX).  Arrays use the reference's layout: C-order
``(x, y, z)`` with z contiguous, ghost widths ``(ngx, ngy, ngz) = (3, 3, 1)``:

* scalars (theta, pi', qv, ...): ``(nx+6, ny+6, nz+2)``
* u: ``(nx+7, ny+6, nz+2)``  (nx+1 interior faces incl. the periodic duplicate)
* v: ``(nx+6, ny+7, nz+2)``
* w: ``(nx+6, ny+6, nz+3)``  (nz+1 interior levels incl. surface and lid)

The dtype of the inputs selects the arithmetic: float64 arrays give the FP64 oracle,
float32 arrays the FP32 oracle (Python-float constants are weak scalars under NumPy 2,
as in JAX, so FP32 stays FP32).

Pinning (see tests/test_oracle_*.py): the reference ships no tests or golden vectors
and JAX is absent, so each function is pinned by closed forms, invariants, brute-force
per-face loops and library routines (scipy) chosen so that a dropped term, a wrong
sign/index or a transposed operand fails.  Functions whose only pin would be an import
of the reference are marked "parity unpinned" in their docstring.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------------
# Physical constants: namelist_n_constants.py:64-75
# ---------------------------------------------------------------------------------
Rd = 287.04
Rv = 461.5
Cp = 1004.64
Cv = Cp - Rd
eps = Rd / Rv
reps = Rv / Rd
repsm1 = Rv / Rd - 1.0
g = 9.80616
p00 = 100000.0
fCor = 1.0e-4

NGX, NGY, NGZ = 3, 3, 1   # namelist_n_constants.py:23-25

# WENO small numbers: one_step_integration.py:896 (momentum), :905 (theta), :913 (qv)
WEPS_MOM = 1.0e-16
WEPS_THETA = 1.0e-17
WEPS_QV = 1.0e-18


@dataclass(frozen=True)
class Grid:
    """Uniform grid (the reference's grids are uniform: setup_lex.py:12-24)."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float

    @property
    def shape_c(self):
        return (self.nx + 6, self.ny + 6, self.nz + 2)

    @property
    def shape_u(self):
        return (self.nx + 7, self.ny + 6, self.nz + 2)

    @property
    def shape_v(self):
        return (self.nx + 6, self.ny + 7, self.nz + 2)

    @property
    def shape_w(self):
        return (self.nx + 6, self.ny + 6, self.nz + 3)


@dataclass
class Options:
    """Mirror of the namelist switches the hot path reads (namelist_n_constants.py:6-50)."""
    dt: float = 1.0
    solver_opt: int = 2
    n_sound: int = 32
    turb_opt: int = 0
    cor_opt: int = 0
    cs: float = 0.18            # turbulence.py:433 (hard-coded there; a runtime scalar here)
    prandtl: float = 1.0 / 3.0  # turbulence.py:432
    bicg_tol: float = 1.0e-4    # one_step_integration.py:475
    bicg_atol: float = 1.0e-8   # one_step_integration.py:476
    bicg_maxiter: int = 100     # one_step_integration.py:485
    damp_opt: bool = False      # upper-level Rayleigh damping (boundary_conditions.py:119-141)
    rad_opt: bool = False       # Newtonian heating (radiation.py:7-16)
    z_damping: float = 9600.0   # namelist_n_constants.py:54
    rd_alpha: float = 1.0 / 120.0  # namelist_n_constants.py:55
    sfc_opt: bool = False       # bulk surface fluxes (boundary_conditions.py:7-75); base[5] = surface T


# ---------------------------------------------------------------------------------
# small slicing helpers
# ---------------------------------------------------------------------------------
def _sl(a, axis, start, stop):
    idx = [slice(None)] * a.ndim
    idx[axis] = slice(start, stop)
    return a[tuple(idx)]


def _cat(parts, axis):
    return np.concatenate(parts, axis=axis)


def _interior(a):
    """[ngx:-ngx, ngy:-ngy, ngz:-ngz] of a scalar-shaped padded array."""
    return a[NGX:-NGX, NGY:-NGY, NGZ:-NGZ]


# ---------------------------------------------------------------------------------
# H1: halo padding — one_step_integration.py:950-985, pressure_equations.py:111-120
# ---------------------------------------------------------------------------------
def _zcopy(a):
    return _cat([a[:, :, :1], a, a[:, :, -1:]], 2)


def padding3_array(arr):
    """Periodic 3-cell x/y halo, z ghost = nearest level (one_step_integration.py:950-960)."""
    a = _cat([arr[-3:], arr, arr[:3]], 0)
    a = _cat([a[:, -3:], a, a[:, :3]], 1)
    return _zcopy(a)


def padding3u_array(arr):
    """u-point halo: x ghosts skip the duplicate face (one_step_integration.py:963-973)."""
    a = _cat([arr[-4:-1], arr, arr[1:4]], 0)
    a = _cat([a[:, -3:], a, a[:, :3]], 1)
    return _zcopy(a)


def padding3v_array(arr):
    """v-point halo: y ghosts skip the duplicate face (one_step_integration.py:976-985)."""
    a = _cat([arr[-3:], arr, arr[:3]], 0)
    a = _cat([a[:, -4:-1], a, a[:, 1:4]], 1)
    return _zcopy(a)


def padding_array(arr):
    """1-cell periodic x/y halo, z copy (pressure_equations.py:111-120)."""
    a = _cat([arr[-1:], arr, arr[:1]], 0)
    a = _cat([a[:, -1:], a, a[:, :1]], 1)
    return _zcopy(a)


def set_w_bc(w):
    """Zero w on padded levels 0,1 and nz+1,nz+2 (boundary_conditions.py:104-109)."""
    w = w.copy()
    w[:, :, 0:2] = 0.0
    w[:, :, -2:] = 0.0
    return w


# ---------------------------------------------------------------------------------
# WENO reconstructions (advection.py:305-416 and siblings)
# Jiang & Shu (1996) candidates and smoothness indicators, Borges et al. (2008)
# WENO-Z weights with squared ratio and a 1e30 cap.
# ---------------------------------------------------------------------------------
def weno5_parts(a, b, c, d, e, weps):
    """Smoothness indicators, candidate values and weights of one WENO5 face.

    advection.py:311-326: Jiang-Shu beta_k, the three quadratic candidates, and
    w_k = d_k (1 + min(1e30, |beta1-beta3|/(beta_k+eps))^2) with d = (0.1, 0.6, 0.3).
    """
    b1 = (13.0 / 12.0) * (a - 2.0 * b + c) ** 2 + 0.25 * (a - 4.0 * b + 3.0 * c) ** 2
    b2 = (13.0 / 12.0) * (b - 2.0 * c + d) ** 2 + 0.25 * (b - d) ** 2
    b3 = (13.0 / 12.0) * (c - 2.0 * d + e) ** 2 + 0.25 * (3.0 * c - 4.0 * d + e) ** 2
    tau = np.abs(b1 - b3)
    w1 = 0.1 * (1.0 + np.fmin(1.0e30, tau / (b1 + weps)) ** 2)
    w2 = 0.6 * (1.0 + np.fmin(1.0e30, tau / (b2 + weps)) ** 2)
    w3 = 0.3 * (1.0 + np.fmin(1.0e30, tau / (b3 + weps)) ** 2)
    q1 = (2.0 / 6.0) * a + (-7.0 / 6.0) * b + (11.0 / 6.0) * c
    q2 = (-1.0 / 6.0) * b + (5.0 / 6.0) * c + (2.0 / 6.0) * d
    q3 = (2.0 / 6.0) * c + (5.0 / 6.0) * d + (-1.0 / 6.0) * e
    return (b1, b2, b3), (q1, q2, q3), (w1, w2, w3)


def weno5(a, b, c, d, e, weps):
    """Face value from the upwind-ordered stencil (a, b, c | d, e), c the upwind cell.

    advection.py:311-326 (upward form; the other orientations pass a mirrored stencil).
    """
    _, (q1, q2, q3), (w1, w2, w3) = weno5_parts(a, b, c, d, e, weps)
    return (w1 * q1 + w2 * q2 + w3 * q3) / (w1 + w2 + w3)


def weno3(a, b, c, weps):
    """Face value from (a, b | c), b the upwind cell (advection.py:329-340)."""
    b1 = (a - b) ** 2
    b2 = (b - c) ** 2
    tau = np.abs(b1 - b2)
    w1 = (1.0 / 3.0) * (1.0 + np.fmin(1.0e30, tau / (b1 + weps)) ** 2)
    w2 = (2.0 / 3.0) * (1.0 + np.fmin(1.0e30, tau / (b2 + weps)) ** 2)
    return (w1 * ((-1.0 / 2.0) * a + (3.0 / 2.0) * b) + w2 * ((1.0 / 2.0) * b + (1.0 / 2.0) * c)) / (w1 + w2)


def weno5_faces(s, axis, weps):
    """Both upwind reconstructions at the L-5 faces of a padded axis of length L.

    Face g lies between s[g+2] and s[g+3] (advection.py:425-470 for x, :477-513 for y).
    Returns (positive-flow value, negative-flow value).
    """
    L = s.shape[axis]
    t = [_sl(s, axis, k, L - 5 + k) for k in range(6)]
    up = weno5(t[0], t[1], t[2], t[3], t[4], weps)
    dn = weno5(t[5], t[4], t[3], t[2], t[1], weps)
    return up, dn


def vertical_faces(s, weps):
    """Vertical reconstruction table on a padded z-column of length L (L-1 faces).

    Face f lies between s[..., f] and s[..., f+1] (advection.py:305-416 for scalars,
    :561-675 u, :816-930 v, :1070-1183 w).  Faces 0 and L-2 are zero (filled later by
    boundary conditions).  Upward flow: f=1 centred, f=2 WENO3, f=3..L-4 WENO5,
    f=L-3 WENO3.  Downward: f=1 WENO3, f=2..L-5 WENO5, f=L-4 WENO3, f=L-3 centred.
    """
    L = s.shape[2]
    p = [s[:, :, k] for k in range(L)]
    zero = np.zeros_like(p[0])
    up = [zero] * (L - 1)
    dn = [zero] * (L - 1)
    up[1] = 0.5 * (p[1] + p[2])
    up[2] = weno3(p[1], p[2], p[3], weps)
    up[L - 3] = weno3(p[L - 4], p[L - 3], p[L - 2], weps)
    dn[1] = weno3(p[3], p[2], p[1], weps)
    dn[L - 4] = weno3(p[L - 2], p[L - 3], p[L - 4], weps)
    dn[L - 3] = 0.5 * (p[L - 2] + p[L - 3])
    # WENO5 interior faces computed as whole slabs, then spliced in
    f_up = list(range(3, L - 3))
    f_dn = list(range(2, L - 4))
    if f_up:
        a = np.stack([p[f - 2] for f in f_up], 2)
        b = np.stack([p[f - 1] for f in f_up], 2)
        c = np.stack([p[f] for f in f_up], 2)
        d = np.stack([p[f + 1] for f in f_up], 2)
        e = np.stack([p[f + 2] for f in f_up], 2)
        v = weno5(a, b, c, d, e, weps)
        for n, f in enumerate(f_up):
            up[f] = v[:, :, n]
    if f_dn:
        a = np.stack([p[f + 3] for f in f_dn], 2)
        b = np.stack([p[f + 2] for f in f_dn], 2)
        c = np.stack([p[f + 1] for f in f_dn], 2)
        d = np.stack([p[f] for f in f_dn], 2)
        e = np.stack([p[f - 1] for f in f_dn], 2)
        v = weno5(a, b, c, d, e, weps)
        for n, f in enumerate(f_dn):
            dn[f] = v[:, :, n]
    return np.stack(up, 2), np.stack(dn, 2)


def _upwind(mass, up, dn):
    """lax.select(mass >= 0, mass*up, mass*dn) (advection.py:412, :467, :513)."""
    return np.where(mass >= 0.0, mass * up, mass * dn)


# ---------------------------------------------------------------------------------
# A1: divergences — advection.py:200-258
# ---------------------------------------------------------------------------------
def _rho_faces_x(rho):
    part = 0.5 * (rho[:-1] + rho[1:])
    return _cat([part[-1:], part, part[:1]], 0)


def _rho_faces_y(rho):
    part = 0.5 * (rho[:, :-1] + rho[:, 1:])
    return _cat([part[:, -1:], part, part[:, :1]], 1)


def _rho_faces_z(rho):
    part = 0.5 * (rho[:, :, :-1] + rho[:, :, 1:])
    z = np.zeros_like(part[:, :, :1])
    return _cat([z, part, z], 2)


def get_divergence(rho0, u, v, w, grid):
    """div(rho0 u) on the whole padded scalar array (advection.py:200-226)."""
    ru = _rho_faces_x(rho0) * u
    rv = _rho_faces_y(rho0) * v
    rw = _rho_faces_z(rho0) * w
    return ((ru[1:] - ru[:-1]) / grid.dx + (rv[:, 1:] - rv[:, :-1]) / grid.dy
            + (rw[:, :, 1:] - rw[:, :, :-1]) / grid.dz)


def get_divergence2(u, v, w, grid):
    """div(u) on the padded scalar array (advection.py:229-236)."""
    return ((u[1:] - u[:-1]) / grid.dx + (v[:, 1:] - v[:, :-1]) / grid.dy
            + (w[:, :, 1:] - w[:, :, :-1]) / grid.dz)


def get_2d_divergence(rho, u, v, grid):
    """Horizontal div(rho u, rho v) (advection.py:239-258)."""
    ru = _rho_faces_x(rho) * u
    rv = _rho_faces_y(rho) * v
    return (ru[1:] - ru[:-1]) / grid.dx + (rv[:, 1:] - rv[:, :-1]) / grid.dy


# ---------------------------------------------------------------------------------
# A2: scalar advection — advection.py:9-34, :305-416, :419-516
# ---------------------------------------------------------------------------------
def vertical_flux_scalar2(weps, rho0, w, s):
    """(advection.py:305-416) -> (nx, ny, nz+1)."""
    up, dn = vertical_faces(s, weps)
    mass = 0.5 * (rho0[:, :, :-1] + rho0[:, :, 1:]) * w[:, :, 1:-1]
    return _upwind(mass, up, dn)[NGX:-NGX, NGY:-NGY, :]


def horizontal_flux_scalar(weps, rho0, u, v, s):
    """(advection.py:419-516) -> (nx+1, ny, nz), (nx, ny+1, nz)."""
    up, dn = weno5_faces(s, 0, weps)
    mass = 0.5 * (rho0[2:-3] + rho0[3:-2]) * u[3:-3]
    fx = _upwind(mass, up, dn)[:, NGY:-NGY, NGZ:-NGZ]
    up, dn = weno5_faces(s, 1, weps)
    mass = 0.5 * (rho0[:, 2:-3] + rho0[:, 3:-2]) * v[:, 3:-3]
    fy = _upwind(mass, up, dn)[NGX:-NGX, :, NGZ:-NGZ]
    return fx, fy


def _flux_div(fx, fy, fz, grid):
    return -((fx[1:] - fx[:-1]) / grid.dx + (fy[:, 1:] - fy[:, :-1]) / grid.dy
             + (fz[:, :, 1:] - fz[:, :, :-1]) / grid.dz)


def advection_scalar(rho0, s, u, v, w, weps, flow_divergence, sfc_flux, grid):
    """[-div F + s div(rho0 u)] / rho0 at interior cells (advection.py:9-34)."""
    fz = vertical_flux_scalar2(weps, rho0, w, s)
    fz[:, :, 0] = sfc_flux
    fx, fy = horizontal_flux_scalar(weps, rho0, u, v, s)
    conv = _flux_div(fx, fy, fz, grid)
    return (conv + _interior(s) * _interior(flow_divergence)) / _interior(rho0)


# ---------------------------------------------------------------------------------
# A3: momentum advection — advection.py:37-149, :561-675, :678-771, :816-1025,
# :1070-1281
# ---------------------------------------------------------------------------------
def vertical_flux_u2(weps, rho0, w, u):
    """(advection.py:561-675) -> (nx+1, ny, nz+1)."""
    up, dn = vertical_faces(u, weps)
    m = 0.5 * (rho0[:, :, :-1] + rho0[:, :, 1:]) * w[:, :, 1:-1]
    m8u = 0.5 * (m[:-1] + m[1:])
    return _upwind(m8u, up[1:-1], dn[1:-1])[NGX - 1:-(NGX - 1), NGY:-NGY, :]


def horizontal_flux_u(weps, rho0, u, v):
    """(advection.py:678-771) -> (nx+2, ny, nz) x-flux at cells, (nx+1, ny+1, nz) y-flux at corners."""
    up, dn = weno5_faces(u, 0, weps)
    mass = 0.5 * (u[2:-3] + u[3:-2]) * rho0[2:-2]
    fx = _upwind(mass, up, dn)[:, NGY:-NGY, NGZ:-NGZ]
    up, dn = weno5_faces(u, 1, weps)
    mv = 0.5 * (rho0[:, 2:-3] + rho0[:, 3:-2]) * v[:, 3:-3]
    mv8u = 0.5 * (mv[:-1] + mv[1:])
    fy = _upwind(mv8u, up[1:-1], dn[1:-1])[NGX - 1:-(NGX - 1), :, NGZ:-NGZ]
    return fx, fy


def advection_u(rho0, u, v, w, weps, flow_divergence, u_sfc_flux, grid):
    """u tendency at the nx+1 interior u faces (advection.py:37-74)."""
    fz = vertical_flux_u2(weps, rho0, w, u)
    tx = _cat([u_sfc_flux[-1:], u_sfc_flux, u_sfc_flux[:1]], 0)
    fz[:, :, 0] = 0.5 * (tx[:-1] + tx[1:])
    fx, fy = horizontal_flux_u(weps, rho0, u, v)
    conv = _flux_div(fx, fy, fz, grid)
    fd = flow_divergence[:, NGY:-NGY, NGZ:-NGZ]
    div8u = 0.5 * (fd[NGX - 1:-NGX] + fd[NGX:-(NGX - 1)])
    r = rho0[:, NGY:-NGY, NGZ:-NGZ]
    rho8u = 0.5 * (r[NGX - 1:-NGX] + r[NGX:-(NGX - 1)])
    return (conv + u[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] * div8u) / rho8u


def vertical_flux_v2(weps, rho0, w, v):
    """(advection.py:816-930) -> (nx, ny+1, nz+1)."""
    up, dn = vertical_faces(v, weps)
    m = 0.5 * (rho0[:, :, :-1] + rho0[:, :, 1:]) * w[:, :, 1:-1]
    m8v = 0.5 * (m[:, :-1] + m[:, 1:])
    return _upwind(m8v, up[:, 1:-1], dn[:, 1:-1])[NGX:-NGX, NGY - 1:-(NGY - 1), :]


def horizontal_flux_v(weps, rho0, u, v):
    """(advection.py:933-1025) -> (nx+1, ny+1, nz) x-flux at corners, (nx, ny+2, nz) y-flux at cells."""
    up, dn = weno5_faces(v, 0, weps)
    mu = 0.5 * (rho0[2:-3] + rho0[3:-2]) * u[3:-3]
    mu8v = 0.5 * (mu[:, :-1] + mu[:, 1:])
    fx = _upwind(mu8v, up[:, 1:-1], dn[:, 1:-1])[:, NGY - 1:-(NGY - 1), NGZ:-NGZ]
    up, dn = weno5_faces(v, 1, weps)
    mass = 0.5 * (v[:, 2:-3] + v[:, 3:-2]) * rho0[:, 2:-2]
    fy = _upwind(mass, up, dn)[NGX:-NGX, :, NGZ:-NGZ]
    return fx, fy


def advection_v(rho0, u, v, w, weps, flow_divergence, v_sfc_flux, grid):
    """v tendency at the ny+1 interior v faces (advection.py:77-114)."""
    fz = vertical_flux_v2(weps, rho0, w, v)
    ty = _cat([v_sfc_flux[:, -1:], v_sfc_flux, v_sfc_flux[:, :1]], 1)
    fz[:, :, 0] = 0.5 * (ty[:, :-1] + ty[:, 1:])
    fx, fy = horizontal_flux_v(weps, rho0, u, v)
    conv = _flux_div(fx, fy, fz, grid)
    fd = flow_divergence[NGX:-NGX, :, NGZ:-NGZ]
    div8v = 0.5 * (fd[:, NGY - 1:-NGY] + fd[:, NGY:-(NGY - 1)])
    r = rho0[NGX:-NGX, :, NGZ:-NGZ]
    rho8v = 0.5 * (r[:, NGY - 1:-NGY] + r[:, NGY:-(NGY - 1)])
    return (conv + v[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] * div8v) / rho8v


def vertical_flux_w2(weps, rho0, w):
    """(advection.py:1070-1183) -> (nx, ny, nz+2) at scalar levels incl. ghosts."""
    up, dn = vertical_faces(w, weps)
    mass = 0.5 * (w[:, :, :-1] + w[:, :, 1:]) * rho0
    return _upwind(mass, up, dn)[NGX:-NGX, NGY:-NGY, :]


def horizontal_flux_w(weps, rho0, u, v, w):
    """(advection.py:1186-1281) -> (nx+1, ny, nz+1), (nx, ny+1, nz+1)."""
    up, dn = weno5_faces(w, 0, weps)
    mu = 0.5 * (rho0[2:-3] + rho0[3:-2]) * u[3:-3]
    mu8w = 0.5 * (mu[:, :, :-1] + mu[:, :, 1:])
    fx = _upwind(mu8w, up[:, :, 1:-1], dn[:, :, 1:-1])[:, NGY:-NGY, :]
    up, dn = weno5_faces(w, 1, weps)
    mv = 0.5 * (rho0[:, 2:-3] + rho0[:, 3:-2]) * v[:, 3:-3]
    mv8w = 0.5 * (mv[:, :, :-1] + mv[:, :, 1:])
    fy = _upwind(mv8w, up[:, :, 1:-1], dn[:, :, 1:-1])[NGX:-NGX, :, :]
    return fx, fy


def advection_w(rho0, u, v, w, weps, flow_divergence, grid):
    """w tendency at the nz+1 interior w levels, zero at surface and lid (advection.py:117-149)."""
    fz = vertical_flux_w2(weps, rho0, w)
    fx, fy = horizontal_flux_w(weps, rho0, u, v, w)
    conv = _flux_div(fx, fy, fz, grid)
    fd = flow_divergence[NGX:-NGX, NGY:-NGY, :]
    div8w = 0.5 * (fd[:, :, :-1] + fd[:, :, 1:])
    r = rho0[NGX:-NGX, NGY:-NGY, :]
    rho8w = 0.5 * (r[:, :, :-1] + r[:, :, 1:])
    t = (conv + w[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] * div8w) / rho8w
    t[:, :, 0] = 0.0
    t[:, :, -1] = 0.0
    return t


# ---------------------------------------------------------------------------------
# A4: low-order pi' advection — advection.py:152-197
# ---------------------------------------------------------------------------------
def advection_pi(rho0, pi0, pip, u, v, w, grid, cc1=0.5, cc2=0.5):
    """-(u.grad pi) with face products averaged to cells (advection.py:152-197)."""
    I = slice(NGX, -NGX)
    J = slice(NGY, -NGY)
    K = slice(NGZ, -NGZ)
    dpi_dx = (pip[NGX:-NGX + 1, J, K] - pip[NGX - 1:-NGX, J, K]) / grid.dx
    dpi_dy = (pip[I, NGY:-NGY + 1, K] - pip[I, NGY - 1:-NGY, K]) / grid.dy
    dpi0_dz = (pi0[I, J, NGZ:] - pi0[I, J, :-NGZ]) / grid.dz
    dpip_dz = (pip[I, J, NGZ:] - pip[I, J, :-NGZ]) / grid.dz
    dpi_dz = dpi0_dz + dpip_dz
    udx = u[I, J, K] * dpi_dx
    vdy = v[I, J, K] * dpi_dy
    wdz = w[I, J, K] * dpi_dz
    wdz[:, :, 0] = 0.0
    wdz[:, :, -1] = 0.0
    r = rho0[I, J, :]
    rho8w = 0.5 * (r[:, :, :-1] + r[:, :, 1:])
    return -(0.5 * (udx[1:] + udx[:-1]) + 0.5 * (vdy[:, 1:] + vdy[:, :-1])
             + (cc1 * wdz[:, :, :-1] * rho8w[:, :, :-1] + cc2 * wdz[:, :, 1:] * rho8w[:, :, 1:])
             / rho0[I, J, K])


# ---------------------------------------------------------------------------------
# B1 / P3: buoyancy, rtt, pressure equation — pressure_equations.py
# ---------------------------------------------------------------------------------
def extrapolate_bottom_top(s):
    """2nd-order Lagrange extrapolation half a level beyond the end levels (pressure_equations.py:141-159)."""
    bottom = 1.875 * s[:, :, 0] + (-1.25) * s[:, :, 1] + 0.375 * s[:, :, 2]
    top = 1.875 * s[:, :, -1] + (-1.25) * s[:, :, -2] + 0.375 * s[:, :, -3]
    return bottom[:, :, None], top[:, :, None]


def interpolate_scalar2w(s):
    """(pressure_equations.py:123-128) (nx,ny,nz) -> (nx,ny,nz+1)."""
    b, t = extrapolate_bottom_top(s)
    return _cat([b, 0.5 * (s[:, :, :-1] + s[:, :, 1:]), t], 2)


def interpolate_scalar2w_0(s):
    """(pressure_equations.py:131-138) zero end levels."""
    z = np.zeros_like(s[:, :, :1])
    return _cat([z, 0.5 * (s[:, :, :-1] + s[:, :, 1:]), z], 2)


def theta_rho_of(theta, qv):
    """Density potential temperature theta(1+Rv/Rd qv)/(1+qv) (pressure_equations.py:167)."""
    return theta * (1.0 + reps * qv) / (1.0 + qv)


def calculate_buoyancy(theta0, theta, qv0, qv):
    """B and B at w levels (pressure_equations.py:162-172)."""
    tr0 = theta0 * (1.0 + reps * qv0) / (1.0 + qv0)
    tr = theta * (1.0 + reps * qv) / (1.0 + qv)
    b = g * (tr - tr0) / tr0
    bi = _interior(b)
    return bi, interpolate_scalar2w(bi)


def calculate_rtt(rho0_theta0, theta, qv, buoyancy):
    """rho0 theta0 theta_rho with extrapolated rho0theta0 B/Cp in the z ghosts (pressure_equations.py:175-195)."""
    part = rho0_theta0[:, :, 1:-1] * theta[:, :, 1:-1] * (1.0 + reps * qv[:, :, 1:-1]) / (1.0 + qv[:, :, 1:-1])
    b, t = extrapolate_bottom_top(_interior(rho0_theta0) * buoyancy / Cp)

    def pad2(p):
        p = p[:, :, 0]
        p = _cat([p[-3:], p, p[:3]], 0)
        p = _cat([p[:, -3:], p, p[:, :3]], 1)
        return p[:, :, None]
    return _cat([pad2(b), part, pad2(t)], 2)


def laplace_of_pressure(rtt, pi, grid):
    """Affine matvec Cp div(rtt_face grad pi) with rtt-ghost boundary z-fluxes (pressure_equations.py:8-48)."""
    pp = padding_array(pi)
    dpx = (pp[1:, 1:-1, 1:-1] - pp[:-1, 1:-1, 1:-1]) / grid.dx
    dpy = (pp[1:-1, 1:, 1:-1] - pp[1:-1, :-1, 1:-1]) / grid.dy
    dpz = (pp[1:-1, 1:-1, 1:] - pp[1:-1, 1:-1, :-1]) / grid.dz
    I = slice(NGX, -NGX)
    J = slice(NGY, -NGY)
    K = slice(NGZ, -NGZ)
    px = 0.5 * (rtt[NGX:-(NGX - 1), J, K] + rtt[NGX - 1:-NGX, J, K]) * dpx
    py = 0.5 * (rtt[I, NGY:-(NGY - 1), K] + rtt[I, NGY - 1:-NGY, K]) * dpy
    pz = 0.5 * (rtt[I, J, NGZ:] + rtt[I, J, :-NGZ]) * dpz
    pz[:, :, 0] = rtt[I, J, 0]
    pz[:, :, -1] = rtt[I, J, -1]
    return Cp * ((px[1:] - px[:-1]) / grid.dx + (py[:, 1:] - py[:, :-1]) / grid.dy
                 + (pz[:, :, 1:] - pz[:, :, :-1]) / grid.dz)


def rhs_of_pressure_equation(rho0_theta0, adv4u, adv4v, adv4w, fu, fv, buoyancy, grid):
    """RHS = div(rt adv) + div_h(rt (fv,-fu)) + d/dz(rt B) (pressure_equations.py:51-77)."""
    rt = _interior(rho0_theta0)
    # get_divergence on un-padded arrays (pressure_equations.py:54-57)
    ru = _rho_faces_x(rt) * adv4u
    rv = _rho_faces_y(rt) * adv4v
    rw = _rho_faces_z(rt) * adv4w
    rhs_adv = ((ru[1:] - ru[:-1]) / grid.dx + (rv[:, 1:] - rv[:, :-1]) / grid.dy
               + (rw[:, :, 1:] - rw[:, :, :-1]) / grid.dz)
    rhs_cor = get_2d_divergence(rt, fv, -fu, grid)
    rtb = interpolate_scalar2w(rt * buoyancy)
    rhs_buoy = (rtb[:, :, 1:] - rtb[:, :, :-1]) / grid.dz
    return rhs_adv + rhs_cor + rhs_buoy


# ---------------------------------------------------------------------------------
# P4: BiCGSTAB with jax.scipy.sparse.linalg.bicgstab semantics (third-party; the
# reference calls it at one_step_integration.py:482-485).  Restated from JAX's public
# _bicgstab_solve: unpreconditioned, atol2 = max(tol^2 |b|^2, atol^2), breakdown codes.
# The exact iterate trajectory in FP32 is "parity unpinned" (reduction order).
# ---------------------------------------------------------------------------------
def _dot(a, b):
    return float(np.dot(a.ravel().astype(np.float64), b.ravel().astype(np.float64)))


def bicgstab(A, b, x0, tol, atol, maxiter):
    """Returns (x, k) with k the iteration count or -10/-11 on breakdown."""
    dt = b.dtype
    bs = _dot(b, b)
    atol2 = max(tol * tol * bs, atol * atol)
    x = x0.copy()
    r = b - A(x)
    rhat = r.copy()
    p = r.copy()
    q = r.copy()
    rho = alpha = omega = np.float64(1.0)
    k = 0
    while _dot(r, r) > atol2 and k < maxiter and k >= 0:
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):  # IEEE inf/nan, as in JAX
            rho_ = np.float64(_dot(rhat, r))
            beta = rho_ / rho * alpha / omega
            p = (r + dt.type(beta) * (p - dt.type(omega) * q)).astype(dt)
            q = A(p)
            alpha_ = rho_ / np.float64(_dot(rhat, q))
            s = (r - dt.type(alpha_) * q).astype(dt)
            early = _dot(s, s) < atol2
            t = A(s)
            omega_ = np.float64(_dot(t, s)) / np.float64(_dot(t, t))
        if early:
            x = (x + dt.type(alpha_) * p).astype(dt)
            r = s
        else:
            x = (x + (dt.type(alpha_) * p + dt.type(omega_) * s)).astype(dt)
            r = (s - dt.type(omega_) * t).astype(dt)
        knew = -11 if (omega_ == 0.0 or alpha_ == 0.0) else k + 1
        if rho_ == 0.0:
            knew = -10
        k = knew
        rho, alpha, omega = rho_, alpha_, omega_
    return x, k


def solve_pres_eqn(pip_prev, rho0_theta0, rtt, adv4u, adv4v, adv4w, fu, fv, buoyancy, grid, opts):
    """(one_step_integration.py:468-487) -> interior pi' and the BiCGSTAB exit code."""
    rhs = rhs_of_pressure_equation(rho0_theta0, adv4u, adv4v, adv4w, fu, fv, buoyancy, grid)
    A = lambda beta: laplace_of_pressure(rtt, beta, grid)  # noqa: E731
    return bicgstab(A, rhs, _interior(pip_prev), opts.bicg_tol, opts.bicg_atol, opts.bicg_maxiter)


# ---------------------------------------------------------------------------------
# P1 / P2: pressure gradient and Coriolis — pressure_gradient_coriolis.py:7-95
# ---------------------------------------------------------------------------------
def pressure_gradient_force(pip, theta_rho, grid):
    """-Cp theta_rho_face dpi'/dn on u, v, w interiors; w zero at surface/lid (pressure_gradient_coriolis.py:7-37)."""
    I = slice(NGX, -NGX)
    J = slice(NGY, -NGY)
    K = slice(NGZ, -NGZ)
    th = theta_rho
    dx = (pip[NGX:-(NGX - 1), J, K] - pip[NGX - 1:-NGX, J, K]) / grid.dx
    pu = -Cp * 0.5 * (th[NGX:-(NGX - 1), J, K] + th[NGX - 1:-NGX, J, K]) * dx
    dy = (pip[I, NGY:-(NGY - 1), K] - pip[I, NGY - 1:-NGY, K]) / grid.dy
    pv = -Cp * 0.5 * (th[I, NGY:-(NGY - 1), K] + th[I, NGY - 1:-NGY, K]) * dy
    dz = (pip[I, J, NGZ:] - pip[I, J, :-NGZ]) / grid.dz
    pw = -Cp * 0.5 * (th[I, J, NGZ:] + th[I, J, :-NGZ]) * dz
    pw[:, :, 0] = 0.0
    pw[:, :, -1] = 0.0
    return pu, pv, pw


def _corner_means(u, v):
    """0.25 x the sums of the four v around each u point and the four u around each v
    point, in the reference's summation order (pressure_gradient_coriolis.py:82-91)."""
    K = slice(NGZ, -NGZ)
    sv = (v[NGX - 1:-NGX, NGY:-(NGY + 1), K] + v[NGX - 1:-NGX, NGY + 1:-NGY, K]
          + v[NGX:-(NGX - 1), NGY:-(NGY + 1), K] + v[NGX:-(NGX - 1), NGY + 1:-NGY, K])
    su = (u[NGX:-(NGX + 1), NGY - 1:-NGY, K] + u[NGX:-(NGX + 1), NGY:-(NGY - 1), K]
          + u[NGX + 1:-NGX, NGY - 1:-NGY, K] + u[NGX + 1:-NGX, NGY:-(NGY - 1), K])
    return su, sv


def calculate_coriolis_force(u, v, f=fCor):
    """fu at v points, fv at u points (pressure_gradient_coriolis.py:80-95)."""
    su, sv = _corner_means(u, v)
    return f * 0.25 * su, f * 0.25 * sv


def calculate_coriolis_force_with_lspgrad(u, v, f=fCor, u_geo=1.0, v_geo=0.0):
    """The Coriolis force with a large-scale pressure gradient in geostrophic balance with
    the wind (u_geo, v_geo) = (1, 0) m/s: fu = f (u_bar - u_geo), fv = f (v_bar - v_geo)
    (pressure_gradient_coriolis.py:98-113; the reference writes the constants inline)."""
    su, sv = _corner_means(u, v)
    return f * (0.25 * su - u_geo), f * (0.25 * sv - v_geo)


# ---------------------------------------------------------------------------------
# G1: Smagorinsky — turbulence.py:9-111, :298-446, :474-504
# ---------------------------------------------------------------------------------
def compute_deformation(rho, u, v, w, grid):
    """rho S_ij at cells and the deformation (turbulence.py:298-388); all padded."""
    I = slice(NGX, -NGX)
    J = slice(NGY, -NGY)
    K = slice(NGZ, -NGZ)
    r = rho[I, J, K]
    s11 = r * (u[NGX + 1:-NGX, J, K] - u[NGX:-(NGX + 1), J, K]) / grid.dx
    s22 = r * (v[I, NGY + 1:-NGY, K] - v[I, NGY:-(NGY + 1), K]) / grid.dy
    s33 = r * (w[I, J, NGZ + 1:-NGZ] - w[I, J, NGZ:-(NGZ + 1)]) / grid.dz
    dudy = (u[I, NGY:-(NGY - 1), K] - u[I, NGY - 1:-NGY, K]) / grid.dy
    dvdx = (v[NGX:-(NGX - 1), NGY:-NGY, K] - v[NGX - 1:-NGX, NGY:-NGY, K]) / grid.dx
    e12 = 0.5 * (dudy + dvdx)
    s12 = r * (0.25 * (e12[:-1, :-1] + e12[1:, 1:] + e12[:-1, 1:] + e12[1:, :-1]))
    dudz = (u[I, J, NGZ + 1:-NGZ] - u[I, J, NGZ:-(NGZ + 1)]) / grid.dz
    dwdx = (w[NGX:-(NGX - 1), J, NGZ + 1:-(NGZ + 1)] - w[NGX - 1:-NGX, J, NGZ + 1:-(NGZ + 1)]) / grid.dx
    e13 = 0.5 * (dudz + dwdx)
    b0, t0 = extrapolate_bottom_top(e13)
    s13 = _cat([0.5 * (b0[1:] + b0[:-1]),
                0.25 * (e13[:-1, :, :-1] + e13[1:, :, 1:] + e13[:-1, :, 1:] + e13[1:, :, :-1]),
                0.5 * (t0[1:] + t0[:-1])], 2)
    s13 = r * s13
    dvdz = (v[I, J, NGZ + 1:-NGZ] - v[I, J, NGZ:-(NGZ + 1)]) / grid.dz
    dwdy = (w[I, NGY:-(NGY - 1), NGZ + 1:-(NGZ + 1)] - w[I, NGY - 1:-NGY, NGZ + 1:-(NGZ + 1)]) / grid.dy
    e23 = 0.5 * (dvdz + dwdy)
    b0, t0 = extrapolate_bottom_top(e23)
    s23 = _cat([0.5 * (b0[:, 1:] + b0[:, :-1]),
                0.25 * (e23[:, :-1, :-1] + e23[:, 1:, 1:] + e23[:, :-1, 1:] + e23[:, 1:, :-1]),
                0.5 * (t0[:, 1:] + t0[:, :-1])], 2)
    s23 = r * s23
    deform = (2.0 * (s11 ** 2 + s22 ** 2 + s33 ** 2) + 4.0 * (s12 ** 2 + s13 ** 2 + s23 ** 2)) / r ** 2
    return tuple(padding3_array(a) for a in (s11, s22, s33, s12, s13, s23, deform))


def compute_nm(theta, qv, grid):
    """N^2 = g d ln(theta(1+repsm1 qv))/dz at cells, extrapolated ends (turbulence.py:391-427)."""
    tr = theta * (1.0 + repsm1 * qv)
    I = slice(NGX, -NGX)
    J = slice(NGY, -NGY)
    n28w = np.log(tr[I, J, NGZ + 1:-NGZ] / tr[I, J, NGZ:-(NGZ + 1)]) / grid.dz * g
    b, t = extrapolate_bottom_top(n28w)
    return padding3_array(_cat([b, 0.5 * (n28w[:, :, :-1] + n28w[:, :, 1:]), t], 2))


def compute_k(s2, n2, grid, cs=0.18, prandtl=1.0 / 3.0):
    """Km = (Cs Delta)^2 sqrt(max(S2 - N2/Pr, 0)), Kh = Km/Pr (turbulence.py:430-446)."""
    gs = np.cbrt(np.asarray(grid.dx * grid.dy * grid.dz, dtype=s2.dtype))
    s2n2 = s2 - n2 / prandtl
    s2n2 = np.where(s2n2 > 0.0, s2n2, 0.0)
    km = (s2.dtype.type(cs) * gs) ** 2 * np.sqrt(s2n2)
    return km, km / prandtl


def stagger_scalar(rho):
    """(turbulence.py:497-504)."""
    J = slice(NGY, -NGY)
    K = slice(NGZ, -NGZ)
    I = slice(NGX, -NGX)
    r8u = 0.5 * (rho[NGX:-(NGX - 1), J, K] + rho[NGX - 1:-NGX, J, K])
    r8v = 0.5 * (rho[I, NGY:-(NGY - 1), K] + rho[I, NGY - 1:-NGY, K])
    r8w = interpolate_scalar2w(rho[I, J, K])
    return r8u, r8v, r8w


def compute_scalar_gradient(s, grid):
    """(turbulence.py:474-494)."""
    I = slice(NGX, -NGX)
    J = slice(NGY, -NGY)
    K = slice(NGZ, -NGZ)
    dx = (s[NGX:-(NGX - 1), J, K] - s[NGX - 1:-NGX, J, K]) / grid.dx
    dy = (s[I, NGY:-(NGY - 1), K] - s[I, NGY - 1:-NGY, K]) / grid.dy
    dz = (s[I, J, NGZ:] - s[I, J, :-NGZ]) / grid.dz
    dz[:, :, 0] = 0.0
    dz[:, :, -1] = 0.0
    return dx, dy, dz


def _scalar_sgs(rho, kh, s, grid, khv=None):
    """SGS diffusion of one scalar (turbulence.py:80-109); with khv the vertical faces take
    khv (compute_smag2, turbulence.py:207-234)."""
    r8u, r8v, r8w = stagger_scalar(rho)
    k8u, k8v, k8w = stagger_scalar(kh)
    if khv is not None:
        k8w = stagger_scalar(khv)[2]
    dx, dy, dz = compute_scalar_gradient(s, grid)
    tx = -r8u * k8u * dx
    ty = -r8v * k8v * dy
    tz = -r8w * k8w * dz
    tz[:, :, 0] = 0.0
    tz[:, :, -1] = 0.0
    return -((tx[1:] - tx[:-1]) / grid.dx + (ty[:, 1:] - ty[:, :-1]) / grid.dy
             + (tz[:, :, 1:] - tz[:, :, :-1]) / grid.dz) / _interior(rho)


def compute_smag(rho, km, kh, s11, s22, s33, s12, s13, s23, scalars, grid):
    """SGS tendencies for u, v, w and each scalar in `scalars` (turbulence.py:9-111)."""
    return _smag_tendencies(rho, km, km, kh, kh, s11, s22, s33, s12, s13, s23, scalars, grid, True)


def compute_smag2(rho, kmh, kmv, khh, khv, s11, s22, s33, s12, s13, s23, scalars, grid):
    """compute_smag with horizontal (kmh, khh) and vertical (kmv, khv) coefficients
    (turbulence.py:114-295); unlike compute_smag it leaves sgs_w at the surface at its
    stencil value (turbulence.py:208 zeroes the lid only)."""
    return _smag_tendencies(rho, kmh, kmv, khh, khv, s11, s22, s33, s12, s13, s23, scalars, grid, False)


def _smag_tendencies(rho, km, kmv, kh, khv, s11, s22, s33, s12, s13, s23, scalars, grid, zero_surface_w):
    I = slice(NGX, -NGX)
    J = slice(NGY, -NGY)
    K = slice(NGZ, -NGZ)
    t11 = -2.0 * km * s11
    t22 = -2.0 * km * s22
    t33 = -2.0 * km * s33
    t12 = -2.0 * km * s12
    t13 = -2.0 * kmv * s13
    t23 = -2.0 * kmv * s23
    d11 = (t11[NGX:-(NGX - 1), J, K] - t11[NGX - 1:-NGX, J, K]) / grid.dx
    d12y = (t12[NGX - 1:-(NGX - 1), NGY:-(NGY - 1), K] - t12[NGX - 1:-(NGX - 1), NGY - 1:-NGY, K]) / grid.dy
    d12 = 0.25 * (d12y[:-1, :-1] + d12y[1:, 1:] + d12y[:-1, 1:] + d12y[1:, :-1])
    t13w = interpolate_scalar2w_0(t13[:, :, K])
    d13s = (t13w[:, :, 1:] - t13w[:, :, :-1]) / grid.dz
    d13 = 0.5 * (d13s[NGX - 1:-NGX, J] + d13s[NGX:-(NGX - 1), J])
    sgs_u = -(d11 + d12 + d13) / (0.5 * (rho[NGX:-(NGX - 1), J, K] + rho[NGX - 1:-NGX, J, K]))

    d22 = (t22[I, NGY:-(NGY - 1), K] - t22[I, NGY - 1:-NGY, K]) / grid.dy
    d12x = (t12[NGX:-(NGX - 1), NGY - 1:-(NGY - 1), K] - t12[NGX - 1:-NGX, NGY - 1:-(NGY - 1), K]) / grid.dx
    d21 = 0.25 * (d12x[:-1, :-1] + d12x[1:, 1:] + d12x[:-1, 1:] + d12x[1:, :-1])
    t23w = interpolate_scalar2w_0(t23[:, :, K])
    d23s = (t23w[:, :, 1:] - t23w[:, :, :-1]) / grid.dz
    d23 = 0.5 * (d23s[I, NGY - 1:-NGY] + d23s[I, NGY:-(NGY - 1)])
    sgs_v = -(d21 + d22 + d23) / (0.5 * (rho[I, NGY:-(NGY - 1), K] + rho[I, NGY - 1:-NGY, K]))

    d33 = (t33[I, J, NGZ:] - t33[I, J, :-NGZ]) / grid.dz
    d13x = (t13[NGX + 1:-(NGX - 1), J, K] - t13[NGX - 1:-(NGX + 1), J, K]) / (2.0 * grid.dx)
    d23y = (t23[I, NGY + 1:-(NGY - 1), K] - t23[I, NGY - 1:-(NGY + 1), K]) / (2.0 * grid.dy)
    r8w = interpolate_scalar2w(rho[I, J, K])
    sgs_w = -(interpolate_scalar2w_0(d13x) + interpolate_scalar2w_0(d23y) + d33) / r8w
    if zero_surface_w:
        sgs_w[:, :, 0] = 0.0
    sgs_w[:, :, -1] = 0.0
    sgs_s = tuple(_scalar_sgs(rho, kh, s, grid, None if khv is kh else khv) for s in scalars)
    return (sgs_u, sgs_v, sgs_w) + sgs_s


def compute_k2(rho, s11, s22, s33, s12, s13, s23, n2, grid, cs=0.18):
    """(kmh, kmv, khh, khv) of the horizontal / vertical closure (turbulence.py:449-471),
    prandtl 1: kmh = cs^2 dx dy Dh (no square root), kmv = cs^2 Lv^2 sqrt(max(Dv - n2, 0))
    with 1/Lv^2 = (0.4 z)^-2 + dz^-2 at the padded cell-centre heights z."""
    prandtl = 1.0
    ft = s11.dtype.type
    z3d = np.linspace(-0.5 * grid.dz, (grid.nz + 0.5) * grid.dz, grid.nz + 2).astype(s11.dtype)
    gh2 = ft(grid.dx) * ft(grid.dy)
    gv2 = 1.0 / ((0.4 * z3d) ** (-2) + ft(grid.dz) ** (-2))
    deform_h = (2.0 * (s11 ** 2 + s22 ** 2 + s33 ** 2) + 4.0 * s12 ** 2) / rho ** 2
    deform_v = 4.0 * (s13 ** 2 + s23 ** 2) / rho ** 2
    kmh = ft(cs) ** 2 * gh2 * deform_h
    kmv = ft(cs) ** 2 * gv2 * np.sqrt(np.fmax(deform_v - n2 / prandtl, 0.0))
    return kmh, kmv, kmh / prandtl, kmv / prandtl


def smagorinsky(rho0, u, v, w, theta, qv, scalars, grid, cs=0.18, prandtl=1.0 / 3.0, turb_opt=1):
    """compute_deformation -> compute_nm -> compute_k -> compute_smag, or with turb_opt 2
    compute_k2 -> compute_smag2 (one_step_integration.py:58-74)."""
    s11, s22, s33, s12, s13, s23, deform = compute_deformation(rho0, u, v, w, grid)
    n2 = compute_nm(theta, qv, grid)
    if turb_opt == 2:
        kmh, kmv, khh, khv = compute_k2(rho0, s11, s22, s33, s12, s13, s23, n2, grid, cs)
        return compute_smag2(rho0, kmh, kmv, khh, khv, s11, s22, s33, s12, s13, s23, scalars, grid)
    km, kh = compute_k(deform, n2, grid, cs, prandtl)
    return compute_smag(rho0, km, kh, s11, s22, s33, s12, s13, s23, scalars, grid)


# ---------------------------------------------------------------------------------
# S1 / M1 / E1: acoustic substeps and Euler updates — one_step_integration.py
# ---------------------------------------------------------------------------------
def sound_rk_step(u_now, v_now, w_now, pi0, pip_now, pip_star, rhs_u, rhs_v, rhs_w,
                  theta_rho, rho0, n, dt_sound, grid):
    """Forward-backward acoustic substeps (one_step_integration.py:519-556, :638-675, :757-794)."""
    piadv = pi0 + pip_star
    c = -Rd / Cv
    for _ in range(n):
        pu, pv, pw = pressure_gradient_force(pip_now, theta_rho, grid)
        un = u_now[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + (pu + rhs_u) * dt_sound
        vn = v_now[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + (pv + rhs_v) * dt_sound
        wn = w_now[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + (pw + rhs_w) * dt_sound
        un = padding3u_array(un)
        vn = padding3v_array(vn)
        wn = set_w_bc(padding3_array(wn))
        div = get_divergence2(un, vn, wn, grid)
        apt = advection_pi(rho0, pi0, pip_star, un, vn, wn, grid)
        dpdt = c * (_interior(piadv) * _interior(div)) + apt
        pn = padding3_array(_interior(pip_now) + 1.1 * dpdt * dt_sound)
        u_now, v_now, w_now, pip_now = un, vn, wn, pn
    return u_now, v_now, w_now, pip_now


ALPHA_IMPLICIT = 0.6  # off-centring of the vertically implicit acoustic step (namelist_n_constants.py:14)


def tridiagonal_solve(a, b, c, d):
    """Solve a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i along the last axis (a_0, c_{n-1}
    unused): the Thomas algorithm, no pivoting.  Restates jax.lax.linalg.tridiagonal_solve
    as called at one_step_integration.py:614, :733, :852 (third-party; the systems there are
    diagonally dominant)."""
    n = d.shape[-1]
    cp = np.empty_like(d)
    dp = np.empty_like(d)
    cp[..., 0] = c[..., 0] / b[..., 0]
    dp[..., 0] = d[..., 0] / b[..., 0]
    for i in range(1, n):
        m = b[..., i] - a[..., i] * cp[..., i - 1]
        cp[..., i] = c[..., i] / m
        dp[..., i] = (d[..., i] - a[..., i] * dp[..., i - 1]) / m
    x = np.empty_like(d)
    x[..., -1] = dp[..., -1]
    for i in range(n - 2, -1, -1):
        x[..., i] = dp[..., i] - cp[..., i] * x[..., i + 1]
    return x


def implicit_w_coefficients(pi0, pip_star, theta_rho, dt_sound, grid, alpha=ALPHA_IMPLICIT):
    """Tridiagonal coefficients (Ak, Bk, Ck) of the w equation at levels 1..nz-1 and the
    rhs prefactor (one_step_integration.py:570-588), uniform vertical spacing."""
    I, J = slice(NGX, -NGX), slice(NGY, -NGY)
    dz = grid.dz
    pa = (pi0 + pip_star)[I, J, :]
    thr8w = 0.5 * (theta_rho[I, J, 2:-1] + theta_rho[I, J, 1:-2])
    lhs = (alpha * dt_sound) ** 2 * Cp * thr8w / dz
    dpdz = (pa[:, :, 1:] - pa[:, :, :-1]) / dz
    ak = -lhs * (Rd / Cv * pa[:, :, 1:-2] / dz - dpdz[:, :, 0:-2] / 2.0)
    ck = -lhs * (Rd / Cv * pa[:, :, 2:-1] / dz + dpdz[:, :, 2:] / 2.0)
    bk = 1.0 + lhs * Rd / Cv * (pa[:, :, 2:-1] / dz + pa[:, :, 1:-2] / dz)
    ak[:, :, 0] = 0.0
    ck[:, :, -1] = 0.0
    return ak, bk, ck, alpha * dt_sound ** 2 * Cp * thr8w


def sound_implicit_rk_step(u_now, v_now, w_now, pi0, pip_now, pip_star, rhs_u, rhs_v, rhs_w,
                           theta_rho, rho0, n, dt_sound, grid, alpha=ALPHA_IMPLICIT):
    """Vertically implicit acoustic substeps, solver_opt 3 (sound_implicit_rk_step0/1/2,
    one_step_integration.py:559-635, :678-754, :797-873): u, v forward; w and pi' coupled
    implicitly in the vertical with off-centring alpha (one tridiagonal solve per column);
    rhs_pi = 0 as in the reference stages (:133)."""
    I, J = slice(NGX, -NGX), slice(NGY, -NGY)
    piadv = pi0 + pip_star
    ak, bk, ck, rhs3 = implicit_w_coefficients(pi0, pip_star, theta_rho, dt_sound, grid, alpha)
    beta = 1.0 - alpha
    c = -Rd / Cv
    for _ in range(n):
        pu, pv, pw = pressure_gradient_force(pip_now, theta_rho, grid)
        un = padding3u_array(u_now[I, J, NGZ:-NGZ] + (pu + rhs_u) * dt_sound)
        vn = padding3v_array(v_now[I, J, NGZ:-NGZ] + (pv + rhs_v) * dt_sound)
        wb = beta * w_now
        div = get_divergence2(un, vn, wb, grid)
        ptm = c * (_interior(piadv) * _interior(div)) + advection_pi(rho0, pi0, pip_star, un, vn, wb, grid)
        dptm = (ptm[:, :, 1:] - ptm[:, :, :-1]) / grid.dz
        dk = w_now[I, J, 2:-2] + dt_sound * (rhs_w[:, :, 1:-1] + pw[:, :, 1:-1]) - rhs3 * dptm
        wi = tridiagonal_solve(ak, bk, ck, dk)
        z0 = np.zeros(wi.shape[:2] + (1,), dtype=wi.dtype)
        wn = padding3_array(np.concatenate([z0, wi, z0], axis=2))
        wmix = alpha * wn + beta * w_now
        div = get_divergence2(un, vn, wmix, grid)
        apt = advection_pi(rho0, pi0, pip_star, un, vn, wmix, grid)
        dpdt = c * (_interior(piadv) * _interior(div)) + apt
        pn = padding3_array(_interior(pip_now) + 1.1 * dpdt * dt_sound)
        u_now, v_now, w_now, pip_now = un, vn, wn, pn
    return u_now, v_now, w_now, pip_now


def update_momentum_eqn_euler(u0, v0, w0, pip, theta_rho, adv4u, adv4v, adv4w, fu, fv, b8w, dt, grid):
    """(one_step_integration.py:876-891)."""
    pu, pv, pw = pressure_gradient_force(pip, theta_rho, grid)
    du = adv4u + pu + fv
    dv = adv4v + pv - fu
    dw = adv4w + pw + b8w
    un = padding3u_array(u0[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + du * dt)
    vn = padding3v_array(v0[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + dv * dt)
    wn = set_w_bc(padding3_array(w0[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + dw * dt))
    return un, vn, wn, du, dv, dw


def update_scalar_euler(rho0, s0, s, u, v, w, flow_divergence, sfc_flux, heating, weps, dt, grid):
    """update_theta_euler / update_qv_euler (one_step_integration.py:450-465, :903-916)."""
    tend = advection_scalar(rho0, s, u, v, w, weps, flow_divergence, sfc_flux, grid) + heating
    return padding3_array(_interior(s0) + tend * dt), tend


# ---------------------------------------------------------------------------------
# R1 / R2: stages and the WS-RK3 step — one_step_integration.py:16-447,
# one_sprint_functions.py:10-30, :109-115
# ---------------------------------------------------------------------------------
def _n_acoustic(stage, dt_stage, opts):
    """Acoustic substeps per stage: 1 (stage 0, dt/3), round(dt_stage/(dt/n_sound)) after.

    one_step_integration.py:121-122, :253-254, :383-384 (Python round is half-to-even).
    """
    if stage == 0:
        return 1, dt_stage
    dts = opts.dt / opts.n_sound
    return int(round(dt_stage / dts)), dts


def setup_damping_tau(grid, z_damping):
    """Ramp of the damping layer (setup_lex.py:149-156), FP64, horizontally uniform:
    tauh on the padded cell levels (nz+2), tauf on the padded w levels (nz+3).  The depth
    runs to the top padded w level, (nz+1) dz."""
    nz, dz = grid.nz, grid.dz
    z3d = np.linspace(-0.5 * dz, (nz + 0.5) * dz, nz + 2)
    z3d4w = np.linspace(-dz, (nz + 1) * dz, nz + 3)
    depth = z3d4w[-1] - z_damping
    tauh = 0.5 * (1.0 - np.cos(np.pi * (z3d - z_damping) / depth))
    tauf = 0.5 * (1.0 - np.cos(np.pi * (z3d4w - z_damping) / depth))
    tauh = np.where(z3d <= z_damping, 0.0, tauh)
    tauf = np.where(z3d4w <= z_damping, 0.0, tauf)
    sh = (grid.nx + 6, grid.ny + 6)
    return (np.broadcast_to(tauh, sh + (nz + 2,)), np.broadcast_to(tauf, sh + (nz + 3,)))


def rayleigh_damping(tauh, tauf, u, v, w, theta, rd_alpha):
    """Damping tendencies of u, v, w (towards zero) and theta (towards its horizontal
    mean per level) (boundary_conditions.py:119-141)."""
    I, J, K = slice(NGX, -NGX), slice(NGY, -NGY), slice(NGZ, -NGZ)
    u_t = -rd_alpha * (0.5 * (tauh[NGX - 1:-NGX, J, K] + tauh[NGX:-(NGX - 1), J, K]) * u[I, J, K])
    v_t = -rd_alpha * (0.5 * (tauh[I, NGY - 1:-NGY, K] + tauh[I, NGY:-(NGY - 1), K]) * v[I, J, K])
    w_t = -rd_alpha * (tauf[I, J, K] * w[I, J, K])
    th_i = theta[I, J, K]
    avg = np.mean(th_i, axis=(0, 1), keepdims=True)
    th_t = -rd_alpha * tauh[I, J, K] * (th_i - avg)
    return u_t, v_t, w_t, th_t


def newton_radiation(theta, theta0_ic):
    """Newtonian relaxation of theta to theta0 over 12 h, capped beyond +-1 K
    (radiation.py:7-16); interior."""
    t = -(theta - theta0_ic) / (12.0 * 3600.0)
    t = np.where(theta - theta0_ic > 1.0, -1.0 / (12.0 * 3600.0), t)
    t = np.where(theta - theta0_ic < -1.0, 1.0 / (12.0 * 3600.0), t)
    return t[NGX:-NGX, NGY:-NGY, NGZ:-NGZ]


# ---------------------------------------------------------------------------------
# Surface fluxes: the bulk atmosphere-ocean scheme of CAM (boundary_conditions.py:7-101)
# ---------------------------------------------------------------------------------
ATM_OCN_MAX_ITER = 8     # namelist_n_constants.py:43
ATM_OCN_MIN_WIND = 0.25  # namelist_n_constants.py:44
SFC_Z_REF = 10.0         # namelist_n_constants.py:45
SFC_Z_T_REF = 2.0        # namelist_n_constants.py:46
KARMAN = 0.4             # namelist_n_constants.py:47
CPWV = 1810.0            # namelist_n_constants.py:56
CPVIR = CPWV / Cp - 1.0  # namelist_n_constants.py:57
LAT_VAP = 2.501e6        # namelist_n_constants.py:60


def q_sat_sfc(tk):
    """Saturation vapour density (kg m-3) over the sea surface (boundary_conditions.py:78-81)."""
    return 640380.0 / np.exp(5107.4 / tk)


def cdn(u_ref):
    """Neutral 10 m drag coefficient (boundary_conditions.py:84-87)."""
    return 0.0027 / u_ref + 0.000142 + 0.0000764 * u_ref


def psi_m_s(hol):
    """Integrated stability functions for momentum and scalars at z/L = hol
    (boundary_conditions.py:90-103): -5 hol when stable, the unstable forms otherwise."""
    stable = -5.0 * hol
    xsq = np.maximum(np.sqrt(np.abs(1.0 - 16.0 * hol)), 1.0)
    xqq = np.sqrt(xsq)
    pm_u = np.log((1.0 + xqq * (2.0 + xqq)) * (1.0 + xqq * xqq) / 8.0) - 2.0 * np.arctan(xqq) + 1.571
    ps_u = 2.0 * np.log((1.0 + xqq * xqq) / 2.0)
    return np.where(hol > 0, stable, pm_u), np.where(hol > 0, stable, ps_u)


def atm_ocn_flux(z_bottom, u_bottom, v_bottom, theta_bottom, q_bottom, rho_bottom, surface_t):
    """Bulk fluxes at the lowest level (boundary_conditions.py:7-75): a fixed eight passes
    of the Monin-Obukhov iteration.  Returns (tau_x, tau_y, sen, evap, t_ref, q_ref, u10n)
    with the reference's final sign flip (upward fluxes positive)."""
    v_mag = np.maximum(np.sqrt(u_bottom ** 2 + v_bottom ** 2), ATM_OCN_MIN_WIND)
    ssq = 0.98 * q_sat_sfc(surface_t) / rho_bottom
    delta_t = theta_bottom - surface_t
    delta_q = q_bottom - ssq
    alz = np.log(z_bottom / SFC_Z_REF)
    cp = Cp * (1.0 + CPVIR * ssq)
    ft = np.result_type(theta_bottom).type   # keep FP32 inputs in FP32
    rdn = np.sqrt(cdn(v_mag))
    rhn = np.where(delta_t > 0.0, ft(0.018), ft(0.0327))
    ren = 0.0346
    u_star, t_star, q_star = rdn * v_mag, rhn * delta_t, ren * delta_q
    for _ in range(ATM_OCN_MAX_ITER):
        hol = KARMAN * g * z_bottom * (t_star / theta_bottom + q_star / (1.0 / repsm1 + q_bottom)) / u_star ** 2
        hol = np.clip(hol, -10.0, 10.0)
        psi_mh, psi_xh = psi_m_s(hol)
        rd = rdn / (1.0 + rdn / KARMAN * (alz - psi_mh))
        u10n = v_mag * rd / rdn
        rdn = np.sqrt(cdn(u10n))
        rhn = np.where(hol > 0.0, ft(0.018), ft(0.0327))
        rd = rdn / (1.0 + rdn / KARMAN * (alz - psi_mh))
        rh = rhn / (1.0 + rhn / KARMAN * (alz - psi_xh))
        re = ren / (1.0 + ren / KARMAN * (alz - psi_xh))
        u_star, t_star, q_star = rd * v_mag, rh * delta_t, re * delta_q
    tau = rho_bottom * u_star * u_star
    tau_x = tau * u_bottom / v_mag
    tau_y = tau * v_bottom / v_mag
    sen = cp * tau * t_star / u_star
    evap = LAT_VAP * tau * q_star / u_star / LAT_VAP
    _, psi_x2 = psi_m_s(hol * SFC_Z_T_REF / z_bottom)
    al2 = math.log(SFC_Z_REF / SFC_Z_T_REF)
    fac = (rh / KARMAN) * (alz + al2 - psi_xh + psi_x2)
    t_ref = theta_bottom - delta_t * fac - 0.01 * SFC_Z_T_REF
    fac = (re / KARMAN) * (alz + al2 - psi_xh + psi_x2)
    q_ref = q_bottom - delta_q * fac
    return -tau_x, -tau_y, -sen, -evap, t_ref, q_ref, u10n


def surface_fluxes(xi, base, grid):
    """atm_ocn_flux on the lowest level of state xi as rk_sub_step0 feeds it
    (one_step_integration.py:37-47): u, v averaged to the cell centre, rho0 of level 1,
    the height of the first cell centre; base[5] is the surface temperature (nx, ny)."""
    th, u, v, _, _, qv = xi[:6]
    rho0, surface_t = base[1], base[5]
    I, J, k = slice(NGX, -NGX), slice(NGY, -NGY), NGZ
    z_bottom = np.full((grid.nx, grid.ny), 0.5 * grid.dz, dtype=rho0.dtype)
    u_b = 0.5 * (u[NGX:-(NGX + 1), J, k] + u[NGX + 1:-NGX, J, k])
    v_b = 0.5 * (v[I, NGY:-(NGY + 1), k] + v[I, NGY + 1:-NGY, k])
    return atm_ocn_flux(z_bottom, u_b, v_b, th[I, J, k], qv[I, J, k], rho0[I, J, k],
                        np.asarray(surface_t, dtype=rho0.dtype))


def rk_sub_step(stage, xi0, xi, base, grid, opts, dt_stage, frozen=None, heating=None, sfc=None):
    """One WS-RK3 stage (rk_sub_step0/1/2, one_step_integration.py:16-188, :191-317, :320-447).

    State tuples are (theta, u, v, w, pip, qv, *passive) with optional passive scalars
    (qc, qr) advected and SGS-diffused like qv.  `frozen` carries the stage-0 SGS
    tendencies (one_step_integration.py:57-76, :245) into stages 1 and 2; `heating` the
    Newtonian heating and `sfc` the surface fluxes (tau_x, tau_y, sen, evap, ...) of the
    step's first stage (one_step_integration.py:31-52, :208-209).
    Returns (xi_next, frozen, info) with info the BiCGSTAB exit code (solver_opt 1).
    """
    rho0_theta0, rho0, theta0, pi0, qv0 = base[:5]
    th0, u0, v0, w0, pip0, qv0_now = xi0[:6]
    th, u, v, w, pip, qv = xi[:6]
    pas0 = tuple(xi0[6:])
    pas = tuple(xi[6:])
    dt_ = rho0.dtype.type
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    zeros2 = np.zeros((nx, ny), dtype=rho0.dtype)
    if stage == 0 and frozen is None:
        if opts.turb_opt in (1, 2):
            frozen = smagorinsky(rho0, u, v, w, th, qv, (th, qv) + pas, grid, opts.cs, opts.prandtl, opts.turb_opt)
    sgs = frozen
    buoy, b8w = calculate_buoyancy(theta0, th, qv0, qv)
    fdiv = get_divergence(rho0, u, v, w, grid)
    tau_x, tau_y, sen, evap = (zeros2,) * 4 if sfc is None else tuple(sfc[:4])
    adv4u = advection_u(rho0, u, v, w, WEPS_MOM, fdiv, tau_x, grid)
    adv4v = advection_v(rho0, u, v, w, WEPS_MOM, fdiv, tau_y, grid)
    adv4w = advection_w(rho0, u, v, w, WEPS_MOM, fdiv, grid)
    if opts.cor_opt == 1:
        fu, fv = calculate_coriolis_force(u, v)
    elif opts.cor_opt == 2:
        fu, fv = calculate_coriolis_force_with_lspgrad(u, v)
    else:
        fu = np.zeros((nx, ny + 1, nz), rho0.dtype)
        fv = np.zeros((nx + 1, ny, nz), rho0.dtype)
    info = None
    if opts.solver_opt == 1:
        rtt = calculate_rtt(rho0_theta0, th, qv, buoy)
        pint, info = solve_pres_eqn(pip0, rho0_theta0, rtt, adv4u, adv4v, adv4w, fu, fv, buoy, grid, opts)
        pip_n = padding3_array(pint)
        trho = theta_rho_of(th, qv)
        un, vn, wn, _, _, _ = update_momentum_eqn_euler(u0, v0, w0, pip_n, trho, adv4u, adv4v, adv4w,
                                                        fu, fv, b8w, dt_stage, grid)
        if sgs is not None:
            un = un + padding3u_array(sgs[0] * dt_stage)
            vn = vn + padding3v_array(sgs[1] * dt_stage)
            wn = wn + padding3_array(sgs[2] * dt_stage)
    else:
        rhs_u = adv4u + fv
        rhs_v = adv4v - fu
        rhs_w = adv4w + b8w
        if sgs is not None:
            rhs_u = rhs_u + sgs[0]
            rhs_v = rhs_v + sgs[1]
            rhs_w = rhs_w + sgs[2]
        trho = theta_rho_of(th, qv)
        n, dts = _n_acoustic(stage, dt_stage, opts)
        step = sound_implicit_rk_step if opts.solver_opt == 3 else sound_rk_step
        un, vn, wn, pip_n = step(u0, v0, w0, pi0, pip0, pip, rhs_u, rhs_v, rhs_w, trho, rho0, n, dts, grid)
    thn, _ = update_scalar_euler(rho0, th0, th, u, v, w, fdiv, sen / Cp, 0.0 if heating is None else heating,
                                 WEPS_THETA, dt_stage, grid)
    qvn, _ = update_scalar_euler(rho0, qv0_now, qv, u, v, w, fdiv, evap, 0.0, WEPS_QV, dt_stage, grid)
    pasn = tuple(update_scalar_euler(rho0, a0, a, u, v, w, fdiv, zeros2, 0.0, WEPS_QV, dt_stage, grid)[0]
                 for a0, a in zip(pas0, pas))
    if sgs is not None:
        thn = thn + padding3_array(sgs[3] * dt_stage)
        qvn = qvn + padding3_array(sgs[4] * dt_stage)
        pasn = tuple(a + padding3_array(t * dt_stage) for a, t in zip(pasn, sgs[5:]))
    if opts.damp_opt:  # one_step_integration.py:164-175
        tauh, tauf = setup_damping_tau(grid, opts.z_damping)
        tauh, tauf = tauh.astype(rho0.dtype), tauf.astype(rho0.dtype)
        ut, vt, wt, tht = rayleigh_damping(tauh, tauf, un, vn, wn, thn, opts.rd_alpha)
        un = un + padding3u_array(ut * dt_stage)
        vn = vn + padding3v_array(vt * dt_stage)
        wn = wn + padding3_array(wt * dt_stage)
        thn = thn + padding3_array(tht * dt_stage)
    del dt_
    return (thn, un, vn, wn, pip_n, qvn) + pasn, frozen, info


def ws_rk3_big_step(xi0, base, grid, opts):
    """Wicker-Skamarock RK3: every stage restarts from xi0 (one_sprint_functions.py:10-30).
    With rad_opt the heating, with sfc_opt the surface fluxes are computed from the stage-0
    state and kept for the step (one_step_integration.py:31-52, :208-209, :338-339)."""
    kw = {"heating": newton_radiation(xi0[0], base[2])} if opts.rad_opt else {}
    if opts.sfc_opt:
        kw["sfc"] = surface_fluxes(xi0, base, grid)
    xi1, frozen, i0 = rk_sub_step(0, xi0, xi0, base, grid, opts, opts.dt / 3.0, **kw)
    xi2, _, i1 = rk_sub_step(1, xi0, xi1, base, grid, opts, opts.dt / 2.0, frozen, **kw)
    xi3, _, i2 = rk_sub_step(2, xi0, xi2, base, grid, opts, opts.dt, frozen, **kw)
    return xi3, (i0, i1, i2)


def ws_rk3_big_step_ext_sgs(xi0, base, grid, opts, sgs):
    """ws_rk3_big_step with caller-supplied SGS tendencies (sgs_u, sgs_v, sgs_w, then one
    per scalar) frozen over the step in place of the stage-0 Smagorinsky ones: the path
    of the learned SGS net (its outputs replace compute_smag's, one_step_integration.py:57-76)."""
    xi1, _, i0 = rk_sub_step(0, xi0, xi0, base, grid, opts, opts.dt / 3.0, tuple(sgs))
    xi2, _, i1 = rk_sub_step(1, xi0, xi1, base, grid, opts, opts.dt / 2.0, tuple(sgs))
    xi3, _, i2 = rk_sub_step(2, xi0, xi2, base, grid, opts, opts.dt, tuple(sgs))
    return xi3, (i0, i1, i2)


def ws_rk3_big_step_stage_sgs(xi0, base, grid, opts, sgs_fn):
    """ws_rk3_big_step with the SGS tendencies recomputed at the start of every stage:
    sgs_fn(stage, xi) returns (sgs_u, sgs_v, sgs_w, then one per scalar) from the stage's
    input state xi (xi0, then the first and second stage results).  No counterpart in the
    reference, which computes its SGS in stage 0 and freezes it (one_step_integration.py:57-76);
    this is the learned SGS net run inside each stage.  A tendency that is a linear
    function of the state then advances by the WS-RK3 polynomial
    1 + L dt + (L dt)^2 / 2 + (L dt)^3 / 6 instead of the frozen 1 + L dt."""
    xi1, _, i0 = rk_sub_step(0, xi0, xi0, base, grid, opts, opts.dt / 3.0, tuple(sgs_fn(0, xi0)))
    xi2, _, i1 = rk_sub_step(1, xi0, xi1, base, grid, opts, opts.dt / 2.0, tuple(sgs_fn(1, xi1)))
    xi3, _, i2 = rk_sub_step(2, xi0, xi2, base, grid, opts, opts.dt, tuple(sgs_fn(2, xi2)))
    return xi3, (i0, i1, i2)


def sgs_features(phys, base, grid):
    """(nx, ny, nz, 11) inputs of the learned SGS net from a padded state: u, v, w at cell
    centres, theta' and qv' against the base columns, and the six deformation components
    S11, S22, S33, S12, S13, S23 of compute_deformation (turbulence.py:298-388) with rho = 1."""
    th, u, v, w, _, qv = phys[:6]
    I = (slice(NGX, -NGX), slice(NGY, -NGY), slice(NGZ, -NGZ))
    uc, vc, wc = unstagger_uvw(u[I], v[I], w[I])
    s = compute_deformation(np.ones_like(th), u, v, w, grid)
    chans = [uc, vc, wc, th[I] - base[2][I], qv[I] - base[4][I]] + [a[I] for a in s[:6]]
    return np.stack(chans, axis=3)


def sgs_apply(y, out_scale):
    """(nx, ny, nz, 5) cell-centred net outputs times out_scale -> the SGS tendencies on
    compute_smag's points (turbulence.py:9-111): sgs_u on the nx+1 x faces and sgs_v on
    the ny+1 y faces as the mean of the two adjacent cells (periodic), sgs_w on the nz+1
    levels as the mean of the cells below and above (zero at the surface and lid),
    sgs_theta and sgs_qv at the cells."""
    t = [y[..., c] * out_scale[c] for c in range(5)]
    ext_x = np.concatenate([t[0][-1:], t[0]], 0)
    ext_x = np.concatenate([ext_x, ext_x[1:2]], 0)
    su = 0.5 * (ext_x[:-1] + ext_x[1:])
    ext_y = np.concatenate([t[1][:, -1:], t[1]], 1)
    ext_y = np.concatenate([ext_y, ext_y[:, 1:2]], 1)
    sv = 0.5 * (ext_y[:, :-1] + ext_y[:, 1:])
    z = np.zeros(y.shape[:2] + (1,), y.dtype)
    sw = np.concatenate([z, 0.5 * (t[2][:, :, :-1] + t[2][:, :, 1:]), z], 2)
    return su, sv, sw, t[3], t[4]


def ws_rk3_sprint(xi, base, grid, opts, nsteps):
    """nsteps WS-RK3 steps (one_sprint_functions.py:109-115 with sprint_n = nsteps)."""
    info = None
    for _ in range(nsteps):
        xi, info = ws_rk3_big_step(xi, base, grid, opts)
    return xi, info


# ---------------------------------------------------------------------------------
# I1: grid, base state and warm-bubble initial conditions — setup_lex.py:9-167
# ---------------------------------------------------------------------------------
def rslf(p, t):
    """Bolton (1980) saturation mixing ratio over liquid (setup_lex.py:159-167)."""
    esl = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))
    esl = np.where(esl > p * 0.5, p * 0.5, esl)
    return eps * esl / (p - esl)


def hydrostatic_base_state(grid, theta0_value, rh_of_z, tol=1e-9, max_iter=200):
    """FP64 hydrostatic base state iterated to qv0 convergence (setup_lex.py:103-137).

    Returns 1-D padded columns (nz+2) of pi0, rho0_theta0, rho0, theta0, qv0 and the
    iteration count.  `rh_of_z(z)` gives the base relative humidity at cell centres.
    """
    nz, dz = grid.nz, grid.dz
    z = (np.arange(nz + 2) - 0.5) * dz     # setup_lex.py:14
    theta0 = np.full(nz + 2, float(theta0_value))
    rh = rh_of_z(z)

    def integrate(qv0):
        dpdz = -g / Cp / (theta0 * (1.0 + reps * qv0) / (1.0 + qv0))
        part = np.cumsum(dpdz[1:-1]) * dz + 1.0
        p8w = np.concatenate([[1.0], part])
        pi0 = np.empty(nz + 2)
        pi0[1:-1] = 0.5 * (p8w[:-1] + p8w[1:])
        pi0[0] = pi0[1]
        pi0[-1] = pi0[-2]
        return pi0

    pi0 = integrate(np.zeros(nz + 2))
    qv0 = rslf(pi0 ** (Cp / Rd) * p00, theta0 * pi0) * rh
    prev = qv0 * 1.5
    it = 0
    while np.max(np.abs(qv0 - prev)) > tol and it < max_iter:
        prev = qv0
        pi0 = integrate(qv0)
        qv0 = rslf(pi0 ** (Cp / Rd) * p00, theta0 * pi0) * rh
        it += 1
    rho0_theta0 = pi0 ** (Cv / Rd) * p00 / Rd
    rho0 = rho0_theta0 / theta0 / ((1.0 + reps * qv0) / (1.0 + qv0))
    # setup_lex.py:64 overwrites rho0_theta0 with rho0 * theta0
    rho0_theta0 = rho0 * theta0
    return dict(pi0=pi0, rho0_theta0=rho0_theta0, rho0=rho0, theta0=theta0, qv0=qv0, iters=it, z=z)


@dataclass
class BubbleCase:
    """Warm-bubble initial-condition recipe (setup_lex.py:74-146, with the workload's sizes)."""
    grid: Grid
    amplitude: float = 2.0
    xr: float = 2000.0
    yr: float = 2000.0
    zr: float = 1400.0
    zc: float = 1400.0
    xc: float | None = None      # default: domain centre
    yc: float | None = None
    rh_surface: float = 0.0      # base RH profile: rh_surface -> rh_top at z_rh_top (linear)
    rh_top: float = 0.0
    z_rh_top: float = 10000.0
    rh_bubble: float = 0.0       # extra RH inside the bubble (times the cos^2 shape)
    theta_noise: tuple = ("none", 0.0, 2025)   # ("uniform"|"normal"|"none", scale, seed)
    uvw_noise: tuple = (0.0, 2026)             # (sigma, first seed; u, v, w use seed, +1, +2)
    passive: int = 0                            # number of passive scalars (qc, qr)
    extra: dict = field(default_factory=dict)


def setup_warm_bubble(case: BubbleCase, dtype=np.float64):
    """Build (phys_state, base_state) as padded arrays (setup_grid_n_ic, setup_lex.py:9-71)."""
    gr = case.grid
    nx, ny, nz = gr.nx, gr.ny, gr.nz

    def rh_of_z(z):
        zz = np.minimum(np.maximum(z, 0.0), case.z_rh_top)
        return case.rh_surface + (case.rh_top - case.rh_surface) * zz / case.z_rh_top

    bs = hydrostatic_base_state(gr, 300.0, rh_of_z)
    col = lambda a: np.broadcast_to(a[None, None, :], gr.shape_c).astype(np.float64)  # noqa: E731
    pi0, rt0, rho0, th0, qv0 = (col(bs[k]) for k in ("pi0", "rho0_theta0", "rho0", "theta0", "qv0"))
    x = (np.arange(nx + 6) - 2.5) * gr.dx
    y = (np.arange(ny + 6) - 2.5) * gr.dy
    z = (np.arange(nz + 2) - 0.5) * gr.dz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    xc = case.xc if case.xc is not None else 0.5 * nx * gr.dx
    yc = case.yc if case.yc is not None else 0.5 * ny * gr.dy
    r = np.sqrt(((X - xc) / case.xr) ** 2 + ((Y - yc) / case.yr) ** 2 + ((Z - case.zc) / case.zr) ** 2)
    shape = np.where(r > 1.0, 0.0, np.cos(r * np.pi / 2.0) ** 2)
    theta_p = case.amplitude * shape
    kind, scale, seed = case.theta_noise
    if kind != "none" and scale != 0.0:
        rng = np.random.default_rng(seed)
        if kind == "uniform":
            nz_ = rng.uniform(-scale, scale, size=(nx, ny, nz))
        else:
            nz_ = rng.normal(0.0, scale, size=(nx, ny, nz))
        theta_p = theta_p + padding3_array(nz_)
    theta = th0 + theta_p
    # qv from the local RH and the estimated total pi (setup_lex.py:140-144)
    rh = rh_of_z(Z) + case.rh_bubble * shape
    pi = (rho0 * theta * (1.0 + reps * qv0) / (1.0 + qv0) * Rd / p00) ** (Rd / Cv)
    qv = rslf(pi ** (Cp / Rd) * p00, theta * pi) * rh
    u = np.zeros(gr.shape_u)
    v = np.zeros(gr.shape_v)
    w = np.zeros(gr.shape_w)
    sig, s0 = case.uvw_noise
    if sig != 0.0:
        ui = np.random.default_rng(s0).normal(0.0, sig, size=(nx, ny, nz))
        u = padding3u_array(np.concatenate([ui, ui[:1]], 0))        # face nx = face 0
        vi = np.random.default_rng(s0 + 1).normal(0.0, sig, size=(nx, ny, nz))
        v = padding3v_array(np.concatenate([vi, vi[:, :1]], 1))
        wi = np.random.default_rng(s0 + 2).normal(0.0, sig, size=(nx, ny, nz + 1))
        w = set_w_bc(padding3_array(wi))
    pip = np.zeros(gr.shape_c)
    phys = [theta, u, v, w, pip, qv] + [np.zeros(gr.shape_c) for _ in range(case.passive)]
    base = (rt0, rho0, th0, pi0, qv0, np.full((nx, ny), 300.0))
    phys = tuple(np.ascontiguousarray(a, dtype=dtype) for a in phys)
    base = tuple(np.ascontiguousarray(a, dtype=dtype) for a in base)
    return phys, base, bs


def case_config0():
    """64x64x40, 600/600/300 m, dt 30 s, pseudo-incompressible, Smagorinsky, dry."""
    gr = Grid(64, 64, 40, 600.0, 600.0, 300.0)
    case = BubbleCase(gr, amplitude=2.0, theta_noise=("uniform", 0.01, 2025))
    opts = Options(dt=30.0, solver_opt=1, n_sound=32, turb_opt=1)
    return case, opts


def case_config1():
    """256x256x128, 100 m, dt 1 s, compressible, n_sound 6, moist, passive qc/qr."""
    gr = Grid(256, 256, 128, 100.0, 100.0, 100.0)
    case = BubbleCase(gr, amplitude=2.0, rh_surface=0.8, rh_top=0.2, z_rh_top=10000.0, rh_bubble=0.1,
                      theta_noise=("normal", 0.05, 2025), uvw_noise=(0.01, 2026), passive=2)
    opts = Options(dt=1.0, solver_opt=2, n_sound=6, turb_opt=1)
    return case, opts


def case_config2():
    """128x128x64, 200 m, dt 2 s, compressible, n_sound 10 (config 2 states no dt)."""
    gr = Grid(128, 128, 64, 200.0, 200.0, 200.0)
    case = BubbleCase(gr, amplitude=2.0, rh_surface=0.8, rh_top=0.2, rh_bubble=0.1,
                      theta_noise=("normal", 0.1, 2025))
    opts = Options(dt=2.0, solver_opt=2, n_sound=10, turb_opt=1)
    return case, opts


def scaled_case(case: BubbleCase, nx, ny, nz):
    """Same recipe on a smaller grid of the same spacing, bubble kept centred."""
    gr = Grid(nx, ny, nz, case.grid.dx, case.grid.dy, case.grid.dz)
    c = BubbleCase(**{**case.__dict__, "grid": gr})
    return c


# ---------------------------------------------------------------------------------
# X5: block-mean coarsening (initial_comparison/plot_theta_pert_coarsened.m:7-8)
# ---------------------------------------------------------------------------------
def coarsen_block_mean(a, fx, fy, fz):
    """Mean over non-overlapping fx*fy*fz blocks of an interior (nx, ny, nz) field."""
    nx, ny, nz = a.shape
    return a.reshape(nx // fx, fx, ny // fy, fy, nz // fz, fz).mean(axis=(1, 3, 5))


def coarsen_face_mean(a, axis, fx, fy, fz):
    """Face-area mean of a staggered interior velocity onto the coarse faces: u (nx+1, ny, nz)
    with axis 0, v (nx, ny+1, nz) with axis 1, w (nx, ny, nz+1) with axis 2.  The coarse
    face n is fine face n*f of the staggered axis, averaged over the f_other1*f_other2
    fine faces it covers.  (No reference routine: the reference coarsens theta only,
    initial_comparison/plot_theta_pert_coarsened.m:7-8; this restriction keeps the
    block-mean of the divergence, see tests.)"""
    f = (fx, fy, fz)
    sl = [slice(None)] * 3
    sl[axis] = slice(0, None, f[axis])
    b = a[tuple(sl)]
    shp = list(b.shape)
    red = []
    new = []
    for ax in range(3):
        if ax == axis:
            new.append(shp[ax])
        else:
            new += [shp[ax] // f[ax], f[ax]]
            red.append(len(new) - 1)
    return b.reshape(new).mean(axis=tuple(red))


# ---------------------------------------------------------------------------------
# The reference's post-step deep-learning correction (one_sprint_functions.py:118-165,
# dl_models.py:125-182).  The network itself is a parameter (net_fn), so these restate
# only the data maps and the correction loop around it.
# ---------------------------------------------------------------------------------
def dl_level_norm(scaling_params):
    """Per-level offsets and denominators of phys_state_tuple2array (dl_models.py:127-151):
    for theta, u, v, w, qv: lo = min over x, y of the squeezed minimum field, den = max
    over x, y of the squeezed maximum field - lo + 0.01 (1e-5 for qv)."""
    mins = scaling_params[0:10:2]
    maxs = scaling_params[1:10:2]
    out = []
    for q, (mn, mx) in enumerate(zip(mins, maxs)):
        lo = np.min(np.squeeze(mn), axis=(0, 1))
        c = 1.0e-5 if q == 4 else 0.01
        den = np.max(np.squeeze(mx), axis=(0, 1)) - lo + c
        out.append((lo, den))
    return out


def unstagger_uvw(u, v, w):
    """Face values to cell centres (dl_models.py:166-171)."""
    return 0.5 * (u[:-1] + u[1:]), 0.5 * (v[:, :-1] + v[:, 1:]), 0.5 * (w[:, :, :-1] + w[:, :, 1:])


def stagger_uvw(ui, vi, wi):
    """Centres back to faces (dl_models.py:174-182): interior faces average their two
    centres; the first x (y) face takes the last interior face's value and the last face
    the first's; w is zero at the surface and the lid."""
    uc = 0.5 * (ui[:-1] + ui[1:])
    vc = 0.5 * (vi[:, :-1] + vi[:, 1:])
    wc = 0.5 * (wi[:, :, :-1] + wi[:, :, 1:])
    u = np.concatenate([uc[-1:], uc, uc[:1]], 0)
    v = np.concatenate([vc[:, -1:], vc, vc[:, :1]], 1)
    z = np.zeros(wi.shape[:2] + (1,), wi.dtype)
    w = np.concatenate([z, wc, z], 2)
    return u, v, w


def phys_state_tuple2array(phys_state, scaling_params):
    """Normalised, centred (nx, ny, nz, 5) network input (dl_models.py:125-155)."""
    th, u, v, w, _, qv = phys_state[:6]
    (lt, dt_), (lu, du), (lv, dv), (lw, dw), (lq, dq) = dl_level_norm(scaling_params)
    thn = (_interior(th) - lt) / dt_
    un = (u[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] - lu) / du
    vn = (v[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] - lv) / dv
    wn = (w[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] - lw) / dw
    qn = (_interior(qv) - lq) / dq
    ui, vi, wi = unstagger_uvw(un, vn, wn)
    return np.stack([thn, ui, vi, wi, qn], axis=3)


def correct_array2tuple(correct):
    """(nx, ny, nz, 5) -> staggered corrections of theta, u, v, w, qv (dl_models.py:158-163)."""
    tc, uic, vic, wic, qc = (correct[..., q] for q in range(5))
    uc, vc, wc = stagger_uvw(uic, vic, wic)
    return tc, uc, vc, wc, qc


# ---------------------------------------------------------------------------------
# AutoEncoder (dl_models.py:33-68): the reference's default post-step correction net
# (namelist_n_constants.py:78).  flax's Conv / ConvTranspose / Dense are third-party;
# their semantics are restated from JAX's public lax (conv_general_dilated with 'SAME'
# strides, conv_transpose with transpose_kernel=False).  Pinned by per-output loops and
# by torch's conv_transpose3d on flipped kernels (tests/test_oracle_autoencoder.py);
# flax's initialisers are not restated (parity unpinned).
# ---------------------------------------------------------------------------------
def same_pads(n, k, s):
    """lax 'SAME' padding (lo, hi) of a window k, stride s on n points: the output has
    ceil(n / s) points and the total padding is split with the smaller half low."""
    out = -(-n // s)
    total = max((out - 1) * s + k - n, 0)
    return total // 2, total - total // 2


def transpose_pads(k, s, padding):
    """The padding lax.conv_transpose gives the input-dilated correlation for 'SAME' and
    'VALID' (JAX's _conv_transpose_padding)."""
    if padding == "SAME":
        pad_len = k + s - 2
        pad_a = k - 1 if s > k - 1 else int(math.ceil(pad_len / 2))
    else:
        pad_len = k + s - 2 + max(k - s, 0)
        pad_a = k - 1
    return pad_a, pad_len - pad_a


def conv3d_general(x, W, b=None, stride=1, pads=((1, 1),) * 3, dil=1, relu=False):
    """Correlation of x (X, Y, Z, ci) with W (k, k, k, ci, co) over x dilated by `dil`
    (dil - 1 zeros between points) and zero-padded by `pads`, window stride `stride`;
    + b; optional ReLU."""
    if dil > 1:
        xd = np.zeros(tuple((n - 1) * dil + 1 for n in x.shape[:3]) + x.shape[3:], x.dtype)
        xd[::dil, ::dil, ::dil] = x
        x = xd
    xp = np.pad(x, tuple(pads) + ((0, 0),))
    k = W.shape[0]
    no = tuple((xp.shape[a] - k) // stride + 1 for a in range(3))
    out = np.zeros(no + (W.shape[4],), np.result_type(x, W))
    for tx in range(k):
        for ty in range(k):
            for tz in range(k):
                sl = xp[tx:tx + stride * (no[0] - 1) + 1:stride, ty:ty + stride * (no[1] - 1) + 1:stride,
                        tz:tz + stride * (no[2] - 1) + 1:stride]
                out = out + sl @ W[tx, ty, tz]
    if b is not None:
        out = out + b
    return np.maximum(out, 0) if relu else out


# (kind, features, stride or padding) in the order of dl_models.py:36-66; ReLU after all
# but the last layer
AUTOENCODER_LAYERS = (("conv", 64, 1), ("conv", 32, 2), ("conv", 32, 2), ("conv", 64, 2), ("conv", 128, 2),
                      ("conv", 256, 2), ("dense", 512, None), ("convT", 256, "SAME"), ("convT", 128, "VALID"),
                      ("convT", 64, "SAME"), ("convT", 32, "SAME"), ("convT", 32, "SAME"), ("conv", 64, 1),
                      ("conv", 5, 1))


def autoencoder_forward(params, x):
    """AutoEncoder (dl_models.py:33-68) on one sample x (X, Y, Z, 5); params: (W, b) per
    layer, conv kernels (3, 3, 3, ci, co), the dense kernel (ci, co)."""
    h = x
    last = len(AUTOENCODER_LAYERS) - 1
    for q, ((kind, _, arg), (W, b)) in enumerate(zip(AUTOENCODER_LAYERS, params)):
        relu = q != last
        if kind == "conv":
            pads = tuple(same_pads(n, 3, arg) for n in h.shape[:3])
            h = conv3d_general(h, W, b, stride=arg, pads=pads, relu=relu)
        elif kind == "convT":
            pads = (transpose_pads(3, 2, arg),) * 3
            h = conv3d_general(h, W, b, stride=1, pads=pads, dil=2, relu=relu)
        else:   # flatten, Dense, reshape to (1, 1, 1, features) (dl_models.py:49-52)
            h = h.reshape(1, 1, 1, -1) @ W + b
            h = np.maximum(h, 0) if relu else h
    return h


def dl_apply_correction(phys_state, correct, scaling_params):
    """The correction step of ws_rk3_sprint_dl (one_sprint_functions.py:137-160): each
    field's interior plus its correction times the scaling range, padded again; pi'
    unchanged."""
    th, u, v, w, pip, qv = phys_state[:6]
    tc, uc, vc, wc, qc = correct_array2tuple(correct)
    mins = scaling_params[0:10:2]
    maxs = scaling_params[1:10:2]
    rng = [np.squeeze(mx - mn) for mn, mx in zip(mins, maxs)]
    thn = _interior(th) + tc * rng[0]
    un = u[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + uc * rng[1]
    vn = v[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + vc * rng[2]
    wn = w[NGX:-NGX, NGY:-NGY, NGZ:-NGZ] + wc * rng[3]
    qn = _interior(qv) + qc * rng[4]
    return (padding3_array(thn), padding3u_array(un), padding3v_array(vn), padding3_array(wn), pip,
            padding3_array(qn)) + tuple(phys_state[6:])


def ws_rk3_sprint_dl(xi, base, grid, opts, net_fn, scaling_params, sprint_n):
    """ws_rk3_sprint_dl (one_sprint_functions.py:118-165): every step followed by the
    network's correction; net_fn maps the (nx, ny, nz, 5) input to the (nx, ny, nz, 5)
    correction."""
    for _ in range(sprint_n):
        xi, _ = ws_rk3_big_step(xi, base, grid, opts)
        corr = net_fn(phys_state_tuple2array(xi, scaling_params))
        xi = dl_apply_correction(xi, corr, scaling_params)
    return xi


def ws_rk3_sprint_dl_train(xi, base, grid, opts, net_fn, scaling_params, sprint_n):
    """ws_rk3_sprint_dl_train (one_sprint_functions.py:34-97): ws_rk3_sprint_dl keeping
    each step's interior theta, u, v, w, pi', qv (after the correction) in stacks of
    sprint_n."""
    stacks = [[] for _ in range(6)]
    for _ in range(sprint_n):
        xi = ws_rk3_sprint_dl(xi, base, grid, opts, net_fn, scaling_params, 1)
        for q in range(6):
            stacks[q].append(xi[q][NGX:-NGX, NGY:-NGY, NGZ:-NGZ])
    return tuple(np.stack(s) for s in stacks), xi


def laplace_of_tensor(s, grid):
    """laplace_of_tensor (dataset_generation.py:271-320) of interior scalar fields with any
    leading axes (..., nx, ny, nz): periodic in x and y, the end levels copied in z.  The
    reference applies it to 5-D stacks, where its s_z[:, :, (0, -1)] = 0 indexes the x
    axis: the z derivative of the first and last x planes is zeroed (no vertical term
    there), kept here on the x axis whatever the leading axes."""
    sp = np.concatenate([s[..., -1:, :, :], s, s[..., :1, :, :]], axis=-3)
    sp = np.concatenate([sp[..., :, -1:, :], sp, sp[..., :, :1, :]], axis=-2)
    sp = np.concatenate([sp[..., :1], sp, sp[..., -1:]], axis=-1)
    sx = (sp[..., 1:, 1:-1, 1:-1] - sp[..., :-1, 1:-1, 1:-1]) / grid.dx
    sy = (sp[..., 1:-1, 1:, 1:-1] - sp[..., 1:-1, :-1, 1:-1]) / grid.dy
    sz = (sp[..., 1:-1, 1:-1, 1:] - sp[..., 1:-1, 1:-1, :-1]) / grid.dz
    sz[..., 0, :, :] = 0.0
    sz[..., -1, :, :] = 0.0
    return ((sx[..., 1:, :, :] - sx[..., :-1, :, :]) / grid.dx + (sy[..., :, 1:, :] - sy[..., :, :-1, :]) / grid.dy
            + (sz[..., 1:] - sz[..., :-1]) / grid.dz)


def dl_loss_weights(scaling_params):
    """The per-point weights of the training loss (lex_dl_train.py:69-96): for each term
    (theta, u, v, w, qv, the Laplacians of theta and qv) the scaling range divided by the
    cube of the level range plus its epsilon (0.01, 1e-5 for qv, 1e-7 and 1e-10 for the
    Laplacians), the level range being the max over x, y of the maximum field minus the
    min over x, y of the minimum field; and the negative-qv weight with the square."""
    eps = (0.01, 0.01, 0.01, 0.01, 1.0e-5, 1.0e-7, 1.0e-10)
    w = []
    for q in range(7):
        mn = np.asarray(scaling_params[2 * q], dtype=np.float64)
        mx = np.asarray(scaling_params[2 * q + 1], dtype=np.float64)
        lev = (np.max(mx, axis=(-3, -2), keepdims=True) - np.min(mn, axis=(-3, -2), keepdims=True) + eps[q])
        w.append((mx - mn) / lev ** 3)
    mn, mx = np.asarray(scaling_params[8], np.float64), np.asarray(scaling_params[9], np.float64)
    lev = np.max(mx, axis=(-3, -2), keepdims=True) - np.min(mn, axis=(-3, -2), keepdims=True) + 1.0e-5
    w.append((mx - mn) / lev ** 2)
    return w


def dl_loss(stacks, truth, scaling_params, grid):
    """_loss_and_aux (lex_dl_train.py:49-99) given the sprint's stacks (theta, u, v, w, pip,
    qv; leading axes batch, step) and the truth (theta, u, v, w, pip, qv, del_th, del_qv):
    weighted l2 terms of theta, u, v, w, qv and of the Laplacians of theta - theta_min and
    qv - qv_min, plus the weighted negative-qv penalty.  FP64."""
    th, u, v, w, _, qv = (np.asarray(a, np.float64) for a in stacks)
    w8 = dl_loss_weights(scaling_params)
    theta_min, qv_min = (np.asarray(scaling_params[i], np.float64) for i in (0, 8))
    del_th = laplace_of_tensor(th - theta_min, grid)
    del_qv = laplace_of_tensor(qv - qv_min, grid)
    preds = (th, u, v, w, qv, del_th, del_qv)
    trues = tuple(np.asarray(truth[i], np.float64) for i in (0, 1, 2, 3, 5, 6, 7))
    loss = sum(float(np.sum(0.5 * (p - t) ** 2 * wq)) for p, t, wq in zip(preds, trues, w8[:7]))
    return loss + float(np.sum(np.where(qv < 0.0, -qv, 0.0) * w8[7]))


def n_acoustic_total(opts):
    """Acoustic substeps per full step: 1 + round(n/2) + n (solver_opt 2)."""
    return 1 + _n_acoustic(1, opts.dt / 2.0, opts)[0] + _n_acoustic(2, opts.dt, opts)[0]


def python_round(x):
    """Python's round (half-to-even), kept explicit for the acoustic counts."""
    return int(round(x))


__all__ = [n for n in dir() if not n.startswith("_")] + ["_n_acoustic", "_interior"]
_ = math
