"""Grid, base state and warm-bubble initial conditions (setup_lex.py:9-167).

The hydrostatic base state is integrated in FP64 on the GPU (lex_base_state); the
3-D initial fields are then built on the host from the FP64 profiles (as the
reference does in NumPy) and uploaded as float32 device tensors.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import namelist_n_constants as nl

Rd, Rv, Cp, p00 = nl.Rd, nl.Rv, nl.Cp, nl.p00
Cv = Cp - Rd
reps = Rv / Rd
eps = Rd / Rv


@dataclass
class BubbleCase:
    """Warm-bubble recipe (setup_ic_option1, setup_lex.py:74-146) with configurable sizes."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    amplitude: float = 2.0
    xr: float = 2000.0
    yr: float = 2000.0
    zr: float = 1400.0
    zc: float = 1400.0
    xc: float | None = None
    yc: float | None = None
    rh_surface: float = 0.0
    rh_top: float = 0.0
    z_rh_top: float = 10000.0
    rh_bubble: float = 0.0
    theta_noise: tuple = ("none", 0.0, 2025)
    uvw_noise: tuple = (0.0, 2026)
    passive: int = 0
    extra: dict = field(default_factory=dict)

    def scaled(self, nx, ny, nz):
        c = BubbleCase(**self.__dict__)
        c.nx, c.ny, c.nz = nx, ny, nz
        return c


def config0():
    """64x64x40, 600/600/300 m, dt 30 s, pseudo-incompressible (solver_opt 1), Smagorinsky, dry."""
    case = BubbleCase(64, 64, 40, 600.0, 600.0, 300.0, amplitude=2.0, theta_noise=("uniform", 0.01, 2025))
    opts = dict(dt=30.0, solver_opt=1, n_sound=32, turb_opt=1, steps=60)
    return case, opts


def config1():
    """256x256x128, 100 m, dt 1 s, compressible (solver_opt 2), n_sound 6, moist, passive qc/qr."""
    case = BubbleCase(256, 256, 128, 100.0, 100.0, 100.0, amplitude=2.0, rh_surface=0.8, rh_top=0.2,
                      z_rh_top=10000.0, rh_bubble=0.1, theta_noise=("normal", 0.05, 2025), uvw_noise=(0.01, 2026),
                      passive=2)
    opts = dict(dt=1.0, solver_opt=2, n_sound=6, turb_opt=1, steps=300)
    return case, opts


def config2():
    """128x128x64, 200 m, compressible; dt 2 s and n_sound 10 (the configuration states no dt)."""
    case = BubbleCase(128, 128, 64, 200.0, 200.0, 200.0, amplitude=2.0, rh_surface=0.8, rh_top=0.2,
                      rh_bubble=0.1, theta_noise=("normal", 0.1, 2025))
    opts = dict(dt=2.0, solver_opt=2, n_sound=10, turb_opt=1, steps=20, checkpoint_every=5)
    return case, opts


def rslf(p, t):
    """Bolton (1980) saturation mixing ratio (setup_lex.py:159-167)."""
    esl = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))
    esl = np.where(esl > p * 0.5, p * 0.5, esl)
    return eps * esl / (p - esl)


def base_state_gpu(nz, dz, theta0, rh_col, device="cuda", tol=1e-9, max_iter=200):
    """FP64 hydrostatic base state on the GPU -> dict of float64 numpy columns (nz+2)."""
    rh = torch.as_tensor(np.asarray(rh_col, dtype=np.float64), device=device).contiguous()
    out = torch.empty(5 * (nz + 2) + 1, dtype=torch.float64, device=device)
    it = ctypes.c_int(0)
    _lib.call("lex_base_state", int(nz), float(dz), float(theta0), rh.data_ptr(), out.data_ptr(), int(max_iter),
              float(tol), ctypes.byref(it), _lib.stream())
    cols = out[: 5 * (nz + 2)].view(5, nz + 2).cpu().numpy()
    return dict(pi0=cols[0], rho0_theta0=cols[1], rho0=cols[2], theta0=cols[3], qv0=cols[4], iters=it.value)


def _pad_np(a, face_axis=None):
    def pad_axis(x, ax, face):
        n = x.shape[ax]
        if face:
            lo = np.take(x, range(n - 4, n - 1), axis=ax)
            hi = np.take(x, range(1, 4), axis=ax)
        else:
            lo = np.take(x, range(n - 3, n), axis=ax)
            hi = np.take(x, range(0, 3), axis=ax)
        return np.concatenate([lo, x, hi], ax)
    a = pad_axis(a, 0, face_axis == 0)
    a = pad_axis(a, 1, face_axis == 1)
    return np.concatenate([a[:, :, :1], a, a[:, :, -1:]], 2)


def build_case(case: BubbleCase, device="cuda"):
    """(phys_state, base_state, grids, profiles) for a warm-bubble case.

    phys_state = (theta, u, v, w, pip, qv, *passive) float32 device tensors,
    base_state = (rho0_theta0, rho0, theta0, pi0, qv0, surface_t) (setup_lex.py:66-68).
    """
    nx, ny, nz = case.nx, case.ny, case.nz
    z = (np.arange(nz + 2) - 0.5) * case.dz

    def rh_of_z(zz):
        zc = np.minimum(np.maximum(zz, 0.0), case.z_rh_top)
        return case.rh_surface + (case.rh_top - case.rh_surface) * zc / case.z_rh_top

    bs = base_state_gpu(nz, case.dz, 300.0, rh_of_z(z), device=device)
    shp = (nx + 6, ny + 6, nz + 2)
    x = (np.arange(nx + 6) - 2.5) * case.dx
    y = (np.arange(ny + 6) - 2.5) * case.dy
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    xc = case.xc if case.xc is not None else 0.5 * nx * case.dx
    yc = case.yc if case.yc is not None else 0.5 * ny * case.dy
    r = np.sqrt(((X - xc) / case.xr) ** 2 + ((Y - yc) / case.yr) ** 2 + ((Z - case.zc) / case.zr) ** 2)
    shape = np.where(r > 1.0, 0.0, np.cos(r * np.pi / 2.0) ** 2)
    del X, Y, r
    theta_p = case.amplitude * shape
    kind, scale, seed = case.theta_noise
    if kind != "none" and scale != 0.0:
        rng = np.random.default_rng(seed)
        nz_ = rng.uniform(-scale, scale, (nx, ny, nz)) if kind == "uniform" else rng.normal(0.0, scale, (nx, ny, nz))
        theta_p = theta_p + _pad_np(nz_)
    th0 = np.broadcast_to(bs["theta0"], shp)
    qv0 = np.broadcast_to(bs["qv0"], shp)
    rho0 = np.broadcast_to(bs["rho0"], shp)
    theta = th0 + theta_p
    rh = rh_of_z(Z) + case.rh_bubble * shape
    del Z, shape
    pi = (rho0 * theta * (1.0 + reps * qv0) / (1.0 + qv0) * Rd / p00) ** (Rd / Cv)
    qv = rslf(pi ** (Cp / Rd) * p00, theta * pi) * rh
    del pi, rh
    u = np.zeros((nx + 7, ny + 6, nz + 2))
    v = np.zeros((nx + 6, ny + 7, nz + 2))
    w = np.zeros((nx + 6, ny + 6, nz + 3))
    sig, s0 = case.uvw_noise
    if sig != 0.0:
        ui = np.random.default_rng(s0).normal(0.0, sig, (nx, ny, nz))
        u = _pad_np(np.concatenate([ui, ui[:1]], 0), face_axis=0)
        vi = np.random.default_rng(s0 + 1).normal(0.0, sig, (nx, ny, nz))
        v = _pad_np(np.concatenate([vi, vi[:, :1]], 1), face_axis=1)
        wi = np.random.default_rng(s0 + 2).normal(0.0, sig, (nx, ny, nz + 1))
        w = _pad_np(wi)
        w[:, :, 0:2] = 0.0
        w[:, :, -2:] = 0.0
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device)  # noqa: E731
    phys = [dev(theta), dev(u), dev(v), dev(w), torch.zeros(shp, dtype=torch.float32, device=device), dev(qv)]
    phys += [torch.zeros(shp, dtype=torch.float32, device=device) for _ in range(case.passive)]
    colt = lambda c: torch.from_numpy(np.ascontiguousarray(c, np.float32)).to(device).view(1, 1, -1).expand(shp).contiguous()  # noqa: E731
    base = (colt(bs["rho0_theta0"]), colt(bs["rho0"]), colt(bs["theta0"]), colt(bs["pi0"]), colt(bs["qv0"]),
            torch.full((nx, ny), 300.0, device=device))
    grids = make_grids(nx, ny, nz, case.dx, case.dy, case.dz)
    return tuple(phys), base, grids, bs


def make_grids(nx, ny, nz, dx, dy, dz):
    """Coordinate arrays as read-only broadcast views (setup_lex.py:12-29, :149-156)."""
    x1 = np.linspace(-(3 - 0.5) * dx, (nx + 3 - 0.5) * dx, nx + 6)
    y1 = np.linspace(-(3 - 0.5) * dy, (ny + 3 - 0.5) * dy, ny + 6)
    z1 = np.linspace(-(1 - 0.5) * dz, (nz + 1 - 0.5) * dz, nz + 2)
    x1u = np.linspace(-3 * dx, (nx + 3) * dx, nx + 7)
    y1v = np.linspace(-3 * dy, (ny + 3) * dy, ny + 7)
    z1w = np.linspace(-1 * dz, (nz + 1) * dz, nz + 3)
    b = np.broadcast_to
    x3d = b(x1[:, None, None], (nx + 6, ny + 6, nz + 2))
    y3d = b(y1[None, :, None], (nx + 6, ny + 6, nz + 2))
    z3d = b(z1[None, None, :], (nx + 6, ny + 6, nz + 2))
    x3d4u = b(x1u[:, None, None], (nx + 7, ny + 6, nz + 2))
    y3d4v = b(y1v[None, :, None], (nx + 6, ny + 7, nz + 2))
    z3d4w = b(z1w[None, None, :], (nx + 6, ny + 6, nz + 3))
    cc1 = np.full((nx, ny, nz), 0.5)
    cc2 = 1.0 - cc1
    tauh = np.zeros((1, 1, nz + 2))
    tauf = np.zeros((1, 1, nz + 3))
    return (x3d, y3d, z3d, x3d4u, y3d4v, z3d4w, cc1, cc2, tauh, tauf)


def setup_grid_n_ic(ic_option, device="cuda"):
    """Reference entry point (setup_lex.py:9-71): grid and warm bubble from the namelist."""
    if ic_option != 1:
        raise RuntimeError("Undefine I.C. option! Exiting ...")
    case = BubbleCase(nl.nx, nl.ny, nl.nz, nl.dx, nl.dy, nl.dz, amplitude=1.0, xc=12000.0, yc=12000.0, zc=2000.0,
                      zr=2000.0, rh_surface=0.1, rh_top=0.1)
    phys, base, grids, _ = build_case(case, device)
    model_opt = (nl.solver_opt, nl.damp_opt, nl.rad_opt, nl.cor_opt, nl.sfc_opt, nl.pic_opt, nl.turb_opt)
    return phys, base, grids, model_opt
