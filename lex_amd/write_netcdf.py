"""Model output in the reference's file layout (write_netcdf.py:8-183).

`save2nc` and `save2nc_base` keep the reference's names, arguments, dimensions
(x, y, z, x4u, y4v, z4w, unlimited time), variable names, long_name attributes, types and
the 1e36 fill value, so the reference's MATLAB analysis reads the files unchanged.  The
reference writes NETCDF4 with zlib level 2 through netCDF4-python; that library is not
available here, so the files are netCDF-3 (64-bit offset) written with
scipy.io.netcdf_file: the same dimensions and variables, uncompressed.  Device tensors are
copied to the host once per call (interior slices only).
"""
from __future__ import annotations

import os

import numpy as np
from scipy.io import netcdf_file

from . import namelist_n_constants as nl

FILL = 1.0e36

# the interior slices of the padded arrays (ghosts 3, 3, 1)
_IX = slice(3, -3)
_IZ = slice(1, -1)


def _host(a):
    """numpy view of a torch tensor (any device) or array."""
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _interior(a):
    return _host(a)[_IX, _IX, _IZ]


def _open(filename):
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    return netcdf_file(filename, mode="w", version=2)


def _var(f, name, dtype, dims, long_name, data, fill=True):
    v = f.createVariable(name, dtype, dims)
    if fill:
        v._FillValue = np.array(FILL, dtype=dtype)
    v.long_name = long_name
    v[:] = np.asarray(data, dtype=dtype)
    return v


def save2nc(phys_state, sfc_others, grids, sprint_i, model_time, filename=None):
    """Write the physical state of one output time (write_netcdf.py:8-119).

    phys_state: (theta, u, v, w, pip, qv[, ...]) in the padded layout; sfc_others: the
    reference's 9-tuple (info, pip_const, tau_x, tau_y, sen, evap, t_ref, q_ref, u10n);
    grids: the reference's 10-tuple.  The file name is nl.file_name_format % sprint_i
    unless `filename` is given.  Returns the file name.
    """
    if filename is None:
        filename = nl.file_name_format % sprint_i
    theta, u, v, w, pip, qv = phys_state[:6]
    info, pip_const, tau_x, tau_y, sen, evap, t_ref, q_ref, u10n = sfc_others
    x3d, y3d, z3d, x3d4u, y3d4v, z3d4w = grids[:6]
    th = _interior(theta)
    nx, ny, nz = th.shape
    f = _open(filename)
    try:
        f.createDimension("time", None)  # netCDF-3: the record dimension is defined first
        f.createDimension("x", nx)
        f.createDimension("y", ny)
        f.createDimension("z", nz)
        f.createDimension("x4u", nx + 1)
        f.createDimension("y4v", ny + 1)
        f.createDimension("z4w", nz + 1)
        f32 = np.float32
        _var(f, "x", f32, ("x",), "x (m)", _host(x3d)[_IX, 0, 0], fill=False)
        _var(f, "y", f32, ("y",), "y (m)", _host(y3d)[0, _IX, 0], fill=False)
        _var(f, "z", f32, ("z",), "z (m)", _host(z3d)[0, 0, _IZ], fill=False)
        _var(f, "x4u", f32, ("x4u",), "x (m) for u points", _host(x3d4u)[_IX, 0, 0], fill=False)
        _var(f, "y4v", f32, ("y4v",), "y (m) for v points", _host(y3d4v)[0, _IX, 0], fill=False)
        _var(f, "z4w", f32, ("z4w",), "z (m) for w points", _host(z3d4w)[0, 0, _IZ], fill=False)
        _var(f, "time", f32, ("time",), "time (s)", [model_time], fill=False)
        cell = ("time", "x", "y", "z")
        _var(f, "theta", f32, cell, "theta for the current time", th[None])
        _var(f, "pip", f32, cell, "pip for the current time", _interior(pip)[None])
        _var(f, "qv", f32, cell, "qv for the current time", _interior(qv)[None])
        _var(f, "u", f32, ("time", "x4u", "y", "z"), "u for the current time", _interior(u)[None])
        _var(f, "v", f32, ("time", "x", "y4v", "z"), "v for the current time", _interior(v)[None])
        _var(f, "w", f32, ("time", "x", "y", "z4w"), "w for the current time", _interior(w)[None])
        info_v = _host(info) if info is not None else np.float32(FILL)
        _var(f, "info", f32, ("time",), "exit code of the Poisson equation solver", np.reshape(info_v, (1,)))
        if nl.pic_opt:
            _var(f, "pip_const", f32, ("time",), "pi' correction constant for the current time",
                 np.reshape(_host(pip_const), (1,)))
        if nl.sfc_opt:
            two = ("time", "x", "y")
            for name, arr, ln in (("tau_x", tau_x, "tau_x for the current time"),
                                  ("tau_y", tau_y, "tau_y for the current time"),
                                  ("sen", sen, "sensible heat for the current time"),
                                  ("evap", evap, "evaporation for the current time"),
                                  ("T_ref", t_ref, "temperature at reference height (2m) for the current time"),
                                  ("q_ref", q_ref, "mixing ratio at reference height (2m) for the current time"),
                                  ("u10", u10n, "wind speed at reference height (10m) for the current time")):
                _var(f, name, f32, two, ln, np.reshape(_host(arr), (1, nx, ny)))
    finally:
        f.close()
    return filename


def save2nc_base(base_state, grids, filename=None):
    """Write the base state (write_netcdf.py:122-183), FP64 like the reference."""
    if filename is None:
        filename = nl.base_file_name
    rho0_theta0, rho0, theta0, pi0, qv0, surface_t = base_state[:6]
    x3d, y3d, z3d = grids[:3]
    th0 = _interior(theta0)
    nx, ny, nz = th0.shape
    f = _open(filename)
    try:
        f.createDimension("time", None)
        f.createDimension("x", nx)
        f.createDimension("y", ny)
        f.createDimension("z", nz)
        f64 = np.float64
        _var(f, "x", f64, ("x",), "x (m)", _host(x3d)[_IX, 0, 0], fill=False)
        _var(f, "y", f64, ("y",), "y (m)", _host(y3d)[0, _IX, 0], fill=False)
        _var(f, "z", f64, ("z",), "z (m)", _host(z3d)[0, 0, _IZ], fill=False)
        _var(f, "time", f64, ("time",), "time (s)", [0.0], fill=False)
        cell = ("time", "x", "y", "z")
        _var(f, "rho0_theta0", f64, cell, "rho0*theta0", _interior(rho0_theta0)[None])
        _var(f, "theta0", f64, cell, "theta0", th0[None])
        _var(f, "rho0", f64, cell, "rho0", _interior(rho0)[None])
        _var(f, "qv0", f64, cell, "qv0", _interior(qv0)[None])
        _var(f, "pi0", f64, cell, "pi0", _interior(pi0)[None])
        sfc = np.zeros((nx, ny)) if surface_t is None else _host(surface_t)
        _var(f, "T_sfc", f64, ("time", "x", "y"), "surface temperature", np.reshape(sfc, (1, nx, ny)))
    finally:
        f.close()
    return filename
