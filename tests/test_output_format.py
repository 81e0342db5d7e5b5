"""Model output files in the reference's layout (write_netcdf.py:8-183): dimensions,
variables, attributes, fill value and contents read back from the written file."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from lex_amd import namelist_n_constants as nl
from lex_amd import write_netcdf as wn
from lex_amd.setup_lex import make_grids

NX, NY, NZ = 7, 5, 4

# (name, dims, long_name) of write_netcdf.py:19-81 (state) and :131-175 (base state)
STATE_VARS = [
    ("x", ("x",), "x (m)"), ("y", ("y",), "y (m)"), ("z", ("z",), "z (m)"),
    ("x4u", ("x4u",), "x (m) for u points"), ("y4v", ("y4v",), "y (m) for v points"),
    ("z4w", ("z4w",), "z (m) for w points"), ("time", ("time",), "time (s)"),
    ("theta", ("time", "x", "y", "z"), "theta for the current time"),
    ("pip", ("time", "x", "y", "z"), "pip for the current time"),
    ("qv", ("time", "x", "y", "z"), "qv for the current time"),
    ("u", ("time", "x4u", "y", "z"), "u for the current time"),
    ("v", ("time", "x", "y4v", "z"), "v for the current time"),
    ("w", ("time", "x", "y", "z4w"), "w for the current time"),
    ("info", ("time",), "exit code of the Poisson equation solver"),
]
BASE_VARS = [
    ("rho0_theta0", "rho0*theta0"), ("theta0", "theta0"), ("rho0", "rho0"), ("qv0", "qv0"), ("pi0", "pi0"),
]


def _state(rng):
    shp = (NX + 6, NY + 6, NZ + 2)
    th = rng.normal(300, 1, shp).astype(np.float32)
    u = rng.normal(0, 1, (NX + 7, NY + 6, NZ + 2)).astype(np.float32)
    v = rng.normal(0, 1, (NX + 6, NY + 7, NZ + 2)).astype(np.float32)
    w = rng.normal(0, 1, (NX + 6, NY + 6, NZ + 3)).astype(np.float32)
    pip = rng.normal(0, 1e-3, shp).astype(np.float32)
    qv = rng.uniform(0, 1e-2, shp).astype(np.float32)
    return th, u, v, w, pip, qv


def test_state_file_layout_and_contents(tmp_path):
    rng = np.random.default_rng(1)
    st = _state(rng)
    grids = make_grids(NX, NY, NZ, 100.0, 120.0, 80.0)
    sfc = (np.float32(3.0), -999.9) + (np.zeros((NX, NY), np.float32),) * 7
    fn = wn.save2nc(st, sfc, grids, 12, 360.0, filename=str(tmp_path / "lex_out_0012.nc"))
    with netcdf_file(fn, "r", mmap=False) as f:
        dims = dict(f.dimensions)
        assert dims == {"x": NX, "y": NY, "z": NZ, "x4u": NX + 1, "y4v": NY + 1, "z4w": NZ + 1, "time": None}
        assert set(f.variables) == {n for n, _, _ in STATE_VARS}
        for name, vd, ln in STATE_VARS:
            var = f.variables[name]
            assert var.dimensions == vd, name
            assert var.long_name.decode() == ln, name
            assert var.data.dtype == np.dtype(">f4"), name
        for name in ("theta", "pip", "qv", "u", "v", "w", "info"):
            assert float(f.variables[name]._FillValue) == np.float32(1.0e36), name
        th, u, v, w, pip, qv = st
        sl = (slice(3, -3), slice(3, -3), slice(1, -1))
        for name, a in (("theta", th), ("pip", pip), ("qv", qv), ("u", u), ("v", v), ("w", w)):
            got = f.variables[name].data
            assert got.shape == (1,) + a[sl].shape, name
            assert np.array_equal(got[0], a[sl]), name
        assert np.array_equal(f.variables["x"].data, grids[0][3:-3, 0, 0].astype(np.float32))
        assert np.array_equal(f.variables["z4w"].data, grids[5][0, 0, 1:-1].astype(np.float32))
        assert float(f.variables["time"].data[0]) == 360.0
        assert float(f.variables["info"].data[0]) == 3.0


def test_surface_and_correction_variables_follow_the_switches(tmp_path, monkeypatch):
    rng = np.random.default_rng(2)
    st = _state(rng)
    grids = make_grids(NX, NY, NZ, 100.0, 100.0, 100.0)
    two = [rng.normal(size=(NX, NY)).astype(np.float32) for _ in range(7)]
    sfc = (np.float32(0.0), np.float32(0.25)) + tuple(two)
    monkeypatch.setattr(nl, "sfc_opt", True)
    monkeypatch.setattr(nl, "pic_opt", True)
    fn = wn.save2nc(st, sfc, grids, 0, 0.0, filename=str(tmp_path / "o.nc"))
    with netcdf_file(fn, "r", mmap=False) as f:
        for name, a in zip(("tau_x", "tau_y", "sen", "evap", "T_ref", "q_ref", "u10"), two):
            assert f.variables[name].dimensions == ("time", "x", "y")
            assert np.array_equal(f.variables[name].data[0], a), name
        assert float(f.variables["pip_const"].data[0]) == np.float32(0.25)


def test_base_state_file(tmp_path):
    rng = np.random.default_rng(3)
    shp = (NX + 6, NY + 6, NZ + 2)
    base = tuple(rng.normal(1, 0.1, shp) for _ in range(5)) + (rng.normal(300, 1, (NX, NY)),)
    grids = make_grids(NX, NY, NZ, 100.0, 100.0, 100.0)
    fn = wn.save2nc_base(base, grids, filename=str(tmp_path / "base.nc"))
    with netcdf_file(fn, "r", mmap=False) as f:
        assert dict(f.dimensions) == {"x": NX, "y": NY, "z": NZ, "time": None}
        sl = (slice(3, -3), slice(3, -3), slice(1, -1))
        for (name, ln), a in zip(BASE_VARS, (base[0], base[2], base[1], base[4], base[3])):
            var = f.variables[name]
            assert var.long_name.decode() == ln
            assert var.data.dtype == np.dtype(">f8")
            assert np.array_equal(var.data[0], a[sl]), name
        assert np.array_equal(f.variables["T_sfc"].data[0], base[5])


def test_default_file_names_follow_the_namelist():
    assert nl.file_name_format % 7 == "experiments/lex_out_0007.nc"
    assert nl.base_file_name == "experiments/lex_reference_state.nc"
