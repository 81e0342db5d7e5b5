"""lex_bicgstab_noalloc (static device scratch, no allocation) against lex_bicgstab
(caller workspace): the same kernels on the same launch shape, so the iterate and the
iteration count must be identical, including on a second, smaller call that finds the
first call's contents in the scratch."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import f32, random_state, to_dev

pytestmark = pytest.mark.gpu


def _system(nx, ny, nz, seed):
    """rtt, rhs and the stage-start pi' of a pressure solve on a random moist state
    (pressure_equations.py:51-77, :175-195 through the oracle in FP32)."""
    gr, phys, base = random_state(nx, ny, nz, seed=seed)
    p32 = tuple(f32(a) for a in phys)
    b32 = tuple(f32(a) for a in base[:5])
    rng = np.random.default_rng(seed + 100)
    au = f32(rng.normal(0, 1e-3, (nx + 1, ny, nz)))
    av = f32(rng.normal(0, 1e-3, (nx, ny + 1, nz)))
    aw = f32(rng.normal(0, 1e-3, (nx, ny, nz + 1)))
    aw[:, :, 0] = 0.0
    aw[:, :, -1] = 0.0
    b = f32(o.calculate_buoyancy(b32[2], p32[0], b32[4], p32[5])[0])
    rtt = f32(o.calculate_rtt(b32[0], p32[0], p32[5], b))
    z_u = np.zeros((nx, ny + 1, nz), np.float32)
    z_v = np.zeros((nx + 1, ny, nz), np.float32)
    rhs = f32(o.rhs_of_pressure_equation(b32[0], au, av, aw, z_u, z_v, b, gr))
    x0 = f32(o._interior(p32[4]))
    return gr, to_dev(rtt), to_dev(rhs), to_dev(x0)


def _solve_both(gr, rtt, rhs, x0, maxiter=100):
    import torch
    from lex_amd import _lib
    g = _lib.grid(gr.nx, gr.ny, gr.nz, gr.dx, gr.dy, gr.dz)
    lib = _lib.load()
    nbytes = lib.lex_bicgstab_workspace_bytes(g)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=rtt.device)
    ws.fill_(0xA5)   # garbage the solver must not depend on
    x_ws = torch.empty_like(x0)
    k_ws = ctypes.c_int(-99)
    _lib.call("lex_bicgstab", g, _lib.ptr(rtt), _lib.ptr(rhs), _lib.ptr(x0), _lib.ptr(x_ws), 1e-4, 1e-8, maxiter,
              _lib.ptr(ws), nbytes, ctypes.byref(k_ws), _lib.stream())
    x_na = torch.empty_like(x0)
    k_na = ctypes.c_int(-99)
    _lib.call("lex_bicgstab_noalloc", g, _lib.ptr(rtt), _lib.ptr(rhs), _lib.ptr(x0), _lib.ptr(x_na), 1e-4, 1e-8,
              maxiter, ctypes.byref(k_na), _lib.stream())
    torch.cuda.synchronize()
    return x_ws, k_ws.value, x_na, k_na.value


def test_noalloc_equals_workspace_version_on_two_calls():
    import torch
    # first call: a larger ragged grid; second: a smaller one, which runs over the
    # first call's leftovers in the shared scratch
    for shape, seed in (((40, 33, 26), 11), ((16, 12, 12), 12)):
        gr, rtt, rhs, x0 = _system(*shape, seed)
        print(f"bicgstab noalloc vs workspace: {shape}", flush=True)
        x_ws, k_ws, x_na, k_na = _solve_both(gr, rtt, rhs, x0)
        assert k_ws == k_na, (shape, k_ws, k_na)
        assert k_ws != 0, "the solve must iterate for the comparison to mean anything"
        assert bool(torch.isfinite(x_ws).all())
        assert torch.equal(x_ws, x_na), (shape, float((x_ws - x_na).abs().max()))
        assert not torch.equal(x_ws, x0)


def test_noalloc_rejects_grids_beyond_its_scratch():
    import torch
    from lex_amd import _lib
    nmax = 1048576   # LEX_BICGSTAB_NOALLOC_MAX_CELLS
    g = _lib.grid(128, 128, 65, 100.0, 100.0, 100.0)   # 1,064,960 cells > nmax
    assert 128 * 128 * 65 > nmax
    t = torch.zeros(8, device="cuda")
    x = torch.zeros(8, device="cuda")
    rc = _lib.load().lex_bicgstab_noalloc(g, _lib.ptr(t), _lib.ptr(t), _lib.ptr(t), _lib.ptr(x), 1e-4, 1e-8, 10,
                                          None, _lib.stream())
    assert rc == -3
    assert b"LEX_BICGSTAB_NOALLOC_MAX_CELLS" in _lib.load().lex_last_error()
