"""Shared builders and comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np

from oracle import lex_oracle as o


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))


def random_state(nx=16, ny=12, nz=12, dx=100.0, dy=120.0, dz=80.0, seed=0, moist=True, vel=3.0, passive=0):
    """A padded, halo-consistent state with both flow signs everywhere (float64)."""
    rng = np.random.default_rng(seed)
    gr = o.Grid(nx, ny, nz, dx, dy, dz)
    case = o.BubbleCase(gr, amplitude=2.0, xr=3 * dx, yr=3 * dy, zr=3 * dz, zc=nz * dz / 3,
                        rh_surface=0.8 if moist else 0.0, rh_top=0.2 if moist else 0.0, z_rh_top=nz * dz,
                        rh_bubble=0.1 if moist else 0.0, theta_noise=("normal", 0.3, seed + 1), passive=passive)
    phys, base, _ = o.setup_warm_bubble(case)
    th, _, _, _, _, qv = phys[:6]
    # halos as any stage leaves them (the set-up evaluates the bubble formula in the
    # ghosts, which is not the periodic image when the bubble reaches a side)
    th = o.padding3_array(o._interior(th))
    qv = o.padding3_array(o._interior(qv))
    u = o.padding3u_array(np.concatenate([ui := rng.normal(0, vel, (nx, ny, nz)), ui[:1]], 0))
    v = o.padding3v_array(np.concatenate([vi := rng.normal(0, vel, (nx, ny, nz)), vi[:, :1]], 1))
    w = o.set_w_bc(o.padding3_array(rng.normal(0, vel, (nx, ny, nz + 1))))
    pip = o.padding3_array(rng.normal(0, 1e-4, (nx, ny, nz)))
    extra = tuple(o.padding3_array(np.abs(rng.normal(1e-4, 5e-5, (nx, ny, nz)))) for _ in range(passive))
    return gr, (th, u, v, w, pip, qv) + extra, base


def f32(a):
    return np.asarray(a, np.float32)


def f64(a):
    return np.asarray(a, np.float64)


def to_dev(a, device="cuda"):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)


def from_dev(t):
    return t.detach().double().cpu().numpy()


def check_fp32(gpu, ref64, ref32, name, factor=8.0, floor=1e-6):
    """GPU error vs FP64 must be within `factor` x the FP32 oracle's own error (or `floor`)."""
    e_gpu = rel_l2(gpu, ref64)
    e_32 = rel_l2(ref32, ref64)
    bound = max(factor * e_32, floor)
    assert e_gpu <= bound, f"{name}: rel-L2 vs FP64 {e_gpu:.3e} > bound {bound:.3e} (FP32 oracle {e_32:.3e})"
    return e_gpu, e_32
