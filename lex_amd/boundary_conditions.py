"""Boundary conditions on the hot path (boundary_conditions.py:104-109)."""
from __future__ import annotations

import torch

from . import _lib
from ._core import grid_from


def set_w_bc(w: torch.Tensor) -> torch.Tensor:
    """Zero w on padded levels 0, 1 and nz+1, nz+2 (boundary_conditions.py:104-109).

    Returns a new tensor, as the reference does.
    """
    out = w.clone()
    nx, ny, nz = w.shape[0] - 6, w.shape[1] - 6, w.shape[2] - 3
    g = _lib.grid(nx, ny, nz, 1.0, 1.0, 1.0)
    _lib.call("lex_set_w_bc", g, _lib.ptr(out), _lib.stream())
    return out


__all__ = ["set_w_bc", "grid_from"]
