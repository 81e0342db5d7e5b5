"""The C++/OpenMP restatement of the forward WS-RK3 step (oracle/lex_cpu_step.cpp, the
all-core CPU baseline) against the NumPy oracle it restates, on the CPU.

The FP64 instantiation keeps the oracle's evaluation order (built without contraction),
so it matches the FP64 oracle to rounding (rel-L2 of the step increments ~1e-15; theta's
~1e-13, its increment being small against 300 K); the FP32 one is another FP32
evaluation, within the FP32 oracle's own distance from the FP64 result.  Moist state
with two passive scalars, ragged grid, two steps at n_sound 7.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import random_state, rel_l2

NAMES = ["theta", "u", "v", "w", "pip", "qv", "qc", "qr"]


@pytest.fixture(scope="module")
def lex_cpu():
    from oracle import lex_cpu as lc
    if not os.path.exists(lc.LIB):
        lc.build()
    lc.load()
    return lc


def _case(shape=(17, 13, 11), seed=5):
    gr, phys, base = random_state(*shape, seed=seed, passive=2, vel=3.0)
    cols = [np.ascontiguousarray(base[i][3, 3, :]) for i in (1, 2, 4, 3)]   # rho0, theta0, qv0, pi0
    return gr, phys, base, cols


def _oracle(phys, base, gr, nsteps, dtype):
    oo = o.Options(dt=1.0, solver_opt=2, n_sound=7, turb_opt=1)
    xi = tuple(a.astype(dtype) for a in phys)
    bs = tuple(b.astype(dtype) for b in base[:5]) + (None,)
    for _ in range(nsteps):
        xi, _ = o.ws_rk3_big_step(xi, bs, gr, oo)
    return xi


def test_fp64_port_reproduces_the_fp64_oracle(lex_cpu):
    gr, phys, base, cols = _case()
    ref = _oracle(phys, base, gr, 2, np.float64)
    got = lex_cpu.rk3_steps(phys, cols, gr, 1.0, 7, 2, np.float64)
    for n, a, b, a0 in zip(NAMES, got, ref, phys):
        e = rel_l2(a - a0, b - a0)
        assert e <= (1e-11 if n == "theta" else 1e-13), (n, e)
    # the input is not modified
    assert all(np.array_equal(a, b) for a, b in zip(phys, _case()[1]))


def test_fp32_port_within_the_fp32_oracle_spread(lex_cpu):
    gr, phys, base, cols = _case(seed=6)
    r64 = _oracle(phys, base, gr, 2, np.float64)
    r32 = _oracle(phys, base, gr, 2, np.float32)
    got = lex_cpu.rk3_steps(phys, cols, gr, 1.0, 7, 2, np.float32)
    for n, a, b64, b32, a0 in zip(NAMES, got, r64, r32, phys):
        d = np.asarray(a, np.float64) - a0
        e, e32 = rel_l2(d, b64 - a0), rel_l2(np.asarray(b32, np.float64) - a0, b64 - a0)
        assert e <= max(3.0 * e32, 1e-6), (n, e, e32)


def test_port_rejects_bad_arguments(lex_cpu):
    gr, phys, base, cols = _case()
    with pytest.raises(ValueError):
        lex_cpu.rk3_steps(phys, cols, o.Grid(3, gr.ny, gr.nz, gr.dx, gr.dy, gr.dz), 1.0, 7, 1, np.float64)
