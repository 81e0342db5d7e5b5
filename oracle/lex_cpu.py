"""ctypes binding of the C++/OpenMP restatement of the forward WS-RK3 step
(oracle/lex_cpu_step.cpp), the all-core CPU baseline.  Test infrastructure: only tests/
and the cpu_baseline leg of bench.py use it."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(_HERE, "lex_cpu_step.cpp")
LIB = os.path.join(_HERE, "_build", "liblexcpu.so")
# -ffp-contract=off keeps the oracle's rounding (no fused multiply-adds); AVX2 is the
# common x86-64 server baseline
CXXFLAGS = ["-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off", "-mavx2"]


def build() -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    r = subprocess.run(["g++", *CXXFLAGS, "-o", tmp, SRC], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build it with oracle.lex_cpu.build()")
        lib = C.CDLL(LIB)
        for name, ct in (("lex_cpu_rk3_steps_f32", C.c_float), ("lex_cpu_rk3_steps_f64", C.c_double)):
            fn = getattr(lib, name)
            P = C.POINTER(ct)
            fn.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int, C.c_double, C.c_double, C.c_int,
                                                             P, P, P, P, C.POINTER(C.c_void_p), C.c_int]
            fn.restype = C.c_int
        lib.lex_cpu_threads.restype = C.c_int
        _lib = lib
    return _lib


def threads() -> int:
    return int(load().lex_cpu_threads())


def rk3_steps(state, base_cols, grid, dt, n_sound, nsteps, dtype=np.float32, cs=0.18, prandtl=1.0 / 3.0):
    """nsteps WS-RK3 steps (solver_opt 2, Smagorinsky) of the padded state tuple
    (theta, u, v, w, pip, qv, *passive); base_cols = (rho0, theta0, qv0, pi0) columns of
    length nz + 2.  Returns new arrays of `dtype`."""
    lib = load()
    arrs = [np.array(a, dtype=dtype, order="C", copy=True) for a in state]   # stepped in place
    cols = [np.ascontiguousarray(a, dtype) for a in base_cols]
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib.lex_cpu_rk3_steps_f32 if dtype == np.float32 else lib.lex_cpu_rk3_steps_f64
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    cp = [c.ctypes.data_as(C.POINTER(ct)) for c in cols]
    rc = fn(grid.nx, grid.ny, grid.nz, grid.dx, grid.dy, grid.dz, dt, n_sound, cs, prandtl, len(arrs) - 6,
            *cp, ptrs, nsteps)
    if rc != 0:
        raise ValueError(f"lex_cpu_rk3_steps failed ({rc})")
    return tuple(arrs)
