#!/usr/bin/env python3
"""Throughput of LEX's WS-RK3 dynamics step on one MI355X.

Workload: the fully compressible moist warm bubble at 256x256x128 cells (dx = dy = dz =
100 m, dt = 1 s, WS-RK3, WENO5 advection, Smagorinsky Cs = 0.18, scalars theta, qv and
passive qc, qr), FP32.  The specified n_sound = 6 (1 + 3 + 6 acoustic substeps) is
acoustically unstable on this grid, so the timed run uses n_sound = 7 (1 + 4 + 7) and the
n_sound = 6 step time is reported beside it.  A "step" is one full WS-RK3 step of the
whole grid; inputs are resident in HBM.  Also measured: the same workload with the
vertically implicit acoustic substeps (solver_opt 3) at the specified n_sound = 6, which is
stable, and the forward and forward+adjoint step times of the differentiable 128x128x64
workload (20 steps, checkpoint every 5).

Prints one JSON line: cell-steps/s from an uninstrumented timed loop; the dominant
kernel's roofline (algorithmic bytes per launch / its average launch time from HIP
events around each launch in a separate profiled loop, against the 8 TB/s HBM peak); the
per-variable error of a GPU run against the FP64 and FP32 NumPy oracle on a sub-grid of
the same recipe; the CPU baseline: a few steps of the full workload in the C++/OpenMP
FP32 restatement of the step on every host core the OpenMP runtime is given (with the
GPU-vs-CPU error of those steps at full size, and the 1-core NumPy oracle's time on the
sub-grid beside it); and the pseudo-incompressible 64x64x40 recipe (BiCGSTAB iterations
per solve and the BiCGSTAB iteration's roofline).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s

# algorithmic bytes per interior cell for each kernel kind: 4 B x distinct FP32 3-D
# fields read or written at the boundary of the reference routine(s) it implements
BYTES_PER_CELL = {
    "k_acoustic": 52,        # u,v,w,pi',rhs x3, theta_rho, pi* in; u,v,w,pi' out
    "k_acoustic2": 104,      # two substeps per launch: 2 x 52
    "k_scalar_update": 32,   # s, s0, u, v, w, flow_div, sgs in; s_next out
    "k_rhs_uv": 32,          # u, v, w, flow_div, sgs_u, sgs_v in; rhs_u, rhs_v out
    "k_rhs_w": 32,           # u, v, w, theta, qv, flow_div, sgs_w in; rhs_w out
    "k_div_rho": 16,
    "k_theta_rho": 12,
    "k_smag_deform": 40,
    "k_smag_nm": 12,
    "k_smag_k": 16,
    "k_smag_uvw": 40,
    "k_smag_scalar": 12,
    "k_copy": 8,
    # fused stage kernels (flow divergence computed inline, never stored)
    "k_stage_scalars": 80,   # per scalar s, s0, sgs in + s_next out (x4: theta, qv, qc, qr); u, v, w in; theta_rho out
    "k_mom_u": 20,           # u, v, w, sgs_u in; rhs_u out
    "k_mom_v": 20,           # u, v, w, sgs_v in; rhs_v out
    "k_mom_w": 28,           # u, v, w, theta, qv, sgs_w in; rhs_w out
    "k_smag_fused": 56,      # u, v, w, theta, qv, qc, qr in; sgs_u, sgs_v, sgs_w and 4 scalar tendencies out
}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rk3_step_256x256x128_hbm_traffic.json")


def measured_traffic(kernel, nx, ny, nz, n_sound):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    measurement (scaled by a copy of known size) when it was taken on this workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if (nx, ny, nz, n_sound) != (256, 256, 128, 7) or kernel not in t.get("kernels", {}):
        return None
    return t["kernels"][kernel]["traffic_bytes"]


ERR_NAMES = ["theta", "u", "v", "w", "pip", "qv", "qc", "qr"]


def _field_errors(got, ref, init, demean_pip=False):
    """Per variable rel-L2 and L-inf of the perturbation (field minus the initial state;
    pi' with its mean removed for the pseudo-incompressible core) against `ref`."""
    import numpy as np
    out = {}
    for n, a, b, a0 in zip(ERR_NAMES, got, ref, init):
        da = np.asarray(a, np.float64) - a0
        db = np.asarray(b, np.float64) - a0
        if n == "pip" and demean_pip:
            da, db = da - da.mean(), db - db.mean()
        nb = float(np.linalg.norm(db))
        if nb == 0.0:
            out[n] = {"rel_l2": 0.0 if float(np.abs(da).max()) == 0.0 else float("inf"), "linf": float(np.abs(da).max()),
                      "note": "field stays zero"}
            continue
        out[n] = {"rel_l2": float(np.linalg.norm(da - db) / nb), "linf": float(np.abs(da - db).max())}
    return out


def errors_and_cpu_baseline(steps=3, n_sound=7, nx=64, ny=64, nz=128, time_cpu=True):
    """The same recipe on a nx x ny x nz sub-grid: `steps` fused GPU steps against the
    FP64 and the FP32 NumPy oracle from the same initial state.  The FP32 oracle run is
    also the CPU baseline (one core, NumPy elementwise is single-threaded)."""
    import numpy as np
    import torch
    from lex_amd import one_sprint_functions as sp
    from lex_amd._core import grid_from
    from lex_amd.setup_lex import build_case, config1
    from oracle import lex_oracle as o
    case, opts = config1()
    case = case.scaled(nx, ny, nz)
    phys, base, grids, _ = build_case(case)
    g = grid_from(base[1], *grids[3:6])
    init = tuple(a.double().cpu().numpy() for a in phys)
    b = tuple(a.double().cpu().numpy() for a in base[:5])
    st = tuple(a.clone() for a in phys)
    mo = (2, False, False, 0, False, False, 1)
    for _ in range(steps):
        sp.fused_step(st, base, g, mo, n_sound=n_sound)
    torch.cuda.synchronize()
    got = tuple(a.double().cpu().numpy() for a in st)
    og = o.Grid(nx, ny, nz, case.dx, case.dy, case.dz)
    oo = o.Options(dt=opts["dt"], solver_opt=2, n_sound=n_sound, turb_opt=1)
    runs, secs = {}, {}
    for name, dt in (("fp32", np.float32), ("fp64", np.float64)):
        xi = tuple(a.astype(dt) for a in init)
        bs = tuple(c.astype(dt) for c in b) + (None,)
        t0 = time.perf_counter()
        for _ in range(steps):
            xi, _ = o.ws_rk3_big_step(xi, bs, og, oo)
        secs[name] = time.perf_counter() - t0
        runs[name] = xi
    errs = {"workload": f"{steps} WS-RK3 steps of the same recipe (n_sound {n_sound}) on a {nx}x{ny}x{nz} sub-grid",
            "fields": "perturbation from the initial state; rel_l2 = |d - d_ref| / |d_ref|",
            "gpu_vs_fp64_oracle": _field_errors(got, runs["fp64"], init),
            "gpu_vs_fp32_oracle": _field_errors(got, runs["fp32"], init),
            "fp32_oracle_vs_fp64_oracle": _field_errors(runs["fp32"], runs["fp64"], init)}
    cpu = None
    if time_cpu:
        cpu = {"value": nx * ny * nz * steps / secs["fp32"], "unit": "cell-steps/s", "cores": 1, "kind": "port",
               "sample": f"{steps} WS-RK3 steps of the same recipe on a {nx}x{ny}x{nz} sub-grid "
                         f"(NumPy FP32 oracle, single-threaded), {secs['fp32']:.1f} s"}
    return errs, cpu


def _cpu_model():
    try:
        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10)
        for ln in r.stdout.splitlines():
            if ln.startswith("Model name:"):
                return ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _merge_cpu(port, numpy_1core):
    """The C++ all-core port is the CPU baseline; the 1-core NumPy oracle time rides along."""
    if port is None:
        return numpy_1core
    out = dict(port)
    if numpy_1core is not None:
        out["numpy_oracle_1_core"] = numpy_1core
    return out


def _merge_errors(errs, full):
    if errs is None:
        return full
    if full is not None:
        errs = dict(errs)
        errs["full_size"] = full
    return errs


def cpu_port_baseline(init, base, case, opts, n_sound, steps):
    """The CPU baseline: `steps` WS-RK3 steps of the full workload from its initial state
    on every host core the OpenMP runtime is given, in the C++ FP32 restatement of the
    step (oracle/lex_cpu_step.cpp, pinned to the NumPy oracle by the CPU tests); the same
    steps on the GPU from the same state give a full-size GPU-vs-CPU comparison."""
    import numpy as np
    import torch
    from lex_amd import _lib, one_sprint_functions as sp
    from oracle import lex_cpu, lex_oracle as o
    host0 = tuple(a.cpu().numpy() for a in init)
    cols = [np.ascontiguousarray(base[i][3, 3, :].cpu().numpy()) for i in (1, 2, 4, 3)]
    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    t0 = time.perf_counter()
    cpu = lex_cpu.rk3_steps(host0, cols, og, opts["dt"], n_sound, steps, np.float32)
    secs = time.perf_counter() - t0
    gg = _lib.grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    st = tuple(a.clone() for a in init)
    for _ in range(steps):
        sp.fused_step(st, base, gg, (2, False, False, 0, False, False, 1), n_sound=n_sound)
    torch.cuda.synchronize()
    gpu = tuple(a.double().cpu().numpy() for a in st)
    ncell = case.nx * case.ny * case.nz
    base_line = {"value": ncell * steps / secs, "unit": "cell-steps/s", "cores": lex_cpu.threads(), "kind": "port",
                 "sample": f"{steps} WS-RK3 steps of the full {case.nx}x{case.ny}x{case.nz} workload (n_sound "
                           f"{n_sound}, theta, qv, qc, qr, Smagorinsky), C++/OpenMP FP32 restatement of the step "
                           f"(oracle/lex_cpu_step.cpp), {secs:.1f} s",
                 "nproc": os.cpu_count(), "cpu_model": _cpu_model()}
    errs = _field_errors(gpu, tuple(np.asarray(a, np.float64) for a in cpu), tuple(np.asarray(a, np.float64) for a in host0))
    return base_line, {"workload": f"{steps} WS-RK3 steps of the full {case.nx}x{case.ny}x{case.nz} workload",
                       "gpu_vs_cpu_port_fp32": errs}


def c0_timing(steps=10, warmup=2):
    """The pseudo-incompressible recipe at its full 64x64x40 (dx = dy = 600 m, dz = 300 m,
    dt = 30 s, BiCGSTAB pressure solve, tol 1e-4, maxiter 100): step time, iterations of
    every solve, and the BiCGSTAB iteration's roofline (88 B/cell per iteration: two
    matvecs plus the s, x, r, p updates, dots fused) from the per-kernel HIP events."""
    import torch
    from lex_amd import _lib, namelist_n_constants as nl, one_sprint_functions as sp
    from lex_amd import one_step_integration as one
    from lex_amd.setup_lex import build_case, config0
    case, opts = config0()
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], opts["n_sound"], 1, 1
    mo = (1, False, False, 0, False, False, 1)
    ks = []
    solve = one.solve_pres_eqn

    def rec(*a, **k):
        x, info = solve(*a, **k)
        ks.append(info)
        return x, info
    one.solve_pres_eqn = rec
    try:
        st = phys
        for _ in range(warmup):
            st, _ = sp.ws_rk3_big_step(st, base, grids, mo)
        torch.cuda.synchronize()
        ks.clear()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            st, _ = sp.ws_rk3_big_step(st, base, grids, mo)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        iters = list(ks)
        # a profiled repetition for the per-kernel times of the solver
        ks.clear()
        _lib.profile_enable(True, 16384)
        for _ in range(2):
            st, _ = sp.ws_rk3_big_step(st, base, grids, mo)
        prof = _lib.profile_read()
        _lib.profile_enable(False)
        piters = sum(max(k, 0) for k in ks)
    finally:
        one.solve_pres_eqn = solve
    ncell = case.nx * case.ny * case.nz
    it_ms = sum(prof.get(k, (0.0, 0))[0] for k in ("k_bicg_pq", "k_bicg_st", "k_bicg_xr"))
    launches = prof.get("k_bicg_pq", (0.0, 0))[1]
    gbs = 88.0 * ncell * piters / (it_ms / 1e3) / 1e9 if it_ms > 0 else None
    return {"workload": f"pseudo-incompressible dry warm bubble {case.nx}x{case.ny}x{case.nz}, dx {case.dx:.0f} m, "
                        f"dz {case.dz:.0f} m, dt {opts['dt']:.0f} s, WS-RK3, BiCGSTAB (tol 1e-4, maxiter 100)",
            "steps": steps, "warmup": warmup, "ms_per_step": ms / steps,
            "value": ncell * steps / (ms / 1e3), "unit": "cell-steps/s",
            "bicgstab_iterations": iters, "solves": len(iters),
            "bicgstab_exit_codes_note": "k = iterations done; every solve stops at maxiter 100 because the affine "
                                        "pressure system is not consistent to the tolerance (as in the reference)",
            "bicgstab": {"bound": "hbm", "bytes_per_cell_per_iteration": 88, "iterations_profiled": piters,
                         "iteration_launches_profiled": launches, "iteration_ms": it_ms / piters if piters else None,
                         "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS if gbs else None},
            "kernels_profiled": {k: {"avg_ms": round(v[0] / v[1], 5), "launches": v[1]} for k, v in prof.items()},
            "state_finite": all(bool(torch.isfinite(a).all()) for a in st)}


def implicit_solver(init, base, g, ncell, n_sound, steps=20, warmup=2):
    """The same workload with solver_opt 3 (vertically implicit acoustic substeps) at the
    specified n_sound: the implicit vertical coupling removes the vertical sound-wave limit
    that makes the explicit substeps unstable there."""
    import torch
    from lex_amd import _lib, one_sprint_functions as sp
    mo = (3, False, False, 0, False, False, 1)
    st = tuple(a.clone() for a in init)
    for _ in range(warmup):
        sp.fused_step(st, base, g, mo, n_sound=n_sound)
    torch.cuda.synchronize()
    _lib.profile_enable(True, 16384)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        sp.fused_step(st, base, g, mo, n_sound=n_sound)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    kvi = prof.get("k_acoustic_vi")
    out = {"solver_opt": 3, "n_sound": n_sound, "steps": steps, "warmup": warmup,
           "value": ncell * steps / (ms / 1e3), "unit": "cell-steps/s", "ms_per_step": ms / steps,
           "state_finite": all(bool(torch.isfinite(a).all()) for a in st)}
    if kvi:
        avg = kvi[0] / kvi[1]
        gbs = BYTES_PER_CELL["k_acoustic"] * ncell / (avg / 1e3) / 1e9
        out["k_acoustic_vi"] = {"avg_launch_ms": avg, "launches_per_step": kvi[1] / steps,
                                "share": kvi[0] / ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
                                "bytes_per_cell": BYTES_PER_CELL["k_acoustic"]}
    return out


def output_timing(reps=3):
    """save2nc of the 256x256x128 state (device tensors -> a netCDF file in the
    reference's layout): seconds per file and MB/s of field data."""
    import tempfile
    import torch
    from lex_amd import write_netcdf as wn
    from lex_amd.one_step_integration import _sfc_etc
    from lex_amd.setup_lex import build_case, config1
    case, _ = config1()
    phys, base, grids, _ = build_case(case)
    torch.cuda.synchronize()
    nbytes = sum((a.shape[0] - 6) * (a.shape[1] - 6) * (a.shape[2] - 2) * 4 for a in phys[:6])
    with tempfile.TemporaryDirectory() as td:
        times = []
        for r in range(reps):
            t0 = time.perf_counter()
            fn = wn.save2nc(phys, _sfc_etc(0, phys[0]), grids, r, float(r), filename=os.path.join(td, f"o{r}.nc"))
            times.append(time.perf_counter() - t0)
            fsize = os.path.getsize(fn)
    t = min(times)
    return {"workload": f"save2nc of the {case.nx}x{case.ny}x{case.nz} state (6 fields, device -> netCDF-3 file)",
            "s_per_file": t, "field_MB": nbytes / 1e6, "file_MB": fsize / 1e6, "MB_per_s": nbytes / 1e6 / t}


def dl_sprint_timing(sprint_n=10):
    """ws_rk3_sprint_dl on the 256x256x128 workload with simpleCNN (5 -> 32 -> 32 -> 32 -> 5,
    BF16 MFMA) correcting the state after every step, against the same steps without it."""
    import torch
    from lex_amd import _lib, dl_models as dlm, namelist_n_constants as nl, one_sprint_functions as sp
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt, nl.sprint_n = opts["dt"], 7, 2, 1, sprint_n
    mo = (2, False, False, 0, False, False, 1)
    g = _lib.grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    sc = []
    for a, q in zip((phys[0], phys[1], phys[2], phys[3], phys[5]), range(5)):
        inner = a[3:-3, 3:-3, 1:-1]
        spread = float(inner.std()) + 1e-3
        lo = (inner - spread).unsqueeze(0).unsqueeze(0).contiguous()
        hi = (inner + spread).unsqueeze(0).unsqueeze(0).contiguous()
        sc += [lo, hi]
    scaling = tuple(sc) + (None,) * 4
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_net = {}
    for name in ("simpleCNN", "Zcnn"):
        net = getattr(dlm, name)(seed=7)
        with torch.no_grad():
            net.weights[-1].mul_(1e-3)
        st = tuple(a.clone() for a in phys)
        sp.ws_rk3_sprint_dl(st, base, grids, mo, net, scaling)
        torch.cuda.synchronize()
        e0.record()
        out, _ = sp.ws_rk3_sprint_dl(st, base, grids, mo, net, scaling)
        e1.record()
        torch.cuda.synchronize()
        t_net[name] = (e0.elapsed_time(e1) / sprint_n, all(bool(torch.isfinite(a).all()) for a in out))
        del net, out, st
    t_dl = t_net["simpleCNN"][0]
    st2 = tuple(a.clone() for a in phys)
    e0.record()
    for _ in range(sprint_n):
        sp.fused_step(st2, base, g, mo)
    e1.record()
    torch.cuda.synchronize()
    t_plain = e0.elapsed_time(e1) / sprint_n
    ncell = case.nx * case.ny * case.nz
    return {"workload": f"ws_rk3_sprint_dl, {case.nx}x{case.ny}x{case.nz}, n_sound 7, sprint_n {sprint_n}, "
                        "post-step correction by simpleCNN and by Zcnn",
            "ms_per_step": t_dl, "ms_per_step_without_correction": t_plain, "correction_ms": t_dl - t_plain,
            "cell_steps_per_s": ncell / (t_dl / 1e3), "state_finite": t_net["simpleCNN"][1],
            "zcnn_ms_per_step": t_net["Zcnn"][0], "zcnn_correction_ms": t_net["Zcnn"][0] - t_plain,
            "zcnn_state_finite": t_net["Zcnn"][1], "loss_terms": dl_loss_timing(case, grids, sprint_n)}


def dl_loss_timing(case, grids, sprint_n):
    """The loss terms of the correction's training (lex_dl_train.py:62-97) after a
    sprint_n-step sprint of one sequence on the 256x256x128 grid: the two Laplacians of
    the theta and qv stacks and the eight weighted FP64 sums (stacks synthetic)."""
    import torch
    from lex_amd import _lib, dl_train as dt
    nx, ny, nz = case.nx, case.ny, case.nz
    gen = torch.Generator(device="cuda").manual_seed(3)
    shp = (1, sprint_n, nx, ny, nz)
    th = 300.0 + torch.randn(shp, device="cuda", generator=gen)
    qv = 1e-2 + 1e-3 * torch.randn(shp, device="cuda", generator=gen)
    wt = torch.rand((nx, ny, nz), device="cuda", generator=gen)
    x3d, y3d, z3d, x3d4u, y3d4v, z3d4w = grids[:6]
    ws = torch.empty(_lib.load().lex_weighted_sum_workspace_bytes(), dtype=torch.uint8, device="cuda")
    out = torch.empty(8, dtype=torch.float64, device="cuda")

    def once():
        lt = dt.laplace_of_tensor(x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, th)
        lq = dt.laplace_of_tensor(x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, qv)
        for q, (a, b) in enumerate(((th, th), (th, lt), (qv, qv), (qv, lq), (th, lt), (qv, lq), (lt, lq), (qv, qv))):
            _lib.call("lex_weighted_sum", a.numel(), 1 if q == 7 else 0, _lib.ptr(a), _lib.ptr(b), _lib.ptr(wt),
                      wt.numel(), _lib.ptr(out[q:q + 1]), _lib.ptr(ws), ws.numel(), _lib.stream())

    once()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        once()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    nbytes = (2 * 2 + 7 * 2 + 1) * th.numel() * 4   # Laplacians: read + write; sums: 2 reads (penalty 1)
    return {"ms_per_loss": ms, "fields": f"2 Laplacians + 8 sums over {sprint_n} x {nx}x{ny}x{nz}",
            "GBs": nbytes / (ms / 1e3) / 1e9}


def dl_train_timing(sprint_n=5, nx=128, ny=128, nz=64):
    """One train_step of the post-step correction (lex_dl_train.py:112-132) with simpleCNN
    on a 128x128x64 slice of the 256x256x128 workload recipe: a sprint_n-step corrected
    sprint recorded, the loss, and its gradient through the solver; against the corrected
    sprint alone (loss_and_aux)."""
    import numpy as np
    import torch
    from lex_amd import dl_models as dlm, dl_train as dt, namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    case = case.scaled(nx, ny, nz)
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    mo = (2, False, False, 0, False, False, 1)
    sc = []
    for a, q in zip((phys[0], phys[1], phys[2], phys[3], phys[5]), range(5)):
        inner = a[3:-3, 3:-3, 1:-1]
        spread = float(inner.std()) + 1e-3
        sc += [(inner - spread)[None, None].contiguous(), (inner + spread)[None, None].contiguous()]
    for a in (phys[0], phys[5]):
        inner = a[3:-3, 3:-3, 1:-1]
        sc += [(torch.zeros_like(inner) - 1e-4)[None, None], (torch.zeros_like(inner) + 1e-4)[None, None]]
    scaling = tuple(sc)
    net = dlm.simpleCNN(seed=7)
    with torch.no_grad():
        net.weights[-1].mul_(1e-3)
    ic = tuple(a.unsqueeze(0).clone() for a in phys[:6])
    st = dt.ws_rk3_sprint_dl_train_vmap(*ic, base, grids, mo, net, scaling, sprint_n=sprint_n)
    lap = lambda a: dt.laplace_of_tensor(*(grids[i] for i in (0, 3, 1, 4, 2, 5)), a)  # noqa: E731
    truth = tuple(a + 0.01 for a in st) + (lap(st[0] - scaling[0][0, 0]), lap(st[5] - scaling[8][0, 0]))
    opt = torch.optim.Adam(net.parameters(), lr=1e-6)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    dt.train_step((net, opt), ic, truth, scaling, base, grids, mo, sprint_n=sprint_n)
    torch.cuda.synchronize()
    e0.record()
    m = dt.train_step((net, opt), ic, truth, scaling, base, grids, mo, sprint_n=sprint_n)
    e1.record()
    torch.cuda.synchronize()
    t_train = e0.elapsed_time(e1)
    e0.record()
    loss = dt.loss_and_aux(net, ic, truth, scaling, base, grids, mo, sprint_n=sprint_n)
    e1.record()
    torch.cuda.synchronize()
    t_fwd = e0.elapsed_time(e1)
    return {"workload": f"train_step of the post-step correction (simpleCNN), {nx}x{ny}x{nz}, n_sound 7, "
                        f"sprint_n {sprint_n}, batch 1",
            "train_step_ms": t_train, "loss_only_ms": t_fwd, "ratio": t_train / t_fwd,
            "loss": float(loss.item()), "grad_norm": float(m["grad_norm"].item()),
            "finite": bool(np.isfinite(float(m["grad_norm"].item())))}


def physics_timing(steps=10):
    """The 256x256x128 step with the upper-level Rayleigh damping (layer above 9.6 km of
    the 12.8 km column), the Newtonian heating and the bulk surface fluxes (sea surface
    at the set-up's 300 K) switched on, and with the horizontal / vertical Smagorinsky
    closure (turb_opt 2), against the plain step (turb_opt 1); and the surface-flux
    kernel alone on the 256x256 columns."""
    import torch
    from lex_amd import _lib, boundary_conditions as bcm, namelist_n_constants as nl, one_sprint_functions as sp
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    g = _lib.grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    out = {}
    reps = 200
    bcm.surface_fluxes(phys, base, g)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        bcm.surface_fluxes(phys, base, g)
    e1.record()
    torch.cuda.synchronize()
    out["surface_fluxes_us_per_call"] = e0.elapsed_time(e1) / reps * 1e3
    for name, mo in (("plain", (2, False, False, 0, False, False, 1)), ("damp_rad", (2, True, True, 0, False, False, 1)),
                     ("sfc", (2, False, False, 0, True, False, 1)), ("damp_rad_sfc", (2, True, True, 0, True, False, 1)),
                     ("turb2", (2, False, False, 0, False, False, 2))):
        st = tuple(a.clone() for a in phys)
        sp.fused_step(st, base, g, mo)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            sp.fused_step(st, base, g, mo)
        e1.record()
        torch.cuda.synchronize()
        out[name + "_ms_per_step"] = e0.elapsed_time(e1) / steps
    out["workload"] = f"{case.nx}x{case.ny}x{case.nz}, n_sound 7, z_damping {nl.z_damping} m, rd_alpha {nl.rd_alpha:.5f}"
    return out


def copy_bandwidth(mib=1024, reps=5):
    """Device-to-device copy of `mib` MiB (read + write), the practical HBM ceiling."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * src.numel() * 4 * reps / (e0.elapsed_time(e1) / 1e3) / 1e9
    del src, dst
    return {"bytes_per_copy": 2 * (mib << 20), "GBs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS}


def fwd_adjoint(steps=20, checkpoint_every=5, nx=128, ny=128, nz=64):
    """Differentiable workload: 128x128x64 at 200 m, dt 2 s, n_sound 10, 20 steps,
    loss 1/2 ||theta'_20 - target||^2, gradients w.r.t. theta'_0 and Cs, checkpoint every 5
    steps.  Returns forward and forward+adjoint step times on the same state."""
    import numpy as np
    import torch
    from lex_amd import adjoint as adj, one_sprint_functions as sp
    from lex_amd._core import grid_from
    from lex_amd.setup_lex import build_case, config2
    torch.cuda.empty_cache()   # the earlier 256x256x128 workloads' blocks
    case, opts = config2()
    case = case.scaled(nx, ny, nz)
    phys, base, grids, _ = build_case(case)
    g = grid_from(base[1], *grids[3:6])
    dt, ns = opts["dt"], opts["n_sound"]
    mo = (2, False, False, 0, False, False, 1)
    th0 = base[2][3:-3, 3:-3, 1:-1]
    tp0 = (phys[0][3:-3, 3:-3, 1:-1] - th0).contiguous()
    target = torch.from_numpy(np.random.default_rng(2029).normal(0, 0.5, (nx, ny, nz)).astype(np.float32)).cuda()
    # forward only
    st = tuple(a.clone() for a in phys)
    sp.fused_step(st, base, g, mo, dt=dt, n_sound=ns)
    st = tuple(a.clone() for a in phys)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        sp.fused_step(st, base, g, mo, dt=dt, n_sound=ns)
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1)

    def fwd_bwd():
        tp = tp0.clone().requires_grad_(True)
        cs = torch.tensor(0.18, device="cuda", requires_grad=True)
        th = adj.sprint_theta(tp, cs, phys, base, g, dt, ns, steps, checkpoint_every)
        loss = 0.5 * ((th - target) ** 2).sum()
        loss.backward()
        return tp.grad, cs.grad, loss

    fwd_bwd()
    torch.cuda.synchronize()
    passes = []
    for _ in range(3):   # the median of three timed passes (a single launch-bound pass varies)
        e0.record()
        gt, gc, loss = fwd_bwd()
        e1.record()
        torch.cuda.synchronize()
        passes.append(e0.elapsed_time(e1))
    tot_ms = sorted(passes)[1]
    ncell = nx * ny * nz
    return {"workload": f"differentiable compressible bubble {nx}x{ny}x{nz}, dx 200 m, dt {dt} s, n_sound {ns}, "
                        f"{steps} steps, checkpoint every {checkpoint_every}",
            "fwd_ms_per_step": fwd_ms / steps, "fwd_adj_ms_per_step": tot_ms / steps,
            "ratio_fwd_adj_over_fwd": tot_ms / fwd_ms, "fwd_adj_pass_ms": passes,
            "fwd_cell_steps_per_s": ncell * steps / (fwd_ms / 1e3),
            "fwd_adj_cell_steps_per_s": ncell * steps / (tot_ms / 1e3),
            "grad_finite": bool(torch.isfinite(gt).all()) and bool(torch.isfinite(gc).all()),
            "loss": float(loss.detach()), "dL_dCs": float(gc)}


def sgs_net_timing(nx=96, ny=96, nz=48, reps=10):
    """The learned-SGS net (11 -> 32 -> 32 -> 32 -> 5, 3x3x3, BF16 MFMA) on the coarse grid:
    forward and forward+backward times and the MFMA throughput of the forward."""
    import torch
    from lex_amd.sgs_net import SgsNet
    net = SgsNet(seed=1)
    g = torch.Generator(device="cuda").manual_seed(2031)
    x = torch.randn(nx, ny, nz, 11, device="cuda", generator=g)
    t = torch.randn(nx, ny, nz, 5, device="cuda", generator=g)
    for _ in range(2):
        ((net(x) - t) ** 2).mean().backward()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    with torch.no_grad():
        for _ in range(reps):
            net(x)
    e1.record()
    torch.cuda.synchronize()
    fwd = e0.elapsed_time(e1) / reps
    e0.record()
    for _ in range(reps):
        ((net(x) - t) ** 2).mean().backward()
    e1.record()
    torch.cuda.synchronize()
    fb = e0.elapsed_time(e1) / reps
    flop = 2.0 * 27 * (11 * 32 + 32 * 32 + 32 * 32 + 32 * 5) * nx * ny * nz
    tf = flop / (fwd / 1e3) / 1e12
    # the 32 -> 32 hidden layer (conv + bias + GELU, BF16 in and out) alone
    from lex_amd import sgs_net as sn, _lib as lib
    ncell = nx * ny * nz
    xl = torch.randn(ncell, 32, device="cuda", generator=g).to(torch.bfloat16)
    wl = torch.randn(27, 32, 32, device="cuda", generator=g) * 0.1
    wp = sn._pack(wl, 0)
    bl = torch.zeros(32, device="cuda")
    yl = torch.empty(ncell, 32, dtype=torch.bfloat16, device="cuda")

    def layer():
        lib.call("lex_conv3d_bf16_fwd", nx, ny, nz, 32, 32, lib.ptr(xl), lib.ptr(wp), lib.ptr(bl), 1, 0,
                 lib.ptr(yl), None, None, lib.stream())
    for _ in range(5):
        layer()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(50):
        layer()
    e1.record()
    torch.cuda.synchronize()
    lus = e0.elapsed_time(e1) / 50 * 1e3
    ltf = 2.0 * 27 * 32 * 32 * ncell / (lus / 1e6) / 1e12
    return {"grid": [nx, ny, nz], "fwd_ms": fwd, "fwd_bwd_ms": fb,
            "roofline": {"bound": "mfma", "achieved": tf, "peak": 2500.0, "unit": "TFLOP/s", "frac": tf / 2500.0,
                         "note": "useful flops of the four convolutions (unpadded channels), BF16 dense peak"},
            "layer_32_32": {"us": lus, "achieved_TFLOPs": ltf, "frac": ltf / 2500.0,
                            "note": "one 3x3x3 convolution 32 -> 32 channels + bias + GELU, 50 calls"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-steps", type=int, default=3)
    ap.add_argument("--profile", action="store_true", help="also print per-kernel times to stderr")
    ap.add_argument("--n-sound", type=int, default=7,
                    help="acoustic substeps per full step (the specified 6 is acoustically unstable at "
                         "dx=100 m, dt=1 s; 7 is the smallest stable value)")
    ap.add_argument("--no-as-specified", action="store_true", help="skip the n_sound=6 timing")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint workload")
    ap.add_argument("--no-implicit", action="store_true",
                    help="skip the solver_opt 3 (vertically implicit) timing at the specified n_sound")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed region, copy a 256 MiB buffer 3 times (a kernel of known HBM "
                         "traffic, used to scale rocprof FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the error report against the FP64/FP32 oracle (and with it the CPU baseline)")
    ap.add_argument("--no-c0", action="store_true", help="skip the pseudo-incompressible 64x64x40 timing")
    ap.add_argument("--profile-steps", type=int, default=10,
                    help="steps of the separate profiled loop that feeds the per-kernel table and the roofline")
    ap.add_argument("--only", choices=("adjoint", "sgs-net", "output", "dl-sprint", "dl-train", "physics", "c0"),
                    default=None,
                    help="run only the forward+adjoint workload, the SGS-net timing, the output-file "
                         "timing, the post-step-correction sprint, its train_step or the physics-option steps and "
                         "print its "
                         "JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the domain-decomposed multi-GPU path is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    from lex_amd import _lib, namelist_n_constants as nl, one_sprint_functions as sp
    from lex_amd.setup_lex import build_case, config1

    _lib.load()
    if args.only is not None:
        fn = {"adjoint": fwd_adjoint, "sgs-net": sgs_net_timing, "output": output_timing,
              "dl-sprint": dl_sprint_timing, "dl-train": dl_train_timing, "physics": physics_timing,
              "c0": c0_timing}[args.only]
        print(json.dumps(fn()))
        return
    case, opts = config1()
    phys, base, grids, _ = build_case(case)
    n_sound = args.n_sound
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], n_sound, 2, 1
    model_opt = (2, False, False, 0, False, False, 1)
    g = _lib.grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    ncell = case.nx * case.ny * case.nz
    init = tuple(a.clone() for a in phys)

    as_specified = None
    if not args.no_as_specified and n_sound != opts["n_sound"]:
        # the specified n_sound=6: time its first 3 steps from the initial state (all
        # finite; the FP64 oracle diverges at step 5), 5 repetitions
        seg, reps, tot = 3, 5, 0.0
        st6 = tuple(a.clone() for a in init)
        sp.fused_step(st6, base, g, model_opt, n_sound=opts["n_sound"])
        for _ in range(reps):
            for a, b in zip(st6, init):
                a.copy_(b)
            torch.cuda.synchronize()
            a0 = torch.cuda.Event(enable_timing=True)
            a1 = torch.cuda.Event(enable_timing=True)
            a0.record()
            for _ in range(seg):
                sp.fused_step(st6, base, g, model_opt, n_sound=opts["n_sound"])
            a1.record()
            torch.cuda.synchronize()
            tot += a0.elapsed_time(a1)
        fin6 = all(bool(torch.isfinite(a).all()) for a in st6)
        as_specified = {"n_sound": opts["n_sound"], "acoustic_substeps_per_step": 10,
                        "value": ncell * seg * reps / (tot / 1e3), "ms_per_step": tot / (seg * reps),
                        "steps_timed": seg * reps, "state_finite": fin6,
                        "note": "first 3 steps from the initial state, repeated 5 times; n_sound=6 is "
                                "acoustically unstable at dx=100 m, dt=1 s (the FP64 oracle reaches |w| ~ 6e18 "
                                "at step 5), so longer runs time NaN arithmetic"}
        del st6

    state = tuple(phys)
    for _ in range(args.warmup):
        sp.fused_step(state, base, g, model_opt)
    torch.cuda.synchronize()

    # headline: the timed steps with no per-kernel instrumentation
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        sp.fused_step(state, base, g, model_opt)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    finite = all(bool(torch.isfinite(a).all()) for a in state)
    ms_step = ms / args.steps
    value = ncell * args.steps / (ms / 1e3)

    # a separate profiled loop (HIP events around every launch, on the launch stream)
    # feeds the per-kernel table and the dominant kernel's roofline
    psteps = max(1, args.profile_steps)
    _lib.profile_enable(True, 16384)
    p0 = torch.cuda.Event(enable_timing=True)
    p1 = torch.cuda.Event(enable_timing=True)
    p0.record()
    for _ in range(psteps):
        sp.fused_step(state, base, g, model_opt)
    p1.record()
    torch.cuda.synchronize()
    pms = p0.elapsed_time(p1)
    prof = _lib.profile_read()
    _lib.profile_enable(False)

    # dominant kernel = largest total time in the profiled loop
    dom, (dom_ms, dom_n) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_ms = dom_ms / dom_n
    bpc = BYTES_PER_CELL.get(dom, 0)
    traffic_alg = bpc * ncell
    achieved = traffic_alg / (avg_ms / 1e3) / 1e9
    kernels = {}
    for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        e = {"avg_ms": round(v[0] / v[1], 5), "launches_per_step": v[1] / psteps, "share": round(v[0] / pms, 4)}
        if k in BYTES_PER_CELL:   # algorithmic bytes of every kernel, not only the dominant one's
            gbs = BYTES_PER_CELL[k] * ncell / (v[0] / v[1] / 1e3) / 1e9
            e.update({"bytes_per_cell": BYTES_PER_CELL[k], "achieved_GBs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)})
        kernels[k] = e
    if args.profile:
        print(json.dumps(kernels, indent=1), file=sys.stderr)

    cpu_port, err_full = None, None
    if not args.no_cpu_baseline:
        cpu_port, err_full = cpu_port_baseline(init, base, case, opts, n_sound, args.cpu_sample_steps)
    if args.copy_probe:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    del state, phys
    impl = None if args.no_implicit else implicit_solver(init, base, g, ncell, opts["n_sound"])
    del init
    cb = copy_bandwidth()
    fa = None if args.no_adjoint else fwd_adjoint()
    sn = None if args.no_adjoint else sgs_net_timing()
    errs, cpu = (None, None) if args.no_errors else errors_and_cpu_baseline(args.cpu_sample_steps, n_sound,
                                                                            time_cpu=not args.no_cpu_baseline)
    c0 = None if args.no_c0 else c0_timing()
    line = {
        "metric": "cell-steps/s (fwd, fwd+adjoint) vs plain at stated error; % HBM roofline",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "fully compressible moist warm bubble, forward WS-RK3 step",
                   "nx": case.nx, "ny": case.ny, "nz": case.nz, "dx_m": case.dx, "dt_s": opts["dt"],
                   "n_sound": n_sound, "n_sound_specified": opts["n_sound"],
                   "acoustic_substeps_per_step": 1 + round((opts["dt"] / 2) / (opts["dt"] / n_sound)) + n_sound,
                   "scalars": ["theta", "qv", "qc", "qr"],
                   "advection": "WENO5 (WENO3/centred near lid and surface)", "sgs": "Smagorinsky Cs=0.18",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, case.nx, case.ny, case.nz, n_sound),
                     "traffic_source": "profiles/rk3_step_256x256x128_hbm_traffic.md (rocprofv3 --pmc FETCH_SIZE, "
                                       "WRITE_SIZE; copy-kernel scaled)",
                     "bytes_per_cell": bpc, "avg_launch_ms": avg_ms,
                     "avg_launch_source": f"HIP events around each launch on its stream, separate loop of {psteps} "
                                          "steps",
                     "copy_GBs": cb["GBs"], "frac_of_copy": achieved / cb["GBs"]},
        "cpu_baseline": _merge_cpu(cpu_port, cpu),
        "errors": _merge_errors(errs, err_full),
        "c0": c0,
        "as_specified": as_specified,
        "implicit_as_specified": impl,
        "fwd_adjoint": fa,
        "sgs_net": sn,
        "kernels": kernels,
        "profiled_ms_per_step": pms / psteps,
        "state_finite": finite,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
