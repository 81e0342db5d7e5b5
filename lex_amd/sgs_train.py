"""Learned SGS net trained on coarsened fine runs that the GPU solver generates.

Workload (the learned-SGS configuration): fine runs at 100 m (576 x 576 x 144, dt 1 s)
of seeded moist warm bubbles are block-mean coarsened 6 x 6 x 3 (k_coarsen; the
reference coarsens its truth the same way, initial_comparison/plot_theta_pert_coarsened.m:7-8)
to 96 x 96 x 48 at 600 m / 300 m.  For snapshot pairs (t, t + dt_c) the target is the
coarse model's missing tendency

    R = (C(truth(t + dt_c)) - M_0(C(truth(t)))) / dt_c

with M_0 one coarse WS-RK3 step without SGS, brought to cell centres.  The net
(simpleCNN shape, dl_models.py:19-30; 11 inputs, 5 outputs) is trained on BF16 MFMA
kernels; at inference its tendencies replace the stage-0 Smagorinsky tendencies and are
frozen over the step (lex_rk3_step_ext_sgs), or the net runs at the start of every
RK3 stage on that stage's state (lex_rk3_step_stage_sgs).  Coarse runs from C(truth(0))
of a held-out member with (a) no SGS, (b) Smagorinsky, (c) the frozen net and (d) the
per-stage net are scored by the RMS of theta' against the coarsened truth; Smagorinsky
stays the fallback when the net is worse.  The full workload scores two held-out members
over 175 coarse steps (700 s) each: the 100 m truth of this recipe (neutral base state,
rigid lid, 1-3 K bubbles) diverges near 750 s, once the updraft passes ~43 m/s, with
either acoustic scheme (the vertically implicit one lasts to ~1250 s).

Every array operation on fields runs in liblexhip; torch holds tensors, runs the
optimiser and computes scalar statistics.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import sys
import time

import numpy as np
import torch

from . import _lib
from ._core import column
from .one_sprint_functions import fused_step
from .setup_lex import BubbleCase, build_case
from .sgs_net import SgsNet

FULL = dict(nx=576, ny=576, nz=144, dx=100.0, dz=100.0, dt=1.0, n_sound=7, f=(6, 6, 3), dtc=4.0, n_sound_c=10,
            members=8, snaps=40, gap=11, eval_steps=175, eval_members=2, epochs=12)
SMALL = dict(nx=72, ny=72, nz=36, dx=100.0, dz=100.0, dt=1.0, n_sound=7, f=(6, 6, 3), dtc=4.0, n_sound_c=10,
             members=2, snaps=4, gap=3, eval_steps=8, eval_members=1, epochs=2)
MO_NONE = (2, False, False, 0, False, False, 0)
MO_SMAG = (2, False, False, 0, False, False, 1)


def member_case(seed, nx, ny, nz, dx, dz):
    """Seeded bubble: amplitude 1-3 K, horizontal radius 1.5-2.5 km, centre jittered +-2 km."""
    rng = np.random.default_rng(seed)
    amp = float(rng.uniform(1.0, 3.0))
    rad = float(rng.uniform(1500.0, 2500.0))
    jx, jy = (float(a) for a in rng.uniform(-2000.0, 2000.0, 2))
    lx, ly = nx * dx, ny * dx
    return BubbleCase(nx, ny, nz, dx, dx, dz, amplitude=amp, xr=rad, yr=rad, zr=1400.0, zc=1400.0,
                      xc=0.5 * lx + jx, yc=0.5 * ly + jy, rh_surface=0.8, rh_top=0.2, z_rh_top=10000.0,
                      rh_bubble=0.1, theta_noise=("normal", 0.05, seed + 1), uvw_noise=(0.01, seed + 2))


def _grid(case):
    return _lib.grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)


def _pad(mode, interior, shape):
    ni, nj, nk = interior.shape
    out = torch.empty(shape, dtype=torch.float32, device=interior.device)
    _lib.call("lex_pad", mode, ni, nj, nk, 3, _lib.ptr(interior.contiguous()), _lib.ptr(out), _lib.stream())
    return out


def coarsen_state(phys, gf, f, gc):
    """Block-mean coarsening of (theta, u, v, w, pip, qv) to a padded coarse state."""
    fx, fy, fz = f
    nx, ny, nz = gc.nx, gc.ny, gc.nz
    dev = phys[0].device
    shapes_int = {0: (nx, ny, nz), 1: (nx + 1, ny, nz), 2: (nx, ny + 1, nz), 3: (nx, ny, nz + 1)}
    out = []
    for q, a in enumerate(phys[:6]):
        mode = {1: 1, 2: 2, 3: 3}.get(q, 0)
        ci = torch.empty(shapes_int[mode], dtype=torch.float32, device=dev)
        _lib.call("lex_coarsen", gf, mode, fx, fy, fz, _lib.ptr(a), _lib.ptr(ci), _lib.stream())
        if mode == 1:
            out.append(_pad(1, ci, (nx + 7, ny + 6, nz + 2)))
        elif mode == 2:
            out.append(_pad(2, ci, (nx + 6, ny + 7, nz + 2)))
        elif mode == 3:
            w = _pad(0, ci, (nx + 6, ny + 6, nz + 3))
            _lib.call("lex_set_w_bc", gc, _lib.ptr(w), _lib.stream())
            out.append(w)
        else:
            out.append(_pad(0, ci, (nx + 6, ny + 6, nz + 2)))
    return tuple(out)


def stage_sgs_step(g, opts, base_cols, state, fn, ws):
    """One step through lex_rk3_step_stage_sgs: fn(stage, stage_ptrs, su, sv, sw, ss_ptrs)
    enqueues the SGS tendencies of each stage on the current stream from the stage's
    input state (device pointers in state order) into the library's slots.  An exception
    raised by fn aborts the step and is re-raised here."""
    err = []

    def tramp(stage, sp, su, sv, sw, ss, _user):
        try:
            fn(stage, sp, su, sv, sw, ss)
            return 0
        except BaseException as e:   # noqa: BLE001 - handed back to the caller below
            err.append(e)
            return 1

    cb = _lib.STAGE_SGS_FN(tramp)
    ptrs = (ctypes.c_void_p * len(state))(*[_lib.ptr(a) for a in state])
    try:
        _lib.call("lex_rk3_step_stage_sgs", g, opts, base_cols, ptrs, cb, None, _lib.ptr(ws), ws.numel(),
                  _lib.stream())
    except _lib.LexError:
        if err:
            raise err[0]
        raise


class CoarseModel:
    """Coarse 96 x 96 x 48 runs with no SGS, Smagorinsky or the net."""

    def __init__(self, case_c, dtc, n_sound_c, device="cuda"):
        _, base, _, _ = build_case(case_c, device=device)
        self.base = base
        self.case = case_c
        self.g = _grid(case_c)
        self.dtc, self.n_sound = dtc, n_sound_c
        self.th0c = column(base[2])
        self.qv0c = column(base[4])
        self.shift = (ctypes.c_float * 11)(*([0.0] * 11))
        self.scale = (ctypes.c_float * 11)(*([1.0] * 11))
        self.out_scale = (ctypes.c_float * 5)(*([1.0] * 5))
        self._ws = None
        self._feat = None

    def step(self, st, mode, net=None):
        """One coarse step: "none", "smag" (stage-0 Smagorinsky, frozen), "net" (the net's
        tendencies from the step's initial state, frozen) or "net_stage" (the net run at the
        start of every stage on that stage's input state)."""
        if mode in ("none", "smag"):
            fused_step(st, self.base, self.g, MO_NONE if mode == "none" else MO_SMAG, dt=self.dtc,
                       n_sound=self.n_sound)
            return st
        cols = [column(a) for a in (self.base[1], self.base[2], self.base[4], self.base[3], self.base[0])]
        bc = _lib.BaseCols(*[c.data_ptr() for c in cols])
        o = _lib.StepOpts(float(self.dtc), int(self.n_sound), 2, 0.18, 1.0 / 3.0, 0, 0)
        if self._ws is None:
            nb = _lib.load().lex_rk3_workspace_bytes(self.g, 0)
            self._ws = torch.empty(nb, dtype=torch.uint8, device=st[0].device)
        if mode == "net_stage":
            stage_sgs_step(self.g, o, bc, st, lambda s, sp, su, sv, sw, ss: self.net_sgs_into(sp, net, su, sv, sw, ss),
                           self._ws)
            return st
        if mode != "net":
            raise ValueError(f"unknown SGS mode {mode!r}")
        sgs = self.net_sgs(st, net)
        ptrs = (ctypes.c_void_p * 6)(*[_lib.ptr(a) for a in st])
        ss = (ctypes.c_void_p * 2)(_lib.ptr(sgs[3]), _lib.ptr(sgs[4]))
        _lib.call("lex_rk3_step_ext_sgs", self.g, o, bc, ptrs, _lib.ptr(sgs[0]), _lib.ptr(sgs[1]), _lib.ptr(sgs[2]),
                  ss, _lib.ptr(self._ws), self._ws.numel(), _lib.stream())
        return st

    def features(self, st):
        g = self.g
        out = torch.empty((g.nx, g.ny, g.nz, 11), dtype=torch.float32, device=st[0].device)
        _lib.call("lex_sgs_features", g, _lib.ptr(self.th0c), _lib.ptr(self.qv0c), _lib.ptr(st[1]), _lib.ptr(st[2]),
                  _lib.ptr(st[3]), _lib.ptr(st[0]), _lib.ptr(st[5]), self.shift, self.scale, _lib.ptr(out),
                  _lib.stream())
        return out

    def net_sgs_into(self, sp, net, su, sv, sw, ss):
        """The net's SGS tendencies from the state at device pointers sp (state order) into
        the device buffers su, sv, sw, ss[0] (theta), ss[1] (qv); work on the current stream."""
        g = self.g
        if self._feat is None:
            self._feat = torch.empty((g.nx, g.ny, g.nz, 11), dtype=torch.float32, device=self.base[0].device)
        _lib.call("lex_sgs_features", g, _lib.ptr(self.th0c), _lib.ptr(self.qv0c), sp[1], sp[2], sp[3], sp[0], sp[5],
                  self.shift, self.scale, _lib.ptr(self._feat), _lib.stream())
        with torch.no_grad():
            y = net(self._feat).contiguous()
        _lib.call("lex_sgs_apply", g, _lib.ptr(y), self.out_scale, su, sv, sw, ss[0], ss[1], _lib.stream())

    def net_sgs(self, st, net):
        g = self.g
        dev = st[0].device
        with torch.no_grad():
            y = net(self.features(st)).contiguous()
        su = torch.empty((g.nx + 1, g.ny, g.nz), dtype=torch.float32, device=dev)
        sv = torch.empty((g.nx, g.ny + 1, g.nz), dtype=torch.float32, device=dev)
        sw = torch.empty((g.nx, g.ny, g.nz + 1), dtype=torch.float32, device=dev)
        sth = torch.empty((g.nx, g.ny, g.nz), dtype=torch.float32, device=dev)
        sqv = torch.empty_like(sth)
        _lib.call("lex_sgs_apply", g, _lib.ptr(y), self.out_scale, _lib.ptr(su), _lib.ptr(sv), _lib.ptr(sw),
                  _lib.ptr(sth), _lib.ptr(sqv), _lib.stream())
        return su, sv, sw, sth, sqv

    def target(self, c0, c1):
        """Cell-centred (nx, ny, nz, 5) missing tendency of one no-SGS coarse step."""
        st = tuple(a.clone() for a in c0)
        self.step(st, "none")
        r = [(b - a) / self.dtc for a, b in zip(st, c1)]
        u = r[1][3:-3, 3:-3, 1:-1]
        v = r[2][3:-3, 3:-3, 1:-1]
        w = r[3][3:-3, 3:-3, 1:-1]
        uc = 0.5 * (u[:-1] + u[1:])
        vc = 0.5 * (v[:, :-1] + v[:, 1:])
        wc = 0.5 * (w[:, :, :-1] + w[:, :, 1:])
        th = r[0][3:-3, 3:-3, 1:-1]
        qv = r[5][3:-3, 3:-3, 1:-1]
        return torch.stack([uc, vc, wc, th, qv], dim=-1).contiguous()

    def theta_p(self, st):
        return st[0][3:-3, 3:-3, 1:-1] - self.base[2][3:-3, 3:-3, 1:-1]


def generate_pairs(cfg, seeds, model, log):
    """Fine runs -> coarse (C(t), C(t + dt_c)) pairs; returns a list of pairs."""
    pairs = []
    nfine_c = int(round(cfg["dtc"] / cfg["dt"]))
    for m, seed in enumerate(seeds):
        case = member_case(seed, cfg["nx"], cfg["nx"], cfg["nz"], cfg["dx"], cfg["dz"])
        phys, base, _, _ = build_case(case)
        gf = _grid(case)
        st = tuple(a.clone() for a in phys)
        t0 = time.time()
        for s in range(cfg["snaps"]):
            for _ in range(cfg["gap"]):
                fused_step(st, base, gf, MO_SMAG, dt=cfg["dt"], n_sound=cfg["n_sound"])
            c0 = coarsen_state(st, gf, cfg["f"], model.g)
            for _ in range(nfine_c):
                fused_step(st, base, gf, MO_SMAG, dt=cfg["dt"], n_sound=cfg["n_sound"])
            c1 = coarsen_state(st, gf, cfg["f"], model.g)
            pairs.append((c0, c1))
        torch.cuda.synchronize()
        log(f"member {m} (seed {seed}): {cfg['snaps']} pairs, {time.time() - t0:.1f} s, "
            f"finite={bool(torch.isfinite(st[0]).all())}")
        del phys, base, st
    return pairs


def truth_trajectory(cfg, seed, model, log):
    """Coarsened truth every coarse step for the held-out member."""
    nfine_c = int(round(cfg["dtc"] / cfg["dt"]))
    case = member_case(seed, cfg["nx"], cfg["nx"], cfg["nz"], cfg["dx"], cfg["dz"])
    phys, base, _, _ = build_case(case)
    gf = _grid(case)
    st = tuple(a.clone() for a in phys)
    traj = [coarsen_state(st, gf, cfg["f"], model.g)]
    for _ in range(cfg["eval_steps"]):
        for _ in range(nfine_c):
            fused_step(st, base, gf, MO_SMAG, dt=cfg["dt"], n_sound=cfg["n_sound"])
        traj.append(tuple(a.clone() for a in coarsen_state(st, gf, cfg["f"], model.g)))
    log(f"held-out truth (seed {seed}): {cfg['eval_steps']} coarse steps")
    return traj


def fit_normalisation(model, pairs, targets):
    feats = torch.stack([model.features(c0) for c0, _ in pairs[:: max(1, len(pairs) // 16)]])
    mu = feats.mean(dim=(0, 1, 2, 3))
    sd = feats.std(dim=(0, 1, 2, 3)).clamp_min(1e-12)
    for c in range(11):
        model.shift[c] = float(mu[c])
        model.scale[c] = float(1.0 / sd[c])
    tsd = torch.stack(targets[:: max(1, len(targets) // 16)]).std(dim=(0, 1, 2, 3)).clamp_min(1e-12)
    for c in range(5):
        model.out_scale[c] = float(tsd[c])
    return [float(a) for a in tsd]


def train(net, model, pairs, targets, epochs, lr, log, seed=2031):
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    osc = torch.tensor([model.out_scale[c] for c in range(5)], device=targets[0].device)
    feats = [model.features(c0) for c0, _ in pairs]
    tnorm = [t / osc for t in targets]
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        tot = 0.0
        for n in order:
            opt.zero_grad(set_to_none=True)
            y = net(feats[n])
            loss = ((y - tnorm[n]) ** 2).mean()
            loss.backward()
            opt.step()
            tot += float(loss.detach())
        hist.append(tot / len(pairs))
        log(f"epoch {ep}: mse {hist[-1]:.4f}")
    return hist


def evaluate(model, net, truth):
    out = {}
    for mode in ("none", "smag", "net", "net_stage"):
        st = tuple(a.clone() for a in truth[0])
        errs = []
        for n in range(1, len(truth)):
            model.step(st, mode, net)
            d = model.theta_p(st) - model.theta_p(truth[n])
            errs.append(float(torch.sqrt((d * d).mean())))
        out[mode] = errs
    return out


def run(cfg, log=print, device="cuda"):
    t_start = time.time()
    nxc, nzc = cfg["nx"] // cfg["f"][0], cfg["nz"] // cfg["f"][2]
    case_c = member_case(0, nxc, nxc, nzc, cfg["dx"] * cfg["f"][0], cfg["dz"] * cfg["f"][2])
    model = CoarseModel(case_c, cfg["dtc"], cfg["n_sound_c"], device=device)
    seeds = [2031 + m for m in range(cfg["members"])]
    pairs = generate_pairs(cfg, seeds, model, log)
    t_gen = time.time() - t_start
    targets = [model.target(c0, c1) for c0, c1 in pairs]
    tsd = fit_normalisation(model, pairs, targets)
    net = SgsNet(seed=2031, device=device)
    t0 = time.time()
    hist = train(net, model, pairs, targets, cfg["epochs"], 1e-3, log)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    # held-out members (seeds after the training ones), each scored from its coarsened
    # initial state over eval_steps coarse steps
    per_member, curves = [], {}
    for e in range(cfg.get("eval_members", 1)):
        seed = 2031 + cfg["members"] + e
        truth = truth_trajectory(cfg, seed, model, log)
        scores = evaluate(model, net, truth)
        del truth
        per_member.append(dict(seed=seed, theta_rmse_final={k: v[-1] for k, v in scores.items()},
                               theta_rmse_mean={k: float(np.mean(v)) for k, v in scores.items()}))
        curves[seed] = {k: v[:: max(1, len(v) // 20)] for k, v in scores.items()}
        log(f"held-out seed {seed}: mean theta' RMSE " +
            ", ".join(f"{k} {v:.5f}" for k, v in per_member[-1]["theta_rmse_mean"].items()))
    modes = list(per_member[0]["theta_rmse_mean"])
    fin = {k: float(np.mean([p["theta_rmse_final"][k] for p in per_member])) for k in modes}
    mean = {k: float(np.mean([p["theta_rmse_mean"][k] for p in per_member])) for k in modes}
    nets = [m for m in ("net", "net_stage") if math.isfinite(mean[m]) and mean[m] <= mean["smag"]]
    chosen = min(nets, key=lambda m: mean[m]) if nets else "smagorinsky"
    return dict(fine=[cfg["nx"], cfg["nx"], cfg["nz"]], coarse=[nxc, nxc, nzc], pairs=len(pairs),
                target_std=tsd, train_mse=hist, eval_seconds=cfg["eval_steps"] * cfg["dtc"],
                theta_rmse_final=fin, theta_rmse_mean=mean, per_member=per_member, chosen=chosen,
                seconds=dict(generate=t_gen, train=t_train, total=time.time() - t_start),
                rmse_curves=curves)


def main(argv=None):
    ap = argparse.ArgumentParser(description="train and score the learned SGS net on coarsened fine runs")
    ap.add_argument("--small", action="store_true", help="72x72x36 fine grid (smoke size)")
    ap.add_argument("--members", type=int)
    ap.add_argument("--snaps", type=int)
    ap.add_argument("--epochs", type=int)
    ap.add_argument("--eval-steps", type=int)
    ap.add_argument("--eval-members", type=int)
    a = ap.parse_args(argv)
    cfg = dict(SMALL if a.small else FULL)
    for k in ("members", "snaps", "epochs"):
        if getattr(a, k) is not None:
            cfg[k] = getattr(a, k)
    if a.eval_steps is not None:
        cfg["eval_steps"] = a.eval_steps
    if a.eval_members is not None:
        cfg["eval_members"] = a.eval_members
    res = run(cfg, log=lambda s: print(s, file=sys.stderr, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
