"""Training the learned SGS net end to end through unrolls of the solver.

Workload (the differentiable-training configuration): the SGS net (simpleCNN shape,
dl_models.py:19-30; 11 inputs, 5 outputs, BF16 MFMA kernels) supplies the SGS tendencies
of every step of a 10-step unroll of the compressible core at 200 m (256 x 128 x 96,
dt 2 s), frozen over each step as the stage-0 Smagorinsky tendencies are.  The loss is
the squared difference, after the unroll, between the coarse state and the truth: a
100 m run (512 x 256 x 192, dt 1 s) of the same seeded bubble block-mean coarsened
2 x 2 x 2.  A batch holds 4 seeded bubbles (amplitude 1-3 K, radius 1.5-2.5 km, centre
jittered +-2 km).

The reference trains its net through its solver with jax.value_and_grad of the loss of
ws_rk3_sprint_dl_train (lex_dl_train.py:111-131, one_sprint_functions.py:118-165).  Here
the gradient is assembled from hand-written pieces, per step and in reverse:
  lex_rk3_step_vjp_taped_ext_sgs  state adjoint -> adjoints of the input state and of the
                                  frozen SGS tendencies
  lex_sgs_apply_vjp               -> adjoint of the net output
  the net's MFMA backward         -> weight gradients and the adjoint of its inputs
  lex_sgs_features_vjp            -> added to the input state's adjoint
with the state kept every `checkpoint_every` steps and each segment re-run with its
records (lex_rk3_step_tape_ext_sgs) before it is reversed.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import _lib
from . import adjoint as adj
from .one_sprint_functions import fused_step
from .setup_lex import build_case
from .sgs_net import SgsNet
from .sgs_train import MO_SMAG, CoarseModel, _grid, coarsen_state, member_case

FULL = dict(nx=256, ny=128, nz=96, dx=200.0, dt=2.0, n_sound=10, f=(2, 2, 2), fine_dt=1.0, fine_n_sound=7,
            nsteps=10, checkpoint_every=5, batch=4, iters=20, lr=1e-3, seed=2041)
SMALL = dict(nx=32, ny=24, nz=16, dx=200.0, dt=2.0, n_sound=10, f=(2, 2, 2), fine_dt=1.0, fine_n_sound=7,
             nsteps=4, checkpoint_every=2, batch=2, iters=3, lr=1e-3, seed=2041)
# loss weights per field of (theta, u, v, w, pip, qv): K, m/s, -, m/s, -, g/kg
LOSS_W = (1.0, 1.0, 1.0, 1.0, 0.0, 1.0e6)
# physical scale of the net's five outputs (u, v, w [m/s^2], theta [K/s], qv [1/s])
OUT_SCALE = (1.0e-3, 1.0e-3, 1.0e-3, 1.0e-3, 1.0e-6)


def _interior(a):
    return a[3:-3, 3:-3, 1:-1]


def field_loss(st, target):
    """L = 1/2 sum_f w_f ||x_f - target_f||^2 over interior points; returns (L, lam) with
    lam the padded adjoint of the state (zero on ghost lines)."""
    tot = torch.zeros((), dtype=torch.float64, device=st[0].device)
    lam = []
    for q, (a, b) in enumerate(zip(st, target)):
        lq = torch.zeros_like(a)
        wq = LOSS_W[q] if q < len(LOSS_W) else 0.0
        if wq != 0.0:
            d = _interior(a) - _interior(b)
            tot = tot + 0.5 * wq * (d.double() ** 2).sum()
            _interior(lq).copy_(wq * d)
        lam.append(lq)
    return tot, tuple(lam)


class NetSolverUnroll:
    """n-step unroll of the coarse core with the net's SGS inside every step, and its
    gradient with respect to the net's parameters."""

    def __init__(self, model: CoarseModel, net: SgsNet, checkpoint_every: int):
        self.m = model
        self.net = net
        self.k = int(checkpoint_every)
        self._tapes = []

    def _sgs(self, st, with_grad):
        """features -> net -> SGS tendencies; with_grad keeps the net's graph."""
        m = self.m
        g = m.g
        feats = m.features(st)
        if with_grad:
            feats.requires_grad_(True)
            y = self.net(feats)
        else:
            with torch.no_grad():
                y = self.net(feats)
        dev = st[0].device
        su = torch.empty((g.nx + 1, g.ny, g.nz), dtype=torch.float32, device=dev)
        sv = torch.empty((g.nx, g.ny + 1, g.nz), dtype=torch.float32, device=dev)
        sw = torch.empty((g.nx, g.ny, g.nz + 1), dtype=torch.float32, device=dev)
        sth = torch.empty((g.nx, g.ny, g.nz), dtype=torch.float32, device=dev)
        sqv = torch.empty_like(sth)
        yc = y.detach().contiguous()
        _lib.call("lex_sgs_apply", g, _lib.ptr(yc), m.out_scale, _lib.ptr(su), _lib.ptr(sv), _lib.ptr(sw),
                  _lib.ptr(sth), _lib.ptr(sqv), _lib.stream())
        return feats, y, (su, sv, sw, sth, sqv)

    def forward(self, st0, nsteps):
        """Returns the final state and the checkpoints {step: state}."""
        m = self.m
        st = tuple(a.clone() for a in st0)
        ckpts = {0: tuple(a.clone() for a in st)}
        for i in range(nsteps):
            m.step(st, "net", self.net)
            if (i + 1) % self.k == 0 and i + 1 < nsteps:
                ckpts[i + 1] = tuple(a.clone() for a in st)
        return st, ckpts

    def backward(self, ckpts, nsteps, lam_final):
        """Accumulates d loss / d params into the net's .grad; returns the adjoint of the
        initial state."""
        m = self.m
        lam = lam_final
        for s in sorted(ckpts.keys(), reverse=True):
            end = min(s + self.k, nsteps)
            recs = []
            st = ckpts[s]
            for t in range(end - s):
                feats, y, sgs = self._sgs(st, with_grad=True)
                tp = self._tapes[t] if t < len(self._tapes) else None
                tp, nxt = adj.step_tape_ext_sgs(st, sgs, m.base, m.g, m.dtc, m.n_sound, tape=tp)
                if t >= len(self._tapes):
                    self._tapes.append(tp)
                recs.append((st, feats, y))
                st = nxt
            for t in reversed(range(end - s)):
                st_t, feats, y = recs[t]
                lam, lsgs = adj.step_vjp_taped_ext_sgs(st_t, self._tapes[t], lam, m.base, m.g, m.dtc, m.n_sound)
                ly = adj.sgs_apply_vjp(m.g, lsgs, m.out_scale)
                y.backward(ly)
                adj.sgs_features_vjp(m.g, feats.grad, m.scale, lam)
                recs[t] = None
        return lam


def truth_final(cfg, seed, gc):
    """Coarsened 100 m truth after nsteps coarse steps, and the coarsened initial state."""
    fx, fy, fz = cfg["f"]
    case = member_case(seed, cfg["nx"] * fx, cfg["ny"] * fy, cfg["nz"] * fz, cfg["dx"] / fx, cfg["dx"] / fz)
    phys, base, _, _ = build_case(case)
    gf = _grid(case)
    st = tuple(a.clone() for a in phys)
    c0 = coarsen_state(st, gf, cfg["f"], gc)
    nfine = int(round(cfg["dt"] / cfg["fine_dt"]))
    for _ in range(cfg["nsteps"] * nfine):
        fused_step(st, base, gf, MO_SMAG, dt=cfg["fine_dt"], n_sound=cfg["fine_n_sound"])
    c1 = coarsen_state(st, gf, cfg["f"], gc)
    del phys, base, st
    return c0, c1


def fit_features(model, states):
    feats = torch.stack([model.features(s) for s in states])
    mu = feats.mean(dim=(0, 1, 2, 3))
    sd = feats.std(dim=(0, 1, 2, 3)).clamp_min(1e-12)
    for c in range(11):
        model.shift[c] = float(mu[c])
        model.scale[c] = float(1.0 / sd[c])
    for c in range(5):
        model.out_scale[c] = OUT_SCALE[c]


def run(cfg, log=print, device="cuda"):
    t_start = time.time()
    case_c = member_case(cfg["seed"], cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dx"])
    model = CoarseModel(case_c, cfg["dt"], cfg["n_sound"], device=device)
    seeds = [cfg["seed"] + 10 * b for b in range(cfg["batch"])]
    held = cfg["seed"] + 10 * cfg["batch"]
    data = []
    for sd in seeds + [held]:
        data.append(truth_final(cfg, sd, model.g))
        log(f"truth for seed {sd} ({cfg['nsteps']} coarse steps)")
    torch.cuda.synchronize()
    t_truth = time.time() - t_start
    fit_features(model, [c0 for c0, _ in data])
    net = SgsNet(seed=2031, device=device)
    with torch.no_grad():  # start from the no-SGS model: zero last layer
        net.weights[-1].zero_()
        net.biases[-1].zero_()
    un = NetSolverUnroll(model, net, cfg["checkpoint_every"])
    opt = torch.optim.Adam(net.parameters(), lr=cfg["lr"])

    def held_out_losses():
        c0, c1 = data[-1]
        out = {}
        for mode in ("none", "smag", "net"):
            st = tuple(a.clone() for a in c0)
            for _ in range(cfg["nsteps"]):
                model.step(st, mode, net)
            out[mode] = float(field_loss(st, c1)[0])
        return out

    before = held_out_losses()
    hist = []
    iter_times = []
    for it in range(cfg["iters"]):
        torch.cuda.synchronize()
        t0 = time.time()
        opt.zero_grad(set_to_none=True)
        tot = 0.0
        for c0, c1 in data[:-1]:
            st, ck = un.forward(c0, cfg["nsteps"])
            L, lam = field_loss(st, c1)
            un.backward(ck, cfg["nsteps"], lam)
            tot += float(L)
        for p in net.parameters():
            if p.grad is not None:
                p.grad /= len(data) - 1
        opt.step()
        torch.cuda.synchronize()
        iter_times.append(time.time() - t0)
        hist.append(tot / (len(data) - 1))
        log(f"iteration {it}: batch loss {hist[-1]:.6g} ({iter_times[-1]:.2f} s)")
    after = held_out_losses()
    return {"workload": f"net trained through {cfg['nsteps']}-step unrolls of the {cfg['nx']}x{cfg['ny']}x{cfg['nz']} "
                        f"compressible core at {cfg['dx']:.0f} m (dt {cfg['dt']} s, n_sound {cfg['n_sound']}), "
                        f"checkpoint every {cfg['checkpoint_every']}, batch {cfg['batch']}; truth: "
                        f"{cfg['f'][0]}x{cfg['f'][1]}x{cfg['f'][2]} coarsened 100 m runs",
            "batch_loss": hist, "held_out_before": before, "held_out_after": after,
            "s_per_iteration": float(np.median(iter_times[1:] if len(iter_times) > 1 else iter_times)),
            "truth_seconds": t_truth, "seconds": time.time() - t_start}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--iters", type=int, default=None)
    a = ap.parse_args(argv)
    cfg = dict(SMALL if a.small else FULL)
    if a.iters is not None:
        cfg["iters"] = a.iters
    res = run(cfg, log=lambda s: print(s, file=sys.stderr, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
