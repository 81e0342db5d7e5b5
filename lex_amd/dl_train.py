"""The loss side of the reference's training of the post-step correction, on the device.

Mirrors, with the reference's names and arguments:

* ``ws_rk3_sprint_dl_train`` (one_sprint_functions.py:34-97): the corrected sprint keeping
  each step's interior theta, u, v, w, pi', qv in stacks;
* ``laplace_of_tensor`` (dataset_generation.py:271-320) on ``lex_laplace``;
* ``loss_and_aux`` / ``eval_step`` (lex_dl_train.py:49-99, :135-139): the weighted l2
  terms and the negative-qv penalty as deterministic FP64 sums (``lex_weighted_sum``);
* ``generate_training_data`` and ``padding3_batch`` / ``padding3u_batch`` /
  ``padding3v_batch`` (dataset_generation.py:115-268): the shuffled 75 / 25 split of a
  master dataset of sequences into padded initial states and truth.

* ``loss_and_grad`` / ``train_step`` (lex_dl_train.py:112-132): the gradient of that loss
  with respect to the network through the solver, which the reference takes with
  jax.value_and_grad, assembled per step in reverse from hand-written pieces:
    lex_weighted_sum_grad, lex_laplace   the loss's gradient with respect to the stacks
    lex_dl_apply_vjp                     -> the adjoint of the network's output
    the network's MFMA backward          -> weight gradients, the adjoint of its input
    lex_dl_pack_vjp                      -> added to the stepped state's adjoint
    lex_rk3_step_vjp_taped               -> the adjoint of the step's input state
  with every step recorded (lex_rk3_step_tape) on the way forward.  The step's adjoint
  covers the compressible core with Smagorinsky (solver_opt 2, turb_opt 1) and the
  physics options off.

The reference's ``dl_params`` (flax parameters) are the network object itself here (a
module mapping the (nx, ny, nz, 5) input to the correction, e.g. dl_models.simpleCNN()),
and its optax state a torch optimizer over the network's parameters.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import dl_models as dlm
from . import namelist_n_constants as nl
from ._core import grid_from
from .one_sprint_functions import fused_step

_EPS = (0.01, 0.01, 0.01, 0.01, 1.0e-5, 1.0e-7, 1.0e-10)


def _interior(a):
    return a[3:-3, 3:-3, 1:-1]


def ws_rk3_sprint_dl_train(theta_now, u_now, v_now, w_now, pip, qv_now, base_state, grids, model_opt, dl_params,
                           scaling_params, sprint_n=None):
    """sprint_n corrected steps from the padded state; returns the stacks (theta, u, v, w,
    pip, qv) of each step's interior fields, (sprint_n, ...) device tensors."""
    n = nl.sprint_n if sprint_n is None else int(sprint_n)
    state = tuple(a.clone() for a in (theta_now, u_now, v_now, w_now, pip, qv_now))
    g = grid_from(base_state[1], *grids[3:6])
    norm = dlm.level_norm(scaling_params)
    ranges = dlm.scaling_ranges(scaling_params, state[0].device)
    stacks = [torch.empty((n,) + tuple(_interior(a).shape), dtype=torch.float32, device=a.device) for a in state]
    with torch.no_grad():
        for i in range(n):
            fused_step(state, base_state, g, model_opt)
            x = dlm.phys_state_tuple2array(state, scaling_params, norm)
            dlm.apply_correction(state, dl_params(x[0]), ranges)
            for q in range(6):
                stacks[q][i].copy_(_interior(state[q]))
    return tuple(stacks)


def ws_rk3_sprint_dl_train_vmap(theta_batch, u_batch, v_batch, w_batch, pip_batch, qv_batch, base_state, grids,
                                model_opt, dl_params, scaling_params, sprint_n=None):
    """The batched sprint (one_sprint_functions.py:100-106): stacks (batch, sprint_n, ...)."""
    outs = [ws_rk3_sprint_dl_train(*(a[b] for a in (theta_batch, u_batch, v_batch, w_batch, pip_batch, qv_batch)),
                                   base_state, grids, model_opt, dl_params, scaling_params, sprint_n)
            for b in range(theta_batch.shape[0])]
    return tuple(torch.stack([o[q] for o in outs]) for q in range(6))


def laplace_of_tensor(x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, s):
    """Laplacian of interior scalar fields with any leading axes (..., nx, ny, nz): periodic
    in x, y, end levels copied in z, no vertical term in the first and last x planes (the
    reference's indexing on its 5-D stacks)."""
    s = s.contiguous().float()
    nx, ny, nz = s.shape[-3:]
    g = grid_from(torch.empty((nx + 6, ny + 6, nz + 2)), x3d4u, y3d4v, z3d4w)
    out = torch.empty_like(s)
    nf = s.numel() // (nx * ny * nz)
    _lib.call("lex_laplace", g, nf, _lib.ptr(s), _lib.ptr(out), _lib.stream())
    return out


def loss_weights(scaling_params, device):
    """Per-point weights of the seven l2 terms (theta, u, v, w, qv, del_th, del_qv) and of the
    negative-qv penalty (lex_dl_train.py:69-96) as device FP32 arrays of the fields' shapes."""
    out = []
    for q in range(8):
        p = 4 if q == 7 else q
        mn = torch.as_tensor(scaling_params[2 * p], dtype=torch.float64, device=device).squeeze()
        mx = torch.as_tensor(scaling_params[2 * p + 1], dtype=torch.float64, device=device).squeeze()
        lev = torch.amax(mx, dim=(0, 1), keepdim=True) - torch.amin(mn, dim=(0, 1), keepdim=True)
        lev = lev + (1.0e-5 if q == 7 else _EPS[q])
        out.append(((mx - mn) / lev ** (2 if q == 7 else 3)).float().contiguous())
    return out


class _Summer:
    """Deterministic FP64 device sums (lex_weighted_sum) into one device accumulator."""

    def __init__(self, device):
        nb = _lib.load().lex_weighted_sum_workspace_bytes()
        self.ws = torch.empty(nb, dtype=torch.uint8, device=device)
        self.terms = []

    def add(self, op, a, b, w):
        a = a.contiguous()
        out = torch.empty(1, dtype=torch.float64, device=a.device)
        _lib.call("lex_weighted_sum", a.numel(), op, _lib.ptr(a), _lib.ptr(b.contiguous()) if b is not None else None,
                  _lib.ptr(w), w.numel(), _lib.ptr(out), _lib.ptr(self.ws), self.ws.numel(), _lib.stream())
        self.terms.append(out)

    def total(self):
        return torch.cat(self.terms).sum()


def loss_and_aux(params, ic, truth, scaling_params, base_state, grids, model_opt, sprint_n=None, weights=None):
    """The training loss of the reference (lex_dl_train.py:49-99): the batched corrected
    sprint from ic, then the weighted l2 terms against truth (theta, u, v, w, pip, qv,
    del_th, del_qv; leading axes batch, step) and the negative-qv penalty.  Returns a
    0-d FP64 device tensor.  (The reference reads base_state, grids and model_opt from
    module globals; they are arguments here.)"""
    th, u, v, w, _, qv = ws_rk3_sprint_dl_train_vmap(*ic[:6], base_state, grids, model_opt, params, scaling_params,
                                                      sprint_n)
    dev = th.device
    wts = loss_weights(scaling_params, dev) if weights is None else weights
    x3d, y3d, z3d, x3d4u, y3d4v, z3d4w = grids[:6]
    theta_min = torch.as_tensor(scaling_params[0], dtype=torch.float32, device=dev).squeeze()
    qv_min = torch.as_tensor(scaling_params[8], dtype=torch.float32, device=dev).squeeze()
    del_th = laplace_of_tensor(x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, th - theta_min)
    del_qv = laplace_of_tensor(x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, qv - qv_min)
    s = _Summer(dev)
    preds = (th, u, v, w, qv, del_th, del_qv)
    trues = tuple(truth[i] for i in (0, 1, 2, 3, 5, 6, 7))
    for p, t, wq in zip(preds, trues, wts[:7]):
        s.add(0, p, torch.as_tensor(t, dtype=torch.float32, device=dev), wq)
    s.add(1, qv, None, wts[7])
    return s.total()


def _wgrad(op, a, b, w, out):
    _lib.call("lex_weighted_sum_grad", a.numel(), op, _lib.ptr(a), _lib.ptr(b) if b is not None else None, _lib.ptr(w),
              w.numel(), _lib.ptr(out), _lib.stream())


def _check_differentiable(model_opt):
    solver_opt, damp_opt, rad_opt, cor_opt, sfc_opt, pic_opt, turb_opt = model_opt
    if solver_opt != 2 or turb_opt != 1 or damp_opt or rad_opt or cor_opt or sfc_opt or pic_opt:
        raise NotImplementedError("the step's adjoint covers solver_opt 2 with turb_opt 1 and the physics options off")


def _stack_grads(stacks, truth, wts, th_min, qv_min, lap, summer):
    """Loss terms of one sequence's stacks (theta, u, v, w, pip, qv; leading axis step)
    added to `summer`, and the loss's gradient with respect to each stack."""
    th, u, v, w, pip, qv = stacks
    grads = [torch.zeros_like(a) for a in stacks]
    for q, (a, t, wq) in enumerate(zip((th, u, v, w, qv), (truth[0], truth[1], truth[2], truth[3], truth[5]),
                                       wts[:5])):
        summer.add(0, a, t, wq)
        _wgrad(0, a.contiguous(), t.contiguous(), wq, grads[q if q < 4 else 5])
    for a, amin, t, wq, gq in ((th, th_min, truth[6], wts[5], grads[0]), (qv, qv_min, truth[7], wts[6], grads[5])):
        la = lap(a - amin)
        summer.add(0, la, t, wq)
        r = torch.zeros_like(la)
        _wgrad(0, la, t.contiguous(), wq, r)
        gq += lap(r)   # the Laplacian is symmetric
    summer.add(1, qv, None, wts[7])
    _wgrad(1, qv.contiguous(), None, wts[7], grads[5])
    return grads


def loss_and_grad(params, ic, truth, scaling_params, base_state, grids, model_opt, sprint_n=None):
    """loss_and_aux and its gradient with respect to the network's parameters, accumulated
    into their .grad (lex_dl_train.py:112-116).  Returns the loss (0-d FP64 device tensor)."""
    from . import adjoint as adj
    _check_differentiable(model_opt)
    n = nl.sprint_n if sprint_n is None else int(sprint_n)
    dev = ic[0].device
    g = grid_from(base_state[1], *grids[3:6])
    norm = dlm.level_norm(scaling_params)
    lo, den = (a.to(dev).contiguous() for a in norm)
    ranges = dlm.scaling_ranges(scaling_params, dev)
    wts = loss_weights(scaling_params, dev)
    x3d, y3d, z3d, x3d4u, y3d4v, z3d4w = grids[:6]
    th_min = torch.as_tensor(scaling_params[0], dtype=torch.float32, device=dev).squeeze()
    qv_min = torch.as_tensor(scaling_params[8], dtype=torch.float32, device=dev).squeeze()

    def lap(a):
        return laplace_of_tensor(x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, a)

    summer = _Summer(dev)
    step_args = (base_state, g, nl.dt, nl.n_sound, nl.cs, nl.prandtl)
    for b in range(ic[0].shape[0]):
        state = tuple(a[b].clone() for a in ic[:6])
        xs, tapes, recs = [], [], []
        stacks = [torch.empty((n,) + tuple(_interior(a).shape), dtype=torch.float32, device=dev) for a in state]
        for i in range(n):
            xs.append(state)
            tape, s_out = adj.step_tape(state, *step_args)
            p = dlm.phys_state_tuple2array(s_out, scaling_params, (lo, den))[0].detach().clone().requires_grad_(True)
            c = params(p)
            nxt = tuple(a.clone() for a in s_out)
            dlm.apply_correction(nxt, c.detach(), ranges)
            for q in range(6):
                stacks[q][i].copy_(_interior(nxt[q]))
            tapes.append(tape)
            recs.append((p, c))
            state = nxt
        grads = _stack_grads(stacks, tuple(torch.as_tensor(t[b], dtype=torch.float32, device=dev) for t in truth),
                             wts, th_min, qv_min, lap, summer)
        lam = tuple(torch.zeros_like(a) for a in state)
        for i in reversed(range(n)):
            for q in (0, 1, 2, 3, 5):
                _interior(lam[q]).add_(grads[q][i])
            p, c = recs[i]
            dc = torch.empty((g.nx, g.ny, g.nz, 5), dtype=torch.float32, device=dev)
            _lib.call("lex_dl_apply_vjp", g, *[_lib.ptr(lam[q]) for q in (0, 1, 2, 3, 5)],
                      *[_lib.ptr(r) for r in ranges], _lib.ptr(dc), _lib.stream())
            c.backward(dc.view_as(c))
            lam_s = tuple(t.clone() for t in lam)
            _lib.call("lex_dl_pack_vjp", g, _lib.ptr(p.grad.contiguous()), _lib.ptr(lo), _lib.ptr(den),
                      *[_lib.ptr(lam_s[q]) for q in (0, 1, 2, 3, 5)], _lib.stream())
            lam, _ = adj.step_vjp_taped(xs[i], tapes[i], lam_s, *step_args)
            recs[i] = None
    return summer.total()


def train_step(dl_state, ic, truth, scaling_params, base_state, grids, model_opt, sprint_n=None):
    """train_step (lex_dl_train.py:112-132): dl_state = (network, torch optimizer over its
    parameters); one gradient step on the loss; returns {"total_mse": loss, "grad_norm": ...}."""
    net, opt = dl_state
    opt.zero_grad(set_to_none=False)
    for prm in net.parameters():
        if prm.grad is None:
            prm.grad = torch.zeros_like(prm)
    loss = loss_and_grad(net, ic, truth, scaling_params, base_state, grids, model_opt, sprint_n)
    gn = torch.sqrt(sum((prm.grad.double() ** 2).sum() for prm in net.parameters()))
    opt.step()
    return {"total_mse": loss, "grad_norm": gn}


def eval_step(dl_state, ic, truth, scaling_params, base_state, grids, model_opt, sprint_n=None):
    """eval_step (lex_dl_train.py:135-139): {"total_mse": loss} (dl_state: the network)."""
    return {"total_mse": loss_and_aux(dl_state, ic, truth, scaling_params, base_state, grids, model_opt, sprint_n)}


def _pad_batch(batch, mode):
    """Pad every sample of a (B, ...) interior batch on the device (mode 0 scalar / w,
    1 u, 2 v)."""
    b = torch.as_tensor(batch, dtype=torch.float32, device="cuda").contiguous()
    ni, nj, nk = b.shape[1:]
    out = torch.empty((b.shape[0], ni + 6, nj + 6, nk + 2), dtype=torch.float32, device=b.device)
    for q in range(b.shape[0]):
        _lib.call("lex_pad", mode, ni, nj, nk, 3, _lib.ptr(b[q]), _lib.ptr(out[q]), _lib.stream())
    return out


def padding3_batch(tensor):
    """(B, nx, ny, nz) -> (B, nx+6, ny+6, nz+2) (dataset_generation.py:234-243)."""
    return _pad_batch(tensor, 0)


def padding3u_batch(tensor):
    """(B, nx+1, ny, nz) -> (B, nx+7, ny+6, nz+2) (dataset_generation.py:245-256)."""
    return _pad_batch(tensor, 1)


def padding3v_batch(tensor):
    """(B, nx, ny+1, nz) -> (B, nx+6, ny+7, nz+2) (dataset_generation.py:258-268)."""
    return _pad_batch(tensor, 2)


def split_indices(num_seq, seed):
    """The shuffled order and the 75 % training count of generate_training_data
    (dataset_generation.py:117-122): numpy's default_rng(seed).choice without replacement."""
    rng = np.random.default_rng(seed)
    heads = rng.choice(np.arange(num_seq), num_seq, replace=False)
    return heads, int(np.int32(num_seq * 0.75))


def generate_training_data(dataset, seed):
    """Shuffle the master dataset (theta, u, v, w, pip, qv, del_th, del_qv; each
    (sequences, steps, ...)) and split it 75 / 25 (dataset_generation.py:115-194):
    returns (train_ic, train_true, test_ic, test_true, epoch_size, test_size) with the
    initial states padded on the device and the truth the later steps."""
    heads, num_train = split_indices(dataset[0].shape[0], seed)
    seqs = [np.asarray(a)[heads] for a in dataset]
    pads = (padding3_batch, padding3u_batch, padding3v_batch, padding3_batch, padding3_batch, padding3_batch)

    def part(sl):
        ic = tuple(p(a[sl, 0]) for p, a in zip(pads, seqs[:6]))
        true = tuple(torch.as_tensor(a[sl, 1:], dtype=torch.float32, device="cuda") for a in seqs)
        return ic, true

    train_ic, train_true = part(slice(0, num_train))
    test_ic, test_true = part(slice(num_train, None))
    return train_ic, train_true, test_ic, test_true, train_ic[0].shape[0], test_ic[0].shape[0]


__all__ = ["ws_rk3_sprint_dl_train", "ws_rk3_sprint_dl_train_vmap", "laplace_of_tensor", "loss_weights",
           "loss_and_aux", "loss_and_grad", "train_step", "eval_step", "padding3_batch", "padding3u_batch", "padding3v_batch", "split_indices",
           "generate_training_data"]
