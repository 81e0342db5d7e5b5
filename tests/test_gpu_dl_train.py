"""The loss side of the post-step correction's training on the device vs the oracle:
laplace_of_tensor (lex_laplace), the deterministic weighted sums (lex_weighted_sum), the
training sprint's stacks and the whole loss_and_aux with simpleCNN, and the dataset split
with its device padding."""
import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import check_fp32, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _grids(gr):
    from lex_amd.setup_lex import make_grids
    return make_grids(gr.nx, gr.ny, gr.nz, gr.dx, gr.dy, gr.dz)


def test_laplace_of_tensor_vs_oracle():
    from lex_amd import dl_train as dt
    gr = o.Grid(20, 12, 16, 100.0, 80.0, 50.0)
    rng = np.random.default_rng(4)
    s = (300.0 + rng.normal(0, 1, (2, 3, gr.nx, gr.ny, gr.nz))).astype(np.float32)
    x3d, y3d, z3d, x3d4u, y3d4v, z3d4w = _grids(gr)[:6]
    got = from_dev(dt.laplace_of_tensor(x3d, x3d4u, y3d, y3d4v, z3d, z3d4w, to_dev(s - np.float32(299.0))))
    r64 = o.laplace_of_tensor(s.astype(np.float64) - 299.0, gr)
    r32 = o.laplace_of_tensor(s - np.float32(299.0), gr)
    check_fp32(got, r64, r32, "laplace", factor=8.0)
    assert np.max(np.abs(got[..., [0, -1], :, :] - r64[..., [0, -1], :, :])) < 1e-4   # the x-plane rule


@pytest.mark.parametrize("op", [0, 1])
def test_weighted_sum_vs_numpy(op):
    import torch
    from lex_amd import _lib
    rng = np.random.default_rng(9 + op)
    n, period = 3 * 1_000_003, 1_000_003
    a = rng.normal(0, 1, n).astype(np.float32)
    b = rng.normal(0, 1, n).astype(np.float32)
    w = rng.uniform(0, 2, period).astype(np.float32)
    ws = torch.empty(_lib.load().lex_weighted_sum_workspace_bytes(), dtype=torch.uint8, device="cuda")
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    da, db, dw = to_dev(a), to_dev(b), to_dev(w)
    _lib.call("lex_weighted_sum", n, op, _lib.ptr(da), _lib.ptr(db), _lib.ptr(dw), period, _lib.ptr(out),
              _lib.ptr(ws), ws.numel(), _lib.stream())
    a64, b64, w64 = a.astype(np.float64), b.astype(np.float64), np.tile(w.astype(np.float64), 3)
    exp = np.sum(w64 * 0.5 * (a64 - b64) ** 2) if op == 0 else np.sum(w64 * np.where(a64 < 0, -a64, 0.0))
    got = float(out.item())
    assert abs(got - exp) <= 1e-10 * abs(exp)
    # deterministic: the same bits again
    _lib.call("lex_weighted_sum", n, op, _lib.ptr(da), _lib.ptr(db), _lib.ptr(dw), period, _lib.ptr(out),
              _lib.ptr(ws), ws.numel(), _lib.stream())
    assert float(out.item()) == got


def _scaling(rng, nx, ny, nz):
    shapes = [(nx, ny, nz), (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz), (nx, ny, nz),
              (nx, ny, nz)]
    scales = [1.0, 1.0, 1.0, 0.5, 1e-3, 1e-4, 1e-8]
    offs = [300.0, 0.0, 0.0, 0.0, 5e-3, 0.0, 0.0]
    out = []
    for shp, sc, off in zip(shapes, scales, offs):
        b = rng.normal(0, sc, shp) + off
        out += [(b - rng.uniform(0.1, 0.5, shp) * sc)[None, None].astype(np.float32),
                (b + rng.uniform(0.1, 0.5, shp) * sc)[None, None].astype(np.float32)]
    return tuple(out)


def test_loss_and_aux_vs_oracle():
    import torch
    from lex_amd import dl_models as dlm, dl_train as dt, namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    case = case.scaled(16, 12, 12)
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    model_opt = (2, False, False, 0, False, False, 1)
    rng = np.random.default_rng(12)
    spr = _scaling(rng, case.nx, case.ny, case.nz)
    spd = tuple(to_dev(a) for a in spr)
    net = dlm.simpleCNN(seed=9)
    # a small correction, so that the steps dominate the stacks' increments (the BF16
    # network sees inputs rounded from slightly different FP32 / FP64 states)
    with torch.no_grad():
        net.weights[-1].mul_(1e-5)
        net.biases[-1].fill_(1e-6)
    B, S = 2, 2
    # two initial states: the bubble and the bubble with a small wind
    ic = [torch.stack([a, a + (0.5 if q == 1 else 0.0)]) for q, a in enumerate(phys[:6])]
    got_stacks = dt.ws_rk3_sprint_dl_train_vmap(*ic, base, grids, model_opt, net, spd, sprint_n=S)

    def net_fn(x):
        with torch.no_grad():
            return from_dev(net(to_dev(x.astype(np.float32))))

    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    oo = o.Options(dt=opts["dt"], solver_opt=2, n_sound=7, turb_opt=1)
    bs = tuple(from_dev(a).astype(np.float64) for a in base[:5]) + (None,)
    sp64 = tuple(a.astype(np.float64) for a in spr)
    bs32 = tuple(from_dev(a).astype(np.float32) for a in base[:5]) + (None,)

    def oracle_stacks(dtype):
        ref = []
        for b in range(B):
            x0 = tuple(from_dev(a[b]).astype(dtype) for a in ic)
            st, _ = o.ws_rk3_sprint_dl_train(x0, bs if dtype == np.float64 else bs32, og, oo, net_fn,
                                             tuple(a.astype(dtype) for a in spr), S)
            ref.append(st)
        return tuple(np.stack([r[q] for r in ref]) for q in range(6))

    ref_stacks, ref32 = oracle_stacks(np.float64), oracle_stacks(np.float32)
    # increments of the stacks over the initial state, vs the FP64 oracle, within 10x the
    # FP32 oracle's own error (theta's increments over two steps are a few ulp of 300 K)
    for q, name in enumerate(("theta", "u", "v", "w", "pip", "qv")):
        i0 = np.stack([np.stack([from_dev(ic[q][b])[3:-3, 3:-3, 1:-1]] * S) for b in range(B)])
        r = ref_stacks[q] - i0
        nr = max(np.linalg.norm(r), 1e-30)
        err = np.linalg.norm(from_dev(got_stacks[q]) - i0 - r) / nr
        e32 = np.linalg.norm(ref32[q].astype(np.float64) - i0 - r) / nr
        assert err <= max(10.0 * e32, 1e-3), f"{name}: rel-L2 of the increments {err:.2e} (FP32 oracle {e32:.2e})"
    # the loss on the device's own stacks: truth = those stacks perturbed, and their
    # Laplacians perturbed
    own = tuple(from_dev(a) for a in got_stacks)
    truth = [r + rng.normal(0, s, r.shape) for r, s in zip(own, (0.05, 0.05, 0.05, 0.05, 1e-4, 1e-4))]
    truth += [o.laplace_of_tensor(own[0] - sp64[0][0, 0], og) + rng.normal(0, 1e-4, own[0].shape),
              o.laplace_of_tensor(own[5] - sp64[8][0, 0], og) + rng.normal(0, 1e-9, own[5].shape)]
    sp_o = tuple(a[0, 0] for a in sp64)
    exp = o.dl_loss(own, truth, sp_o, og)
    got = float(dt.loss_and_aux(net, tuple(ic), tuple(to_dev(t.astype(np.float32)) for t in truth), spd, base,
                                grids, model_opt, sprint_n=S).item())
    assert abs(got - exp) <= 1e-3 * exp, (got, exp)
    ev = dt.eval_step(net, tuple(ic), tuple(to_dev(t.astype(np.float32)) for t in truth), spd, base, grids,
                      model_opt, sprint_n=S)
    assert float(ev["total_mse"].item()) == got


def test_generate_training_data_split_and_padding():
    from lex_amd import dl_train as dt
    nx, ny, nz, n_seq, n_step = 8, 6, 5, 8, 3
    rng = np.random.default_rng(1)
    shapes = [(nx, ny, nz), (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz), (nx, ny, nz),
              (nx, ny, nz), (nx, ny, nz)]
    data = tuple(rng.normal(size=(n_seq, n_step) + s).astype(np.float32) for s in shapes)
    tr_ic, tr_true, te_ic, te_true, n_tr, n_te = dt.generate_training_data(data, 5)
    heads, n_train = dt.split_indices(n_seq, 5)
    assert (n_tr, n_te) == (n_train, n_seq - n_train) == (6, 2)
    pads = (o.padding3_array, o.padding3u_array, o.padding3v_array, o.padding3_array, o.padding3_array,
            o.padding3_array)
    for q in range(6):
        for b in range(n_tr):
            exp = pads[q](data[q][heads[b], 0].astype(np.float64))
            assert np.array_equal(from_dev(tr_ic[q][b]), exp.astype(np.float32).astype(np.float64)), (q, b)
        assert np.array_equal(from_dev(te_ic[q][0]), pads[q](data[q][heads[n_train], 0].astype(np.float64)))
    for q in range(8):
        assert np.array_equal(from_dev(tr_true[q]), data[q][heads[:n_train], 1:].astype(np.float64))
        assert np.array_equal(from_dev(te_true[q]), data[q][heads[n_train:], 1:].astype(np.float64))


def _rand_padded(rng, gr):
    return (o.padding3_array(rng.normal(0, 1, (gr.nx, gr.ny, gr.nz))).astype(np.float32),
            o.padding3u_array(rng.normal(0, 1, (gr.nx + 1, gr.ny, gr.nz))).astype(np.float32),
            o.padding3v_array(rng.normal(0, 1, (gr.nx, gr.ny + 1, gr.nz))).astype(np.float32),
            o.padding3_array(rng.normal(0, 1, (gr.nx, gr.ny, gr.nz + 1))).astype(np.float32),
            o.padding3_array(rng.normal(0, 1, (gr.nx, gr.ny, gr.nz))).astype(np.float32),
            o.padding3_array(rng.normal(0, 1, (gr.nx, gr.ny, gr.nz))).astype(np.float32))


def _interior_of(a, q):
    return a[3:-3, 3:-3, 1:-1]


def test_correction_reverses_are_transposes():
    """<A c, lam> = <c, A^T lam> for the correction (lex_dl_apply on a zero state) and
    <pack(x), dp> = <x, pack^T dp> for the packing with zero offsets, on the interior points
    each map reads or writes (lam folded: zero ghosts)."""
    import torch
    from lex_amd import _lib
    gr = o.Grid(12, 10, 8, 100.0, 100.0, 100.0)
    g = _lib.grid(gr.nx, gr.ny, gr.nz, gr.dx, gr.dy, gr.dz)
    rng = np.random.default_rng(21)
    c = rng.normal(0, 1, (gr.nx, gr.ny, gr.nz, 5)).astype(np.float32)
    rs = [rng.uniform(0.5, 2, s).astype(np.float32) for s in ((gr.nx, gr.ny, gr.nz), (gr.nx + 1, gr.ny, gr.nz),
                                                              (gr.nx, gr.ny + 1, gr.nz), (gr.nx, gr.ny, gr.nz + 1),
                                                              (gr.nx, gr.ny, gr.nz))]
    # device copies held in names: a pointer taken from a temporary tensor would dangle
    zero = [to_dev(np.zeros_like(a)) for a in _rand_padded(rng, gr)]
    c_d = to_dev(c)
    rs_d = [to_dev(r) for r in rs]
    _lib.call("lex_dl_apply", g, _lib.ptr(c_d), *[_lib.ptr(r) for r in rs_d],
              *[_lib.ptr(zero[q]) for q in (0, 1, 2, 3, 5)], _lib.stream())
    lam = [o.padding3_array(rng.normal(0, 1, (gr.nx, gr.ny, gr.nz))), None, None, None, None, None]
    lam[1] = np.zeros((gr.nx + 7, gr.ny + 6, gr.nz + 2))
    lam[1][3:-3, 3:-3, 1:-1] = rng.normal(0, 1, (gr.nx + 1, gr.ny, gr.nz))
    lam[2] = np.zeros((gr.nx + 6, gr.ny + 7, gr.nz + 2))
    lam[2][3:-3, 3:-3, 1:-1] = rng.normal(0, 1, (gr.nx, gr.ny + 1, gr.nz))
    lam[3] = np.zeros((gr.nx + 6, gr.ny + 6, gr.nz + 3))
    lam[3][3:-3, 3:-3, 1:-1] = rng.normal(0, 1, (gr.nx, gr.ny, gr.nz + 1))
    lam[5] = np.zeros((gr.nx + 6, gr.ny + 6, gr.nz + 2))
    lam[5][3:-3, 3:-3, 1:-1] = rng.normal(0, 1, (gr.nx, gr.ny, gr.nz))
    lam[0][:3] = lam[0][-3:] = 0.0
    lam[0][:, :3] = lam[0][:, -3:] = 0.0
    lam[0][:, :, 0] = lam[0][:, :, -1] = 0.0
    lam32 = [None if a is None else a.astype(np.float32) for a in lam]
    lhs = sum(float(np.sum(from_dev(zero[q]) * lam32[q].astype(np.float64))) for q in (0, 1, 2, 3, 5))
    dc = torch.empty((gr.nx, gr.ny, gr.nz, 5), dtype=torch.float32, device="cuda")
    lam_d = [None if a is None else to_dev(a) for a in lam32]
    _lib.call("lex_dl_apply_vjp", g, *[_lib.ptr(lam_d[q]) for q in (0, 1, 2, 3, 5)],
              *[_lib.ptr(r) for r in rs_d], _lib.ptr(dc), _lib.stream())
    rhs = float(np.sum(from_dev(dc) * c.astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1.0), (lhs, rhs)
    # the packing (lo = 0, so linear)
    x = [to_dev(a) for a in _rand_padded(rng, gr)]
    lo = torch.zeros(5 * gr.nz + 1, dtype=torch.float32, device="cuda")
    den = to_dev(rng.uniform(0.5, 2.0, 5 * gr.nz + 1).astype(np.float32))
    out = torch.empty((gr.nx, gr.ny, gr.nz, 5), dtype=torch.float32, device="cuda")
    _lib.call("lex_dl_pack", g, *[_lib.ptr(x[q]) for q in (0, 1, 2, 3, 5)], _lib.ptr(lo), _lib.ptr(den), _lib.ptr(out),
              _lib.stream())
    dp = rng.normal(0, 1, (gr.nx, gr.ny, gr.nz, 5)).astype(np.float32)
    lhs = float(np.sum(from_dev(out) * dp.astype(np.float64)))
    adjs = [torch.zeros_like(x[q]) for q in range(6)]
    dp_d = to_dev(dp)
    _lib.call("lex_dl_pack_vjp", g, _lib.ptr(dp_d), _lib.ptr(lo), _lib.ptr(den),
              *[_lib.ptr(adjs[q]) for q in (0, 1, 2, 3, 5)], _lib.stream())
    rhs = sum(float(np.sum(from_dev(adjs[q]) * from_dev(x[q]))) for q in (0, 1, 2, 3, 5))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1.0), (lhs, rhs)
    for q in (0, 1, 2, 3, 5):   # only interior points receive adjoint
        a = from_dev(adjs[q])
        assert not a[:3].any() and not a[-3:].any() and not a[:, :3].any() and not a[:, -3:].any()


def test_loss_gradient_of_the_last_bias():
    """One corrected step: a constant on the last layer's bias of channel q shifts field q
    by the bias times its range on every point the correction reaches (stagger_uvw's end
    faces average two equal values; w's surface and lid stay), so d loss / d bias_q is the
    sum over those points of range x d loss / d field, with the loss's gradient in FP64
    from the device's own stacks (l2 terms, the symmetric Laplacian's chain, the
    negative-qv penalty).  Tolerance: the network's BF16 backward (2^-7 of the summed
    magnitudes)."""
    import torch
    from lex_amd import dl_models as dlm, dl_train as dt, namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    case = case.scaled(16, 12, 12)
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    model_opt = (2, False, False, 0, False, False, 1)
    rng = np.random.default_rng(31)
    spd = tuple(to_dev(a) for a in _scaling(rng, case.nx, case.ny, case.nz))
    net = dlm.simpleCNN(seed=11)
    with torch.no_grad():
        net.weights[-1].mul_(1e-3)
        net.biases[-1].fill_(1e-4)
    ic = tuple(a.unsqueeze(0).clone() for a in phys[:6])
    stacks = dt.ws_rk3_sprint_dl_train_vmap(*ic, base, grids, model_opt, net, spd, sprint_n=1)
    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    own = tuple(from_dev(a) for a in stacks)
    sp0 = tuple(from_dev(a)[0, 0] for a in spd)
    truth = [r + rng.normal(0, s, r.shape) for r, s in zip(own, (0.05, 0.05, 0.05, 0.05, 1e-4, 1e-4))]
    truth += [o.laplace_of_tensor(own[0] - sp0[0], og) + rng.normal(0, 1e-4, own[0].shape),
              o.laplace_of_tensor(own[5] - sp0[8], og) + rng.normal(0, 1e-9, own[5].shape)]
    truth_d = tuple(to_dev(t.astype(np.float32)) for t in truth)
    for prm in net.parameters():
        prm.grad = torch.zeros_like(prm)
    loss = float(dt.loss_and_grad(net, ic, truth_d, spd, base, grids, model_opt, sprint_n=1).item())
    assert abs(loss - o.dl_loss(own, truth, sp0, og)) <= 1e-5 * loss
    gb = from_dev(net.biases[-1].grad)
    w8 = o.dl_loss_weights(sp0)
    tr = [np.asarray(t, np.float64) for t in truth]
    g = [w8[0] * (own[0] - tr[0]), w8[1] * (own[1] - tr[1]), w8[2] * (own[2] - tr[2]), w8[3] * (own[3] - tr[3]),
         w8[4] * (own[5] - tr[5])]
    g[0] = g[0] + o.laplace_of_tensor(w8[5] * (o.laplace_of_tensor(own[0] - sp0[0], og) - tr[6]), og)
    g[4] = g[4] + o.laplace_of_tensor(w8[6] * (o.laplace_of_tensor(own[5] - sp0[8], og) - tr[7]), og)
    g[4] = g[4] - np.where(own[5] < 0, w8[7], 0.0)
    for q in range(5):
        r = sp0[2 * q + 1] - sp0[2 * q]
        terms = g[q] * r
        if q == 3:   # w: the surface and lid levels receive no correction
            terms = terms[..., 1:-1]
        exact, mag = float(terms.sum()), float(np.abs(terms).sum())
        print(f"channel {q}: adjoint {gb[q]:.6e} exact {exact:.6e} (sum of magnitudes {mag:.3e})")
        assert abs(gb[q] - exact) <= 2.0 ** -7 * mag + 1e-6 * abs(exact), q
    assert all(np.isfinite(from_dev(prm.grad)).all() for prm in net.parameters())


def test_two_step_gradient_is_finite_and_uses_the_solver():
    """Two corrected steps: the gradient passes through the step's adjoint (the first
    step's bias gradient differs from a one-step sprint's) and stays finite."""
    import torch
    from lex_amd import dl_models as dlm, dl_train as dt, namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    case = case.scaled(16, 12, 12)
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    model_opt = (2, False, False, 0, False, False, 1)
    rng = np.random.default_rng(33)
    spd = tuple(to_dev(a) for a in _scaling(rng, case.nx, case.ny, case.nz))
    net = dlm.simpleCNN(seed=12)
    with torch.no_grad():
        net.weights[-1].mul_(1e-3)
    ic = tuple(a.unsqueeze(0).clone() for a in phys[:6])
    stacks = dt.ws_rk3_sprint_dl_train_vmap(*ic, base, grids, model_opt, net, spd, sprint_n=2)
    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    own = tuple(from_dev(a) for a in stacks)
    sp0 = tuple(from_dev(a)[0, 0] for a in spd)
    truth = [r + 0.01 for r in own] + [o.laplace_of_tensor(own[0] - sp0[0], og), o.laplace_of_tensor(own[5] - sp0[8], og)]
    truth_d = tuple(to_dev(t.astype(np.float32)) for t in truth)
    grads = []
    for n in (1, 2):
        for prm in net.parameters():
            prm.grad = torch.zeros_like(prm)
        tr = tuple(t[:, :n].contiguous() for t in truth_d)
        dt.loss_and_grad(net, ic, tr, spd, base, grids, model_opt, sprint_n=n)
        grads.append([from_dev(prm.grad) for prm in net.parameters()])
    assert all(np.isfinite(a).all() for a in grads[1])
    assert any(not np.allclose(a, b) for a, b in zip(grads[0], grads[1]))


def test_two_step_bias_gradient_vs_central_differences():
    """Two corrected steps: d loss / d (last-layer bias) from loss_and_grad (the chain of
    k_dl_apply_vjp, the network's BF16 backward, k_dl_pack_vjp and the step's adjoint
    lex_rk3_step_vjp_taped across the first step) against central differences of the
    FP64 oracle sprint's loss, with the network's forward restated in FP64 (torch conv3d
    on the CPU, the same weights).  The bias of step 1 reaches the loss through the
    carried state and through the second step's network input, so a dropped or
    mis-signed term of the chain moves the result by a large fraction.  The residuals
    are one-signed (no cancellation in the sum over cells), so the tolerance is 2 % of
    the finite difference: BF16 rounding of the per-cell gradients (2^-8) plus the
    FP32 / BF16 trajectory against the FP64 one."""
    import torch
    from lex_amd import dl_models as dlm, dl_train as dt, namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    case = case.scaled(16, 12, 12)
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    model_opt = (2, False, False, 0, False, False, 1)
    rng = np.random.default_rng(35)
    spr = _scaling(rng, case.nx, case.ny, case.nz)
    spd = tuple(to_dev(a) for a in spr)
    net = dlm.simpleCNN(seed=13)
    with torch.no_grad():
        net.weights[-1].mul_(1e-2)
        net.biases[-1].fill_(1e-4)
    ic = tuple(a.unsqueeze(0).clone() for a in phys[:6])
    S = 2
    stacks = dt.ws_rk3_sprint_dl_train_vmap(*ic, base, grids, model_opt, net, spd, sprint_n=S)
    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    own = tuple(from_dev(a) for a in stacks)
    sp0 = tuple(from_dev(a)[0, 0] for a in spd)
    # one-signed residuals: every cell's loss gradient has the same sign per field, so the
    # bias gradient is a sum without cancellation and the BF16 rounding of the per-cell
    # gradients stays a relative error of the result
    truth = [r - sc * (1.0 + 0.5 * rng.random(r.shape)) for r, sc in zip(own, (0.05, 0.05, 0.05, 0.05, 1e-4, 1e-4))]
    truth += [o.laplace_of_tensor(own[0] - sp0[0], og) - 1e-4, o.laplace_of_tensor(own[5] - sp0[8], og) - 1e-9]
    truth_d = tuple(to_dev(t.astype(np.float32)) for t in truth)
    for prm in net.parameters():
        prm.grad = torch.zeros_like(prm)
    dt.loss_and_grad(net, ic, truth_d, spd, base, grids, model_opt, sprint_n=S)
    gb = from_dev(net.biases[-1].grad)

    # FP64 restatement of the network's forward (the same weights; the last bias varies)
    ws = [w.detach().double().cpu() for w in net.weights]
    bs = [b.detach().double().cpu() for b in net.biases]

    def tw(W):
        cin, cout = W.shape[1], W.shape[2]
        return W.reshape(3, 3, 3, cin, cout).permute(4, 3, 0, 1, 2).contiguous()

    def net64(x, b_last):
        h = torch.from_numpy(np.asarray(x, np.float64)).permute(3, 0, 1, 2).unsqueeze(0)
        for li, (w, b) in enumerate(zip(ws, bs)):
            h = torch.nn.functional.conv3d(h, tw(w), b_last if li == len(ws) - 1 else b, padding=1)
            if li < len(ws) - 1:
                h = torch.nn.functional.gelu(h, approximate="tanh")
        return h[0].permute(1, 2, 3, 0).numpy()

    oo = o.Options(dt=opts["dt"], solver_opt=2, n_sound=7, turb_opt=1)
    bs64 = tuple(from_dev(a).astype(np.float64) for a in base[:5]) + (None,)
    sp64 = tuple(a.astype(np.float64) for a in spr)
    x0 = tuple(from_dev(a[0]) for a in ic)

    def loss64(b_last):
        st, _ = o.ws_rk3_sprint_dl_train(x0, bs64, og, oo, lambda x: net64(x, b_last), sp64, S)
        return o.dl_loss(tuple(a[None] for a in st), truth, sp0, og)

    h = 2e-5
    for q in range(5):
        bp, bm = bs[-1].clone(), bs[-1].clone()
        bp[q] += h
        bm[q] -= h
        fd = (loss64(bp) - loss64(bm)) / (2 * h)
        print(f"channel {q}: adjoint {gb[q]:.6e} central difference {fd:.6e}")
        assert abs(gb[q] - fd) <= 2e-2 * abs(fd), (q, gb[q], fd)


def test_train_step_lowers_the_loss():
    import torch
    from lex_amd import dl_models as dlm, dl_train as dt, namelist_n_constants as nl
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    case = case.scaled(16, 12, 12)
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    model_opt = (2, False, False, 0, False, False, 1)
    rng = np.random.default_rng(41)
    spd = tuple(to_dev(a) for a in _scaling(rng, case.nx, case.ny, case.nz))
    net = dlm.simpleCNN(seed=13)
    with torch.no_grad():
        net.weights[-1].mul_(1e-3)
    ic = tuple(a.unsqueeze(0).clone() for a in phys[:6])
    stacks = dt.ws_rk3_sprint_dl_train_vmap(*ic, base, grids, model_opt, net, spd, sprint_n=1)
    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    own = tuple(from_dev(a) for a in stacks)
    sp0 = tuple(from_dev(a)[0, 0] for a in spd)
    truth = [own[0] + 0.01, own[1] - 0.02, own[2], own[3], own[4], own[5]]
    truth += [o.laplace_of_tensor(own[0] - sp0[0], og), o.laplace_of_tensor(own[5] - sp0[8], og)]
    truth = tuple(to_dev(t.astype(np.float32)) for t in truth)
    # Adam's first steps move every parameter by about lr along minus the gradient's sign:
    # a small lr keeps them in the loss's first-order neighbourhood
    opt = torch.optim.Adam(net.parameters(), lr=1e-6)
    losses = [float(dt.train_step((net, opt), ic, truth, spd, base, grids, model_opt, sprint_n=1)["total_mse"].item())
              for _ in range(3)]
    assert losses[1] < losses[0] and losses[2] < losses[1], losses
