"""Adjoint parity: the hand-written VJP kernels (through the C ABI) against FP64 central
finite differences of the oracle, by the dot-product identity

    <lam, (F(x + eps v) - F(x - eps v)) / (2 eps)>  ==  <J^T lam, v>

for random lam and random directions v.  The right side is the GPU adjoint (FP32), the
left side the FP64 oracle.  Per-operator checks perturb every padded position of one
input at a time; whole-step checks use halo-consistent directions (the adjoint of a
step input is returned folded onto the interior).  Tolerances are written in each test:
1e-4 relative for the smooth operators, looser where the Smagorinsky square root and
the many accumulated FP32 terms of a full step enter.
"""
import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import f32, f64, from_dev, random_state, to_dev

pytestmark = pytest.mark.gpu

NX, NY, NZ = 16, 12, 12
EPS = 1e-6


def _dot(a, b):
    return float(np.sum(np.asarray(a, np.float64) * np.asarray(b, np.float64)))


def _fd(fun, x, v, lam, eps=EPS):
    """<lam, directional derivative> by FP64 central differences; fun returns a tuple."""
    fp = fun(*[a + eps * b for a, b in zip(x, v)])
    fm = fun(*[a - eps * b for a, b in zip(x, v)])
    return sum(_dot(l, (p - m) / (2 * eps)) for l, p, m in zip(lam, fp, fm))


def _check(gpu, fd, tol, what):
    err = abs(gpu - fd) / max(abs(fd), 1e-30)
    assert err < tol, f"{what}: adjoint {gpu:.9e} vs FD {fd:.9e} (rel {err:.2e} > {tol})"
    return err


def _make_state(nx, ny, nz, seed):
    import torch  # noqa: F401
    gr, phys, base = random_state(nx, ny, nz, seed=seed)
    p32 = tuple(f32(a) for a in phys)
    b32 = tuple(f32(a) for a in base[:5])
    from lex_amd import _lib
    g = _lib.grid(nx, ny, nz, gr.dx, gr.dy, gr.dz)
    cols = {n: to_dev(np.ascontiguousarray(b32[i][3, 3, :])) for n, i in
            (("rho0", 1), ("theta0", 2), ("pi0", 3), ("qv0", 4))}
    return dict(gr=gr, g=g, p=tuple(f64(a) for a in p32), b=tuple(f64(a) for a in b32), dev=tuple(to_dev(a) for a in p32),
                cols=cols, rng=np.random.default_rng(11), shape=(nx, ny, nz))


@pytest.fixture(scope="module")
def st():
    import torch  # noqa: F401
    gr, phys, base = random_state(NX, NY, NZ, seed=7)
    p32 = tuple(f32(a) for a in phys)
    b32 = tuple(f32(a) for a in base[:5])
    from lex_amd import _lib
    g = _lib.grid(NX, NY, NZ, gr.dx, gr.dy, gr.dz)
    cols = {n: to_dev(np.ascontiguousarray(b32[i][3, 3, :])) for n, i in
            (("rho0", 1), ("theta0", 2), ("pi0", 3), ("qv0", 4))}
    return dict(gr=gr, g=g, p=tuple(f64(a) for a in p32), b=tuple(f64(a) for a in b32), dev=tuple(to_dev(a) for a in p32),
                cols=cols, rng=np.random.default_rng(11))


def _scale(a):
    s = float(np.std(a))
    return s if s > 0 else 1.0


def test_adv_scalar_vjp(st):
    import torch
    from lex_amd import adjoint as adj
    gr, p, b, dv, rng = st["gr"], st["p"], st["b"], st["dev"], st["rng"]
    rho0 = b[1]
    z2 = np.zeros((NX, NY))
    weps = 1e-17

    def F(s, u, v, w):
        fd = o.get_divergence(rho0, u, v, w, gr)
        return (o.advection_scalar(rho0, s, u, v, w, weps, fd, z2, gr),)

    lam = rng.normal(size=(NX, NY, NZ))
    x = (p[0], p[1], p[2], p[3])
    outs = [torch.zeros_like(t) for t in (dv[0], dv[1], dv[2], dv[3])]
    adj.adv_scalar_vjp(st["g"], st["cols"]["rho0"], dv[0], dv[1], dv[2], dv[3], to_dev(lam), weps, *outs)
    grads = [from_dev(t) for t in outs]
    for q, name in enumerate(["s", "u", "v", "w"]):
        vq = [np.zeros_like(a) for a in x]
        vq[q] = rng.normal(size=x[q].shape) * _scale(x[q][3:-3, 3:-3, 1:-1] if q == 0 else x[q])
        _check(_dot(grads[q], vq[q]), _fd(F, x, vq, (lam,)), 1e-4, f"adv_scalar d/d{name}")


@pytest.mark.parametrize("which", ["u", "v", "w"])
def test_adv_momentum_vjp(st, which):
    import torch
    from lex_amd import adjoint as adj
    gr, p, b, dv, rng = st["gr"], st["p"], st["b"], st["dev"], st["rng"]
    rho0 = b[1]

    def F(u, v, w):
        fd = o.get_divergence(rho0, u, v, w, gr)
        if which == "u":
            return (o.advection_u(rho0, u, v, w, 1e-16, fd, np.zeros((NX, NY)), gr),)
        if which == "v":
            return (o.advection_v(rho0, u, v, w, 1e-16, fd, np.zeros((NX, NY)), gr),)
        return (o.advection_w(rho0, u, v, w, 1e-16, fd, gr),)

    x = (p[1], p[2], p[3])
    lam = rng.normal(size=F(*x)[0].shape)
    outs = [torch.zeros_like(t) for t in (dv[1], dv[2], dv[3])]
    adj.adv_momentum_vjp(which, st["g"], st["cols"]["rho0"], dv[1], dv[2], dv[3], to_dev(lam), *outs)
    grads = [from_dev(t) for t in outs]
    for q, name in enumerate(["u", "v", "w"]):
        vq = [np.zeros_like(a) for a in x]
        vq[q] = rng.normal(size=x[q].shape) * _scale(x[q])
        _check(_dot(grads[q], vq[q]), _fd(F, x, vq, (lam,)), 1e-4, f"adv_{which} d/d{name}")


def test_adv_vjp_deep_column():
    """The advection reverses on a 70-level column: deeper than one 64-lane wave, so the
    column-stencil contributions go through slot planes and the gathers (a column of at
    most 64 levels sums them inside its wave with shuffles, which the 12-level tests above
    cover)."""
    import torch
    from lex_amd import adjoint as adj
    sd = _make_state(8, 6, 70, seed=21)
    nx, ny, nz = sd["shape"]
    gr, p, b, dv, rng = sd["gr"], sd["p"], sd["b"], sd["dev"], sd["rng"]
    rho0 = b[1]
    z2 = np.zeros((nx, ny))

    def Fs(s_, u, v, w):
        fd = o.get_divergence(rho0, u, v, w, gr)
        return (o.advection_scalar(rho0, s_, u, v, w, 1e-17, fd, z2, gr),)

    lam = rng.normal(size=(nx, ny, nz))
    x = (p[0], p[1], p[2], p[3])
    outs = [torch.zeros_like(t) for t in (dv[0], dv[1], dv[2], dv[3])]
    adj.adv_scalar_vjp(sd["g"], sd["cols"]["rho0"], dv[0], dv[1], dv[2], dv[3], to_dev(lam), 1e-17, *outs)
    grads = [from_dev(t) for t in outs]
    for q, name in enumerate(["s", "u", "v", "w"]):
        vq = [np.zeros_like(a) for a in x]
        vq[q] = rng.normal(size=x[q].shape) * _scale(x[q][3:-3, 3:-3, 1:-1] if q == 0 else x[q])
        _check(_dot(grads[q], vq[q]), _fd(Fs, x, vq, (lam,)), 1e-4, f"deep adv_scalar d/d{name}")
    for which in ("u", "v", "w"):
        def F(u, v, w):
            fd = o.get_divergence(rho0, u, v, w, gr)
            if which == "u":
                return (o.advection_u(rho0, u, v, w, 1e-16, fd, z2, gr),)
            if which == "v":
                return (o.advection_v(rho0, u, v, w, 1e-16, fd, z2, gr),)
            return (o.advection_w(rho0, u, v, w, 1e-16, fd, gr),)
        xm = (p[1], p[2], p[3])
        lam = rng.normal(size=F(*xm)[0].shape)
        outs = [torch.zeros_like(t) for t in (dv[1], dv[2], dv[3])]
        adj.adv_momentum_vjp(which, sd["g"], sd["cols"]["rho0"], dv[1], dv[2], dv[3], to_dev(lam), *outs)
        grads = [from_dev(t) for t in outs]
        for q, name in enumerate(["u", "v", "w"]):
            vq = [np.zeros_like(a) for a in xm]
            vq[q] = rng.normal(size=xm[q].shape) * _scale(xm[q])
            _check(_dot(grads[q], vq[q]), _fd(F, xm, vq, (lam,)), 1e-4, f"deep adv_{which} d/d{name}")


def test_buoyancy_and_theta_rho_vjp(st):
    import torch
    from lex_amd import adjoint as adj
    gr, p, b, dv, rng = st["gr"], st["p"], st["b"], st["dev"], st["rng"]

    def Fb(th, qv):
        return (o.calculate_buoyancy(b[2], th, b[4], qv)[1],)

    def Ft(th, qv):
        return (o.theta_rho_of(th, qv),)

    x = (p[0], p[5])
    lam_b = rng.normal(size=(NX, NY, NZ + 1))
    lam_t = rng.normal(size=p[0].shape)
    gb = [torch.zeros_like(dv[0]), torch.zeros_like(dv[5])]
    adj.buoyancy_vjp(st["g"], st["cols"]["theta0"], st["cols"]["qv0"], dv[0], dv[5], to_dev(lam_b), *gb)
    gt = [torch.zeros_like(dv[0]), torch.zeros_like(dv[5])]
    adj.theta_rho_vjp(dv[0], dv[5], to_dev(lam_t), *gt)
    for q, name in enumerate(["theta", "qv"]):
        vq = [np.zeros_like(a) for a in x]
        vq[q] = rng.normal(size=x[q].shape) * (1.0 if q == 0 else 1e-3)
        _check(_dot(from_dev(gb[q]), vq[q]), _fd(Fb, x, vq, (lam_b,)), 1e-4, f"b8w d/d{name}")
        _check(_dot(from_dev(gt[q]), vq[q]), _fd(Ft, x, vq, (lam_t,)), 1e-4, f"theta_rho d/d{name}")


def test_acoustic_substep_vjp(st):
    import torch
    from lex_amd import adjoint as adj
    gr, p, b, dv, rng = st["gr"], st["p"], st["b"], st["dev"], st["rng"]
    rho0, pi0 = b[1], b[3]
    dts = 0.25
    ru = rng.normal(0, 0.05, (NX + 1, NY, NZ))
    rv = rng.normal(0, 0.05, (NX, NY + 1, NZ))
    rw = rng.normal(0, 0.05, (NX, NY, NZ + 1))
    ps = o.padding3_array(rng.normal(0, 1e-4, (NX, NY, NZ)))
    tr = o.theta_rho_of(p[0], p[5])
    ru, rv, rw, ps, tr = (f64(f32(a)) for a in (ru, rv, rw, ps, tr))

    def F(u, v, w, pp, ps_, ru_, rv_, rw_, tr_):
        return o.sound_rk_step(u, v, w, pi0, pp, ps_, ru_, rv_, rw_, tr_, rho0, 1, dts, gr)

    x = (p[1], p[2], p[3], p[4], ps, ru, rv, rw, tr)
    outs = F(*x)
    lam = [rng.normal(size=a.shape) for a in outs]
    lam[2] = o.set_w_bc(lam[2])
    lam[3] = lam[3] * 1e4
    xd = [to_dev(a) for a in x]
    lo = [to_dev(a) for a in lam]
    g = [torch.zeros_like(t) for t in xd]
    adj.acoustic_substep_vjp(st["g"], st["cols"]["rho0"], st["cols"]["pi0"], xd[0], xd[1], xd[2], xd[3], xd[4],
                             xd[5], xd[6], xd[7], xd[8], dts, lo[0], lo[1], lo[2], lo[3], g[0], g[1], g[2], g[3], g[4],
                             g[8], g[5], g[6], g[7])
    names = ["u", "v", "w", "pip", "pi*", "rhs_u", "rhs_v", "rhs_w", "theta_rho"]
    for q, name in enumerate(names):
        vq = [np.zeros_like(a) for a in x]
        vq[q] = rng.normal(size=x[q].shape) * _scale(x[q])
        _check(_dot(from_dev(g[q]), vq[q]), _fd(F, x, vq, lam), 1e-4, f"acoustic d/d{name}")


def test_smagorinsky_vjp(st):
    import torch
    from lex_amd import adjoint as adj
    gr, p, b, dv, rng = st["gr"], st["p"], st["b"], st["dev"], st["rng"]
    rho0 = b[1]
    cs, pr = 0.18, 1.0 / 3.0

    def F(u, v, w, th, qv, csv):
        return o.smagorinsky(rho0, u, v, w, th, qv, (th, qv), gr, float(csv[0]), pr)

    x = (p[1], p[2], p[3], p[0], p[5], np.array([cs]))
    outs = F(*x)
    lam = [rng.normal(size=a.shape) for a in outs]
    lam[2][:, :, 0] = 0.0
    lam[2][:, :, -1] = 0.0
    ld = [to_dev(a) for a in lam]
    g = [torch.zeros_like(t) for t in (dv[1], dv[2], dv[3], dv[0], dv[5])]
    lcs = torch.zeros(1, dtype=torch.float64, device=dv[0].device)
    adj.smag_vjp(st["g"], st["cols"]["rho0"], dv[1], dv[2], dv[3], dv[0], dv[5], (dv[0], dv[5]), cs, pr,
                 ld[0], ld[1], ld[2], (ld[3], ld[4]), g[0], g[1], g[2], g[3], g[4], (g[3], g[4]), lcs)
    grads = [from_dev(t) for t in g] + [np.array([float(lcs.item())])]
    for q, name in enumerate(["u", "v", "w", "theta", "qv", "cs"]):
        vq = [np.zeros_like(a) for a in x]
        vq[q] = rng.normal(size=x[q].shape) * (_scale(x[q]) if q < 5 else 1.0)
        _check(_dot(grads[q], vq[q]), _fd(F, x, vq, lam, eps=1e-7), 5e-4, f"smagorinsky d/d{name}")


# ------------------------------------------------------------------ whole step
def _step_setup(nx=16, ny=16, nz=16, seed=3, turb=1):
    from lex_amd.setup_lex import build_case, config2
    case, opts = config2()
    case = case.scaled(nx, ny, nz)
    phys, base, grids, _ = build_case(case)
    og = o.Grid(nx, ny, nz, case.dx, case.dy, case.dz)
    oo = o.Options(dt=opts["dt"], solver_opt=2, n_sound=opts["n_sound"], turb_opt=turb)
    rng = np.random.default_rng(seed)
    # give the state some flow so every term of the step is active
    st = [from_dev(a) for a in phys]
    st[1] = o.padding3u_array(np.concatenate([ui := rng.normal(0, 1.0, (nx, ny, nz)), ui[:1]], 0))
    st[2] = o.padding3v_array(np.concatenate([vi := rng.normal(0, 1.0, (nx, ny, nz)), vi[:, :1]], 1))
    st[3] = o.set_w_bc(o.padding3_array(rng.normal(0, 0.5, (nx, ny, nz + 1))))
    st = [f64(f32(a)) for a in st]
    return case, opts, og, oo, st, base, grids


def _pad_like(q, a_int):
    if q == 1:
        return o.padding3u_array(a_int)
    if q == 2:
        return o.padding3v_array(a_int)
    if q == 3:
        return o.set_w_bc(o.padding3_array(a_int))
    return o.padding3_array(a_int)


def test_step_vjp_vs_fd():
    import torch
    from lex_amd import adjoint as adj, _lib
    from lex_amd._core import grid_from
    case, opts, og, oo, st, base, grids = _step_setup()
    nf = len(st)
    bs = tuple(from_dev(a) for a in base[:5]) + (None,)
    rng = np.random.default_rng(5)

    def F(*xs):
        return o.ws_rk3_big_step(tuple(xs), bs, og, oo)[0]

    out = F(*st)
    lam = [rng.normal(size=a.shape) for a in out]
    lam[4] = lam[4] * 1e3
    g = grid_from(base[1], *grids[3:6])
    dev = tuple(to_dev(a) for a in st)
    li, dcs = adj.step_vjp(dev, tuple(to_dev(a) for a in lam), base, g, opts["dt"], opts["n_sound"], 0.18, 1.0 / 3.0)
    grads = [from_dev(t) for t in li]
    shapes = [(og.nx, og.ny, og.nz), (og.nx + 1, og.ny, og.nz), (og.nx, og.ny + 1, og.nz), (og.nx, og.ny, og.nz + 1),
              (og.nx, og.ny, og.nz), (og.nx, og.ny, og.nz)]
    scales = [0.3, 0.5, 0.5, 0.3, 1e-4, 1e-4]
    for q, name in enumerate(["theta", "u", "v", "w", "pip", "qv"]):
        vi = rng.normal(0, scales[q], shapes[q])
        if q in (1, 2):  # periodic duplicate face carries the same value
            sl = [slice(None)] * 3
            sl[q - 1] = slice(-1, None)
            sl0 = list(sl)
            sl0[q - 1] = slice(0, 1)
            vi[tuple(sl)] = vi[tuple(sl0)]
        vq = [np.zeros_like(a) for a in st]
        vq[q] = _pad_like(q, vi)
        _check(_dot(grads[q], vq[q]), _fd(F, st, vq, lam), 2e-3, f"step d/d{name}")

    # Cs
    def Fc(c):
        oc = o.Options(dt=oo.dt, solver_opt=2, n_sound=oo.n_sound, turb_opt=1, cs=float(c))
        return o.ws_rk3_big_step(tuple(st), bs, og, oc)[0]

    e = 1e-6
    fdc = sum(_dot(l, (a - b) / (2 * e)) for l, a, b in zip(lam, Fc(0.18 + e), Fc(0.18 - e)))
    _check(float(dcs.item()), fdc, 2e-3, "step d/dCs")
    assert nf == len(grads)


def test_step_vjp_is_bit_reproducible():
    """The step's reverse uses no atomics (the advection reverses write slot planes that
    fixed-order gathers add; buoyancy and N^2 reverses gather per cell), so two calls on
    the same inputs give bit-identical adjoints and dL/dCs, on a grid large enough that
    float atomics would reorder their additions from run to run."""
    import torch
    from lex_amd import adjoint as adj
    from lex_amd._core import grid_from
    case, opts, og, oo, st, base, grids = _step_setup(nx=40, ny=32, nz=24, seed=12)
    rng = np.random.default_rng(13)
    g = grid_from(base[1], *grids[3:6])
    dev = tuple(to_dev(a) for a in st)
    lam = tuple(to_dev(rng.normal(size=a.shape)) for a in st)
    runs = [adj.step_vjp(dev, lam, base, g, opts["dt"], opts["n_sound"], 0.18, 1.0 / 3.0) for _ in range(3)]
    torch.cuda.synchronize()
    (li0, dcs0) = runs[0]
    assert all(bool(torch.isfinite(t).all()) for t in li0)
    assert any(float(t.abs().max()) > 0 for t in li0)
    for li, dcs in runs[1:]:
        for q, (a, b) in enumerate(zip(li0, li)):
            assert torch.equal(a, b), f"adjoint of state field {q} differs between runs"
        assert torch.equal(dcs0, dcs)


def test_sprint_gradients_with_checkpoints():
    """Differentiable 4-step run (checkpoint every 2): d loss/d theta'_0 along random
    directions and d loss/d Cs vs FP64 central differences of the oracle; the gradient
    does not depend on the checkpoint spacing."""
    import torch
    from lex_amd import adjoint as adj
    from lex_amd._core import grid_from
    case, opts, og, oo, st, base, grids = _step_setup(seed=4)
    nsteps = 4
    bs = tuple(from_dev(a) for a in base[:5]) + (None,)
    th0 = bs[2]
    rng = np.random.default_rng(2029)
    target = rng.normal(0, 0.5, (og.nx, og.ny, og.nz))
    tp0 = st[0][3:-3, 3:-3, 1:-1] - th0[3:-3, 3:-3, 1:-1]

    def loss_oracle(tp, cs):
        oc = o.Options(dt=oo.dt, solver_opt=2, n_sound=oo.n_sound, turb_opt=1, cs=float(cs))
        x = (o.padding3_array(th0[3:-3, 3:-3, 1:-1] + tp),) + tuple(st[1:])
        for _ in range(nsteps):
            x = o.ws_rk3_big_step(x, bs, og, oc)[0]
        d = x[0][3:-3, 3:-3, 1:-1] - th0[3:-3, 3:-3, 1:-1] - target
        return 0.5 * float(np.sum(d * d))

    g = grid_from(base[1], *grids[3:6])
    state0 = tuple(to_dev(a) for a in st)
    grads = {}
    for k in (2, 1, 4):
        tp = to_dev(tp0).requires_grad_(True)
        cs = torch.tensor(0.18, dtype=torch.float32, device=tp.device, requires_grad=True)
        th = adj.sprint_theta(tp, cs, state0, base, g, opts["dt"], opts["n_sound"], nsteps, k)
        loss = 0.5 * ((th - to_dev(target)) ** 2).sum()
        loss.backward()
        grads[k] = (from_dev(tp.grad), float(cs.grad.item()))
    gth, gcs = grads[2]
    for k in (1, 4):
        a, c = grads[k]
        assert np.linalg.norm(a - gth) <= 1e-5 * np.linalg.norm(gth), k
        assert abs(c - gcs) <= 1e-5 * abs(gcs) + 1e-12, k
    for d in range(2):
        v = rng.normal(0, 1.0, tp0.shape)
        e = 1e-5
        fd = (loss_oracle(tp0 + e * v, 0.18) - loss_oracle(tp0 - e * v, 0.18)) / (2 * e)
        _check(_dot(gth, v), fd, 2e-3, f"sprint d/dtheta' direction {d}")
    e = 1e-5
    fdc = (loss_oracle(tp0, 0.18 + e) - loss_oracle(tp0, 0.18 - e)) / (2 * e)
    _check(gcs, fdc, 5e-3, "sprint d/dCs")


def test_sprint_gradient_along_16_seeded_directions():
    """The differentiable workload's check protocol at a reduced size: a 6-step run with
    a checkpoint every 2 steps, loss 1/2||theta'_6 - target||^2 (target from rng 2029);
    the adjoint's d loss/d theta'_0 against FP64 central differences of the oracle along
    16 seeded random directions (rng 2030)."""
    import torch
    from lex_amd import adjoint as adj
    from lex_amd._core import grid_from
    case, opts, og, oo, st, base, grids = _step_setup(nx=12, ny=12, nz=12, seed=6)
    nsteps, every = 6, 2
    bs = tuple(from_dev(a) for a in base[:5]) + (None,)
    th0 = bs[2]
    target = np.random.default_rng(2029).normal(0, 0.5, (og.nx, og.ny, og.nz))
    tp0 = st[0][3:-3, 3:-3, 1:-1] - th0[3:-3, 3:-3, 1:-1]

    def loss_oracle(tp):
        x = (o.padding3_array(th0[3:-3, 3:-3, 1:-1] + tp),) + tuple(st[1:])
        for _ in range(nsteps):
            x = o.ws_rk3_big_step(x, bs, og, oo)[0]
        d = x[0][3:-3, 3:-3, 1:-1] - th0[3:-3, 3:-3, 1:-1] - target
        return 0.5 * float(np.sum(d * d))

    g = grid_from(base[1], *grids[3:6])
    tp = to_dev(tp0).requires_grad_(True)
    cs = torch.tensor(0.18, dtype=torch.float32, device=tp.device, requires_grad=True)
    th = adj.sprint_theta(tp, cs, tuple(to_dev(a) for a in st), base, g, opts["dt"], opts["n_sound"], nsteps, every)
    (0.5 * ((th - to_dev(target)) ** 2).sum()).backward()
    gth = from_dev(tp.grad)
    rng = np.random.default_rng(2030)
    errs = []
    for d in range(16):
        v = rng.normal(0, 1.0, tp0.shape)
        e = 1e-5
        fd = (loss_oracle(tp0 + e * v) - loss_oracle(tp0 - e * v)) / (2 * e)
        errs.append(_check(_dot(gth, v), fd, 2e-3, f"direction {d}"))
    print(f"16 directions: max rel err {max(errs):.2e}")
