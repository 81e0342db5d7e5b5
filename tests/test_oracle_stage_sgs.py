"""Pins of the oracle's learned-SGS data paths and of the step with the SGS recomputed in
every stage, on the CPU.

  * ws_rk3_big_step_stage_sgs with a state-independent SGS function is the frozen
    step ws_rk3_big_step_ext_sgs, bit for bit;
  * a uniform wind damped by sgs_u = -k u over a dry rest state advances by the WS-RK3
    polynomial of -k dt (1 - k dt + (k dt)^2 / 2 - (k dt)^3 / 6); the frozen step gives
    1 - k dt, and a stage function that is handed the step's initial state in every stage
    is rejected;
  * sgs_apply: a sinusoid sampled at cell centres comes out on the faces as
    cos(pi / n) times the sinusoid at the faces (the mean of the samples half a cell
    either side, periodic in x and y), and w is zero at the surface and lid;
  * sgs_features: for a wind linear in x, y, z the centre values and the deformation
    components are the closed forms (S11 = du/dx, S12 = (du/dy + dv/dx) / 2, ...);
each pin is shown to fail under a planted mistake.
"""
from __future__ import annotations

import inspect
import textwrap

import numpy as np

from oracle import lex_oracle as o


def _mutant(name, old, new):
    src = textwrap.dedent(inspect.getsource(getattr(o, name)))
    assert old in src, (name, old)
    ns = dict(vars(o))
    exec(compile(src.replace(old, new), f"<mutant {name}>", "exec"), ns)
    return ns[name]


def _interior_shapes(gr):
    return ((gr.nx + 1, gr.ny, gr.nz), (gr.nx, gr.ny + 1, gr.nz), (gr.nx, gr.ny, gr.nz + 1),
            (gr.nx, gr.ny, gr.nz), (gr.nx, gr.ny, gr.nz))


def test_state_independent_stage_sgs_is_the_frozen_step():
    gr = o.Grid(10, 8, 8, 200.0, 200.0, 200.0)
    case = o.BubbleCase(gr, amplitude=2.0)
    phys, base, _ = o.setup_warm_bubble(case)
    opts = o.Options(dt=2.0, solver_opt=2, n_sound=6, turb_opt=1)
    rng = np.random.default_rng(4)
    sgs = tuple(rng.normal(0, 1e-3, s) for s in _interior_shapes(gr))
    a, _ = o.ws_rk3_big_step_ext_sgs(phys[:6], base, gr, opts, sgs)
    b, _ = o.ws_rk3_big_step_stage_sgs(phys[:6], base, gr, opts, lambda s, xi: sgs)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def _damped(k, dt, step):
    gr = o.Grid(8, 8, 8, 200.0, 200.0, 200.0)
    case = o.BubbleCase(gr, amplitude=0.0)
    phys, base, _ = o.setup_warm_bubble(case)
    th, u, v, w, pip, qv = phys[:6]
    u0 = 1.7
    u = np.full_like(u, u0)
    opts = o.Options(dt=dt, solver_opt=2, n_sound=6, turb_opt=1)
    zero = [np.zeros(s) for s in _interior_shapes(gr)]

    def fn(stage, xi):
        return (-k * xi[1][3:-3, 3:-3, 1:-1],) + tuple(zero[1:])
    xi, _ = step((th, u, v, w, pip, qv), base, gr, opts, fn)
    return xi, u0


def test_linear_damping_follows_the_ws_rk3_polynomial():
    k, dt = 1e-2, 20.0
    xi, u0 = _damped(k, dt, o.ws_rk3_big_step_stage_sgs)
    z = -k * dt
    np.testing.assert_allclose(xi[1], u0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-12)
    assert np.abs(xi[3]).max() < 1e-12
    # the frozen step: 1 + z only
    frozen = lambda x0, b, gr, op, fn: o.ws_rk3_big_step_ext_sgs(x0, b, gr, op, fn(0, x0))  # noqa: E731
    xf, _ = _damped(k, dt, frozen)
    np.testing.assert_allclose(xf[1], u0 * (1 + z), rtol=1e-12)
    # a stage function handed the step's initial state in every stage is the frozen step
    mut = _mutant("ws_rk3_big_step_stage_sgs", "tuple(sgs_fn(1, xi1))", "tuple(sgs_fn(1, xi0))")
    xm, _ = _damped(k, dt, mut)
    assert not np.allclose(xm[1], u0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-6)


def test_sgs_apply_sinusoid_closed_form():
    nx, ny, nz = 12, 10, 8
    y = np.zeros((nx, ny, nz, 5))
    xc = (np.arange(nx) + 0.5) / nx
    yc = (np.arange(ny) + 0.5) / ny
    zc = (np.arange(nz) + 0.5) / nz
    y[..., 0] = np.sin(2 * np.pi * xc)[:, None, None]
    y[..., 1] = np.cos(2 * np.pi * yc)[None, :, None]
    y[..., 2] = np.cos(2 * np.pi * zc)[None, None, :]
    y[..., 3] = 2.0
    y[..., 4] = -1.0
    scale = (2.0, 3.0, 0.5, 1.5, 0.25)
    su, sv, sw, sth, sqv = o.sgs_apply(y, scale)
    assert su.shape == (nx + 1, ny, nz) and sv.shape == (nx, ny + 1, nz) and sw.shape == (nx, ny, nz + 1)
    xf = np.arange(nx + 1) / nx
    yf = np.arange(ny + 1) / ny
    zf = np.arange(nz + 1) / nz
    np.testing.assert_allclose(su, 2.0 * np.cos(np.pi / nx) * np.sin(2 * np.pi * xf)[:, None, None]
                               * np.ones((1, ny, nz)), atol=1e-14)
    np.testing.assert_allclose(sv, 3.0 * np.cos(np.pi / ny) * np.cos(2 * np.pi * yf)[None, :, None]
                               * np.ones((nx, 1, nz)), atol=1e-14)
    wref = 0.5 * np.cos(np.pi / nz) * np.cos(2 * np.pi * zf)
    wref[0] = wref[-1] = 0.0
    np.testing.assert_allclose(sw, wref[None, None, :] * np.ones((nx, ny, 1)), atol=1e-14)
    np.testing.assert_array_equal(sth, 3.0)
    np.testing.assert_array_equal(sqv, -0.25)
    # faces averaging cells i and i+1 (half a cell off) fail the closed form
    mut = _mutant("sgs_apply", "ext_x = np.concatenate([t[0][-1:], t[0]], 0)",
                  "ext_x = np.concatenate([t[0], t[0][:1]], 0)")
    assert not np.allclose(mut(y, scale)[0][:nx], su[:nx], atol=1e-3)


def test_sgs_features_linear_wind_closed_form():
    gr = o.Grid(8, 7, 6, 50.0, 70.0, 40.0)
    case = o.BubbleCase(gr, amplitude=0.0)
    phys, base, _ = o.setup_warm_bubble(case)
    a, b, c, d, e, f, gg, h = 0.3, -0.2, 0.05, 0.7, 0.11, -0.04, 0.02, -0.03
    P = lambda n: np.arange(n)[:, None, None]   # noqa: E731
    Jy = lambda n: np.arange(n)[None, :, None]  # noqa: E731
    Kz = lambda n: np.arange(n)[None, None, :]  # noqa: E731
    # padded index -> coordinate: x centre (P - 3 + 1/2) dx, x face (P - 3) dx; z centre
    # (K - 1 + 1/2) dz, w level (K - 1) dz
    xc = lambda p: (p - 3 + 0.5) * gr.dx   # noqa: E731
    xf = lambda p: (p - 3) * gr.dx         # noqa: E731
    yc = lambda q: (q - 3 + 0.5) * gr.dy   # noqa: E731
    yf = lambda q: (q - 3) * gr.dy         # noqa: E731
    zc = lambda r: (r - 1 + 0.5) * gr.dz   # noqa: E731
    u = a * xf(P(gr.nx + 7)) + b * yc(Jy(gr.ny + 6)) + c * zc(Kz(gr.nz + 2))
    v = d * xc(P(gr.nx + 6)) + e * yf(Jy(gr.ny + 7)) + f * zc(Kz(gr.nz + 2))
    w = gg * xc(P(gr.nx + 6)) + h * yc(Jy(gr.ny + 6)) + 0.0 * Kz(gr.nz + 3)
    th = phys[0] + 0.3
    qv = phys[5] - 1e-4
    st = (th, u, v, w, phys[4], qv)
    feat = o.sgs_features(st, base, gr)
    assert feat.shape == (gr.nx, gr.ny, gr.nz, 11)
    ic, jc, kc = np.meshgrid(np.arange(gr.nx) + 3, np.arange(gr.ny) + 3, np.arange(gr.nz) + 1, indexing="ij")
    ref = [a * xc(ic) + b * yc(jc) + c * zc(kc), d * xc(ic) + e * yc(jc) + f * zc(kc), gg * xc(ic) + h * yc(jc),
           np.full(ic.shape, 0.3), np.full(ic.shape, -1e-4), np.full(ic.shape, a), np.full(ic.shape, e),
           np.zeros(ic.shape), np.full(ic.shape, 0.5 * (b + d)), np.full(ic.shape, 0.5 * (c + gg)),
           np.full(ic.shape, 0.5 * (f + h))]
    for ch, r in enumerate(ref):
        np.testing.assert_allclose(feat[..., ch], r, rtol=1e-10, atol=1e-12, err_msg=f"channel {ch}")
    # the perturbation taken against rho0 instead of theta0 fails channel 3
    mut = _mutant("sgs_features", "th[I] - base[2][I]", "th[I] - base[1][I]")
    assert not np.allclose(mut(st, base, gr)[..., 3], 0.3)
