"""Pins of the oracle's v-, w- and pi'-advection (advection.py:77-197, :816-1281) and of
the acoustic substep with a nonzero pi* (one_step_integration.py:519-556), on the CPU.

  * advection_v equals advection_u on the x<->y transposed state (the reference's two
    routines are mirror images: advection.py:37-114, :678-771 vs :816-1025);
  * a constant advected field gives a zero tendency for u, v and w under arbitrary
    advecting winds, when the flow divergence is the one get_divergence computes: this
    fixes every corner / level average of the mass fluxes (a wrong average leaves a
    residual c (div8 - div of the averaged mass flux));
  * advection_w: its scalar-level vertical flux against an independent per-face loop;
    flux-form conservation over the interior w levels; surface and lid tendencies
    exactly zero; mirror symmetry z -> -z, w -> -w with a uniform rho column;
  * advection_pi against the closed form for a linear pi*, uniform winds and a quadratic
    pi0 column, with cc1 != cc2 so the lower and upper w-face terms are told apart;
  * one acoustic substep from rest with uniform forcing and a linear pi* against its
    closed form (the pi' tendency then consists of the pi* advection and the divergence
    of the lid-bounded w).
Each pin is shown to fail under a plausible mistake planted in the oracle's source.
"""
from __future__ import annotations

import inspect
import textwrap

import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import random_state


def _mutant(name, old, new):
    """The oracle function `name` with `old` replaced by `new` in its source, compiled in
    the oracle module's namespace (a planted mistake, for showing that a pin bites)."""
    src = textwrap.dedent(inspect.getsource(getattr(o, name)))
    assert old in src, (name, old)
    ns = dict(vars(o))
    exec(compile(src.replace(old, new), f"<mutant {name}>", "exec"), ns)
    return ns[name]


def _transpose_xy(a):
    return np.ascontiguousarray(np.swapaxes(a, 0, 1))


# ------------------------------------------------------------------ advection_v
def _uv_case(seed=3):
    gr, phys, base = random_state(14, 11, 9, dx=100.0, dy=130.0, dz=70.0, seed=seed)
    th, u, v, w = phys[:4]
    rho0 = base[1]
    fd = o.get_divergence(rho0, u, v, w, gr)
    rng = np.random.default_rng(seed + 50)
    su = rng.normal(0, 0.1, (gr.nx, gr.ny))
    sv = rng.normal(0, 0.1, (gr.nx, gr.ny))
    return gr, rho0, u, v, w, fd, su, sv


def _v_vs_transposed_u():
    gr, rho0, u, v, w, fd, su, sv = _uv_case()
    tv = o.advection_v(rho0, u, v, w, 1e-16, fd, sv, gr)
    grT = o.Grid(gr.ny, gr.nx, gr.nz, gr.dy, gr.dx, gr.dz)
    tuT = o.advection_u(_transpose_xy(rho0), _transpose_xy(v), _transpose_xy(u), _transpose_xy(w), 1e-16,
                        _transpose_xy(fd), _transpose_xy(sv), grT)
    return tv, _transpose_xy(tuT)


def test_advection_v_is_advection_u_transposed():
    tv, tu_t = _v_vs_transposed_u()
    assert tv.shape == tu_t.shape
    np.testing.assert_allclose(tv, tu_t, rtol=1e-10, atol=1e-12 * np.abs(tv).max())


def test_advection_v_transposition_pin_bites(monkeypatch):
    monkeypatch.setattr(o, "horizontal_flux_v", _mutant("horizontal_flux_v", "mu8v = 0.5 * (mu[:, :-1] + mu[:, 1:])",
                                                         "mu8v = mu[:, 1:]"))
    tv, tu_t = _v_vs_transposed_u()
    assert not np.allclose(tv, tu_t, rtol=1e-10, atol=1e-12 * np.abs(tv).max())


def test_advection_v_surface_stress_term_is_transposed_u_term():
    """The bottom-face stress enters v as the y-average of v_sfc_flux (advection.py:83-88)."""
    gr, rho0, u, v, w, fd, su, sv = _uv_case(seed=8)
    t1 = o.advection_v(rho0, u, v, w, 1e-16, fd, sv, gr)
    t0 = o.advection_v(rho0, u, v, w, 1e-16, fd, np.zeros_like(sv), gr)
    d = t1 - t0
    # only the lowest level changes, by the periodic y-average of the flux / (rho8v dz)
    assert not d[:, :, 1:].any()
    ty = np.concatenate([sv[:, -1:], sv, sv[:, :1]], 1)
    avg = 0.5 * (ty[:, :-1] + ty[:, 1:])
    r = rho0[3:-3, :, 1]
    rho8v = 0.5 * (r[:, 2:-3] + r[:, 3:-2])
    np.testing.assert_allclose(d[:, :, 0], avg / gr.dz / rho8v, rtol=1e-9)


# --------------------------------------------------------- constant-field pins
def _linear_rho_case(nx=12, ny=10, nz=10, seed=6):
    """Random winds, and a base density linear in z (so the discrete vertical mass-flux
    divergence of a constant field equals its flow-divergence term)."""
    gr, phys, base = random_state(nx, ny, nz, dx=90.0, dy=110.0, dz=60.0, seed=seed)
    u, v, w = phys[1:4]
    K = np.arange(nz + 2, dtype=np.float64)
    rho0 = np.broadcast_to((1.2 - 0.01 * K)[None, None, :], (nx + 6, ny + 6, nz + 2)).copy()
    return gr, rho0, u, v, w


def _const_u_tendency(gr, rho0, u, v, w, c=4.0):
    uc = np.full_like(u, c)
    fd = o.get_divergence(rho0, uc, v, w, gr)
    return o.advection_u(rho0, uc, v, w, 1e-16, fd, np.zeros((gr.nx, gr.ny)), gr), c


def _const_v_tendency(gr, rho0, u, v, w, c=-3.0):
    vc = np.full_like(v, c)
    fd = o.get_divergence(rho0, u, vc, w, gr)
    return o.advection_v(rho0, u, vc, w, 1e-16, fd, np.zeros((gr.nx, gr.ny)), gr), c


def _const_w_tendency(gr, rho0, u, v, w, c=0.7):
    wc = np.full_like(w, c)
    fd = o.get_divergence(rho0, u, v, wc, gr)
    return o.advection_w(rho0, u, v, wc, 1e-16, fd, gr), c


def test_constant_momentum_fields_have_zero_tendency():
    """u, v or w constant, the other winds random, rho linear in z: zero tendency at every
    level, the ones next to the surface and lid included (the zero-flux face slots and
    get_divergence's zero density beyond the lids are consistent)."""
    gr, rho0, u, v, w = _linear_rho_case()
    scale = 10.0 * max(np.abs(u).max(), np.abs(v).max()) / min(gr.dx, gr.dy)
    for f in (_const_u_tendency, _const_v_tendency, _const_w_tendency):
        t, c = f(gr, rho0, u, v, w)
        assert np.abs(t).max() < 1e-12 * scale * abs(c), (f.__name__, np.abs(t).max())
    # a non-constant field under the same winds has an O(1) tendency
    fd = o.get_divergence(rho0, u, v, w, gr)
    assert np.abs(o.advection_w(rho0, u, v, w, 1e-16, fd, gr)).max() > 1e-3


@pytest.mark.parametrize("func,old,new,which", [
    ("horizontal_flux_w", "mu8w = 0.5 * (mu[:, :, :-1] + mu[:, :, 1:])", "mu8w = mu[:, :, 1:]", "w"),
    ("horizontal_flux_w", "mv8w = 0.5 * (mv[:, :, :-1] + mv[:, :, 1:])", "mv8w = mv[:, :, :-1]", "w"),
    ("horizontal_flux_v", "mu8v = 0.5 * (mu[:, :-1] + mu[:, 1:])", "mu8v = mu[:, 1:]", "v"),
    ("vertical_flux_v2", "m8v = 0.5 * (m[:, :-1] + m[:, 1:])", "m8v = m[:, 1:]", "v"),
    ("horizontal_flux_u", "mv8u = 0.5 * (mv[:-1] + mv[1:])", "mv8u = mv[1:]", "u"),
])
def test_constant_field_pin_bites(monkeypatch, func, old, new, which):
    monkeypatch.setattr(o, func, _mutant(func, old, new))
    gr, rho0, u, v, w = _linear_rho_case()
    t, c = {"u": _const_u_tendency, "v": _const_v_tendency, "w": _const_w_tendency}[which](gr, rho0, u, v, w)
    assert np.abs(t).max() > 1e-6 * abs(c)


# -------------------------------------------------------------- advection_w
def _w_flux_face_loop(rho0, w, weps):
    """vertical_flux_w2 face by face: at every scalar level s (incl. the two ghosts) the
    mass 0.5 (w[s] + w[s+1]) rho0[s] picks the upwind reconstruction of the face between
    w levels s and s+1 from the reference's face table (advection.py:1070-1183)."""
    from tests.test_oracle_weno import _vertical_table_loop
    nxp, nyp, L = w.shape
    out = np.zeros((nxp - 6, nyp - 6, L - 1))
    for i in range(3, nxp - 3):
        for j in range(3, nyp - 3):
            col = w[i, j]
            up = _vertical_table_loop(col, True, weps)
            dn = _vertical_table_loop(col, False, weps)
            for s in range(L - 1):
                m = 0.5 * (col[s] + col[s + 1]) * rho0[i, j, s]
                out[i - 3, j - 3, s] = m * (up[s] if m >= 0.0 else dn[s])
    return out


def test_vertical_flux_w_matches_face_loop():
    gr, phys, base = random_state(7, 6, 11, seed=12)
    w, rho0 = phys[3], base[1]
    np.testing.assert_allclose(o.vertical_flux_w2(1e-16, rho0, w), _w_flux_face_loop(rho0, w, 1e-16),
                               rtol=1e-12, atol=1e-14)


def test_vertical_flux_w_face_loop_pin_bites(monkeypatch):
    mut = _mutant("vertical_flux_w2", "mass = 0.5 * (w[:, :, :-1] + w[:, :, 1:]) * rho0", "mass = w[:, :, :-1] * rho0")
    gr, phys, base = random_state(7, 6, 11, seed=12)
    w, rho0 = phys[3], base[1]
    assert not np.allclose(mut(1e-16, rho0, w), _w_flux_face_loop(rho0, w, 1e-16), rtol=1e-12, atol=1e-14)


def _w_conservation(gr, rho0, u, v, w):
    """sum over interior w levels of (rho8w t - w div8w) dV vs the vertical flux through
    the first and last interior scalar levels (the face loop's fluxes)."""
    fd = o.get_divergence(rho0, u, v, w, gr)
    t = o.advection_w(rho0, u, v, w, 1e-16, fd, gr)
    r = rho0[3:-3, 3:-3, :]
    rho8w = 0.5 * (r[:, :, :-1] + r[:, :, 1:])
    f = fd[3:-3, 3:-3, :]
    div8w = 0.5 * (f[:, :, :-1] + f[:, :, 1:])
    wi = w[3:-3, 3:-3, 1:-1]
    lhs = np.sum((rho8w * t - wi * div8w)[:, :, 1:-1])
    fz = _w_flux_face_loop(rho0, w, 1e-16)
    rhs = -np.sum(fz[:, :, gr.nz] - fz[:, :, 1]) / gr.dz
    return lhs, rhs, np.sum(np.abs(wi * div8w)) + np.sum(np.abs(fz)) / gr.dz


def test_advection_w_flux_form_conservation():
    gr, phys, base = random_state(10, 9, 12, seed=14)
    lhs, rhs, scale = _w_conservation(gr, base[1], *phys[1:4])
    assert abs(lhs - rhs) < 1e-11 * scale, (lhs, rhs)


def test_advection_w_lid_and_surface_exactly_zero():
    gr, phys, base = random_state(10, 9, 12, seed=15)
    rho0, u, v, w = base[1], *phys[1:4]
    fd = o.get_divergence(rho0, u, v, w, gr)
    t = o.advection_w(rho0, u, v, w, 1e-16, fd, gr)
    assert not t[:, :, 0].any() and not t[:, :, -1].any()
    assert np.abs(t[:, :, 1:-1]).min() >= 0.0 and np.abs(t[:, :, 1:-1]).max() > 0.0
    mut = _mutant("advection_w", "t[:, :, -1] = 0.0", "pass")
    assert mut(rho0, u, v, w, 1e-16, fd, gr)[:, :, -1].any()


def test_advection_w_mirror_symmetry_z():
    """z -> -z with w -> -w (uniform rho column): the w tendency flips and changes sign
    (the face table is reflection-symmetric, the surface and lid are treated alike)."""
    gr, phys, base = random_state(8, 7, 12, seed=16)
    u, v, w = phys[1:4]
    rho0 = np.full_like(base[1], 1.1)
    # a state symmetric in construction: every field reflected about mid-column
    um, vm = u[:, :, ::-1].copy(), v[:, :, ::-1].copy()
    wm = -w[:, :, ::-1].copy()
    fd = o.get_divergence(rho0, u, v, w, gr)
    fdm = o.get_divergence(rho0, um, vm, wm, gr)
    t = o.advection_w(rho0, u, v, w, 1e-16, fd, gr)
    tm = o.advection_w(rho0, um, vm, wm, 1e-16, fdm, gr)
    np.testing.assert_allclose(tm, -t[:, :, ::-1], rtol=1e-9, atol=1e-12 * np.abs(t).max())


# ------------------------------------------------------------- advection_pi
def _pi_case(nx=6, ny=5, nz=8, dx=80.0, dy=120.0, dz=50.0):
    gr = o.Grid(nx, ny, nz, dx, dy, dz)
    I, J, K = np.meshgrid(np.arange(nx + 6), np.arange(ny + 6), np.arange(nz + 2), indexing="ij")
    a, b, c = 2e-6, -3e-6, 5e-6
    pip = a * I * dx + b * J * dy + c * (K - 0.5) * dz
    zc = (np.arange(nz + 2) - 0.5) * dz
    p1, p2 = -3.4e-5, 2.0e-9
    pi0 = np.broadcast_to((1.0 + p1 * zc + p2 * zc * zc)[None, None, :], (nx + 6, ny + 6, nz + 2)).copy()
    rcol = 1.2 * np.exp(-zc / 8000.0)
    rho0 = np.broadcast_to(rcol[None, None, :], (nx + 6, ny + 6, nz + 2)).copy()
    U, V, W = 3.0, -2.0, 0.8
    u = np.full((nx + 7, ny + 6, nz + 2), U)
    v = np.full((nx + 6, ny + 7, nz + 2), V)
    w = np.full((nx + 6, ny + 6, nz + 3), W)
    return gr, dict(a=a, b=b, c=c, p1=p1, p2=p2, zc=zc, rcol=rcol, U=U, V=V, W=W), rho0, pi0, pip, u, v, w


def _pi_closed_form(gr, P, cc1, cc2):
    """-(U a + V b + [cc1 W (dpi0/dz + c) rho8w below + cc2 W (...) rho8w above] / rho)
    per level: w-level k lies between padded scalar levels k and k+1; its vertical term is
    dropped at the surface (k = 0) and the lid (k = nz)."""
    nz, dz = gr.nz, gr.dz
    zc, rcol = P["zc"], P["rcol"]
    out = np.empty(nz)
    for k in range(nz):
        terms = []
        for lev, wt in ((k, cc1), (k + 1, cc2)):      # the cell's lower and upper w levels
            if lev in (0, nz):
                terms.append(0.0)
                continue
            dpi0 = P["p1"] + P["p2"] * (zc[lev] + zc[lev + 1])    # exact difference quotient of a quadratic
            rho8w = 0.5 * (rcol[lev] + rcol[lev + 1])
            terms.append(wt * P["W"] * (dpi0 + P["c"]) * rho8w)
        out[k] = -(P["U"] * P["a"] + P["V"] * P["b"] + (terms[0] + terms[1]) / rcol[k + 1])
    return out


def test_advection_pi_closed_form():
    gr, P, rho0, pi0, pip, u, v, w = _pi_case()
    for cc1, cc2 in ((0.5, 0.5), (0.3, 0.7)):
        t = o.advection_pi(rho0, pi0, pip, u, v, w, gr, cc1, cc2)
        ref = _pi_closed_form(gr, P, cc1, cc2)
        np.testing.assert_allclose(t, np.broadcast_to(ref, t.shape), rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("old,new", [
    ("(cc1 * wdz[:, :, :-1] * rho8w[:, :, :-1] + cc2 * wdz[:, :, 1:] * rho8w[:, :, 1:])",
     "(cc2 * wdz[:, :, :-1] * rho8w[:, :, :-1] + cc1 * wdz[:, :, 1:] * rho8w[:, :, 1:])"),   # swapped weights
    ("wdz[:, :, -1] = 0.0", "pass"),                                                       # lid not zeroed
    ("wdz[:, :, 0] = 0.0", "pass"),                                                        # surface not zeroed
    ("rho8w = 0.5 * (r[:, :, :-1] + r[:, :, 1:])", "rho8w = r[:, :, 1:]"),                 # wrong level average
    ("dpi_dz = dpi0_dz + dpip_dz", "dpi_dz = dpip_dz"),                                    # pi0 dropped
])
def test_advection_pi_closed_form_pin_bites(old, new):
    gr, P, rho0, pi0, pip, u, v, w = _pi_case()
    mut = _mutant("advection_pi", old, new)
    t = mut(rho0, pi0, pip, u, v, w, gr, 0.3, 0.7)
    ref = _pi_closed_form(gr, P, 0.3, 0.7)
    assert not np.allclose(t, np.broadcast_to(ref, t.shape), rtol=1e-9, atol=1e-15)


# --------------------------------------------------------- acoustic substep
def _substep_closed_form_case():
    gr, P, rho0, pi0, pistar, _, _, _ = _pi_case(nx=7, ny=6, nz=9)
    U0, V0, W0 = 0.4, -0.3, 0.25        # uniform forcing (rhs of u, v, w)
    dts = 0.2
    th = 301.0
    u = np.zeros((gr.nx + 7, gr.ny + 6, gr.nz + 2))
    v = np.zeros((gr.nx + 6, gr.ny + 7, gr.nz + 2))
    w = np.zeros((gr.nx + 6, gr.ny + 6, gr.nz + 3))
    p = np.zeros(gr.shape_c)
    ru = np.full((gr.nx + 1, gr.ny, gr.nz), U0)
    rv = np.full((gr.nx, gr.ny + 1, gr.nz), V0)
    rw = np.full((gr.nx, gr.ny, gr.nz + 1), W0)
    trho = np.full(gr.shape_c, th)
    args = (u, v, w, pi0, p, pistar, ru, rv, rw, trho, rho0, 1, dts, gr)
    # closed form of one substep from rest: u' = dts U0, v' = dts V0, w' = dts W0 except
    # at the surface and lid levels (set_w_bc); D = dw/dz is nonzero only in the first and
    # last cell; pi' = 1.1 dts (-Rd/Cv (pi0 + pi*) D + adv_pi(u', v', w'; pi*))
    P2 = dict(P, U=dts * U0, V=dts * V0, W=dts * W0)
    apt = _pi_closed_form(gr, P2, 0.5, 0.5)
    D = np.zeros(gr.nz)
    D[0] = dts * W0 / gr.dz
    D[-1] = -dts * W0 / gr.dz
    piadv = o._interior(pi0 + pistar)
    pexp = 1.1 * dts * (-(o.Rd / o.Cv) * piadv * D[None, None, :] + apt[None, None, :])
    return args, dict(u=dts * U0, v=dts * V0, w=dts * W0, p=pexp)


def test_acoustic_substep_with_linear_pistar_closed_form():
    args, exp = _substep_closed_form_case()
    un, vn, wn, pn = o.sound_rk_step(*args)
    np.testing.assert_allclose(un, exp["u"], rtol=1e-12)
    np.testing.assert_allclose(vn, exp["v"], rtol=1e-12)
    wi = wn[3:-3, 3:-3, 1:-1]
    np.testing.assert_allclose(wi[:, :, 1:-1], exp["w"], rtol=1e-12)
    assert not wi[:, :, [0, -1]].any()
    np.testing.assert_allclose(o._interior(pn), exp["p"], rtol=1e-9, atol=1e-18)
    # pi* matters: the same substep with pi* = 0 differs
    args0 = list(args)
    args0[5] = np.zeros_like(args[5])
    _, _, _, pn0 = o.sound_rk_step(*args0)
    assert not np.allclose(o._interior(pn0), exp["p"], rtol=1e-6)


def test_acoustic_substep_closed_form_pin_bites(monkeypatch):
    """Planted mistakes: pi* left out of the divergence coefficient; advection_pi fed the
    new pi' instead of pi*."""
    args, exp = _substep_closed_form_case()
    for old, new in (("piadv = pi0 + pip_star", "piadv = pi0"),
                     ("apt = advection_pi(rho0, pi0, pip_star,", "apt = advection_pi(rho0, pi0, pip_now,")):
        mut = _mutant("sound_rk_step", old, new)
        _, _, _, pn = mut(*args)
        assert not np.allclose(o._interior(pn), exp["p"], rtol=1e-9, atol=1e-18), old
