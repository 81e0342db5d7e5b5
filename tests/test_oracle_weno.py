"""Pins of the oracle's WENO reconstructions and advection operators (CPU).

The reference ships no tests or vectors, so each piece is checked against what the
mathematics fixes, independently of the oracle's own formulas:
  * Jiang-Shu smoothness indicators against their integral definition (numerical
    quadrature of the fitted candidate polynomials);
  * candidates against cell-average reconstruction of quadratics (exactness);
  * ideal weights d_k against 5th-order convergence on smooth data;
  * the non-oscillatory (ENO) weighting at a discontinuity;
  * the vertical near-boundary table against a per-face loop from the reference's
    face list (advection.py:330-409) and against reflection symmetry;
  * advection operators: constant fields, flux-form conservation and mirror symmetry.
"""
import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import random_state


def _cell_avg_poly(coef, xl, xr):
    """Average of the polynomial sum coef[k] x^k over [xl, xr]."""
    P = np.polynomial.polynomial
    integ = P.polyint(coef)
    return (P.polyval(xr, integ) - P.polyval(xl, integ)) / (xr - xl)


def test_candidates_exact_for_quadratics():
    rng = np.random.default_rng(0)
    for _ in range(50):
        coef = rng.normal(size=3)
        h = 0.7
        centers = (np.arange(5) - 2) * h          # cells i-2..i+2, upwind cell i at 0
        avgs = [_cell_avg_poly(coef, c - h / 2, c + h / 2) for c in centers]
        exact = np.polynomial.polynomial.polyval(h / 2, coef)   # face i+1/2
        _, qs, _ = o.weno5_parts(*avgs, 1e-16)
        for q in qs:
            assert abs(q - exact) < 1e-12 * max(1, abs(exact))
        assert abs(o.weno5(*avgs, 1e-16) - exact) < 1e-12 * max(1, abs(exact))


def test_smoothness_indicators_match_integral_definition():
    """beta_k = sum_{l=1,2} h^(2l-1) int_cell (d^l p_k/dx^l)^2 over the upwind cell."""
    P = np.polynomial.polynomial
    rng = np.random.default_rng(1)
    h = 1.0
    for _ in range(20):
        v = rng.normal(size=5)
        betas, _, _ = o.weno5_parts(*v, 1e-16)
        for k, cells in enumerate([(0, 1, 2), (1, 2, 3), (2, 3, 4)]):
            # quadratic whose cell averages over cells (centres (j-2)h) match v[cells]
            A = np.array([[_cell_avg_poly(np.eye(3)[m], (j - 2) * h - h / 2, (j - 2) * h + h / 2) for m in range(3)]
                          for j in cells])
            coef = np.linalg.solve(A, v[list(cells)])
            ref = 0.0
            for lvl in (1, 2):
                d = P.polyder(coef, lvl)
                sq = P.polymul(d, d)
                iq = P.polyint(sq)
                ref += h ** (2 * lvl - 1) * (P.polyval(h / 2, iq) - P.polyval(-h / 2, iq))
            assert abs(betas[k] - ref) < 1e-10 * max(1.0, abs(ref)), (k, betas[k], ref)


def test_linear_data_gives_ideal_weights_and_midpoint():
    rng = np.random.default_rng(2)
    for _ in range(20):
        a0, s = rng.normal(size=2)
        v = a0 + s * np.arange(5)
        _, _, ws = o.weno5_parts(*v, 1e-16)
        wsum = sum(ws)
        for w, d in zip(ws, (0.1, 0.6, 0.3)):
            assert abs(w / wsum - d) < 1e-12
        assert abs(o.weno5(*v, 1e-16) - 0.5 * (v[2] + v[3])) < 1e-12 * (1 + abs(v[2]))
        assert abs(o.weno3(v[1], v[2], v[3], 1e-16) - 0.5 * (v[2] + v[3])) < 1e-12 * (1 + abs(v[2]))


def test_fifth_order_convergence_on_smooth_data():
    """Ideal weights d_k = (0.1, 0.6, 0.3) make the face value 5th-order accurate."""
    errs = []
    for n in (20, 40, 80):
        h = 2 * np.pi / n / 4
        x0 = 0.3
        centers = x0 + (np.arange(5) - 2) * h
        avgs = [(np.cos(c - h / 2) - np.cos(c + h / 2)) / h for c in centers]   # averages of sin
        errs.append(abs(o.weno5(*avgs, 1e-40) - np.sin(x0 + h / 2)))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > 4.5, rates


def test_eno_property_at_a_jump():
    v = np.array([0.0, 0.0, 0.0, 1.0, 1.0])
    _, _, ws = o.weno5_parts(*v, 1e-16)
    assert ws[0] / sum(ws) > 0.999            # the smooth stencil takes the weight
    assert abs(o.weno5(*v, 1e-16)) < 1e-3      # no overshoot on the upwind side
    assert abs(o.weno3(0.0, 0.0, 1.0, 1e-16)) < 1e-3


def _vertical_table_loop(col, up, weps):
    """Per-face loop of the reference's vertical face list (advection.py:330-409)."""
    L = len(col)
    out = np.zeros(L - 1)
    for f in range(L - 1):
        if f in (0, L - 2):
            continue
        if up:
            if f == 1:
                out[f] = 0.5 * (col[1] + col[2])
            elif f == 2:
                out[f] = o.weno3(col[1], col[2], col[3], weps)
            elif f == L - 3:
                out[f] = o.weno3(col[L - 4], col[L - 3], col[L - 2], weps)
            else:
                out[f] = o.weno5(*col[f - 2:f + 3], weps)
        else:
            if f == 1:
                out[f] = o.weno3(col[3], col[2], col[1], weps)
            elif f == L - 3:
                out[f] = 0.5 * (col[L - 2] + col[L - 3])
            elif f == L - 4:
                out[f] = o.weno3(col[L - 2], col[L - 3], col[L - 4], weps)
            else:
                out[f] = o.weno5(*col[f + 3:f - 2:-1] if f - 2 >= 0 else col[f + 3::-1], weps)
    return out


@pytest.mark.parametrize("L", [8, 9, 14])
def test_vertical_table_matches_face_loop(L):
    rng = np.random.default_rng(L)
    s = rng.normal(size=(2, 3, L))
    up, dn = o.vertical_faces(s, 1e-16)
    for i in range(2):
        for j in range(3):
            np.testing.assert_allclose(up[i, j], _vertical_table_loop(s[i, j], True, 1e-16), rtol=1e-13, atol=1e-15)
            np.testing.assert_allclose(dn[i, j], _vertical_table_loop(s[i, j], False, 1e-16), rtol=1e-13, atol=1e-15)


def test_vertical_table_reflection_symmetry():
    """Flipping the column maps downward reconstructions onto upward ones."""
    rng = np.random.default_rng(9)
    s = rng.normal(size=(1, 1, 12))
    up, dn = o.vertical_faces(s, 1e-16)
    up_f, dn_f = o.vertical_faces(s[:, :, ::-1].copy(), 1e-16)
    np.testing.assert_allclose(dn[0, 0], up_f[0, 0, ::-1], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(up[0, 0], dn_f[0, 0, ::-1], rtol=1e-13, atol=1e-15)


def _state():
    gr, phys, base = random_state(12, 10, 10, seed=4, passive=0)
    return gr, phys, base


def test_constant_field_has_zero_tendency():
    gr, phys, base = _state()
    th, u, v, w = phys[:4]
    rho0 = base[1]
    fd = o.get_divergence(rho0, u, v, w, gr)
    c = np.full_like(th, 7.5)
    z2 = np.zeros((gr.nx, gr.ny))
    t = o.advection_scalar(rho0, c, u, v, w, 1e-17, fd, z2, gr)
    assert np.max(np.abs(t)) < 1e-9 * 7.5 * np.max(np.abs(u)) / gr.dx * 10


def test_flux_form_conservation():
    """sum rho0 tend dV = sum s div(rho0 u) dV with periodic sides and closed lids."""
    gr, phys, base = _state()
    th, u, v, w, _, qv = phys[:6]
    rho0 = base[1]
    fd = o.get_divergence(rho0, u, v, w, gr)
    z2 = np.zeros((gr.nx, gr.ny))
    for s in (th, qv):
        t = o.advection_scalar(rho0, s, u, v, w, 1e-17, fd, z2, gr)
        lhs = np.sum(o._interior(rho0) * t)
        rhs = np.sum(o._interior(s) * o._interior(fd))
        assert abs(lhs - rhs) < 1e-10 * np.sum(np.abs(o._interior(s) * o._interior(fd)))
    # momentum: sum over u faces 0..nx-1 (face nx duplicates face 0)
    tu = o.advection_u(rho0, u, v, w, 1e-16, fd, z2, gr)
    r = o._interior(rho0)
    fdi = fd[:, 3:-3, 1:-1]
    div8u = 0.5 * (fdi[2:-3] + fdi[3:-2])
    lhs = np.sum(r * tu[:-1])
    rhs = np.sum(u[3:-4, 3:-3, 1:-1] * div8u[:-1])
    assert abs(lhs - rhs) < 1e-9 * (np.sum(np.abs(u[3:-4, 3:-3, 1:-1] * div8u[:-1])) + 1e-30)


def test_mirror_symmetry_x():
    """x -> -x with u -> -u mirrors the scalar and u tendencies (upwind orientation)."""
    gr, phys, base = _state()
    th, u, v, w = phys[:4]
    rho0 = base[1]
    z2 = np.zeros((gr.nx, gr.ny))
    thi = o._interior(th)[::-1].copy()
    ui = -u[3:-3, 3:-3, 1:-1][::-1].copy()      # faces nx..0 -> 0..nx
    vi = v[3:-3, 3:-3, 1:-1][::-1].copy()
    wi = w[3:-3, 3:-3, 1:-1][::-1].copy()
    thm = o.padding3_array(thi)
    um = o.padding3u_array(ui)
    vm = o.padding3v_array(vi)
    wm = o.set_w_bc(o.padding3_array(wi))
    fd = o.get_divergence(rho0, u, v, w, gr)
    fdm = o.get_divergence(rho0, um, vm, wm, gr)
    t = o.advection_scalar(rho0, th, u, v, w, 1e-17, fd, z2, gr)
    tm = o.advection_scalar(rho0, thm, um, vm, wm, 1e-17, fdm, z2, gr)
    np.testing.assert_allclose(tm, t[::-1], rtol=1e-9, atol=1e-12)
    tu = o.advection_u(rho0, u, v, w, 1e-16, fd, z2, gr)
    tum = o.advection_u(rho0, um, vm, wm, 1e-16, fdm, z2, gr)
    np.testing.assert_allclose(tum, -tu[::-1], rtol=1e-9, atol=1e-12)


def test_horizontal_advection_converges_to_analytic():
    """Uniform u, rho: -u ds/dx is reproduced at 5th order for a smooth periodic field."""
    errs = []
    for nx in (16, 32):
        gr = o.Grid(nx, 6, 6, 1.0 / nx, 1.0, 1.0)
        x = (np.arange(nx) + 0.5) * gr.dx
        # cell averages of sin(2 pi x)
        sa = (np.cos(2 * np.pi * (x - gr.dx / 2)) - np.cos(2 * np.pi * (x + gr.dx / 2))) / (2 * np.pi * gr.dx)
        s = o.padding3_array(np.broadcast_to(sa[:, None, None], (nx, 6, 6)).copy())
        rho0 = np.ones(gr.shape_c)
        u = np.full(gr.shape_u, 0.8)
        v = np.zeros(gr.shape_v)
        w = np.zeros(gr.shape_w)
        fd = o.get_divergence(rho0, u, v, w, gr)
        t = o.advection_scalar(rho0, s, u, v, w, 1e-40, fd, np.zeros((nx, 6)), gr)[:, 2, 2]
        exact = -0.8 * (np.sin(2 * np.pi * (x + gr.dx / 2)) - np.sin(2 * np.pi * (x - gr.dx / 2))) / gr.dx
        errs.append(np.max(np.abs(t - exact)))
    assert np.log2(errs[0] / errs[1]) > 4.5, errs
