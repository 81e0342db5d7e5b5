"""Pins of the oracle's Coriolis options (pressure_gradient_coriolis.py:80-113) and of
their place in the stage (one_step_integration.py:83-89, :126-131), on the CPU.

  * uniform wind: fu = f u, fv = f v (cor_opt 1) and fu = f (u - 1), fv = f v (cor_opt 2,
    geostrophic wind (1, 0) m/s); the geostrophic wind itself feels no force;
  * a wind linear in x and y: the four-point means equal the wind at the other grid's
    points exactly (fixes the stencil's offsets);
  * a whole WS-RK3 step of a uniform wind over a dry rest state: the stages only see the
    Coriolis rhs, so (u, v) - (u_g, v_g) is multiplied by 1 + A + A^2/2 + A^3/6,
    A = f dt [[0, 1], [-1, 0]] (the WS-RK3 polynomial of the f-plane rotation);
each pin is shown to fail under a planted mistake.
"""
from __future__ import annotations

import inspect
import textwrap

import numpy as np
import pytest

from oracle import lex_oracle as o


def _mutant(name, old, new):
    src = textwrap.dedent(inspect.getsource(getattr(o, name)))
    assert old in src, (name, old)
    ns = dict(vars(o))
    exec(compile(src.replace(old, new), f"<mutant {name}>", "exec"), ns)
    return ns[name]


def _grid():
    return o.Grid(8, 7, 6, 50.0, 70.0, 40.0)


def test_uniform_wind_closed_form_both_options():
    gr = _grid()
    u = np.full(gr.shape_u, 3.0)
    v = np.full(gr.shape_v, -2.0)
    fu, fv = o.calculate_coriolis_force(u, v)
    np.testing.assert_allclose(fu, o.fCor * 3.0, rtol=1e-15)
    np.testing.assert_allclose(fv, o.fCor * -2.0, rtol=1e-15)
    fu, fv = o.calculate_coriolis_force_with_lspgrad(u, v)
    np.testing.assert_allclose(fu, o.fCor * (3.0 - 1.0), rtol=1e-15)
    np.testing.assert_allclose(fv, o.fCor * -2.0, rtol=1e-15)
    assert fu.shape == (gr.nx, gr.ny + 1, gr.nz) and fv.shape == (gr.nx + 1, gr.ny, gr.nz)
    # the geostrophic wind is in balance with the large-scale pressure gradient
    fu, fv = o.calculate_coriolis_force_with_lspgrad(np.ones(gr.shape_u), np.zeros(gr.shape_v))
    assert not fu.any() and not fv.any()


def test_linear_wind_corner_means_are_exact():
    """u = a x + b y at u points, v = c x + d y at v points: fv at the u point (x_{i-1/2},
    y_j) is f v(x_{i-1/2}, y_j), fu at the v point (x_i, y_{j-1/2}) is f u(x_i, y_{j-1/2})."""
    gr = _grid()
    a, b, c, dd = 0.3, -0.2, 0.7, 0.11
    # padded index P along x: scalar centre x = (P - 3 + 0.5) dx, u face x = (P - 3) dx
    Pu = np.arange(gr.nx + 7)[:, None, None]
    Ju = np.arange(gr.ny + 6)[None, :, None]
    u = a * (Pu - 3) * gr.dx + b * (Ju - 3 + 0.5) * gr.dy + 0 * np.arange(gr.nz + 2)
    Pv = np.arange(gr.nx + 6)[:, None, None]
    Jv = np.arange(gr.ny + 7)[None, :, None]
    v = c * (Pv - 3 + 0.5) * gr.dx + dd * (Jv - 3) * gr.dy + 0 * np.arange(gr.nz + 2)
    fu, fv = o.calculate_coriolis_force(u, v)
    iu_, ju_ = np.meshgrid(np.arange(gr.nx + 1), np.arange(gr.ny), indexing="ij")
    v_at_u = c * iu_ * gr.dx + dd * (ju_ + 0.5) * gr.dy
    np.testing.assert_allclose(fv, o.fCor * v_at_u[:, :, None] * np.ones(gr.nz), rtol=1e-12, atol=1e-18)
    iv_, jv_ = np.meshgrid(np.arange(gr.nx), np.arange(gr.ny + 1), indexing="ij")
    u_at_v = a * (iv_ + 0.5) * gr.dx + b * jv_ * gr.dy
    np.testing.assert_allclose(fu, o.fCor * u_at_v[:, :, None] * np.ones(gr.nz), rtol=1e-12, atol=1e-18)


def _inertial_step(cor_opt, u0, v0, dt=20.0):
    """One WS-RK3 step of a uniform wind over the dry rest state of a small grid."""
    gr = o.Grid(8, 8, 8, 200.0, 200.0, 200.0)
    case = o.BubbleCase(gr, amplitude=0.0)
    phys, base, _ = o.setup_warm_bubble(case)
    th, u, v, w, pip, qv = phys[:6]
    u = np.full_like(u, u0)
    v = np.full_like(v, v0)
    opts = o.Options(dt=dt, solver_opt=2, n_sound=6, turb_opt=1, cor_opt=cor_opt)
    xi, _ = o.ws_rk3_big_step((th, u, v, w, pip, qv), base, gr, opts)
    return xi, (th, u, v, w, pip, qv), dt


def _rk3_rotation(fdt):
    A = fdt * np.array([[0.0, 1.0], [-1.0, 0.0]])
    return np.eye(2) + A + A @ A / 2 + A @ A @ A / 6


@pytest.mark.parametrize("cor_opt,ug", [(1, 0.0), (2, 1.0)])
def test_inertial_oscillation_follows_the_ws_rk3_polynomial(cor_opt, ug):
    u0, v0 = 1.7, -0.6
    xi, init, dt = _inertial_step(cor_opt, u0, v0)
    M = _rk3_rotation(o.fCor * dt)
    y = M @ np.array([u0 - ug, v0])
    un, vn = xi[1], xi[2]
    np.testing.assert_allclose(un - u0, (y[0] + ug) - u0, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(vn - v0, y[1] - v0, rtol=1e-9, atol=1e-15)
    # nothing else moves
    assert np.abs(xi[3]).max() < 1e-12 and np.abs(xi[4]).max() < 1e-15
    np.testing.assert_array_equal(xi[0], init[0])


def test_geostrophic_wind_is_steady_with_cor_opt_2():
    xi, init, _ = _inertial_step(2, 1.0, 0.0)
    np.testing.assert_allclose(xi[1], 1.0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(xi[2], 0.0, rtol=0, atol=1e-14)


def test_coriolis_pins_bite(monkeypatch):
    gr = _grid()
    u = np.full(gr.shape_u, 3.0)
    v = np.full(gr.shape_v, -2.0)
    mut = _mutant("calculate_coriolis_force_with_lspgrad", "f * (0.25 * su - u_geo), f * (0.25 * sv - v_geo)",
                  "f * (0.25 * su - v_geo), f * (0.25 * sv - u_geo)")
    fu, fv = mut(u, v)
    assert not np.allclose(fu, o.fCor * 2.0)
    # fu entering the v equation with the wrong sign rotates the wind the other way
    monkeypatch.setattr(o, "rk_sub_step", _mutant("rk_sub_step", "rhs_v = adv4v - fu", "rhs_v = adv4v + fu"))
    xi, _, dt = _inertial_step(1, 1.7, -0.6)
    y = _rk3_rotation(o.fCor * dt) @ np.array([1.7, -0.6])
    assert not np.allclose(xi[2] - (-0.6), y[1] - (-0.6), rtol=1e-3)
