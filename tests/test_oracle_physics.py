"""Oracle restatement of the physics options Rayleigh damping (boundary_conditions.py:
119-141, setup_lex.py:149-156) and Newtonian heating (radiation.py:7-16), pinned by
closed forms."""
import numpy as np

from oracle import lex_oracle as o


def test_damping_ramp_closed_form():
    g = o.Grid(4, 3, 20, 100.0, 100.0, 100.0)
    zd = 1200.0
    tauh, tauf = o.setup_damping_tau(g, zd)
    zc = (np.arange(g.nz + 2) - 0.5) * g.dz
    zw = (np.arange(g.nz + 3) - 1.0) * g.dz
    depth = (g.nz + 1) * g.dz - zd
    assert not tauh[0, 0][zc <= zd].any() and not tauf[0, 0][zw <= zd].any()
    assert np.allclose(tauh[0, 0][zc > zd], 0.5 * (1 - np.cos(np.pi * (zc[zc > zd] - zd) / depth)), atol=1e-15)
    assert tauf[0, 0, -1] == 1.0          # the top padded w level ends the ramp
    assert (np.diff(tauf[0, 0]) >= 0).all() and (np.diff(tauh[0, 0]) >= 0).all()
    assert (tauh == tauh[:1, :1]).all()   # horizontally uniform


def test_damping_tendencies():
    rng = np.random.default_rng(1)
    g = o.Grid(5, 4, 10, 100.0, 100.0, 100.0)
    tauh, tauf = o.setup_damping_tau(g, 400.0)
    u = o.padding3u_array(rng.normal(size=(g.nx + 1, g.ny, g.nz)))
    v = o.padding3v_array(rng.normal(size=(g.nx, g.ny + 1, g.nz)))
    w = o.padding3_array(rng.normal(size=(g.nx, g.ny, g.nz + 1)))
    # theta horizontally uniform per level: no theta tendency
    th = o.padding3_array(np.broadcast_to(300.0 + np.arange(g.nz), (g.nx, g.ny, g.nz)).copy())
    a = 1.0 / 120.0
    ut, vt, wt, tt = o.rayleigh_damping(tauh, tauf, u, v, w, th, a)
    assert not tt.any()
    th_col = tauh[0, 0, 1:-1]
    assert np.allclose(ut, -a * th_col * u[3:-3, 3:-3, 1:-1], rtol=1e-15, atol=0)
    assert np.allclose(vt, -a * th_col * v[3:-3, 3:-3, 1:-1], rtol=1e-15, atol=0)
    assert np.allclose(wt, -a * tauf[0, 0, 1:-1] * w[3:-3, 3:-3, 1:-1], rtol=1e-15, atol=0)
    # theta relaxes to the level mean: the tendency sums to zero per level
    th2 = o.padding3_array(300.0 + rng.normal(size=(g.nx, g.ny, g.nz)))
    tt2 = o.rayleigh_damping(tauh, tauf, u, v, w, th2, a)[3]
    assert np.allclose(tt2.sum(axis=(0, 1)), 0.0, atol=1e-12)


def test_newtonian_heating_closed_form_and_caps():
    g = o.Grid(4, 4, 4, 100.0, 100.0, 100.0)
    th0 = o.padding3_array(np.full((g.nx, g.ny, g.nz), 300.0))
    d = np.array([-3.0, -1.0, -0.25, 0.0, 0.5, 1.0, 2.0])
    th = th0.copy()
    th[3:-3, 3:-3, 1:-1] = 300.0 + np.resize(d, (g.nx, g.ny, g.nz))
    h = o.newton_radiation(th, th0)
    dd = th[3:-3, 3:-3, 1:-1] - 300.0
    tau = 12.0 * 3600.0
    exp = np.where(dd > 1.0, -1.0 / tau, np.where(dd < -1.0, 1.0 / tau, -dd / tau))
    assert np.allclose(h, exp, rtol=1e-15, atol=0)


def test_step_with_damping_only_changes_the_layer():
    """A warm-bubble step with the damping layer above all the motion and a uniform
    theta there equals the undamped step (below the layer nothing changes), and damping
    a uniform upper-level wind removes kinetic energy."""
    case, opts = o.case_config1()
    c = o.scaled_case(case, 12, 10, 16)
    phys, base, _ = o.setup_warm_bubble(c, dtype=np.float64)
    oo = o.Options(dt=1.0, solver_opt=2, n_sound=7, turb_opt=1)
    od = o.Options(dt=1.0, solver_opt=2, n_sound=7, turb_opt=1, damp_opt=True, z_damping=1400.0)
    u = phys[1].copy()
    u[:, :, 12:] += 3.0   # wind in the upper levels
    st = (phys[0], u) + tuple(phys[2:])
    a, _ = o.ws_rk3_big_step(st, base, c.grid, oo)
    b, _ = o.ws_rk3_big_step(st, base, c.grid, od)
    ke = lambda x: float(np.sum(x[1][3:-3, 3:-3, 1:-1] ** 2))  # noqa: E731
    assert ke(b) < ke(a)
    assert np.allclose(b[1][3:-3, 3:-3, 1:8], a[1][3:-3, 3:-3, 1:8], rtol=0, atol=1e-9)
