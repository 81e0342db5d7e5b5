"""Oracle restatement of the bulk atmosphere-ocean fluxes (atm_ocn_flux, q_sat_sfc, cdn,
psi_m_s: boundary_conditions.py:7-103) and their use in the step (one_step_integration.py:
37-52; advection.py:9-15, :37-48, :77-88), pinned by textbook forms and Monin-Obukhov
identities that hold whatever the iteration does:

* psi_m_s against the Paulson (1970) / Dyer unstable forms with x = (1 - 16 z/L)^(1/4);
* cdn against the Large & Yeager (2004) neutral drag law 1e-3 (2.7/U + 0.142 + U/13.09);
* q_sat_sfc against Bolton's saturation pressure over the ideal gas law;
* the fluxes against the similarity profiles their outputs must satisfy (the 10 m neutral
  wind, the 2 m temperature and humidity, the neutral drag law at 10 m) with u*, theta*
  and q* recovered from the returned stress, heat and moisture fluxes;
* the step's use: the bottom-face fluxes change only the lowest level, by exactly the
  flux over rho0 dz, with the stress averaged onto the u and v faces periodically.
"""
import numpy as np

from oracle import lex_oracle as o


def _psi_textbook(zeta):
    zeta = np.asarray(zeta, dtype=np.float64)
    x = np.abs(1.0 - 16.0 * zeta) ** 0.25
    pm = 2 * np.log((1 + x) / 2) + np.log((1 + x * x) / 2) - 2 * np.arctan(x) + np.pi / 2
    ps = 2 * np.log((1 + x * x) / 2)
    return np.where(zeta > 0, -5 * zeta, pm), np.where(zeta > 0, -5 * zeta, ps)


def test_stability_functions_textbook():
    zeta = np.array([-10.0, -3.0, -1.0, -0.3, -0.05, -1e-4, 0.0, 1e-3, 0.2, 1.0, 10.0])
    pm, ps = o.psi_m_s(zeta)
    em, es = _psi_textbook(zeta)
    # the reference writes pi/2 as 1.571
    assert np.allclose(pm, em, rtol=0, atol=3e-4)
    assert np.allclose(ps, es, rtol=0, atol=1e-12)
    assert np.allclose(pm[zeta > 0], -5 * zeta[zeta > 0], rtol=0, atol=0)
    assert (np.diff(ps[zeta <= 0]) < 0).all()   # unstable: psi grows with instability


def test_neutral_drag_and_saturation_textbook():
    u = np.array([0.5, 2.0, 5.0, 10.0, 20.0, 30.0])
    assert np.allclose(o.cdn(u), 1e-3 * (2.7 / u + 0.142 + u / 13.09), rtol=1e-4, atol=0)
    assert abs(o.cdn(10.0) - 1.176e-3) < 1e-9
    t = np.array([275.0, 285.0, 295.0, 300.0, 305.0])
    es = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))   # Bolton (1980), Pa
    rho_v = es / (o.Rv * t)
    assert np.allclose(o.q_sat_sfc(t), rho_v, rtol=0.03, atol=0)


def test_neutral_case_closed_form():
    """Delta theta = 0 and saturated air at the 10 m reference height: z/L = 0, so the
    stress is the neutral drag law at the lowest-level wind, with no heat or moisture flux;
    winds under 0.25 m/s count as 0.25 m/s."""
    rho = 1.15
    ts = np.full(5, 300.0)
    ssq = 0.98 * o.q_sat_sfc(ts) / rho
    u = np.array([3.0, -7.0, 0.1, 12.0, 0.0])
    v = np.array([4.0, 2.0, 0.05, -5.0, 0.0])
    tx, ty, sen, evap, t_ref, q_ref, u10 = o.atm_ocn_flux(np.full(5, 10.0), u, v, ts.copy(), ssq, np.full(5, rho), ts)
    vm = np.maximum(np.hypot(u, v), 0.25)
    tau = rho * o.cdn(vm) * vm * vm
    assert np.allclose(tx, -tau * u / vm, rtol=2e-4, atol=1e-15)
    assert np.allclose(ty, -tau * v / vm, rtol=2e-4, atol=1e-15)
    assert not sen.any() and not evap.any()
    assert np.allclose(u10, vm, rtol=1e-4)   # psi_m(0) = 1.571 - pi/2, not 0
    assert np.allclose(t_ref, ts - 0.02, rtol=0, atol=1e-12) and np.allclose(q_ref, ssq, rtol=0, atol=1e-15)


def _cases(dth=(-3.0, -0.2)):
    rng = np.random.default_rng(7)
    n = 64
    z = rng.uniform(5.0, 60.0, n)
    u = rng.uniform(-15.0, 15.0, n)
    v = rng.uniform(-15.0, 15.0, n)
    ts = rng.uniform(285.0, 303.0, n)
    th = ts + rng.uniform(*dth, n)
    rho = rng.uniform(1.1, 1.2, n)
    q = 0.98 * o.q_sat_sfc(ts) / rho * rng.uniform(0.5, 1.0, n)
    return z, u, v, th, q, rho, ts


def test_similarity_profiles_hold():
    """Over a sea warmer than the air (z/L < 0, where the eight passes converge), with
    u* = sqrt(|tau|/rho), theta* = -sen/(cp rho u*), q* = -evap/(rho u*) and z/L from
    them, the outputs satisfy Monin-Obukhov similarity:
      U10N = |U| - (u*/k)(ln(z/10) - psi_m(z/L))       (10 m neutral wind)
      u*   = sqrt(cd_n(U10N)) U10N                      (neutral drag law at 10 m)
      T2   = theta - (theta*/k)(ln(z/2) - psi_h(z/L) + psi_h(2/L)) - 0.02
      q2   = q - (q*/k)(ln(z/2) - psi_h(z/L) + psi_h(2/L))
    and the Large & Yeager (2004) unstable neutral Stanton and Dalton numbers 32.7e-3 and
    34.6e-3 (per sqrt(cd_n)) shifted to height z and stability:
      theta*/dtheta = 0.0327 / (1 + 0.0327/k (ln(z/10) - psi_h(z/L))), q*/dq alike."""
    z, u, v, th, q, rho, ts = _cases()
    tx, ty, sen, evap, t_ref, q_ref, u10 = o.atm_ocn_flux(z, u, v, th, q, rho, ts)
    vm = np.maximum(np.hypot(u, v), 0.25)
    tau = np.hypot(tx, ty)
    us = np.sqrt(tau / rho)
    ssq = 0.98 * o.q_sat_sfc(ts) / rho
    cp = o.Cp * (1 + (1810.0 / o.Cp - 1.0) * ssq)
    ts_ = -sen / (cp * rho * us)
    qs_ = -evap / (rho * us)
    k = 0.4
    hol = np.clip(k * o.g * z * (ts_ / th + qs_ / (1 / o.repsm1 + q)) / us ** 2, -10, 10)
    pm, ph = _psi_textbook(hol)
    _, ph2 = _psi_textbook(hol * 2.0 / z)
    assert np.allclose(u10, vm - us / k * (np.log(z / 10) - pm), rtol=2e-3)
    assert np.allclose(us, np.sqrt(1e-3 * (2.7 / u10 + 0.142 + u10 / 13.09)) * u10, rtol=2e-3)
    prof = np.log(z / 2) - ph + ph2
    assert np.allclose(t_ref, th - ts_ / k * prof - 0.02, rtol=0, atol=2e-3)
    assert np.allclose(q_ref, q - qs_ / k * prof, rtol=0, atol=2e-6)
    ssq = 0.98 * o.q_sat_sfc(ts) / rho
    shift = np.log(z / 10) - ph
    assert np.allclose(ts_ / (th - ts), 0.0327 / (1 + 0.0327 / k * shift), rtol=2e-3)
    assert np.allclose(qs_ / (q - ssq), 0.0346 / (1 + 0.0346 / k * shift), rtol=2e-3)
    # stress opposes the wind; a warmer sea heats the air, a subsaturated one moistens it
    assert np.allclose(tx * vm, -tau * u, rtol=1e-12) and np.allclose(ty * vm, -tau * v, rtol=1e-12)
    assert (sen > 0).all() and (evap > 0).all()
    assert (hol < 0).all()


def test_stable_side():
    """Air warmer than the sea: downward heat flux, and a drag below the neutral one at the
    same 10 m wind that shrinks as the air-sea difference grows.  (Here the eight passes
    need not converge: the neutral coefficient alternates between passes as the reference's
    fixed iteration does.)"""
    z, u, v, th, q, rho, ts = _cases((0.2, 3.0))
    tx, ty, sen, evap, _, _, u10 = o.atm_ocn_flux(z, u, v, th, q, rho, ts)
    vm = np.maximum(np.hypot(u, v), 0.25)
    tau = np.hypot(tx, ty)
    assert (sen < 0).all()
    z10 = np.full_like(z, 10.0)
    tn = np.hypot(*o.atm_ocn_flux(z10, u, v, ts, q, rho, ts)[:2])
    t1 = np.hypot(*o.atm_ocn_flux(z10, u, v, ts + 0.5, q, rho, ts)[:2])
    t2 = np.hypot(*o.atm_ocn_flux(z10, u, v, ts + 2.0, q, rho, ts)[:2])
    assert (t2 < t1).all() and (t1 < tn * 1.0001).all()
    assert np.isfinite(tau).all() and (tau > 0).all() and np.isfinite(u10).all()


def test_reference_height_two_metres():
    """At z = 2 m the 2 m diagnostics are the lowest level itself (minus the 0.01 K/m
    potential-temperature correction)."""
    _, u, v, th, q, rho, ts = _cases()
    z = np.full_like(u, 2.0)
    _, _, _, _, t_ref, q_ref, _ = o.atm_ocn_flux(z, u, v, th, q, rho, ts)
    assert np.allclose(t_ref, th - 0.02, rtol=0, atol=1e-12)
    assert np.allclose(q_ref, q, rtol=0, atol=1e-15)


def test_step_uses_bottom_fluxes_on_the_lowest_level():
    """In the first stage (one acoustic substep, velocities updated before pi') the fluxes
    move only the lowest level: theta by dt sen/(Cp rho0 dz), qv by dt evap/(rho0 dz), u on
    face f by dt (tau_x[f-1] + tau_x[f])/2 / (rho0 dz) with periodic neighbours, v alike."""
    case, _ = o.case_config1()
    c = o.scaled_case(case, 8, 6, 10)
    phys, base, _ = o.setup_warm_bubble(c, dtype=np.float64)
    gr = c.grid
    rng = np.random.default_rng(3)
    sfc = tuple(rng.normal(scale=s, size=(gr.nx, gr.ny)) for s in (0.1, 0.1, 50.0, 1e-4))
    op = o.Options(dt=1.0, solver_opt=2, n_sound=6)
    dts = op.dt / 3.0
    a, _, _ = o.rk_sub_step(0, phys, phys, base, gr, op, dts)
    b, _, _ = o.rk_sub_step(0, phys, phys, base, gr, op, dts, sfc=sfc)
    r1 = base[1][3, 3, 1]
    tau_x, tau_y, sen, evap = sfc
    dth = b[0][3:-3, 3:-3, 1:-1] - a[0][3:-3, 3:-3, 1:-1]
    dqv = b[5][3:-3, 3:-3, 1:-1] - a[5][3:-3, 3:-3, 1:-1]
    assert np.allclose(dth[:, :, 0], dts * sen / o.Cp / (r1 * gr.dz), rtol=1e-9, atol=1e-12)
    assert np.allclose(dqv[:, :, 0], dts * evap / (r1 * gr.dz), rtol=1e-9, atol=1e-16)
    assert not dth[:, :, 1:].any() and not dqv[:, :, 1:].any()
    du = b[1][3:-3, 3:-3, 1:-1] - a[1][3:-3, 3:-3, 1:-1]
    dv = b[2][3:-3, 3:-3, 1:-1] - a[2][3:-3, 3:-3, 1:-1]
    fu = np.empty((gr.nx + 1, gr.ny))
    for f in range(gr.nx + 1):
        fu[f] = 0.5 * (tau_x[(f - 1) % gr.nx] + tau_x[f % gr.nx])
    fv = np.empty((gr.nx, gr.ny + 1))
    for f in range(gr.ny + 1):
        fv[:, f] = 0.5 * (tau_y[:, (f - 1) % gr.ny] + tau_y[:, f % gr.ny])
    assert np.allclose(du[:, :, 0], dts * fu / (r1 * gr.dz), rtol=1e-9, atol=1e-13)
    assert np.allclose(dv[:, :, 0], dts * fv / (r1 * gr.dz), rtol=1e-9, atol=1e-13)
    assert np.abs(du[:, :, 1:]).max() < 1e-13 and np.abs(dv[:, :, 1:]).max() < 1e-13


def test_step_with_surface_fluxes_runs_from_the_base_surface_temperature():
    """ws_rk3_big_step with sfc_opt computes the fluxes from the step's input state and the
    base state's surface temperature; a sea 2 K warmer than the air warms the lowest level."""
    case, _ = o.case_config1()
    c = o.scaled_case(case, 8, 6, 10)
    phys, base, _ = o.setup_warm_bubble(c, dtype=np.float64)
    gr = c.grid
    u = phys[1].copy()
    u += 5.0
    st = (phys[0], u) + tuple(phys[2:])
    ts = phys[0][3:-3, 3:-3, 1] + 2.0
    base6 = tuple(base[:5]) + (ts,)
    off, _ = o.ws_rk3_big_step(st, base6, gr, o.Options(dt=1.0, solver_opt=2, n_sound=6))
    on, _ = o.ws_rk3_big_step(st, base6, gr, o.Options(dt=1.0, solver_opt=2, n_sound=6, sfc_opt=True))
    d = on[0][3:-3, 3:-3, 1] - off[0][3:-3, 3:-3, 1]
    assert (d > 0).all()
    fl = o.surface_fluxes(st, base6, gr)
    assert (fl[2] > 0).all() and (fl[0] < 0).all()   # upward heat flux, stress against +u
    # the lowest level: theta, qv and rho0 of the first cell, whose centre lies midway
    # between the surface (w level 0) and w level 1; u, v averaged onto the cell centre
    zc = 0.5 * (0.0 + gr.dz)
    uc = 0.5 * (u[3:-4, 3:-3, 1] + u[4:-3, 3:-3, 1])
    vc = 0.5 * (phys[2][3:-3, 3:-4, 1] + phys[2][3:-3, 4:-3, 1])
    ex = o.atm_ocn_flux(np.full((gr.nx, gr.ny), zc), uc, vc, phys[0][3:-3, 3:-3, 1], phys[5][3:-3, 3:-3, 1],
                        base[1][3:-3, 3:-3, 1], ts)
    for a, b in zip(fl, ex):
        assert np.array_equal(a, b)
