"""Oracle restatement of the horizontal / vertical Smagorinsky closure (turb_opt 2:
compute_k2, turbulence.py:449-471; compute_smag2, :114-295), pinned by discrete closed
forms and by which coefficient each stress and flux takes:

* a vertical shear u = a z, v = c z gives Kmh = 0 and Kmv = Cs^2 Lv^2 sqrt(a^2 + c^2 - N^2)
  with 1/Lv^2 = (0.4 z)^-2 + dz^-2;
* a horizontal shear u = U sin(k y) gives Kmv = 0 and Kmh = Cs^2 dx dy (U sin(k dy)/dy)^2
  cos^2(k y) (the discrete S12 of the staggered grid, squared: no square root);
* with Kmh = Kmv and Khh = Khv the tendencies equal compute_smag's (turb_opt 1) except the
  surface w level, which compute_smag2 leaves at its stencil value;
* zeroing Kmv removes the vertical-shear stresses, zeroing Kmh the horizontal ones, and
  likewise Khv / Khh for a scalar varying only in z / only in x.
"""
import numpy as np

from oracle import lex_oracle as o

GR = o.Grid(8, 8, 10, 100.0, 120.0, 50.0)
Z = (np.arange(GR.nz + 2) - 0.5) * GR.dz


def _zeros():
    return np.zeros(GR.shape_u), np.zeros(GR.shape_v), o.set_w_bc(np.zeros(GR.shape_w))


def _neutral():
    return np.full(GR.shape_c, 300.0), np.zeros(GR.shape_c)


def test_vertical_shear_closed_form():
    a, c = 0.02, -0.01
    u, v, w = _zeros()
    u[:] = a * Z[None, None, :]
    v[:] = c * Z[None, None, :]
    rho = np.full(GR.shape_c, 1.1)
    th, qv = _neutral()
    s = o.compute_deformation(rho, u, v, w, GR)
    n2 = o.compute_nm(th, qv, GR)
    kmh, kmv, khh, khv = o.compute_k2(rho, *s[:6], n2, GR, cs=0.18)
    assert np.max(np.abs(o._interior(kmh))) < 1e-18
    lv2 = 1.0 / ((0.4 * Z) ** -2 + GR.dz ** -2)
    ref = 0.18 ** 2 * lv2[1:-1] * np.sqrt(a * a + c * c)
    np.testing.assert_allclose(o._interior(kmv), np.broadcast_to(ref, (GR.nx, GR.ny, GR.nz)), rtol=1e-10)
    assert abs(lv2[1] - GR.dz ** 2 / 26.0) < 1e-9   # first cell centre at dz/2
    assert khh is not None and np.array_equal(khv, kmv) and np.array_equal(khh, kmh)


def test_vertical_shear_with_stratification():
    """N^2 > 0 reduces Kmv as sqrt(max(Dv - N^2, 0)) (prandtl 1), and switches it off when
    the stratification wins."""
    a = 0.02
    u, v, w = _zeros()
    u[:] = a * Z[None, None, :]
    rho = np.full(GR.shape_c, 1.1)
    qv = np.zeros(GR.shape_c)
    s = o.compute_deformation(rho, u, v, w, GR)
    lv2 = 1.0 / ((0.4 * Z) ** -2 + GR.dz ** -2)
    for gamma in (0.003, 0.02):
        th = np.broadcast_to((300.0 + gamma * Z)[None, None, :], GR.shape_c).copy()
        n2 = o.compute_nm(th, qv, GR)
        _, kmv, _, _ = o.compute_k2(rho, *s[:6], n2, GR)
        ref = 0.18 ** 2 * lv2[None, None, 1:-1] * np.sqrt(np.maximum(a * a - o._interior(n2), 0.0))
        np.testing.assert_allclose(o._interior(kmv), ref, rtol=1e-10, atol=1e-18)
    assert np.max(o._interior(kmv)) == 0.0   # g/theta 0.02 ~ 6.5e-4 > a^2 = 4e-4


def test_horizontal_shear_closed_form():
    U = 1.5
    ky = 2 * np.pi / (GR.ny * GR.dy)
    u, v, w = _zeros()
    y_u = (np.arange(GR.ny + 6) - 3 + 0.5) * GR.dy        # u rows sit at cell centres in y
    u[:] = U * np.sin(ky * y_u)[None, :, None]
    rho = np.full(GR.shape_c, 1.2)
    th, qv = _neutral()
    s = o.compute_deformation(rho, u, v, w, GR)
    n2 = o.compute_nm(th, qv, GR)
    kmh, kmv, _, _ = o.compute_k2(rho, *s[:6], n2, GR)
    yj = (np.arange(GR.ny) + 0.5) * GR.dy
    s12 = 0.5 * U * np.sin(ky * GR.dy) / GR.dy * np.cos(ky * yj)
    ref = 0.18 ** 2 * GR.dx * GR.dy * 4.0 * s12 ** 2
    np.testing.assert_allclose(o._interior(kmh), np.broadcast_to(ref[None, :, None], (GR.nx, GR.ny, GR.nz)),
                               rtol=1e-9, atol=1e-15)
    assert np.max(np.abs(o._interior(kmv))) < 1e-15


def _random_flow(seed=4):
    rng = np.random.default_rng(seed)
    u = o.padding3u_array(np.concatenate([ui := rng.normal(0, 1, (GR.nx, GR.ny, GR.nz)), ui[:1]], 0))
    v = o.padding3v_array(np.concatenate([vi := rng.normal(0, 1, (GR.nx, GR.ny, GR.nz)), vi[:, :1]], 1))
    w = o.set_w_bc(o.padding3_array(rng.normal(0, 1, (GR.nx, GR.ny, GR.nz + 1))))
    th = o.padding3_array(300.0 + rng.normal(0, 1, (GR.nx, GR.ny, GR.nz)))
    qv = o.padding3_array(0.01 + 1e-3 * rng.normal(0, 1, (GR.nx, GR.ny, GR.nz)))
    rho = np.broadcast_to(np.linspace(1.2, 0.9, GR.nz + 2)[None, None, :], GR.shape_c).copy()
    return rho, u, v, w, th, qv


def test_equal_coefficients_reduce_to_compute_smag():
    rho, u, v, w, th, qv = _random_flow()
    s = o.compute_deformation(rho, u, v, w, GR)
    n2 = o.compute_nm(th, qv, GR)
    km, kh = o.compute_k(s[6], n2, GR)
    a = o.compute_smag(rho, km, kh, *s[:6], (th, qv), GR)
    b = o.compute_smag2(rho, km, km, kh, kh, *s[:6], (th, qv), GR)
    for n, x, y in zip(["u", "v", "w", "theta", "qv"], a, b):
        if n == "w":
            assert np.array_equal(x[:, :, 1:], y[:, :, 1:])
            assert not x[:, :, 0].any()
        else:
            assert np.array_equal(x, y), n
    # with its own coefficients (rho0 differs between the ghost and the first level)
    # compute_smag2 leaves a non-zero stencil value at the surface w level
    kmh, kmv, khh, khv = o.compute_k2(rho, *s[:6], n2, GR)
    assert o.compute_smag2(rho, kmh, kmv, khh, khv, *s[:6], (), GR)[2][:, :, 0].any()


def test_each_stress_and_flux_takes_its_coefficient():
    rho, u, v, w, th, qv = _random_flow(5)
    s = o.compute_deformation(rho, u, v, w, GR)
    k = np.abs(np.random.default_rng(6).normal(1.0, 0.2, GR.shape_c))
    zero = np.zeros(GR.shape_c)
    # vertical shear only (s13): the stress uses Kmv, nothing with Kmv = 0
    only13 = [zero, zero, zero, zero, s[4], zero]
    assert all(not np.any(t) for t in o.compute_smag2(rho, k, zero, k, k, *only13, (), GR)[:3])
    assert any(np.any(t) for t in o.compute_smag2(rho, zero, k, k, k, *only13, (), GR)[:3])
    # horizontal strain only (s11, s12): Kmh
    only_h = [s[0], zero, zero, s[3], zero, zero]
    assert all(not np.any(t) for t in o.compute_smag2(rho, zero, k, k, k, *only_h, (), GR)[:3])
    assert any(np.any(t) for t in o.compute_smag2(rho, k, zero, k, k, *only_h, (), GR)[:3])
    # a scalar varying only in z diffuses with Khv, one varying only in x with Khh
    sz = np.broadcast_to((300.0 + 1e-4 * Z ** 2)[None, None, :], GR.shape_c).copy()
    x = (np.arange(GR.nx + 6) - 3 + 0.5) * GR.dx
    sx = np.broadcast_to((300.0 + np.sin(2 * np.pi * x / (GR.nx * GR.dx)))[:, None, None], GR.shape_c).copy()
    assert not np.any(o.compute_smag2(rho, k, k, k, zero, *s[:6], (sz,), GR)[3])
    assert np.any(o.compute_smag2(rho, k, k, zero, k, *s[:6], (sz,), GR)[3])
    assert np.max(np.abs(o.compute_smag2(rho, k, k, zero, k, *s[:6], (sx,), GR)[3])) < 1e-12
    assert np.any(o.compute_smag2(rho, k, k, k, zero, *s[:6], (sx,), GR)[3])


def test_step_with_turb_opt_2_runs_and_differs():
    case, _ = o.case_config1()
    c = o.scaled_case(case, 8, 6, 10)
    phys, base, _ = o.setup_warm_bubble(c, dtype=np.float64)
    a, _ = o.ws_rk3_big_step(phys, base, c.grid, o.Options(dt=1.0, solver_opt=2, n_sound=6, turb_opt=1))
    b, _ = o.ws_rk3_big_step(phys, base, c.grid, o.Options(dt=1.0, solver_opt=2, n_sound=6, turb_opt=2))
    assert all(np.isfinite(x).all() for x in b)
    assert not np.array_equal(a[0], b[0])
