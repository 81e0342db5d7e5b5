"""Pins of the oracle's halos, divergences, PGF/Coriolis, acoustic substeps, WS-RK3
driver, Smagorinsky, base state and coarsening (CPU)."""
import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import random_state


# ----------------------------------------------------------------- H1 halos
def test_padding_index_maps_brute_force():
    rng = np.random.default_rng(0)
    nx, ny, nz = 7, 6, 5
    a = rng.normal(size=(nx, ny, nz))
    p = o.padding3_array(a)
    for I in range(nx + 6):
        for J in range(ny + 6):
            for K in range(nz + 2):
                assert p[I, J, K] == a[(I - 3) % nx, (J - 3) % ny, min(max(K - 1, 0), nz - 1)]
    au = rng.normal(size=(nx + 1, ny, nz))
    pu = o.padding3u_array(au)
    for I in range(nx + 7):
        f = I - 3
        src = f + nx if f < 0 else (f - nx if f > nx else f)   # ghosts skip the duplicate face
        for J in range(ny + 6):
            assert np.array_equal(pu[I, J, 1:-1], au[src, (J - 3) % ny])
    av = rng.normal(size=(nx, ny + 1, nz))
    pv = o.padding3v_array(av)
    for J in range(ny + 7):
        f = J - 3
        src = f + ny if f < 0 else (f - ny if f > ny else f)
        assert np.array_equal(pv[3:-3, J, 1:-1], av[:, src])
    p1 = o.padding_array(a)
    assert np.array_equal(p1[0, 1:-1, 1:-1], a[-1]) and np.array_equal(p1[-1, 1:-1, 1:-1], a[0])
    w = o.set_w_bc(rng.normal(size=(nx + 6, ny + 6, nz + 3)))
    assert not w[:, :, [0, 1, -2, -1]].any()


# ------------------------------------------------------------- A1 divergence
def test_divergence_of_linear_fields_is_exact():
    gr = o.Grid(8, 8, 6, 50.0, 70.0, 40.0)
    a, b, c = 0.01, -0.02, 0.03
    xu = (np.arange(gr.nx + 7) - 3) * gr.dx
    yv = (np.arange(gr.ny + 7) - 3) * gr.dy
    zw = (np.arange(gr.nz + 3) - 1) * gr.dz
    u = np.broadcast_to(a * xu[:, None, None], gr.shape_u).copy()
    v = np.broadcast_to(b * yv[None, :, None], gr.shape_v).copy()
    w = np.broadcast_to(c * zw[None, None, :], gr.shape_w).copy()
    d = o.get_divergence2(u, v, w, gr)
    np.testing.assert_allclose(d, a + b + c, rtol=1e-12)
    rho = np.full(gr.shape_c, 1.2)
    dr = o.get_divergence(rho, u, v, w, gr)
    np.testing.assert_allclose(o._interior(dr)[:, :, 1:-1], 1.2 * (a + b + c), rtol=1e-12)


def test_mass_divergence_integrates_to_zero():
    gr, phys, base = random_state(10, 8, 8, seed=3)
    u, v, w = phys[1:4]
    fd = o.get_divergence(base[1], u, v, w, gr)
    tot = np.abs(o._interior(fd)).sum()
    assert abs(o._interior(fd).sum()) < 1e-12 * tot


# ------------------------------------------------------------ P1/P2
def test_pgf_of_linear_pressure_is_constant():
    gr = o.Grid(8, 8, 6, 50.0, 70.0, 40.0)
    gx, gy, gz = 1e-5, -2e-5, 3e-5
    I, J, K = np.meshgrid(np.arange(gr.nx + 6), np.arange(gr.ny + 6), np.arange(gr.nz + 2), indexing="ij")
    pip = gx * I * gr.dx + gy * J * gr.dy + gz * K * gr.dz
    th = np.full(gr.shape_c, 301.0)
    pu, pv, pw = o.pressure_gradient_force(pip, th, gr)
    np.testing.assert_allclose(pu, -o.Cp * 301.0 * gx, rtol=1e-9)
    np.testing.assert_allclose(pv, -o.Cp * 301.0 * gy, rtol=1e-9)
    np.testing.assert_allclose(pw[:, :, 1:-1], -o.Cp * 301.0 * gz, rtol=1e-9)
    assert not pw[:, :, [0, -1]].any()


def test_coriolis_uniform_wind():
    gr = o.Grid(8, 8, 6, 50.0, 70.0, 40.0)
    u = np.full(gr.shape_u, 3.0)
    v = np.full(gr.shape_v, -2.0)
    fu, fv = o.calculate_coriolis_force(u, v)
    np.testing.assert_allclose(fu, o.fCor * 3.0)
    np.testing.assert_allclose(fv, o.fCor * -2.0)


# ------------------------------------------------------------ S1 acoustic
def test_acoustic_single_mode_matches_two_by_two_recurrence():
    """1-D acoustic mode exp(i k x) in the forward-backward substep, uniform theta, pi0:
    u' = u - Cp th dt D+ p ;  p' = p - 1.1 dt (Rd/Cv) pi0 D- u'  (pi* = 0, w = 0)."""
    gr = o.Grid(16, 6, 6, 100.0, 100.0, 100.0)
    kx = 2 * np.pi * 3 / gr.nx
    th = 300.0
    pi0c = 0.95
    dt = 0.05
    x = np.arange(gr.nx)
    p = o.padding3_array(np.broadcast_to((1e-4 * np.cos(kx * x))[:, None, None], (gr.nx, 6, 6)).copy())
    u = np.zeros(gr.shape_u)
    v = np.zeros(gr.shape_v)
    w = np.zeros(gr.shape_w)
    pi0 = np.full(gr.shape_c, pi0c)
    rho0 = np.ones(gr.shape_c)
    trho = np.full(gr.shape_c, th)
    ru = np.zeros((gr.nx + 1, 6, 6))
    rv = np.zeros((gr.nx, 7, 6))
    rw = np.zeros((gr.nx, 6, 7))
    n = 7
    un, vn, wn, pn = o.sound_rk_step(u, v, w, pi0, p, np.zeros_like(p), ru, rv, rw, trho, rho0, n, dt, gr)
    # Fourier recurrence: u on faces is shifted by half a cell
    s = 2j * np.sin(kx / 2) / gr.dx          # symbol of the staggered difference
    M = np.array([[1.0, -o.Cp * th * dt * s], [0.0, 1.0]], complex)
    Mp = np.array([[1.0, 0.0], [-1.1 * dt * (o.Rd / o.Cv) * pi0c * s, 1.0]], complex)
    st = np.array([0.0, 1e-4], complex)
    for _ in range(n):
        st = Mp @ (M @ st)
    pred_p = np.real(st[1] * np.exp(1j * kx * x))
    np.testing.assert_allclose(o._interior(pn)[:, 2, 2], pred_p, rtol=1e-8, atol=1e-14)
    pred_u = np.real(st[0] * np.exp(1j * kx * (np.arange(gr.nx + 1) - 0.5)))
    np.testing.assert_allclose(un[3:-3, 5, 3], pred_u, rtol=1e-8, atol=1e-12)


def test_rest_state_stays_at_rest():
    gr = o.Grid(8, 8, 10, 100.0, 100.0, 100.0)
    case = o.BubbleCase(gr, amplitude=0.0)
    phys, base, _ = o.setup_warm_bubble(case)
    opts = o.Options(dt=1.0, solver_opt=2, n_sound=6, turb_opt=1)
    xi, _ = o.ws_rk3_big_step(phys, base, gr, opts)
    for a, b in zip(xi, phys):
        assert np.array_equal(a, b)


# ------------------------------------------------------------ R2 WS-RK3
def test_ws_rk3_stability_polynomial(monkeypatch):
    """Stub every stage by xi0 + dt_s * lam * xi: one step = (1 + z + z^2/2 + z^3/6) xi0."""
    lam = -0.7 + 0.4j

    def stub(stage, xi0, xi, base, grid, opts, dt_stage, frozen=None):
        return tuple(a0 + dt_stage * lam * a for a0, a in zip(xi0, xi)), frozen, None

    monkeypatch.setattr(o, "rk_sub_step", stub)
    x0 = (np.array([1.0 + 0j, 2.0 - 1j]),)
    for dt in (0.1, 0.5, 1.3):
        xi, _ = o.ws_rk3_big_step(x0, None, None, o.Options(dt=dt))
        z = lam * dt
        np.testing.assert_allclose(xi[0], (1 + z + z ** 2 / 2 + z ** 3 / 6) * x0[0], rtol=1e-13)


def test_acoustic_substep_counts():
    assert o._n_acoustic(1, 0.5, o.Options(dt=1.0, n_sound=6))[0] == 3
    assert o._n_acoustic(1, 0.5, o.Options(dt=1.0, n_sound=5))[0] == 2      # round half to even
    assert o._n_acoustic(1, 0.5, o.Options(dt=1.0, n_sound=7))[0] == 4
    assert o._n_acoustic(2, 1.0, o.Options(dt=1.0, n_sound=6))[0] == 6
    assert o._n_acoustic(0, 1 / 3, o.Options(dt=1.0, n_sound=6)) == (1, 1 / 3)


# ------------------------------------------------------------ G1 Smagorinsky
def test_smagorinsky_uniform_shear():
    """u = a z: S13 = a/2, deformation = a^2 at every level (extrapolation exact for linear)."""
    gr = o.Grid(8, 8, 10, 100.0, 100.0, 50.0)
    a = 0.02
    z = (np.arange(gr.nz + 2) - 0.5) * gr.dz
    u = np.broadcast_to(a * z[None, None, :], gr.shape_u).copy()
    v = np.zeros(gr.shape_v)
    w = np.zeros(gr.shape_w)
    rho = np.full(gr.shape_c, 1.1)
    s11, s22, s33, s12, s13, s23, df = o.compute_deformation(rho, u, v, w, gr)
    np.testing.assert_allclose(o._interior(s13), 1.1 * a / 2, rtol=1e-10)
    np.testing.assert_allclose(o._interior(df), a * a, rtol=1e-10)
    th = np.broadcast_to((300.0 + 0.003 * z)[None, None, :], gr.shape_c).copy()
    qv = np.zeros(gr.shape_c)
    n2 = o.compute_nm(th, qv, gr)
    km, kh = o.compute_k(df, n2, gr, cs=0.18)
    delta = (gr.dx * gr.dy * gr.dz) ** (1 / 3)
    ref = (0.18 * delta) ** 2 * np.sqrt(np.maximum(a * a - o._interior(n2) * 3.0, 0.0))
    np.testing.assert_allclose(o._interior(km), ref, rtol=1e-10)
    np.testing.assert_allclose(kh, 3.0 * km, rtol=1e-12)
    # N^2 of theta = 300 + 0.003 z: g d ln(theta)/dz between levels, averaged to cells
    thi = 300.0 + 0.003 * z[1:-1]
    n28w = o.g * np.log(thi[1:] / thi[:-1]) / gr.dz
    got = o._interior(n2)[:, :, 1:-1]
    np.testing.assert_allclose(got, np.broadcast_to(0.5 * (n28w[:-1] + n28w[1:]), got.shape), rtol=1e-10)


def test_smagorinsky_uniform_flow_and_linear_scalar_give_no_tendency():
    gr = o.Grid(8, 8, 10, 100.0, 100.0, 50.0)
    u = np.full(gr.shape_u, 2.0)
    v = np.full(gr.shape_v, -1.0)
    w = o.set_w_bc(np.zeros(gr.shape_w))
    rho = np.full(gr.shape_c, 1.1)
    th = np.full(gr.shape_c, 300.0)
    qv = np.zeros(gr.shape_c)
    out = o.smagorinsky(rho, u, v, w, th, qv, (th, qv), gr)
    for a in out:
        assert np.max(np.abs(a)) == 0.0


# ------------------------------------------------------------ I1 base state
def test_dry_isentropic_base_state():
    """Constant theta0, dry: pi at w levels is 1 - g z/(cp theta0) exactly; cells average it."""
    gr = o.Grid(4, 4, 30, 100.0, 100.0, 150.0)
    bs = o.hydrostatic_base_state(gr, 300.0, lambda z: 0.0 * z)
    zw = np.arange(gr.nz + 1) * gr.dz
    piw = 1.0 - o.g * zw / (o.Cp * 300.0)
    np.testing.assert_allclose(bs["pi0"][1:-1], 0.5 * (piw[:-1] + piw[1:]), rtol=1e-13)
    assert bs["pi0"][0] == bs["pi0"][1] and bs["pi0"][-1] == bs["pi0"][-2]
    np.testing.assert_allclose(bs["rho0_theta0"], bs["pi0"] ** (o.Cv / o.Rd) * o.p00 / o.Rd, rtol=1e-13)


def test_rslf_matches_bolton_celsius_form():
    """Bolton (1980): e_s = 6.112 hPa exp(17.67 T_c / (T_c + 243.5))."""
    for tc, p in [(0.0, 1.0e5), (25.0, 9.0e4), (-20.0, 5.0e4)]:
        es = 611.2 * np.exp(17.67 * tc / (tc + 243.5))
        ref = (o.Rd / o.Rv) * es / (p - es)
        assert abs(o.rslf(p, tc + 273.15) - ref) < 1e-12 * ref


def test_moist_base_state_converges_and_is_hydrostatic():
    gr = o.Grid(4, 4, 40, 100.0, 100.0, 250.0)
    rh = lambda z: 0.8 - 0.6 * np.clip(z, 0, 1e4) / 1e4  # noqa: E731
    bs = o.hydrostatic_base_state(gr, 300.0, rh)
    assert 0 < bs["iters"] < 50
    # discrete hydrostatic balance: d(pi)/dz at cell faces = -g / (cp theta_rho0)
    trho = bs["theta0"] * (1 + o.reps * bs["qv0"]) / (1 + bs["qv0"])
    dpi = np.diff(bs["pi0"][1:-1]) / gr.dz
    ref = -o.g / o.Cp * 0.5 * (1 / trho[1:-2] + 1 / trho[2:-1])
    np.testing.assert_allclose(dpi, ref, rtol=1e-3)


# ------------------------------------------------------------ X5 coarsening
def test_block_mean_coarsening():
    rng = np.random.default_rng(5)
    a = rng.normal(size=(12, 6, 9))
    c = o.coarsen_block_mean(a, 6, 3, 3)
    for I in range(2):
        for J in range(2):
            for K in range(3):
                assert abs(c[I, J, K] - a[6 * I:6 * I + 6, 3 * J:3 * J + 3, 3 * K:3 * K + 3].mean()) < 1e-14


def test_face_mean_coarsening_preserves_block_mean_divergence():
    """div_coarse(face-mean(u, v, w)) == block-mean(div_fine(u, v, w)) exactly (telescoping
    sums of the fine fluxes across each coarse block), for a non-periodic w with lids."""
    rng = np.random.default_rng(6)
    nx, ny, nz, f = 12, 18, 9, (6, 3, 3)
    dx, dy, dz = 100.0, 120.0, 80.0
    u = rng.normal(size=(nx + 1, ny, nz))
    v = rng.normal(size=(nx, ny + 1, nz))
    w = rng.normal(size=(nx, ny, nz + 1))
    div_f = np.diff(u, axis=0) / dx + np.diff(v, axis=1) / dy + np.diff(w, axis=2) / dz
    uc = o.coarsen_face_mean(u, 0, *f)
    vc = o.coarsen_face_mean(v, 1, *f)
    wc = o.coarsen_face_mean(w, 2, *f)
    assert uc.shape == (nx // 6 + 1, ny // 3, nz // 3) and wc.shape == (nx // 6, ny // 3, nz // 3 + 1)
    div_c = np.diff(uc, axis=0) / (6 * dx) + np.diff(vc, axis=1) / (3 * dy) + np.diff(wc, axis=2) / (3 * dz)
    np.testing.assert_allclose(div_c, o.coarsen_block_mean(div_f, *f), rtol=0, atol=1e-12)
    # the coarse face value is the plain mean of the fine faces it covers
    assert abs(uc[1, 2, 1] - u[6, 6:9, 3:6].mean()) < 1e-14


@pytest.mark.slow
def test_config1_n_sound_6_is_unstable_and_7_is_not():
    """The acoustic Courant number of config 1 (dx=100 m, dt=1 s) at n_sound=6 is ~1.0:
    the FP64 oracle diverges within 6 steps; n_sound=7 stays bounded."""
    case, opts = o.case_config1()
    c = o.scaled_case(case, 8, 8, 64)
    phys, base, _ = o.setup_warm_bubble(c)
    with np.errstate(all="ignore"):
        for ns, expect_blowup in ((6, True), (7, False)):
            opts.n_sound = ns
            xi = phys
            wmax = 0.0
            for _ in range(6):
                xi, _ = o.ws_rk3_big_step(xi, base, c.grid, opts)
                wmax = float(np.nanmax(np.abs(xi[3]))) if np.isfinite(xi[3]).all() else np.inf
            assert (wmax > 100.0) == expect_blowup, (ns, wmax)


# ---------------------------------------------------------------- solver_opt 3
def test_tridiagonal_solve_matches_banded_solver():
    """Thomas restatement of jax.lax.linalg.tridiagonal_solve vs scipy's banded LU."""
    from scipy.linalg import solve_banded
    rng = np.random.default_rng(8)
    n = 37
    a = rng.normal(size=(3, 4, n))
    c = rng.normal(size=(3, 4, n))
    b = 3.0 + np.abs(a) + np.abs(c) + rng.uniform(size=(3, 4, n))
    d = rng.normal(size=(3, 4, n))
    a[..., 0] = 0.0
    c[..., -1] = 0.0
    x = o.tridiagonal_solve(a, b, c, d)
    for i in range(3):
        for j in range(4):
            ab = np.zeros((3, n))
            ab[0, 1:] = c[i, j, :-1]
            ab[1] = b[i, j]
            ab[2, :-1] = a[i, j, 1:]
            ref = solve_banded((1, 1), ab, d[i, j])
            assert np.allclose(x[i, j], ref, rtol=1e-12, atol=1e-12)


def _acoustic_inputs(nx=8, ny=6, nz=12, seed=3):
    gr = o.Grid(nx, ny, nz, 200.0, 200.0, 200.0)
    rng = np.random.default_rng(seed)
    z = (np.arange(nz + 2) - 0.5) * gr.dz
    th0 = np.broadcast_to(300.0 + 0.003 * z, (nx + 6, ny + 6, nz + 2)).copy()
    pi0 = np.broadcast_to(1.0 - 9.81 * z / (1004.64 * 300.0), (nx + 6, ny + 6, nz + 2)).copy()
    rho0 = np.broadcast_to(1.1 - 1e-4 * z, (nx + 6, ny + 6, nz + 2)).copy()
    u = o.padding3u_array(rng.normal(0, 1.0, (nx + 1, ny, nz)))
    v = o.padding3v_array(rng.normal(0, 1.0, (nx, ny + 1, nz)))
    wi = rng.normal(0, 0.5, (nx, ny, nz + 1))
    wi[:, :, 0] = wi[:, :, -1] = 0.0
    w = o.padding3_array(wi)
    pip = o.padding3_array(rng.normal(0, 1e-4, (nx, ny, nz)))
    pstar = o.padding3_array(rng.normal(0, 1e-4, (nx, ny, nz)))
    ru = rng.normal(0, 1e-2, (nx + 1, ny, nz))
    rv = rng.normal(0, 1e-2, (nx, ny + 1, nz))
    rw = rng.normal(0, 1e-2, (nx, ny, nz + 1))
    trho = th0 + o.padding3_array(rng.normal(0, 0.5, (nx, ny, nz)))
    return gr, (u, v, w, pi0, pip, pstar, ru, rv, rw, trho, rho0)


def test_implicit_acoustic_rest_state_stays_at_rest():
    gr, (u, v, w, pi0, pip, pstar, ru, rv, rw, trho, rho0) = _acoustic_inputs()
    z = [np.zeros_like(a) for a in (u, v, w, pip, ru, rv, rw)]
    out = o.sound_implicit_rk_step(z[0], z[1], z[2], pi0, z[3], z[3], z[4], z[5], z[6], trho, rho0, 3, 0.2, gr)
    assert all(np.max(np.abs(a)) == 0.0 for a in out)


def test_implicit_acoustic_substep_is_consistent_with_the_explicit_one():
    """One substep of the vertically implicit scheme differs from the forward-backward one
    by O(dts^2) (the off-centred implicit terms): halving dts divides the difference by ~4.
    A dropped or misplaced O(dts) term (wrong slice, missing rhs) breaks the ratio."""
    gr, (u, v, w, pi0, pip, pstar, ru, rv, rw, trho, rho0) = _acoustic_inputs()
    diffs = []
    for dts in (0.02, 0.01, 0.005):
        ei = o.sound_implicit_rk_step(u, v, w, pi0, pip, pstar, ru, rv, rw, trho, rho0, 1, dts, gr)
        ee = o.sound_rk_step(u, v, w, pi0, pip, pstar, ru, rv, rw, trho, rho0, 1, dts, gr)
        diffs.append((np.max(np.abs(ei[2] - ee[2])), np.max(np.abs(ei[3] - ee[3]))))
        # the explicit parts (u, v) are identical
        assert np.array_equal(ei[0], ee[0]) and np.array_equal(ei[1], ee[1])
    # w: ratio 4; pi' carries the O(dts^2) w difference one substep further, with O(dts^3)
    # terms still visible at these steps (ratio 4.5-5.1): bounds (3.5, 6)
    for q, hi in ((0, 4.5), (1, 6.0)):
        r1 = diffs[0][q] / diffs[1][q]
        r2 = diffs[1][q] / diffs[2][q]
        assert 3.5 < r1 < hi and 3.5 < r2 < hi, (q, diffs)


def test_config1_n_sound_6_is_stable_with_the_implicit_solver():
    """The vertically implicit acoustic step (solver_opt 3) removes the vertical acoustic
    Courant limit that makes config 1 (dx = dz = 100 m, dt = 1 s) unstable at n_sound 6
    with solver_opt 2 (same reduced column set as the solver_opt 2 check above)."""
    case, opts = o.case_config1()
    c = o.scaled_case(case, 8, 8, 64)
    phys, base, _ = o.setup_warm_bubble(c)
    opts.n_sound = 6
    opts.solver_opt = 3
    xi = phys
    with np.errstate(all="ignore"):
        for _ in range(6):
            xi, _ = o.ws_rk3_big_step(xi, base, c.grid, opts)
    assert np.isfinite(xi[3]).all() and float(np.max(np.abs(xi[3]))) < 100.0
