"""Pins of the oracle's buoyancy, pressure-equation operator, RHS and BiCGSTAB (CPU)."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import lex_oracle as o
from tests._helpers import random_state


def test_extrapolation_exact_for_quadratics():
    """(1.875, -1.25, 0.375) is the Lagrange value half a level beyond the end level."""
    rng = np.random.default_rng(0)
    c = rng.normal(size=3)
    z = np.arange(6) + 0.5                         # levels at 0.5, 1.5, ...
    f = c[0] + c[1] * z + c[2] * z ** 2
    b, t = o.extrapolate_bottom_top(f[None, None, :])
    assert abs(b[0, 0, 0] - c[0]) < 1e-12                               # z = 0
    zt = 6.0
    assert abs(t[0, 0, 0] - (c[0] + c[1] * zt + c[2] * zt ** 2)) < 1e-12  # z = 6


def test_buoyancy_closed_forms():
    gr, phys, base = random_state(8, 8, 8, seed=1)
    th0, qv0 = base[2], base[4]
    b, b8w = o.calculate_buoyancy(th0, th0.copy(), qv0, qv0.copy())
    assert np.max(np.abs(b)) == 0.0 and np.max(np.abs(b8w)) == 0.0
    # dry, small perturbation: B ~ g theta'/theta0 to first order
    z = np.zeros_like(th0)
    tp = 1e-3 * np.ones_like(th0)
    b, _ = o.calculate_buoyancy(np.full_like(th0, 300.0), 300.0 + tp, z, z)
    np.testing.assert_allclose(b, o.g * 1e-3 / 300.0, rtol=1e-9)   # (300+1e-3)-300 loses 5 digits


def test_matvec_linear_part_symmetric_and_affine_constant():
    gr, phys, base = random_state(10, 8, 9, seed=2)
    th, qv = phys[0], phys[5]
    b, _ = o.calculate_buoyancy(base[2], th, base[4], qv)
    rtt = o.calculate_rtt(base[0], th, qv, b)
    A = lambda x: o.laplace_of_pressure(rtt, x, gr)  # noqa: E731
    c = A(np.zeros((gr.nx, gr.ny, gr.nz)))
    # constant part: only the boundary z-fluxes taken from the rtt ghosts
    assert np.max(np.abs(c[:, :, 1:-1])) == 0.0
    np.testing.assert_allclose(c[:, :, 0], -o.Cp * rtt[3:-3, 3:-3, 0] / gr.dz, rtol=1e-12)
    np.testing.assert_allclose(c[:, :, -1], o.Cp * rtt[3:-3, 3:-3, -1] / gr.dz, rtol=1e-12)
    rng = np.random.default_rng(3)
    x, y = rng.normal(size=(2, gr.nx, gr.ny, gr.nz))
    Lx, Ly = A(x) - c, A(y) - c
    assert abs(np.vdot(Lx, y) - np.vdot(x, Ly)) < 1e-9 * abs(np.vdot(Lx, y))
    # constants are in the null space of the linear part
    assert np.max(np.abs(A(np.ones_like(x)) - c)) < 1e-9 * np.max(np.abs(Lx))


def test_matvec_uniform_rtt_fourier_eigenvalues():
    gr = o.Grid(12, 8, 6, 100.0, 150.0, 80.0)
    R = 3.0e4
    rtt = np.full(gr.shape_c, R)
    rtt[:, :, 0] = 0.0
    rtt[:, :, -1] = 0.0
    i = np.arange(gr.nx)[:, None, None]
    j = np.arange(gr.ny)[None, :, None]
    for m, n in [(1, 0), (2, 3), (5, 1)]:
        x = np.cos(2 * np.pi * (m * i / gr.nx + n * j / gr.ny)) * np.ones((1, 1, gr.nz))
        lam = o.Cp * R * ((2 * np.cos(2 * np.pi * m / gr.nx) - 2) / gr.dx ** 2
                          + (2 * np.cos(2 * np.pi * n / gr.ny) - 2) / gr.dy ** 2)
        np.testing.assert_allclose(o.laplace_of_pressure(rtt, x, gr), lam * x, rtol=1e-10, atol=1e-9 * abs(lam))


def test_rhs_pieces():
    gr, phys, base = random_state(8, 8, 8, seed=5)
    rt = base[0]
    z = lambda *s: np.zeros(s)  # noqa: E731
    # uniform adv4u, zero elsewhere -> zero divergence
    au = np.full((gr.nx + 1, gr.ny, gr.nz), 0.3)
    r = o.rhs_of_pressure_equation(rt, au, z(gr.nx, gr.ny + 1, gr.nz), z(gr.nx, gr.ny, gr.nz + 1),
                                   z(gr.nx, gr.ny + 1, gr.nz), z(gr.nx + 1, gr.ny, gr.nz), z(gr.nx, gr.ny, gr.nz), gr)
    assert np.max(np.abs(r)) < 1e-9
    # the buoyancy term telescopes to the extrapolated end values
    rng = np.random.default_rng(6)
    B = rng.normal(size=(gr.nx, gr.ny, gr.nz))
    r = o.rhs_of_pressure_equation(rt, z(gr.nx + 1, gr.ny, gr.nz), z(gr.nx, gr.ny + 1, gr.nz),
                                   z(gr.nx, gr.ny, gr.nz + 1), z(gr.nx, gr.ny + 1, gr.nz), z(gr.nx + 1, gr.ny, gr.nz),
                                   B, gr)
    rb = o._interior(rt) * B
    bot, top = o.extrapolate_bottom_top(rb)
    np.testing.assert_allclose(r.sum(axis=2) * gr.dz, (top - bot)[:, :, 0], rtol=1e-10, atol=1e-12)


def _poisson_matrix(gr, R):
    """Dense/sparse 7-point operator of the linear part with uniform rtt (checks against scipy)."""
    n = gr.nx * gr.ny * gr.nz
    idx = np.arange(n).reshape(gr.nx, gr.ny, gr.nz)
    rows, cols, vals = [], [], []
    for (di, dj, dk, h) in [(1, 0, 0, gr.dx), (-1, 0, 0, gr.dx), (0, 1, 0, gr.dy), (0, -1, 0, gr.dy),
                            (0, 0, 1, gr.dz), (0, 0, -1, gr.dz)]:
        src = idx
        dst = np.roll(idx, (-di, -dj, -dk), axis=(0, 1, 2))
        mask = np.ones_like(idx, bool)
        if dk == 1:
            mask[:, :, -1] = False
        if dk == -1:
            mask[:, :, 0] = False
        c = o.Cp * R / h ** 2
        rows += [src[mask].ravel(), src[mask].ravel()]
        cols += [dst[mask].ravel(), src[mask].ravel()]
        vals += [np.full(mask.sum(), c), np.full(mask.sum(), -c)]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def test_matvec_equals_assembled_operator():
    gr = o.Grid(6, 6, 6, 100.0, 120.0, 90.0)
    R = 2.0e4
    rtt = np.full(gr.shape_c, R)
    rtt[:, :, 0] = 0.0
    rtt[:, :, -1] = 0.0
    M = _poisson_matrix(gr, R)
    x = np.random.default_rng(7).normal(size=(gr.nx, gr.ny, gr.nz))
    np.testing.assert_allclose(o.laplace_of_pressure(rtt, x, gr).ravel(), M @ x.ravel(), rtol=1e-10, atol=1e-8)


def test_bicgstab_solves_consistent_linear_system():
    """Linear (non-affine) operator: the JAX-semantics BiCGSTAB matches scipy's direct solve."""
    gr = o.Grid(6, 6, 6, 100.0, 120.0, 90.0)
    R = 2.0e4
    rtt = np.full(gr.shape_c, R)
    rtt[:, :, 0] = 0.0
    rtt[:, :, -1] = 0.0
    A = lambda x: o.laplace_of_pressure(rtt, x, gr)  # noqa: E731
    rng = np.random.default_rng(8)
    xt = rng.normal(size=(gr.nx, gr.ny, gr.nz))
    xt -= xt.mean()
    b = A(xt)
    x, k = o.bicgstab(A, b, np.zeros_like(b), 1e-10, 0.0, 500)
    assert k > 0
    x -= x.mean()
    np.testing.assert_allclose(x, xt, rtol=0, atol=1e-6 * np.abs(xt).max())
    # gauge-fixed direct solve with scipy agrees
    M = _poisson_matrix(gr, R).tolil()
    M[0, :] = 0.0
    M[0, 0] = 1.0
    bb = b.ravel().copy()
    bb[0] = xt.ravel()[0]
    xd = spla.spsolve(M.tocsr(), bb).reshape(b.shape)
    np.testing.assert_allclose(xd - xd.mean(), xt, atol=1e-8 * np.abs(xt).max())


def test_bicgstab_jax_exit_semantics():
    gr = o.Grid(6, 6, 6, 100.0, 100.0, 100.0)
    rtt = np.full(gr.shape_c, 1.0e4)
    A = lambda x: o.laplace_of_pressure(rtt, x, gr)  # noqa: E731
    b = np.zeros((gr.nx, gr.ny, gr.nz))
    x0 = np.zeros_like(b)
    # |r0|^2 = |A(0)|^2 vs atol: with rtt ghosts nonzero the loop runs, maxiter caps it
    x, k = o.bicgstab(A, b, x0, 1e-4, 1e-8, 3)
    assert k in (3, -10, -11)
    # maxiter = 0: no iteration, x0 returned
    x, k = o.bicgstab(A, b + 1.0, x0, 1e-4, 1e-8, 0)
    assert k == 0 and np.array_equal(x, x0)
    # converged start: no iteration
    rtt[:, :, 0] = rtt[:, :, -1] = 0.0
    x, k = o.bicgstab(A, b, x0, 1e-4, 1e-8, 10)
    assert k == 0
