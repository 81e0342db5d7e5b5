"""Oracle restatement of the reference's post-step deep-learning correction maps
(dl_models.py:125-182, one_sprint_functions.py:137-160), pinned by closed forms."""
import numpy as np

from oracle import lex_oracle as o


def _scaling(shapes, rng):
    out = []
    for shp in shapes:
        base = rng.normal(0, 1, shp)
        out += [(base - rng.uniform(0.1, 0.5, shp))[None, None], (base + rng.uniform(0.1, 0.5, shp))[None, None]]
    return tuple(out) + (None,) * 4


def test_unstagger_of_linear_fields_is_exact():
    nx, ny, nz = 6, 5, 4
    xf = np.arange(nx + 1)[:, None, None] * np.ones((1, ny, nz))
    yf = np.arange(ny + 1)[None, :, None] * np.ones((nx, 1, nz))
    zf = np.arange(nz + 1)[None, None, :] * np.ones((nx, ny, 1))
    ui, vi, wi = o.unstagger_uvw(2.0 * xf + 1.0, -3.0 * yf, 0.5 * zf + 2.0)
    xc = np.arange(nx)[:, None, None] + 0.5
    assert np.array_equal(ui, np.broadcast_to(2.0 * xc + 1.0, (nx, ny, nz)))
    assert np.array_equal(vi, np.broadcast_to(-3.0 * (np.arange(ny)[None, :, None] + 0.5), (nx, ny, nz)))
    assert np.array_equal(wi, np.broadcast_to(0.5 * (np.arange(nz)[None, None, :] + 0.5) + 2.0, (nx, ny, nz)))


def test_stagger_end_faces_and_lids():
    """Interior faces average their centres; the end faces copy the reference's choice
    (first face = the last interior face, last face = the first interior face); w is zero
    at the surface and the lid."""
    rng = np.random.default_rng(1)
    nx, ny, nz = 5, 4, 6
    ui, vi, wi = (rng.normal(size=(nx, ny, nz)) for _ in range(3))
    u, v, w = o.stagger_uvw(ui, vi, wi)
    assert u.shape == (nx + 1, ny, nz) and v.shape == (nx, ny + 1, nz) and w.shape == (nx, ny, nz + 1)
    for f in range(1, nx):
        assert np.array_equal(u[f], 0.5 * (ui[f - 1] + ui[f]))
    assert np.array_equal(u[0], u[nx - 1]) and np.array_equal(u[nx], u[1])
    assert np.array_equal(v[:, 0], v[:, ny - 1]) and np.array_equal(v[:, ny], v[:, 1])
    assert not w[:, :, 0].any() and not w[:, :, nz].any()
    assert np.array_equal(w[:, :, 3], 0.5 * (wi[:, :, 2] + wi[:, :, 3]))


def test_normalisation_maps_the_level_extremes():
    rng = np.random.default_rng(2)
    nx, ny, nz = 4, 3, 5
    shapes = [(nx, ny, nz), (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz)]
    sp = _scaling(shapes, rng)
    norm = o.dl_level_norm(sp)
    for q, (lo, den) in enumerate(norm):
        mn, mx = np.squeeze(sp[2 * q]), np.squeeze(sp[2 * q + 1])
        assert np.array_equal(lo, mn.min(axis=(0, 1)))
        c = 1e-5 if q == 4 else 0.01
        assert np.allclose(den, mx.max(axis=(0, 1)) - lo + c, rtol=0, atol=1e-15)
    # a state sitting at the per-level minimum packs to zero (the centred velocities too)
    lo = [n[0] for n in norm]
    th = o.padding3_array(np.broadcast_to(lo[0], (nx, ny, nz)).copy())
    u = o.padding3u_array(np.broadcast_to(lo[1], (nx + 1, ny, nz)).copy())
    v = o.padding3v_array(np.broadcast_to(lo[2], (nx, ny + 1, nz)).copy())
    w = o.padding3_array(np.broadcast_to(lo[3][:nz], (nx, ny + 0, nz)).copy()[:, :, list(range(nz)) + [nz - 1]])
    w[3:-3, 3:-3, 1:-1] = lo[3]
    qv = o.padding3_array(np.broadcast_to(lo[4], (nx, ny, nz)).copy())
    x = o.phys_state_tuple2array((th, u, v, w, th * 0, qv), sp)
    assert x.shape == (nx, ny, nz, 5)
    assert np.abs(x).max() == 0.0


def test_correction_adds_correction_times_range():
    rng = np.random.default_rng(3)
    nx, ny, nz = 4, 3, 5
    shapes = [(nx, ny, nz), (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz)]
    sp = _scaling(shapes, rng)
    st = (o.padding3_array(rng.normal(300, 1, (nx, ny, nz))), o.padding3u_array(rng.normal(size=(nx + 1, ny, nz))),
          o.padding3v_array(rng.normal(size=(nx, ny + 1, nz))),
          o.set_w_bc(o.padding3_array(rng.normal(size=(nx, ny, nz + 1)))), o.padding3_array(rng.normal(size=(nx, ny, nz))),
          o.padding3_array(rng.uniform(0, 1e-2, (nx, ny, nz))))
    zero = o.dl_apply_correction(st, np.zeros((nx, ny, nz, 5)), sp)
    for a, b in zip(st, zero):
        assert np.array_equal(a, b)
    one_theta = np.zeros((nx, ny, nz, 5))
    one_theta[..., 0] = 1.0
    out = o.dl_apply_correction(st, one_theta, sp)
    rng_th = np.squeeze(sp[1] - sp[0])
    assert np.allclose(out[0][3:-3, 3:-3, 1:-1], st[0][3:-3, 3:-3, 1:-1] + rng_th, rtol=0, atol=1e-12)
    for a, b in zip(st[1:], out[1:]):
        assert np.array_equal(a, b)
