"""The reference's post-step deep-learning correction on the GPU (ws_rk3_sprint_dl,
one_sprint_functions.py:118-165; dl_models.py:125-182) against the oracle."""
import numpy as np
import pytest

from oracle import lex_oracle as o
from tests._helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu

NX, NY, NZ = 12, 10, 14


def _scaling(rng, nx, ny, nz):
    shapes = [(nx, ny, nz), (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz)]
    scales = [1.0, 1.0, 1.0, 0.5, 1e-3]
    out = []
    for shp, sc in zip(shapes, scales):
        base = rng.normal(0, sc, shp)
        out += [(base - rng.uniform(0.1, 0.5, shp) * sc)[None, None].astype(np.float32),
                (base + rng.uniform(0.1, 0.5, shp) * sc)[None, None].astype(np.float32)]
    out[0] = out[0] + np.float32(300.0)
    out[1] = out[1] + np.float32(300.0)
    out[8] = out[8] + np.float32(5e-3)
    out[9] = out[9] + np.float32(5e-3)
    return tuple(out) + (None,) * 4


def _state(rng, nx, ny, nz):
    return (o.padding3_array(rng.normal(300, 0.5, (nx, ny, nz))).astype(np.float32),
            o.padding3u_array(rng.normal(0, 1, (nx + 1, ny, nz))).astype(np.float32),
            o.padding3v_array(rng.normal(0, 1, (nx, ny + 1, nz))).astype(np.float32),
            o.set_w_bc(o.padding3_array(rng.normal(0, 0.5, (nx, ny, nz + 1)))).astype(np.float32),
            o.padding3_array(rng.normal(0, 1e-3, (nx, ny, nz))).astype(np.float32),
            o.padding3_array(rng.uniform(0, 1e-2, (nx, ny, nz))).astype(np.float32))


def test_pack_stagger_and_apply_match_the_oracle():
    import torch
    from lex_amd import dl_models as dlm
    rng = np.random.default_rng(11)
    sp = _scaling(rng, NX, NY, NZ)
    st = _state(rng, NX, NY, NZ)
    dev = tuple(to_dev(a) for a in st)
    spd = tuple(None if a is None else to_dev(a) for a in sp)
    # phys_state_tuple2array: the same FP32 operations (within one rounding of the division)
    x = dlm.phys_state_tuple2array(dev, spd)
    xr = o.phys_state_tuple2array(st, sp)
    assert x.shape == (1, NX, NY, NZ, 5)
    xg = from_dev(x[0])
    for q in range(5):
        d = np.abs(xg[..., q] - xr[..., q].astype(np.float64))
        tol = 2.0 ** -23 * (np.abs(xr[..., q]) + np.abs(xr[..., q]).max())
        assert (d <= tol).all(), f"channel {q}: max |diff| {d.max():.3e}"
    # correct_array2tuple: averages only -> bit-equal
    corr = rng.normal(0, 0.1, (NX, NY, NZ, 5)).astype(np.float32)
    got = dlm.correct_array2tuple(to_dev(corr))
    ref = o.correct_array2tuple(corr)
    for name, a, b in zip(("theta", "u", "v", "w", "qv"), got, ref):
        assert a.shape == b.shape, name
        assert np.array_equal(from_dev(a).astype(np.float32), b.astype(np.float32)), name
    # the correction step: interior + correction x range, ghosts refilled (the product and
    # sum may fuse into one FMA: 1-ulp tolerance)
    ranges = dlm.scaling_ranges(spd, dev[0].device)
    dlm.apply_correction(dev, to_dev(corr), ranges)
    torch.cuda.synchronize()
    exp = o.dl_apply_correction(st, corr, sp)
    for name, a, b in zip(("theta", "u", "v", "w", "pip", "qv"), dev, exp):
        a = from_dev(a)
        assert a.shape == b.shape, name
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        assert err <= 2.0 ** -23, f"{name}: {err:.2e}"


def test_autoencoder_matches_the_oracle():
    """AutoEncoder (dl_models.py:33-68) in FP32 on the device vs the oracle's FP64
    restatement with the same weights, on a random 40^3 sample."""
    import torch
    from lex_amd import dl_models as dlm
    net = dlm.AutoEncoder(seed=3)
    rng = np.random.default_rng(8)
    with torch.no_grad():
        for b in net.biases:   # nonzero biases so every bias path is exercised
            b.copy_(torch.from_numpy(rng.normal(0, 0.05, b.shape[0]).astype(np.float32)))
    x = rng.normal(0, 1, (40, 40, 40, 5)).astype(np.float32)
    y = from_dev(net(to_dev(x)))
    params = [(from_dev(w), from_dev(b)) for w, b in zip(net.weights, net.biases)]
    ref = o.autoencoder_forward(params, x.astype(np.float64))
    assert y.shape == ref.shape == (40, 40, 40, 5)
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"AutoEncoder rel-L2 vs FP64 oracle: {err:.2e}")
    assert err < 1e-5, err
    with pytest.raises(ValueError):
        net(to_dev(x[:32]))


@pytest.mark.parametrize("model", ["simpleCNN", "Zcnn", "AutoEncoder"])
def test_sprint_dl_matches_the_oracle_loop(model):
    """Two steps of ws_rk3_sprint_dl (simpleCNN, random weights, correction about a tenth
    of the step's increment) vs the oracle's loop (FP64 steps) with the same network as
    net_fn: increments over the sprint agree to the FP32/BF16 budget."""
    import torch
    from lex_amd import dl_models as dlm, namelist_n_constants as nl, one_sprint_functions as sp
    from lex_amd.setup_lex import build_case, config1
    case, opts = config1()
    # the AutoEncoder's dense layer fixes its sample to 40^3
    case = case.scaled(40, 40, 40) if model == "AutoEncoder" else case.scaled(16, 12, 12)
    phys, base, grids, _ = build_case(case)
    nl.dt, nl.n_sound, nl.solver_opt, nl.turb_opt = opts["dt"], 7, 2, 1
    model_opt = (2, False, False, 0, False, False, 1)
    rng = np.random.default_rng(5)
    st0 = tuple(from_dev(a).astype(np.float32) for a in phys)
    spr = _scaling(rng, case.nx, case.ny, case.nz)
    spd = tuple(None if a is None else to_dev(a) for a in spr)
    net = getattr(dlm, model)(seed=7)
    with torch.no_grad():
        net.weights[-1].mul_(1e-3)
        net.biases[-1].fill_(1e-4)

    def net_fn(x):
        with torch.no_grad():
            return from_dev(net(to_dev(x.astype(np.float32))))

    old = nl.sprint_n
    nl.sprint_n = 2
    try:
        out, _ = sp.ws_rk3_sprint_dl(phys, base, grids, model_opt, net, spd)
    finally:
        nl.sprint_n = old
    torch.cuda.synchronize()
    og = o.Grid(case.nx, case.ny, case.nz, case.dx, case.dy, case.dz)
    oo = o.Options(dt=opts["dt"], solver_opt=2, n_sound=7, turb_opt=1)
    bs = tuple(from_dev(a).astype(np.float64) for a in base[:5]) + (None,)
    x64 = tuple(a.astype(np.float64) for a in st0)
    sp64 = tuple(None if a is None else a.astype(np.float64) for a in spr)
    ref = o.ws_rk3_sprint_dl(x64, bs, og, oo, net_fn, sp64, 2)
    # the same sprint in the FP32 oracle: its drift from FP64 is the floor of any FP32 run
    bs32 = tuple(None if a is None else a.astype(np.float32) for a in bs)
    sp32 = tuple(None if a is None else a.astype(np.float32) for a in spr)
    ref32 = o.ws_rk3_sprint_dl(tuple(a.astype(np.float32) for a in st0), bs32, og, oo, net_fn, sp32, 2)
    for name, a, r, r32, a0 in zip(("theta", "u", "v", "w", "pip", "qv"), out, ref, ref32, x64):
        d = from_dev(a).astype(np.float64) - a0
        rd = r - a0
        nrm = max(np.linalg.norm(rd), 1e-30)
        err = np.linalg.norm(d - rd) / nrm
        err32 = np.linalg.norm(np.asarray(r32, np.float64) - a0 - rd) / nrm
        print(f"{model} {name}: rel-L2 of the increment {err:.2e} (FP32 oracle {err32:.2e})")
        assert err < max(5e-3, 3 * err32), f"{name}: rel-L2 of the increment {err:.2e} (FP32 oracle {err32:.2e})"
