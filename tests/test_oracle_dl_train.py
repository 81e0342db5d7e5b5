"""Oracle restatement of the loss side of the post-step correction's training
(lex_dl_train.py:49-99, dataset_generation.py:115-320, one_sprint_functions.py:34-97),
pinned by closed forms:

* laplace_of_tensor: the discrete eigenvalues of periodic cosine modes in x, y and of
  the copy-padded (zero-flux) cosine modes in z, with the vertical term absent in the
  first and last x planes (the reference's indexing on its 5-D stacks); constants map to
  zero; the operator is symmetric;
* the loss: a constant offset c per field gives 0.5 c^2 times the summed weights, truth
  Laplacians equal to the prediction's give no Laplacian term, and a constant negative qv
  gives its magnitude times the summed penalty weights; the weights are the scaling
  range over the cubed (squared for the penalty) level range;
* the split: a permutation from numpy's default_rng, 75 % for training.
"""
import numpy as np

from oracle import lex_oracle as o

GR = o.Grid(10, 8, 12, 100.0, 80.0, 50.0)


def test_laplacian_eigenmodes():
    nx, ny, nz = GR.nx, GR.ny, GR.nz
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    mx, my, mz = 2, 3, 4
    ax = 2 * np.pi * mx / nx
    ay = 2 * np.pi * my / ny
    az = np.pi * mz / nz
    # x, y modes (z constant): eigenvalue everywhere
    s = np.cos(ax * i) * np.cos(ay * j)
    lam = -(4 / GR.dx ** 2) * np.sin(ax / 2) ** 2 - (4 / GR.dy ** 2) * np.sin(ay / 2) ** 2
    np.testing.assert_allclose(o.laplace_of_tensor(s[None, None], GR)[0, 0], lam * s, atol=1e-12)
    # a zero-flux z mode: eigenvalue except the first and last x planes, which lose the term
    s = np.cos(az * (k + 0.5))
    lz = -(4 / GR.dz ** 2) * np.sin(az / 2) ** 2
    got = o.laplace_of_tensor(s[None, None], GR)[0, 0]
    np.testing.assert_allclose(got[1:-1], lz * s[1:-1], atol=1e-12)
    assert np.max(np.abs(got[[0, -1]])) < 1e-12
    assert np.max(np.abs(o.laplace_of_tensor(np.full((2, 3, nx, ny, nz), 7.0), GR))) == 0.0


def test_laplacian_is_symmetric():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(GR.nx, GR.ny, GR.nz))
    b = rng.normal(size=(GR.nx, GR.ny, GR.nz))
    la = o.laplace_of_tensor(a, GR)
    lb = o.laplace_of_tensor(b, GR)
    assert abs(np.sum(la * b) - np.sum(a * lb)) < 1e-10 * np.sum(np.abs(la * b))


def _scaling(rng):
    nx, ny, nz = GR.nx, GR.ny, GR.nz
    shapes = [(nx, ny, nz), (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz), (nx, ny, nz),
              (nx, ny, nz)]
    out = []
    for shp in shapes:
        b = rng.normal(0, 1, shp)
        out += [b - rng.uniform(0.1, 0.5, shp), b + rng.uniform(0.1, 0.5, shp)]
    return tuple(out)


def test_loss_weights_and_closed_forms():
    rng = np.random.default_rng(3)
    sp = _scaling(rng)
    w = o.dl_loss_weights(sp)
    eps = (0.01, 0.01, 0.01, 0.01, 1e-5, 1e-7, 1e-10)
    for q in range(7):
        mn, mx = sp[2 * q], sp[2 * q + 1]
        lev = mx.max(axis=(0, 1)) - mn.min(axis=(0, 1)) + eps[q]
        np.testing.assert_allclose(w[q], (mx - mn) / lev ** 3, rtol=1e-14)
    lev = sp[9].max(axis=(0, 1)) - sp[8].min(axis=(0, 1)) + 1e-5
    np.testing.assert_allclose(w[7], (sp[9] - sp[8]) / lev ** 2, rtol=1e-14)
    B, S = 2, 3
    shapes = [(GR.nx, GR.ny, GR.nz), (GR.nx + 1, GR.ny, GR.nz), (GR.nx, GR.ny + 1, GR.nz),
              (GR.nx, GR.ny, GR.nz + 1), (GR.nx, GR.ny, GR.nz), (GR.nx, GR.ny, GR.nz)]
    stacks = [rng.normal(size=(B, S) + shp) for shp in shapes]
    stacks[5] = np.abs(stacks[5]) * 1e-3           # qv >= 0: no penalty
    c = (0.3, -0.2, 0.1, 0.05, 1e-4)
    truth = [stacks[0] + c[0], stacks[1] + c[1], stacks[2] + c[2], stacks[3] + c[3], stacks[4], stacks[5] + c[4]]
    truth += [o.laplace_of_tensor(stacks[0] - sp[0], GR), o.laplace_of_tensor(stacks[5] - sp[8], GR)]
    got = o.dl_loss(stacks, truth, sp, GR)
    exp = sum(0.5 * cq ** 2 * B * S * np.sum(w[q]) for q, cq in zip((0, 1, 2, 3, 4), c))
    assert abs(got - exp) < 1e-10 * exp
    # the Laplacian terms: an offset of the truth Laplacian by d costs 0.5 d^2 sum(w)
    truth2 = list(truth)
    truth2[6] = truth[6] + 2.0
    assert abs(o.dl_loss(stacks, truth2, sp, GR) - got - 0.5 * 4.0 * B * S * np.sum(w[5])) < 1e-8 * got
    # a constant negative qv: the penalty adds |qv| sum(w_neg), the l2 term is unchanged
    st3 = list(stacks)
    st3[5] = np.full_like(stacks[5], -2e-4)
    tr3 = list(truth)
    tr3[5] = st3[5] + c[4]
    tr3[7] = o.laplace_of_tensor(st3[5] - sp[8], GR)
    exp3 = exp + 2e-4 * B * S * np.sum(w[7])
    assert abs(o.dl_loss(st3, tr3, sp, GR) - exp3) < 1e-10 * exp3


def test_training_split():
    from lex_amd.dl_train import split_indices
    heads, n_train = split_indices(24, 7)
    assert sorted(heads.tolist()) == list(range(24)) and n_train == 18
    assert heads.tolist() != list(range(24))
    h2, _ = split_indices(24, 7)
    assert np.array_equal(heads, h2)
