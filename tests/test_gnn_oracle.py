"""Pins the learned-preconditioner oracle (oracle/gnn.py) on the CPU:
central finite differences of its backward pass, the Neumann closed form of
its apply, and its features on a hand-computed case."""
import numpy as np
import pytest

import owtk
from oracle import gnn


def small_graph(nx=6, ny=4, b=3, seed=5, law=owtk.LAW_LOGNORMAL_FIELD, sigma=1.5, shift=0.05):
    s = owtk.generate_split(nx, ny, b, seed, ordering=owtk.ORDER_RCM, law=law, sigma=sigma, correlation=2.0,
                            shift=shift)
    return s, gnn.Graph(s.nodes, b, s.offsets, s.columns, s.diagonal, s.off_diagonal)


def pairs(g, count, seed):
    rng = np.random.default_rng(seed)
    res, err = [], []
    for _ in range(count):
        e = rng.uniform(-1, 1, (g.n, g.b))
        err.append(e)
        res.append(g.apply(e))
    return res, err


def test_layout_count():
    lay = gnn.Layout(3, 4)
    w, f0, o = gnn.WIDTH, gnn.feature_count(3), 15
    assert lay.count() == w * f0 + w + 3 * (2 * w * w + w) + o * w + 2 * o
    flat = np.arange(lay.count(), dtype=np.float64)
    assert np.array_equal(lay.pack(lay.unpack(flat)), flat)


def test_features_hand_case():
    # 2 vertices, b = 1: a01 = -2, a10 = -1, d = (4, 3)
    g = gnn.Graph(2, 1, [0, 1, 2], [1, 0], [4.0, 3.0], [-2.0, -1.0])
    x, w = gnn.features(g)
    assert x.shape == (2, gnn.feature_count(1))
    assert np.allclose(x[:, 0], [0.5, 1.0 / 3.0])
    assert np.allclose(x[:, 1], np.log([0.5, 2.0 / 3.0]) / 8)
    assert np.allclose(x[:, 2], [0.0, 0.0])        # one edge per row: no anisotropy
    assert np.allclose(x[:, 3], [1 / 8, 1 / 8])
    assert np.allclose(w, [0.5, 1.0 / 3.0])
    assert np.array_equal(g.mirror, [1, 0])
    # a path 0-1-2, b = 1: vertex 1 has edges of magnitude 1 and 3
    g = gnn.Graph(3, 1, [0, 1, 3, 4], [1, 0, 2, 1], [2.0, 5.0, 4.0], [-1.0, -1.0, -3.0, -3.0])
    x, _ = gnn.features(g)
    assert np.allclose(x[:, 2], [0.0, np.log(2 * 3 / 4) / 2, 0.0])
    assert np.allclose(x[:, 0], [0.5, 0.8, 0.75])
    assert np.allclose(x[:, 3], [1 / 8, 2 / 8, 1 / 8])


def test_neumann_closed_form():
    s, g = small_graph(law=owtk.LAW_UNIFORM, sigma=0.0, shift=2.0)
    r = np.random.default_rng(1).uniform(-1, 1, (g.n, g.b))
    # truncated Neumann series of the Jacobi splitting -> A^-1 r
    z = gnn.apply(g, np.ones((400, g.n, g.b)), r)
    for c in range(g.b):
        dense = np.diag(g.diag[:, c])
        np.add.at(dense, (g.rows, g.columns), g.off[:, c])
        ref = np.linalg.solve(dense, r[:, c])
        assert np.allclose(z[:, c], ref, rtol=0, atol=1e-9 * np.abs(ref).max())
    # degree 1 by hand
    z1 = gnn.apply(g, np.stack([np.full((g.n, g.b), 2.0), np.full((g.n, g.b), 3.0)]), r)
    u0 = r / g.diag
    assert np.allclose(z1, 2 * u0 - 3 * g.off_apply(u0) / g.diag)


def test_transpose_operators():
    s, g = small_graph()
    rng = np.random.default_rng(2)
    x, y = rng.normal(size=(g.n, g.b)), rng.normal(size=(g.n, g.b))
    assert np.isclose((g.apply(x) * y).sum(), (x * g.apply_t(y)).sum())
    _, w = gnn.features(g)
    hx, hy = rng.normal(size=(g.n, 5)), rng.normal(size=(g.n, 5))
    assert np.isclose((g.aggregate(w, hx) * hy).sum(), (hx * g.aggregate_t(w, hy)).sum())


@pytest.mark.parametrize("degree", [2, 4])
def test_backward_finite_differences(degree):
    s, g = small_graph()
    flat = gnn.init_parameters(g.b, degree, seed=3)
    lay = gnn.Layout(g.b, degree)
    p = lay.unpack(flat)
    rng = np.random.default_rng(4)
    # a non-trivial point: decoder and g away from their initial values
    p["w_out"] = rng.normal(0, 0.3, p["w_out"].shape)
    p["b_out"] = rng.normal(0, 0.3, p["b_out"].shape)
    p["g"] = rng.uniform(0.5, 2.0, p["g"].shape)
    for layer in range(gnn.LAYERS):
        p[f"b{layer}"] = rng.normal(0, 0.2, p[f"b{layer}"].shape)
    flat = lay.pack(p)
    res, err = pairs(g, 3, 6)
    loss, grad = gnn.loss_and_grad(g, flat, degree, res, err)
    assert np.isfinite(loss) and loss > 0
    idx = rng.choice(flat.size, 60, replace=False)
    # every parameter block gets at least one probe
    starts = np.cumsum([0] + [int(np.prod(sh)) for _, sh in lay.blocks()])
    idx = np.unique(np.concatenate([idx, starts[:-1]]))
    h = 1e-6
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        lp, _ = gnn.loss_and_grad(g, fp, degree, res, err)
        lm, _ = gnn.loss_and_grad(g, fm, degree, res, err)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7, (i, fd, grad[i])


def test_exact_error_pairs_zero_loss_for_exact_inverse():
    # with e = A^-1 r and z = e the loss vanishes: check loss_and_dgamma's
    # target and normalization on a degree-0 diagonal system
    g = gnn.Graph(3, 1, [0, 0, 0, 0], [], [2.0, 4.0, 5.0], np.zeros((0, 1)))
    r = [np.array([[1.0], [2.0], [3.0]])]
    e = [r[0] / g.diag]
    loss, dg = gnn.loss_and_dgamma(g, np.ones((1, 3, 1)), r, e)
    assert loss == 0.0 and np.all(dg == 0.0)
    loss, _ = gnn.loss_and_dgamma(g, np.full((1, 3, 1), 2.0), r, e)
    assert np.isclose(loss, 1.0)


def test_bf16_round():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0], np.float32)
    out = gnn.bf16_round(vals)
    assert out[0] == 1.0 and out[1] == 1.0  # tie to even
    assert out[2] == 1.0078125
    assert abs(out[3] + 3.140625) < 1e-12


def test_committed_lognormal_model_modulates_per_vertex():
    """The committed c3 model (owt-krylov_amd/owtk/models/gnn_lognormal.npz)
    gives coefficients that vary across the vertices of an unseen lognormal
    field (the network is not a constant polynomial), evaluated by the
    oracle's forward pass."""
    import pathlib
    path = pathlib.Path(owtk.__file__).resolve().parent / "models" / "gnn_lognormal.npz"
    with np.load(path) as f:
        params, degree = f["params"], int(f["degree"])
    assert params.size == gnn.Layout(3, degree).count()
    s = owtk.generate_split(128, 64, 3, 2000, ordering=owtk.ORDER_HILBERT, law=owtk.LAW_LOGNORMAL_FIELD, sigma=1.5,
                            correlation=16.0)
    g = gnn.Graph(s.nodes, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    gam, _ = gnn.forward(g, gnn.Layout(3, degree).unpack(gnn.bf16_round(params)), degree)
    spread = gam.std(axis=1) / np.abs(gam.mean(axis=1))  # (degree + 1, b)
    print("relative spread per degree:", spread.mean(axis=1))
    assert np.all(spread[2:] > 0.1)
    # and the decoder carries the modulation: zeroing it leaves one polynomial g
    lay = gnn.Layout(3, degree)
    p = lay.unpack(params)
    p["w_out"][:] = 0.0
    p["b_out"][:] = 0.0
    flat = gnn.init_parameters(3, degree) * 0 + lay.pack(p)
    gam0, _ = gnn.forward(g, lay.unpack(gnn.bf16_round(flat)), degree)
    assert np.all(gam0.std(axis=1) == 0.0)
