"""Learned vertex-block preconditioner on the GPU against the CPU oracle
(oracle/gnn.py): network inference, loss and gradient of the HIP training
kernels, the preconditioner apply, training progress, and use inside FGMRES."""
import pathlib

import numpy as np
import pytest

import owtk
from oracle import gnn

pytestmark = pytest.mark.gpu

MODELS = pathlib.Path(owtk.__file__).resolve().parent / "models"


def system(nx=24, ny=16, seed=7, sigma=1.5, law=owtk.LAW_LOGNORMAL_FIELD):
    return owtk.generate_split(nx, ny, 3, seed, ordering=owtk.ORDER_HILBERT, law=law, sigma=sigma, correlation=4.0)


def oracle_graph(s):
    return gnn.Graph(s.nodes, s.block, s.offsets, s.columns, s.diagonal, s.off_diagonal)


def perturbed_parameters(b, degree, seed):
    lay = gnn.Layout(b, degree)
    p = lay.unpack(gnn.init_parameters(b, degree, seed))
    rng = np.random.default_rng(seed + 1)
    p["w_out"] = rng.normal(0, 0.3, p["w_out"].shape)
    p["b_out"] = rng.normal(0, 0.3, p["b_out"].shape)
    p["g"] = rng.uniform(0.5, 2.0, p["g"].shape)
    for layer in range(gnn.LAYERS):
        p[f"b{layer}"] = rng.normal(0, 0.2, p[f"b{layer}"].shape)
    return lay.pack(p).astype(np.float32)


@pytest.fixture(scope="module")
def lognormal(ctx):
    s = system()
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m.add_storage(owtk.F32)
    m.add_storage(owtk.BF16)
    return s, m


@pytest.mark.parametrize("degree", [2, 4])
def test_coefficients_match_oracle(ctx, lognormal, degree):
    s, m = lognormal
    params = perturbed_parameters(3, degree, 11)
    model = owtk.GnnModel(ctx, 3, degree)
    model.set_parameters(params)
    gam = model.coefficients(m).astype(np.float64)
    g = oracle_graph(s)
    ref, _ = gnn.forward(g, gnn.Layout(3, degree).unpack(gnn.bf16_round(params)), degree)
    err = np.abs(gam - ref).max() / np.abs(ref).max()
    print(f"degree {degree}: max relative coefficient error {err:.2e}")
    assert err < 2e-5


def test_fresh_model_is_neumann(ctx, lognormal):
    s, m = lognormal
    model = owtk.GnnModel(ctx, 3, 3)
    assert np.all(model.coefficients(m) == 1.0)


@pytest.mark.parametrize("degree", [1, 4])
def test_loss_and_gradient_match_oracle(ctx, lognormal, degree):
    s, m = lognormal
    g = oracle_graph(s)
    n = s.nodes * 3
    rng = np.random.default_rng(21)
    errs = [rng.uniform(-1, 1, (s.nodes, 3)) for _ in range(3)]
    ress = [g.apply(e) for e in errs]
    R = ctx.upload(np.concatenate([r.ravel() for r in ress]))
    E = ctx.upload(np.concatenate([e.ravel() for e in errs]))
    params = perturbed_parameters(3, degree, 5)
    model = owtk.GnnModel(ctx, 3, degree)
    model.set_parameters(params)
    prob = model.problem(m, R, E, 3)
    model.zero_grad()
    model.accumulate(prob, 1.0)
    loss, grad = model.read()
    ref_loss, ref_grad = gnn.loss_and_grad(g, gnn.bf16_round(params), degree, ress, errs)
    rel_loss = abs(loss - ref_loss) / ref_loss
    lay = gnn.Layout(3, degree)
    gg, rg = lay.unpack(grad.astype(np.float64)), lay.unpack(ref_grad)
    worst = max(np.linalg.norm(gg[k] - rg[k]) / max(np.linalg.norm(rg[k]), 1e-30) for k in rg)
    total = np.linalg.norm(grad - ref_grad) / np.linalg.norm(ref_grad)
    print(f"degree {degree}: loss {loss:.6e} oracle {ref_loss:.6e} (rel {rel_loss:.1e}); "
          f"gradient rel error total {total:.2e}, worst block {worst:.2e}")
    assert rel_loss < 1e-4
    assert total < 1e-3
    assert worst < 5e-3
    # weights add up: two accumulations at weight 1/2 equal one at weight 1
    model.zero_grad()
    model.accumulate(prob, 0.5)
    model.accumulate(prob, 0.5)
    loss2, grad2 = model.read()
    assert abs(loss2 - loss) <= 1e-6 * loss
    assert np.allclose(grad2, grad, rtol=1e-4, atol=1e-6 * np.abs(grad).max())


def test_learned_apply_matches_oracle(ctx, lognormal):
    s, m = lognormal
    g = oracle_graph(s)
    degree = 4
    model = owtk.GnnModel(ctx, 3, degree)
    model.set_parameters(perturbed_parameters(3, degree, 3))
    gam = model.coefficients(m).astype(np.float64)
    r = np.random.default_rng(4).uniform(-1, 1, s.nodes * 3)
    ref = gnn.apply(g, gam, r.reshape(-1, 3)).ravel()
    pc64 = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model)
    x, y = ctx.upload(r), ctx.empty(r.size)
    pc64.apply(x, y)
    err64 = np.abs(y.download() - ref).max() / np.abs(ref).max()
    pc64.apply(x, x)  # in place
    assert np.array_equal(x.download(), y.download())
    for storage, tol in ((owtk.F32, 2e-5), (owtk.BF16, 2e-2)):
        pc = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model, storage=storage)
        x32, y32 = ctx.upload(r.astype(np.float32)), ctx.empty(r.size, np.float32)
        pc.apply(x32, y32)
        err = np.abs(y32.download() - ref).max() / np.abs(ref).max()
        print(f"storage {storage}: relative error {err:.2e}")
        assert err < tol
    print(f"FP64 apply relative error {err64:.2e}")
    assert err64 < 1e-13
    # numeric update after a value change keeps matching
    assert pc64.info().apply_bytes > 0


def test_training_reduces_loss(ctx):
    from owtk import learn
    model = owtk.GnnModel(ctx, 3, 3)
    problems = learn.build_problems(ctx, model, 4, 48, 24, 1.5, 100, correlation=4.0)
    held = learn.build_problems(ctx, model, 1, 48, 24, 1.5, 200, correlation=4.0)
    before = learn.evaluation_loss(model, held)
    hist, record = learn.train(model, problems, epochs=40, lr=1e-2, batch=2, lr_final=1e-3)
    after = learn.evaluation_loss(model, held)
    print(f"train loss {hist[0]:.4f} -> {hist[-1]:.4f}; held-out {before:.4f} -> {after:.4f}")
    assert hist[-1] < 0.95 * hist[0]
    assert after < 0.95 * before
    assert record == []


def test_training_selects_best_scored_checkpoint(ctx):
    """train(select=...) restores the parameters of the best score: with a
    score that is the held-out loss itself, the returned model's loss is the
    minimum of the recorded scores."""
    from owtk import learn
    model = owtk.GnnModel(ctx, 3, 3)
    problems = learn.build_problems(ctx, model, 3, 48, 24, 1.5, 300, correlation=4.0)
    held = learn.build_problems(ctx, model, 1, 48, 24, 1.5, 400, correlation=4.0)
    _, record = learn.train(model, problems, epochs=12, lr=1e-2, batch=2, lr_final=1e-3,
                            select=lambda mdl: learn.evaluation_loss(mdl, held), select_every=3)
    scores = [r["score"] for r in record if "epoch" in r]
    chosen = record[-1]
    print(record)
    assert len(scores) == 5 and chosen["score"] == min(scores)
    assert learn.evaluation_loss(model, held) == pytest.approx(chosen["score"], rel=1e-6)


def test_learned_preconditioned_fgmres(ctx, lognormal):
    s, m = lognormal
    model = owtk.GnnModel(ctx, 3, 3)
    b = ctx.upload(s.rhs)
    x = ctx.empty(s.nodes * 3)
    r = ctx.empty(s.nodes * 3)
    for storage in (owtk.F64, owtk.F32):
        pc = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model, storage=storage)
        if storage == owtk.F64:
            x.zero()
            res = owtk.fgmres(m, pc, b, x, owtk.SolverOptions(relative_tolerance=1e-10, restart=30))
        else:
            x.zero()
            res = owtk.fgmres_refined(m, pc, b, x, owtk.refinement_options(relative_tolerance=1e-10,
                                                                          inner_storage=owtk.F32))
        m.residual(b, x, r)
        rel = ctx.norm(r) / ctx.norm(b)
        print(f"storage {storage}: {res['status_name']} after {res['iterations']} iterations, true rel {rel:.2e}")
        assert res["status_name"] == "converged"
        assert rel <= 1.5e-10


def test_trained_network_gains_on_unseen_full_size_fields(ctx):
    """c3 evaluation size (1024x512, 524,288 vertices), two unseen lognormal
    fields: the committed network against the same model with its decoder
    zeroed (gamma = g, the learned global polynomial) and against its own
    coefficients averaged over the vertices (one polynomial, no per-vertex
    modulation), all at identical apply cost, FGMRES(30) to rtol 1e-8."""
    from owtk import learn
    model = owtk.GnnModel.load(ctx, MODELS / "gnn_lognormal.npz")
    for seed in (2000, 2001):
        s, m = learn.instance(ctx, 1024, 512, seed, 1.5)
        row = learn.compare(ctx, model, s, m)
        its = {k: v["iterations"] for k, v in row.items()}
        print(seed, its, {k: round(v["ms"], 1) for k, v in row.items()})
        for name in ("learned", "vertex_average", "no_network"):
            assert row[name]["status"] == "converged"
            assert row[name]["true_rel"] <= 1e-8
        assert its["learned"] <= 0.9 * its["no_network"]
        assert its["learned"] < its["vertex_average"]
        assert its["learned"] < its["ilu0"]


def test_fallback_to_ilu_when_learned_fails(ctx, lognormal):
    """A network whose coefficients are all zero cannot precondition: the true
    FP64 residual rejects its step and block-Jacobi ILU(0) takes over."""
    s, m = lognormal
    n = s.nodes * 3
    b = ctx.upload(s.rhs)
    x, r = ctx.empty(n), ctx.empty(n)
    ilu = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=128, storage=owtk.F32)
    opts = owtk.refinement_options(relative_tolerance=1e-9, inner_storage=owtk.F32)
    good = owtk.GnnModel(ctx, 3, 2)
    pc = owtk.Preconditioner(m, owtk.PC_LEARNED, model=good, storage=owtk.F32)
    res = owtk.fgmres_refined(m, pc, b, x, opts, fallback=ilu)
    assert res["status_name"] == "converged" and res["fallback_refinements"] == 0
    bad = owtk.GnnModel(ctx, 3, 2)
    p = bad.get_parameters()
    p[-9:] = 0.0  # g = 0: gamma = 0
    bad.set_parameters(p)
    assert np.all(bad.coefficients(m) == 0.0)
    pc_bad = owtk.Preconditioner(m, owtk.PC_LEARNED, model=bad, storage=owtk.F32)
    x.zero()
    res = owtk.fgmres_refined(m, pc_bad, b, x, opts, fallback=ilu)
    m.residual(b, x, r)
    rel = ctx.norm(r) / ctx.norm(b)
    print(res["status_name"], res["refinement_steps"], res["fallback_refinements"], rel)
    assert res["status_name"] == "converged"
    assert res["fallback_refinements"] >= 1
    assert rel <= 1.5e-9


@pytest.mark.parametrize("storage,tol", [(owtk.F64, 1e-13), (owtk.F32, 2e-5), (owtk.BF16, 2e-2)])
def test_learned_apply_follows_value_update(ctx, storage, tol):
    """After new matrix values and update_values(), the apply (column-scaled
    values and scaled coefficients rebuilt) matches the oracle on the new
    values."""
    s = system(seed=11)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    if storage != owtk.F64:
        m.add_storage(storage)
    degree = 3
    model = owtk.GnnModel(ctx, 3, degree)
    model.set_parameters(perturbed_parameters(3, degree, 5))
    pc = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model, storage=storage)
    diag2, off2 = s.diagonal * 1.3, s.off_diagonal * 0.8
    m.update_values(diag2, off2)
    pc.update_values()
    g = gnn.Graph(s.nodes, s.block, s.offsets, s.columns, diag2, off2)
    gam = model.coefficients(m).astype(np.float64)
    r = np.random.default_rng(12).uniform(-1, 1, s.nodes * 3)
    ref = gnn.apply(g, gam, r.reshape(-1, 3)).ravel()
    dt = np.float64 if storage == owtk.F64 else np.float32
    x, y = ctx.upload(r.astype(dt)), ctx.empty(r.size, dt)
    pc.apply(x, y)
    err = np.abs(y.download() - ref).max() / np.abs(ref).max()
    print(f"storage {storage}: relative error after the value update {err:.2e}")
    assert err < tol
