"""Device-resident solvers against the oracle and long-double recurrences."""
import numpy as np
import pytest

import owtk
import refcases as rc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def solve(ctx, fn, m, pc, b, x0=None, **kw):
    bd = ctx.upload(b)
    xd = ctx.upload(np.zeros(m.local_size) if x0 is None else np.asarray(x0, float))
    res = fn(m, pc, bd, xd, owtk.SolverOptions(**kw))
    return xd.download(), res


@pytest.mark.parametrize("use_ilu", [False, True])
def test_bicgstab_recurrence_audit(ctx, use_ilu):
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    m = owtk.Matrix.csr(ctx, 12, 0, 1, offsets, columns, values)
    pattern = [[a[i, j] != 0 for j in range(12)] for i in range(12)]
    lower, upper = rc.doolittle(a, pattern)
    inverse = (lambda r: rc.lu_solve(lower, upper, r)) if use_ilu else (lambda r: r.copy())
    history = rc.dense_bicgstab(a, rhs, x0, inverse, 4)
    pc = owtk.Preconditioner(m, owtk.PC_ILU) if use_ilu else None
    for count in range(1, 5):
        x, res = solve(ctx, owtk.bicgstab, m, pc, rhs, x0, relative_tolerance=0.0, maximum_iterations=count,
                       convergence_check_interval=1, residual_replacement_interval=0)
        assert res["status_name"] == "maximum_iterations" and res["iterations"] == count
        assert np.max(np.abs(x - history[count - 1].astype(float))) < 2e-11
        assert res["preconditioner_applications"] == 2 * count
        assert res["global_reductions"] == 2 + 6 * count + 1


def test_bicgstab_exact_preconditioner(ctx):
    m = owtk.Matrix.csr(ctx, 3, 0, 1, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [4.0, -1, -2, 5, -1, -1, 3])
    om = orc.OMatrix.csr(3, 0, 1, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [4.0, -1, -2, 5, -1, -1, 3])
    rhs = om.spmv([1.0, 2.0, 3.0])
    pc = owtk.Preconditioner(m, owtk.PC_ILU)
    x, res = solve(ctx, owtk.bicgstab, m, pc, rhs, relative_tolerance=32 * EPS, convergence_check_interval=1)
    assert res["status_name"] == "converged" and res["iterations"] == 1
    assert res["preconditioner_applications"] == 1
    assert np.max(np.abs(x - [1, 2, 3])) < 128 * EPS


def test_bicgstab_breakdown_zero_operator(ctx):
    m = owtk.Matrix.csr(ctx, 1, 0, 1, [0, 1], [0], [0.0])
    _, res = solve(ctx, owtk.bicgstab, m, None, np.array([1.0]), convergence_check_interval=1)
    assert res["status_name"] == "breakdown" and res["reason_name"] == "alpha_denominator"


def test_bicgstab_nonfinite_rhs(ctx):
    m = owtk.Matrix.csr(ctx, 1, 0, 2, [0, 1], [0], [1.0, 1.0])
    _, res = solve(ctx, owtk.bicgstab, m, None, np.array([np.inf, 1.0]))
    assert res["status_name"] == "breakdown" and res["reason_name"] == "non_finite_scalar"


def test_bicgstab_c0_jacobi_matches_oracle(ctx, c0):
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    xo, ro = orc.bicgstab(om, orc.OPreconditioner(om, 1), s.rhs)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    x, res = solve(ctx, owtk.bicgstab, m, owtk.Preconditioner(m, owtk.PC_JACOBI), s.rhs)
    print("oracle", ro["iterations"], ro["relative_residual_norm"], "gpu", res["iterations"],
          res["relative_residual_norm"])
    assert res["status_name"] == "converged" and ro["status"] == 0
    assert abs(res["iterations"] - ro["iterations"]) <= 1
    true_rel = np.linalg.norm(s.rhs - om.spmv(x)) / np.linalg.norm(s.rhs)
    assert true_rel <= 1e-8
    assert abs(res["relative_residual_norm"] - true_rel) <= 1e-3 * true_rel
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-7
    if res["iterations"] == ro["iterations"]:
        for key in ("operator_applications", "preconditioner_applications", "global_reductions"):
            assert res[key] == ro[key], key


@pytest.mark.parametrize("pc_kind", ["ilu", "ssor"])
def test_bicgstab_c0_preconditioned_matches_oracle(ctx, c0, pc_kind):
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    if pc_kind == "ilu":
        opc, pc = orc.OPreconditioner(om, 3, block_nodes=512), owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=512)
    else:
        opc, pc = orc.OPreconditioner(om, 2), owtk.Preconditioner(m, owtk.PC_SSOR)
    xo, ro = orc.bicgstab(om, opc, s.rhs)
    x, res = solve(ctx, owtk.bicgstab, m, pc, s.rhs)
    assert res["status_name"] == "converged"
    assert abs(res["iterations"] - ro["iterations"]) <= 1
    assert np.linalg.norm(s.rhs - om.spmv(x)) / np.linalg.norm(s.rhs) <= 1e-8


COUNTERS = ("operator_applications", "preconditioner_applications", "global_reductions")


@pytest.mark.parametrize("use_ilu", [False, True])
def test_pipelined_bicgstab_recurrence(ctx, use_ilu):
    """Device pipelined BiCGSTAB on the 12 x 12 audit matrix against the
    oracle (pinned by a long-double dense restatement): the first four
    iterates agree and every counter matches (pipelined_bicgstab.hpp:86-244)."""
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    m = owtk.Matrix.csr(ctx, 12, 0, 1, offsets, columns, values)
    om = orc.OMatrix.csr(12, 0, 1, offsets, columns, values)
    pc = owtk.Preconditioner(m, owtk.PC_ILU) if use_ilu else None
    opc = orc.OPreconditioner(om, 3) if use_ilu else None
    for count in range(1, 5):
        kw = dict(relative_tolerance=0.0, maximum_iterations=count, residual_replacement_interval=0)
        x, res = solve(ctx, owtk.pipelined_bicgstab, m, pc, rhs, x0, **kw)
        xo, ro = orc.pipelined_bicgstab(om, opc, rhs, x0=x0, **kw)
        assert res["status_name"] == "maximum_iterations" and res["iterations"] == count
        assert np.max(np.abs(x[:12] - xo[:12])) < 2e-11, count
        for key in COUNTERS:
            assert res[key] == ro[key], (count, key)
        # initial fused norms + rho, 3 per iteration, final true residual norm
        assert res["global_reductions"] == 2 + 3 * count + 1


def test_pipelined_bicgstab_exact_preconditioner(ctx):
    """Exact ILU on a tridiagonal 3 x 3: the intermediate residual vanishes in
    the first iteration and the early exit converges on the true residual."""
    vals = [4.0, -1, -2, 5, -1, -1, 3]
    m = owtk.Matrix.csr(ctx, 3, 0, 1, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], vals)
    om = orc.OMatrix.csr(3, 0, 1, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], vals)
    rhs = om.spmv([1.0, 2.0, 3.0])
    x, res = solve(ctx, owtk.pipelined_bicgstab, m, owtk.Preconditioner(m, owtk.PC_ILU), rhs,
                   relative_tolerance=32 * EPS)
    xo, ro = orc.pipelined_bicgstab(om, orc.OPreconditioner(om, 3), rhs, relative_tolerance=32 * EPS)
    assert res["status_name"] == "converged" and res["iterations"] == 1 == ro["iterations"]
    for key in COUNTERS:
        assert res[key] == ro[key], key
    assert np.max(np.abs(x[:3] - [1, 2, 3])) < 128 * EPS


def test_pipelined_bicgstab_breakdowns(ctx):
    m = owtk.Matrix.csr(ctx, 1, 0, 1, [0, 1], [0], [0.0])
    _, res = solve(ctx, owtk.pipelined_bicgstab, m, None, np.array([1.0]))
    assert res["status_name"] == "breakdown" and res["reason_name"] == "alpha_denominator"
    om = orc.OMatrix.csr(1, 0, 1, [0, 1], [0], [0.0])
    _, ro = orc.pipelined_bicgstab(om, None, np.array([1.0]))
    for key in COUNTERS:
        assert res[key] == ro[key], key
    m = owtk.Matrix.csr(ctx, 1, 0, 2, [0, 1], [0], [1.0, 1.0])
    _, res = solve(ctx, owtk.pipelined_bicgstab, m, None, np.array([np.inf, 1.0]))
    assert res["status_name"] == "breakdown" and res["reason_name"] == "non_finite_scalar"


@pytest.mark.parametrize("pc_kind", ["jacobi", "ssor", "ilu"])
@pytest.mark.parametrize("replace", [50, 7])
def test_pipelined_bicgstab_c0_matches_oracle(ctx, c0, pc_kind, replace):
    """c0 (the reference's CPU case) with each preconditioner, including a
    short replacement interval that restarts the recurrence several times."""
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    kind = {"jacobi": (1, owtk.PC_JACOBI), "ssor": (2, owtk.PC_SSOR), "ilu": (3, owtk.PC_ILU)}[pc_kind]
    extra = {"block_nodes": 512} if pc_kind == "ilu" else {}
    opc, pc = orc.OPreconditioner(om, kind[0], **extra), owtk.Preconditioner(m, kind[1], **extra)
    xo, ro = orc.pipelined_bicgstab(om, opc, s.rhs, residual_replacement_interval=replace)
    x, res = solve(ctx, owtk.pipelined_bicgstab, m, pc, s.rhs, residual_replacement_interval=replace)
    print(pc_kind, replace, "oracle", ro["iterations"], ro["relative_residual_norm"], "gpu", res["iterations"],
          res["relative_residual_norm"], flush=True)
    assert res["status_name"] == "converged" and ro["status"] == 0
    assert abs(res["iterations"] - ro["iterations"]) <= 1
    true_rel = np.linalg.norm(s.rhs - om.spmv(x)) / np.linalg.norm(s.rhs)
    assert true_rel <= 1e-8
    assert abs(res["relative_residual_norm"] - true_rel) <= 1e-3 * true_rel
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6
    if res["iterations"] == ro["iterations"]:
        for key in COUNTERS:
            assert res[key] == ro[key], key


@pytest.mark.parametrize("use_ilu", [False, True])
def test_communication_hiding_recurrence(ctx, use_ilu):
    """Device communication-hiding BiCGSTAB on the 12 x 12 audit matrix:
    the first four iterates agree with the oracle (itself pinned by the
    long-double dense BiCGSTAB), every counter matches."""
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    m = owtk.Matrix.csr(ctx, 12, 0, 1, offsets, columns, values)
    om = orc.OMatrix.csr(12, 0, 1, offsets, columns, values)
    pc = owtk.Preconditioner(m, owtk.PC_ILU) if use_ilu else None
    opc = orc.OPreconditioner(om, 3) if use_ilu else None
    for count in range(1, 5):
        kw = dict(relative_tolerance=0.0, maximum_iterations=count, residual_replacement_interval=0)
        x, res = solve(ctx, owtk.communication_hiding_bicgstab, m, pc, rhs, x0, **kw)
        xo, ro = orc.communication_hiding_bicgstab(om, opc, rhs, x0=x0, **kw)
        assert res["status_name"] == "maximum_iterations" and res["iterations"] == count
        assert np.max(np.abs(x[:12] - xo[:12])) < 2e-11, count
        for key in COUNTERS:
            assert res[key] == ro[key], (count, key)


def test_communication_hiding_exact_preconditioner_and_breakdowns(ctx):
    vals = [4.0, -1, -2, 5, -1, -1, 3]
    m = owtk.Matrix.csr(ctx, 3, 0, 1, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], vals)
    om = orc.OMatrix.csr(3, 0, 1, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], vals)
    rhs = om.spmv([1.0, 2.0, 3.0])
    x, res = solve(ctx, owtk.communication_hiding_bicgstab, m, owtk.Preconditioner(m, owtk.PC_ILU), rhs,
                   relative_tolerance=32 * EPS)
    xo, ro = orc.communication_hiding_bicgstab(om, orc.OPreconditioner(om, 3), rhs, relative_tolerance=32 * EPS)
    assert res["status_name"] == "converged" and res["iterations"] == ro["iterations"]
    for key in COUNTERS:
        assert res[key] == ro[key], key
    assert np.max(np.abs(x[:3] - [1, 2, 3])) < 128 * EPS
    # zero operator: the first restart's <A M r, rp> vanishes
    m0 = owtk.Matrix.csr(ctx, 1, 0, 1, [0, 1], [0], [0.0])
    _, res = solve(ctx, owtk.communication_hiding_bicgstab, m0, None, np.array([1.0]))
    _, ro = orc.communication_hiding_bicgstab(orc.OMatrix.csr(1, 0, 1, [0, 1], [0], [0.0]), None, np.array([1.0]))
    assert res["status_name"] == "breakdown" and res["reason_name"] == "biorthogonality_loss"
    for key in COUNTERS:
        assert res[key] == ro[key], key
    m1 = owtk.Matrix.csr(ctx, 1, 0, 2, [0, 1], [0], [1.0, 1.0])
    _, res = solve(ctx, owtk.communication_hiding_bicgstab, m1, None, np.array([np.inf, 1.0]))
    assert res["status_name"] == "breakdown" and res["reason_name"] == "non_finite_scalar"


@pytest.mark.parametrize("pc_kind", ["jacobi", "ssor", "ilu"])
@pytest.mark.parametrize("replace", [50, 7])
def test_communication_hiding_c0_matches_oracle(ctx, c0, pc_kind, replace):
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    kind = {"jacobi": (1, owtk.PC_JACOBI), "ssor": (2, owtk.PC_SSOR), "ilu": (3, owtk.PC_ILU)}[pc_kind]
    extra = {"block_nodes": 512} if pc_kind == "ilu" else {}
    opc, pc = orc.OPreconditioner(om, kind[0], **extra), owtk.Preconditioner(m, kind[1], **extra)
    xo, ro = orc.communication_hiding_bicgstab(om, opc, s.rhs, residual_replacement_interval=replace)
    x, res = solve(ctx, owtk.communication_hiding_bicgstab, m, pc, s.rhs, residual_replacement_interval=replace)
    print(pc_kind, replace, "oracle", ro["iterations"], ro["relative_residual_norm"], "gpu", res["iterations"],
          res["relative_residual_norm"], flush=True)
    assert res["status_name"] == "converged" and ro["status"] == 0
    assert abs(res["iterations"] - ro["iterations"]) <= 1
    true_rel = np.linalg.norm(s.rhs - om.spmv(x)) / np.linalg.norm(s.rhs)
    assert true_rel <= 1e-8
    assert abs(res["relative_residual_norm"] - true_rel) <= 1e-3 * true_rel
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6
    if res["iterations"] == ro["iterations"]:
        for key in COUNTERS:
            assert res[key] == ro[key], key


@pytest.mark.parametrize("sh", [2, 4])
def test_idrs_recurrence_vs_oracle(ctx, sh):
    """Device IDR(s) (idrs.hpp:128-505) on the audit's 12 x 12 matrix: the
    iterates of the first steps agree with the oracle (same shadow draws),
    counters match, and the solve terminates within n + n/s steps (the IDR
    theorem) at the dense solution."""
    import math
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    m = owtk.Matrix.csr(ctx, 12, 0, 1, offsets, columns, values)
    om = orc.OMatrix.csr(12, 0, 1, offsets, columns, values)
    for count in range(1, 7):
        kw = dict(relative_tolerance=0.0, maximum_iterations=count, residual_replacement_interval=0,
                  idr_shadow_space=sh)
        x, res = solve(ctx, owtk.idrs, m, None, rhs, x0, **kw)
        xo, ro = orc.idrs(om, None, rhs, x0=x0, **kw)
        assert res["status_name"] == "maximum_iterations" and res["iterations"] == count == ro["iterations"]
        assert np.max(np.abs(x[:12] - xo[:12])) < 1e-10 * np.max(np.abs(xo[:12])), count
        for key in COUNTERS:
            assert res[key] == ro[key], (count, key)
    exact = np.linalg.solve(a.astype(np.float64), rhs)
    x, res = solve(ctx, owtk.idrs, m, None, rhs, x0, relative_tolerance=1e-13, idr_shadow_space=sh,
                   residual_replacement_interval=0)
    assert res["status_name"] == "converged" and res["iterations"] <= 12 + math.ceil(12 / sh)
    assert np.max(np.abs(x[:12] - exact)) <= 1e-12 * np.max(np.abs(exact))


def test_idrs_edge_cases(ctx, c0):
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    m = owtk.Matrix.csr(ctx, 12, 0, 1, offsets, columns, values)
    # IDR(1) is the BiCGSTAB recurrence
    x1, r1 = solve(ctx, owtk.idrs, m, None, rhs, x0, idr_shadow_space=1, relative_tolerance=1e-12)
    xb, rb = solve(ctx, owtk.bicgstab, m, None, rhs, x0, relative_tolerance=1e-12)
    assert np.array_equal(x1, xb) and r1["iterations"] == rb["iterations"]
    # s beyond the system dimension is invalid input; beyond 16 unsupported on the device
    _, r = solve(ctx, owtk.idrs, m, None, rhs, x0, idr_shadow_space=13)
    assert r["status_name"] == "invalid_input"
    s = c0
    ms = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    with pytest.raises(owtk.UnsupportedError):
        solve(ctx, owtk.idrs, ms, None, s.rhs, idr_shadow_space=17)
    # the reference's scaled right-hand side (test_core.cpp:784-798) and non-finite data
    pc = owtk.Preconditioner(ms, owtk.PC_JACOBI)
    _, r = solve(ctx, owtk.idrs, ms, pc, s.rhs * 1e-18, idr_shadow_space=2)
    assert r["status_name"] == "converged"
    one = owtk.Matrix.csr(ctx, 2, 0, 1, [0, 1, 2], [0, 1], [1.0, 1.0])
    for bad in (np.nan, np.inf):
        _, r = solve(ctx, owtk.idrs, one, None, np.array([bad, 1.0]), idr_shadow_space=2, maximum_iterations=3)
        assert r["status_name"] != "converged" and not np.isfinite(r["initial_residual_norm"])


@pytest.mark.parametrize("pc_kind", ["jacobi", "ssor", "ilu"])
@pytest.mark.parametrize("sh", [2, 4, 8])
def test_idrs_c0_matches_oracle(ctx, c0, pc_kind, sh):
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    kind = {"jacobi": (1, owtk.PC_JACOBI), "ssor": (2, owtk.PC_SSOR), "ilu": (3, owtk.PC_ILU)}[pc_kind]
    extra = {"block_nodes": 512} if pc_kind == "ilu" else {}
    opc, pc = orc.OPreconditioner(om, kind[0], **extra), owtk.Preconditioner(m, kind[1], **extra)
    xo, ro = orc.idrs(om, opc, s.rhs, idr_shadow_space=sh)
    x, res = solve(ctx, owtk.idrs, m, pc, s.rhs, idr_shadow_space=sh)
    print(pc_kind, sh, "oracle", ro["iterations"], ro["relative_residual_norm"], "gpu", res["iterations"],
          res["relative_residual_norm"], flush=True)
    assert res["status_name"] == "converged" and ro["status"] == 0
    # the same recurrence and shadow draws; reductions round differently
    assert abs(res["iterations"] - ro["iterations"]) <= max(2, ro["iterations"] // 20)
    true_rel = np.linalg.norm(s.rhs - om.spmv(x)) / np.linalg.norm(s.rhs)
    assert true_rel <= 1e-8
    assert abs(res["relative_residual_norm"] - true_rel) <= 1e-3 * true_rel
    if res["iterations"] == ro["iterations"]:
        for key in COUNTERS:
            assert res[key] == ro[key], key


def test_fgmres_direct_solve(ctx):
    a, offsets, columns, values, rhs, _ = rc.recurrence_matrix()
    m = owtk.Matrix.csr(ctx, 12, 0, 1, offsets, columns, values)
    exact = np.linalg.solve(a.astype(float), rhs)
    om = orc.OMatrix.csr(12, 0, 1, offsets, columns, values)
    for mode in (0, 1, 2):
        x, res = solve(ctx, owtk.fgmres, m, None, rhs, relative_tolerance=1e-13, restart=12,
                       gmres_orthogonalization=mode)
        xo, ro = orc.fgmres(om, None, rhs, relative_tolerance=1e-13, restart=12, gmres_orthogonalization=mode)
        print(f"mode {mode}: gpu it={res['iterations']} rel={res['relative_residual_norm']:.3e} "
              f"reorth={res['reorthogonalizations']} verif={res['arnoldi_norm_verifications']} | oracle "
              f"it={ro['iterations']} rel={ro['relative_residual_norm']:.3e}", flush=True)
        assert res["status_name"] == "converged"
        # one Krylov space of dimension 12 spans the solution; the one-pass
        # variants may need a short second cycle under different rounding
        assert res["iterations"] <= (12 if mode == 0 else 24)
        assert np.max(np.abs(x - exact)) < 1e-11


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("pc_kind", ["jacobi", "ilu"])
def test_fgmres_c0_matches_oracle(ctx, c0, pc_kind, mode):
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    if pc_kind == "ilu":
        opc, pc = orc.OPreconditioner(om, 3, block_nodes=512), owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=512)
    else:
        opc, pc = orc.OPreconditioner(om, 1), owtk.Preconditioner(m, owtk.PC_JACOBI)
    xo, ro = orc.fgmres(om, opc, s.rhs, gmres_orthogonalization=mode)
    x, res = solve(ctx, owtk.fgmres, m, pc, s.rhs, gmres_orthogonalization=mode)
    print("mode", mode, "oracle", ro["iterations"], "gpu", res["iterations"])
    assert res["status_name"] == "converged" and ro["status"] == 0
    assert abs(res["iterations"] - ro["iterations"]) <= 1
    assert np.linalg.norm(s.rhs - om.spmv(x)) / np.linalg.norm(s.rhs) <= 1e-8
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-7
    if res["iterations"] == ro["iterations"]:
        for key in ("operator_applications", "preconditioner_applications", "global_reductions",
                    "reorthogonalizations"):
            assert res[key] == ro[key], key


@pytest.mark.parametrize("storage", [owtk.F32, owtk.BF16])
def test_refined_fgmres_reaches_fp64_accuracy(ctx, c0, storage):
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    xo, _ = orc.bicgstab(om, orc.OPreconditioner(om, 3), s.rhs, relative_tolerance=1e-13)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m.add_storage(storage)
    pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=512, storage=storage)
    bd, xd = ctx.upload(s.rhs), ctx.upload(np.zeros(m.local_size))
    res = owtk.fgmres_refined(m, pc, bd, xd, owtk.refinement_options(inner_storage=storage))
    x = xd.download()
    true_rel = np.linalg.norm(s.rhs - om.spmv(x)) / np.linalg.norm(s.rhs)
    print("refined", res["refinement_steps"], res["inner_iterations"], true_rel)
    assert res["status_name"] == "converged" and true_rel <= 1e-8
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6


def test_dense_bsr_bf16_refined_fgmres_with_ssor(ctx):
    """Dense 4x4 BSR system (the c2 generator on a small grid): BF16 matrix and
    multicolour SSOR storage inside FP64-refined FGMRES(40) to rtol 1e-10."""
    n, off, col, val, rhs = owtk.generate_bsr(96, 96, 4, 0, ordering=owtk.ORDER_HILBERT)
    m = owtk.Matrix.bsr(ctx, n, 0, 4, off, col, val)
    m.add_storage(owtk.BF16)
    pc = owtk.Preconditioner(m, owtk.PC_SSOR, storage=owtk.BF16)
    b = ctx.upload(rhs)
    x = ctx.empty(rhs.size)
    res = owtk.fgmres_refined(m, pc, b, x, owtk.refinement_options(relative_tolerance=1e-10, inner_storage=owtk.BF16,
                                                                  inner_restart=40))
    om = orc.OMatrix.bsr(n, 0, 4, off, col, val)
    rel = np.linalg.norm(rhs - om.spmv(x.download())) / np.linalg.norm(rhs)
    print(res["status_name"], res["inner_iterations"], res["refinement_steps"], rel)
    assert res["status_name"] == "converged"
    assert rel <= 1e-10


def test_cold_front_golden_fgmres(ctx):
    """The reference tests' real-data cold-front matrix (golden fixture): GPU
    FGMRES reaches the test's absolute tolerance on the true residual with
    the oracle's iteration count, in FP64 and through FP32 refinement."""
    from golden.fixtures import cold_front
    n, off, col, val, rhs, restart, maxit, atol = cold_front()
    om = orc.OMatrix.csr(n, 0, 1, off, col, val)
    _, ro = orc.fgmres(om, None, rhs, relative_tolerance=0.0, absolute_tolerance=atol, restart=restart,
                       maximum_iterations=maxit)
    m = owtk.Matrix.csr(ctx, n, 0, 1, off, col, val)
    b = ctx.upload(rhs)
    x = ctx.empty(n)
    res = owtk.fgmres(m, None, b, x, owtk.SolverOptions(relative_tolerance=0.0, absolute_tolerance=atol,
                                                        restart=restart, maximum_iterations=maxit))
    true = np.linalg.norm(rhs - om.spmv(x.download()))
    print("gpu", res["iterations"], true, "oracle", ro["iterations"], ro["true_residual_norm"])
    assert res["status_name"] == "converged" and true <= atol
    assert abs(res["iterations"] - ro["iterations"]) <= 1
    assert abs(res["true_residual_norm"] - true) <= 1e-12 + 1e-6 * true
    x.zero()
    rr = owtk.fgmres_refined(m, None, b, x, owtk.refinement_options(relative_tolerance=0.0, absolute_tolerance=atol,
                                                                   inner_storage=owtk.F32, inner_restart=restart))
    true = np.linalg.norm(rhs - om.spmv(x.download()))
    print("refined", rr["status_name"], rr["refinement_steps"], rr["inner_iterations"], true)
    assert rr["status_name"] == "converged" and true <= atol


def _fgmres_on(ctx, s, refined):
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    bd, xd = ctx.upload(s.rhs), ctx.empty(m.local_size)
    xd.zero()
    if refined:
        m.add_storage(owtk.F32)
        pc = owtk.Preconditioner(m, owtk.PC_JACOBI, storage=owtk.F32)
        res = owtk.fgmres_refined(m, pc, bd, xd, owtk.refinement_options(relative_tolerance=1e-9))
    else:
        pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
        res = owtk.fgmres(m, pc, bd, xd, owtk.SolverOptions(relative_tolerance=1e-9, restart=20))
    return xd.download(), res["iterations"]


@pytest.mark.parametrize("refined", [False, True])
def test_fgmres_workspace_reuse_across_sizes(ctx, refined):
    """The context keeps FGMRES and refinement workspaces between solves and
    re-zeroes them when the vector layout changes: alternating systems of
    different sizes on one context give bit-for-bit the results of fresh
    contexts."""
    small = owtk.generate_split(40, 24, 3, 3, ordering=owtk.ORDER_HILBERT)
    large = owtk.generate_split(64, 40, 3, 4, ordering=owtk.ORDER_HILBERT)
    fresh = {}
    for name, s in (("small", small), ("large", large)):
        with owtk.Context(0) as c:
            fresh[name] = _fgmres_on(c, s, refined)
    for name, s in (("large", large), ("small", small), ("large", large), ("small", small)):
        x, its = _fgmres_on(ctx, s, refined)
        print(name, "iterations", its, "fresh", fresh[name][1], flush=True)
        assert its == fresh[name][1]
        assert np.array_equal(x, fresh[name][0])
