"""FGMRES's Arnoldi snapshot (ArnoldiSnapshot, core.hpp:462-534, filled by
fgmres: krylov_solvers.hpp:171, 252-253, 361-366, 420-437) and its
application convergence test (krylov_solvers.hpp:398-427) on the GPU.

The snapshot is compared with the oracle's on c0 (same cycles, same column
count), and both are held to the Arnoldi relation A Z = V H with V
orthonormal.  The application test is held to the reference's own contract
(test_core.cpp:934-955) and to the fact that the iterate it sees after k
columns is the solution fgmres returns with maximum_iterations = k."""
import ctypes as C

import numpy as np
import pytest

import owtk
from owtk import _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _system(ctx, c0):
    s = c0
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    return s, m, om


def _download(ctx, ptr, count):
    host = np.empty(count)
    _lib.check(_lib.lib().owtk_memcpy_dtoh(ctx.handle, host.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), count * 8))
    return host


@pytest.mark.parametrize("restart,limit", [(5, 12), (30, 30)])
def test_snapshot_matches_oracle_and_arnoldi_relation(ctx, c0, restart, limit):
    """Runs end on the iteration limit (tolerance out of reach), so both sides
    stop after the same cycles: the last one has limit - restart * k columns."""
    s, m, om = _system(ctx, c0)
    n = s.nodes * 3
    pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
    opc = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
    snap = owtk.ArnoldiSnapshot(ctx)
    b, x = ctx.upload(s.rhs), ctx.empty(n)
    x.zero()
    opts = owtk.SolverOptions(restart=restart, maximum_iterations=limit, relative_tolerance=1e-15)
    res = owtk.fgmres(m, pc, b, x, opts, snapshot=snap)
    _, ores, (H, Z) = orc.fgmres_snapshot(om, opc, s.rhs, restart=restart, maximum_iterations=limit,
                                          relative_tolerance=1e-15)
    k = snap.columns()
    expect = limit - restart * ((limit - 1) // restart)
    print(f"restart {restart}: status {res['status_name']}, {res['iterations']} iterations, snapshot columns {k} "
          f"(oracle {H.shape[1]})")
    assert res["status_name"] == "maximum_iterations" and ores["status"] == 1
    assert k == expect == H.shape[1]
    Hg = snap.hessenberg()
    scale = np.max(np.abs(H))
    assert np.max(np.abs(Hg - H)) <= 1e-10 * scale
    # subdiagonal: the norms h(j+1, j) > 0, everything below it zero
    assert np.all(np.diag(Hg, -1) > 0) and np.all(np.tril(Hg, -2) == 0)
    Zg = np.array([snap.preconditioned_basis(j) for j in range(k)])
    assert np.max(np.abs(Zg - Z)) <= 1e-9 * np.max(np.abs(Z))
    # A Z = V H with V orthonormal: (A Z)^T (A Z) = H^T H, and V = D Z for
    # Jacobi (z = D^-1 v per component), so (D Z)^T (D Z) = I on the first k
    AZ = np.array([om.spmv(z) for z in Zg])
    G = AZ @ AZ.T
    assert np.max(np.abs(G - Hg.T @ Hg)) <= 1e-10 * scale * scale
    V = Zg * s.diagonal.reshape(-1)[None, :]
    assert np.max(np.abs(V @ V.T - np.eye(k))) <= 1e-10
    # the address is the device vector itself (owned values, no ghosts here)
    assert np.array_equal(_download(ctx, snap.basis_pointer(k - 1), n), Zg[k - 1])
    with pytest.raises(owtk.InvalidArgumentError):
        snap.h(0, k)
    with pytest.raises(owtk.InvalidArgumentError):
        snap.preconditioned_basis(k)


def test_snapshot_identity_preconditioner_is_orthonormal_basis(ctx, c0):
    """Without a preconditioner Z is V: Z^T Z = I, Z^T A Z is the square part
    of H (this fails for a rotated H), and Z_0 = b / |b| from x0 = 0."""
    s, m, om = _system(ctx, c0)
    n = s.nodes * 3
    snap = owtk.ArnoldiSnapshot(ctx)
    b, x = ctx.upload(s.rhs), ctx.empty(n)
    x.zero()
    owtk.fgmres(m, None, b, x, owtk.SolverOptions(restart=20, maximum_iterations=20, relative_tolerance=1e-15),
                snapshot=snap)
    k = snap.columns()
    assert k == 20
    Z = np.array([snap.preconditioned_basis(j) for j in range(k)])
    H = snap.hessenberg()
    assert np.max(np.abs(Z @ Z.T - np.eye(k))) <= 1e-10
    AZ = np.array([om.spmv(z) for z in Z])
    assert np.max(np.abs(Z @ AZ.T - H[:k, :])) <= 1e-10 * np.max(np.abs(H))
    assert np.max(np.abs(Z[0] - s.rhs / np.linalg.norm(s.rhs))) <= 1e-15


def test_snapshot_cleared_when_no_cycle_runs(ctx, c0):
    """clear() on a zero restart and on an already converged start
    (krylov_solvers.hpp:175-179, 224-229)."""
    s, m, om = _system(ctx, c0)
    n = s.nodes * 3
    snap = owtk.ArnoldiSnapshot(ctx)
    b, x = ctx.upload(s.rhs), ctx.empty(n)
    x.zero()
    owtk.fgmres(m, None, b, x, owtk.SolverOptions(restart=4, maximum_iterations=4), snapshot=snap)
    assert snap.columns() == 4
    owtk.fgmres(m, None, b, x, owtk.SolverOptions(restart=0), snapshot=snap)
    assert snap.columns() == 0
    owtk.fgmres(m, None, b, x, owtk.SolverOptions(restart=4, maximum_iterations=4), snapshot=snap)
    assert snap.columns() == 4
    zero = ctx.empty(n)
    zero.zero()
    x.zero()
    res = owtk.fgmres(m, None, zero, x, owtk.SolverOptions(restart=4), snapshot=snap)
    assert res["status_name"] == "converged" and res["iterations"] == 0 and snap.columns() == 0
    snap.close()


def test_application_test_contract(ctx, c0):
    """test_core.cpp:934-955: a stop at iteration 1 keeps a one-column
    snapshot; a stop at iteration 0 returns at once and clears it."""
    s, m, om = _system(ctx, c0)
    n = s.nodes * 3
    snap = owtk.ArnoldiSnapshot(ctx)
    b, x, r = ctx.upload(s.rhs), ctx.empty(n), ctx.empty(n)
    x.zero()
    res = owtk.fgmres(m, None, b, x, owtk.SolverOptions(convergence_test=lambda i, p, c: i == 1), snapshot=snap)
    m.residual(b, x, r)
    rel = ctx.norm(r) / ctx.norm(b)
    print(res["status_name"], res["iterations"], res["converged_by_application"], snap.columns(), rel)
    assert res["converged_by_application"] == 1 and res["status_name"] == "converged"
    assert res["iterations"] == 1 and snap.columns() == 1
    assert res["relative_residual_norm"] == pytest.approx(rel, rel=1e-12)
    res = owtk.fgmres(m, None, b, x, owtk.SolverOptions(convergence_test=lambda i, p, c: True), snapshot=snap)
    assert res["converged_by_application"] == 1 and res["iterations"] == 0 and snap.columns() == 0


def test_application_iterates_are_truncated_solves(ctx, c0):
    """The iterate the test sees after k columns (across restarts) is the
    solution of fgmres with maximum_iterations = k, bit for bit; the test sees
    iterations 0..limit, and a test that never fires leaves the solve to end on
    its limit."""
    s, m, om = _system(ctx, c0)
    n = s.nodes * 3
    pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
    b, x = ctx.upload(s.rhs), ctx.empty(n)
    seen, iterates = [], {}

    def test(i, ptr, count):
        seen.append(i)
        iterates[i] = _download(ctx, ptr, count)
        return False

    x.zero()
    limit = 7
    res = owtk.fgmres(m, pc, b, x, owtk.SolverOptions(restart=3, maximum_iterations=limit, convergence_test=test))
    assert seen == list(range(limit + 1))
    assert res["status_name"] == "maximum_iterations" and res["converged_by_application"] == 0
    assert np.array_equal(iterates[limit], x.download())
    for k in range(1, limit):
        x.zero()
        owtk.fgmres(m, pc, b, x, owtk.SolverOptions(restart=3, maximum_iterations=k, relative_tolerance=1e-15))
        got = x.download()
        print(k, np.max(np.abs(got - iterates[k])))
        assert np.array_equal(got, iterates[k]), k
