"""The device reductions selectable like the reference's Reduction argument
(reduction.hpp:103-299) and the application convergence test
(SolverOptions::convergence_test, core.hpp:100-108; krylov_solvers.hpp:
122-143) on the GPU.

The compensated (double-double) dot is checked against the exact rational
sum of the products on an ill-conditioned case, where FP64 tree sums lose
most digits; the solvers are checked by their true residuals and counters.
"""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

import owtk
from owtk import _lib

pytestmark = pytest.mark.gpu


def exact_dot(a, b):
    return float(sum((Fraction(float(x)) * Fraction(float(y)) for x, y in zip(a, b)), Fraction(0)))


def ill_conditioned_pair(n, seed):
    """Products of magnitude up to 1e9 that cancel to a sum of order 1e-3
    (condition number about 1e15)."""
    rng = np.random.default_rng(seed)
    big = rng.uniform(-1, 1, n // 2) * 10.0 ** rng.uniform(3, 9, n // 2)
    small = rng.uniform(-1e-3, 1e-3, n - 2 * (n // 2))
    a = np.concatenate([big, -big, small])
    b = np.concatenate([np.ones(n // 2), np.ones(n // 2), np.ones(small.size)])
    # scale b by exactly representable factors so the products carry rounding
    f = rng.uniform(0.5, 1.5, n)
    perm = rng.permutation(n)
    a, b = a[perm] * f[perm], b[perm] / f[perm]
    return a, b


def test_compensated_dot_ill_conditioned(ctx):
    n = 60_001
    a, b = ill_conditioned_pair(n, 3)
    exact = exact_dot(a, b)
    da, db = ctx.upload(a), ctx.upload(b)
    serial = ctx.dot(da, db, reduction=owtk.REDUCTION_SERIAL)
    comp = ctx.dot(da, db, reduction=owtk.REDUCTION_COMPENSATED)
    mixed = ctx.dot(da, db, reduction=owtk.REDUCTION_MIXED_PRECISION)
    scale = float(np.sum(np.abs(a * b)))
    print(f"exact {exact:.17e}; serial {serial:.17e} (error {abs(serial - exact):.2e}); compensated {comp:.17e} "
          f"(error {abs(comp - exact):.2e}); sum |a b| {scale:.3e}")
    # double-double: error of order u^2 sum |a b| plus the final rounding
    assert abs(comp - exact) <= 2 * np.finfo(float).eps * abs(exact) + 1e-28 * scale
    assert comp == mixed  # FP64 vectors: the same double-double sums
    # the plain FP64 tree cannot resolve the cancellation this well
    assert abs(serial - exact) > abs(comp - exact)
    # and the plain dot entry point is the serial reduction
    assert ctx.dot(da, db) == serial


def test_compensated_norm_and_extremes(ctx):
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, 200_003) * 10.0 ** rng.uniform(-8, 8, 200_003)
    exact = float(sum(Fraction(float(v)) ** 2 for v in x[:20_000]))
    d = ctx.upload(x[:20_000])
    comp = ctx.norm(d, reduction=owtk.REDUCTION_COMPENSATED)
    assert abs(comp - np.sqrt(exact)) <= 2 * np.finfo(float).eps * np.sqrt(exact)
    # the norm's overflow and underflow handling is shared with the serial path
    for v, want in (([3e300, 4e300], 5e300), ([3e-310, 4e-310], 5e-310), ([0.0, 0.0], 0.0)):
        dv = ctx.upload(np.array(v))
        got = ctx.norm(dv, reduction=owtk.REDUCTION_COMPENSATED)
        assert abs(got - want) <= 4 * np.finfo(float).eps * want, (v, got)
    assert np.isnan(ctx.norm(ctx.upload(np.array([1.0, np.nan])), reduction=owtk.REDUCTION_COMPENSATED))
    with pytest.raises(owtk.InvalidArgumentError):
        ctx.dot(d, d, reduction=7)


def _c0_solver(ctx, c0):
    s = c0
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=512)
    b = ctx.upload(s.rhs)
    return s, m, pc, b


@pytest.mark.parametrize("reduction", [owtk.REDUCTION_COMPENSATED, owtk.REDUCTION_MIXED_PRECISION])
def test_bicgstab_with_compensated_reduction(ctx, c0, reduction):
    """bicgstab with CompensatedSerialReduction / MixedPrecisionSerialReduction
    (bicgstab_stable, bicgstab_mixed_precision: krylov_solvers.hpp:766-809)."""
    s, m, pc, b = _c0_solver(ctx, c0)
    n = s.nodes * 3
    x, r = ctx.empty(n), ctx.empty(n)
    serial = owtk.bicgstab(m, pc, b, x, owtk.SolverOptions(relative_tolerance=1e-10))
    x.zero()
    comp = owtk.bicgstab(m, pc, b, x, owtk.SolverOptions(relative_tolerance=1e-10, reduction=reduction))
    m.residual(b, x, r)
    rel = ctx.norm(r) / ctx.norm(b)
    print(f"serial: {serial['iterations']} iterations, {serial['global_reductions']} reductions; "
          f"reduction {reduction}: {comp['iterations']} iterations, {comp['global_reductions']} reductions, "
          f"true rel {rel:.2e}")
    assert comp["status_name"] == "converged"
    assert rel <= 1e-10
    assert abs(comp["iterations"] - serial["iterations"]) <= 1
    if comp["iterations"] == serial["iterations"]:
        assert comp["global_reductions"] == serial["global_reductions"]
        assert comp["operator_applications"] == serial["operator_applications"]


def test_bicgstab_application_convergence_test(ctx, c0):
    """The test sees iterations 0, 1, 2, 3 and the current solution; it stops
    the solve at 3, which reports the true residual and
    converged_by_application."""
    s, m, pc, b = _c0_solver(ctx, c0)
    n = s.nodes * 3
    x, r = ctx.empty(n), ctx.empty(n)
    seen, norms = [], []

    def test(iteration, ptr, count):
        seen.append(iteration)
        host = np.empty(count)
        _lib.check(_lib.lib().owtk_memcpy_dtoh(ctx.handle, host.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                               count * 8))
        norms.append(float(np.linalg.norm(host)))
        return iteration >= 3

    res = owtk.bicgstab(m, pc, b, x, owtk.SolverOptions(convergence_test=test))
    m.residual(b, x, r)
    rel = ctx.norm(r) / ctx.norm(b)
    print(res["status_name"], res["iterations"], res["converged_by_application"], seen, rel)
    assert seen == [0, 1, 2, 3]
    assert norms[0] == 0.0 and norms[-1] == pytest.approx(float(np.linalg.norm(x.download())), rel=1e-15)
    assert res["status_name"] == "converged" and res["converged_by_application"] == 1
    assert res["iterations"] == 3
    assert res["relative_residual_norm"] == pytest.approx(rel, rel=1e-12)
    # a test that never fires replaces the tolerance: the solve runs to the limit
    seen.clear()
    x.zero()
    res = owtk.bicgstab(m, pc, b, x, owtk.SolverOptions(maximum_iterations=12,
                                                          convergence_test=lambda i, p, c: seen.append(i) and False))
    assert res["status_name"] == "maximum_iterations" and res["iterations"] == 12
    assert seen == list(range(13)) and res["converged_by_application"] == 0


def test_other_solvers_refuse_reductions_and_tests(ctx, c0):
    s, m, pc, b = _c0_solver(ctx, c0)
    x = ctx.empty(s.nodes * 3)
    for fn in (owtk.fgmres, owtk.pipelined_bicgstab, owtk.communication_hiding_bicgstab, owtk.idrs):
        with pytest.raises(owtk.UnsupportedError):
            fn(m, pc, b, x, owtk.SolverOptions(reduction=owtk.REDUCTION_COMPENSATED))
    # fgmres takes the application test too (tests/test_gpu_snapshot.py)
    for fn in (owtk.pipelined_bicgstab, owtk.communication_hiding_bicgstab, owtk.idrs):
        with pytest.raises(owtk.UnsupportedError):
            fn(m, pc, b, x, owtk.SolverOptions(convergence_test=lambda i, p, c: True))
