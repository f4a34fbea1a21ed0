"""Device aggregation multigrid (multigrid.hpp:41-421) against the oracle."""
import numpy as np
import pytest

import owtk
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
COUNTERS = ("operator_applications", "global_reductions")
REF_OFF, REF_COL = [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2]
REF_VAL = [4.0, 5, -1, -1, -1, -1, 4, 5, -1, -1, -1, -1, 3, 4]


def run(ctx, mg, b, cycle, **kw):
    bd = ctx.upload(b)
    xd = ctx.empty(len(b))
    res = mg.solve(bd, xd, owtk.SolverOptions(**kw), cycle=cycle)
    return xd.download(), res


@pytest.mark.parametrize("aggregate", [2, 4])
def test_coarse_level_matches_oracle(ctx, c0, aggregate):
    s = c0
    co, cc, cv = s.csr()
    m = owtk.Matrix.csr(ctx, s.nodes, 0, 3, co, cc, cv)
    om = orc.OMatrix.csr(s.nodes, 0, 3, co, cc, cv)
    got = owtk.Multigrid(m, target_aggregate_size=aggregate).coarse()
    ref = orc.OMultigrid(om, aggregate=aggregate).coarse()
    assert got[0] == ref[0]
    for a, b in zip(got[1:], ref[1:]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("smoother", [0, 1, 2])
@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_reference_block_system(ctx, smoother, cycle):
    """test_core.cpp:800-844 on the device: every cycle and smoother
    converges; iterations, counters and iterates follow the oracle."""
    m = owtk.Matrix.csr(ctx, 3, 0, 2, REF_OFF, REF_COL, REF_VAL)
    om = orc.OMatrix.csr(3, 0, 2, REF_OFF, REF_COL, REF_VAL)
    exact = np.arange(1, 7, dtype=float)
    b = om.spmv(exact)
    mg = owtk.Multigrid(m, target_aggregate_size=2, coarse_smoothing_steps=40, smoother=smoother)
    x, res = run(ctx, mg, b, cycle, relative_tolerance=1e-9, maximum_iterations=100)
    xo, ro = orc.OMultigrid(om, aggregate=2, coarse_steps=40, smoother=smoother).solve(
        b, cycle=cycle, relative_tolerance=1e-9, maximum_iterations=100)
    assert res["status_name"] == "converged" and ro["status"] == 0
    assert res["iterations"] == ro["iterations"]
    for key in COUNTERS:
        assert res[key] == ro[key], key
    assert np.max(np.abs(x - xo)) <= 1e-13 * np.max(np.abs(xo))
    assert np.max(np.abs(x - exact)) < 1e-8


@pytest.mark.parametrize("smoother,cycle", [(0, 0), (0, 2), (1, 0), (1, 1), (2, 0)])
def test_c0_matches_oracle(ctx, c0, smoother, cycle):
    s = c0
    co, cc, cv = s.csr()
    m = owtk.Matrix.csr(ctx, s.nodes, 0, 3, co, cc, cv)
    om = orc.OMatrix.csr(s.nodes, 0, 3, co, cc, cv)
    mg = owtk.Multigrid(m, smoother=smoother)
    kw = dict(relative_tolerance=1e-8, maximum_iterations=400)
    x, res = run(ctx, mg, s.rhs, cycle, **kw)
    xo, ro = orc.OMultigrid(om, smoother=smoother).solve(s.rhs, cycle=cycle, **kw)
    print(smoother, cycle, "oracle", ro["status"], ro["iterations"], "gpu", res["status_name"], res["iterations"],
          flush=True)
    assert res["status"] == ro["status"]
    if ro["status"] == 0:
        assert abs(res["iterations"] - ro["iterations"]) <= 1
        assert np.linalg.norm(s.rhs - om.spmv(x)) <= 1e-8 * np.linalg.norm(s.rhs)
        if res["iterations"] == ro["iterations"]:
            for key in COUNTERS:
                assert res[key] == ro[key], key
    else:
        # Chebyshev with the reference's fixed spectral bound diverges here, on both
        assert res["reason_name"] == "non_finite_scalar" and ro["breakdown_reason"] == 9


def test_rejects_unsupported_input(ctx, c0):
    s = c0
    ms = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    with pytest.raises(owtk.InvalidArgumentError):
        owtk.Multigrid(ms)
    m = owtk.Matrix.csr(ctx, 3, 0, 2, REF_OFF, REF_COL, REF_VAL)
    with pytest.raises(owtk.InvalidArgumentError):
        owtk.Multigrid(m, target_aggregate_size=0)
