"""The reference tests' real-data fixture (tests/golden/cold_front.json, from
test_gmres_precision.cpp:19-55) on the CPU oracle: structure, a dense solve,
and FGMRES reaching the test's absolute tolerance on the true residual."""
import numpy as np

from golden.fixtures import cold_front
from oracle import oracle as orc


def test_cold_front_fixture_structure():
    n, off, col, val, rhs, restart, maxit, atol = cold_front()
    assert n == 25 and off[-1] == 57
    for i in range(n):
        assert i in col[off[i]:off[i + 1]], f"row {i} has no diagonal"
    assert rhs.sum() == 5.0 and restart == 25 and maxit == 200


def test_cold_front_oracle_fgmres_meets_true_tolerance():
    n, off, col, val, rhs, restart, maxit, atol = cold_front()
    dense = np.zeros((n, n))
    for i in range(n):
        for e in range(off[i], off[i + 1]):
            dense[i, col[e]] += val[e]
    exact = np.linalg.solve(dense, rhs)
    om = orc.OMatrix.csr(n, 0, 1, off, col, val)
    x, res = orc.fgmres(om, None, rhs, relative_tolerance=0.0, absolute_tolerance=atol, restart=restart,
                        maximum_iterations=maxit)
    true = np.linalg.norm(rhs - dense @ x)
    assert res["status"] == 0
    assert true <= atol
    # the reported true residual is the recomputed one (test_gmres_precision.cpp:48)
    assert abs(res["true_residual_norm"] - true) <= 1e-12 + 1e-6 * true
    assert np.linalg.norm(x - exact) <= 1e-3 * np.linalg.norm(exact)
