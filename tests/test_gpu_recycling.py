"""Device recycled FGMRES and recycle space (recycling.hpp:24-360) against the oracle."""
import numpy as np
import pytest

import owtk
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
COUNTERS = ("operator_applications", "preconditioner_applications", "global_reductions")


def systems(ctx, s):
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    return om, m


def solve(ctx, m, pc, b, space, update=True, x0=None, **kw):
    bd = ctx.upload(b)
    xd = ctx.upload(np.zeros(m.local_size) if x0 is None else x0)
    res = owtk.recycled_fgmres(m, pc, bd, xd, space, owtk.SolverOptions(**kw), update)
    return xd.download(), res


def test_sequence_matches_oracle(ctx, c0):
    """A sequence of related right-hand sides through one recycle space:
    the device follows the oracle solve by solve (iterations, counters), the
    first solve with an empty space is plain FGMRES, and re-solving a spanned
    right-hand side needs no Arnoldi step."""
    s = c0
    om, m = systems(ctx, s)
    n = s.nodes * 3
    opc, pc = orc.OPreconditioner(om, 1), owtk.Preconditioner(m, owtk.PC_JACOBI)
    ospace, space = orc.ORecycleSpace(3), owtk.RecycleSpace(m, 3)
    rng = np.random.default_rng(7)
    rhs = [s.rhs, s.rhs, s.rhs * (1 + 0.01 * rng.standard_normal(n)), s.rhs * (1 + 0.02 * rng.standard_normal(n)),
           rng.standard_normal(n)]
    plain_x, plain = solve(ctx, m, pc, s.rhs, owtk.RecycleSpace(m, 3), update=False)
    xf = ctx.upload(np.zeros(m.local_size))
    ref = owtk.fgmres(m, pc, ctx.upload(s.rhs), xf)
    assert np.array_equal(plain_x, xf.download())
    for key in COUNTERS + ("iterations",):
        assert plain[key] == ref[key], key
    for k, b in enumerate(rhs):
        x, res = solve(ctx, m, pc, b, space)
        xo, ro = orc.recycled_fgmres(om, opc, b, ospace)
        print(k, "oracle", ro["iterations"], "gpu", res["iterations"], "space", space.size(), ospace.size(),
              flush=True)
        assert res["status_name"] == "converged" and ro["status"] == 0
        assert abs(res["iterations"] - ro["iterations"]) <= 1
        if res["iterations"] == ro["iterations"]:
            for key in COUNTERS:
                assert res[key] == ro[key], (k, key)
        assert space.size() == ospace.size()
        assert np.linalg.norm(b - om.spmv(x)) <= 1e-8 * np.linalg.norm(b)
        if k == 1:
            assert res["iterations"] == 0
    # the space's invariants: C orthonormal, A U = C
    pairs = [space.get(ctx, i) for i in range(space.size())]
    G = np.array([[c1 @ c2 for _, c2 in pairs] for _, c1 in pairs])
    assert np.max(np.abs(G - np.eye(len(pairs)))) < 1e-13
    for u, c in pairs:
        assert np.max(np.abs(om.spmv(u) - c)) < 1e-12


def test_candidates_and_capacity(ctx, c0):
    s = c0
    om, m = systems(ctx, s)
    n = s.nodes * 3
    space = owtk.RecycleSpace(m, 2)
    rng = np.random.default_rng(9)
    first = rng.standard_normal(n)
    added, tel = space.add_candidate(ctx.upload(first))
    assert added and space.size() == 1
    assert tel["operator_applications"] == 1 and tel["global_reductions"] == 2
    added, _ = space.add_candidate(ctx.upload(2.5 * first))  # contained in the space
    assert not added and space.size() == 1
    for _ in range(2):
        added, tel = space.add_candidate(ctx.upload(rng.standard_normal(n)))
        assert added
    assert space.size() == 2
    assert tel["global_reductions"] == 4  # original norm, two projection passes, final norm
    # the first pair was dropped: the oldest image left was orthogonalised against it
    _, c0v = space.get(ctx, 0)
    af = om.spmv(first)
    assert abs(c0v @ af) / np.linalg.norm(af) < 1e-10
    with pytest.raises(owtk.InvalidArgumentError):
        owtk.RecycleSpace(m, 0)
