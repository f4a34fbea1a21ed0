"""GCRO-DR on the device (gcrodr.hpp:86-212): harmonic-Ritz extraction from
an Arnoldi snapshot held to the reference's own contract cases
(review_harmonic.cpp:11-138) and to the oracle's extraction (LAPACK ggev via
scipy), and gcrodr held to test_core.cpp:611-634 and to the oracle's gcrodr
solve by solve (iterations, counters, recycle-space spans)."""
import numpy as np
import pytest

import owtk
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
COUNTERS = ("operator_applications", "preconditioner_applications", "global_reductions")


def csr(ctx, owned, block, offsets, columns, values):
    return (owtk.Matrix.csr(ctx, owned, 0, block, offsets, columns, values),
            orc.OMatrix.csr(owned, 0, block, offsets, columns, values))


def manual_snapshot(ctx, m, columns, basis):
    """begin_cycle / record_column / finish_cycle (core.hpp:465-502)."""
    snap = owtk.ArnoldiSnapshot(ctx)
    snap.begin_cycle(m, len(columns))
    for j, col in enumerate(columns):
        snap.record_column(j, col)
    bufs = [ctx.upload(np.asarray(z, np.float64)) for z in basis]
    snap.finish_cycle(bufs, len(basis))
    return snap, bufs


def space_arrays(ctx, space):
    pairs = [space.get(ctx, i) for i in range(space.size())]
    return np.array([u for u, _ in pairs]), np.array([c for _, c in pairs])


def max_angle(a, b):
    """Largest principal angle between the row spans of a and b (equal
    dimensions), from its sine, which keeps small angles accurate."""
    qa, _ = np.linalg.qr(np.asarray(a).T)
    qb, _ = np.linalg.qr(np.asarray(b).T)
    s = np.linalg.norm(qb - qa @ (qa.T @ qb), 2)
    return float(np.arcsin(min(s, 1.0)))


@pytest.mark.parametrize("magnitude", [1e-20, 1.0, 1e20])
def test_contract_scaled_diagonal_and_pair(ctx, magnitude):
    """review_harmonic.cpp:23-47 for capacities 1, 2, 3: sizes 1, 1, 3,
    A U = C to 512 eps, C orthonormal, telemetry and the kept span equal
    to the oracle's."""
    mg = magnitude
    m, om = csr(ctx, 3, 1, [0, 1, 3, 5], [0, 1, 2, 1, 2], [0.1 * mg, mg, -mg, mg, mg])
    cols = [[0.1 * mg, 0.0], [0.0, mg, mg], [0.0, -mg, mg, 0.0]]
    snap, _ = manual_snapshot(ctx, m, cols, np.eye(3))
    assert snap.columns() == 3
    for j, col in enumerate(cols):
        for i, v in enumerate(col):
            assert snap.h(i, j) == v
    tol = 512 * EPS
    for capacity in (1, 2, 3):
        space = owtk.RecycleSpace(m, capacity)
        size, tel = owtk.extract_harmonic_ritz(m, snap, space)
        ospace, _, otel = orc.extract_harmonic_ritz(om, np.eye(3), capacity)
        print(f"magnitude {mg:g} capacity {capacity}: kept {size} (oracle {ospace.size()}), telemetry "
              f"{tel['operator_applications']}/{tel['global_reductions']}", flush=True)
        assert size == space.size() == ospace.size() == (3 if capacity == 3 else 1)
        for key in otel:
            assert tel[key] == otel[key], key
        U, Cm = space_arrays(ctx, space)
        for u, c in zip(U, Cm):
            assert np.linalg.norm(om.spmv(u) - c) <= tol
        assert np.max(np.abs(Cm @ Cm.T - np.eye(size))) <= tol
        OU, _ = zip(*(ospace.get(i, 3) for i in range(ospace.size())))
        assert max_angle(U, np.array(OU)) <= 1e-12


def test_contract_flexible_nonorthogonal_basis(ctx):
    """review_harmonic.cpp:48-70: Z = [(2,0,0), (0.25,1,0)]; the kept u
    satisfies (AZ)^T (A u - theta u) = 0."""
    m, om = csr(ctx, 3, 1, [0, 1, 3, 5], [0, 0, 1, 1, 2], [1.0, 0.5, 3.0, 1.0, 4.0])
    Z = np.array([[2.0, 0.0, 0.0], [0.25, 1.0, 0.0]])
    snap, _ = manual_snapshot(ctx, m, [[2.0, 1.0], [0.25, 3.125, 1.0]], Z)
    space = owtk.RecycleSpace(m, 1)
    size, _ = owtk.extract_harmonic_ritz(m, snap, space)
    assert size == 1
    (u,), _ = space_arrays(ctx, space)
    AZ = np.array([om.spmv(z) for z in Z])
    g, b = AZ @ om.spmv(u), AZ @ u
    theta = (b @ g) / (b @ b)
    assert np.linalg.norm(g - theta * b) <= 512 * EPS * np.linalg.norm(g)
    ospace, _, _ = orc.extract_harmonic_ritz(om, Z, 1)
    assert max_angle([u], [ospace.get(0, 3)[0]]) <= 1e-12


def test_contract_infinite_value_empties_the_space(ctx):
    """review_harmonic.cpp:71-89: (AZ)^T Z = 0 gives only an infinite
    harmonic value; the replacement leaves the space empty, whatever it held."""
    m, _ = csr(ctx, 2, 1, [0, 1, 2], [1, 0], [-1.0, 1.0])
    snap, _ = manual_snapshot(ctx, m, [[0.0, 1.0]], [[1.0, 0.0]])
    space = owtk.RecycleSpace(m, 1)
    added, _ = space.add_candidate(ctx.upload(np.array([1.0, 2.0])))
    assert added and space.size() == 1
    size, _ = owtk.extract_harmonic_ritz(m, snap, space)
    assert size == 0 and space.size() == 0


@pytest.mark.parametrize("pair", [False, True])
def test_review_main_cases(ctx, pair):
    """review_harmonic.cpp:102-138: capacity 1 keeps nothing of a complex
    pair (both directions must fit) and a harmonic Ritz vector otherwise;
    capacity 2 keeps the whole pair."""
    if pair:
        m, om = csr(ctx, 2, 1, [0, 2, 4], [0, 1, 0, 1], [1.0, -1.0, 1.0, 1.0])
        Z = np.eye(2)
        cols = [[1.0, 1.0], [-1.0, 1.0, 0.0]]
    else:
        m, om = csr(ctx, 3, 1, [0, 1, 3, 5], [0, 0, 1, 1, 2], [1.0, 0.5, 3.0, 1.0, 4.0])
        Z = np.array([[2.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
        cols = [[2.0, 1.0], [0.0, 3.0, 1.0]]
    snap, _ = manual_snapshot(ctx, m, cols, Z)
    space = owtk.RecycleSpace(m, 1)
    size, _ = owtk.extract_harmonic_ritz(m, snap, space)
    if pair:
        assert size == 0
        wide = owtk.RecycleSpace(m, 2)
        assert owtk.extract_harmonic_ritz(m, snap, wide)[0] == 2
        return
    assert size == 1
    (u,), _ = space_arrays(ctx, space)
    AZ = np.array([om.spmv(z) for z in Z])
    g, b = AZ @ om.spmv(u), AZ @ u
    theta = (b @ g) / (b @ b)
    assert np.linalg.norm(g - theta * b) <= 1e-12 * np.linalg.norm(g)


def test_dependent_basis_columns(ctx):
    """A snapshot whose Z repeats a column: (AZ)^T AZ is singular; the
    indeterminate value is skipped and the two determinate harmonic Ritz
    vectors (those of the pencil over the independent columns) are kept:
    U spans them, (AZ)^T (A U - U T) = 0 for a 2 x 2 T with the pencil's
    values (13 +- sqrt 46) / 6, and A U = C with C orthonormal."""
    m, om = csr(ctx, 3, 1, [0, 1, 3, 5], [0, 0, 1, 1, 2], [1.0, 0.5, 3.0, 1.0, 4.0])
    Z = np.array([[1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    snap, _ = manual_snapshot(ctx, m, [[1.0, 0.5], [0.0, 1.0, 0.5], [0.0, 0.0, 3.0, 1.0]], Z)
    space = owtk.RecycleSpace(m, 3)
    size, _ = owtk.extract_harmonic_ritz(m, snap, space)
    assert size == 2
    U, Cm = space_arrays(ctx, space)
    AZ = np.array([om.spmv(z) for z in Z])
    for u, c in zip(U, Cm):
        assert abs(u[2]) <= 1e-14 * np.linalg.norm(u)
        assert np.linalg.norm(om.spmv(u) - c) <= 1e-14
    assert np.max(np.abs(Cm @ Cm.T - np.eye(2))) <= 1e-14
    AU = np.array([om.spmv(u) for u in U])
    G, B = AZ @ AU.T, AZ @ U.T
    T = np.linalg.lstsq(B, G, rcond=None)[0]
    assert np.linalg.norm(G - B @ T) <= 1e-12 * np.linalg.norm(G)
    want = np.sort([(13 - np.sqrt(46)) / 6, (13 + np.sqrt(46)) / 6])
    assert np.allclose(np.sort(np.linalg.eigvals(T).real), want, rtol=1e-12)


def test_snapshot_recording_errors(ctx):
    """record_column / finish_cycle argument checks (core.hpp:481-496)."""
    m, _ = csr(ctx, 2, 1, [0, 1, 2], [0, 1], [1.0, 1.0])
    snap = owtk.ArnoldiSnapshot(ctx)
    snap.begin_cycle(m, 2)
    with pytest.raises(owtk.InvalidArgumentError):
        snap.record_column(0, [1.0, 2.0, 3.0])  # column 0 holds two values
    with pytest.raises(owtk.InvalidArgumentError):
        snap.record_column(2, [1.0, 2.0, 3.0, 4.0])  # beyond maximum_columns
    with pytest.raises(owtk.InvalidArgumentError):
        snap.finish_cycle([ctx.upload(np.ones(2))] * 3, 3)
    with pytest.raises(owtk.UnsupportedError):
        snap.begin_cycle(m, 61)


def test_gcrodr_reference_block_system(ctx):
    """test_core.cpp:611-634 on the device, solve for solve with the oracle:
    capacity 2, a first solve keeping two harmonic Ritz vectors with a
    recorded snapshot, then a related system converging keeping two."""
    offsets, columns = [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2]
    values = [4, 5, -1, -1, -1, -1, 4, 5, -1, -1, -1, -1, 3, 4]
    m, om = csr(ctx, 3, 2, offsets, columns, values)
    kw = dict(relative_tolerance=1e-12, convergence_check_interval=1, residual_replacement_interval=0,
              maximum_iterations=100, restart=8)
    pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
    opc = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
    exact = np.arange(1, 7, dtype=float)
    related = exact.copy()
    related[0] += 0.25
    related[5] -= 0.5
    space, ospace = owtk.RecycleSpace(m, 2), orc.ORecycleSpace(2)
    snap = owtk.ArnoldiSnapshot(ctx)
    for k, want in enumerate((exact, related)):
        b = om.spmv(want)
        x = ctx.upload(np.zeros(6))
        res = owtk.gcrodr(m, pc, ctx.upload(b), x, space, owtk.SolverOptions(**kw), snapshot=snap if k == 0 else None)
        xo, ro, ospace, (H, _) = orc.gcrodr(om, opc, b, ospace, 2, **kw)
        print(f"system {k}: {res['status_name']} after {res['iterations']} (oracle {ro['iterations']}), "
              f"space {space.size()}", flush=True)
        assert res["status_name"] == "converged" and ro["status"] == 0
        assert space.size() == ospace.size() == 2
        assert res["iterations"] == ro["iterations"]
        for key in COUNTERS:
            assert res[key] == ro[key], key
        assert np.max(np.abs(x.download() - want)) < 1e-9
        if k == 0:
            assert snap.columns() == H.shape[1] > 0
            assert np.max(np.abs(snap.hessenberg() - H)) <= 1e-12 * np.max(np.abs(H))
        U, _ = space_arrays(ctx, space)
        OU = np.array([ospace.get(i, 6)[0] for i in range(ospace.size())])
        assert max_angle(U, OU) <= 1e-8


@pytest.mark.parametrize("kind", ["jacobi", "ilu0"])
def test_gcrodr_sequence_matches_oracle(ctx, c0, kind):
    """Three related right-hand sides of c0 through one capacity-4 space:
    every solve converges with the oracle's iteration count (within one), the
    same space size and, where the counts agree, the same counters and a
    recycle space spanning the oracle's.  With block-Jacobi ILU(0) the
    preconditioned basis Z = M^-1 V is far from orthonormal, the case the
    harmonic pencil (AZ)^T AZ y = theta (AZ)^T Z y is built for."""
    s = c0
    n = s.nodes * 3
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    if kind == "jacobi":
        pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
        opc = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
    else:
        pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=512)
        opc = orc.OPreconditioner(om, orc.OPreconditioner.ILU, block_nodes=512)
    space, ospace = owtk.RecycleSpace(m, 4), orc.ORecycleSpace(4)
    rng = np.random.default_rng(11)
    kw = dict(restart=20)
    for k in range(3):
        b = s.rhs if k == 0 else s.rhs * (1 + 0.01 * rng.standard_normal(n))
        x = ctx.upload(np.zeros(m.local_size))
        res = owtk.gcrodr(m, pc, ctx.upload(b), x, space, owtk.SolverOptions(**kw))
        _, ro, ospace, _ = orc.gcrodr(om, opc, b, ospace, 4, **kw)
        print(f"rhs {k}: gpu {res['iterations']} iterations, oracle {ro['iterations']}, space {space.size()} "
              f"(oracle {ospace.size()})", flush=True)
        assert res["status_name"] == "converged" and ro["status"] == 0
        assert abs(res["iterations"] - ro["iterations"]) <= 1
        assert space.size() == ospace.size()
        assert np.linalg.norm(b - om.spmv(x.download())) <= 1e-8 * np.linalg.norm(b) * 1.01
        if res["iterations"] == ro["iterations"]:
            for key in COUNTERS:
                assert res[key] == ro[key], (k, key)
            U, Cm = space_arrays(ctx, space)
            OU = np.array([ospace.get(i, n)[0] for i in range(ospace.size())])
            assert max_angle(U, OU) <= 1e-6
            assert np.max(np.abs(Cm @ Cm.T - np.eye(len(Cm)))) < 1e-12
        if k == 0:
            plain = owtk.fgmres(m, pc, ctx.upload(b), ctx.upload(np.zeros(m.local_size)), owtk.SolverOptions(**kw))
            assert res["iterations"] == plain["iterations"]  # an empty space is plain FGMRES


def test_gcrodr_rejects_distributed_reductions(ctx, c0):
    s = c0
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    space = owtk.RecycleSpace(m, 2)
    opts = owtk.SolverOptions(reduction=owtk.REDUCTION_COMPENSATED)
    with pytest.raises(owtk.UnsupportedError):
        owtk.gcrodr(m, None, ctx.upload(s.rhs), ctx.upload(np.zeros(m.local_size)), space, opts)
