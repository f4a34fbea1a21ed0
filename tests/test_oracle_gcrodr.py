"""The oracle's GCRO-DR (gcrodr.hpp:86-212): harmonic-Ritz extraction pinned
by the reference's own contract cases (review_harmonic.cpp:11-100, 104-138)
and the defining property of a harmonic Ritz vector, and gcrodr on c0."""
import numpy as np
import pytest

from oracle import oracle as orc

EPS = np.finfo(np.float64).eps


def unit(n, i, value=1.0):
    v = np.zeros(n)
    v[i] = value
    return v


def space_vectors(space, n):
    return [space.get(i, n) for i in range(space.size())]


@pytest.mark.parametrize("magnitude", [1e-20, 1.0, 1e20])
@pytest.mark.parametrize("capacity", [1, 2, 3])
def test_contract_scaled_diagonal_and_pair(magnitude, capacity):
    """review_harmonic.cpp:23-47: A = [[0.1 m, 0, 0], [0, m, -m], [0, m, m]],
    Z = I: harmonic values 0.1 m and m(1 +- i).  Capacity 3 keeps all three,
    capacities 1 and 2 keep the real one only (the pair must fit whole); every
    kept u has A u = c and C is orthonormal."""
    m = magnitude
    om = orc.OMatrix.csr(3, 0, 1, [0, 1, 3, 5], [0, 1, 2, 1, 2], [0.1 * m, m, -m, m, m])
    Z = np.eye(3)
    space, _, tel = orc.extract_harmonic_ritz(om, Z, capacity)
    assert space.size() == (3 if capacity == 3 else 1)
    # images (3 applications, 1 reduction) and each tried candidate (1
    # application; 2 reductions into an empty space, 4 otherwise)
    assert tel == ({"operator_applications": 6, "global_reductions": 11} if capacity == 3
                   else {"operator_applications": 4, "global_reductions": 3})
    tol = 512 * EPS
    vecs = space_vectors(space, 3)
    for u, c in vecs:
        assert np.linalg.norm(om.spmv(u) - c) <= tol * max(1.0, np.linalg.norm(om.spmv(u)))
    Cm = np.array([c for _, c in vecs])
    assert np.max(np.abs(Cm @ Cm.T - np.eye(len(vecs)))) <= tol
    if capacity < 3:  # the real harmonic value 0.1 m: u along e0
        u = vecs[0][0]
        assert abs(u[1]) + abs(u[2]) <= tol * abs(u[0])


def test_contract_flexible_nonorthogonal_basis():
    """review_harmonic.cpp:48-70: Z = [(2,0,0), (0.25,1,0)] for the lower
    bidiagonal A of rows {1}, {0.5, 3}, {1, 4}: the kept u satisfies
    (AZ)^T (A u - theta u) = 0 for the theta of least squares."""
    om = orc.OMatrix.csr(3, 0, 1, [0, 1, 3, 5], [0, 0, 1, 1, 2], [1.0, 0.5, 3.0, 1.0, 4.0])
    Z = np.array([[2.0, 0.0, 0.0], [0.25, 1.0, 0.0]])
    space, _, _ = orc.extract_harmonic_ritz(om, Z, 1)
    assert space.size() == 1
    u, _ = space.get(0, 3)
    au = om.spmv(u)
    AZ = np.array([om.spmv(z) for z in Z])
    g, b = AZ @ au, AZ @ u
    theta = (b @ g) / (b @ b)
    assert np.linalg.norm(g - theta * b) <= 512 * EPS * np.linalg.norm(g)


def test_contract_infinite_value_and_unfitting_pair():
    """review_harmonic.cpp:71-80 (nonsingular rotation, (AZ)^T Z = 0: the
    only harmonic value is infinite, nothing is kept) and 104-119 (a complex
    pair with capacity 1: nothing is kept)."""
    rot = orc.OMatrix.csr(2, 0, 1, [0, 1, 2], [1, 0], [-1.0, 1.0])
    space, _, _ = orc.extract_harmonic_ritz(rot, np.array([unit(2, 0)]), 1)
    assert space.size() == 0
    pair = orc.OMatrix.csr(2, 0, 1, [0, 2, 4], [0, 1, 0, 1], [1.0, -1.0, 1.0, 1.0])
    space, _, _ = orc.extract_harmonic_ritz(pair, np.eye(2), 1)
    assert space.size() == 0
    space, _, _ = orc.extract_harmonic_ritz(pair, np.eye(2), 2)
    assert space.size() == 2


def test_harmonic_ritz_property_on_c0_snapshot(c0):
    """On a real FGMRES snapshot, the kept candidates U (real vectors and the
    parts of conjugate pairs) span a harmonic Ritz subspace: the residual
    A U - U T is orthogonal to A Z for a small real T, which a dropped image,
    a transposed pencil or a wrong eigenvector column breaks.  The kept
    directions come from the smallest |theta| groups."""
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    _, _, (H, Z) = orc.fgmres_snapshot(om, orc.OPreconditioner(om, orc.OPreconditioner.JACOBI), s.rhs, restart=12,
                                      maximum_iterations=12, relative_tolerance=1e-15)
    space, cands, _ = orc.extract_harmonic_ritz(om, Z, 4)
    assert 1 <= space.size() <= 4 and len(cands) == space.size()
    n = om.owned * om.block
    AZ = np.array([om.spmv(z)[:n] for z in Z])
    U = np.array(cands)[:, :n]
    AU = np.array([om.spmv(c)[:n] for c in cands])
    G, B = AZ @ AU.T, AZ @ U.T
    T = np.linalg.lstsq(B, G, rcond=None)[0]
    assert np.linalg.norm(G - B @ T) <= 1e-9 * np.linalg.norm(G)
    # the kept harmonic values are the smallest in magnitude of the pencil
    import scipy.linalg as sla
    w = sla.eigvals(AZ @ AZ.T, AZ @ Z[:, :n].T)
    kept = np.sort(np.abs(np.linalg.eigvals(T)))
    allv = np.sort(np.abs(w[np.isfinite(w)]))
    assert np.allclose(kept, allv[:len(kept)], rtol=1e-6)


def test_gcrodr_sequence_on_c0(c0):
    """test_core.cpp:611-634: gcrodr converges and keeps its recycle space;
    a related system then converges with the space."""
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
    space = orc.ORecycleSpace(4)
    x, r, space, (H, Z) = orc.gcrodr(om, pc, s.rhs, space, 4, restart=20)
    assert r["status"] == 0 and space.size() == 4
    n = s.nodes * 3
    assert np.linalg.norm(s.rhs - om.spmv(x)[:n]) <= 1e-8 * np.linalg.norm(s.rhs) * 1.01
    b2 = s.rhs * (1 + 0.01 * np.random.default_rng(5).standard_normal(n))
    _, plain = orc.fgmres(om, pc, b2, restart=20)
    x2, r2, space2, _ = orc.gcrodr(om, pc, b2, space, 4, restart=20)
    print(plain["iterations"], r2["iterations"], space2.size())
    assert r2["status"] == 0 and r2["iterations"] < plain["iterations"]


def core_block_system():
    """test_core.cpp:474-507: 3 nodes x 2 components, exact solution 1..6,
    Jacobi, rtol 1e-12, check interval 1, restart 8, no residual
    replacement."""
    om = orc.OMatrix.csr(3, 0, 2, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2],
                         [4, 5, -1, -1, -1, -1, 4, 5, -1, -1, -1, -1, 3, 4])
    kw = dict(relative_tolerance=1e-12, convergence_check_interval=1, residual_replacement_interval=0,
              maximum_iterations=100, restart=8)
    return om, kw


def test_gcrodr_reference_block_system():
    """test_core.cpp:611-634: gcrodr with capacity 2 converges, keeps two
    harmonic Ritz vectors and records a snapshot; a related system (the
    exact solution moved at two entries) then converges keeping two."""
    om, kw = core_block_system()
    pc = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
    exact = np.arange(1, 7, dtype=float)
    x, r, space, (H, Z) = orc.gcrodr(om, pc, om.spmv(exact), orc.ORecycleSpace(2), 2, **kw)
    assert r["status"] == 0 and space.size() == 2 and H.shape[1] > 0
    assert np.max(np.abs(x - exact)) < 1e-9
    related = exact.copy()
    related[0] += 0.25
    related[5] -= 0.5
    x2, r2, space2, _ = orc.gcrodr(om, pc, om.spmv(related), space, 2, **kw)
    assert r2["status"] == 0 and space2.size() == 2
    assert np.max(np.abs(x2 - related)) < 1e-9


@pytest.mark.parametrize("pair", [False, True])
def test_reference_review_main_cases(pair):
    """review_harmonic.cpp:102-138: a two-column snapshot with capacity 1; the
    complex pair keeps nothing, the real case keeps a vector satisfying the
    harmonic-Ritz relation."""
    if pair:
        om = orc.OMatrix.csr(2, 0, 1, [0, 2, 4], [0, 1, 0, 1], [1.0, -1.0, 1.0, 1.0])
        Z = np.eye(2)
    else:
        om = orc.OMatrix.csr(3, 0, 1, [0, 1, 3, 5], [0, 0, 1, 1, 2], [1.0, 0.5, 3.0, 1.0, 4.0])
        Z = np.array([[2.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    space, _, _ = orc.extract_harmonic_ritz(om, Z, 1)
    if pair:
        assert space.size() == 0
        return
    assert space.size() == 1
    u, _ = space.get(0, 3)
    AZ = np.array([om.spmv(z) for z in Z])
    g, b = AZ @ om.spmv(u), AZ @ u
    theta = (b @ g) / (b @ b)
    assert np.linalg.norm(g - theta * b) <= 1e-12 * np.linalg.norm(g)
