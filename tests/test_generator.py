"""Synthetic grids and orderings (host code of the library, no GPU)."""
import numpy as np
import pytest

import owtk


def edges(nx, ny):
    return (nx - 1) * ny + nx * (ny - 1) + (nx - 1) * (ny - 1)


@pytest.mark.parametrize("nx,ny", [(64, 32), (17, 9), (2, 2)])
def test_sizes_and_structure(nx, ny):
    s = owtk.generate_split(nx, ny, 3, 7)
    assert s.nodes == nx * ny
    assert s.offsets[-1] == 2 * edges(nx, ny)
    rows = np.repeat(np.arange(s.nodes), np.diff(s.offsets))
    # no self loops, sorted unique columns, symmetric pattern
    assert np.all(s.columns != rows)
    for r in range(0, s.nodes, max(1, s.nodes // 50)):
        seg = s.columns[s.offsets[r]:s.offsets[r + 1]]
        assert np.all(np.diff(seg) > 0)
    pairs = set(zip(rows.tolist(), s.columns.tolist()))
    assert all((c, r) in pairs for r, c in pairs)
    # diagonal = row |off-diagonal| sum + 0.05 per component; off-diagonals negative in [-1.3, -0.07]
    absum = np.zeros((s.nodes, 3))
    np.add.at(absum, rows, np.abs(s.off_diagonal.reshape(-1, 3)))
    np.testing.assert_allclose(s.diagonal.reshape(-1, 3), absum + 0.05, rtol=1e-14)
    assert np.all(s.off_diagonal < -0.1 * 0.7 + 1e-15) and np.all(s.off_diagonal > -1.3)
    assert np.all(np.abs(s.rhs) <= 1.0)
    assert sorted(s.final_to_natural.tolist()) == list(range(s.nodes))


def test_operator_invariant_under_ordering():
    """A_final = P A_natural P^T: the same physical operator under RCM, Hilbert
    and identity numberings."""
    nat = owtk.generate_split(20, 12, 3, 5, ordering=owtk.ORDER_IDENTITY, permute=False)
    assert np.array_equal(nat.final_to_natural, np.arange(nat.nodes))
    for ordering in (owtk.ORDER_RCM, owtk.ORDER_HILBERT):
        s = owtk.generate_split(20, 12, 3, 5, ordering=ordering)
        f2n = s.final_to_natural
        for f in range(s.nodes):
            u = f2n[f]
            np.testing.assert_array_equal(s.diagonal[3 * f:3 * f + 3], nat.diagonal[3 * u:3 * u + 3])
            np.testing.assert_array_equal(s.rhs[3 * f:3 * f + 3], nat.rhs[3 * u:3 * u + 3])
            got = {int(f2n[c]): s.off_diagonal[3 * e:3 * e + 3].tolist() for e, c in
                   zip(range(s.offsets[f], s.offsets[f + 1]), s.columns[s.offsets[f]:s.offsets[f + 1]])}
            want = {int(c): nat.off_diagonal[3 * e:3 * e + 3].tolist() for e, c in
                    zip(range(nat.offsets[u], nat.offsets[u + 1]), nat.columns[nat.offsets[u]:nat.offsets[u + 1]])}
            assert got == want


def test_hilbert_index_closed_form():
    # classic curve on a 2x2 and 4x4 grid (ordering.hpp:64-77 semantics)
    pts = np.array([[0, 0], [0, 1], [1, 1], [1, 0]], float)
    assert owtk.order_hilbert(pts).tolist() == [0, 1, 2, 3]
    xs, ys = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
    pts = np.stack([xs.ravel(), ys.ravel()], 1).astype(float)
    order = owtk.order_hilbert(pts)
    # consecutive points along the curve are grid neighbours
    walk = pts[order]
    assert np.all(np.abs(np.diff(walk, axis=0)).sum(1) == 1)
    assert tuple(walk[0]) == (0, 0) and tuple(walk[-1]) == (3, 0)


def test_rcm_path_graph():
    # tridiagonal 4 x 4 (test_core.cpp:382-401): degrees 1,2,2,1 -> BFS 0,1,2,3, reversed
    order = owtk.order_rcm(4, [0, 2, 5, 8, 10], [0, 1, 0, 1, 2, 1, 2, 3, 2, 3])
    assert order.tolist() == [3, 2, 1, 0]


def test_rcm_reduces_bandwidth():
    perm = owtk.generate_split(64, 32, 1, 3, ordering=owtk.ORDER_IDENTITY, permute=True)
    rcm = owtk.generate_split(64, 32, 1, 3, ordering=owtk.ORDER_RCM)

    def bandwidth(s):
        rows = np.repeat(np.arange(s.nodes), np.diff(s.offsets))
        return int(np.max(np.abs(rows - s.columns)))

    assert bandwidth(rcm) < 100 < bandwidth(perm)


def test_lognormal_field_statistics():
    s = owtk.generate_split(256, 128, 1, 11, law=owtk.LAW_LOGNORMAL_FIELD, sigma=1.5, correlation=16.0,
                            ordering=owtk.ORDER_IDENTITY, permute=False)
    base = owtk.generate_split(256, 128, 1, 11, ordering=owtk.ORDER_IDENTITY, permute=False)
    log_ratio = np.log(s.off_diagonal / base.off_diagonal)   # = 1.5 (g_u + g_v) / 2
    assert 0.5 < np.std(log_ratio) < 1.6
    assert abs(np.mean(log_ratio)) < 0.6


def test_bsr_generator():
    n, off, col, val, rhs = owtk.generate_bsr(16, 8, 4, 0, ordering=owtk.ORDER_HILBERT)
    assert n == 128 and off[-1] == 2 * edges(16, 8) + n
    blocks = val.reshape(-1, 4, 4)
    rows = np.repeat(np.arange(n), np.diff(off))
    diag = blocks[col == rows]
    offd = blocks[col != rows]
    assert diag.shape[0] == n
    assert abs(np.mean(np.diagonal(diag, axis1=1, axis2=2)) - 6.0) < 0.5
    assert np.std(offd) < 0.15


def test_split_to_csr_round_trip():
    s = owtk.generate_split(9, 7, 2, 1)
    co, cc, cv = s.csr()
    assert co[-1] == s.offsets[-1] + s.nodes
    for r in range(s.nodes):
        cols = cc[co[r]:co[r + 1]]
        assert np.all(np.diff(cols) > 0) and r in cols
        d = np.where(cols == r)[0][0] + co[r]
        np.testing.assert_array_equal(cv[2 * d:2 * d + 2], s.diagonal[2 * r:2 * r + 2])
