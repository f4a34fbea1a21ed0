"""The harmonic-Ritz eigenproblem of GCRO-DR (gcrodr.hpp:30-78, 105-135), on
the host: owtk_harmonic_ggev against LAPACK's dggev (scipy.linalg.eig with
two matrices), which is the routine the reference calls.  Eigenvalues are
compared as multisets, eigenvectors through their residuals, and the output
follows dggev's convention (conjugate pairs consecutive, imag > 0 first,
vectors as real/imaginary columns, beta = 0 for infinite values)."""
import numpy as np
import pytest
import scipy.linalg as sla

import owtk


def spd(rng, n, cond=1e3):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return q @ np.diag(np.logspace(0, np.log10(cond), n)) @ q.T


def finite_thetas(re, im, be):
    keep = be != 0
    return (re[keep] + 1j * im[keep]) / be[keep]


def same_multiset(got, want, rtol=1e-8):
    """Each computed value matched to a distinct reference value."""
    want = list(want)
    if len(got) != len(want):
        return False
    scale = max(1.0, max(abs(w) for w in want)) if want else 1.0
    for g in got:
        k = int(np.argmin([abs(g - w) for w in want]))
        if abs(g - want[k]) > rtol * max(abs(want[k]), 1e-2 * scale):
            return False
        want.pop(k)
    return True


def check_vectors(G, M, re, im, be, right):
    n = G.shape[0]
    j = 0
    scale = np.linalg.norm(G, 2) + np.linalg.norm(M, 2)
    while j < n:
        if be[j] == 0:
            j += 1
            continue
        theta = (re[j] + 1j * im[j]) / be[j]
        if im[j] > 0:
            y = right[:, j] + 1j * right[:, j + 1]
            assert im[j + 1] == pytest.approx(-im[j]) and re[j + 1] == pytest.approx(re[j])
            step = 2
        else:
            assert im[j] == 0
            y = right[:, j].astype(complex)
            step = 1
        res = np.linalg.norm(G @ y - theta * (M @ y))
        assert res <= 1e-9 * scale * max(1.0, abs(theta)) * np.linalg.norm(y), (j, theta, res)
        j += step


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 20])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_pencils_match_dggev(n, seed):
    rng = np.random.default_rng(100 * n + seed)
    G = spd(rng, n)
    M = rng.standard_normal((n, n))
    re, im, be, right = owtk.harmonic_ggev(G, M)
    want = sla.eig(G, M, right=False)
    want = want[np.isfinite(want)]
    got = finite_thetas(re, im, be)
    assert same_multiset(got, want), (got, want)
    check_vectors(G, M, re, im, be, right)


def test_harmonic_ritz_pencil_from_a_basis():
    """G = (AZ)^T AZ and M = (AZ)^T Z for a nonsymmetric A and a flexible
    basis Z, as extract_harmonic_ritz forms them."""
    rng = np.random.default_rng(7)
    A = np.diag(np.linspace(1, 10, 30)) + 0.3 * rng.standard_normal((30, 30))
    Z = rng.standard_normal((30, 9))
    AZ = A @ Z
    G, M = AZ.T @ AZ, AZ.T @ Z
    re, im, be, right = owtk.harmonic_ggev(G, M)
    want = sla.eig(G, M, right=False)
    assert same_multiset(finite_thetas(re, im, be), want)
    check_vectors(G, M, re, im, be, right)


def test_singular_metric_gives_infinite_values():
    """review_harmonic.cpp:90-100: (AZ)^T Z = 0 gives an infinite harmonic
    value (beta = 0); a rank-one metric gives n - 1 of them."""
    re, im, be, right = owtk.harmonic_ggev(np.array([[1.0]]), np.array([[0.0]]))
    assert be[0] == 0
    rng = np.random.default_rng(3)
    G = spd(rng, 4)
    u = rng.standard_normal(4)
    M = np.outer(u, rng.standard_normal(4))
    re, im, be, right = owtk.harmonic_ggev(G, M)
    assert np.sum(be == 0) == 3
    theta = finite_thetas(re, im, be)
    want = sla.eig(G, M, right=False)
    assert same_multiset(theta, want[np.isfinite(want)])


def test_complex_pair_convention():
    """A rotation-like pencil: one conjugate pair, imag > 0 first, the vector
    as (real, imaginary) columns."""
    G = np.eye(2)
    M = np.array([[1.0, -1.0], [1.0, 1.0]])  # eigenvalues of M^-1 are complex
    re, im, be, right = owtk.harmonic_ggev(G, M)
    assert im[0] > 0 and im[1] == pytest.approx(-im[0]) and np.all(be == 1)
    check_vectors(G, M, re, im, be, right)


def test_indefinite_first_matrix_is_refused():
    """(AZ)^T AZ is a Gram matrix; an indefinite one is not such an input."""
    with pytest.raises(owtk.InvalidArgumentError):
        owtk.harmonic_ggev(np.array([[1.0, 2.0], [2.0, 1.0]]), np.eye(2))


@pytest.mark.parametrize("dependent", ["repeat", "multiple", "zero"])
@pytest.mark.parametrize("seed", range(4))
def test_dependent_basis_columns(seed, dependent):
    """Z with a dependent column (a repeat, a multiple that rounds to a tiny
    positive Cholesky pivot, or zero): G = (AZ)^T AZ is singular, the pencil
    has one indeterminate value (0/0 in dggev), reported last with beta = 0
    and a zero vector; the determinate values are those of the pencil over
    the independent columns and their vectors satisfy G y = theta M y."""
    rng = np.random.default_rng(seed)
    n, k = 9, 5
    A = rng.standard_normal((n, n)) + 4 * np.eye(n)
    Z = rng.standard_normal((n, k))
    Z[:, 3] = {"repeat": 1.0, "multiple": -3.7, "zero": 0.0}[dependent] * Z[:, 1]
    AZ = A @ Z
    G, M = AZ.T @ AZ, AZ.T @ Z
    re, im, be, right = owtk.harmonic_ggev(G, M)
    assert be[-1] == 0 and re[-1] == 0 and im[-1] == 0 and not np.any(right[:, -1])
    assert np.all(be[:-1] == 1)
    keep = [0, 1, 2, 4]
    w = sla.eigvals(G[np.ix_(keep, keep)], M[np.ix_(keep, keep)])
    assert same_multiset(finite_thetas(re, im, be), w)
    check_vectors(G, M, re, im, be, right)


@pytest.mark.parametrize("seed", range(4))
def test_small_independent_column_is_kept(seed):
    """A column of Z a million times shorter than the others is independent:
    no value is dropped, all match dggev on the full pencil."""
    rng = np.random.default_rng(seed)
    n, k = 9, 5
    A = rng.standard_normal((n, n)) + 4 * np.eye(n)
    Z = rng.standard_normal((n, k))
    Z[:, 2] *= 1e-6
    AZ = A @ Z
    G, M = AZ.T @ AZ, AZ.T @ Z
    re, im, be, right = owtk.harmonic_ggev(G, M)
    assert np.all(be == 1)
    assert same_multiset(finite_thetas(re, im, be), sla.eigvals(G, M), rtol=1e-6)
    check_vectors(G, M, re, im, be, right)
