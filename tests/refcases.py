"""Independent references that pin the oracle and the GPU path.

Known-answer fixtures restate the inputs and expected values of the
reference's own tests (cited per function); the long-double dense
algorithms follow the reference's audit method (dense symbolic elimination,
dot-product Doolittle, dense BiCGSTAB recurrence: tests/ilu_audit_reference.hpp
and tests/test_bicgstab_ilu_audit.cpp) and are written independently of the
sparse code they check.
"""
from __future__ import annotations

import numpy as np

LD = np.longdouble


# ---------------------------------------------------------------- SpMV cases
def core_csr_case():
    """test_core.cpp:31-58: 2 nodes x 3 components, exact integers."""
    offsets = [0, 2, 4]
    columns = [0, 1, 0, 1]
    values = [4, 5, 6, 1, 1, 1, 2, 2, 2, 3, 4, 5]
    x = [1, 2, 3, 4, 5, 6]
    expected = [8, 15, 24, 14, 24, 36]
    return offsets, columns, values, x, expected


def core_bsr_case():
    """test_core.cpp:176-195: dense 2x2 blocks."""
    offsets = [0, 2, 4]
    columns = [0, 1, 0, 1]
    values = [2, 1, 0, 3, -1, 0, 1, -1, 1, 0, 0, -1, 3, 2, 1, 4]
    x = [1, 2, 3, 4]
    expected = [1, 5, 18, 17]
    return offsets, columns, values, x, expected


def core_split_case():
    """test_core.cpp:233-279: split view, 2 owned + 1 ghost, block 2, then a
    numeric update diag[0]=8, edge[0]=-2 giving out[0] = 2."""
    offsets = [0, 1, 2]
    columns = [1, 2]
    diagonal = [4, 5, 6, 7]
    off_diagonal = [-1, -2, -3, -4]
    x = [1, 2, 3, 4, 5, 6]
    expected = [1, 2, 3, 4]
    return offsets, columns, diagonal, off_diagonal, x, expected


def dense_from_split(n, ghost, bs, offsets, columns, diagonal, off_diagonal, component):
    """Dense (n x (n+ghost)) long-double matrix of one component."""
    a = np.zeros((n, n + ghost), dtype=LD)
    for row in range(n):
        a[row, row] = diagonal[row * bs + component]
        for e in range(offsets[row], offsets[row + 1]):
            a[row, columns[e]] += off_diagonal[e * bs + component]
    return a


# --------------------------------------------------------- preconditioners
def ssor_review_case(block, omega):
    """review_contracts.cpp:29-96: native vs split SSOR closed form, 2 owned
    nodes + 1 ghost (ghost couplings 19 must be ignored)."""
    diag = np.zeros(2 * block)
    edges = np.full(4 * block, 19.0)
    rhs = np.full(3 * block, 77.0)
    for c in range(block):
        diag[c] = 4 + c % 3 + 0.25
        diag[block + c] = 3 + c % 5 + 0.25
        edges[c] = 1.0
        edges[2 * block + c] = 2.0
        rhs[c] = 1 + (c % 7) / 10
        rhs[block + c] = 2.0
    split = dict(offsets=[0, 2, 4], columns=[1, 2, 0, 2], diagonal=diag, off_diagonal=edges)
    csr_values = np.zeros(6 * block)
    for c in range(block):
        csr_values[c] = diag[c]
        csr_values[block + c] = edges[c]
        csr_values[2 * block + c] = 19.0
        csr_values[3 * block + c] = edges[2 * block + c]
        csr_values[4 * block + c] = diag[block + c]
        csr_values[5 * block + c] = 19.0
    csr = dict(offsets=[0, 3, 6], columns=[0, 1, 2, 0, 1, 2], values=csr_values)
    expected = np.zeros(2 * block)
    for c in range(block):
        d0, d1 = LD(diag[c]), LD(diag[block + c])
        y0 = LD(rhs[c]) / d0
        y1 = (LD(rhs[block + c]) - LD(omega) * 2 * y0) / d1
        z0 = (d0 * y0 - LD(omega) * y1) / d0
        expected[c] = float(LD(omega) * (2 - LD(omega)) * z0)
        expected[block + c] = float(LD(omega) * (2 - LD(omega)) * y1)
    return split, csr, rhs, expected


def ilu_audit_matrix():
    """test_bicgstab_ilu_audit.cpp:14-25: 5 owned nodes + 1 ghost, 3
    components, deliberately unsorted columns."""
    n, comps = 5, 3
    offsets = [0, 4, 7, 11, 14, 18]
    columns = [2, 5, 0, 1, 3, 1, 0, 4, 2, 0, 3, 3, 4, 1, 2, 5, 4, 3]
    values = np.zeros(18 * comps)
    for row in range(n):
        for e in range(offsets[row], offsets[row + 1]):
            for c in range(comps):
                values[e * comps + c] = (3 + c) if columns[e] == row else -float(LD(0.15) * (1 + (row + c + e) % 5))
    return n, comps, offsets, columns, values


def dense_symbolic(pattern, fill):
    """Pivot-major dense symbolic ILU(k) levels (test_bicgstab_ilu_audit.cpp:46-56)."""
    n = len(pattern)
    big = n * n
    levels = [[0 if pattern[i][j] else big for j in range(n)] for i in range(n)]
    for k in range(n):
        for i in range(k + 1, n):
            if levels[i][k] <= fill:
                for j in range(k + 1, n):
                    if levels[k][j] <= fill:
                        levels[i][j] = min(levels[i][j], levels[i][k] + levels[k][j] + 1)
    return [[levels[i][j] <= fill for j in range(n)] for i in range(n)]


def doolittle(a, pattern):
    """Dense dot-product Doolittle on a retained pattern, long double
    (ilu_audit_reference.hpp:63-90)."""
    n = len(a)
    lower = np.zeros((n, n), dtype=LD)
    upper = np.zeros((n, n), dtype=LD)
    for i in range(n):
        lower[i, i] = 1
        for j in range(i):
            if not pattern[i][j]:
                continue
            v = LD(a[i][j])
            for k in range(j):
                v -= lower[i, k] * upper[k, j]
            lower[i, j] = v / upper[j, j]
        for j in range(i, n):
            if not pattern[i][j]:
                continue
            v = LD(a[i][j])
            for k in range(i):
                v -= lower[i, k] * upper[k, j]
            upper[i, j] = v
    return lower, upper


def lu_solve(lower, upper, rhs):
    x = np.array(rhs, dtype=LD)
    n = len(x)
    for i in range(n):
        for j in range(i):
            x[i] -= lower[i, j] * x[j]
    for i in reversed(range(n)):
        for j in range(i + 1, n):
            x[i] -= upper[i, j] * x[j]
        x[i] /= upper[i, i]
    return x


def recurrence_matrix():
    """test_bicgstab_ilu_audit.cpp:104-131: 12 x 12 cyclic nonsymmetric."""
    n = 12
    a = np.zeros((n, n), dtype=LD)
    for i in range(n):
        a[i, i] = LD(2.1) + LD(0.03) * i
        a[i, (i + 1) % n] = LD(-0.8)
        a[i, (i + 5) % n] = LD(-0.65)
        a[i, (i + n - 2) % n] = LD(-0.3)
    offsets, columns, values = [0], [], []
    for i in range(n):
        for j in range(n):
            if a[i, j] != 0:
                columns.append(j)
                values.append(float(a[i, j]))
        offsets.append(len(columns))
    rhs = np.array([float(LD(0.7) + LD(0.13) * (i % 5)) for i in range(n)])
    x0 = np.array([float(LD(0.04) * (i % 4)) for i in range(n)])
    return a, offsets, columns, values, rhs, x0


def dense_bicgstab(a, b, x0, inverse, count):
    """Long-double dense BiCGSTAB recurrence (ilu_audit_reference.hpp:119-156).
    Returns the iterate after each of `count` iterations."""
    a = np.asarray(a, dtype=LD)
    x = np.array(x0, dtype=LD)
    r = np.array(b, dtype=LD) - a @ x
    shadow = r.copy()
    p = np.zeros_like(r)
    v = np.zeros_like(r)
    rho_prev = alpha = omega = LD(1)
    history = []
    for it in range(count):
        rho = shadow @ r
        beta = (rho / rho_prev) * (alpha / omega) if it else LD(0)
        p = r + beta * (p - omega * v)
        z = inverse(p)
        v = a @ z
        alpha = rho / (shadow @ v)
        s = r - alpha * v
        y = inverse(s)
        t = a @ y
        omega = (t @ s) / (t @ t)
        x = x + alpha * z + omega * y
        r = s - omega * t
        history.append(x.copy())
        rho_prev = rho
    return history
