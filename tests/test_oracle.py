"""Pin the CPU oracle against the reference's known answers, long-double dense
references and closed forms (no GPU)."""
import numpy as np
import pytest

import refcases as rc
from oracle import oracle as orc

EPS = np.finfo(np.float64).eps


def test_spmv_csr_known_answer():
    o, c, v, x, exp = rc.core_csr_case()
    m = orc.OMatrix.csr(2, 0, 3, o, c, v)
    assert m.spmv(x).tolist() == exp


def test_spmv_bsr_known_answer():
    o, c, v, x, exp = rc.core_bsr_case()
    m = orc.OMatrix.bsr(2, 0, 2, o, c, v)
    assert m.spmv(x).tolist() == exp


def test_spmv_split_known_answer_and_update():
    o, c, d, e, x, exp = rc.core_split_case()
    d = np.array(d, float)
    e = np.array(e, float)
    m = orc.OMatrix.split(2, 1, 2, o, c, d, e)
    assert m.spmv(x).tolist() == exp
    d[0], e[0] = 8, -2
    m2 = orc.OMatrix.split(2, 1, 2, o, c, d, e)
    assert m2.spmv(x)[0] == 2


def test_spmv_split_matches_long_double_dense(c0):
    s = c0
    x = np.random.default_rng(1).uniform(-1, 1, s.nodes * 3)
    y = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal).spmv(x)
    for comp in range(3):
        a = rc.dense_from_split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal, comp)
        exact = a @ x[comp::3].astype(rc.LD)
        assert np.max(np.abs(y[comp::3] - exact.astype(float))) <= 8 * EPS * np.max(np.abs(exact.astype(float)))


def test_csr_and_split_agree_bitwise(c0):
    s = c0
    co, cc, cv = s.csr()
    x = np.random.default_rng(2).uniform(-1, 1, s.nodes * 3)
    y_csr = orc.OMatrix.csr(s.nodes, 0, 3, co, cc, cv).spmv(x)
    y_split = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal).spmv(x)
    # different accumulation start (split: diag*x first, csr: sorted order) -> close, not identical
    np.testing.assert_allclose(y_csr, y_split, rtol=1e-14, atol=1e-14)


# ------------------------------------------- FP32 / BF16 SpMV oracle pinning
def _bf16(a):
    """Round to bfloat16 (nearest even) by bit manipulation, as FP32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


CASTS = {"f32": lambda a: np.asarray(a, np.float32), "bf16": _bf16}


@pytest.mark.parametrize("cast", ["f32", "bf16"])
def test_spmv_f32_known_answers(cast):
    """test_core.cpp:31-58, 176-195 and 233-279 run in float: every value is a
    small integer (exact in FP32 and BF16), so every sum is exact."""
    f = CASTS[cast]
    o, c, v, x, exp = rc.core_csr_case()
    m = orc.OMatrix.csr(2, 0, 3, o, c, v)
    assert m.spmv_f32(f(v), None, f(x)).tolist() == exp
    o, c, v, x, exp = rc.core_bsr_case()
    m = orc.OMatrix.bsr(2, 0, 2, o, c, v)
    assert m.spmv_f32(f(v), None, f(x)).tolist() == exp
    o, c, d, e, x, exp = rc.core_split_case()
    d, e = np.array(d, float), np.array(e, float)
    m = orc.OMatrix.split(2, 1, 2, o, c, d, e)
    assert m.spmv_f32(f(e), f(d), f(x)).tolist() == exp
    d[0], e[0] = 8, -2
    m2 = orc.OMatrix.split(2, 1, 2, o, c, d, e)
    assert m2.spmv_f32(f(e), f(d), f(x))[0] == 2


def test_bf16_cast_known_values():
    """The BF16 cast used against the oracle rounds to nearest, ties to even
    (7 stored mantissa bits: the spacing at 1 is 2^-7)."""
    v = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, -(1.0 + 2.0**-9), 1.0 + 2.0**-7 + 2.0**-8, 2.0**-130],
                 np.float32)
    expected = [1.0, 1.0, 1.0 + 2.0**-6, -1.0, 1.0 + 2.0**-6, 2.0**-130]
    assert _bf16(v).tolist() == expected


@pytest.mark.parametrize("cast", ["f32", "bf16"])
@pytest.mark.parametrize("kind", ["split", "csr", "bsr"])
def test_spmv_f32_long_double_bound(c0, cast, kind):
    """On c0 (and a dense 3x3-block BSR of the same graph): the FP32 result
    of the storage-rounded values differs from the long-double product of the
    same rounded values by at most (row entries + 1) u sum |a||x|, u = 2^-24
    (FP32 accumulation), and is not the FP64 result rounded once (a
    computation in FP64 would be more accurate than the bound detects, so the
    test also requires some rows to differ from it)."""
    f = CASTS[cast]
    s = c0
    rng = np.random.default_rng(11)
    bs = 3
    x = rng.uniform(-1, 1, s.nodes * bs).astype(np.float32)
    if kind == "split":
        vals, diag = f(s.off_diagonal), f(s.diagonal)
        m = orc.OMatrix.split(s.nodes, 0, bs, s.offsets, s.columns, s.diagonal, s.off_diagonal)
        y = m.spmv_f32(vals, diag, x)
        offsets, columns = np.asarray(s.offsets), np.asarray(s.columns)
        width = 1
    elif kind == "csr":
        co, cc, cv = s.csr()
        vals, diag = f(cv), None
        m = orc.OMatrix.csr(s.nodes, 0, bs, co, cc, cv)
        y = m.spmv_f32(vals, None, x)
        offsets, columns = np.asarray(co), np.asarray(cc)
        width = 1
    else:
        co, cc, _ = s.csr()
        blocks = rng.normal(0.0, 1.0, len(cc) * bs * bs)
        vals, diag = f(blocks), None
        m = orc.OMatrix.bsr(s.nodes, 0, bs, co, cc, blocks)
        y = m.spmv_f32(vals, None, x)
        offsets, columns = np.asarray(co), np.asarray(cc)
        width = bs
    xl = x.astype(rc.LD)
    vl = np.asarray(vals, np.float64).astype(rc.LD)
    worst = 0.0
    differs_from_rounded_exact = 0
    u = 2.0**-24
    for row in range(s.nodes):
        lo, hi = offsets[row], offsets[row + 1]
        for r in range(bs):
            exact = rc.LD(0)
            mag = rc.LD(0)
            if kind == "split":
                t = rc.LD(float(diag[row * bs + r])) * xl[row * bs + r]
                exact += t
                mag += abs(t)
            for e in range(lo, hi):
                col = columns[e]
                if width == 1:
                    t = vl[e * bs + r] * xl[col * bs + r]
                    exact += t
                    mag += abs(t)
                else:
                    for cc_ in range(bs):
                        t = vl[(e * bs + r) * bs + cc_] * xl[col * bs + cc_]
                        exact += t
                        mag += abs(t)
            terms = (hi - lo) * width + 2
            err = abs(float(y[row * bs + r]) - float(exact))
            bound = terms * u * float(mag)
            assert err <= bound, (row, r, err, bound)
            worst = max(worst, err / bound if bound > 0 else 0.0)
            if float(y[row * bs + r]) != float(np.float32(float(exact))):
                differs_from_rounded_exact += 1
    assert differs_from_rounded_exact > 0
    print(f"{kind}/{cast}: worst error {worst:.3f} of the bound; {differs_from_rounded_exact} rows carry FP32 rounding")


def test_norm_extremes():
    # test_foundation.cpp:71-107
    tiny = np.finfo(np.float64).tiny
    for mag in (1.0, np.finfo(np.float64).max / 8, tiny * 4, 5e-324 * 4):
        v = np.array([3 * mag, -4 * mag])
        assert abs(orc.norm(v) - 5 * mag) <= 4 * EPS * 5 * mag
    assert orc.norm(np.zeros(2)) == 0.0
    assert np.isinf(orc.norm(np.full(2, np.finfo(np.float64).max)))
    assert np.isinf(orc.norm(np.array([np.inf, 0.0])))
    assert np.isnan(orc.norm(np.array([np.inf, np.nan])))
    assert orc.norm(np.array([3.0, 4.0])) == 5.0


def test_dot_matches_exact_sum():
    rng = np.random.default_rng(3)
    a, b = rng.uniform(-1, 1, 1001), rng.uniform(-1, 1, 1001)
    exact = float(np.sum(a.astype(rc.LD) * b.astype(rc.LD)))
    assert abs(orc.dot(a, b) - exact) <= 64 * EPS * np.sum(np.abs(a * b))


def test_jacobi_and_ilu_updates_known_answer():
    # test_core.cpp:403-442 after the numeric update (values 5, -1, -1, 4)
    m = orc.OMatrix.csr(2, 0, 1, [0, 2, 4], [0, 1, 0, 1], [5.0, -1, -1, 4.0])
    rhs = m.spmv([2.0, 3.0])
    jac = orc.OPreconditioner(m, orc.OPreconditioner.JACOBI)
    assert np.allclose(jac.apply(rhs), rhs / np.array([5.0, 4.0]), rtol=0, atol=1e-13)
    ilu = orc.OPreconditioner(m, orc.OPreconditioner.ILU)
    assert np.allclose(ilu.apply(rhs), [2.0, 3.0], rtol=0, atol=1e-13)


def test_ssor_closed_form_review_case():
    # review_cases.cpp:95-113, omega = 0.5
    m = orc.OMatrix.csr(2, 0, 1, [0, 2, 4], [0, 1, 0, 1], [4.0, 1, 2, 3.0])
    w = 0.5
    out = orc.OPreconditioner(m, orc.OPreconditioner.SSOR, omega=w).apply([1.0, 1.0])
    y0 = 1.0 / 4
    y1 = (1 - w * 2 * y0) / 3
    z0 = (4 * y0 - w * y1) / 4
    assert abs(out[0] - w * (2 - w) * z0) < 1e-14 and abs(out[1] - w * (2 - w) * y1) < 1e-14


@pytest.mark.parametrize("block", [1, 5, 1296])
@pytest.mark.parametrize("omega", [0.5, 1.0, 1.5])
def test_ssor_native_and_split_closed_form(block, omega):
    # review_contracts.cpp:29-96: ghost couplings ignored, both layouts agree
    split, csr, rhs, expected = rc.ssor_review_case(block, omega)
    ms = orc.OMatrix.split(2, 1, block, split["offsets"], split["columns"], split["diagonal"], split["off_diagonal"])
    mc = orc.OMatrix.csr(2, 1, block, csr["offsets"], csr["columns"], csr["values"])
    for m in (ms, mc):
        pc = orc.OPreconditioner(m, orc.OPreconditioner.SSOR, omega=omega)
        out = pc.apply(rhs)
        assert np.all(np.abs(out / expected - 1) < 64 * EPS)
        assert np.array_equal(pc.apply_inplace(rhs[: 2 * block]), out)


@pytest.mark.parametrize("fill", [0, 1, 2])
def test_ilu_k_against_dense_doolittle(fill):
    # test_bicgstab_ilu_audit.cpp:14-91
    n, comps, offsets, columns, values = rc.ilu_audit_matrix()
    for update in range(2):
        vals = values.copy()
        if update:
            vals = vals * np.array([1.1 + 0.02 * (e % 3) for e in range(vals.size)])
        m = orc.OMatrix.csr(n, 1, comps, offsets, columns, vals)
        pc = orc.OPreconditioner(m, orc.OPreconditioner.ILU, fill_level=fill)
        off, col, fval, inv = pc.ilu_factors()
        for c in range(comps):
            dense = [[0.0] * n for _ in range(n)]
            pattern = [[False] * n for _ in range(n)]
            for row in range(n):
                for e in range(offsets[row], offsets[row + 1]):
                    if columns[e] < n:
                        dense[row][columns[e]] = vals[e * comps + c]
                        pattern[row][columns[e]] = True
            pat = rc.dense_symbolic(pattern, fill)
            for i in range(n):
                expected_cols = [j for j in range(n) if pat[i][j]]
                assert col[off[i]:off[i + 1]].tolist() == expected_cols
            lower, upper = rc.doolittle(dense, pat)
            prod = lower @ upper
            if fill == 0:
                assert abs(prod[1, 2]) > 0.001
            else:
                assert abs(prod[1, 2]) < 1e-16
            for i in range(n):
                for j in range(n):
                    if pat[i][j]:
                        assert abs(prod[i, j] - rc.LD(dense[i][j])) < 1e-15
            # factors equal the dense Doolittle factors
            for i in range(n):
                for e in range(off[i], off[i + 1]):
                    j = col[e]
                    ref = lower[i, j] if j < i else upper[i, j]
                    assert abs(fval[e * comps + c] - float(ref)) <= 1e-14 * max(1.0, abs(float(ref)))
            # inverse on every LU basis column
            for basis in range(n):
                rhs = np.zeros(n * comps)
                for i in range(n):
                    rhs[i * comps + c] = float(prod[i, basis])
                out = pc.apply(rhs)
                exp = np.zeros(n * comps)
                exp[basis * comps + c] = 1.0
                assert np.max(np.abs(out - exp)) < 32 * EPS * 4


def test_ilu_zero_pivots_rejected():
    # review_contracts.cpp:152-177, review_cases.cpp:127-139
    for pivot in (0.0, np.nan, np.inf):
        m = orc.OMatrix.csr(1, 0, 1, [0, 1], [0], [pivot])
        with pytest.raises(orc.OracleError) as err:
            orc.OPreconditioner(m, orc.OPreconditioner.ILU)
        assert err.value.code == -2
    m = orc.OMatrix.csr(2, 0, 1, [0, 2, 4], [0, 1, 0, 1], [1.0, 1, 1, 1])
    with pytest.raises(orc.OracleError):
        orc.OPreconditioner(m, orc.OPreconditioner.ILU)


@pytest.mark.parametrize("use_ilu", [False, True])
def test_bicgstab_recurrence_audit(use_ilu):
    # test_bicgstab_ilu_audit.cpp:104-183: 1-4 iterations against a long-double recurrence
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    m = orc.OMatrix.csr(12, 0, 1, offsets, columns, values)
    pattern = [[a[i, j] != 0 for j in range(12)] for i in range(12)]
    lower, upper = rc.doolittle(a, pattern)
    inverse = (lambda r: rc.lu_solve(lower, upper, r)) if use_ilu else (lambda r: r.copy())
    history = rc.dense_bicgstab(a, rhs, x0, inverse, 4)
    pc = orc.OPreconditioner(m, orc.OPreconditioner.ILU) if use_ilu else None
    for count in range(1, 5):
        x, res = orc.bicgstab(m, pc, rhs, x0, relative_tolerance=0.0, maximum_iterations=count,
                              convergence_check_interval=1, residual_replacement_interval=0)
        assert res["status"] == 1
        assert np.max(np.abs(x - history[count - 1].astype(float))) < 2e-11


def test_bicgstab_exact_preconditioner_one_iteration():
    # test_bicgstab_ilu_audit.cpp:185-200
    m = orc.OMatrix.csr(3, 0, 1, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [4.0, -1, -2, 5, -1, -1, 3])
    rhs = m.spmv([1.0, 2.0, 3.0])
    pc = orc.OPreconditioner(m, orc.OPreconditioner.ILU)
    x, res = orc.bicgstab(m, pc, rhs, relative_tolerance=32 * EPS, convergence_check_interval=1)
    assert res["status"] == 0 and res["iterations"] == 1 and res["preconditioner_applications"] == 1
    assert np.max(np.abs(x - [1, 2, 3])) < 32 * EPS * 4


def test_bicgstab_zero_operator_breakdown():
    # test_core.cpp:444-469
    m = orc.OMatrix.csr(1, 0, 1, [0, 1], [0], [0.0])
    _, res = orc.bicgstab(m, None, [1.0], convergence_check_interval=1)
    assert res["status"] == 2 and res["breakdown_reason"] == 4


def test_fgmres_equals_direct_solve():
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    m = orc.OMatrix.csr(12, 0, 1, offsets, columns, values)
    exact = np.linalg.solve(a.astype(float), rhs)
    for mode in (0, 1, 2):
        x, res = orc.fgmres(m, None, rhs, relative_tolerance=1e-13, restart=12, gmres_orthogonalization=mode)
        assert res["status"] == 0 and res["iterations"] <= 12
        assert np.max(np.abs(x - exact)) < 1e-11


def test_solvers_on_block_system():
    # test_core.cpp:471-560: 3 nodes x 2 components, Jacobi, rtol 1e-12
    m = orc.OMatrix.csr(3, 0, 2, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2],
                        [4, 5, -1, -1, -1, -1, 4, 5, -1, -1, -1, -1, 3, 4])
    exact = np.arange(1, 7, dtype=float)
    rhs = m.spmv(exact)
    jac = orc.OPreconditioner(m, orc.OPreconditioner.JACOBI)
    for solver in (orc.fgmres, orc.bicgstab):
        x, res = solver(m, jac, rhs, relative_tolerance=1e-12, convergence_check_interval=1,
                        residual_replacement_interval=0, maximum_iterations=100, restart=8)
        assert res["status"] == 0
        assert np.max(np.abs(x - exact)) < 1e-10
    x, res = orc.fgmres(m, jac, rhs, relative_tolerance=1e-12, restart=8)
    assert res["global_reductions"] <= 2 * res["iterations"] + 4 + res["arnoldi_norm_verifications"]


def test_c0_bicgstab_jacobi_converges(c0):
    s = c0
    m = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = orc.OPreconditioner(m, orc.OPreconditioner.JACOBI)
    x, res = orc.bicgstab(m, pc, s.rhs)
    assert res["status"] == 0
    r = s.rhs - m.spmv(x)
    assert np.linalg.norm(r) / np.linalg.norm(s.rhs) <= 1e-8
    assert res["relative_residual_norm"] <= 1e-8


def _greedy_colours(n, offsets, columns):
    adj = [set() for _ in range(n)]
    for i in range(n):
        for j in columns[offsets[i]:offsets[i + 1]]:
            if j != i and j < n:
                adj[i].add(int(j))
                adj[int(j)].add(i)
    colour = [-1] * n
    for i in range(n):
        used = {colour[j] for j in adj[i] if colour[j] >= 0}
        c = 0
        while c in used:
            c += 1
        colour[i] = c
    return np.array(colour)


@pytest.mark.parametrize("omega", [0.7, 1.0, 1.4])
def test_dense_ssor_textbook_form(omega):
    """Dense-block multicolour SSOR equals s (D + w U)^-1 D (D + w L)^-1 r of
    the colour-permuted matrix (dense solve)."""
    import owtk
    bs = 4
    n, off, col, val, rhs = owtk.generate_bsr(6, 4, bs, 3, ordering=owtk.ORDER_RCM)
    colour = _greedy_colours(n, off, col)
    order = np.argsort(colour, kind="stable")
    pos = np.empty(n, int)
    pos[order] = np.arange(n)
    dense = np.zeros((n * bs, n * bs))
    for i in range(n):
        for e in range(off[i], off[i + 1]):
            j = col[e]
            dense[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = val[e * bs * bs:(e + 1) * bs * bs].reshape(bs, bs)
    perm = np.concatenate([np.arange(o * bs, (o + 1) * bs) for o in order])
    P = dense[np.ix_(perm, perm)]
    D = np.zeros_like(P)
    L = np.zeros_like(P)
    U = np.zeros_like(P)
    for bi in range(n):
        for bj in range(n):
            blk = P[bi * bs:(bi + 1) * bs, bj * bs:(bj + 1) * bs]
            (D if bi == bj else (L if bj < bi else U))[bi * bs:(bi + 1) * bs, bj * bs:(bj + 1) * bs] = blk
    r = np.random.default_rng(5).uniform(-1, 1, n * bs)
    s = omega * (2 - omega)
    zp = s * np.linalg.solve(D + omega * U, D @ np.linalg.solve(D + omega * L, r[perm]))
    ref = np.empty_like(zp)
    ref[perm] = zp
    om = orc.OMatrix.bsr(n, 0, bs, off, col, val)
    pc = orc.OPreconditioner(om, orc.OPreconditioner.SSOR, omega=omega)
    z = pc.apply(r)
    assert np.max(np.abs(z - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_pipelined_bicgstab_dense_recurrence():
    """Oracle pipelined BiCGSTAB on the audit's 12 x 12 cyclic matrix
    (test_bicgstab_ilu_audit.cpp:104-131) against a long-double dense
    restatement of the pipelined recurrence (pipelined_bicgstab.hpp:86-244):
    the first four iterates agree, and the solve converges."""
    import refcases as rc
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    nrow = a.shape[0]
    A = a.astype(np.longdouble)
    x = np.array(x0, np.longdouble)
    r = rhs.astype(np.longdouble) - A @ x
    shadow = r.copy()
    p = r.copy()
    rho = shadow @ r
    xs = []
    for _ in range(4):
        v = A @ p
        alpha = rho / (shadow @ v)
        s = r - alpha * v
        t = A @ s
        omega = (t @ s) / (t @ t)
        x = x + alpha * p + omega * s
        r = s - omega * t
        prev, rho = rho, shadow @ r
        beta = (rho / prev) * (alpha / omega)
        p = r + beta * (p - omega * v)
        xs.append(x.astype(np.float64).copy())
    om = orc.OMatrix.csr(nrow, 0, 1, np.array(offsets, np.int32), np.array(columns, np.int32), np.array(values))
    for k, xr in enumerate(xs, start=1):
        xo, res = orc.pipelined_bicgstab(om, None, rhs, x0=x0, relative_tolerance=0.0, maximum_iterations=k,
                                         residual_replacement_interval=0)
        assert res["iterations"] == k
        assert np.allclose(xo[:nrow], xr, rtol=1e-12, atol=1e-14), k
    xo, res = orc.pipelined_bicgstab(om, None, rhs, x0=x0, relative_tolerance=1e-12)
    assert res["status"] == 0
    assert np.linalg.norm(rhs - a.astype(np.float64) @ xo[:nrow]) <= 1.1e-12 * np.linalg.norm(rhs)


@pytest.mark.parametrize("use_ilu", [False, True])
def test_communication_hiding_bicgstab_equals_bicgstab_recurrence(use_ilu):
    """The communication-hiding recurrence (pipelined_bicgstab.hpp:317-597)
    carries auxiliary vectors (s = A M p, z = A M s, ...) but in exact
    arithmetic its iterates are BiCGSTAB's.  The oracle's first four iterates
    on the audit's 12 x 12 matrix agree with the long-double dense BiCGSTAB
    of ilu_audit_reference.hpp:119-156, unpreconditioned and with the exact
    ILU(0) of that matrix as M; the solve then converges with the counters
    the recurrence implies."""
    import refcases as rc
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    n = a.shape[0]
    om = orc.OMatrix.csr(n, 0, 1, np.array(offsets, np.int32), np.array(columns, np.int32), np.array(values))
    if use_ilu:
        pattern = [[a[i, j] != 0 for j in range(n)] for i in range(n)]
        lower, upper = rc.doolittle(a, pattern)
        inverse = lambda r: rc.lu_solve(lower, upper, r)  # noqa: E731
        pc = orc.OPreconditioner(om, 3)
    else:
        inverse, pc = (lambda r: r.copy()), None
    history = rc.dense_bicgstab(a, rhs, x0, inverse, 4)
    for k in range(1, 5):
        xo, res = orc.communication_hiding_bicgstab(om, pc, rhs, x0=x0, relative_tolerance=0.0, maximum_iterations=k,
                                                    residual_replacement_interval=0)
        assert res["iterations"] == k and res["status"] == 1
        assert np.max(np.abs(xo[:n] - history[k - 1].astype(np.float64))) < 1e-12, k
        # initial norms, the restart's rho and <w, rp>, 2 per iteration, final norm
        assert res["global_reductions"] == 1 + 2 + 2 * k + 1
        # restart: 2 applications of each; iteration: 2 of each; final true residual: 1 op
        assert res["preconditioner_applications"] == 2 + 2 * k
        assert res["operator_applications"] == 1 + 2 + 2 * k + 1
    xo, res = orc.communication_hiding_bicgstab(om, pc, rhs, x0=x0, relative_tolerance=1e-12)
    assert res["status"] == 0
    assert np.linalg.norm(rhs - a.astype(np.float64) @ xo[:n]) <= 1.1e-12 * np.linalg.norm(rhs)


def test_communication_hiding_bicgstab_breakdowns():
    # zero operator: <r, rp> is fine but <A M r, rp> = 0 stops the first restart
    om = orc.OMatrix.csr(1, 0, 1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([0.0]))
    _, res = orc.communication_hiding_bicgstab(om, None, np.array([1.0]))
    assert res["status"] == 2 and res["breakdown_reason"] == 3
    assert res["global_reductions"] == 3 and res["operator_applications"] == 2
    om = orc.OMatrix.csr(1, 0, 2, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([1.0, 1.0]))
    _, res = orc.communication_hiding_bicgstab(om, None, np.array([np.inf, 1.0]))
    assert res["status"] == 2 and res["breakdown_reason"] == 9


@pytest.mark.parametrize("s", [2, 3, 4, 6, 12])
def test_idrs_finite_termination(s):
    """IDR(s) (idrs.hpp:128-505) reaches the exact solution of an n x n
    system within n + n/s matrix-vector steps (Sonneveld & van Gijzen, SIAM J.
    Sci. Comput. 31 (2008), the IDR theorem): the oracle on the audit's 12 x 12
    nonsymmetric matrix stops within that bound at the solution of
    numpy.linalg.solve."""
    import math
    import refcases as rc
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    n = a.shape[0]
    om = orc.OMatrix.csr(n, 0, 1, np.array(offsets, np.int32), np.array(columns, np.int32), np.array(values))
    exact = np.linalg.solve(a.astype(np.float64), rhs)
    x, res = orc.idrs(om, None, rhs, x0=x0, relative_tolerance=1e-13, idr_shadow_space=s,
                      residual_replacement_interval=0)
    assert res["status"] == 0
    assert res["iterations"] <= n + math.ceil(n / s)
    assert np.max(np.abs(x[:n] - exact)) <= 1e-13 * np.max(np.abs(exact))


def test_idrs_one_is_bicgstab():
    """IDR(1) is routed through the BiCGSTAB recurrence (idrs.hpp:165-172)."""
    import refcases as rc
    a, offsets, columns, values, rhs, x0 = rc.recurrence_matrix()
    om = orc.OMatrix.csr(12, 0, 1, np.array(offsets, np.int32), np.array(columns, np.int32), np.array(values))
    xb, rb = orc.bicgstab(om, None, rhs, x0=x0, relative_tolerance=1e-13)
    x1, r1 = orc.idrs(om, None, rhs, x0=x0, relative_tolerance=1e-13, idr_shadow_space=1)
    assert np.array_equal(xb, x1) and rb == r1


def test_idrs_reference_cases(c0):
    """The reference's own IDR cases (test_core.cpp:776-798,
    test_foundation.cpp:110-145): IDR(2) + Jacobi converges; the same with the
    right-hand side scaled by 1e-18 (pivots are relative to their column);
    non-finite data is never accepted."""
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = orc.OPreconditioner(om, 1)
    x, res = orc.idrs(om, pc, s.rhs, idr_shadow_space=2)
    assert res["status"] == 0
    assert np.linalg.norm(s.rhs - om.spmv(x)) <= 1e-8 * np.linalg.norm(s.rhs)
    xs, rs = orc.idrs(om, pc, s.rhs * 1e-18, idr_shadow_space=2)
    assert rs["status"] == 0
    one = orc.OMatrix.csr(2, 0, 1, np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, 1.0]))
    for bad in (np.nan, np.inf):
        for in_rhs in (True, False):
            b = np.ones(2)
            x0 = np.ones(2)
            (b if in_rhs else x0)[0] = bad
            _, r = orc.idrs(one, None, b, x0=x0, idr_shadow_space=2, maximum_iterations=3)
            assert r["status"] != 0 and not np.isfinite(r["initial_residual_norm"])


def _dense_components(n, bs, off, col, val):
    a = np.zeros((bs, n, n))
    for r in range(n):
        for e in range(off[r], off[r + 1]):
            a[:, r, col[e]] += val[e * bs:(e + 1) * bs]
    return a


@pytest.mark.parametrize("aggregate", [1, 2, 4])
def test_multigrid_coarse_matrix_closed_form(c0, aggregate):
    """AggregationMultigrid's coarse operator (multigrid.hpp:264-305) is the
    row-averaged Galerkin product C = S^-1 R^T A R per component, R the 0/1
    aggregate indicator and S the aggregate sizes; aggregates partition the
    nodes, respect the target size and are seeded in row order."""
    s = c0
    co, cc, cv = s.csr()
    om = orc.OMatrix.csr(s.nodes, 0, 3, co, cc, cv)
    mg = orc.OMultigrid(om, aggregate=aggregate)
    nc, off, col, val, f2c = mg.coarse()
    assert f2c.min() == 0 and f2c.max() == nc - 1
    sizes = np.bincount(f2c, minlength=nc)
    assert sizes.max() <= aggregate and sizes.min() >= 1
    first = np.array([np.flatnonzero(f2c == k)[0] for k in range(nc)])
    assert np.all(np.diff(first) > 0)  # aggregate k is seeded before aggregate k + 1
    R = np.zeros((s.nodes, nc))
    R[np.arange(s.nodes), f2c] = 1.0
    A = _dense_components(s.nodes, 3, co, cc, cv)
    got = _dense_components(nc, 3, off, col, val)
    for c in range(3):
        ref = (R.T @ A[c] @ R) / sizes[:, None]
        assert np.max(np.abs(got[c] - ref)) <= 1e-14 * np.max(np.abs(ref))
    if aggregate == 1:
        assert np.array_equal(f2c, np.arange(s.nodes))


@pytest.mark.parametrize("smoother", [0, 1, 2])
@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_multigrid_reference_block_system(smoother, cycle):
    """The reference test's 3-node block system (test_core.cpp:474-490,
    800-844): aggregates of 2, 40 coarse smoothing steps; V, W and full cycles
    with Jacobi, red-black Gauss-Seidel and Chebyshev converge to 1e-9 at the
    exact solution 1..6."""
    om = orc.OMatrix.csr(3, 0, 2, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2],
                         [4, 5, -1, -1, -1, -1, 4, 5, -1, -1, -1, -1, 3, 4])
    exact = np.arange(1, 7, dtype=float)
    mg = orc.OMultigrid(om, aggregate=2, coarse_steps=40, smoother=smoother)
    nc, off, col, val, f2c = mg.coarse()
    assert f2c.tolist() == [0, 0, 1]
    assert val.tolist() == [3.0, 4.0, -0.5, -0.5, -1.0, -1.0, 3.0, 4.0]
    x, res = mg.solve(om.spmv(exact), cycle=cycle, relative_tolerance=1e-9, maximum_iterations=100)
    assert res["status"] == 0
    assert np.max(np.abs(x - exact)) < 1e-8


def test_recycled_fgmres_properties(c0):
    """recycled_fgmres / RecycleSpace (recycling.hpp:24-360):
    - with an empty space and no update it is fgmres, bit for bit;
    - the space keeps C = A U orthonormal;
    - re-solving a right-hand side the space already spans needs no Arnoldi
      step: x0 + U C^T r0 is the solution (GCRO's exact projection);
    - a nearby right-hand side needs fewer iterations than plain FGMRES;
    - at capacity the oldest pair is dropped."""
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    n = s.nodes * 3
    pc = orc.OPreconditioner(om, 1)
    space = orc.ORecycleSpace(2)
    xf, rf = orc.fgmres(om, pc, s.rhs)
    x0, r0 = orc.recycled_fgmres(om, pc, s.rhs, space, update=False)
    assert np.array_equal(xf, x0) and rf == r0 and space.size() == 0
    x1, r1 = orc.recycled_fgmres(om, pc, s.rhs, space)
    assert r1["status"] == 0 and space.size() == 1
    x2, r2 = orc.recycled_fgmres(om, pc, s.rhs, space)
    assert r2["status"] == 0 and r2["iterations"] == 0
    b3 = s.rhs * (1 + 0.01 * np.random.default_rng(3).standard_normal(n))
    _, rp = orc.fgmres(om, pc, b3)
    _, r3 = orc.recycled_fgmres(om, pc, b3, space)
    assert r3["status"] == 0 and r3["iterations"] < rp["iterations"]
    assert space.size() == 2
    U, Cm = zip(*(space.get(i, n) for i in range(space.size())))
    G = np.array([[c1 @ c2 for c2 in Cm] for c1 in Cm])
    assert np.max(np.abs(G - np.eye(len(Cm)))) < 1e-13
    for u, c in zip(U, Cm):
        assert np.max(np.abs(om.spmv(u) - c)) < 1e-12
    # capacity 2: a third independent candidate evicts the oldest
    u_old = U[0]
    added, _ = space.add_candidate(om, np.random.default_rng(4).standard_normal(n))
    assert added and space.size() == 2
    assert not np.array_equal(space.get(0, n)[0], u_old)
    # a candidate already in the space is rejected
    added, _ = space.add_candidate(om, space.get(1, n)[0])
    assert not added


def test_fgmres_snapshot_arnoldi_relation(c0):
    """fgmres's ArnoldiSnapshot (core.hpp:462-534; krylov_solvers.hpp:252-253,
    361-366, 435-437), pinned by the Arnoldi relation A Z = V H with V
    orthonormal:
    - without a preconditioner Z = V, so Z^T Z = I and Z^T A Z is the square
      part of H (a stored rotated H or a transposed H fails this);
    - with Jacobi, V = D Z (z = D^-1 v per component), so (D Z)^T (D Z) = I
      and (A Z)^T (A Z) = H^T H;
    - Z_0 = b / |b| from x0 = 0, and H's subdiagonal holds the norms (> 0);
    - the column count is the last cycle's, and an empty run clears it."""
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    _, res, (H, Z) = orc.fgmres_snapshot(om, None, s.rhs, restart=20, maximum_iterations=20,
                                         relative_tolerance=1e-15)
    k = H.shape[1]
    assert k == 20 and H.shape == (21, 20) and Z.shape[0] == 20
    assert np.max(np.abs(Z @ Z.T - np.eye(k))) < 1e-12
    AZ = np.array([om.spmv(z) for z in Z])
    assert np.max(np.abs(Z @ AZ.T - H[:k, :])) < 1e-12 * np.max(np.abs(H))
    assert np.max(np.abs(Z[0] - s.rhs / np.linalg.norm(s.rhs))) < 1e-15
    assert np.all(np.diag(H, -1) > 0) and np.all(np.tril(H, -2) == 0)
    jac = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
    _, res, (H, Z) = orc.fgmres_snapshot(om, jac, s.rhs, restart=5, maximum_iterations=12, relative_tolerance=1e-15)
    assert res["status"] == 1 and H.shape == (3, 2)  # cycles of 5, 5, 2
    V = Z * s.diagonal.reshape(-1)[None, :]
    assert np.max(np.abs(V @ V.T - np.eye(2))) < 1e-12
    AZ = np.array([om.spmv(z) for z in Z])
    assert np.max(np.abs(AZ @ AZ.T - H.T @ H)) < 1e-12 * np.max(np.abs(H)) ** 2
    _, _, (H, Z) = orc.fgmres_snapshot(om, None, np.zeros_like(s.rhs), restart=5)
    assert H.shape == (1, 0) and Z.shape[0] == 0
