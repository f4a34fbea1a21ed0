"""GPU parity of the kernels (SpMV, vector ops, preconditioners) against the
CPU oracle and the reference's known answers.  Integer/ordering work and the
FMA-explicit per-row kernels are compared bit for bit; reductions (different
summation tree) within 1e-13 relative."""
import numpy as np
import pytest

import owtk
import refcases as rc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def gpu_spmv(ctx, m, x, dtype=owtk.F64, out_dtype=np.float64):
    xd = ctx.upload(np.asarray(x, out_dtype))
    yd = ctx.empty(m.owned_size, out_dtype)
    m.spmv(xd, yd, dtype)
    return yd.download()


def test_device_properties(ctx):
    assert owtk.device_count() >= 1


def test_spmv_known_answers(ctx):
    o, c, v, x, exp = rc.core_csr_case()
    assert gpu_spmv(ctx, owtk.Matrix.csr(ctx, 2, 0, 3, o, c, v), x).tolist() == exp
    o, c, v, x, exp = rc.core_bsr_case()
    assert gpu_spmv(ctx, owtk.Matrix.bsr(ctx, 2, 0, 2, o, c, v), x).tolist() == exp
    o, c, d, e, x, exp = rc.core_split_case()
    m = owtk.Matrix.split(ctx, 2, 1, 2, o, c, d, e)
    assert gpu_spmv(ctx, m, x).tolist() == exp
    # explicit numeric update contract of the borrowed split view
    d2, e2 = np.array(d, float), np.array(e, float)
    d2[0], e2[0] = 8, -2
    m.update_values(d2, e2)
    assert gpu_spmv(ctx, m, x)[0] == 2


def test_spmv_c0_bitwise(ctx, c0):
    s = c0
    x = np.random.default_rng(4).uniform(-1, 1, s.nodes * 3)
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    assert np.array_equal(gpu_spmv(ctx, m, x), om.spmv(x))
    co, cc, cv = s.csr()
    omc = orc.OMatrix.csr(s.nodes, 0, 3, co, cc, cv)
    mc = owtk.Matrix.csr(ctx, s.nodes, 0, 3, co, cc, cv)
    assert np.array_equal(gpu_spmv(ctx, mc, x), omc.spmv(x))


@pytest.mark.parametrize("bs", [1, 2, 4, 5])
def test_spmv_block_sizes_bitwise(ctx, bs):
    s = owtk.generate_split(40, 24, bs, 9, ordering=owtk.ORDER_HILBERT)
    x = np.random.default_rng(bs).uniform(-1, 1, s.nodes * bs)
    om = orc.OMatrix.split(s.nodes, 0, bs, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, bs, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    assert np.array_equal(gpu_spmv(ctx, m, x), om.spmv(x))


def test_spmv_large_block_and_ghosts(ctx):
    split, csr, rhs, _ = rc.ssor_review_case(1296, 1.0)
    x = np.random.default_rng(5).uniform(-1, 1, 3 * 1296)
    om = orc.OMatrix.split(2, 1, 1296, split["offsets"], split["columns"], split["diagonal"], split["off_diagonal"])
    m = owtk.Matrix.split(ctx, 2, 1, 1296, split["offsets"], split["columns"], split["diagonal"], split["off_diagonal"])
    assert np.array_equal(gpu_spmv(ctx, m, x), om.spmv(x))
    omc = orc.OMatrix.csr(2, 1, 1296, csr["offsets"], csr["columns"], csr["values"])
    mc = owtk.Matrix.csr(ctx, 2, 1, 1296, csr["offsets"], csr["columns"], csr["values"])
    assert np.array_equal(gpu_spmv(ctx, mc, x), omc.spmv(x))


@pytest.mark.parametrize("bs", [2, 3, 4])
def test_spmv_bsr_bitwise(ctx, bs):
    n, off, col, val, rhs = owtk.generate_bsr(64, 32, bs, 0, ordering=owtk.ORDER_HILBERT)
    x = np.random.default_rng(6).uniform(-1, 1, n * bs)
    om = orc.OMatrix.bsr(n, 0, bs, off, col, val)
    m = owtk.Matrix.bsr(ctx, n, 0, bs, off, col, val)
    assert np.array_equal(gpu_spmv(ctx, m, x), om.spmv(x))


def test_spmv_c1_full_size_bitwise(ctx):
    """2048 x 1024 grid, 2,097,152 vertices, block 3, Hilbert order."""
    s = owtk.generate_split(2048, 1024, 3, 0, ordering=owtk.ORDER_HILBERT)
    assert s.nodes == 2097152 and s.offsets[-1] == 12570626
    x = np.random.default_rng(7).uniform(-1, 1, s.nodes * 3)
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    assert np.array_equal(gpu_spmv(ctx, m, x), om.spmv(x))
    info = m.info()
    assert info.spmv_bytes_f64 == 12570626 * 28 + 2097152 * 28 + 2 * 2097152 * 24


def _bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("dtype", [owtk.F32, owtk.BF16])
def test_spmv_low_precision_bitwise(ctx, c0, dtype):
    s = c0
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m.add_storage(dtype)
    x = np.random.default_rng(8).uniform(-1, 1, s.nodes * 3).astype(np.float32)
    y = gpu_spmv(ctx, m, x, dtype, np.float32)
    cast = (lambda a: np.asarray(a, np.float32)) if dtype == owtk.F32 else _bf16
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    ref = om.spmv_f32(cast(s.off_diagonal), cast(s.diagonal), x)
    assert np.array_equal(y, ref)


def test_vector_kernels(ctx):
    rng = np.random.default_rng(9)
    n = 1_000_003
    a, b, c = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    da, db, dc = ctx.upload(a), ctx.upload(b), ctx.upload(c)
    exact = float(np.sum(a.astype(rc.LD) * b.astype(rc.LD)))
    assert abs(ctx.dot(da, db) - exact) <= 1e-13 * np.sum(np.abs(a * b))
    assert abs(ctx.norm(da) - orc.norm(a)) <= 1e-14 * orc.norm(a)
    out = ctx.axpy_dot(0.5, da, db, dc)
    yb = np.fma(0.5, a, b) if hasattr(np, "fma") else 0.5 * a + b
    assert np.allclose(db.download(), yb, rtol=0, atol=1e-16)
    assert abs(out - float(np.dot(c, db.download()))) <= 1e-12 * np.sum(np.abs(c * yb))
    basis = [ctx.upload(rng.uniform(-1, 1, n)) for _ in range(7)]
    md = ctx.multidot(basis, da)
    for j, v in enumerate(basis):
        assert abs(md[j] - float(np.dot(a, v.download()))) <= 1e-12 * n
    assert abs(md[7] - float(np.dot(a, a))) <= 1e-12 * n


def test_norm_extremes(ctx):
    tiny = np.finfo(np.float64).tiny
    for mag in (1.0, np.finfo(np.float64).max / 8, tiny * 4, 5e-324 * 4):
        v = np.zeros(1000)
        v[0], v[999] = 3 * mag, -4 * mag
        assert abs(ctx.norm(ctx.upload(v)) - 5 * mag) <= 8 * EPS * 5 * mag
    assert ctx.norm(ctx.upload(np.zeros(10))) == 0.0
    assert np.isinf(ctx.norm(ctx.upload(np.full(4, np.finfo(np.float64).max))))
    assert np.isinf(ctx.norm(ctx.upload(np.array([np.inf, 1.0]))))
    assert np.isnan(ctx.norm(ctx.upload(np.array([np.inf, np.nan, 1.0]))))


# ---------------------------------------------------------- preconditioners
def assert_parity(got, ref, schedule, tol=1e-14):
    """Level sweeps reproduce the oracle bit for bit; chain sweeps accumulate
    backward entries in the order they become available (a few ulp)."""
    if schedule == 1:
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))


def gpu_apply(ctx, pc, r, dtype=np.float64):
    rd = ctx.upload(np.asarray(r, dtype))
    od = ctx.empty(rd.count, dtype)
    pc.apply(rd, od)
    return od.download()


def test_jacobi_bitwise_and_update(ctx, c0):
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    r = np.random.default_rng(10).uniform(-1, 1, s.nodes * 3)
    pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
    assert np.array_equal(gpu_apply(ctx, pc, r), orc.OPreconditioner(om, 1).apply(r))
    d2 = s.diagonal * 1.5
    m.update_values(d2, s.off_diagonal)
    pc.update_values()
    assert pc.numeric_updates() == 2
    assert np.array_equal(gpu_apply(ctx, pc, r), r * (1.0 / d2))


def test_jacobi_singular_diagonal_rejected(ctx):
    m = owtk.Matrix.csr(ctx, 2, 0, 1, [0, 1, 2], [0, 1], [1.0, 0.0])
    with pytest.raises(owtk.InvalidArgumentError):
        owtk.Preconditioner(m, owtk.PC_JACOBI)


@pytest.mark.parametrize("block", [1, 5, 1296])
@pytest.mark.parametrize("omega", [0.5, 1.0, 1.5])
def test_ssor_closed_form(ctx, block, omega):
    split, csr, rhs, expected = rc.ssor_review_case(block, omega)
    ms = owtk.Matrix.split(ctx, 2, 1, block, split["offsets"], split["columns"], split["diagonal"],
                           split["off_diagonal"])
    mc = owtk.Matrix.csr(ctx, 2, 1, block, csr["offsets"], csr["columns"], csr["values"])
    for m in (ms, mc):
        pc = owtk.Preconditioner(m, owtk.PC_SSOR, omega=omega)
        out = gpu_apply(ctx, pc, rhs)
        assert np.all(np.abs(out[:2 * block] / expected - 1) < 64 * EPS)
        assert np.all(out[2 * block:] == 0)  # output ghosts untouched (zero-initialised buffer)


@pytest.mark.parametrize("schedule", [1, 2])
@pytest.mark.parametrize("omega", [1.0, 1.3])
@pytest.mark.parametrize("block_nodes", [0, 512])
def test_ssor_c0_bitwise(ctx, c0, omega, block_nodes, schedule):
    s = c0
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    r = np.random.default_rng(11).uniform(-1, 1, s.nodes * 3)
    ref = orc.OPreconditioner(om, 2, omega=omega, block_nodes=block_nodes).apply(r)
    pc = owtk.Preconditioner(m, owtk.PC_SSOR, omega=omega, block_nodes=block_nodes, schedule=schedule)
    assert_parity(gpu_apply(ctx, pc, r), ref, schedule)


@pytest.mark.parametrize("schedule", [1, 2])
@pytest.mark.parametrize("fill", [0, 1, 2])
def test_ilu_audit_factors_bitwise(ctx, fill, schedule):
    n, comps, offsets, columns, values = rc.ilu_audit_matrix()
    for update in range(2):
        vals = values if not update else values * np.array([1.1 + 0.02 * (e % 3) for e in range(values.size)])
        om = orc.OMatrix.csr(n, 1, comps, offsets, columns, vals)
        opc = orc.OPreconditioner(om, 3, fill_level=fill)
        m = owtk.Matrix.csr(ctx, n, 1, comps, offsets, columns, vals)
        pc = owtk.Preconditioner(m, owtk.PC_ILU, fill_level=fill, schedule=schedule)
        for a, b in zip(pc.ilu_factors(), opc.ilu_factors()):
            assert np.array_equal(a, b)
        r = np.random.default_rng(12).uniform(-1, 1, (n + 1) * comps)
        assert_parity(gpu_apply(ctx, pc, r)[:n * comps], opc.apply(r), schedule)


def test_ilu_numeric_update_reuses_symbolic(ctx, c0):
    s = c0
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=256, schedule=1)
    d2, e2 = s.diagonal * 1.1, s.off_diagonal * 0.9
    m.update_values(d2, e2)
    pc.update_values()
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, d2, e2)
    r = np.random.default_rng(13).uniform(-1, 1, s.nodes * 3)
    assert np.array_equal(gpu_apply(ctx, pc, r), orc.OPreconditioner(om, 3, block_nodes=256).apply(r))
    assert pc.numeric_updates() == 2


def test_ilu_zero_pivot(ctx):
    m = owtk.Matrix.csr(ctx, 2, 0, 1, [0, 2, 4], [0, 1, 0, 1], [1.0, 1, 1, 1])
    with pytest.raises(owtk.ZeroPivotError):
        owtk.Preconditioner(m, owtk.PC_ILU)
    m2 = owtk.Matrix.csr(ctx, 2, 0, 1, [0, 2, 4], [0, 1, 0, 1], [2.0, 1, 1, 2])
    pc = owtk.Preconditioner(m2, owtk.PC_ILU)
    m2.update_values([1.0, 1, 1, 1])
    with pytest.raises(owtk.ZeroPivotError):
        pc.update_values()
    m2.update_values([2.0, 1, 1, 2])
    pc.update_values()
    assert np.allclose(gpu_apply(ctx, pc, [3.0, 3.0]), [1.0, 1.0], atol=1e-14)


@pytest.mark.parametrize("schedule", [1, 2])
def test_ilu_block_jacobi_medium_bitwise(ctx, schedule):
    """512 x 256 grid, Hilbert order, 4096-vertex block-Jacobi ILU(0)."""
    s = owtk.generate_split(512, 256, 3, 1, ordering=owtk.ORDER_HILBERT)
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=4096, schedule=schedule)
    opc = orc.OPreconditioner(om, 3, block_nodes=4096)
    r = np.random.default_rng(14).uniform(-1, 1, s.nodes * 3)
    assert_parity(gpu_apply(ctx, pc, r), opc.apply(r), schedule)
    info = pc.info()
    assert info.blocks == 32 and info.forward_levels > 1


def test_ilu_low_precision_storage(ctx, c0):
    s = c0
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    om = orc.OMatrix.split(s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    r = np.random.default_rng(15).uniform(-1, 1, s.nodes * 3)
    ref = orc.OPreconditioner(om, 3, block_nodes=512).apply(r)
    for storage, tol in ((owtk.F32, 1e-5), (owtk.BF16, 5e-2)):
        pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=512, storage=storage, schedule=1)
        out = gpu_apply(ctx, pc, r.astype(np.float32), np.float32)
        assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("storage", [owtk.F32, owtk.BF16])
def test_chain_and_level_agree_low_precision(ctx, c0, storage):
    s = c0
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    r = np.random.default_rng(16).uniform(-1, 1, s.nodes * 3).astype(np.float32)
    for kind in (owtk.PC_ILU, owtk.PC_SSOR):
        a = owtk.Preconditioner(m, kind, block_nodes=512, storage=storage, schedule=1)
        b = owtk.Preconditioner(m, kind, block_nodes=512, storage=storage, schedule=2)
        assert_parity(gpu_apply(ctx, b, r, np.float32), gpu_apply(ctx, a, r, np.float32), 2, tol=1e-6)


@pytest.mark.parametrize("bs", [2, 3, 4])
@pytest.mark.parametrize("omega", [1.0, 1.3])
def test_dense_ssor_matches_oracle(ctx, bs, omega):
    n, off, col, val, rhs = owtk.generate_bsr(48, 32, bs, 1, ordering=owtk.ORDER_HILBERT)
    om = orc.OMatrix.bsr(n, 0, bs, off, col, val)
    ref = orc.OPreconditioner(om, orc.OPreconditioner.SSOR, omega=omega).apply(rhs)
    m = owtk.Matrix.bsr(ctx, n, 0, bs, off, col, val)
    m.add_storage(owtk.F32)
    m.add_storage(owtk.BF16)
    pc = owtk.Preconditioner(m, owtk.PC_SSOR, omega=omega)
    x, y = ctx.upload(rhs), ctx.empty(rhs.size)
    pc.apply(x, y)
    got = y.download()
    err64 = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"bs {bs} omega {omega}: FP64 max relative difference {err64:.2e}, bitwise {np.array_equal(got, ref)}")
    assert err64 <= 1e-13
    for storage, tol in ((owtk.F32, 1e-5), (owtk.BF16, 5e-2)):
        pcl = owtk.Preconditioner(m, owtk.PC_SSOR, omega=omega, storage=storage)
        xl, yl = ctx.upload(rhs.astype(np.float32)), ctx.empty(rhs.size, np.float32)
        pcl.apply(xl, yl)
        err = np.max(np.abs(yl.download() - ref)) / np.max(np.abs(ref))
        print(f"  storage {storage}: {err:.2e}")
        assert err <= tol


def test_dense_ssor_numeric_update(ctx):
    """The multicolour SSOR keeps colour-ordered copies of the blocks: after
    a matrix value change and update_values the apply matches the oracle on
    the new values, for FP64 and BF16 storage."""
    bs = 4
    n, off, col, val, rhs = owtk.generate_bsr(40, 24, bs, 2, ordering=owtk.ORDER_HILBERT)
    m = owtk.Matrix.bsr(ctx, n, 0, bs, off, col, val)
    m.add_storage(owtk.BF16)
    pc = owtk.Preconditioner(m, owtk.PC_SSOR, omega=1.2)
    pcl = owtk.Preconditioner(m, owtk.PC_SSOR, omega=1.2, storage=owtk.BF16)
    val2 = np.asarray(val, float) * (1.0 + 0.25 * np.sin(np.arange(len(val))))
    m.update_values(val2)
    pc.update_values()
    pcl.update_values()
    om = orc.OMatrix.bsr(n, 0, bs, off, col, val2)
    ref = orc.OPreconditioner(om, orc.OPreconditioner.SSOR, omega=1.2).apply(rhs)
    x, y = ctx.upload(rhs), ctx.empty(rhs.size)
    pc.apply(x, y)
    err = np.max(np.abs(y.download() - ref)) / np.max(np.abs(ref))
    xl, yl = ctx.upload(rhs.astype(np.float32)), ctx.empty(rhs.size, np.float32)
    pcl.apply(xl, yl)
    errl = np.max(np.abs(yl.download() - ref)) / np.max(np.abs(ref))
    print(f"after update: FP64 {err:.2e}, BF16 {errl:.2e}")
    assert err <= 1e-13 and errl <= 5e-2


def test_ilu_factor_file_matches_oracle(ctx, tmp_path):
    """System read from a reference-format file, ILU(0) factored on the GPU and
    written in the ilu_fill.cpp output format; the file's factors equal the
    oracle's (IluLevelPreconditioner restated)."""
    s = owtk.generate_split(40, 24, 3, 5, ordering=owtk.ORDER_RCM)
    o, c, v = s.csr()
    sysfile = tmp_path / "system.bin"
    owtk.write_block_csr(sysfile, s.nodes, 3, o, c, v)
    n, b, o2, c2, v2 = owtk.read_block_csr(sysfile)
    m = owtk.Matrix.csr(ctx, n, 0, b, o2, c2, v2)
    pc = owtk.Preconditioner(m, owtk.PC_ILU)
    factfile = tmp_path / "factors.bin"
    pc.write_factors(factfile)
    raw = factfile.read_bytes()
    hn, hb, nnz = np.frombuffer(raw[:24], np.uint64).tolist()
    pos = 24
    off = np.frombuffer(raw[pos:pos + 8 * (hn + 1)], np.uint64); pos += 8 * (hn + 1)
    col = np.frombuffer(raw[pos:pos + 8 * nnz], np.uint64); pos += 8 * nnz
    val = np.frombuffer(raw[pos:pos + 8 * nnz * hb], np.float64); pos += 8 * nnz * hb
    inv = np.frombuffer(raw[pos:], np.float64)
    om = orc.OMatrix.csr(n, 0, b, o2, c2, v2)
    ref = orc.OPreconditioner(om, orc.OPreconditioner.ILU)
    ro, rc_, rv, ri = ref.ilu_factors()
    assert (hn, hb) == (n, b)
    assert np.array_equal(off, ro) and np.array_equal(col, rc_)
    assert np.max(np.abs(val - rv)) <= 1e-13 * np.max(np.abs(rv))
    assert np.max(np.abs(inv - ri)) <= 1e-13 * np.max(np.abs(ri))


def test_destroy_order_independent(ctx, c0):
    """A garbage collector finalizes objects of a reference cycle in no fixed
    order: destroying a matrix before its preconditioner or multigrid
    hierarchy must not touch the freed matrix."""
    s = c0
    m = owtk.Matrix.split(ctx, s.nodes, 0, 3, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pcs = [owtk.Preconditioner(m, kind, block_nodes=512) for kind in (owtk.PC_JACOBI, owtk.PC_SSOR, owtk.PC_ILU)]
    co, cc, cv = s.csr()
    mc = owtk.Matrix.csr(ctx, s.nodes, 0, 3, co, cc, cv)
    mg = owtk.Multigrid(mc)
    m.close()
    mc.close()
    for pc in pcs:
        pc.close()
    mg.close()
    ctx.synchronize()
