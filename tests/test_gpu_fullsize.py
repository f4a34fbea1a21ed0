"""Solves at the configurations' full sizes on the GPU, checked by
size-independent properties: the true relative residual recomputed on the
host in FP64 from the FP64 matrix (scipy, independent of the library), the
error against an FP64 solve to 1e-12, convergence of every batched system,
and the iteration counts printed for the record.

Sizes: c1 2048x1024 (2,097,152 vertices, block 3, split view, Hilbert), c2
1024x1024 dense 4x4 BSR (1,048,576 vertices), c4 512 uneven systems
(32x16 .. 512x256), and a 4096x2048 grid (8,388,608 vertices) beyond the
4,194,304 rows one reduction pass of 16,384 blocks covers.
"""
import time

import numpy as np
import pytest
import scipy.sparse as sp

import owtk

pytestmark = pytest.mark.gpu

BS = 3


def host_operator(s):
    """FP64 block operator of a split system as one scipy matrix over the
    interleaved [node][component] unknowns."""
    o, c, v = s.csr()
    o, c, v = np.asarray(o, np.int64), np.asarray(c, np.int64), np.asarray(v, np.float64)
    n = s.nodes
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(o))
    r_idx = (rows[:, None] * BS + np.arange(BS)[None, :]).ravel()
    c_idx = (c[:, None] * BS + np.arange(BS)[None, :]).ravel()
    return sp.csr_matrix((v.reshape(-1, BS).ravel(), (r_idx, c_idx)), shape=(n * BS, n * BS))


def host_bsr_operator(n, off, col, val, bs):
    return sp.bsr_matrix((np.asarray(val, np.float64).reshape(-1, bs, bs), np.asarray(col, np.int64),
                          np.asarray(off, np.int64)), shape=(n * bs, n * bs)).tocsr()


def true_rel(a, b, x):
    return float(np.linalg.norm(b - a @ x) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def c1():
    t = time.perf_counter()
    s = owtk.generate_split(2048, 1024, BS, 0, ordering=owtk.ORDER_HILBERT)
    a = host_operator(s)
    print(f"c1 generated in {time.perf_counter() - t:.1f}s: {s.nodes} vertices", flush=True)
    return s, a


@pytest.fixture(scope="module")
def c1_reference(ctx, c1):
    """FP64 BiCGSTAB + 4096-block ILU(0) to 1e-12: the FP64 solution the
    other c1 results are measured against."""
    s, a = c1
    m = owtk.Matrix.split(ctx, s.nodes, 0, BS, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=4096)
    b, x = ctx.upload(s.rhs), ctx.empty(s.nodes * BS)
    res = owtk.bicgstab(m, pc, b, x, owtk.SolverOptions(relative_tolerance=1e-12, maximum_iterations=2000))
    xr = x.download()
    print(f"c1 FP64 reference to 1e-12: {res['status_name']} after {res['iterations']} iterations, "
          f"host true rel {true_rel(a, s.rhs, xr):.2e}", flush=True)
    assert res["status_name"] == "converged"
    return xr


def test_c1_bicgstab_ilu0_full_size(ctx, c1, c1_reference):
    """c1: FP64 BiCGSTAB + block-Jacobi ILU(0), 4096-vertex blocks, rtol 1e-8."""
    s, a = c1
    m = owtk.Matrix.split(ctx, s.nodes, 0, BS, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=4096)
    b, x = ctx.upload(s.rhs), ctx.empty(s.nodes * BS)
    res = owtk.bicgstab(m, pc, b, x, owtk.SolverOptions(relative_tolerance=1e-8))
    xs = x.download()
    rel = true_rel(a, s.rhs, xs)
    err = np.linalg.norm(xs - c1_reference) / np.linalg.norm(c1_reference)
    print(f"c1 BiCGSTAB+ILU(0): {res['status_name']} iterations {res['iterations']} "
          f"(applies {res['operator_applications']}/{res['preconditioner_applications']}), "
          f"host true rel {rel:.2e}, error vs FP64 1e-12 solve {err:.2e}")
    assert res["status_name"] == "converged"
    assert rel <= 1e-8
    assert abs(rel - res["relative_residual_norm"]) <= 1e-3 * rel


@pytest.mark.parametrize("storage", [owtk.F32, owtk.BF16])
def test_c1_fgmres_refined_full_size(ctx, c1, c1_reference, storage):
    """c1: FGMRES(30) with FP32 (or BF16) matrix and ILU(0) storage inside
    FP64 iterative refinement, rtol 1e-8, error against the FP64 solution."""
    s, a = c1
    m = owtk.Matrix.split(ctx, s.nodes, 0, BS, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m.add_storage(owtk.F32)
    if storage == owtk.BF16:
        m.add_storage(owtk.BF16)
    pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=4096, storage=storage)
    b, x = ctx.upload(s.rhs), ctx.empty(s.nodes * BS)
    ropts = owtk.refinement_options(inner_storage=owtk.F32, inner_restart=30, relative_tolerance=1e-8)
    res = owtk.fgmres_refined(m, pc, b, x, ropts)
    xs = x.download()
    rel = true_rel(a, s.rhs, xs)
    err = np.linalg.norm(xs - c1_reference) / np.linalg.norm(c1_reference)
    print(f"c1 FGMRES(30)-IR, {'FP32' if storage == owtk.F32 else 'BF16'} ILU(0): {res['status_name']}, "
          f"{res['refinement_steps']} refinements, {res['inner_iterations']} inner iterations, "
          f"host true rel {rel:.2e}, error vs FP64 {err:.2e}")
    assert res["status_name"] == "converged"
    assert rel <= 1e-8
    assert err <= 1e-6


def test_c2_bsr_bf16_fgmres_refined_full_size(ctx):
    """c2: 1024x1024 dense 4x4 BSR, BF16 matrix and multicolour block SSOR
    storage, FP32 accumulation, FP64-refined FGMRES(40), rtol 1e-10."""
    t = time.perf_counter()
    n, off, col, val, rhs = owtk.generate_bsr(1024, 1024, 4, 0, ordering=owtk.ORDER_HILBERT)
    a = host_bsr_operator(n, off, col, val, 4)
    print(f"c2 generated in {time.perf_counter() - t:.1f}s: {n} vertices, {off[-1]} blocks", flush=True)
    m = owtk.Matrix.bsr(ctx, n, 0, 4, off, col, val)
    m.add_storage(owtk.BF16)
    b, x = ctx.upload(rhs), ctx.empty(rhs.size)
    pc64 = owtk.Preconditioner(m, owtk.PC_SSOR)
    ref = owtk.fgmres(m, pc64, b, x, owtk.SolverOptions(relative_tolerance=1e-12, restart=40,
                                                       maximum_iterations=5000))
    x64 = x.download()
    assert ref["status_name"] == "converged"
    x.zero()
    pc16 = owtk.Preconditioner(m, owtk.PC_SSOR, storage=owtk.BF16)
    ropts = owtk.refinement_options(inner_storage=owtk.BF16, inner_restart=40, inner_relative_tolerance=1e-4,
                                    relative_tolerance=1e-10)
    res = owtk.fgmres_refined(m, pc16, b, x, ropts)
    xs = x.download()
    rel = true_rel(a, rhs, xs)
    err = np.linalg.norm(xs - x64) / np.linalg.norm(x64)
    print(f"c2 BF16 FGMRES(40)-IR + block SSOR: {res['status_name']}, {res['refinement_steps']} refinements, "
          f"{res['inner_iterations']} inner iterations, host true rel {rel:.2e}, error vs FP64 {err:.2e} "
          f"(FP64 reference {ref['iterations']} iterations, host true rel {true_rel(a, rhs, x64):.2e})")
    assert res["status_name"] == "converged"
    assert rel <= 1e-10
    assert err <= 1e-8


def test_c4_all_512_systems_converge(ctx):
    """c4: 512 uneven systems, batched BiCGSTAB + Jacobi to rtol 1e-8; every
    system converged and its host-recomputed true residual within 1e-8."""
    from owtk import workloads

    t = time.perf_counter()
    systems, rhs, grids = workloads.batched_systems(512)
    print(f"c4 generated in {time.perf_counter() - t:.1f}s: {sum(s[0] for s in systems)} vertices", flush=True)
    bt = owtk.Batch(ctx, BS, systems)
    b, x = ctx.upload(rhs), ctx.empty(rhs.size)
    tile, _ = bt.autotune(b, iterations=4)
    x.zero()
    res = bt.bicgstab(b, x, owtk.SolverOptions(relative_tolerance=1e-8, maximum_iterations=20000))
    xs = x.download()
    off, worst, its = 0, 0.0, []
    for (n, o, c, v), r in zip(systems, res):
        seg = slice(off, off + n * BS)
        o, c, v = np.asarray(o, np.int64), np.asarray(c, np.int64), np.asarray(v, np.float64)
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(o))
        r_idx = (rows[:, None] * BS + np.arange(BS)[None, :]).ravel()
        c_idx = (c[:, None] * BS + np.arange(BS)[None, :]).ravel()
        a = sp.csr_matrix((v.reshape(-1, BS).ravel(), (r_idx, c_idx)), shape=(n * BS, n * BS))
        rel = true_rel(a, rhs[seg], xs[seg])
        assert r["status_name"] == "converged", r
        assert rel <= 1e-8, (n, rel)
        worst = max(worst, rel)
        its.append(r["iterations"])
        off += n * BS
    print(f"c4: 512 of 512 converged (tile {tile}), worst host true rel {worst:.2e}, iterations "
          f"median {np.median(its):.0f} max {max(its)}")


def test_grid_above_one_pass_reduction_limit(ctx):
    """4096x2048 grid (8,388,608 vertices): BiCGSTAB + Jacobi, whose fused
    SpMV reductions span 32,768 blocks."""
    t = time.perf_counter()
    s = owtk.generate_split(4096, 2048, BS, 0, ordering=owtk.ORDER_HILBERT)
    a = host_operator(s)
    print(f"4096x2048 generated in {time.perf_counter() - t:.1f}s: {s.nodes} vertices", flush=True)
    m = owtk.Matrix.split(ctx, s.nodes, 0, BS, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
    b, x = ctx.upload(s.rhs), ctx.empty(s.nodes * BS)
    res = owtk.bicgstab(m, pc, b, x, owtk.SolverOptions(relative_tolerance=1e-8))
    rel = true_rel(a, s.rhs, x.download())
    print(f"4096x2048 BiCGSTAB+Jacobi: {res['status_name']} iterations {res['iterations']}, host true rel {rel:.2e}")
    assert res["status_name"] == "converged"
    assert rel <= 1e-8
