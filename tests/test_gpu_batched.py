"""Batched BiCGSTAB + Jacobi on the GPU against the single-system solver and
the CPU oracle (the reference's bicgstab restated)."""
import numpy as np
import pytest

import owtk
from owtk import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch(ctx):
    systems, rhs, grids = workloads.batched_systems(count=12, nx_min=8, nx_max=64, seed0=7000)
    return systems, rhs, grids


def oracle_envelope(system, rhs, tol=1e-8):
    """Iteration counts of the oracle BiCGSTAB + Jacobi (the reference's
    recurrence) under perturbations of b at the 1e-15 level.  High-contrast
    systems move by several iterations under such perturbations, so a
    different (equally valid) reduction order may too."""
    n, o, c, v = system
    om = orc.OMatrix.csr(n, 0, 3, o, c, v)
    pc = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
    its = []
    for eps in (0.0, 1e-15, -1e-15, 3e-15, -3e-15, 1e-14, -1e-14, 3e-14):
        _, r = orc.bicgstab(om, pc, rhs * (1 + eps * np.sin(np.arange(rhs.size))), relative_tolerance=tol)
        its.append(r["iterations"])
    return min(its), max(its)


def single_solve(ctx, system, rhs, opts):
    n, o, c, v = system
    m = owtk.Matrix.csr(ctx, n, 0, 3, o, c, v)
    pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
    b = ctx.upload(rhs)
    x = ctx.empty(rhs.size)
    res = owtk.bicgstab(m, pc, b, x, opts)
    return res, x.download()


@pytest.mark.parametrize("tile", [64, 256, 512])
def test_batched_matches_single_system(ctx, small_batch, tile):
    systems, rhs, grids = small_batch
    bt = owtk.Batch(ctx, 3, systems)
    bt.set_tile(tile)
    b = ctx.upload(rhs)
    x = ctx.empty(rhs.size)
    opts = owtk.SolverOptions(relative_tolerance=1e-8)
    results = bt.bicgstab(b, x, opts, graph_iterations=4)
    xb = x.download()
    off = 0
    singles = []
    for k, (sysk, res) in enumerate(zip(systems, results)):
        n = sysk[0] * 3
        single, xs = single_solve(ctx, sysk, rhs[off:off + n], opts)
        singles.append(single["iterations"])
        assert res["status_name"] == "converged", (k, res)
        assert res["relative_residual_norm"] <= 1e-8
        # same recurrence and cadence; only the reduction order differs
        lo, hi = oracle_envelope(sysk, rhs[off:off + n])
        assert lo - 3 <= res["iterations"] <= hi + 3, (k, res["iterations"], single["iterations"], lo, hi)
        if grids[k][2] < 0.6:  # low contrast: counts are stable under rounding
            assert abs(res["iterations"] - single["iterations"]) <= 1, (k, res["iterations"], single["iterations"])
        err = np.linalg.norm(xb[off:off + n] - xs) / np.linalg.norm(xs)
        assert err < 1e-6, (k, err)
        off += n
    print(f"tile {tile}: iterations", [r["iterations"] for r in results])
    print("single-system iterations", singles)


def test_batched_matches_oracle_iterations(ctx, small_batch):
    systems, rhs, grids = small_batch
    bt = owtk.Batch(ctx, 3, systems[:4])
    n_all = sum(s[0] for s in systems[:4]) * 3
    b = ctx.upload(rhs[:n_all])
    x = ctx.empty(n_all)
    results = bt.bicgstab(b, x, owtk.SolverOptions(relative_tolerance=1e-8))
    off = 0
    for sysk, res in zip(systems[:4], results):
        n, o, c, v = sysk
        om = orc.OMatrix.csr(n, 0, 3, o, c, v)
        pc = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
        xo, ro = orc.bicgstab(om, pc, rhs[off:off + n * 3], relative_tolerance=1e-8)
        lo, hi = oracle_envelope(sysk, rhs[off:off + n * 3])
        print(res["iterations"], ro["iterations"], (lo, hi))
        assert ro["status"] == 0 and res["status_name"] == "converged"
        assert lo - 3 <= res["iterations"] <= hi + 3
        off += n * 3


def test_masks_and_zero_rhs(ctx, small_batch):
    systems, rhs, grids = small_batch
    rhs = rhs.copy()
    n0 = systems[0][0] * 3
    rhs[:n0] = 0.0  # system 0: b = 0, x0 = 0 -> converged at iteration 0
    bt = owtk.Batch(ctx, 3, systems)
    b = ctx.upload(rhs)
    x = ctx.empty(rhs.size)
    results = bt.bicgstab(b, x, owtk.SolverOptions(relative_tolerance=1e-8, maximum_iterations=5),
                          graph_iterations=2)
    assert results[0]["status_name"] == "converged" and results[0]["iterations"] == 0
    assert np.all(x.download()[:n0] == 0.0)
    # an iteration cap of 5 stops the others with maximum_iterations after exactly 5
    for r in results[1:]:
        assert r["status_name"] in ("maximum_iterations", "converged")
        if r["status_name"] == "maximum_iterations":
            assert r["iterations"] == 5


def test_autotune_picks_a_valid_tile(ctx, small_batch):
    systems, rhs, grids = small_batch
    bt = owtk.Batch(ctx, 3, systems)
    b = ctx.upload(rhs)
    tile, ms = bt.autotune(b, iterations=4)
    assert tile in (64, 128, 256, 512) and ms > 0
    x = ctx.empty(rhs.size)
    results = bt.bicgstab(b, x, owtk.SolverOptions(relative_tolerance=1e-8))
    assert all(r["status_name"] == "converged" for r in results)


@pytest.mark.parametrize("threads", [256, 512, 1024])
def test_group_solves_match_oracle(ctx, small_batch, threads):
    """Every system solved whole by one workgroup (limit above the largest
    system): same recurrence and stages as the graph path, so each count stays
    inside the oracle's rounding envelope and the solutions match the
    single-system solver's."""
    systems, rhs, grids = small_batch
    bt = owtk.Batch(ctx, 3, systems)
    bt.set_group_limit(1 << 20, threads)
    b = ctx.upload(rhs)
    x = ctx.empty(rhs.size)
    opts = owtk.SolverOptions(relative_tolerance=1e-8)
    results = bt.bicgstab(b, x, opts, graph_iterations=4)
    xb = x.download()
    off = 0
    for k, (sysk, res) in enumerate(zip(systems, results)):
        n = sysk[0] * 3
        assert res["status_name"] == "converged", (k, res)
        assert res["relative_residual_norm"] <= 1e-8
        assert res["stop_ms"] > 0.0
        lo, hi = oracle_envelope(sysk, rhs[off:off + n])
        assert lo - 3 <= res["iterations"] <= hi + 3, (k, res["iterations"], lo, hi)
        single, xs = single_solve(ctx, sysk, rhs[off:off + n], opts)
        if grids[k][2] < 0.6:
            assert abs(res["iterations"] - single["iterations"]) <= 1, (k, res["iterations"], single["iterations"])
        assert np.linalg.norm(xb[off:off + n] - xs) / np.linalg.norm(xs) < 1e-6, k
        # the true residual, recomputed on the host
        nn, o, c, v = sysk
        om = orc.OMatrix.csr(nn, 0, 3, o, c, v)
        r = rhs[off:off + n] - om.spmv(xb[off:off + n])[:n]
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(rhs[off:off + n]), k
        off += n
    print(f"group {threads}: iterations", [r["iterations"] for r in results])


def test_split_between_groups_and_graph(ctx, small_batch):
    """Half the systems on workgroups, half on the graph replays, at the same
    time: every system converges, and each path gives the iteration counts it
    gives alone."""
    systems, rhs, grids = small_batch
    sizes = sorted(s[0] for s in systems)
    limit = sizes[len(sizes) // 2]
    opts = owtk.SolverOptions(relative_tolerance=1e-8)
    b = ctx.upload(rhs)
    runs = {}
    for name, lim in (("graph", 0), ("groups", 1 << 20), ("split", limit)):
        bt = owtk.Batch(ctx, 3, systems)
        bt.set_group_limit(lim, 512)
        x = ctx.empty(rhs.size)
        runs[name] = (bt.bicgstab(b, x, opts, graph_iterations=4), x.download())
    for k, sysk in enumerate(systems):
        res, _ = runs["split"]
        ref, _ = runs["groups" if sysk[0] <= limit else "graph"]
        assert res[k]["status_name"] == "converged", (k, res[k])
        assert res[k]["iterations"] == ref[k]["iterations"], (k, res[k]["iterations"], ref[k]["iterations"])
        assert res[k]["true_residual_norm"] == ref[k]["true_residual_norm"], k
    assert np.array_equal(runs["split"][1], np.where(
        np.repeat([s[0] <= limit for s in systems], [s[0] * 3 for s in systems]), runs["groups"][1], runs["graph"][1]))


def test_group_masks_zero_rhs_and_iteration_cap(ctx, small_batch):
    """The stages of the graph path on a workgroup: b = 0 stops at iteration 0
    with x untouched; a cap of 5 stops the rest after exactly 5."""
    systems, rhs, grids = small_batch
    rhs = rhs.copy()
    n0 = systems[0][0] * 3
    rhs[:n0] = 0.0
    bt = owtk.Batch(ctx, 3, systems)
    bt.set_group_limit(1 << 20, 256)
    b = ctx.upload(rhs)
    x = ctx.empty(rhs.size)
    results = bt.bicgstab(b, x, owtk.SolverOptions(relative_tolerance=1e-8, maximum_iterations=5))
    assert results[0]["status_name"] == "converged" and results[0]["iterations"] == 0
    assert np.all(x.download()[:n0] == 0.0)
    for r in results[1:]:
        assert r["status_name"] in ("maximum_iterations", "converged")
        if r["status_name"] == "maximum_iterations":
            assert r["iterations"] == 5


def test_group_autotune(ctx, small_batch):
    systems, rhs, grids = small_batch
    bt = owtk.Batch(ctx, 3, systems)
    b = ctx.upload(rhs)
    opts = owtk.SolverOptions(relative_tolerance=1e-8)
    limit, threads, ms = bt.autotune_groups(b, opts, graph_iterations=4)
    assert limit == 0 or limit >= 1024
    assert threads in (256, 512, 1024) and ms > 0
    x = ctx.empty(rhs.size)
    results = bt.bicgstab(b, x, opts)
    assert all(r["status_name"] == "converged" for r in results)
    print("group limit", limit, "threads", threads, f"{ms:.2f} ms")


@pytest.mark.parametrize("block", [1, 2, 4])
def test_group_and_graph_paths_other_block_sizes(ctx, block):
    """Block sizes 1, 2 and 4: the one-workgroup solves and the graph path
    both converge every system to 1e-8 (true residual recomputed on the host)
    and agree on the solutions."""
    systems, rhs, grids = workloads.batched_systems(count=6, nx_min=8, nx_max=48, block=block, seed0=7100)
    b = ctx.upload(rhs)
    opts = owtk.SolverOptions(relative_tolerance=1e-8)
    xs = {}
    for name, lim in (("graph", 0), ("groups", 1 << 20)):
        bt = owtk.Batch(ctx, block, systems)
        bt.set_group_limit(lim, 256)
        x = ctx.empty(rhs.size)
        res = bt.bicgstab(b, x, opts, graph_iterations=4)
        assert all(r["status_name"] == "converged" for r in res), (name, [r["status_name"] for r in res])
        xs[name] = x.download()
        off = 0
        for (n, o, c, v) in systems:
            om = orc.OMatrix.csr(n, 0, block, o, c, v)
            r = rhs[off:off + n * block] - om.spmv(xs[name][off:off + n * block])[:n * block]
            assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(rhs[off:off + n * block]), (name, n)
            off += n * block
    assert np.linalg.norm(xs["groups"] - xs["graph"]) <= 1e-6 * np.linalg.norm(xs["graph"])
