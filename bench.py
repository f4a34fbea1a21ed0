"""Headline benchmark: time to true relative residual 1e-8 on the 2M-vertex
block-3 grid, learned-preconditioner + mixed-precision FGMRES vs plain FP64
BiCGSTAB + block-Jacobi ILU(0), plus the split SpMV's HBM bandwidth.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--storage f32|bf16]
    python bench.py --kernel spmv|copy|learned|bsr|learned-update --steps K   (kernel loops for rocprofv3)
    python bench.py --workload batched                   (512 uneven systems, batched BiCGSTAB)
    python bench.py --workload batched [--batch-max-vertices N] [--group-limit L]
                                                         (only the systems of <= N vertices; systems of
                                                          <= L vertices on one workgroup each, default autotuned)
    python bench.py --workload bsr                       (dense 4x4 BSR, BF16 FGMRES(40)-IR + SSOR)

One "step" is one full solve of the headline method from x0 = 0 on the
resident 2048 x 1024 system; the FP64 baseline solve is timed the same way in
the same run.  Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import platform
import subprocess
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "owt-krylov_amd"))

import numpy as np  # noqa: E402

import owtk  # noqa: E402

METRIC = "ms to true rel. residual 1e-8: learned+mixed vs plain FP64; SpMV HBM GB/s"
PEAK_HBM_GBPS = 8000.0
NX, NY, BS, SEED = 2048, 1024, 3, 0
BLOCK_NODES = 4096
BATCH_SYSTEMS = 512
BSR_GRID = 1024
BATCH_GRAPH_ITERATIONS = 8
RTOL = 1e-8
MODEL = ROOT / "owt-krylov_amd" / "owtk" / "models" / "gnn_uniform.npz"
# HBM traffic per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on the
# `--kernel` loops (FETCH_SIZE scaled by 2: it reads half of a plain copy
# kernel's known input; WRITE_SIZE by 1), on the 2048x1024 grid (2,097,152
# vertices, block 3).
TRAFFIC_BYTES = {("spmv", "f64"): 513_325_466, ("learned", "bf16"): 261_427_405}
KERNEL_NAMES = {
    "learned_bf16": "learned series step, BF16 column-scaled off-diagonal values: SpMV + fused update "
                    "(spmv_sell_cd<3,bf16,float,false,EpiSeries>)",
    "spmv_f64": "split SpMV, FP64 (spmv_sell_cd<3,double,double,true,EpiStore>)",
    "spmv_f32": "split SpMV, FP32 storage (spmv_sell_cd<3,float,float,...>)",
    "spmv_bf16": "split SpMV, BF16 storage (spmv_sell_cd<3,bf16,float,...>)",
    "precond_f64": "block-Jacobi ILU(0) triangular sweeps, FP64 (k_tri_ell)",
    "precond_low": "block-Jacobi ILU(0) triangular sweeps, low precision (k_tri_ell)",
    "learned": "learned series step, FP32 column-scaled off-diagonal values: SpMV + fused update "
               "(spmv_sell_cd<3,float,float,false,EpiSeries>)",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return model, os.cpu_count()


def build_system(ctx):
    t0 = time.perf_counter()
    s = owtk.generate_split(NX, NY, BS, SEED, ordering=owtk.ORDER_HILBERT)
    m = owtk.Matrix.split(ctx, s.nodes, 0, BS, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    m.add_storage(owtk.F32)
    m.add_storage(owtk.BF16)
    log(f"[bench] system {s.nodes} vertices, {s.offsets[-1]} edges, set-up {time.perf_counter() - t0:.1f}s")
    return s, m


def true_rel(ctx, m, b, x, r, bnorm):
    m.residual(b, x, r)
    return ctx.norm(r, m.owned_size) / bnorm


def timed_solves(ctx, label, steps, warmup, run, x):
    """Warm-up solves, `steps` timed solves (no kernel timing events in the
    timed region), then one extra solve with per-launch HIP-event timing for
    the per-kernel breakdown (not part of the timing)."""
    for _ in range(warmup):
        x.zero()
        run()
    ctx.synchronize()
    times, results = [], []
    for k in range(steps):
        x.zero()
        ctx.synchronize()
        t0 = time.perf_counter()
        res = run()
        ctx.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        results.append(res)
        log(f"[bench] {label} step {k + 1}/{steps}: {times[-1]:.2f} ms, it={res['iterations']}, "
            f"status={res['status_name']}, fallback steps={res.get('fallback_refinements', 0)}")
    x.zero()
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
    run()
    ctx.synchronize()
    ctx.profile(False)
    prof = {name: ctx.profile_read(cat) for name, cat in (
        ("spmv_f64", ctx.PROF_SPMV_F64), ("spmv_f32", ctx.PROF_SPMV_F32), ("spmv_bf16", ctx.PROF_SPMV_BF16),
        ("precond_f64", ctx.PROF_PRECOND_F64), ("precond_low", ctx.PROF_PRECOND_LOW), ("learned", ctx.PROF_LEARNED),
        ("learned_ends", ctx.PROF_LEARNED_ENDS))}
    return times, results, prof


def spmv_timing(ctx, m, n, reps=50):
    """Back-to-back split SpMVs on the context stream, timed by one HIP event
    pair around the launches: average time per launch (kernel plus the
    launch gap) and algorithmic GB/s per storage type."""
    x64 = ctx.upload(np.random.default_rng(1).uniform(-1, 1, n))
    y64 = ctx.empty(n)
    x32 = ctx.upload(np.random.default_rng(1).uniform(-1, 1, n).astype(np.float32))
    y32 = ctx.empty(n, np.float32)
    out = {}
    for name, dtype, cat, x, y in (("f64", owtk.F64, ctx.PROF_SPMV_F64, x64, y64),
                                   ("f32", owtk.F32, ctx.PROF_SPMV_F32, x32, y32),
                                   ("bf16", owtk.BF16, ctx.PROF_SPMV_BF16, x32, y32)):
        for _ in range(5):
            m.spmv(x, y, dtype)
        ctx.synchronize()
        # algorithmic bytes of one launch (from the library's own accounting)
        ctx.profile_reset()
        ctx.profile(True)
        m.spmv(x, y, dtype)
        ctx.profile(False)
        nbytes = ctx.profile_read(cat)["bytes"]
        ctx.synchronize()
        ctx.region_begin()
        for _ in range(reps):
            m.spmv(x, y, dtype)
        avg = ctx.region_end() / reps
        out[name] = {"avg_ms": avg, "gbps": nbytes / (avg * 1e-3) / 1e9, "bytes": nbytes, "launches": reps}
    return out


def stream_ceiling(ctx, reps=20):
    """The on-box streaming rate of a plain vector kernel, beside the 8 TB/s
    spec: owtk_axpy (y += a x) over two 512 MiB FP64 vectors (1 GiB read,
    512 MiB written per launch; the working set exceeds the 256 MB MALL),
    one HIP event pair around back-to-back launches."""
    count = 64 << 20
    x = ctx.upload(np.full(count, 1e-3))
    y = ctx.upload(np.zeros(count))
    for _ in range(3):
        ctx.axpy(1.0, x, y)
    ctx.synchronize()
    ctx.region_begin()
    for _ in range(reps):
        ctx.axpy(1.0, x, y)
    ms = ctx.region_end() / reps
    nbytes = 3 * count * 8
    gbps = nbytes / (ms * 1e-3) / 1e9
    del x, y
    return {"kernel": "owtk_axpy, y += a x over 2 x 512 MiB FP64", "bytes_per_launch": nbytes,
            "avg_ms": round(ms, 4), "gbps": round(gbps, 1), "frac_of_peak": round(gbps / PEAK_HBM_GBPS, 4),
            "note": "measured streaming rate on this box; roofline fractions use the 8.0 TB/s spec"}


def learned_timing(ctx, pc, n, reps=30):
    """The learned preconditioner applied back to back outside a solver (no
    stopped launches): per-launch HIP-event time of its middle series-step
    kernel, and the whole apply (first, middle and last steps) timed by one
    event pair."""
    xf = ctx.upload(np.random.default_rng(2).uniform(-1, 1, n).astype(np.float32))
    yf = ctx.empty(n, np.float32)
    for _ in range(3):
        pc.apply(xf, yf)
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(reps):
        pc.apply(xf, yf)
    ctx.profile(False)
    p = ctx.profile_read(ctx.PROF_LEARNED)
    ctx.region_begin()
    for _ in range(reps):
        pc.apply(xf, yf)
    apply_ms = ctx.region_end() / reps
    return {"step_avg_ms": p["avg_ms"], "step_gbps": p["gbps"], "step_bytes": p["bytes"] // max(p["launches"], 1),
            "step_launches": p["launches"], "apply_ms": apply_ms,
            "apply_gbps": pc.info().apply_bytes / (apply_ms * 1e-3) / 1e9, "apply_bytes": pc.info().apply_bytes}


def long_double_true_rel(s, x):
    """|b - A x| / |b| in long double from the FP64 split matrix (an
    independent check of a CPU solve, benchmark_split.cpp:90-99)."""
    ld = np.longdouble
    n = s.nodes
    xl = np.asarray(x, ld).reshape(n, BS)
    off = np.asarray(s.offsets, np.int64)
    cols = np.asarray(s.columns, np.int64)
    vals = np.asarray(s.off_diagonal, ld).reshape(-1, BS)
    prod = vals * xl[cols]
    ax = np.asarray(s.diagonal, ld).reshape(n, BS) * xl
    nonempty = np.diff(off) > 0
    sums = np.add.reduceat(prod, off[:-1][nonempty], axis=0) if prod.shape[0] else np.zeros((0, BS), ld)
    ax[nonempty] += sums
    bl = np.asarray(s.rhs, ld).reshape(n, BS)
    return float(np.sqrt(np.sum((bl - ax) ** 2)) / np.sqrt(np.sum(bl ** 2)))


def usable_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit():
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline(s, x_ref, samples=10, sample_iterations=3):
    """The oracle (this project's CPU restatement of the reference path, not
    the reference itself) on this workload, on this host, as the reference's
    benchmark does it (benchmark_split.cpp:135-213): set-up excluded and
    timed apart, a warm-up, then `samples` alternating samples of the two
    variants (1 core, the serial restatement; all cores, OpenMP over rows,
    vector loops and the 512 independent ILU blocks), each sample
    `sample_iterations` BiCGSTAB + 4096-block ILU(0) iterations; then one full
    solve of each variant to rtol 1e-8, each verified by a long-double true
    residual and compared with the FP64 GPU solution to 1e-12."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc

    march = orc.select_build()
    threads = usable_cores()
    om = orc.OMatrix.split(s.nodes, 0, BS, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    t0 = time.perf_counter()
    pc = orc.OPreconditioner(om, orc.OPreconditioner.ILU, block_nodes=BLOCK_NODES)
    setup_s = time.perf_counter() - t0
    variants = (("one_core", 1), ("all_cores", threads))

    def run(nthreads, iterations, rtol=RTOL):
        orc.set_threads(nthreads)
        t = time.perf_counter()
        x, res = orc.bicgstab(om, pc, s.rhs, relative_tolerance=rtol, maximum_iterations=iterations)
        return (time.perf_counter() - t) * 1e3, x, res

    for _, nt in variants:  # warm-up
        run(nt, 1)
    per = {name: [] for name, _ in variants}
    for k in range(samples):
        order = variants if k % 2 == 0 else variants[::-1]
        for name, nt in order:
            ms, _, _ = run(nt, sample_iterations)
            per[name].append(ms)
    full = {}
    for name, nt in variants:
        ms, x, res = run(nt, 10000)
        rel_ld = long_double_true_rel(s, x)
        err = float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))
        if res["status"] != 0 or not rel_ld <= 1.05 * RTOL:
            raise RuntimeError(f"CPU baseline solve ({name}) failed the long-double residual check: {rel_ld}")
        full[name] = {"solve_ms": round(ms, 1), "iterations": res["iterations"],
                      "true_rel_long_double": rel_ld, "error_vs_fp64": err}
    orc.set_threads(1)
    model, nproc = cpu_info()
    out = {"set_up_s": round(setup_s, 3), "march": march, "cpu": model, "nproc": nproc, "threads": threads}
    for name, nt in variants:
        v = np.array(per[name])
        out[name] = dict(full[name], threads=nt, sample_ms={
            "median": round(float(np.median(v)), 2), "min": round(float(v.min()), 2),
            "max": round(float(v.max()), 2), "samples": len(v), "iterations_per_sample": sample_iterations})
        log(f"[bench] cpu {name} ({nt} threads, {march}): {sample_iterations}-iteration sample median "
            f"{np.median(v):.0f} ms [{v.min():.0f}, {v.max():.0f}]; full solve {full[name]['solve_ms']:.0f} ms, "
            f"{full[name]['iterations']} iterations, long-double true rel {full[name]['true_rel_long_double']:.2e}")
    return out


def bsr_kernel_loop(steps):
    """The dense 4x4 BSR SpMV of the 1024x1024 grid (FP64 values, then BF16
    values with FP32 vectors), `steps` launches each, for rocprofv3 runs."""
    with owtk.Context(0) as ctx:
        n, off, col, val, _ = owtk.generate_bsr(BSR_GRID, BSR_GRID, 4, SEED, ordering=owtk.ORDER_HILBERT)
        m = owtk.Matrix.bsr(ctx, n, 0, 4, off, col, val)
        m.add_storage(owtk.BF16)
        x = ctx.upload(np.random.default_rng(0).uniform(-1, 1, n * 4))
        y = ctx.empty(n * 4)
        xf = ctx.upload(np.random.default_rng(0).uniform(-1, 1, n * 4).astype(np.float32))
        yf = ctx.empty(n * 4, np.float32)
        for _ in range(steps):
            m.spmv(x, y)
        for _ in range(steps):
            m.spmv(xf, yf, owtk.BF16)
        ctx.synchronize()
        info = m.info()
        print(json.dumps({"kernel": "bsr", "launches": steps, "vertices": n,
                          "spmv_bytes_f64": info.spmv_bytes_f64, "spmv_bytes_bf16": info.spmv_bytes_bf16}))


def learned_update_loop(steps, pc_storage="bf16"):
    """The learned preconditioner's numeric update on the headline system
    (network inference, coefficient and column scaling), `steps` times after
    two untimed ones; prints the mean host-timed update."""
    with owtk.Context(0) as ctx:
        s, m = build_system(ctx)
        st = {"f32": owtk.F32, "bf16": owtk.BF16}[pc_storage]
        m.add_storage(st)
        model = owtk.GnnModel.load(ctx, MODEL)
        pc = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model, storage=st)
        for _ in range(2):
            pc.update_values()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pc.update_values()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / max(steps, 1)
        print(json.dumps({"kernel": "learned-update", "launches": steps, "vertices": s.nodes,
                          "storage": pc_storage, "update_ms": round(ms, 4)}))


def kernel_loop(kind, steps, pc_storage="f32"):
    """Kernel loops for rocprofv3 counter runs: FP64 split SpMV, the learned
    preconditioner apply (FP32 storage), or a copy kernel of known size
    (FP64 -> FP32 conversion: 8 B read, 4 B written)."""
    if kind == "bsr":
        bsr_kernel_loop(steps)
        return
    if kind == "learned-update":
        learned_update_loop(steps, pc_storage)
        return
    with owtk.Context(0) as ctx:
        s, m = build_system(ctx)
        n = s.nodes * BS
        x = ctx.upload(np.random.default_rng(0).uniform(-1, 1, n))
        y = ctx.empty(n)
        yf = ctx.empty(n, np.float32)
        pc = None
        if kind == "learned":
            st = {"f32": owtk.F32, "bf16": owtk.BF16}[pc_storage]
            m.add_storage(st)
            model = owtk.GnnModel.load(ctx, MODEL)
            pc = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model, storage=st)
            xf = ctx.upload(np.random.default_rng(0).uniform(-1, 1, n).astype(np.float32))
        for k in range(steps):
            if kind == "spmv":
                m.spmv(x, y)
            elif kind == "learned":
                pc.apply(xf, yf)
            else:
                owtk.check(owtk.lib().owtk_convert(ctx.handle, n, owtk.F32, yf.ptr, owtk.F64, x.ptr))
        ctx.synchronize()
        info = m.info()
        print(json.dumps({"kernel": kind, "launches": steps, "elements": n,
                          "spmv_bytes_f64": info.spmv_bytes_f64, "copy_bytes": 12 * n}))


def batched_roofline(systems, vertices, ms_it):
    """Algorithmic HBM bytes of one all-active batched BiCGSTAB + Jacobi
    iteration over the autotuner's time for one. Its two SpMVs each read every
    stored entry (3 FP64 values + a 4-byte column) and the row lengths. Its
    five streaming kernels move 25 block vectors of 3 FP64 values per vertex:
    p update 4 read + 2 written, v = A p^ 3 (p^ gathered once, r^, v),
    s update 3 + 2, t = A s^ 3, x/r update 6 + 2. Finishers and padding slots
    are overhead, not counted."""
    entries = sum(int(s[1][-1]) for s in systems)
    nbytes = 2 * (entries * (BS * 8 + 4) + vertices * 4) + 25 * vertices * BS * 8
    gbps = nbytes / (ms_it * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbps, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(gbps / PEAK_HBM_GBPS, 4), "traffic": None,
            "scope": "one all-active iteration (5 streaming kernels + 5 finishers), autotuned tile",
            "bytes_per_iteration": nbytes, "entries": entries}


def batched_bench(args):
    """512 independent block-3 systems, BiCGSTAB + Jacobi to rtol 1e-8: the
    small systems each solved whole by one workgroup in one launch, the others
    in graph-replayed iterations beside it (split autotuned; the all-graph
    time is reported too); systems/s and per-system stop times."""
    from owtk import workloads

    t0 = time.perf_counter()
    systems, rhs, grids = workloads.batched_systems(BATCH_SYSTEMS)
    if args.batch_max_vertices > 0:  # only the systems of at most this many vertices
        keep = [k for k, s in enumerate(systems) if s[0] <= args.batch_max_vertices]
        offs = np.concatenate([[0], np.cumsum([s[0] * BS for s in systems])])
        rhs = np.concatenate([rhs[offs[k]:offs[k + 1]] for k in keep])
        systems = [systems[k] for k in keep]
        grids = [grids[k] for k in keep]
    vertices = sum(s[0] for s in systems)
    log(f"[bench] batched: {len(systems)} systems, {vertices} vertices, generated in {time.perf_counter() - t0:.1f}s")
    opts = owtk.SolverOptions(relative_tolerance=RTOL, maximum_iterations=20000)
    with owtk.Context(0) as ctx:
        t0 = time.perf_counter()
        bt = owtk.Batch(ctx, BS, systems)
        ctx.synchronize()
        log(f"[bench] batch upload {time.perf_counter() - t0:.1f}s")
        b = ctx.upload(rhs)
        x = ctx.empty(rhs.size)
        tile, ms_it = bt.autotune(b, iterations=8)
        log(f"[bench] autotuner: tile {tile} rows, {ms_it:.3f} ms per all-active iteration")

        def timed(label):
            for _ in range(args.warmup):
                x.zero()
                bt.bicgstab(b, x, opts, graph_iterations=BATCH_GRAPH_ITERATIONS)
            times, res = [], None
            for k in range(args.steps):
                x.zero()
                ctx.synchronize()
                t1 = time.perf_counter()
                res = bt.bicgstab(b, x, opts, graph_iterations=BATCH_GRAPH_ITERATIONS)
                ctx.synchronize()
                times.append((time.perf_counter() - t1) * 1e3)
                log(f"[bench] batched solve ({label}) {k + 1}/{args.steps}: {times[-1]:.1f} ms")
            return float(np.median(times)), res

        # every system on the graph replays, then the autotuned split between
        # one-workgroup solves of the small systems and the graph replays
        if args.group_limit >= 0:
            # a fixed split, set before any solve
            limit, threads = args.group_limit, 1024
            bt.set_group_limit(limit, threads)
            graph_ms = None
        else:
            graph_ms, _ = timed("graph replays only")
            limit, threads, tune_ms = bt.autotune_groups(b, opts, graph_iterations=BATCH_GRAPH_ITERATIONS)
            log(f"[bench] group autotuner: {tune_ms:.1f} ms per solve")
        group_systems = sum(s[0] <= limit for s in systems)
        log(f"[bench] systems of <= {limit} vertices on one workgroup of {threads} threads ({group_systems} systems)")
        ms, last = timed(f"group limit {limit}")
        if graph_ms is None and group_systems == 0:
            graph_ms = ms
        stream = stream_ceiling(ctx)
        stop = np.array([r["stop_ms"] for r in last])
        its = np.array([r["iterations"] for r in last])
        converged = sum(r["status_name"] == "converged" for r in last)
        worst = max(r["relative_residual_norm"] for r in last)
        # the same systems one at a time through the single-system solver
        t1 = time.perf_counter()
        seq_its, off = [], 0
        for (n, o, c, v) in systems:
            m = owtk.Matrix.csr(ctx, n, 0, BS, o, c, v)
            pc = owtk.Preconditioner(m, owtk.PC_JACOBI)
            bs_ = ctx.upload(rhs[off:off + n * BS])
            xs = ctx.empty(n * BS)
            seq_its.append(owtk.bicgstab(m, pc, bs_, xs, opts)["iterations"])
            off += n * BS
        ctx.synchronize()
        seq_ms = (time.perf_counter() - t1) * 1e3
        cpu = None
        if not args.no_cpu:
            sys.path.insert(0, str(ROOT))
            from oracle import oracle as orc
            order = np.argsort([s[0] for s in systems])
            sample = [int(i) for i in order[::max(1, len(order) // 24)]][:24]
            offs = np.concatenate([[0], np.cumsum([s[0] * BS for s in systems])])
            t1 = time.perf_counter()
            for i in sample:
                n, o, c, v = systems[i]
                om = orc.OMatrix.csr(n, 0, BS, o, c, v)
                opc = orc.OPreconditioner(om, orc.OPreconditioner.JACOBI)
                orc.bicgstab(om, opc, rhs[offs[i]:offs[i + 1]], relative_tolerance=RTOL, maximum_iterations=20000)
            cpu_s = time.perf_counter() - t1
            frac = sum(systems[i][0] for i in sample) / vertices
            cpu = {"value": round(len(sample) / cpu_s * frac * len(systems) / len(sample), 3), "unit": "systems/s",
                   "cores": 1, "kind": "port",
                   "sample": f"oracle BiCGSTAB + Jacobi on {len(sample)} of the systems spread over the size range "
                             f"({cpu_s:.1f}s), scaled by their share ({frac:.3f}) of all vertices"}
        out = {
            "metric": "batched uneven systems: systems/s (BiCGSTAB + Jacobi to rtol 1e-8)", "value": round(
                len(systems) / (ms / 1e3), 2), "unit": "systems/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{len(systems)} systems, grids log-uniform 32x16..512x256, block 3, "
                                   f"edge-contrast sigma U(0,1.5), seeds 3000+, {vertices} vertices in all"
                                   + (f" (the systems of at most {args.batch_max_vertices} vertices of the 512)"
                                      if args.batch_max_vertices > 0 else ""),
                       "tile_rows": tile, "graph_iterations": BATCH_GRAPH_ITERATIONS,
                       "group_limit_vertices": limit, "group_threads": threads, "group_systems": group_systems},
            "graph_only_ms": round(graph_ms, 3) if graph_ms is not None else None,
            "graph_only_systems_per_s": round(len(systems) / (graph_ms / 1e3), 2) if graph_ms is not None else None,
            "per_system_ms": {"p50": float(np.percentile(stop, 50)), "p99": float(np.percentile(stop, 99)),
                              "max": float(stop.max())},
            "iterations": {"p50": float(np.median(its)), "max": int(its.max())},
            "converged": int(converged), "max_true_rel": worst,
            "sequential_single_system_ms": round(seq_ms, 1),
            "sequential_systems_per_s": round(len(systems) / (seq_ms / 1e3), 2),
            "iteration_count_max_diff_vs_single": int(np.max(np.abs(np.array(seq_its) - its))),
            "autotune_ms_per_iteration": ms_it, "roofline": batched_roofline(systems, vertices, ms_it),
            "hbm_stream": stream, "cpu_baseline": cpu,
        }
        print(json.dumps(out))
    return 0


def bsr_bench(args):
    """Dense 4x4-block BSR on the 1024x1024 grid: BF16 matrix and multicolour
    SSOR storage, FP32 accumulation, FP64-refined FGMRES(40) to rtol 1e-10;
    error against an FP64 GPU solve."""
    t0 = time.perf_counter()
    n, off, col, val, rhs = owtk.generate_bsr(BSR_GRID, BSR_GRID, 4, SEED, ordering=owtk.ORDER_HILBERT)
    log(f"[bench] bsr: {n} vertices, {off[-1]} blocks, generated in {time.perf_counter() - t0:.1f}s")
    with owtk.Context(0) as ctx:
        m = owtk.Matrix.bsr(ctx, n, 0, 4, off, col, val)
        m.add_storage(owtk.BF16)
        b = ctx.upload(rhs)
        x = ctx.empty(rhs.size)
        r = ctx.empty(rhs.size)
        bn = ctx.norm(b)
        # FP64 reference solution: FP64 FGMRES(40) + FP64 SSOR to 1e-12
        pc64 = owtk.Preconditioner(m, owtk.PC_SSOR)
        ref = owtk.fgmres(m, pc64, b, x, owtk.SolverOptions(relative_tolerance=1e-12, restart=40,
                                                           maximum_iterations=5000))
        x64 = x.download()
        log(f"[bench] bsr FP64 reference: {ref['status_name']} after {ref['iterations']} iterations")
        pc16 = owtk.Preconditioner(m, owtk.PC_SSOR, storage=owtk.BF16)
        ropts = owtk.refinement_options(inner_storage=owtk.BF16, inner_restart=40, inner_relative_tolerance=1e-4,
                                        relative_tolerance=1e-10)
        times, res = [], None
        for k in range(args.warmup + args.steps):
            x.zero()
            ctx.synchronize()
            t1 = time.perf_counter()
            res = owtk.fgmres_refined(m, pc16, b, x, ropts)
            ctx.synchronize()
            if k >= args.warmup:
                times.append((time.perf_counter() - t1) * 1e3)
                log(f"[bench] bsr bf16 fgmres-ir+ssor step {len(times)}/{args.steps}: {times[-1]:.1f} ms, "
                    f"inner it={res['inner_iterations']}, refinements={res['refinement_steps']}")
        m.residual(b, x, r)
        rel = ctx.norm(r) / bn
        err = float(np.linalg.norm(x.download() - x64) / np.linalg.norm(x64))
        info = m.info()
        sp = {}
        xs = ctx.upload(np.random.default_rng(1).uniform(-1, 1, rhs.size))
        ys = ctx.empty(rhs.size)
        xf = ctx.upload(np.random.default_rng(1).uniform(-1, 1, rhs.size).astype(np.float32))
        yf = ctx.empty(rhs.size, np.float32)
        for name, dt, cat, xv, yv in (("f64", owtk.F64, ctx.PROF_SPMV_F64, xs, ys),
                                      ("bf16", owtk.BF16, ctx.PROF_SPMV_BF16, xf, yf)):
            for _ in range(5):
                m.spmv(xv, yv, dt)
            ctx.synchronize()
            ctx.profile_reset()
            ctx.profile(True)
            for _ in range(30):
                m.spmv(xv, yv, dt)
            ctx.profile(False)
            p = ctx.profile_read(cat)
            sp[name] = {"avg_ms": p["avg_ms"], "gbps": p["gbps"], "bytes": p["bytes"] // max(p["launches"], 1)}
        log(f"[bench] bsr spmv f64 {sp['f64']['avg_ms'] * 1e3:.1f} us {sp['f64']['gbps']:.0f} GB/s, "
            f"bf16 {sp['bf16']['avg_ms'] * 1e3:.1f} us {sp['bf16']['gbps']:.0f} GB/s")
        ms = float(np.median(times))
        out = {
            "metric": "ms to true rel. residual 1e-10 (dense 4x4 BSR, BF16 storage)", "value": round(ms, 3),
            "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 storage, f32 accumulation, f64 refinement", "data": "synthetic",
            "config": {"workload": f"{BSR_GRID}x{BSR_GRID} triangulated grid, {n} vertices, dense 4x4 blocks "
                                   "(diag N(0,1)+6I, off N(0,0.15)U(0.1,1)), Hilbert order, seed 0",
                       "method": "FP64-refined FGMRES(40), BF16 matrix + BF16 multicolour block SSOR"},
            "inner_iterations": res["inner_iterations"], "refinements": res["refinement_steps"],
            "status": res["status_name"], "true_rel": rel, "error_vs_fp64": err,
            "fp64_reference_iterations": ref["iterations"], "spmv": sp,
            "spmv_bytes": {"f64": info.spmv_bytes_f64, "bf16": info.spmv_bytes_bf16},
        }
        print(json.dumps(out))
    return 0


def pipelined_bench(args):
    """The reference's production selection (split SSOR + pipelined BiCGSTAB)
    on the 2048x1024 grid: FP64 pipelined and communication-hiding BiCGSTAB
    beside the standard recurrence and IDR(4), all with block-Jacobi SSOR
    (4096-vertex blocks, the one-GPU analogue of the per-rank SSOR), rtol
    1e-8."""
    with owtk.Context(0) as ctx:
        s, m = build_system(ctx)
        n = s.nodes * BS
        b = ctx.upload(s.rhs)
        x = ctx.empty(n)
        r = ctx.empty(n)
        bnorm = ctx.norm(b)
        pc = owtk.Preconditioner(m, owtk.PC_SSOR, block_nodes=BLOCK_NODES)
        opts = owtk.SolverOptions(relative_tolerance=RTOL, maximum_iterations=2000)
        out_methods = {}
        idr_opts = owtk.SolverOptions(relative_tolerance=RTOL, maximum_iterations=2000, idr_shadow_space=4)
        for name, fn, o in (("pipelined_bicgstab", owtk.pipelined_bicgstab, opts),
                            ("communication_hiding_bicgstab", owtk.communication_hiding_bicgstab, opts),
                            ("bicgstab", owtk.bicgstab, opts),
                            ("idrs4", owtk.idrs, idr_opts)):
            times, results, prof = timed_solves(ctx, f"{name}+ssor", args.steps, args.warmup,
                                                lambda f=fn, oo=o: f(m, pc, b, x, oo), x)
            res = results[-1]
            rel = true_rel(ctx, m, b, x, r, bnorm)
            ms = float(np.median(times))
            out_methods[name] = {
                "ms": round(ms, 3), "iterations": res["iterations"], "status": res["status_name"],
                "ms_per_iteration": round(ms / max(res["iterations"], 1), 4), "true_rel": rel,
                "reductions": res["global_reductions"], "operator_applications": res["operator_applications"],
                "preconditioner_applications": res["preconditioner_applications"],
                "ssor_apply_ms": round(prof["precond_f64"]["avg_ms"], 4),
                "spmv_ms": round(prof["spmv_f64"]["avg_ms"], 4)}
            if name.startswith("idrs"):
                out_methods[name]["note"] = ("shadow-space draws (seeded std::mt19937_64, as the reference) are "
                                             "generated on the host in the warm-up solve and kept per context")
        head = out_methods["pipelined_bicgstab"]
        cpu = None
        if not args.no_cpu:
            sys.path.insert(0, str(ROOT))
            from oracle import oracle as orc
            om = orc.OMatrix.split(s.nodes, 0, BS, s.offsets, s.columns, s.diagonal, s.off_diagonal)
            opc = orc.OPreconditioner(om, orc.OPreconditioner.SSOR, block_nodes=BLOCK_NODES)
            its = 3
            t0 = time.perf_counter()
            orc.pipelined_bicgstab(om, opc, s.rhs, relative_tolerance=RTOL, maximum_iterations=its)
            iter_s = (time.perf_counter() - t0) / its
            log(f"[bench] cpu: pipelined bicgstab + ssor iteration {iter_s * 1e3:.0f} ms")
            model_name, cores = cpu_info()
            cpu = {"value": round(iter_s * head["iterations"] * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                   "sample": (f"oracle FP64 pipelined BiCGSTAB + 4096-block SSOR on this grid, {its} timed "
                              f"iterations ({iter_s * 1e3:.0f} ms each) x the {head['iterations']} iterations the "
                              f"solve needs; host {model_name}, {cores} hardware threads")}
        out = {
            "metric": "ms to true rel. residual 1e-8 (pipelined BiCGSTAB + SSOR, FP64)", "value": head["ms"],
            "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{NX}x{NY} triangulated grid, {s.nodes:,} vertices, block 3, split view, "
                                   "Hilbert order, seed 0; rtol 1e-8",
                       "method": "FP64 pipelined BiCGSTAB (pipelined_bicgstab.hpp:20-244) + block-Jacobi "
                                 "SSOR, 4096-vertex blocks"},
            "methods": out_methods, "cpu_baseline": cpu,
        }
        print(json.dumps(out))
    return 0


def multigrid_bench(args):
    """Two-level aggregation multigrid (multigrid.hpp) on the 512x256 grid,
    CSR form (BlockCsrMatrix with the diagonal inline), rtol 1e-8: V cycles
    with Jacobi and red-black Gauss-Seidel smoothing (aggregates of 4, 2+2
    fine and 24 coarse smoothing steps)."""
    nx, ny = 512, 256
    s = owtk.generate_split(nx, ny, BS, SEED, ordering=owtk.ORDER_HILBERT)
    co, cc, cv = s.csr()
    out_methods = {}
    with owtk.Context(0) as ctx:
        m = owtk.Matrix.csr(ctx, s.nodes, 0, BS, co, cc, cv)
        b = ctx.upload(s.rhs)
        x = ctx.empty(s.nodes * BS)
        r = ctx.empty(s.nodes * BS)
        bnorm = ctx.norm(b)
        opts = owtk.SolverOptions(relative_tolerance=RTOL, maximum_iterations=5000)
        for name, smoother, cycle in (("v_jacobi", owtk.MG_JACOBI, owtk.MG_V),
                                      ("v_red_black", owtk.MG_RED_BLACK, owtk.MG_V),
                                      ("w_red_black", owtk.MG_RED_BLACK, owtk.MG_W)):
            t0 = time.perf_counter()
            mg = owtk.Multigrid(m, smoother=smoother)
            ctx.synchronize()
            setup_s = time.perf_counter() - t0
            times, results, _ = timed_solves(ctx, f"multigrid {name}", args.steps, args.warmup,
                                             lambda g=mg, c=cycle: g.solve(b, x, opts, cycle=c), x)
            res = results[-1]
            ms = float(np.median(times))
            out_methods[name] = {"ms": round(ms, 2), "iterations": res["iterations"], "status": res["status_name"],
                                 "ms_per_cycle": round(ms / max(res["iterations"], 1), 4),
                                 "true_rel": true_rel(ctx, m, b, x, r, bnorm), "setup_s": round(setup_s, 3),
                                 "coarse_nodes": mg.coarse()[0]}
            if smoother == owtk.MG_RED_BLACK:
                out_methods[name]["note"] = ("coarse Gauss-Seidel keeps the reference's sequential row order: "
                                             "1,026 dependency levels among the black rows at this size")
            mg.close()
    head = out_methods["v_jacobi"]
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT))
        from oracle import oracle as orc
        om = orc.OMatrix.csr(s.nodes, 0, BS, co, cc, cv)
        omg = orc.OMultigrid(om, smoother=orc.OMultigrid.JACOBI)
        cycles = 5
        t0 = time.perf_counter()
        omg.solve(s.rhs, cycle=orc.OMultigrid.V, relative_tolerance=0.0, maximum_iterations=cycles)
        cyc_s = (time.perf_counter() - t0) / cycles
        model_name, cores = cpu_info()
        cpu = {"value": round(cyc_s * head["iterations"] * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
               "sample": (f"oracle multigrid, Jacobi V cycles on this grid: {cycles} timed cycles "
                          f"({cyc_s * 1e3:.1f} ms each) x the {head['iterations']} cycles the solve needs; "
                          f"host {model_name}, {cores} hardware threads")}
    out = {
        "metric": "ms to true rel. residual 1e-8 (aggregation multigrid, FP64)", "value": head["ms"], "unit": "ms",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{nx}x{ny} triangulated grid, {s.nodes:,} vertices, block 3, CSR form, Hilbert "
                               "order, seed 0; rtol 1e-8",
                   "method": "AggregationMultigrid (multigrid.hpp), V cycles, Jacobi smoothing (the default)"},
        "methods": out_methods, "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


def recycling_bench(args):
    """A sequence of related systems (recycling.hpp): the 2048x1024 grid, five
    right-hand sides (the generated one and four perturbed by 1% noise),
    FP64 FGMRES(30) + 4096-block ILU(0) to rtol 1e-8; plain FGMRES for each
    against recycled FGMRES and GCRO-DR (gcrodr.hpp:185-212: harmonic Ritz
    vectors of each solve's last cycle), each through one recycle space of 4
    vectors."""
    with owtk.Context(0) as ctx:
        s, m = build_system(ctx)
        n = s.nodes * BS
        rng = np.random.default_rng(11)
        rhs = [s.rhs] + [s.rhs * (1 + 0.01 * rng.standard_normal(n)) for _ in range(4)]
        bs_dev = [ctx.upload(b) for b in rhs]
        x = ctx.empty(n)
        pc = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=BLOCK_NODES)
        opts = owtk.SolverOptions(relative_tolerance=RTOL, maximum_iterations=5000, restart=30)

        def sequence(method):
            space = owtk.RecycleSpace(m, 4) if method != "fgmres" else None
            its, ms = [], []
            for b in bs_dev:
                x.zero()
                ctx.synchronize()
                t0 = time.perf_counter()
                if method == "fgmres":
                    res = owtk.fgmres(m, pc, b, x, opts)
                elif method == "recycled_fgmres":
                    res = owtk.recycled_fgmres(m, pc, b, x, space, opts)
                else:
                    res = owtk.gcrodr(m, pc, b, x, space, opts)
                ctx.synchronize()
                ms.append((time.perf_counter() - t0) * 1e3)
                its.append(res["iterations"])
                assert res["status_name"] == "converged", res
            if space is not None:
                space.close()
            return its, ms

        out_methods = {}
        for name in ("fgmres", "recycled_fgmres", "gcrodr"):
            for _ in range(args.warmup):
                sequence(name)
            runs = [sequence(name) for _ in range(args.steps)]
            its = runs[-1][0]
            totals = [sum(ms) for _, ms in runs]
            out_methods[name] = {"total_ms": round(float(np.median(totals)), 2), "iterations": its,
                                 "per_solve_ms": [round(v, 2) for v in runs[-1][1]]}
            log(f"[bench] {name}: {out_methods[name]}")
        head = out_methods["recycled_fgmres"]
        out = {
            "metric": "ms for a sequence of 5 related solves to true rel. residual 1e-8 (recycled FGMRES, FP64)",
            "value": head["total_ms"], "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": head["total_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{NX}x{NY} triangulated grid, {s.nodes:,} vertices, block 3, split view, Hilbert "
                                   "order, seed 0; 5 right-hand sides (4 with 1% noise); rtol 1e-8",
                       "method": "recycled_fgmres (recycling.hpp), FGMRES(30) + 4096-block ILU(0), 4 recycled "
                                 "vectors"},
            "methods": out_methods,
        }
        print(json.dumps(out))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", choices=["spmv", "copy", "learned", "bsr", "learned-update"], default=None)
    ap.add_argument("--storage", choices=["f32", "bf16"], default="f32",
                    help="matrix storage of the inner FGMRES (SpMV and ILU fallback)")
    ap.add_argument("--pc-storage", choices=["f32", "bf16"], default="bf16",
                    help="matrix storage the learned preconditioner's series reads")
    ap.add_argument("--orth", type=int, default=2,
                    help="Gram-Schmidt of the inner FGMRES: 0 iterated CGS, 1 adaptive, 2 one-synchronization "
                         "(GmresOrthogonalization, core.hpp:32-39); inner products stay FP64")
    ap.add_argument("--model", default=str(MODEL), help="learned preconditioner parameters (.npz)")
    ap.add_argument("--batch-max-vertices", type=int, default=0,
                    help="batched workload: keep only the systems of at most this many vertices (0: all 512)")
    ap.add_argument("--group-limit", type=int, default=-1,
                    help="batched workload: systems of at most this many vertices on one workgroup each "
                         "(0: none; default: autotuned)")
    ap.add_argument("--workload", choices=["headline", "batched", "bsr", "pipelined", "multigrid", "recycling"],
                    default="headline")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a multi-GPU solve would shard by "
              "vertex partition, with a ghost halo per SpMV and one allreduce per scalar batch)", file=sys.stderr)
        return 2
    if args.kernel:
        kernel_loop(args.kernel, args.steps, args.pc_storage)
        return 0
    if args.workload == "batched":
        return batched_bench(args)
    if args.workload == "bsr":
        return bsr_bench(args)
    if args.workload == "pipelined":
        return pipelined_bench(args)
    if args.workload == "multigrid":
        return multigrid_bench(args)
    if args.workload == "recycling":
        return recycling_bench(args)

    with owtk.Context(0) as ctx:
        s, m = build_system(ctx)
        n = s.nodes * BS
        b = ctx.upload(s.rhs)
        x = ctx.empty(n)
        r = ctx.empty(n)
        bnorm = ctx.norm(b)
        storage = {"f32": owtk.F32, "bf16": owtk.BF16}[args.storage]
        pc_storage = {"f32": owtk.F32, "bf16": owtk.BF16}[args.pc_storage]

        setup = {}
        t0 = time.perf_counter()
        ilu64 = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=BLOCK_NODES)
        ctx.synchronize()
        setup["ilu0_fp64_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        ilu_low = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=BLOCK_NODES, storage=storage)
        ctx.synchronize()
        setup["ilu0_low_s"] = time.perf_counter() - t0
        model_path = pathlib.Path(args.model)
        model = owtk.GnnModel.load(ctx, model_path)
        t0 = time.perf_counter()
        learned = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model, storage=pc_storage)
        ctx.synchronize()
        setup["learned_s"] = time.perf_counter() - t0
        plain = owtk.Preconditioner(m, owtk.PC_LEARNED, model=owtk.GnnModel(ctx, BS, model.degree),
                                    storage=pc_storage)
        info = ilu64.info()
        log(f"[bench] set-up: ILU(0) fp64 {setup['ilu0_fp64_s']:.2f}s, low {setup['ilu0_low_s']:.2f}s "
            f"({info.blocks} blocks, levels fwd {info.forward_levels}); learned (network inference) "
            f"{setup['learned_s'] * 1e3:.1f} ms, degree {model.degree}")

        # numeric updates (same pattern, new values): the symbolic set-up is reused
        upd = {}
        for name, pcu in (("ilu0_fp64", ilu64), ("ilu0_low", ilu_low), ("learned", learned)):
            pcu.update_values()
            ctx.synchronize()
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                pcu.update_values()
                ctx.synchronize()
                ts.append((time.perf_counter() - t0) * 1e3)
            upd[name] = round(float(np.median(ts)), 3)
        log(f"[bench] numeric update_values (ms): {upd}")

        # the FP64 solution every result's error is measured against: FP64
        # BiCGSTAB + ILU(0) to rtol 1e-12
        x.zero()
        ref = owtk.bicgstab(m, ilu64, b, x, owtk.SolverOptions(relative_tolerance=1e-12, maximum_iterations=2000))
        x_ref = x.download()
        log(f"[bench] FP64 reference to 1e-12: {ref['status_name']}, {ref['iterations']} iterations")

        def error_vs_ref(v):
            return float(np.linalg.norm(v - x_ref) / np.linalg.norm(x_ref))

        # plain FP64 baseline: BiCGSTAB + block-Jacobi ILU(0)
        opts64 = owtk.SolverOptions(relative_tolerance=RTOL)
        t64, res64, prof64 = timed_solves(ctx, "fp64 bicgstab+ilu0", args.steps, args.warmup,
                                          lambda: owtk.bicgstab(m, ilu64, b, x, opts64), x)
        rel64 = true_rel(ctx, m, b, x, r, bnorm)
        x64 = x.download()

        def refined(label, pc, fallback=None):
            ropts = owtk.refinement_options(inner_storage=storage, inner_restart=30, inner_relative_tolerance=1e-4,
                                            relative_tolerance=RTOL)
            ropts.outer.gmres_orthogonalization = args.orth
            t, res, prof = timed_solves(ctx, label, args.steps, args.warmup,
                                        lambda: owtk.fgmres_refined(m, pc, b, x, ropts, fallback=fallback), x)
            rel = true_rel(ctx, m, b, x, r, bnorm)
            err = error_vs_ref(x.download())
            last = res[-1]
            return {"ms": round(float(np.median(t)), 3), "ms_min": round(min(t), 3), "ms_max": round(max(t), 3),
                    "inner_iterations": last["inner_iterations"], "refinements": last["refinement_steps"],
                    "fallback_refinements": last["fallback_refinements"], "status": last["status_name"],
                    "true_rel": rel, "error_vs_fp64": err}, prof

        # headline: learned preconditioner + mixed precision (ILU(0) fallback)
        head, profh = refined(f"learned+mixed ({args.storage})", learned, fallback=ilu_low)
        # the network's gamma averaged over the vertices: one polynomial, no
        # per-vertex modulation
        average = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model.constant_polynomial(m), storage=pc_storage)
        vavg, _ = refined(f"vertex-averaged gamma+mixed ({args.storage})", average, fallback=ilu_low)
        # the learned global polynomial without the network (decoder zeroed: gamma = g)
        polynomial = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model.without_network(), storage=pc_storage)
        nonet, _ = refined(f"learned polynomial+mixed ({args.storage}), network zeroed", polynomial,
                           fallback=ilu_low)
        # the plain truncated series (untrained model: gamma = 1)
        noet, _ = refined(f"series+mixed ({args.storage}), no network", plain, fallback=ilu_low)
        # mixed precision with the fallback preconditioner alone
        milu, _ = refined(f"ilu0+mixed ({args.storage})", ilu_low)

        # split SpMV: the path's data-parallel hot kernel and the metric's "SpMV HBM GB/s"
        sp = spmv_timing(ctx, m, n)
        stream = stream_ceiling(ctx)
        log(f"[bench] stream ceiling: {stream}")
        log(f"[bench] spmv f64 {sp['f64']['avg_ms'] * 1e3:.1f} us {sp['f64']['gbps']:.0f} GB/s, "
            f"f32 {sp['f32']['avg_ms'] * 1e3:.1f} us {sp['f32']['gbps']:.0f} GB/s, "
            f"bf16 {sp['bf16']['avg_ms'] * 1e3:.1f} us {sp['bf16']['gbps']:.0f} GB/s")
        ms64 = float(np.median(t64))
        headline = head["ms"]
        # The headline solve's dominant kernel is the learned series step.  Its
        # in-solve event times include launches made no-ops by the solver's stop
        # flag, so the roofline is taken from back-to-back applies instead.
        lt = learned_timing(ctx, learned, n)
        log(f"[bench] learned series step {lt['step_avg_ms'] * 1e3:.1f} us ({lt['step_gbps']:.0f} GB/s), "
            f"apply {lt['apply_ms'] * 1e3:.1f} us ({lt['apply_gbps']:.0f} GB/s)")
        traffic = TRAFFIC_BYTES.get(("learned", args.pc_storage))
        roof = {
            "kernel": KERNEL_NAMES["learned" if args.pc_storage == "f32" else "learned_bf16"]
                      + " (largest share of the headline solve)",
            "bound": "hbm", "achieved": round(lt["step_gbps"], 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(lt["step_gbps"] / PEAK_HBM_GBPS, 4), "traffic": traffic,
            "avg_launch_ms": round(lt["step_avg_ms"], 5), "bytes_per_launch": lt["step_bytes"],
            "timing": "HIP event pair per launch, 30 back-to-back applies outside the solver",
            "apply_ms": round(lt["apply_ms"], 5), "apply_gbps": round(lt["apply_gbps"], 1),
            "in_solve_launches": profh["learned"]["launches"],
        }
        spmv_roof = {"kernel": KERNEL_NAMES["spmv_f64"] + " (dedicated loop)", "achieved": round(sp["f64"]["gbps"], 1),
                     "peak": PEAK_HBM_GBPS, "frac": round(sp["f64"]["gbps"] / PEAK_HBM_GBPS, 4),
                     "avg_launch_ms": round(sp["f64"]["avg_ms"], 5), "bytes_per_launch": sp["f64"]["bytes"],
                     "traffic": TRAFFIC_BYTES.get(("spmv", "f64")),
                     "timing": "one HIP event pair around 50 back-to-back launches"}
        cpu = None
        if not args.no_cpu:
            cb = cpu_baseline(s, x_ref)
            one = cb["one_core"]
            cpu = {"value": one["solve_ms"], "unit": "ms", "cores": 1, "kind": "port",
                   "sample": (f"the oracle (CPU restatement of the reference path, -march={cb['march']}) "
                              f"FP64 BiCGSTAB + 4096-block ILU(0) on this grid to rtol 1e-8: one full solve per "
                              f"variant ({one['iterations']} iterations on 1 core; long-double true residual "
                              f"checked) plus {one['sample_ms']['samples']} alternating 3-iteration samples per "
                              f"variant; ILU set-up {cb['set_up_s']:.1f}s excluded; host {cb['cpu']}, nproc "
                              f"{cb['nproc']}, all-cores variant on {cb['threads']} threads"),
                   **cb}
        out = {
            "metric": METRIC, "value": round(headline, 3), "unit": "ms", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(headline, 3),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": f"{args.storage} matrix / {args.pc_storage} preconditioner storage, f32 inner "
                     f"arithmetic, f64 refinement", "data": "synthetic",
            "config": {"workload": "2048x1024 triangulated grid, 2,097,152 vertices, block 3, split view, "
                                   "Hilbert order, seed 0; rtol 1e-8 (true FP64 residual)",
                       "headline": f"FP64-refined FGMRES(30) (Gram-Schmidt mode {args.orth}, FP64 inner products), "
                                   f"{args.storage} matrix storage, learned vertex-block preconditioner on "
                                   f"{args.pc_storage} matrix storage (degree {model.degree}, {model_path.name}), "
                                   f"ILU(0) fallback",
                       "baseline": "FP64 BiCGSTAB + FP64 block-Jacobi ILU(0), 4096-vertex blocks"},
            "fp64_bicgstab_ilu0_ms": round(ms64, 3), "speedup_vs_fp64": round(ms64 / headline, 3),
            "fp64_iterations": res64[-1]["iterations"], "fp64_true_rel": rel64,
            "fp64_error_vs_fp64_1e-12": error_vs_ref(x64),
            "error_reference": "FP64 BiCGSTAB + ILU(0) to rtol 1e-12 on the GPU "
                               f"({ref['iterations']} iterations)",
            "learned_mixed": head, "learned_vertex_average": vavg, "learned_no_network": nonet,
            "series_mixed_no_network": noet,
            "ilu0_mixed": milu,
            "setup_s": {k: round(v, 4) for k, v in setup.items()}, "update_values_ms": upd,
            "spmv": sp, "spmv_roofline": spmv_roof, "hbm_stream": stream, "roofline": roof, "cpu_baseline": cpu,
            "kernels": {"fp64": prof64, "learned_mixed": profh},
        }
        print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
