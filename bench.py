"""Headline benchmark: time to true relative residual 1e-8 on the 2M-vertex
block-3 grid, mixed-precision FGMRES vs plain FP64 BiCGSTAB + block-Jacobi
ILU(0), plus the split SpMV's HBM bandwidth.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
    python bench.py --kernel spmv|copy --steps K     (single-kernel loops for rocprofv3 counters)

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
RTOL = 1e-8


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
    for _ in range(warmup):
        x.zero()
        run()
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
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
            f"status={res['status_name']}")
    ctx.profile(False)
    prof = {name: ctx.profile_read(cat) for name, cat in (
        ("spmv_f64", ctx.PROF_SPMV_F64), ("spmv_f32", ctx.PROF_SPMV_F32), ("precond_f64", ctx.PROF_PRECOND_F64),
        ("precond_low", ctx.PROF_PRECOND_LOW))}
    return times, results, prof


def spmv_timing(ctx, m, n, reps=50):
    """Back-to-back split SpMVs on the context stream, timed with HIP events
    (one event pair per launch): average kernel time and algorithmic GB/s per
    storage type."""
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
        ctx.profile_reset()
        ctx.profile(True)
        for _ in range(reps):
            m.spmv(x, y, dtype)
        ctx.profile(False)
        p = ctx.profile_read(cat)
        out[name] = {"avg_ms": p["avg_ms"], "gbps": p["gbps"], "bytes": p["bytes"] // max(p["launches"], 1),
                     "launches": p["launches"]}
    return out


def cpu_baseline(s, gpu_iterations):
    """Oracle (CPU restatement of the reference) on this workload, 1 core:
    split SpMV GB/s and FP64 BiCGSTAB + 4096-block ILU(0) iterations."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc

    om = orc.OMatrix.split(s.nodes, 0, BS, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    x = np.random.default_rng(0).uniform(-1, 1, s.nodes * BS)
    om.spmv(x)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        om.spmv(x)
    spmv_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    pc = orc.OPreconditioner(om, orc.OPreconditioner.ILU, block_nodes=BLOCK_NODES)
    setup_s = time.perf_counter() - t0
    its = 3
    t0 = time.perf_counter()
    _, res = orc.bicgstab(om, pc, s.rhs, relative_tolerance=RTOL, maximum_iterations=its)
    iter_s = (time.perf_counter() - t0) / its
    log(f"[bench] cpu: spmv {spmv_s * 1e3:.1f} ms, ILU setup {setup_s:.1f}s, iteration {iter_s * 1e3:.0f} ms")
    return spmv_s, setup_s, iter_s


def kernel_loop(kind, steps):
    """Single-kernel loop (rocprofv3 counter runs): FP64 split SpMV or a copy
    kernel of known size (FP64 -> FP32 conversion: 8 B read, 4 B written)."""
    with owtk.Context(0) as ctx:
        s, m = build_system(ctx)
        n = s.nodes * BS
        x = ctx.upload(np.random.default_rng(0).uniform(-1, 1, n))
        y = ctx.empty(n)
        yf = ctx.empty(n, np.float32)
        for k in range(steps):
            if kind == "spmv":
                m.spmv(x, y)
            else:
                owtk.check(owtk.lib().owtk_convert(ctx.handle, n, owtk.F32, yf.ptr, owtk.F64, x.ptr))
        ctx.synchronize()
        info = m.info()
        print(json.dumps({"kernel": kind, "launches": steps, "elements": n,
                          "spmv_bytes_f64": info.spmv_bytes_f64, "copy_bytes": 12 * n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", choices=["spmv", "copy"], default=None)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the solve shards by vertex partition; see DESIGN.md)", file=sys.stderr)
        return 2
    if args.kernel:
        kernel_loop(args.kernel, args.steps)
        return 0

    with owtk.Context(0) as ctx:
        s, m = build_system(ctx)
        n = s.nodes * BS
        b = ctx.upload(s.rhs)
        x = ctx.empty(n)
        r = ctx.empty(n)
        bnorm = ctx.norm(b)

        t0 = time.perf_counter()
        ilu64 = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=BLOCK_NODES)
        ctx.synchronize()
        setup64 = time.perf_counter() - t0
        t0 = time.perf_counter()
        ilu32 = owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=BLOCK_NODES, storage=owtk.F32)
        ctx.synchronize()
        setup32 = time.perf_counter() - t0
        info = ilu64.info()
        log(f"[bench] ILU(0) set-up fp64 {setup64:.2f}s fp32 {setup32:.2f}s, {info.blocks} blocks, "
            f"levels fwd {info.forward_levels} bwd {info.backward_levels}")

        # plain FP64 baseline: BiCGSTAB + block-Jacobi ILU(0)
        opts64 = owtk.SolverOptions(relative_tolerance=RTOL)
        t64, res64, prof64 = timed_solves(ctx, "fp64 bicgstab+ilu0", args.steps, args.warmup,
                                          lambda: owtk.bicgstab(m, ilu64, b, x, opts64), x)
        rel64 = true_rel(ctx, m, b, x, r, bnorm)
        x64 = x.download()

        # mixed precision: FP64 refinement around FP32-storage FGMRES(30) + FP32 ILU(0)
        ropts = owtk.refinement_options(inner_storage=owtk.F32, inner_restart=30, inner_relative_tolerance=1e-4,
                                        relative_tolerance=RTOL)
        tmx, resmx, profmx = timed_solves(ctx, "mixed fgmres-ir+ilu0(f32)", args.steps, args.warmup,
                                          lambda: owtk.fgmres_refined(m, ilu32, b, x, ropts), x)
        relmx = true_rel(ctx, m, b, x, r, bnorm)
        xmx = x.download()
        err = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))

        # split SpMV: the path's data-parallel hot kernel and the metric's "SpMV HBM GB/s"
        sp = spmv_timing(ctx, m, n)
        log(f"[bench] spmv f64 {sp['f64']['avg_ms'] * 1e3:.1f} us {sp['f64']['gbps']:.0f} GB/s, "
            f"f32 {sp['f32']['avg_ms'] * 1e3:.1f} us {sp['f32']['gbps']:.0f} GB/s, "
            f"bf16 {sp['bf16']['avg_ms'] * 1e3:.1f} us {sp['bf16']['gbps']:.0f} GB/s")
        ms64 = float(np.median(t64))
        msmx = float(np.median(tmx))
        headline = msmx
        roof = {
            "kernel": "split SpMV, FP64 values and vectors (spmv_sell_cd<3,double,double,true,EpiStore>)",
            "bound": "hbm", "achieved": round(sp["f64"]["gbps"], 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(sp["f64"]["gbps"] / PEAK_HBM_GBPS, 4), "traffic": None,
            "avg_launch_ms": round(sp["f64"]["avg_ms"], 5), "bytes_per_launch": sp["f64"]["bytes"],
        }
        kern = max(profmx.items(), key=lambda kv: kv[1]["total_ms"])
        dominant = {"kernel": f"{kern[0]} (largest share of the headline solve)",
                    "achieved_gbps": round(kern[1]["gbps"], 1), "avg_launch_ms": round(kern[1]["avg_ms"], 5),
                    "share_ms": round(kern[1]["total_ms"] / max(args.steps, 1), 3)}
        cpu = None
        if not args.no_cpu:
            spmv_s, setup_s, iter_s = cpu_baseline(s, res64[-1]["iterations"])
            model, cores = cpu_info()
            est_ms = iter_s * res64[-1]["iterations"] * 1e3
            cpu = {"value": round(est_ms, 1), "unit": "ms", "cores": 1, "kind": "port",
                   "sample": (f"oracle FP64 BiCGSTAB + 4096-block ILU(0) on this grid, 3 timed iterations "
                              f"({iter_s * 1e3:.0f} ms each) x the {res64[-1]['iterations']} iterations the "
                              f"solve needs; ILU set-up {setup_s:.1f}s excluded; split SpMV "
                              f"{m.info().spmv_bytes_f64 / spmv_s / 1e9:.1f} GB/s; host {model}, "
                              f"{cores} hardware threads")}
        out = {
            "metric": METRIC, "value": round(headline, 3), "unit": "ms", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(headline, 3),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 inner / f64 refinement", "data": "synthetic",
            "config": {"workload": "2048x1024 triangulated grid, 2,097,152 vertices, block 3, split view, "
                                   "Hilbert order, seed 0; rtol 1e-8; block-Jacobi ILU(0) on 4096-vertex blocks",
                       "headline": "FP64-refined FGMRES(30), FP32 matrix + FP32 ILU(0) storage",
                       "baseline": "FP64 BiCGSTAB + FP64 ILU(0)"},
            "fp64_bicgstab_ilu0_ms": round(ms64, 3), "speedup_vs_fp64": round(ms64 / headline, 3),
            "fp64_iterations": res64[-1]["iterations"], "fp64_true_rel": rel64,
            "mixed_inner_iterations": resmx[-1]["inner_iterations"],
            "mixed_refinements": resmx[-1]["refinement_steps"], "mixed_true_rel": relmx,
            "mixed_error_vs_fp64": err,
            "spmv": sp, "roofline": roof, "dominant_kernel": dominant, "cpu_baseline": cpu,
            "kernels": {"fp64": prof64, "mixed": profmx},
        }
        print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
