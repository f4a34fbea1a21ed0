"""Learned vertex-block preconditioner: training-data generation, training and
evaluation on the device (the network's forward, backward and Adam kernels
are the HIP ones behind owtk_gnn_*).

Training data.  For each random instance (a triangulated grid with a
lognormal coefficient field) the solver produces its own (residual, error)
pairs: x* is solved to rtol 1e-11, then FGMRES(30) preconditioned by the
untrained model (the plain truncated Neumann series) is stopped after m
iterations for a ladder of m, and each stop gives r = b - A x_m and
e = x* - x_m.  Two probe pairs (e uniform in [-1, 1], r = A e) cover the
high-frequency end of the spectrum.  All pairs stay resident on the device
together with their basis vectors u_k(r).

    python -m owtk.learn train --out model.npz [--instances 200 --nx 512 --ny 256 --sigma 1.5]
    python -m owtk.learn eval --model model.npz [--fields 20 --nx 1024 --ny 512]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import owtk

LADDER = (0, 2, 5, 10, 20, 35, 60, 100)
PROBES = 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def instance(ctx, nx, ny, seed, sigma, correlation=16.0, block=3, law=owtk.LAW_LOGNORMAL_FIELD,
             ordering=owtk.ORDER_HILBERT, low_storage=(owtk.F32,)):
    s = owtk.generate_split(nx, ny, block, seed, ordering=ordering, law=law, sigma=sigma, correlation=correlation)
    m = owtk.Matrix.split(ctx, s.nodes, 0, block, s.offsets, s.columns, s.diagonal, s.off_diagonal)
    for st in low_storage:
        m.add_storage(st)
    return s, m


def solver_pairs(ctx, s, m, degree, ladder=LADDER, probes=PROBES, seed=0):
    """(residual, error) pairs produced by the solver on one instance, as two
    device FP64 buffers of pairs x n."""
    n = s.nodes * s.block
    b = ctx.upload(s.rhs)
    xs = ctx.empty(n)
    jac = owtk.Preconditioner(m, owtk.PC_JACOBI)
    res = owtk.bicgstab(m, jac, b, xs, owtk.SolverOptions(relative_tolerance=1e-11, maximum_iterations=20000))
    if res["status_name"] != "converged":
        raise RuntimeError(f"reference solve did not converge: {res}")
    x_star = xs.download()
    fresh = owtk.GnnModel(ctx, s.block, degree)
    pc = owtk.Preconditioner(m, owtk.PC_LEARNED, model=fresh)
    x = ctx.empty(n)
    r = ctx.empty(n)
    R, E = [], []
    for steps in ladder:
        x.zero()
        if steps > 0:
            owtk.fgmres(m, pc, b, x, owtk.SolverOptions(relative_tolerance=1e-30, restart=30, maximum_iterations=steps))
        m.residual(b, x, r)
        R.append(r.download())
        E.append(x_star - x.download())
    rng = np.random.default_rng(seed)
    e = ctx.empty(n)
    for _ in range(probes):
        ev = rng.uniform(-1.0, 1.0, n)
        e.upload(ev)
        m.spmv(e, r)
        R.append(r.download())
        E.append(ev)
    keep = [k for k in range(len(R)) if np.linalg.norm(E[k]) > 1e-10 * np.linalg.norm(x_star)]
    R = np.concatenate([R[k] for k in keep])
    E = np.concatenate([E[k] for k in keep])
    return ctx.upload(R), ctx.upload(E), len(keep)


def build_problems(ctx, model, count, nx, ny, sigma, seed0, correlation=16.0, law=owtk.LAW_LOGNORMAL_FIELD):
    problems = []
    t0 = time.perf_counter()
    for k in range(count):
        s, m = instance(ctx, nx, ny, seed0 + k, sigma, correlation, law=law, low_storage=())
        R, E, P = solver_pairs(ctx, s, m, model.degree, seed=seed0 + k)
        problems.append(model.problem(m, R, E, P))
        del R, E, m
        if (k + 1) % 10 == 0 or k + 1 == count:
            log(f"[learn] training data {k + 1}/{count} instances ({time.perf_counter() - t0:.1f}s)")
    return problems


def train(model, problems, epochs=30, lr=3e-3, batch=4, seed=0, lr_final=3e-4, select=None, select_every=5):
    """Adam with a cosine learning-rate decay.  With `select` (a function of
    the model returning a score, lower is better, e.g. the mean FGMRES
    iteration count on held-out systems), the parameters are scored every
    `select_every` epochs and at the end, and the best scored parameters are
    restored on return (the training loss is only a proxy for the solver's
    convergence).  Returns the loss history and the selection record."""
    rng = np.random.default_rng(seed)
    history, record = [], []
    best = None
    steps = epochs * ((len(problems) + batch - 1) // batch)
    step = 0
    t0 = time.perf_counter()

    def score(epoch):
        nonlocal best
        if select is None:
            return
        val = float(select(model))
        record.append({"epoch": epoch, "score": val})
        log(f"[learn] selection score after epoch {epoch}: {val:.2f}")
        if best is None or val < best[0]:
            best = (val, model.get_parameters().copy(), epoch)

    score(0)
    for epoch in range(epochs):
        order = rng.permutation(len(problems))
        total = 0.0
        for i in range(0, len(order), batch):
            idx = order[i:i + batch]
            model.zero_grad()
            for j in idx:
                model.accumulate(problems[j], 1.0 / len(idx))
            # cosine decay from lr to lr_final
            cur = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * step / max(steps - 1, 1)))
            total += model.adam_step(cur) * len(idx)
            step += 1
        history.append(total / len(problems))
        log(f"[learn] epoch {epoch + 1}/{epochs}: loss {history[-1]:.5f} ({time.perf_counter() - t0:.1f}s)")
        if (epoch + 1) % select_every == 0 or epoch + 1 == epochs:
            score(epoch + 1)
    if best is not None:
        model.set_parameters(best[1])
        record.append({"selected_epoch": best[2], "score": best[0]})
    return history, record


def iteration_score(ctx, systems, restart=30, rtol=1e-8):
    """Mean FGMRES(restart) iteration count to rtol over (system, matrix)
    pairs with the model's preconditioner: the selection score of train()."""
    def score(model):
        its = []
        for s, m in systems:
            b = ctx.upload(s.rhs)
            x = ctx.empty(s.nodes * s.block)
            pc = owtk.Preconditioner(m, owtk.PC_LEARNED, model=model)
            res, _ = fgmres_run(ctx, m, pc, b, x, restart, rtol)
            its.append(res["iterations"] if res["status_name"] == "converged" else 10 * res["iterations"])
            pc.close()
        return float(np.mean(its))
    return score


def evaluation_loss(model, problems):
    model.zero_grad()
    for p in problems:
        model.accumulate(p, 1.0 / len(problems))
    loss, _ = model.read()
    model.zero_grad()
    return loss


def fgmres_run(ctx, m, pc, b, x, restart=30, rtol=1e-8, maximum_iterations=3000):
    x.zero()
    ctx.synchronize()
    t0 = time.perf_counter()
    res = owtk.fgmres(m, pc, b, x, owtk.SolverOptions(relative_tolerance=rtol, restart=restart,
                                                      maximum_iterations=maximum_iterations))
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    return res, ms


def compare(ctx, model, s, m, restart=30, rtol=1e-8, ilu_block=4096):
    """FGMRES(restart) to rtol with the learned network; the network's
    coefficients averaged over the vertices (one polynomial: no per-vertex
    modulation); the learned global polynomial without the network (decoder
    zeroed: gamma = g); the plain truncated series (untrained model: gamma =
    1); Jacobi and block-Jacobi ILU(0)."""
    n = s.nodes * s.block
    b = ctx.upload(s.rhs)
    x = ctx.empty(n)
    r = ctx.empty(n)
    bn = ctx.norm(b)
    fresh = owtk.GnnModel(ctx, s.block, model.degree)
    polynomial = model.without_network()
    average = model.constant_polynomial(m)
    out = {}
    for name, make in (("learned", lambda: owtk.Preconditioner(m, owtk.PC_LEARNED, model=model)),
                       ("vertex_average", lambda: owtk.Preconditioner(m, owtk.PC_LEARNED, model=average)),
                       ("no_network", lambda: owtk.Preconditioner(m, owtk.PC_LEARNED, model=polynomial)),
                       ("neumann", lambda: owtk.Preconditioner(m, owtk.PC_LEARNED, model=fresh)),
                       ("jacobi", lambda: owtk.Preconditioner(m, owtk.PC_JACOBI)),
                       ("ilu0", lambda: owtk.Preconditioner(m, owtk.PC_ILU, block_nodes=ilu_block))):
        pc = make()
        fgmres_run(ctx, m, pc, b, x, restart, rtol)  # warm-up
        res, ms = fgmres_run(ctx, m, pc, b, x, restart, rtol)
        m.residual(b, x, r)
        out[name] = {"iterations": res["iterations"], "status": res["status_name"], "ms": round(ms, 3),
                     "true_rel": ctx.norm(r) / bn}
    return out


def gamma_spread(model, m):
    """Per (k, c): standard deviation of gamma over the vertices relative to
    its mean magnitude (0: the network's output is one polynomial)."""
    gam = model.coefficients(m).astype(np.float64)
    return [[float(gam[k, :, c].std() / max(abs(gam[k, :, c].mean()), 1e-30)) for c in range(gam.shape[2])]
            for k in range(gam.shape[0])]


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m owtk.learn")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--instances", type=int, default=200)
    t.add_argument("--nx", type=int, default=512)
    t.add_argument("--ny", type=int, default=256)
    t.add_argument("--sigma", type=float, default=1.5)
    t.add_argument("--law", choices=["lognormal", "uniform"], default="lognormal")
    t.add_argument("--degree", type=int, default=4)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--batch", type=int, default=4)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--seed", type=int, default=1000)
    t.add_argument("--init", default=None, help="start from these parameters")
    t.add_argument("--select-fields", type=int, default=0,
                   help="score checkpoints by FGMRES iterations on this many held-out systems")
    t.add_argument("--select-nx", type=int, default=1024)
    t.add_argument("--select-ny", type=int, default=512)
    t.add_argument("--select-every", type=int, default=5)
    e = sub.add_parser("eval")
    e.add_argument("--model", required=True)
    e.add_argument("--fields", type=int, default=20)
    e.add_argument("--nx", type=int, default=1024)
    e.add_argument("--ny", type=int, default=512)
    e.add_argument("--sigma", type=float, default=1.5)
    e.add_argument("--law", choices=["lognormal", "uniform"], default="lognormal")
    e.add_argument("--seed", type=int, default=2000)
    args = ap.parse_args(argv)
    law = owtk.LAW_LOGNORMAL_FIELD if args.law == "lognormal" else owtk.LAW_UNIFORM
    with owtk.Context(0) as ctx:
        if args.cmd == "train":
            model = owtk.GnnModel(ctx, 3, args.degree)
            if args.init:
                model.set_parameters(owtk.GnnModel.load(ctx, args.init).get_parameters())
            problems = build_problems(ctx, model, args.instances, args.nx, args.ny, args.sigma, args.seed, law=law)
            held = problems[-max(1, len(problems) // 10):]
            train_set = problems[:len(problems) - len(held)] or problems
            before = evaluation_loss(model, held)
            select = None
            if args.select_fields > 0:
                # held-out systems at a larger size, seeds apart from training and evaluation
                systems = [instance(ctx, args.select_nx, args.select_ny, args.seed + 50_000 + k, args.sigma, law=law)
                           for k in range(args.select_fields)]
                select = iteration_score(ctx, systems)
            hist, record = train(model, train_set, args.epochs, args.lr, args.batch, seed=args.seed, select=select,
                                 select_every=args.select_every)
            after = evaluation_loss(model, held)
            model.save(args.out)
            print(json.dumps({"trained": args.out, "instances": args.instances, "grid": [args.nx, args.ny],
                              "law": args.law, "sigma": args.sigma, "degree": args.degree, "epochs": args.epochs,
                              "batch": args.batch, "lr": args.lr, "seed": args.seed, "init": args.init,
                              "selection": {"fields": args.select_fields, "grid": [args.select_nx, args.select_ny],
                                            "record": record},
                              "held_out_loss_before": before, "held_out_loss_after": after,
                              "held_out_loss_without_network": evaluation_loss(model.without_network(), held),
                              "train_loss": hist}))
        else:
            model = owtk.GnnModel.load(ctx, args.model)
            rows = []
            for k in range(args.fields):
                s, m = instance(ctx, args.nx, args.ny, args.seed + k, args.sigma, law=law)
                row = compare(ctx, model, s, m)
                rows.append(row)
                row["gamma_spread"] = gamma_spread(model, m)
                log(f"[learn] field {k + 1}/{args.fields}: " + ", ".join(
                    f"{name} {v['iterations']} it {v['ms']:.1f} ms" for name, v in row.items() if name != "gamma_spread"))
            methods = [k for k in rows[0] if k != "gamma_spread"]
            summary = {name: {"median_iterations": float(np.median([r[name]["iterations"] for r in rows])),
                              "median_ms": float(np.median([r[name]["ms"] for r in rows])),
                              "max_true_rel": float(max(r[name]["true_rel"] for r in rows)),
                              "converged": sum(r[name]["status"] == "converged" for r in rows)}
                       for name in methods}
            summary["iteration_ratio_vertex_average_over_learned"] = float(np.median(
                [r["vertex_average"]["iterations"] / r["learned"]["iterations"] for r in rows]))
            summary["iteration_ratio_no_network_over_learned"] = float(np.median(
                [r["no_network"]["iterations"] / r["learned"]["iterations"] for r in rows]))
            print(json.dumps({"model": args.model, "fields": args.fields, "grid": [args.nx, args.ny],
                              "summary": summary, "per_field": rows}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
