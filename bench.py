#!/usr/bin/env python3
"""Benchmark of the wave-equation hot path on one MI355X.

Workload: 1024 x 1024 jittered Delaunay channel (1,048,576 nodes, RCM order), dx = 1 km,
log-normal bathymetry 5-4000 m (correlation 20 dx), 31 sigma levels, dt = 150 s, theta = 0.6,
10 ELM substeps. One step = one FP64 multicolour-SSOR(omega = 1) PCG solve of the assembled
wave-equation system to rtol 1e-12 (checked every 10 iterations), cold start, matrix and
right-hand side resident in HBM. The reference's Jacobi PCG runs on the same system beside it.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed_solves(ctx, solver, steps, rtol, maxiter, check_every):
    iters = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.reset_history()
        _, it, _ = solver.solve(None, rtol, maxiter, check_every, want_x=False)
        iters.append(it)
    ctx.synchronize()
    return time.perf_counter() - t0, float(np.mean(iters))


def ssor_bytes(mesh, solver):
    """Algorithmic bytes per iteration of the SSOR launch kinds (real rows, compulsory traffic):
    backward sweeps read D~, r, p_old and the upper part (8 B value + 4 B column per entry), gather
    t once, write p and t (colour 0 also s and q); forward sweeps (colours 1..C-1) read p, t, D~
    and the lower part, gather s once, write s and q; the update reads y, p, r, q, D~ and writes
    y and r."""
    n, nnz = mesh.n, mesh.nnz
    nc, _, col = solver.ssor_info()
    half = (nnz - n) // 2
    n0 = int(np.count_nonzero(col == 0))
    nf = n - n0
    bwd = 12 * half + 4 * n + 48 * n + 16 * n0
    fwd = 12 * half + 4 * nf + 48 * nf
    upd = 56 * n
    return {"bwd": bwd, "fwd": fwd, "upd": upd, "ncolours": nc, "colour0_rows": n0}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--ssor-rows-per-thread", type=int, default=2)
    ap.add_argument("--jacobi-rows-per-thread", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-jacobi", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0 = all visible cores, max 16)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the solver runs on a single MI355X; multi-GPU sharding is not built)", file=sys.stderr)
        return 2

    from gschism_amd import _lib as L
    from gschism_amd import cases, core

    t0 = time.time()
    spec = cases.spec_gpu_1m(args.nx, args.ny)
    data = cases.make_case(spec)
    log(f"[bench] generated {spec.name}: n={data.xy.shape[0]} ne={data.elnode.shape[0]} in {time.time()-t0:.1f}s")
    ctx = core.Context(0)
    t1 = time.time()
    mesh, p, _ = cases.load_mesh(ctx, data)
    n, nnz = mesh.n, mesh.nnz
    log(f"[bench] mesh on device: n={n} ns={mesh.ns} nnz={nnz} in {time.time()-t1:.1f}s")

    def timed(fn):
        ctx.synchronize()
        a = time.perf_counter()
        fn()
        ctx.synchronize()
        return (time.perf_counter() - a) * 1e3

    # ---- one hot-path step up to the system: backtracking, RHS producers, assembly
    mesh.backtrack(p)  # warm
    ms_bt = min(timed(lambda: mesh.backtrack(p)) for _ in range(3))
    walks = mesh.ns * mesh.nvrt
    ms_rhs = timed(lambda: mesh.rhs_producers(p))
    solver = core.Solver(mesh, L.PREC_FP64, precond=L.PRECOND_SSOR)
    solver.set_tuning(args.ssor_rows_per_thread)
    ms_asm = min(timed(lambda: solver.assemble(p)) for _ in range(3))
    log(f"[bench] backtrack {ms_bt:.2f} ms ({walks} side-levels), rhs {ms_rhs:.2f} ms, assembly {ms_asm:.2f} ms")

    # ---- headline: FP64 multicolour SSOR PCG
    rtol, maxiter, check_every = 1e-12, 20000, 10
    for w in range(args.warmup):
        solver.reset_history()
        _, it, st = solver.solve(None, rtol, maxiter, check_every, want_x=False)
        log(f"[bench] warmup {w}: {it} SSOR iterations, status {st}")
    total, it_mean = timed_solves(ctx, solver, args.steps, rtol, maxiter, check_every)
    solves_per_s = args.steps / total
    log(f"[bench] {args.steps} FP64 SSOR solves: {solves_per_s:.2f} solves/s, {it_mean:.0f} iterations, "
        f"{1e3*total/args.steps:.2f} ms/solve")
    solver.reset_history()
    x_ssor, it_ssor, _ = solver.solve(None, rtol, maxiter, check_every)
    rel_ssor = solver.check()

    # ---- kernel durations per launch kind (kernel-stamped HIP events on the solver stream)
    solver.reset_history()
    ms_kind, _ = solver.profile(rtol, maxiter, check_every)
    kinds = ssor_bytes(mesh, solver)
    nc = kinds["ncolours"]
    per = {}
    for i, (k, kern, launches) in enumerate((("bwd", "k_ssor_bwd", nc), ("fwd", "k_ssor_fwd", nc - 1),
                                              ("upd", "k_ssor_update", 1))):
        ms = float(ms_kind[i])
        per[k] = {"kernel": kern, "launches_per_iter": launches, "us_per_iter": round(1e3 * ms, 2),
                  "bytes_per_iter": kinds[k], "gbs": round(kinds[k] / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
    dom = max(per, key=lambda k: per[k]["us_per_iter"])
    d = per[dom]
    us_launch = d["us_per_iter"] / d["launches_per_iter"]
    log("[bench] SSOR per iteration: " + ", ".join(f"{k} {v['us_per_iter']} us ({v['gbs']} GB/s)"
                                                   for k, v in per.items()))

    cp_ms = C_double()
    copy_bytes = 1 << 30
    L.check(L.lib().gsh_copy_probe(ctx.h, copy_bytes, 20, cp_ms))
    copy_gbs = 2 * copy_bytes / (cp_ms.value * 1e-3) / 1e9
    log(f"[bench] copy probe: {copy_gbs:.0f} GB/s (read+write of 1 GiB)")

    # ---- the reference's Jacobi PCG on the same system
    jac = None
    x_jac = None
    if not args.no_jacobi:
        js = core.Solver(mesh, L.PREC_FP64)
        js.set_tuning(args.jacobi_rows_per_thread)
        js.assemble(p)
        js.reset_history()
        js.solve(None, rtol, maxiter, check_every, want_x=False)
        jt, jit = timed_solves(ctx, js, args.steps, rtol, maxiter, check_every)
        js.reset_history()
        x_jac, _, _ = js.solve(None, rtol, maxiter, check_every)
        js.reset_history()
        ms_ab, _ = js.profile(rtol, maxiter, check_every)
        ip_h, ix_h = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
        cdel = ix_h - (np.repeat(np.arange(n, dtype=np.int64), np.diff(ip_h)) & ~63)
        col_bytes = 2 if (cdel.min() >= -32768 and cdel.max() <= 32767) else 4
        del cdel
        bytes_a = (8 + col_bytes) * nnz + 4 * (n + 1) + 32 * n   # values, columns, row pointers, z, p in; p, q out
        bytes_b = 64 * n                                          # x, p, r, q, invD in; x, r, z out
        jac = {"solves_per_s": round(args.steps / jt, 3), "iters": jit,
               "ms_per_solve": round(1e3 * jt / args.steps, 3),
               "phase_a": {"us": round(1e3 * ms_ab[0], 2), "bytes": bytes_a,
                           "gbs": round(bytes_a / (ms_ab[0] * 1e-3) / 1e9, 1), "column_bytes": col_bytes},
               "phase_b": {"us": round(1e3 * ms_ab[1], 2), "bytes": bytes_b,
                           "gbs": round(bytes_b / (ms_ab[1] * 1e-3) / 1e9, 1)},
               "ssor_speedup": round(solves_per_s / (args.steps / jt), 3),
               "rel_l2_ssor_vs_jacobi": float(np.linalg.norm(x_ssor - x_jac) / np.linalg.norm(x_jac))}
        log(f"[bench] Jacobi: {jac}")

    # ---- mixed precision (FP32 inner Jacobi PCG, FP64 residual and refinement) on the same system
    mixed = None
    if not args.no_mixed and x_jac is not None:
        ms_solver = core.Solver(mesh, L.PREC_MIXED)
        ms_solver.set_tuning(args.jacobi_rows_per_thread, 1e-5)
        ms_solver.assemble(p)
        ms_solver.solve(None, rtol, maxiter, check_every, want_x=False)  # warm
        tm = []
        for i in range(max(2, args.steps // 2)):
            a = time.perf_counter()
            ms_solver.reset_history()
            xm, itm, stm = ms_solver.solve(None, rtol, maxiter, check_every, want_x=(i == 0))
            tm.append(time.perf_counter() - a)
            if i == 0:
                xmix = xm
                st_m = ms_solver.stats()
        mixed = {"solves_per_s": round(len(tm) / sum(tm), 3), "inner_iters": int(st_m.iters),
                 "outer": int(st_m.outer),
                 "rel_l2_vs_fp64": float(np.linalg.norm(xmix - x_jac) / np.linalg.norm(x_jac))}
        log(f"[bench] mixed: {mixed}")

    # ---- CPU baseline: the oracle's FP64 SSOR PCG on the same system, all host cores
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as O

        ip_h, ix_h = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
        val = solver.get_values()
        bvec = solver.get_rhs()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        osr = O.OracleSsor(ip_h, ix_h, args.ssor_rows_per_thread)
        reps, spent, xo, ito = 0, 0.0, None, 0
        while reps < 10 and (spent < 10.0 or reps < 2):
            a = time.perf_counter()
            xo, ito, _, _ = osr.solve(val, bvec, rtol=rtol, maxiter=maxiter, check_every=check_every,
                                      nthreads=threads)
            spent += time.perf_counter() - a
            reps += 1
        osr.close()
        cpu = {"value": round(reps / spent, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"{reps} full FP64 SSOR PCG solves of the same {n}-node system ({ito} iterations each), "
                         f"oracle C restatement, OpenMP over rows, {cpu_model()}",
               "seconds": round(spent, 3),
               "gpu_bitwise_equal": bool(np.array_equal(xo, x_ssor)), "iters_equal": bool(ito == it_ssor)}
        log(f"[bench] cpu baseline: {cpu}")

    line = {
        "metric": "PCG solves/s + HBM GB/s at 1M nodes; mixed vs FP64; served p50/p99 latency",
        "value": round(solves_per_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{spec.name}: FP64 multicolour SSOR(1.0) PCG rtol 1e-12, cold start",
                   "n": n, "nnz": nnz, "ns": mesh.ns, "nvrt": mesh.nvrt, "iters": it_mean, "colours": nc,
                   "rows_per_thread": args.ssor_rows_per_thread, "dt": p.dt, "theta": p.theta,
                   "true_rel_residual": rel_ssor},
        "roofline": {"bound": "hbm", "kernel": d["kernel"], "achieved": d["gbs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(d["gbs"] / PEAK_HBM_GBS, 4), "traffic": None,
                     "bytes_per_launch": round(d["bytes_per_iter"] / d["launches_per_iter"]),
                     "us_per_launch": round(us_launch, 2)},
        "ssor_kernels": per,
        "copy_probe_gbs": round(copy_gbs, 1),
        "jacobi": jac,
        "backtrack": {"ms": round(ms_bt, 3), "side_levels": walks, "per_s": round(walks / (ms_bt * 1e-3), 1)},
        "assembly_ms": round(ms_asm, 3),
        "rhs_ms": round(ms_rhs, 3),
        "mixed": mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


from ctypes import c_double as C_double  # noqa: E402


if __name__ == "__main__":
    sys.exit(main())
