#!/usr/bin/env python3
"""Benchmark of the wave-equation hot path on one MI355X.

Workload: 1024 x 1024 jittered Delaunay channel (1,048,576 nodes, RCM order), dx = 1 km,
log-normal bathymetry 5-4000 m (correlation 20 dx), 31 sigma levels, dt = 150 s, theta = 0.6.
One step = one FP64 Jacobi PCG solve of the assembled wave-equation system to rtol 1e-12
(checked every 10 iterations), cold start, matrix and right-hand side resident in HBM.

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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--rows-per-thread", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
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
    log(f"[bench] mesh on device: n={mesh.n} ns={mesh.ns} nnz={mesh.nnz} in {time.time()-t1:.1f}s")

    # ---- one full hot-path step to build the system (and time its pieces)
    def timed(fn):
        ctx.synchronize()
        a = time.perf_counter()
        fn()
        ctx.synchronize()
        return (time.perf_counter() - a) * 1e3

    mesh.backtrack(p)  # warm
    ms_bt = min(timed(lambda: mesh.backtrack(p)) for _ in range(3))
    walks = mesh.ns * mesh.nvrt
    ms_rhs = timed(lambda: mesh.rhs_producers(p))
    solver = core.Solver(mesh, L.PREC_FP64)
    solver.set_tuning(args.rows_per_thread)
    ms_asm = min(timed(lambda: solver.assemble(p)) for _ in range(3))
    log(f"[bench] backtrack {ms_bt:.2f} ms ({walks} side-levels), rhs {ms_rhs:.2f} ms, assembly {ms_asm:.2f} ms")

    rtol, maxiter, check_every = 1e-12, 20000, 10
    for w in range(args.warmup):
        solver.reset_history()
        _, it, st = solver.solve(None, rtol, maxiter, check_every, want_x=False)
        log(f"[bench] warmup {w}: {it} iterations, status {st}")

    times, iters = [], []
    ctx.synchronize()
    t_all = time.perf_counter()
    for s in range(args.steps):
        a = time.perf_counter()
        solver.reset_history()
        _, it, st = solver.solve(None, rtol, maxiter, check_every, want_x=False)
        times.append(time.perf_counter() - a)
        iters.append(it)
    ctx.synchronize()
    total = time.perf_counter() - t_all
    solves_per_s = args.steps / total
    it_mean = float(np.mean(iters))
    log(f"[bench] {args.steps} FP64 solves: {solves_per_s:.2f} solves/s, {it_mean:.0f} iterations, "
        f"{1e3*total/args.steps:.2f} ms/solve")

    # ---- per-phase kernel durations (HIP events on the solver stream) for the roofline
    solver.reset_history()
    ms_phase, _ = solver.profile(rtol, maxiter, check_every)
    ms_a, ms_b = float(ms_phase[0]), float(ms_phase[1])
    n, nnz = mesh.n, mesh.nnz
    ip_h, ix_h = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    cdel = ix_h - (np.repeat(np.arange(n, dtype=np.int64), np.diff(ip_h)) & ~63)
    col_bytes = 2 if (cdel.min() >= -32768 and cdel.max() <= 32767) else 4   # the solver's column format
    del cdel
    # phase A compulsory bytes: values + columns, row pointers, read z and p_old, write p_new and q
    bytes_a = (8 + col_bytes) * nnz + 4 * (n + 1) + 32 * n
    bytes_b = 64 * n                               # read x r p q invD, write x r z
    bytes_iter_survey = 12 * nnz + 4 * (n + 1) + 104 * n
    gbs_a = bytes_a / (ms_a * 1e-3) / 1e9
    gbs_b = bytes_b / (ms_b * 1e-3) / 1e9
    ms_iter = 1e3 * total / args.steps / max(it_mean, 1.0)
    gbs_iter = bytes_iter_survey / (ms_iter * 1e-3) / 1e9
    log(f"[bench] phase A {1e3*ms_a:.1f} us ({gbs_a:.0f} GB/s), phase B {1e3*ms_b:.1f} us ({gbs_b:.0f} GB/s), "
        f"iteration {1e3*ms_iter:.1f} us ({gbs_iter:.0f} GB/s of 192n-byte iterations)")

    cp_ms = C_double()
    copy_bytes = 1 << 30
    L.check(L.lib().gsh_copy_probe(ctx.h, copy_bytes, 20, cp_ms))
    copy_gbs = 2 * copy_bytes / (cp_ms.value * 1e-3) / 1e9
    log(f"[bench] copy probe: {copy_gbs:.0f} GB/s (read+write of 1 GiB)")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pcg_phase_a_traffic.json")
    if os.path.exists(tfile):
        try:
            with open(tfile) as f:
                tj = json.load(f)
            if tj.get("n") == n and tj.get("nnz") == nnz:
                traffic = tj.get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    # ---- mixed precision on the same system
    mixed = None
    x64 = None
    if not args.no_mixed:
        solver.reset_history()
        x64, it64, _ = solver.solve(None, rtol, maxiter, check_every)
        ms_solver = core.Solver(mesh, L.PREC_MIXED)
        ms_solver.set_tuning(args.rows_per_thread, 1e-5)
        # the mixed solver assembles the same system from the same mesh state
        ms_solver.assemble(p)
        ms_solver.solve(None, rtol, maxiter, check_every, want_x=False)  # warm
        tm = []
        for _ in range(max(2, args.steps // 2)):
            a = time.perf_counter()
            ms_solver.reset_history()
            xm, itm, stm = ms_solver.solve(None, rtol, maxiter, check_every, want_x=(_ == 0))
            tm.append(time.perf_counter() - a)
            if _ == 0:
                xmix = xm
                st_m = ms_solver.stats()
        err = float(np.linalg.norm(xmix - x64) / np.linalg.norm(x64))
        mixed = {"solves_per_s": round(len(tm) / sum(tm), 3), "inner_iters": int(st_m.iters),
                 "outer": int(st_m.outer), "rel_l2_vs_fp64": err, "fp64_iters": int(it64),
                 "speedup_vs_fp64": round((len(tm) / sum(tm)) / solves_per_s, 3)}
        log(f"[bench] mixed: {mixed}")

    # ---- CPU baseline: the oracle (FP64 C restatement) on the same system
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as O

        ip, ix = ip_h, ix_h
        val = solver.get_values()
        om_b = solver.get_rhs()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        a = time.perf_counter()
        xo, ito, _, _ = O.pcg(ip, ix, val, om_b, rtol=rtol, maxiter=maxiter, check_every=check_every,
                              rows_per_thread=args.rows_per_thread, nthreads=threads)
        cpu_s = time.perf_counter() - a
        cpu = {"value": round(1.0 / cpu_s, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"1 full FP64 Jacobi PCG solve of the same {n}-node system ({ito} iterations), "
                         f"oracle C port, OpenMP, {cpu_model()}",
               "seconds": round(cpu_s, 3)}
        if x64 is not None:
            cpu["gpu_fp64_bitwise_equal"] = bool(np.array_equal(xo, x64))
            cpu["iters_equal"] = bool(ito == it64)
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
        "config": {"workload": f"{spec.name}: FP64 Jacobi PCG rtol 1e-12, cold start", "n": n, "nnz": nnz,
                   "ns": mesh.ns, "nvrt": mesh.nvrt, "iters": it_mean, "rows_per_thread": args.rows_per_thread,
                   "dt": p.dt, "theta": p.theta},
        "roofline": {"bound": "hbm", "kernel": "k_pcg_phase_a", "achieved": round(gbs_a, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(gbs_a / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": bytes_a, "column_bytes": col_bytes, "us_per_launch": round(1e3 * ms_a, 2)},
        "phase_b": {"achieved_gbs": round(gbs_b, 1), "bytes_per_launch": bytes_b, "us_per_launch": round(1e3 * ms_b, 2)},
        "iteration": {"us": round(1e3 * ms_iter, 2), "gbs_192n": round(gbs_iter, 1)},
        "copy_probe_gbs": round(copy_gbs, 1),
        "backtrack": {"ms": round(ms_bt, 3), "side_levels": walks, "per_s": round(walks / (ms_bt * 1e-3), 1)},
        "assembly_ms": round(ms_asm, 3),
        "rhs_ms": round(ms_rhs, 3),
        "mixed": mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


from ctypes import c_double as C_double, c_int as C_int  # noqa: E402


if __name__ == "__main__":
    sys.exit(main())
