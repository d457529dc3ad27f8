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
PEAK_FP64_TFLOPS = 78.6  # MI355X vector FP64


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


def counter_traffic(kernel: str, rows_per_thread: int):
    """HBM bytes per launch of `kernel` at this rows-per-thread setting, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (profiles/), or None. The summary
    holds one line 'traffic <kernel> rows_per_thread=<R> bytes_per_launch=<B>' per setting,
    already scaled by the copy-kernel factors it states."""
    path = os.path.join(ROOT, "profiles", "ssor_counters_channel-1024x1024.md")
    try:
        with open(path) as f:
            for line in f:
                w = line.strip().strip("`").split()
                if len(w) == 4 and w[0] == "traffic" and w[1] == kernel and w[2] == f"rows_per_thread={rows_per_thread}":
                    return int(w[3].split("=", 1)[1])
    except OSError:
        pass
    return None


def backtrack_bound():
    """Per-launch FP64 operations, HBM bytes, VALU busy fraction and occupancy of k_backtrack on
    this workload, from the committed rocprofv3 counter summary (profiles/), or None. The summary
    holds one line 'bound k_backtrack fp64_ops_per_launch=<F> hbm_bytes_per_launch=<B>
    valu_busy=<V> waves_per_simd=<W>'."""
    path = os.path.join(ROOT, "profiles", "backtrack_counters_channel-1024x1024.md")
    try:
        with open(path) as f:
            for line in f:
                w = line.strip().strip("`").split()
                if len(w) == 6 and w[0] == "bound" and w[1] == "k_backtrack":
                    return {k: float(v) for k, v in (x.split("=", 1) for x in w[2:])}
    except OSError:
        pass
    return None


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
    """Algorithmic bytes per iteration of the two SSOR launch kinds (real rows, compulsory
    traffic; 8 B value + 4 B column per off-diagonal entry, each half of them in one sweep).
    Backward sweeps read D~, r, p, q, s and write p, q, r for every row, read and write y every
    other iteration (y takes its updates in pairs: 8 B per row on average), gather the t of
    colours 1..C-1 once and write it; colour 0 writes u and s instead of t. Forward sweeps
    (colours 1..C-1) read D~, t, r and write s for their rows, gather the u of colours 0..C-2
    once and write u for colours 1..C-2 (no higher colour gathers the last colour's u)."""
    n, nnz = mesh.n, mesh.nnz
    nc, _, col = solver.ssor_info()
    half = (nnz - n) // 2
    n0 = int(np.count_nonzero(col == 0))
    nl = int(np.count_nonzero(col == nc - 1)) if nc > 1 else 0
    nf = n - n0
    bwd = 12 * half + 64 * n + 8 * n + 16 * (n - n0) + 16 * n0
    fwd = 12 * half + 32 * nf + 8 * (n - nl) + 8 * (nf - nl) if nc > 1 else 0
    return {"bwd": bwd, "fwd": fwd, "ncolours": nc, "colour0_rows": n0}


def batched_leg(ctx, L, cases, core, members, nx, ny, rtol, mixed_rtol, check_every):
    """Ensemble: `members` systems on one nx x ny mesh, bathymetry h (1 + 0.05 N(0,1)) per member
    (seeds 100 + m), one random right-hand side each (seeds 200 + m). Plain = one FP64 Jacobi solve
    after another; batched = all members in one lock-step launch per PCG phase; block = all members
    in one member-interleaved Jacobi PCG over the shared CSR pattern (gsh_block); mixed = FP32 inner
    iterations with FP64 residual and refinement."""
    spec = cases.CaseSpec(f"ensemble-{nx}x{ny}", nx, ny, 1000.0, nvrt=11, depth=None, substeps=10)
    data = cases.make_case(spec)
    mesh, p, _ = cases.load_mesh(ctx, data)
    depth0 = mesh.topology(L.T_DEPTH)
    s64, smx = [], []
    for m in range(members):
        mesh.set(L.F_DEPTH, depth0 * (1.0 + 0.05 * cases.normal(100 + m, mesh.n)))
        mesh.update_levels(p)
        mesh.backtrack(p)
        mesh.rhs_producers(p)
        b = cases.normal(200 + m, mesh.n)
        a = core.Solver(mesh)
        a.assemble(p)
        a.set_rhs(b)
        c = core.Solver(mesh, L.PREC_MIXED)
        c.set_values(a.get_values())
        c.set_rhs(b)
        s64.append(a)
        smx.append(c)
    mesh.set(L.F_DEPTH, depth0)
    mesh.update_levels(p)

    def seq(solvers, tol):
        for s in solvers:
            s.reset_history()
        ctx.synchronize()
        t = time.perf_counter()
        its = [s.solve(None, tol, 20000, check_every, want_x=False)[1] for s in solvers]
        ctx.synchronize()
        return time.perf_counter() - t, its

    def bat(solvers, tol):
        for s in solvers:
            s.reset_history()
        ctx.synchronize()
        t = time.perf_counter()
        its = core.batch_solve(solvers, tol, 20000, check_every)
        ctx.synchronize()
        return time.perf_counter() - t, its

    seq(s64[:2], rtol)  # warm
    t_plain, it_plain = seq(s64, rtol)
    x64 = [s.get_solution() for s in s64]
    t_b64, it_b64 = bat(s64, rtol)
    same = all(np.array_equal(s.get_solution(), x) for s, x in zip(s64, x64))
    bat(smx[:2], mixed_rtol)  # warm
    t_bmx, it_bmx = bat(smx, mixed_rtol)
    err = max(float(np.linalg.norm(s.get_solution() - x) / np.linalg.norm(x)) for s, x in zip(smx, x64))
    t_smx, _ = seq(smx, mixed_rtol)
    # multicolour SSOR (the 1M headline's method): plain FP64 one member after another on the
    # first 8 members (the per-solve rate does not depend on how many are timed), then all members
    # batched in FP64 and in mixed precision (FP32 SSOR inner solves)
    k_ssor = min(8, members)
    sss, ssm = [], []
    for a in s64:
        for prec, lst in ((L.PREC_FP64, sss), (L.PREC_MIXED, ssm)):
            d = core.Solver(mesh, prec, precond=L.PRECOND_SSOR)
            d.set_tuning(2)
            d.set_values(a.get_values())
            d.set_rhs(a.get_rhs())
            lst.append(d)
    seq(sss[:1], rtol)  # warm
    t_ssor, it_ssor = seq(sss[:k_ssor], rtol)
    ssor_rate = k_ssor / t_ssor
    xs = [d.get_solution() for d in sss[:k_ssor]]
    bat(sss[:2], rtol)  # warm
    t_bss, it_bss = bat(sss, rtol)
    same_ssor = all(np.array_equal(d.get_solution(), x) for d, x in zip(sss, xs)) and it_bss[:k_ssor] == it_ssor
    bat(ssm[:2], mixed_rtol)  # warm
    t_bsm, it_bsm = bat(ssm, mixed_rtol)
    err_ssm = max(float(np.linalg.norm(d.get_solution() - x) / np.linalg.norm(x)) for d, x in zip(ssm, x64))
    for d in sss + ssm:
        d.close()
    # the multi-RHS block: all members in one member-interleaved Jacobi PCG over the shared pattern
    blk = core.Block(mesh, members)
    for m, a in enumerate(s64):
        blk.set_member(m, a.get_values(), a.get_rhs())
    blk.solve(rtol, 20000, check_every)  # warm
    ctx.synchronize()
    t = time.perf_counter()
    it_blk, st_blk = blk.solve(rtol, 20000, check_every)
    ctx.synchronize()
    t_blk = time.perf_counter() - t
    err_blk = max(float(np.linalg.norm(blk.solution(m) - x) / np.linalg.norm(x)) for m, x in enumerate(x64))
    blk.close()
    out = {"workload": f"{members} members on {spec.name} (n={mesh.n}), per-member bathymetry and RHS",
           "block_fp64_solves_per_s": round(members / t_blk, 3),
           "block_fp64_vs_plain": round(t_plain / t_blk, 3),
           "block_fp64_vs_batched_fp64": round(t_b64 / t_blk, 3),
           "block_iters_equal_to_plain": bool(list(map(int, it_blk)) == list(map(int, it_plain))),
           "block_max_rel_l2_vs_plain": err_blk,
           "plain_fp64_solves_per_s": round(members / t_plain, 3),
           "plain_fp64_ssor_solves_per_s": round(ssor_rate, 3),
           "batched_fp64_solves_per_s": round(members / t_b64, 3),
           "mixed_solves_per_s": round(members / t_smx, 3),
           "batched_mixed_solves_per_s": round(members / t_bmx, 3),
           "batched_mixed_vs_plain": round(t_plain / t_bmx, 3),
           "batched_mixed_vs_plain_ssor": round(members / t_bmx / ssor_rate, 3),
           "batched_fp64_ssor_solves_per_s": round(members / t_bss, 3),
           "batched_mixed_ssor_solves_per_s": round(members / t_bsm, 3),
           "batched_mixed_ssor_vs_plain_ssor": round(members / t_bsm / ssor_rate, 3),
           "fp64_iters_mean": float(np.mean(it_plain)), "mixed_inner_iters_mean": float(np.mean(it_bmx)),
           "ssor_iters_mean": float(np.mean(it_bss)), "mixed_ssor_inner_iters_mean": float(np.mean(it_bsm)),
           "batched_fp64_bitwise_equal_to_plain": bool(same and it_b64 == it_plain),
           "batched_fp64_ssor_bitwise_equal_to_plain": bool(same_ssor),
           "mixed_rtol": mixed_rtol, "mixed_max_rel_l2_vs_fp64": err, "mixed_ssor_max_rel_l2_vs_fp64": err_ssm}
    for s in s64 + smx:
        s.close()
    mesh.close()
    return out


def mixed_4m_leg(ctx, L, cases, core, steps, rtol, outer_rtol, check_every):
    """Configuration 3: the 2048 x 2048 log-normal channel (4.2M nodes). FP64 PCG (multicolour
    SSOR and the reference's Jacobi) to rtol 1e-12, FP64 SSOR also to the outer tolerance, against
    mixed precision (FP32 inner SSOR or Jacobi PCG, FP64 dot products, true residual and
    refinement) to the outer tolerance; relative L2 error against the FP64 SSOR solution (rtol
    1e-12) and iteration counts, and mixed SSOR over FP64 SSOR at the same tolerance."""
    spec = cases.spec_mixed_4m()
    data = cases.make_case(spec)
    mesh, p, _ = cases.load_mesh(ctx, data)
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    out = {"workload": f"{spec.name} (n={mesh.n}, nnz={mesh.nnz})", "outer_rtol": outer_rtol}
    ref = None
    for name, prec, pc, tol in (("fp64_ssor", L.PREC_FP64, L.PRECOND_SSOR, rtol),
                                ("fp64_jacobi", L.PREC_FP64, L.PRECOND_JACOBI, rtol),
                                ("fp64_ssor_outer_rtol", L.PREC_FP64, L.PRECOND_SSOR, outer_rtol),
                                ("mixed_ssor", L.PREC_MIXED, L.PRECOND_SSOR, outer_rtol),
                                ("mixed_jacobi", L.PREC_MIXED, L.PRECOND_JACOBI, outer_rtol)):
        s = core.Solver(mesh, prec, precond=pc)
        s.set_tuning(2 if pc == L.PRECOND_SSOR else 4)
        s.assemble(p)
        s.reset_history()
        x, it, st = s.solve(None, tol, 40000, check_every)
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            s.reset_history()
            s.solve(None, tol, 40000, check_every, want_x=False)
        ctx.synchronize()
        dt = (time.perf_counter() - t) / steps
        rec = {"solves_per_s": round(1.0 / dt, 3), "ms_per_solve": round(1e3 * dt, 3), "iters": it, "status": st}
        if prec == L.PREC_MIXED:
            stt = s.stats()
            rec["outer"] = int(stt.outer)
        if ref is None:
            ref = x
        else:
            rec["rel_l2_vs_fp64_ssor"] = float(np.linalg.norm(x - ref) / np.linalg.norm(ref))
        out[name] = rec
        s.close()
    # mixed precision against FP64 SSOR stopped at the same (outer) tolerance
    if out.get("fp64_ssor_outer_rtol", {}).get("solves_per_s"):
        out["mixed_ssor_vs_fp64_ssor_same_rtol"] = round(
            out["mixed_ssor"]["solves_per_s"] / out["fp64_ssor_outer_rtol"]["solves_per_s"], 3)
    mesh.close()
    return out


def many_meshes_leg(ctx, L, cases, core, count, rtol, mixed_rtol, check_every):
    """Configuration 4B: `count` independent meshes with n log-uniform in [5e3, 2e5] (sizes from
    seed 7, bathymetry seeds 3 + k), one system each. Plain = one FP64 Jacobi solve after
    another; batched = all systems in one launch per PCG phase (FP64, then mixed)."""
    from math import erfc, sqrt

    z = cases.normal(cases.SEED_ARRIVALS + 1000, count)
    sizes = [int(round(5e3 * 40 ** (0.5 * erfc(-v / sqrt(2.0))))) for v in z]
    meshes, s64, smx = [], [], []
    for k, n in enumerate(sizes):
        nx = max(8, int(round(sqrt(2 * n))))
        ny = max(8, n // nx)
        spec = cases.CaseSpec(f"m{k}-{nx}x{ny}", nx, ny, 1000.0, nvrt=2, depth=None, substeps=10)
        data = cases.make_case(spec)
        data.depth = cases.bathymetry(nx, ny, 20.0, 5.0, 4000.0, seed=3 + k)
        mesh, p, _ = cases.load_mesh(ctx, data)
        mesh.backtrack(p)
        mesh.rhs_producers(p)
        a = core.Solver(mesh)
        a.assemble(p)
        c = core.Solver(mesh, L.PREC_MIXED)
        c.set_values(a.get_values())
        c.set_rhs(a.get_rhs())
        meshes.append(mesh)
        s64.append(a)
        smx.append(c)

    def run(solvers, tol, batched):
        for s in solvers:
            s.reset_history()
        ctx.synchronize()
        t = time.perf_counter()
        if batched:
            core.batch_solve(solvers, tol, 20000, check_every)
        else:
            for s in solvers:
                s.solve(None, tol, 20000, check_every, want_x=False)
        ctx.synchronize()
        return time.perf_counter() - t

    run(s64[:4], rtol, False)
    t_plain = run(s64, rtol, False)
    x64 = [s.get_solution() for s in s64]
    run(s64[:4], rtol, True)
    t_b64 = run(s64, rtol, True)
    same = all(np.array_equal(s.get_solution(), x) for s, x in zip(s64, x64))
    run(smx[:4], mixed_rtol, True)
    t_bmx = run(smx, mixed_rtol, True)
    err = max(float(np.linalg.norm(s.get_solution() - x) / np.linalg.norm(x)) for s, x in zip(smx, x64))
    # multicolour SSOR (the 1M headline's method): plain FP64 one mesh after another, then all
    # meshes batched in FP64 and in mixed precision (FP32 SSOR inner solves)
    sss, ssm = [], []
    for m, a in zip(meshes, s64):
        for prec, lst in ((L.PREC_FP64, sss), (L.PREC_MIXED, ssm)):
            d = core.Solver(m, prec, precond=L.PRECOND_SSOR)
            d.set_tuning(2)
            d.set_values(a.get_values())
            d.set_rhs(a.get_rhs())
            lst.append(d)
    run(sss[:4], rtol, False)
    t_ssor = run(sss, rtol, False)
    xs = [d.get_solution() for d in sss]
    run(sss[:4], rtol, True)
    t_bss = run(sss, rtol, True)
    same_ssor = all(np.array_equal(d.get_solution(), x) for d, x in zip(sss, xs))
    run(ssm[:4], mixed_rtol, True)
    t_bsm = run(ssm, mixed_rtol, True)
    err_ssm = max(float(np.linalg.norm(d.get_solution() - x) / np.linalg.norm(x)) for d, x in zip(ssm, x64))
    for d in sss + ssm:
        d.close()
    out = {"workload": f"{count} meshes, n log-uniform in [5e3, 2e5] (total {sum(m.n for m in meshes)} nodes)",
           "plain_fp64_solves_per_s": round(count / t_plain, 3),
           "plain_fp64_ssor_solves_per_s": round(count / t_ssor, 3),
           "batched_fp64_solves_per_s": round(count / t_b64, 3),
           "batched_mixed_solves_per_s": round(count / t_bmx, 3), "batched_mixed_vs_plain": round(t_plain / t_bmx, 3),
           "batched_mixed_vs_plain_ssor": round(t_ssor / t_bmx, 3),
           "batched_fp64_ssor_solves_per_s": round(count / t_bss, 3),
           "batched_mixed_ssor_solves_per_s": round(count / t_bsm, 3),
           "batched_mixed_ssor_vs_plain_ssor": round(t_ssor / t_bsm, 3),
           "batched_fp64_bitwise_equal_to_plain": bool(same),
           "batched_fp64_ssor_bitwise_equal_to_plain": bool(same_ssor), "mixed_rtol": mixed_rtol,
           "mixed_max_rel_l2_vs_fp64": err, "mixed_ssor_max_rel_l2_vs_fp64": err_ssm}
    for s in s64 + smx:
        s.close()
    for m in meshes:
        m.close()
    return out


def served_leg(ctx, L, cases, core, n_req, rates, rtol, mixed_rtol, check_every):
    """Served solves: a pool of 6 meshes with n log-uniform in [1e4, 2e6] (random bathymetry,
    seeds 3 + k), requests drawn uniformly from the pool with Poisson arrivals (seed 7) and
    right-hand sides N(0,1) (seeds 1000 + j); each request backtracks on its mesh and solves.
    Plain = the same requests one after another in FP64 (Jacobi, then multicolour SSOR); served =
    the continuous-batching server in mixed precision at each arrival rate, with FP32 Jacobi
    ("rates") and FP32 multicolour SSOR ("ssor_rates") inner solves."""
    from math import erfc, log, sqrt

    sizes = [int(round(1e4 * 200 ** (k / 5))) for k in range(6)]
    pool = []
    for k, n in enumerate(sizes):
        nx = int(round(sqrt(2 * n)))
        ny = max(3, n // nx)
        spec = cases.CaseSpec(f"pool{k}-{nx}x{ny}", nx, ny, 1000.0, nvrt=11, depth=None, substeps=10,
                              corr_cells=20.0)
        data = cases.make_case(spec)
        data.depth = cases.bathymetry(nx, ny, 20.0, 5.0, 4000.0, seed=3 + k)
        mesh, _, _ = cases.load_mesh(ctx, data)
        pool.append(mesh)
    pd = L.default_step_params()
    z = cases.normal(cases.SEED_ARRIVALS, 2 * n_req)
    u = np.array([0.5 * erfc(-v / sqrt(2.0)) for v in z])
    which = np.minimum((u[:n_req] * len(pool)).astype(int), len(pool) - 1)
    expo = np.array([-log(max(1e-300, 1.0 - v)) for v in u[n_req:]])
    rhs = [cases.normal(1000 + j, pool[k].n) for j, k in enumerate(which)]
    # plain: one FP64 request after another
    plain = [core.Solver(m) for m in pool]
    for rep in range(2):  # the first pass warms the solvers
        ctx.synchronize()
        t = time.perf_counter()
        x64 = []
        for j, k in enumerate(which):
            m = pool[k]
            m.backtrack(pd)
            m.rhs_producers(pd)
            plain[k].assemble(pd)
            plain[k].reset_history()
            x, _, _ = plain[k].solve(rhs[j], rtol, 20000, check_every)
            x64.append(x)
        ctx.synchronize()
        t_plain = time.perf_counter() - t
    plain_rate = n_req / t_plain
    for s in plain:
        s.close()
    # the same requests with plain FP64 multicolour SSOR solves (the 1M headline's method)
    plain = [core.Solver(m, L.PREC_FP64, precond=L.PRECOND_SSOR) for m in pool]
    for s in plain:
        s.set_tuning(2)
    for rep_ in range(2):  # the first pass warms the solvers
        ctx.synchronize()
        t = time.perf_counter()
        for j, k in enumerate(which):
            m = pool[k]
            m.backtrack(pd)
            m.rhs_producers(pd)
            plain[k].assemble(pd)
            plain[k].reset_history()
            plain[k].solve(rhs[j], rtol, 20000, check_every, want_x=False)
        ctx.synchronize()
        t_plain_ssor = time.perf_counter() - t
    plain_ssor_rate = n_req / t_plain_ssor
    for s in plain:
        s.close()
    res = {"workload": f"{n_req} requests over 6 meshes, n = {sizes}", "plain_fp64_solves_per_s": round(plain_rate, 3),
           "plain_fp64_ssor_solves_per_s": round(plain_ssor_rate, 3),
           "mixed_rtol": mixed_rtol, "rates": []}
    def serve(precond):
        out = []
        srv, ids, warm = None, None, True
        for mult in (rates[-1],) + tuple(rates):  # the first pass warms the server's solver pools
            lam = mult * plain_rate
            arrive = np.cumsum(expo) / lam
            if srv is None:
                srv = core.Server(ctx, rtol=mixed_rtol, maxiter=20000, check_every=check_every,
                                  precision=L.PREC_MIXED, precond=precond)
                ids = [srv.add_mesh(m) for m in pool]
            tickets, by_ticket = {}, {}
            done = {}
            t0 = time.perf_counter()
            nxt = 0
            while len(done) < n_req:
                now = time.perf_counter() - t0
                while nxt < n_req and arrive[nxt] <= now:
                    tickets[nxt] = srv.submit(ids[which[nxt]], rhs[nxt], with_backtrack=True)
                    by_ticket[tickets[nxt]] = nxt
                    nxt += 1
                a, q = srv.step()
                for tk in srv.finished():
                    j = by_ticket[tk]
                    st, x, it, lat = srv.poll(tk, pool[which[j]].n)
                    if st != 0:
                        raise RuntimeError(f"request {j} failed with status {st}")
                    srv.release(tk)
                    done[j] = (x, it, lat, time.perf_counter() - t0)
                if a == 0 and q == 0 and nxt < n_req:
                    time.sleep(max(0.0, min(0.01, arrive[nxt] - (time.perf_counter() - t0))))
            t_end = max(v[3] for v in done.values())
            lats = np.array([done[j][2] for j in range(n_req)])
            errs = [float(np.linalg.norm(done[j][0] - x64[j]) / np.linalg.norm(x64[j])) for j in range(n_req)]
            if warm:
                warm = False
                continue
            out.append({"arrival_rate_per_s": round(lam, 3),
                        "solves_per_s": round(n_req / (t_end - arrive[0]), 3),
                        "p50_ms": round(float(np.percentile(lats, 50)), 3),
                        "p99_ms": round(float(np.percentile(lats, 99)), 3),
                        "max_rel_l2_vs_fp64": max(errs),
                        "mean_inner_iters": float(np.mean([done[j][1] for j in range(n_req)]))})
        srv.close()
        return out

    res["rates"] = serve(L.PRECOND_JACOBI)
    res["ssor_rates"] = serve(L.PRECOND_SSOR)

    # Solve only: the same requests without backtracking (assembly + solve of the given RHS).
    # Plain = FP64 SSOR one request after another; served = mixed SSOR with every request queued
    # at once (saturation). With backtracking on both sides the served gain is bounded by the
    # backtracking share of a plain request, which runs at the same speed in both forms.
    def plain_solve_only():
        sol = [core.Solver(m, L.PREC_FP64, precond=L.PRECOND_SSOR) for m in pool]
        for s in sol:
            s.set_tuning(2)
        for rep_ in range(2):  # the first pass warms the solvers
            ctx.synchronize()
            t = time.perf_counter()
            for j, k in enumerate(which):
                sol[k].assemble(pd)
                sol[k].reset_history()
                sol[k].solve(rhs[j], rtol, 20000, check_every, want_x=False)
            ctx.synchronize()
            dt = time.perf_counter() - t
        for s in sol:
            s.close()
        return n_req / dt

    def served_saturated():
        srv = core.Server(ctx, rtol=mixed_rtol, maxiter=20000, check_every=check_every,
                          precision=L.PREC_MIXED, precond=L.PRECOND_SSOR)
        ids = [srv.add_mesh(m) for m in pool]
        for rep_ in range(2):  # the first pass warms the server's solver pools
            ctx.synchronize()
            t = time.perf_counter()
            left = {srv.submit(ids[k], rhs[j], with_backtrack=False) for j, k in enumerate(which)}
            while left:
                srv.step()
                for tk in srv.finished():
                    st_, _, _, _ = srv.poll(tk)
                    if st_ != 0:
                        raise RuntimeError(f"solve-only request failed with status {st_}")
                    srv.release(tk)
                    left.discard(tk)
            dt = time.perf_counter() - t
        srv.close()
        return n_req / dt

    p_so = plain_solve_only()
    s_so = served_saturated()
    res["solve_only"] = {"plain_fp64_ssor_solves_per_s": round(p_so, 3),
                         "served_mixed_ssor_saturated_solves_per_s": round(s_so, 3),
                         "served_vs_plain": round(s_so / p_so, 3),
                         "backtrack_share_of_plain_ssor_request": round(1.0 - plain_ssor_rate / p_so, 3)}
    for m in pool:
        m.close()
    return res


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--ssor-rows-per-thread", type=int, default=1)
    ap.add_argument("--jacobi-rows-per-thread", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-jacobi", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0 = all visible cores, max 16)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--meshes", type=int, default=256)
    ap.add_argument("--ensemble-nx", type=int, default=708)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--requests", type=int, default=120)
    ap.add_argument("--mixed-rtol", type=float, default=1e-9)
    ap.add_argument("--no-mixed-4m", action="store_true")
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
    # the LDS-staged element walk (same results) beside the default cache-only walk
    mesh.set_backtrack_staging(True)
    ms_bt_staged = min(timed(lambda: mesh.backtrack(p)) for _ in range(3))
    mesh.set_backtrack_staging(False)
    walks = mesh.ns * mesh.nvrt
    # the state and results of these backtracks, for the CPU baseline of backtracking below
    bt_host = None
    if not args.no_cpu_baseline:
        bt_host = {f: mesh.get(f) for f in (L.F_ETA2, L.F_NODE_UVW, L.F_SIDE_U2, L.F_FOOT, L.F_BTU, L.F_BSTATUS)}
    ms_rhs = min(timed(lambda: mesh.rhs_producers(p)) for _ in range(3))
    solver = core.Solver(mesh, L.PREC_FP64, precond=L.PRECOND_SSOR)
    solver.set_tuning(args.ssor_rows_per_thread)
    ms_asm = min(timed(lambda: solver.assemble(p)) for _ in range(3))
    # the step before backtracking: nodal velocities from the side velocities and element w
    # (zero here); the system above is already assembled, so replacing the nodal field is harmless
    ms_nv = min(timed(lambda: mesh.nodal_velocity(p)) for _ in range(3))
    nv_bytes = (16 * mesh.ns + 8 * mesh.ne + 32 * n) * mesh.nvrt
    # the step after the solve: side velocities from the momentum solve (zc, btu in, u2 out per
    # side-level; about 300 B of element and node gathers per side for the elevation gradients)
    # and the Shapiro filter (u2 read, the filtered u2 written per side-level)
    ms_mom = min(timed(lambda: mesh.momentum(p)) for _ in range(3))
    mom_bytes = 40 * mesh.ns * mesh.nvrt + 300 * mesh.ns + (32 * mesh.ns * mesh.nvrt if p.ishapiro == 1 else 0)
    # element w from the new side velocities: u2 and zc rows of every side read, w written
    ms_vv = min(timed(lambda: mesh.vertical_velocity(p)) for _ in range(3))
    vv_bytes = (24 * mesh.ns + 8 * mesh.ne) * mesh.nvrt + 64 * n
    ndelt = int(round(p.dt / p.dtb_max))
    # backtracking is VALU-issue bound: its FP64 rate and HBM rate against the peaks, from the
    # counter summary's per-launch figures over this run's time
    bt_rec = {"ms": round(ms_bt, 3), "side_levels": walks, "per_s": round(walks / (ms_bt * 1e-3), 1),
              "substeps": ndelt, "ms_lds_staged_walk": round(ms_bt_staged, 3)}
    bb = backtrack_bound()
    if bb is not None:
        bt_rec["roofline"] = {"bound": "valu", "valu_busy": bb["valu_busy"], "waves_per_simd": bb["waves_per_simd"],
                              "fp64_tflops": round(bb["fp64_ops_per_launch"] / (ms_bt * 1e-3) / 1e12, 2),
                              "fp64_peak_tflops": PEAK_FP64_TFLOPS,
                              "fp64_frac": round(bb["fp64_ops_per_launch"] / (ms_bt * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 3),
                              "hbm_gbs": round(bb["hbm_bytes_per_launch"] / (ms_bt * 1e-3) / 1e9, 1),
                              "hbm_frac": round(bb["hbm_bytes_per_launch"] / (ms_bt * 1e-3) / 1e9 / PEAK_HBM_GBS, 3)}
    log(f"[bench] backtrack {ms_bt:.2f} ms ({walks} side-levels; LDS-staged walk {ms_bt_staged:.2f} ms), rhs {ms_rhs:.2f} ms, assembly {ms_asm:.2f} ms, "
        f"nodal velocity {ms_nv:.2f} ms ({nv_bytes / (ms_nv * 1e-3) / 1e9:.0f} GB/s), "
        f"momentum {ms_mom:.2f} ms ({mom_bytes / (ms_mom * 1e-3) / 1e9:.0f} GB/s), "
        f"vertical velocity {ms_vv:.2f} ms ({vv_bytes / (ms_vv * 1e-3) / 1e9:.0f} GB/s)")

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
    # the same solves with b handed over from host memory and x returned to it (the boundary's
    # host-buffer form): the rate with the PCIe copies, reported beside the value, never as it
    b_host = solver.get_rhs()
    ctx.synchronize()
    a = time.perf_counter()
    for _ in range(args.steps):
        solver.reset_history()
        solver.solve(b_host, rtol, maxiter, check_every)
    ctx.synchronize()
    host_buf_rate = args.steps / (time.perf_counter() - a)
    log(f"[bench] with b from and x to host memory: {host_buf_rate:.2f} solves/s")

    # ---- kernel durations per launch kind (kernel-stamped HIP events on the solver stream)
    solver.reset_history()
    ms_kind, _ = solver.profile(rtol, maxiter, check_every)
    kinds = ssor_bytes(mesh, solver)
    nc = kinds["ncolours"]
    per = {}
    for i, (k, kern, launches) in enumerate((("bwd", "k_ssor_bwd", nc), ("fwd", "k_ssor_fwd", nc - 1))):
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

    # ---- mixed precision (FP32 inner Jacobi / SSOR PCG, FP64 residual and refinement) on the same system
    mixed = None
    if not args.no_mixed and x_jac is not None:
        ms_solver = core.Solver(mesh, L.PREC_MIXED)
        ms_solver.set_tuning(args.jacobi_rows_per_thread, 1e-5)
        ms_solver.assemble(p)
        ms_solver.reset_history()  # warm, and the solution for the error (untimed: the copy to the host)
        xmix, itm, stm = ms_solver.solve(None, rtol, maxiter, check_every)
        st_m = ms_solver.stats()
        k_m = max(2, args.steps // 2)  # timed as the FP64 solves (timed_solves: no copy of x)
        t_m, _ = timed_solves(ctx, ms_solver, k_m, rtol, maxiter, check_every)
        mixed = {"solves_per_s": round(k_m / t_m, 3), "inner_iters": int(st_m.iters),
                 "outer": int(st_m.outer),
                 "rel_l2_vs_fp64": float(np.linalg.norm(xmix - x_jac) / np.linalg.norm(x_jac))}
        ms_solver.close()
        # the same with FP32 multicolour SSOR inner solves
        ms_solver = core.Solver(mesh, L.PREC_MIXED, precond=L.PRECOND_SSOR)
        ms_solver.set_tuning(args.ssor_rows_per_thread, 1e-5)
        ms_solver.assemble(p)
        ms_solver.reset_history()
        xmix, itm, stm = ms_solver.solve(None, rtol, maxiter, check_every)
        st_m = ms_solver.stats()
        t_m, _ = timed_solves(ctx, ms_solver, k_m, rtol, maxiter, check_every)
        mixed["ssor"] = {"solves_per_s": round(k_m / t_m, 3), "inner_iters": int(st_m.iters),
                         "outer": int(st_m.outer),
                         "rel_l2_vs_fp64_ssor": float(np.linalg.norm(xmix - x_ssor) / np.linalg.norm(x_ssor)),
                         "vs_fp64_ssor": round(k_m / t_m / solves_per_s, 3)}
        ms_solver.close()
        log(f"[bench] mixed: {mixed}")

    # ---- CPU baseline: the oracle's FP64 SSOR PCG on the same system, all host cores
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as O

        ip_h, ix_h = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
        val = solver.get_values()
        bvec = solver.get_rhs()
        # the host cores this process may use: the CPU share the GPU box gives one GPU
        # (OMP_NUM_THREADS there), else every core of the affinity mask
        affinity = len(os.sched_getaffinity(0))
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = args.cpu_threads or (min(share, affinity) if share > 0 else affinity)
        osr = O.OracleSsor(ip_h, ix_h, args.ssor_rows_per_thread)
        reps, spent, xo, ito = 0, 0.0, None, 0
        while reps < 40 and (spent < 10.0 or reps < 2):
            a = time.perf_counter()
            xo, ito, _, _ = osr.solve(val, bvec, rtol=rtol, maxiter=maxiter, check_every=check_every,
                                      nthreads=threads)
            spent += time.perf_counter() - a
            reps += 1
        # the same solve on one core (one solve)
        a = time.perf_counter()
        x1, it1, _, _ = osr.solve(val, bvec, rtol=rtol, maxiter=maxiter, check_every=check_every, nthreads=1)
        one_core = time.perf_counter() - a
        osr.close()
        cpu = {"value": round(reps / spent, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"{reps} full FP64 SSOR PCG solves of the same {n}-node system ({ito} iterations each), "
                         f"oracle C restatement, OpenMP over rows, {cpu_model()}",
               "cpu_model": cpu_model(), "nproc_affinity": affinity, "os_cpu_count": os.cpu_count(),
               "omp_num_threads_env": share,
               "seconds": round(spent, 3),
               "gpu_bitwise_equal": bool(np.array_equal(xo, x_ssor)), "iters_equal": bool(ito == it_ssor),
               "single_core": {"value": round(1.0 / one_core, 4), "unit": "solves/s", "cores": 1,
                               "sample": "1 solve", "bitwise_equal": bool(np.array_equal(x1, x_ssor))}}
        log(f"[bench] cpu baseline: {cpu}")
        # backtracking on the host: the oracle's backtrackAllNodes restatement, all host cores, on
        # every 4th side (all levels) of the same mesh and state, compared with the GPU's feet,
        # velocities and status flags for those start points
        om = O.OracleMesh(mesh.topology(L.T_XY), mesh.topology(L.T_ELNODE), mesh.topology(L.T_DEPTH),
                          mesh.topology(L.T_OPEN_BND), data.sigma, kb=mesh.topology(L.T_KB))
        for f in (L.F_ETA2, L.F_NODE_UVW, L.F_SIDE_U2):
            om.set(f, bt_host[f])
        om.levels(p)
        nv_ = mesh.nvrt
        sides = np.repeat(np.arange(0, mesh.ns, 4, dtype=np.int32), nv_)
        levs = np.tile(np.arange(nv_, dtype=np.int32), (mesh.ns + 3) // 4)
        a = time.perf_counter()
        om.backtrack(p, sides, levs, nthreads=threads)
        bt_cpu_s = time.perf_counter() - a
        at = sides.astype(np.int64) * nv_ + levs
        same = all(np.array_equal(om.get(f).reshape(mesh.ns * nv_, -1)[at], bt_host[f].reshape(mesh.ns * nv_, -1)[at])
                   for f in (L.F_FOOT, L.F_BTU, L.F_BSTATUS))
        bt_rec["cpu_baseline"] = {"value": round(sides.size / bt_cpu_s, 1), "unit": "side-levels/s", "cores": threads,
                                  "kind": "port",
                                  "sample": f"{sides.size} (side, level) start points (every 4th side, all levels, "
                                            f"{ndelt} substeps), oracle C restatement, OpenMP over start points",
                                  "seconds": round(bt_cpu_s, 3), "gpu_bitwise_equal": bool(same)}
        log(f"[bench] backtrack cpu baseline: {bt_rec['cpu_baseline']}")
        del om, bt_host

    # ---- whole time steps of the rebuilt path on the same mesh (after every measurement above
    # that needs the original state): backtracking, RHS producers, assembly, FP64 SSOR solve,
    # elevation update, momentum, element w, levels, nodal velocities
    def one_step():
        mesh.backtrack(p)
        mesh.rhs_producers(p)
        solver.assemble(p)
        _, it_s, st_s = solver.solve(None, rtol, maxiter, check_every, want_x=False)
        solver.update_elevation()
        mesh.momentum(p)
        mesh.vertical_velocity(p)
        mesh.update_levels(p)
        mesh.nodal_velocity(p)
        return it_s, st_s

    one_step()  # warm
    ctx.synchronize()
    t_st = time.perf_counter()
    step_its = [one_step() for _ in range(3)]
    ctx.synchronize()
    ms_step = (time.perf_counter() - t_st) * 1e3 / 3
    full_step = {"ms": round(ms_step, 3), "steps_per_s": round(1e3 / ms_step, 3),
                 "solver_iters": [int(i) for i, _ in step_its], "status": [int(s) for _, s in step_its]}
    log(f"[bench] full step: {full_step}")
    if not args.no_cpu_baseline:
        # the same step on the host: the oracle's restatement of every piece (backtracking,
        # assembly, the SSOR solve and the nodal velocities on all host cores; RHS producers,
        # momentum, vertical velocity and levels on one), from the GPU's state before one more
        # step with a zero solver history, compared bitwise with the GPU's state after it
        from oracle import pyoracle as O

        solver.reset_history()
        start = {f: mesh.get(f) for f in (L.F_ETA1, L.F_ETA2, L.F_NODE_UVW, L.F_SIDE_U2)}
        it_g, _ = one_step()
        after = {f: mesh.get(f) for f in (L.F_ETA2, L.F_SIDE_U2, L.F_ELEM_W, L.F_NODE_UVW)}
        om = O.OracleMesh(mesh.topology(L.T_XY), mesh.topology(L.T_ELNODE), mesh.topology(L.T_DEPTH),
                          mesh.topology(L.T_OPEN_BND), data.sigma, kb=mesh.topology(L.T_KB))
        for f, v in start.items():
            om.set(f, v)
        om.levels(p)
        ip_o, ix_o = om.topology(L.T_CSR_INDPTR), om.topology(L.T_CSR_INDICES)
        a = time.perf_counter()
        om.backtrack(p, nthreads=threads)
        om.rhs(p)
        val_o, b_o = om.assemble(p, nthreads=threads)
        x_o, it_o, _, _, _ = O.pcg_ssor(ip_o, ix_o, val_o, b_o, hist=[np.zeros(n)] * 3, omega=1.0, rtol=rtol,
                                        maxiter=maxiter, check_every=check_every,
                                        rows_per_thread=args.ssor_rows_per_thread, nthreads=threads)
        om.set(L.F_ETA1, om.get(L.F_ETA2))
        om.set(L.F_ETA2, x_o)
        om.momentum(p)
        om.vertical_velocity()
        om.levels(p)
        om.nodal_velocity(p, nthreads=threads)
        cpu_step_s = time.perf_counter() - a
        same = all(np.array_equal(om.get(f), v) for f, v in after.items()) and int(it_o) == int(it_g)
        full_step["cpu_baseline"] = {"value": round(1.0 / cpu_step_s, 4), "unit": "steps/s", "cores": threads,
                                     "kind": "port",
                                     "sample": "one whole step of the same mesh and state (oracle C restatement; "
                                               "backtracking, assembly, SSOR solve and nodal velocities on all "
                                               "cores, the other pieces on one)",
                                     "seconds": round(cpu_step_s, 3), "gpu_bitwise_equal": bool(same)}
        log(f"[bench] full step cpu baseline: {full_step['cpu_baseline']}")
        del om

    mixed4m = None
    if not args.no_mixed_4m:
        mixed4m = mixed_4m_leg(ctx, L, cases, core, 3, rtol, 1e-10, check_every)
        log(f"[bench] mixed 4M: {mixed4m}")
    batched = None
    if not args.no_batched:
        batched = batched_leg(ctx, L, cases, core, args.members, args.ensemble_nx, args.ensemble_nx, rtol,
                              args.mixed_rtol, check_every)
        log(f"[bench] batched: {batched}")
        batched["many_meshes"] = many_meshes_leg(ctx, L, cases, core, args.meshes, rtol, args.mixed_rtol,
                                                 check_every)
        log(f"[bench] batched many meshes: {batched['many_meshes']}")
    served = None
    if not args.no_served:
        served = served_leg(ctx, L, cases, core, args.requests, (1.0, 2.0, 4.0), rtol, args.mixed_rtol,
                            check_every)
        log(f"[bench] served: {served}")

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
                     "unit": "GB/s", "frac": round(d["gbs"] / PEAK_HBM_GBS, 4),
                     "traffic": counter_traffic(d["kernel"], args.ssor_rows_per_thread),
                     "bytes_per_launch": round(d["bytes_per_iter"] / d["launches_per_iter"]),
                     "us_per_launch": round(us_launch, 2)},
        "ssor_kernels": per,
        "copy_probe_gbs": round(copy_gbs, 1),
        "jacobi": jac,
        "backtrack": bt_rec,
        "nodal_velocity": {"ms": round(ms_nv, 3), "bytes": nv_bytes, "gbs": round(nv_bytes / (ms_nv * 1e-3) / 1e9, 1)},
        "momentum": {"ms": round(ms_mom, 3), "bytes": mom_bytes, "gbs": round(mom_bytes / (ms_mom * 1e-3) / 1e9, 1)},
        "full_step": full_step,
        "vertical_velocity": {"ms": round(ms_vv, 3), "bytes": vv_bytes,
                              "gbs": round(vv_bytes / (ms_vv * 1e-3) / 1e9, 1)},
        "assembly_ms": round(ms_asm, 3),
        "rhs_ms": round(ms_rhs, 3),
        "mixed": mixed,
        "mixed_4m": mixed4m,
        "batched": batched,
        "served": served,
        "host_buffers": {"solves_per_s": round(host_buf_rate, 3),
                         "note": "the same solves with b copied from and x copied to host memory per solve"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


from ctypes import c_double as C_double  # noqa: E402


if __name__ == "__main__":
    sys.exit(main())
