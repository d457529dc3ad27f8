"""GPU parity of the multicolour SSOR solver against the oracle, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import assert_bitwise
from gschism_amd import _lib as L
from gschism_amd import cases, core
from oracle import pyoracle as O
from test_gpu_parity import _setup

pytestmark = pytest.mark.gpu


def _assembled(ctx, spec):
    mesh, om, p = _setup(ctx, spec)
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    om.set(L.F_BTU, mesh.get(L.F_BTU))
    om.rhs(p)
    val, b = om.assemble(p, nthreads=8)
    return mesh, om, p, val, b


def _compare(solver, ip, ix, val, b, hist=None, omega=1.0, rpt=1, rtol=1e-12, maxiter=3000, check_every=10,
             what=""):
    x, it, st = solver.solve(None, rtol=rtol, maxiter=maxiter, check_every=check_every)
    xo, ito, rr, at, nc = O.pcg_ssor(ip, ix, val, b, hist=hist, omega=omega, rtol=rtol, maxiter=maxiter,
                                     check_every=check_every, rows_per_thread=rpt, nthreads=8)
    assert it == ito, (what, it, ito)
    assert_bitwise(x, xo, f"{what} SSOR solution")
    s = solver.stats()
    assert s.atol2 == at and s.rr == rr, what
    return x, it, st


@pytest.mark.parametrize("rpt", [1, 4])
def test_ssor_channel_200x80(ctx, rpt):
    spec = cases.spec_cpu_reference()
    print(f"[gpu] SSOR {spec.name} rpt={rpt}", flush=True)
    mesh, om, p, val, b = _assembled(ctx, spec)
    solver = core.Solver(mesh, precond=L.PRECOND_SSOR)
    solver.set_tuning(rpt)
    solver.assemble(p)
    assert_bitwise(solver.get_values(), val, "CSR values")
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    nc, npad, col = solver.ssor_info()
    ocol, onc = O.colour(ip, ix)
    assert nc == onc and np.array_equal(col, ocol)
    assert npad % (256 * rpt) == 0
    x, it, st = _compare(solver, ip, ix, val, b, rpt=rpt, what=f"rpt={rpt}")
    assert st == 0
    assert solver.check() < 1e-10
    # fewer iterations than the Jacobi solver on the same system
    jac = core.Solver(mesh)
    jac.assemble(p)
    _, itj, _ = jac.solve(None, rtol=1e-12, maxiter=3000)
    assert it < itj
    print(f"[gpu] SSOR {it} iterations, Jacobi {itj}, {nc} colours", flush=True)


def test_ssor_history_omega_and_steps(ctx):
    spec = cases.CaseSpec("channel-60x30", 60, 30, 500.0, nvrt=8, depth=None, substeps=5)
    mesh, om, p, val, b = _assembled(ctx, spec)
    solver = core.Solver(mesh, precond=L.PRECOND_SSOR)
    solver.set_omega(1.3)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    hist = [np.zeros(mesh.n)] * 3
    for step in range(3):
        print(f"[gpu] SSOR omega 1.3 step {step}", flush=True)
        if step > 0:
            mesh.backtrack(p)
            mesh.rhs_producers(p)
            om.set(L.F_BTU, mesh.get(L.F_BTU))
            om.rhs(p)
            val, b = om.assemble(p)
        solver.assemble(p)
        # windows of 10, 7 and 5 iterations: the solve ends on odd and even iterations, so the
        # y update left pending after an even one is exercised
        ce = (10, 7, 5)[step]
        x, it, st = _compare(solver, ip, ix, val, b, hist=hist, omega=1.3, check_every=ce,
                             what=f"step {step} check_every {ce}")
        hist = [x, hist[0], hist[1]]
        solver.update_elevation()
        om.set(L.F_ETA2, x)
        mesh.update_levels(p)
        om.levels(p)


def test_ssor_standalone_csr_and_special_cases(ctx):
    rng = np.random.default_rng(5)
    n = 5000
    A = sp.random(n, n, density=5.0 / n, random_state=6, format="csr")
    A = A + A.T
    A = (A + sp.diags(np.abs(A).sum(1).A1 + 0.5)).tocsr()
    A.sort_indices()
    b = rng.normal(size=n)
    print("[gpu] SSOR standalone CSR", flush=True)
    s = core.Solver(None, ctx=ctx, indptr=A.indptr, indices=A.indices, precond=L.PRECOND_SSOR)
    s.set_values(A.data)
    x, it, st = s.solve(b, rtol=1e-12, maxiter=2000)
    xo, ito, _, _, _ = O.pcg_ssor(A.indptr, A.indices, A.data, b, rtol=1e-12, maxiter=2000, rows_per_thread=1)
    assert it == ito
    assert_bitwise(x, xo, "standalone SSOR")
    # zero right-hand side: atol = 0, x = 0, no iterations
    x, it, st = s.solve(np.zeros(n), rtol=1e-12, maxiter=100)
    assert it == 0 and st == 0 and not x.any()
    # maxiter reached without convergence: status flag, iterations = maxiter
    s.reset_history()
    x, it, st = s.solve(b, rtol=1e-30, maxiter=25, check_every=10)
    assert it == 25 and st & L.GSH_W_MAXITER
    xo, ito, _, _, _ = O.pcg_ssor(A.indptr, A.indices, A.data, b, rtol=1e-30, maxiter=25, check_every=10,
                                  rows_per_thread=1)
    assert ito == 25
    assert_bitwise(x, xo, "SSOR at maxiter")
    with pytest.raises(L.GshError):
        s.set_omega(2.5)
    sm = core.Solver(None, L.PREC_MIXED, ctx=ctx, indptr=A.indptr, indices=A.indices, precond=L.PRECOND_SSOR)
    sm.set_values(A.data)
    xm, itm, stm = sm.solve(b, rtol=1e-10, maxiter=2000)
    assert stm == 0 and np.linalg.norm(b - A @ xm) / np.linalg.norm(b) <= 1e-10 * (1 + 1e-6)
    sm.close()


def _tiny_systems():
    """(name, CSR matrix) of small and degenerate SPD systems: one row, a diagonal matrix (one
    colour), a 2 x 2 system, a tridiagonal chain (two colours) and a 3 x 3 dense block (three)."""
    rng = np.random.default_rng(11)
    out = [("n1", sp.csr_matrix(np.array([[4.0]])))]
    out.append(("diagonal-300", sp.diags(rng.uniform(1.0, 5.0, 300)).tocsr()))
    out.append(("2x2", sp.csr_matrix(np.array([[4.0, -1.0], [-1.0, 3.0]]))))
    n = 700
    out.append(("chain-700", sp.diags([-np.ones(n - 1), 2.5 + rng.uniform(0, 1, n), -np.ones(n - 1)],
                                     [-1, 0, 1]).tocsr()))
    out.append(("dense-3", sp.csr_matrix(np.array([[5.0, 1.0, 1.0], [1.0, 4.0, 1.0], [1.0, 1.0, 6.0]]))))
    return out


@pytest.mark.parametrize("name", [t[0] for t in _tiny_systems()])
def test_tiny_and_degenerate_systems(ctx, name):
    """Tiny systems through both solvers against the oracle, bit for bit, with a cold start, a
    repeat and a zero right-hand side."""
    A = dict(_tiny_systems())[name]
    A.sort_indices()
    n = A.shape[0]
    b = np.random.default_rng(3).normal(size=n)
    for precond in (L.PRECOND_SSOR, L.PRECOND_JACOBI):
        print(f"[gpu] tiny system {name} n={n} precond={precond}", flush=True)
        s = core.Solver(None, ctx=ctx, indptr=A.indptr, indices=A.indices, precond=precond)
        s.set_values(A.data)
        for rep in range(2):
            s.reset_history()
            x, it, st = s.solve(b, rtol=1e-12, maxiter=500)
            if precond == L.PRECOND_SSOR:
                xo, ito, _, _, nc = O.pcg_ssor(A.indptr, A.indices, A.data, b, rtol=1e-12, maxiter=500,
                                               rows_per_thread=1)
            else:
                xo, ito, _, _ = O.pcg(A.indptr, A.indices, A.data, b, rtol=1e-12, maxiter=500, rows_per_thread=4)
            assert st == 0 and it == ito, (name, precond, it, ito, st)
            assert_bitwise(x, xo, f"{name} precond={precond} solve {rep}")
        assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
        x, it, st = s.solve(np.zeros(n), rtol=1e-12, maxiter=100)
        assert it == 0 and st == 0 and not x.any()
        s.close()


def test_autotune_keeps_results_exact(ctx):
    """The autotuner only picks the tile size; the solve afterwards still matches the oracle run
    with that tile size (Jacobi and SSOR)."""
    spec = cases.CaseSpec("tune-90x60", 90, 60, 600.0, nvrt=4, depth=None, substeps=4)
    mesh, om, p, val, b = _assembled(ctx, spec)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    for precond in (L.PRECOND_JACOBI, L.PRECOND_SSOR):
        s = core.Solver(mesh, precond=precond)
        s.assemble(p)
        best, us = s.autotune(12)
        print(f"[gpu] autotune precond={precond}: best rows_per_thread {best}, us/iter {us}", flush=True)
        assert best in (1, 2, 4, 8) and (us > 0).all()
        x, it, st = s.solve(None, rtol=1e-12, maxiter=3000)
        if precond == L.PRECOND_SSOR:
            xo, ito, _, _, _ = O.pcg_ssor(ip, ix, val, b, rtol=1e-12, maxiter=3000, rows_per_thread=best)
        else:
            xo, ito, _, _ = O.pcg(ip, ix, val, b, rtol=1e-12, maxiter=3000, rows_per_thread=best)
        assert it == ito
        assert_bitwise(x, xo, f"solution after autotune (precond {precond})")


def test_ssor_2m_matches_oracle(ctx):
    """Two million rows: thousands of workgroups per sweep, so every partial a sweep reads must be
    the previous iteration's (the sweeps write the next iteration's into the other slot)."""
    spec = cases.CaseSpec("channel-1448x1448-lognormal", 1448, 1448, 1000.0, nvrt=4, depth=None, substeps=10)
    data = cases.make_case(spec)
    print(f"[gpu] SSOR 2M: building {spec.name}", flush=True)
    mesh, p, _ = cases.load_mesh(ctx, data)
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    for rpt in (1, 2):
        print(f"[gpu] SSOR 2M rows_per_thread={rpt}", flush=True)
        s = core.Solver(mesh, precond=L.PRECOND_SSOR)
        s.set_tuning(rpt)
        s.assemble(p)
        val, b = s.get_values(), s.get_rhs()
        x, it, st = s.solve(None, rtol=1e-12, maxiter=5000)
        assert st == 0
        A = sp.csr_matrix((val, ix, ip), shape=(mesh.n, mesh.n))
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10
        xo, ito, _, _, _ = O.pcg_ssor(ip, ix, val, b, rtol=1e-12, maxiter=5000, rows_per_thread=rpt, nthreads=16)
        assert it == ito
        assert_bitwise(x, xo, f"2M SSOR solution (rows_per_thread {rpt})")
        s.close()


def test_mixed_4m_against_oracle(ctx):
    """Configuration 3: the 2048 x 2048 log-normal channel (4.2M nodes), mixed precision (FP32 SSOR
    inner solves, FP64 residual and refinement) to the outer tolerance 1e-10, against the oracle's
    FP64 SSOR PCG of the same system (rtol 1e-12): relative L2 error <= 1e-8, the FP64 true
    residual within the outer tolerance; iteration counts printed beside the FP64 solver's."""
    spec = cases.spec_mixed_4m()
    data = cases.make_case(spec)
    print(f"[gpu] mixed 4M: building {spec.name}", flush=True)
    mesh, p, _ = cases.load_mesh(ctx, data)
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    sm = core.Solver(mesh, L.PREC_MIXED, precond=L.PRECOND_SSOR)
    sm.set_tuning(1)
    sm.assemble(p)
    val, b = sm.get_values(), sm.get_rhs()
    xm, itm, stm = sm.solve(None, rtol=1e-10, maxiter=5000, check_every=10)
    stt = sm.stats()
    assert stm == 0 and stt.rr <= stt.atol2
    print(f"[gpu] mixed 4M: oracle FP64 SSOR solve (n={mesh.n})", flush=True)
    xo, ito, _, _, _ = O.pcg_ssor(ip, ix, val, b, rtol=1e-12, maxiter=5000, rows_per_thread=1, nthreads=16)
    err = np.linalg.norm(xm - xo) / np.linalg.norm(xo)
    A = sp.csr_matrix((val, ix, ip), shape=(mesh.n, mesh.n))
    rel = np.linalg.norm(b - A @ xm) / np.linalg.norm(b)
    print(f"[gpu] mixed 4M: {stt.outer} outer, {itm} inner FP32 iterations; oracle FP64 SSOR {ito} iterations; "
          f"rel L2 vs oracle {err:.3e}, true rel residual {rel:.3e}", flush=True)
    assert err <= 1e-8, err
    assert rel <= 1.01e-10, rel
    sm.close()


def test_window_graphs_replay_exactly(ctx):
    """With window graphs on, repeated solves (graph replays with eager work in between) give the
    oracle's solutions, for Jacobi and SSOR."""
    spec = cases.CaseSpec("graph-70x40", 70, 40, 600.0, nvrt=4, depth=None, substeps=4)
    mesh, om, p, val, b = _assembled(ctx, spec)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    for precond in (L.PRECOND_JACOBI, L.PRECOND_SSOR):
        s = core.Solver(mesh, precond=precond)
        s.set_graphs(True)
        s.assemble(p)
        for rep in range(3):
            rhs = cases.normal(50 + rep, mesh.n)
            s.reset_history()
            x, it, st = s.solve(rhs, rtol=1e-12, maxiter=3000)
            if precond == L.PRECOND_SSOR:
                xo, ito, _, _, _ = O.pcg_ssor(ip, ix, val, rhs, rtol=1e-12, maxiter=3000, rows_per_thread=1)
            else:
                xo, ito, _, _ = O.pcg(ip, ix, val, rhs, rtol=1e-12, maxiter=3000)
            assert it == ito
            assert_bitwise(x, xo, f"graph replay {rep} (precond {precond})")
            assert s.check() < 1e-10
        s.close()


def test_mixed_precision_ssor(ctx):
    """FP32 SSOR inner solves with FP64 residual and refinement: the outer test is met on the
    FP64 residual, the solution is within the tolerance of the FP64 SSOR solution, repeated
    solves are bitwise equal, and the FP64 solver of the same mesh is unaffected."""
    spec = cases.CaseSpec("channel-64x40", 64, 40, 500.0, nvrt=6, depth=None, substeps=4)
    mesh, om, p, val, b = _assembled(ctx, spec)
    print(f"[gpu] mixed SSOR n={mesh.n}", flush=True)
    s64 = core.Solver(mesh, L.PREC_FP64, precond=L.PRECOND_SSOR)
    s64.assemble(p)
    x64, it64, st64 = s64.solve(None, rtol=1e-12, maxiter=5000, check_every=10)
    assert st64 == 0
    sm = core.Solver(mesh, L.PREC_MIXED, precond=L.PRECOND_SSOR)
    sm.assemble(p)
    x1, it1, st1 = sm.solve(None, rtol=1e-10, maxiter=5000, check_every=10)
    stt = sm.stats()
    assert st1 == 0 and stt.outer >= 2 and it1 > 0, (st1, stt.outer, it1)
    assert stt.rr <= stt.atol2
    assert np.linalg.norm(x1 - x64) / np.linalg.norm(x64) < 1e-8
    assert sm.check() < 1e-9
    sm.reset_history()
    x2, it2, _ = sm.solve(None, rtol=1e-10, maxiter=5000, check_every=10)
    assert it2 == it1
    assert_bitwise(x2, x1, "repeated mixed SSOR solve")
    # the mixed solver needs no more inner iterations than twice the FP64 SSOR count
    assert it1 <= 2 * it64 + 50, (it1, it64)
    sm.close()
    s64.close()


def _csr_system(ctx, kind, n, seed, precision=L.PREC_FP64):
    """A standalone SSOR solver on an SPD matrix: 'random' (several colours), 'tridiag' (two),
    'diag' (one colour)."""
    if kind == "random":
        A = sp.random(n, n, density=5.0 / n, random_state=seed, format="csr")
        A = A + A.T
        A = A + sp.diags(np.abs(A).sum(1).A1 + 0.5)
    elif kind == "tridiag":
        A = sp.diags([-np.ones(n - 1), 2.5 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    else:
        A = sp.diags(1.0 + np.arange(n) % 7)
    A = A.tocsr()
    A.sort_indices()
    s = core.Solver(None, precision, ctx=ctx, indptr=A.indptr, indices=A.indices, precond=L.PRECOND_SSOR)
    s.set_values(A.data)
    return s, A


def test_batched_ssor_equals_solo(ctx):
    """Batched FP64 SSOR: systems with different sizes and colour counts (one, two, several) and a
    zero right-hand side advance in lock-step; each gets its solo iteration count and solution
    bit for bit, and the solo results match the oracle."""
    rng = np.random.default_rng(11)
    systems = []
    for k, (nx, ny) in enumerate(((40, 30), (64, 20))):
        spec = cases.CaseSpec(f"sb{k}", nx, ny, 500.0, nvrt=5, depth=None, substeps=4)
        mesh, om, p, val, b = _assembled(ctx, spec)
        s = core.Solver(mesh, precond=L.PRECOND_SSOR)
        s.assemble(p)
        ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
        systems.append((s, ip, ix, val, None))
    for kind, n, seed in (("random", 3000, 21), ("tridiag", 2000, 0), ("diag", 700, 0), ("random", 900, 22)):
        s, A = _csr_system(ctx, kind, n, seed)
        b = np.zeros(n) if seed == 22 else rng.normal(size=n)
        systems.append((s, A.indptr, A.indices, A.data, b))
    solo = []
    for s, ip, ix, val, b in systems:
        if b is not None:
            s.set_rhs(b)
        x, it, st = s.solve(None, rtol=1e-12, maxiter=3000)
        solo.append((x, it))
        s.reset_history()
    ncols = [s.ssor_info()[0] for s, *_ in systems]
    print(f"[gpu] batched SSOR of {len(systems)}, colours {ncols}", flush=True)
    assert 1 in ncols and 2 in ncols and max(ncols) > 2
    its = core.batch_solve([s for s, *_ in systems], rtol=1e-12, maxiter=3000)
    for (s, ip, ix, val, b), (x, it), itb in zip(systems, solo, its):
        assert itb == it
        assert_bitwise(s.get_solution(), x, "batched SSOR solution")
    assert its[-1] == 0 and not solo[-1][0].any()
    s, ip, ix, val, b = systems[2]
    xo, ito, _, _, _ = O.pcg_ssor(ip, ix, val, b, rtol=1e-12, maxiter=3000, rows_per_thread=1)
    assert ito == its[2]
    assert_bitwise(solo[2][0], xo, "solo SSOR vs oracle")
    # every system at maxiter: the same count and the oracle's iterate
    for s, *_ in systems:
        s.reset_history()
    its = core.batch_solve([s for s, *_ in systems[:3]], rtol=1e-30, maxiter=25)
    assert its == [25, 25, 25]
    xo, _, _, _, _ = O.pcg_ssor(ip, ix, val, b, rtol=1e-30, maxiter=25, rows_per_thread=1)
    assert_bitwise(systems[2][0].get_solution(), xo, "batched SSOR at maxiter")
    # tile size 4 for all: still the solo results; mixed tile sizes or preconditioners are errors
    for s, *_ in systems:
        s.set_tuning(4)
        s.reset_history()
    its4 = core.batch_solve([s for s, *_ in systems[:3]], rtol=1e-12, maxiter=3000)
    xo4, ito4, _, _, _ = O.pcg_ssor(ip, ix, val, b, rtol=1e-12, maxiter=3000, rows_per_thread=4)
    assert its4[2] == ito4
    assert_bitwise(systems[2][0].get_solution(), xo4, "batched SSOR, 4 rows per thread")
    systems[1][0].set_tuning(1)
    with pytest.raises(L.GshError):
        core.batch_solve([s for s, *_ in systems[:2]], rtol=1e-12, maxiter=100)
    jac = core.Solver(None, ctx=ctx, indptr=systems[2][1], indices=systems[2][2])
    jac.set_values(systems[2][3])
    with pytest.raises(L.GshError):
        core.batch_solve([systems[2][0], jac], rtol=1e-12, maxiter=100)
    for s, *_ in systems:
        s.close()
    jac.close()


def test_mixed_batched_ssor_equals_solo(ctx):
    """Mixed precision with FP32 SSOR inner solves in lock-step: each system's result is its solo
    result bit for bit and within 1e-8 of its FP64 SSOR solution."""
    rng = np.random.default_rng(12)
    systems = []
    for k, (nx, ny) in enumerate(((50, 30), (28, 61))):
        spec = cases.CaseSpec(f"ms{k}", nx, ny, 800.0, nvrt=4, depth=None, substeps=4)
        mesh, om, p, val, b = _assembled(ctx, spec)
        f64 = core.Solver(mesh, precond=L.PRECOND_SSOR)
        f64.assemble(p)
        x64, _, _ = f64.solve(None, rtol=1e-12, maxiter=3000)
        f64.close()
        mx = core.Solver(mesh, L.PREC_MIXED, precond=L.PRECOND_SSOR)
        mx.assemble(p)
        systems.append((mx, x64))
    for kind, n, seed in (("random", 4000, 31), ("tridiag", 1500, 0)):
        s64, A = _csr_system(ctx, kind, n, seed)
        b = rng.normal(size=n)
        x64, _, _ = s64.solve(b, rtol=1e-12, maxiter=3000)
        s64.close()
        mx, _ = _csr_system(ctx, kind, n, seed, L.PREC_MIXED)
        mx.set_rhs(b)
        systems.append((mx, x64))
    solo = []
    for mx, x64 in systems:
        xm, itm, stm = mx.solve(None, rtol=1e-12, maxiter=3000)
        assert stm == 0
        assert np.linalg.norm(xm - x64) / np.linalg.norm(x64) < 1e-8
        solo.append((xm, itm, mx.stats().outer))
        mx.reset_history()
    print(f"[gpu] mixed batched SSOR of {len(systems)}", flush=True)
    its = core.batch_solve([m for m, _ in systems], rtol=1e-12, maxiter=3000)
    for (mx, _), (xm, itm, outer), itb in zip(systems, solo, its):
        assert itb == itm and mx.stats().outer == outer
        assert_bitwise(mx.get_solution(), xm, "mixed batched SSOR solution")
    for mx, _ in systems:
        mx.close()


def test_ssor_natural_and_shuffled_numbering(ctx):
    """The same 5-point operator on a 400 x 300 grid in natural numbering (neighbours within a few
    hundred rows) and in a random numbering (neighbours anywhere in colour blocks of 60k rows):
    both give the oracle's solution bit for bit, alone and batched together."""
    nx, ny = 400, 300
    n = nx * ny
    T = sp.diags([-np.ones(nx - 1), 2.0 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    A0 = (sp.kron(sp.eye(ny), T) + sp.kron(sp.diags([-np.ones(ny - 1), 2.0 * np.ones(ny), -np.ones(ny - 1)],
                                                     [-1, 0, 1]), sp.eye(nx)) + 0.01 * sp.eye(n)).tocsr()
    perm = np.random.default_rng(4).permutation(n)
    systems = []
    for name, A in (("natural", A0), ("shuffled", A0[perm][:, perm].tocsr())):
        A.sort_indices()
        s = core.Solver(None, ctx=ctx, indptr=A.indptr, indices=A.indices, precond=L.PRECOND_SSOR)
        s.set_values(A.data)
        b = np.random.default_rng(9).normal(size=n)
        s.set_rhs(b)
        systems.append((name, s, A, b))
    print(f"[gpu] SSOR natural and shuffled numbering, colours {[s.ssor_info()[0] for _, s, _, _ in systems]}",
          flush=True)
    solo = []
    for name, s, A, b in systems:
        x, it, st = s.solve(None, rtol=1e-12, maxiter=3000)
        xo, ito, _, _, _ = O.pcg_ssor(A.indptr, A.indices, A.data, b, rtol=1e-12, maxiter=3000, rows_per_thread=1,
                                      nthreads=8)
        assert st == 0 and it == ito, (name, it, ito)
        assert_bitwise(x, xo, f"{name} numbering SSOR vs oracle")
        solo.append((x, it))
        s.reset_history()
    its = core.batch_solve([s for _, s, _, _ in systems], rtol=1e-12, maxiter=3000)
    for (name, s, _, _), (x, it), itb in zip(systems, solo, its):
        assert itb == it
        assert_bitwise(s.get_solution(), x, f"{name} numbering batched SSOR")
