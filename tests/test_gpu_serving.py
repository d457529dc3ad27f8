"""Serving: continuous batching admits requests as slots free up; each result equals the
same request solved alone, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from gschism_amd import _lib as L
from gschism_amd import cases, core

pytestmark = pytest.mark.gpu


def _mesh(ctx, nx, ny, seed_shift, precision=L.PREC_FP64):
    spec = cases.CaseSpec(f"srv{nx}x{ny}", nx, ny, 700.0, nvrt=4, depth=None, substeps=4)
    data = cases.make_case(spec)
    mesh, p, _ = cases.load_mesh(ctx, data)
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    s = core.Solver(mesh, precision)
    s.assemble(p)
    return mesh, s


def test_alone_vs_mix(ctx):
    meshes = [_mesh(ctx, 40, 30, 0), _mesh(ctx, 90, 25, 1), _mesh(ctx, 31, 57, 2)]
    rhs = [cases.normal(300 + k, m.n) for k, (m, _) in enumerate(meshes)]
    # alone: each request through the plain solver
    alone = []
    for (m, s), b in zip(meshes, rhs):
        s.reset_history()
        x, it, st = s.solve(b, rtol=1e-12, maxiter=3000, check_every=10)
        alone.append((x, it))
    srv = core.Server(ctx, rtol=1e-12, maxiter=3000, check_every=10, max_active=2)
    ids = [srv.add_mesh(m) for m, _ in meshes]
    # the request mix: staggered submissions; the mesh matrices are assembled by the server
    # from each mesh's state, so the right-hand sides are passed explicitly here
    tickets = []
    srv_vals = [s.get_values() for _, s in meshes]
    order = [0, 1, 2, 1, 0, 2]
    for j, k in enumerate(order):
        tickets.append((k, srv.submit(ids[k], rhs[k])))
        if j % 2 == 1:
            srv.step()
    for _ in range(2000):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    for k, t in tickets:
        st, x, it, lat = srv.poll(t, meshes[k][0].n)
        assert st == 0, st
        assert it == alone[k][1], (k, it, alone[k][1])
        assert np.array_equal(x, alone[k][0]), f"request on mesh {k} differs from the solo solve"
        assert lat > 0.0
    # released requests return their page-locked vectors to the pool; a second round reuses
    # them (right-hand sides overwrite the old contents) with the same results
    for _, t in tickets:
        srv.release(t)
        with pytest.raises(L.GshError) as e:
            srv.poll(t, 1)
        assert e.value.code == L.GSH_E_NOTFOUND
    again = [(k, srv.submit(ids[k], rhs[k])) for k in (2, 0, 1)]
    with pytest.raises(L.GshError) as e:  # queued: not releasable
        srv.release(again[0][1])
    assert e.value.code == L.GSH_E_STATE
    for _ in range(2000):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    for k, t in again:
        st, x, it, lat = srv.poll(t, meshes[k][0].n)
        assert st == 0 and it == alone[k][1]
        assert np.array_equal(x, alone[k][0]), f"request on mesh {k} differs after buffer reuse"
        srv.release(t)
    del srv_vals


def test_admission_budget(ctx):
    m, s = _mesh(ctx, 30, 20, 3)
    srv = core.Server(ctx, rtol=1e-12, maxiter=1000, check_every=10, hbm_budget_bytes=1024)
    mid = srv.add_mesh(m)
    t = srv.submit(mid, cases.normal(5, m.n))
    srv.step()
    st, x, it, lat = srv.poll(t, m.n)
    assert st == L.GSH_E_ADMISSION
    srv.release(t)  # a rejected request is released like a finished one


def test_mixed_alone_vs_mix(ctx):
    """Mixed-precision serving: every result equals the request's solo mixed solve bit for bit
    and is within 1e-8 of the FP64 solution."""
    meshes = [_mesh(ctx, 44, 30, 0, L.PREC_MIXED), _mesh(ctx, 80, 26, 1, L.PREC_MIXED),
              _mesh(ctx, 35, 52, 2, L.PREC_MIXED)]
    rhs = [cases.normal(400 + k, m.n) for k, (m, _) in enumerate(meshes)]
    alone = []
    for (m, s), b in zip(meshes, rhs):
        s.reset_history()
        x, it, st = s.solve(b, rtol=1e-12, maxiter=3000, check_every=10)
        assert st == 0
        f64 = core.Solver(m)
        f64.set_values(s.get_values())
        x64, _, _ = f64.solve(b, rtol=1e-12, maxiter=3000)
        assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-8
        alone.append((x, it))
    srv = core.Server(ctx, rtol=1e-12, maxiter=3000, check_every=10, max_active=2, precision=L.PREC_MIXED)
    ids = [srv.add_mesh(m) for m, _ in meshes]
    tickets = []
    for j, k in enumerate([2, 0, 1, 0, 2, 1, 1]):
        tickets.append((k, srv.submit(ids[k], rhs[k])))
        if j % 3 == 2:
            srv.step()
    for _ in range(5000):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    for k, t in tickets:
        st, x, it, lat = srv.poll(t, meshes[k][0].n)
        assert st == 0, st
        assert it == alone[k][1], (k, it, alone[k][1])
        assert np.array_equal(x, alone[k][0]), f"mixed request on mesh {k} differs from the solo solve"


def _ssor_alone(ctx, meshes, rhs, precision):
    """Solo SSOR solves of each request (2 rows per thread, as the server's solvers)."""
    alone = []
    for (m, s), b in zip(meshes, rhs):
        d = core.Solver(m, precision, precond=L.PRECOND_SSOR)
        d.set_tuning(2)
        d.set_values(s.get_values())
        x, it, st = d.solve(b, rtol=1e-12, maxiter=3000, check_every=10)
        assert st == 0
        alone.append((x, it))
        d.close()
    return alone


def _serve(srv, ids, rhs, order, every):
    tickets = []
    for j, k in enumerate(order):
        tickets.append((k, srv.submit(ids[k], rhs[k])))
        if j % every == every - 1:
            srv.step()
    for _ in range(5000):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    return tickets


def test_ssor_alone_vs_mix(ctx):
    """FP64 SSOR serving: each request, set up on the staging stream and advanced in the batched
    colour-step launches with the others, equals its solo SSOR solve bit for bit; a zero
    right-hand side returns x = 0 at once."""
    meshes = [_mesh(ctx, 40, 30, 0), _mesh(ctx, 90, 25, 1), _mesh(ctx, 31, 57, 2)]
    rhs = [cases.normal(500 + k, m.n) for k, (m, _) in enumerate(meshes)]
    rhs.append(np.zeros(meshes[1][0].n))
    alone = _ssor_alone(ctx, meshes + [meshes[1]], rhs, L.PREC_FP64)
    assert alone[3][1] == 0
    srv = core.Server(ctx, rtol=1e-12, maxiter=3000, check_every=10, max_active=3, precond=L.PRECOND_SSOR)
    ids = [srv.add_mesh(m) for m, _ in meshes]
    ids.append(ids[1])
    print("[gpu] SSOR server, 7 requests", flush=True)
    tickets = _serve(srv, ids, rhs, [0, 1, 3, 2, 1, 0, 2], 2)
    for k, t in tickets:
        st, x, it, lat = srv.poll(t, meshes[1 if k == 3 else k][0].n)
        assert st == 0, st
        assert it == alone[k][1], (k, it, alone[k][1])
        assert np.array_equal(x, alone[k][0]), f"SSOR request {k} differs from the solo solve"
    srv.close()


def test_mixed_ssor_alone_vs_mix(ctx):
    """Mixed-precision SSOR serving (FP32 SSOR inner windows, FP64 residual checks): every result
    equals the request's solo mixed SSOR solve bit for bit and is within 1e-8 of FP64."""
    meshes = [_mesh(ctx, 44, 30, 0, L.PREC_MIXED), _mesh(ctx, 80, 26, 1, L.PREC_MIXED),
              _mesh(ctx, 35, 52, 2, L.PREC_MIXED)]
    rhs = [cases.normal(600 + k, m.n) for k, (m, _) in enumerate(meshes)]
    alone = _ssor_alone(ctx, meshes, rhs, L.PREC_MIXED)
    ref = _ssor_alone(ctx, meshes, rhs, L.PREC_FP64)
    for (x, _), (x64, _) in zip(alone, ref):
        assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-8
    srv = core.Server(ctx, rtol=1e-12, maxiter=3000, check_every=10, max_active=2, precision=L.PREC_MIXED,
                      precond=L.PRECOND_SSOR)
    ids = [srv.add_mesh(m) for m, _ in meshes]
    print("[gpu] mixed SSOR server, 7 requests", flush=True)
    tickets = _serve(srv, ids, rhs, [2, 0, 1, 0, 2, 1, 1], 3)
    for k, t in tickets:
        st, x, it, lat = srv.poll(t, meshes[k][0].n)
        assert st == 0, st
        assert it == alone[k][1], (k, it, alone[k][1])
        assert np.array_equal(x, alone[k][0]), f"mixed SSOR request on mesh {k} differs from the solo solve"
    srv.close()


def test_served_request_uses_mesh_step_params(ctx):
    """A served request (backtracking, RHS producers and assembly at admission, no RHS given) with
    non-default step parameters set for its mesh equals the solo sequence with the same
    parameters bit for bit, and both equal the oracle; open nodes of type 0 bring in the
    open-boundary integral."""
    from conftest import assert_bitwise, oracle_mesh_for
    from oracle import pyoracle as O

    spec = cases.CaseSpec("srvp-36x24", 36, 24, 600.0, nvrt=5, depth=None, substeps=4)
    data = cases.make_case(spec)
    mesh, _, (eta, uvw, su2) = cases.load_mesh(ctx, data)
    p = L.default_step_params(dt=90.0, theta=0.7, dtb_min=15.0, dtb_max=30.0, iettype=2)
    opn = np.flatnonzero(mesh.topology(L.T_OPEN_BND))
    etype = np.full(mesh.n, -1, dtype=np.int8)
    etype[opn[::2]] = 0
    eth = np.where(mesh.topology(L.T_OPEN_BND) != 0, 0.07, 0.0)
    om = oracle_mesh_for(mesh, data.sigma)
    for m in (mesh, om):
        m.set(L.F_ETYPE, etype)
        m.set(L.F_ETH, eth)
    om.set(L.F_ETA2, eta)
    om.set(L.F_NODE_UVW, uvw)
    om.set(L.F_SIDE_U2, su2)
    om.levels(p)
    # solo
    solo = core.Solver(mesh)
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    solo.assemble(p)
    x_solo, it_solo, st = solo.solve(None, rtol=1e-12, maxiter=3000, check_every=10)
    assert st == 0
    # oracle
    om.backtrack(p, nthreads=8)
    om.rhs(p)
    val, b = om.assemble(p)
    assert_bitwise(solo.get_values(), val, "matrix with the request's step parameters")
    assert_bitwise(solo.get_rhs(), b, "RHS with the request's step parameters")
    xo, ito, _, _ = O.pcg(mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES), val, b, rtol=1e-12,
                          maxiter=3000, check_every=10)
    assert ito == it_solo
    assert_bitwise(x_solo, xo, "solo solution")
    # served, twice (the second on a pooled solver)
    srv = core.Server(ctx, rtol=1e-12, maxiter=3000, check_every=10, max_active=2)
    mid = srv.add_mesh(mesh, params=p)
    print("[gpu] served request with mesh step parameters", flush=True)
    for _ in range(2):
        t = srv.submit(mid, None, with_backtrack=True)
        for _ in range(2000):
            a, q = srv.step()
            if a == 0 and q == 0:
                break
        st, x, it, lat = srv.poll(t, mesh.n)
        assert st == 0 and it == it_solo
        assert_bitwise(x, x_solo, "served solution with the mesh's step parameters")
    bad = L.default_step_params(iettype=9)
    with pytest.raises(L.GshError) as ei:
        srv.set_mesh_params(mid, bad)
    assert ei.value.code == L.GSH_E_ARG
    srv.close()
    solo.close()
    mesh.close()


def _pool_mesh(ctx, k):
    """Mesh k of the served configuration's pool (n log-uniform from 1e4; seeds 3 + k)."""
    from math import sqrt

    n = int(round(1e4 * 200 ** (k / 5)))
    nx = int(round(sqrt(2 * n)))
    ny = max(3, n // nx)
    spec = cases.CaseSpec(f"pool{k}-{nx}x{ny}", nx, ny, 1000.0, nvrt=11, depth=None, substeps=10, corr_cells=20.0)
    data = cases.make_case(spec)
    data.depth = cases.bathymetry(nx, ny, 20.0, 5.0, 4000.0, seed=3 + k)
    mesh, _, state = cases.load_mesh(ctx, data)
    return mesh, data, state


@pytest.mark.parametrize("precond", [L.PRECOND_JACOBI, L.PRECOND_SSOR])
def test_served_pool_against_oracle(ctx, precond):
    """The served configuration's two smallest pool meshes (10k and 29k nodes, 11 levels, 10
    backtracking substeps): requests with backtracking, served in a mix, in FP64 equal the
    oracle's backtracking + assembly + PCG bit for bit (Jacobi at 4 rows per thread, SSOR at 2),
    and in mixed precision are within 1e-8 of it."""
    from conftest import oracle_mesh_for
    from oracle import pyoracle as O

    pd = L.default_step_params()
    pool, ref = [], []
    for k in (0, 1):
        mesh, data, (eta, uvw, su2) = _pool_mesh(ctx, k)
        om = oracle_mesh_for(mesh, data.sigma)
        om.set(L.F_ETA2, eta)
        om.set(L.F_NODE_UVW, uvw)
        om.set(L.F_SIDE_U2, su2)
        om.levels(pd)
        print(f"[gpu] pool mesh {k}: n={mesh.n}, oracle backtracking", flush=True)
        om.backtrack(pd, nthreads=16)
        om.rhs(pd)
        val, _ = om.assemble(pd)
        ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
        pool.append(mesh)
        ref.append((ip, ix, val))
    which = [0, 1, 1, 0, 1]
    rhs = [cases.normal(1000 + j, pool[k].n) for j, k in enumerate(which)]
    xref = []
    for j, k in enumerate(which):
        ip, ix, val = ref[k]
        if precond == L.PRECOND_JACOBI:
            x, it, _, _ = O.pcg(ip, ix, val, rhs[j], rtol=1e-12, maxiter=20000, check_every=10, rows_per_thread=4,
                                nthreads=16)
        else:
            x, it, _, _, _ = O.pcg_ssor(ip, ix, val, rhs[j], rtol=1e-12, maxiter=20000, check_every=10,
                                        rows_per_thread=2, nthreads=16)
        xref.append((x, it))
    for precision in (L.PREC_FP64, L.PREC_MIXED):
        srv = core.Server(ctx, rtol=1e-12 if precision == L.PREC_FP64 else 1e-10, maxiter=20000, check_every=10,
                          max_active=3, precision=precision, precond=precond)
        ids = [srv.add_mesh(m) for m in pool]
        print(f"[gpu] served pool, precision {precision}, precond {precond}", flush=True)
        tickets = []
        for j, k in enumerate(which):
            tickets.append(srv.submit(ids[k], rhs[j], with_backtrack=True))
            if j % 2 == 1:
                srv.step()
        for _ in range(20000):
            a, q = srv.step()
            if a == 0 and q == 0:
                break
        for j, t in enumerate(tickets):
            st, x, it, lat = srv.poll(t, pool[which[j]].n)
            assert st == 0, (j, st)
            x64, it64 = xref[j]
            rel = np.linalg.norm(x - x64) / np.linalg.norm(x64)
            print(f"[gpu]   request {j} (mesh {which[j]}): iterations {it} (oracle FP64 {it64}), rel L2 {rel:.2e}",
                  flush=True)
            if precision == L.PREC_FP64:
                assert it == it64
                assert np.array_equal(x, x64), f"served FP64 request {j} differs from the oracle"
            else:
                assert rel <= 1e-8
        srv.close()
    for m in pool:
        m.close()
