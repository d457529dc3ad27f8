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
    del srv_vals


def test_admission_budget(ctx):
    m, s = _mesh(ctx, 30, 20, 3)
    srv = core.Server(ctx, rtol=1e-12, maxiter=1000, check_every=10, hbm_budget_bytes=1024)
    mid = srv.add_mesh(m)
    t = srv.submit(mid, cases.normal(5, m.n))
    srv.step()
    st, x, it, lat = srv.poll(t, m.n)
    assert st == L.GSH_E_ADMISSION


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
