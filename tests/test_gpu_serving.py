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
