"""GPU parity: every stage of the hot path through the C ABI against the CPU oracle, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_bitwise, oracle_mesh_for
from gschism_amd import _lib as L
from gschism_amd import cases, core
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _setup(ctx, spec, kb_fn=None, depth_fn=None):
    data = cases.make_case(spec)
    if depth_fn is not None:
        data.depth = depth_fn(data.xy)
    kb = None if kb_fn is None else kb_fn(data.depth)
    mesh = core.Mesh(ctx, data.xy, data.elnode, data.depth, data.sigma, open_bnd=data.open_bnd, kb=kb)
    xy = mesh.topology(L.T_XY)
    eta, uvw, su2 = cases.initial_state(spec, xy, cases.side_centres(mesh), data.sigma)
    eta = eta + cases.normal(cases.SEED_ETA, spec.n, 0.0, spec.eta_noise)[mesh.node_perm]
    p = cases.step_params(spec)
    mesh.set(L.F_ETA2, eta)
    mesh.set(L.F_NODE_UVW, uvw)
    mesh.set(L.F_SIDE_U2, su2)
    mesh.update_levels(p)
    om = oracle_mesh_for(mesh, data.sigma)
    om.set(L.F_ETA2, eta)
    om.set(L.F_NODE_UVW, uvw)
    om.set(L.F_SIDE_U2, su2)
    om.levels(p)
    return mesh, om, p


def _full_chain(ctx, mesh, om, p, rows_per_thread=4):
    print(f"[gpu] chain n={mesh.n} ns={mesh.ns} nvrt={mesh.nvrt}", flush=True)
    assert_bitwise(mesh.get(L.F_ZC), om.get(L.F_ZC), "zc")
    assert_bitwise(mesh.get(L.F_WET), om.get(L.F_WET), "wet flags")
    mesh.backtrack(p)
    om.backtrack(p, nthreads=8)
    assert_bitwise(mesh.get(L.F_FOOT), om.get(L.F_FOOT), "foot elements")
    assert_bitwise(mesh.get(L.F_BTU), om.get(L.F_BTU), "backtracked velocity")
    assert_bitwise(mesh.get(L.F_BSTATUS), om.get(L.F_BSTATUS), "backtrack status")
    mesh.rhs_producers(p)
    om.rhs(p)
    for f, name in ((L.F_SIDE_U, "U"), (L.F_SIDE_HHAT, "hhat"), (L.F_SIDE_CHI, "chi"), (L.F_GHAT1, "ghat1"),
                    (L.F_ELBC, "elbc")):
        assert_bitwise(mesh.get(f), om.get(f), name)
    solver = core.Solver(mesh)
    solver.set_tuning(rows_per_thread)
    solver.assemble(p)
    val, b = om.assemble(p)
    assert_bitwise(solver.get_values(), val, "CSR values")
    assert_bitwise(solver.get_rhs(), b, "RHS qel")
    x, it, st = solver.solve(None, rtol=1e-12, maxiter=1500, check_every=10)
    xo, ito, rr, at = O.pcg(mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES), val, b, rtol=1e-12,
                            maxiter=1500, check_every=10, rows_per_thread=rows_per_thread)
    assert st == 0 and it == ito, (it, ito, st)
    assert_bitwise(x, xo, "PCG solution")
    rel = solver.check()
    assert rel < 1e-10
    stt = solver.stats()
    assert stt.atol2 == at and stt.rr == rr
    # the step after the solve: new elevation, then the momentum solve for the side velocities
    eta_old = mesh.get(L.F_ETA2)
    solver.update_elevation()
    om.set(L.F_ETA1, eta_old)
    om.set(L.F_ETA2, xo)
    mesh.momentum(p)
    om.momentum(p)
    assert_bitwise(mesh.get(L.F_SIDE_U2), om.get(L.F_SIDE_U2), "side velocity from the momentum step")
    mesh.vertical_velocity(p)
    om.vertical_velocity()
    assert_bitwise(mesh.get(L.F_ELEM_W), om.get(L.F_ELEM_W), "element vertical velocity")
    # the levels of the new elevation, then the nodal velocities the next step backtracks through
    mesh.update_levels(p)
    om.levels(p)
    assert_bitwise(mesh.get(L.F_ZC), om.get(L.F_ZC), "zc of the new elevation")
    mesh.nodal_velocity(p)
    om.nodal_velocity(p, nthreads=4)
    assert_bitwise(mesh.get(L.F_NODE_UVW), om.get(L.F_NODE_UVW), "nodal velocities of the next step")
    return solver, x, val, b


def test_chain_channel_200x80(ctx):
    spec = cases.spec_cpu_reference()
    mesh, om, p = _setup(ctx, spec)
    _full_chain(ctx, mesh, om, p)
    st = mesh.get(L.F_BSTATUS)
    assert (st & L.BT_OPEN_EXIT).any() and (st & L.BT_SLIDE).any()


def _island_depth(xy):
    cx, cy = xy[:, 0].mean(), xy[:, 1].mean()
    r2 = ((xy[:, 0] - cx) ** 2 + (xy[:, 1] - cy) ** 2) / (4000.0 ** 2)
    return 12.0 - 16.0 * np.exp(-r2)


def _lsc_kb(depth):
    return np.clip((6 - depth).astype(np.int32), 0, 6)


def test_chain_island_dry_and_kb(ctx):
    """Dry elements (an island above the water) and node-varying bottom levels."""
    spec = cases.CaseSpec("island-48x32", 48, 32, 400.0, nvrt=12, depth=10.0, substeps=6)
    mesh, om, p = _setup(ctx, spec, kb_fn=_lsc_kb, depth_fn=_island_depth)
    wet = mesh.get(L.F_WET)
    assert (wet == 0).any(), "the island should leave dry entities"
    assert (mesh.topology(L.T_KB) > 0).any()
    _full_chain(ctx, mesh, om, p)


def test_nodal_velocity_then_backtrack(ctx):
    """Nodal velocities from random side velocities and element w (dry island, node-varying
    bottom levels, clamping active) bit for bit, then backtracking through them."""
    spec = cases.CaseSpec("island-48x32", 48, 32, 400.0, nvrt=12, depth=10.0, substeps=6)
    mesh, om, p = _setup(ctx, spec, kb_fn=_lsc_kb, depth_fn=_island_depth)
    rng = np.random.default_rng(21)
    su2 = rng.normal(0.0, 0.5, (mesh.ns, mesh.nvrt, 2))
    ew = rng.normal(0.0, 1e-3, (mesh.ne, mesh.nvrt))
    p.rmaxvel = 0.9
    for m in (mesh, om):
        m.set(L.F_SIDE_U2, su2)
        m.set(L.F_ELEM_W, ew)
    print(f"[gpu] nodal velocity n={mesh.n} nvrt={mesh.nvrt}", flush=True)
    mesh.nodal_velocity(p)
    om.nodal_velocity(p, nthreads=4)
    uvw = mesh.get(L.F_NODE_UVW)
    assert_bitwise(uvw, om.get(L.F_NODE_UVW), "nodal velocities")
    assert np.abs(uvw[..., 0]).max() == 0.9
    mesh.backtrack(p)
    om.backtrack(p, nthreads=8)
    assert_bitwise(mesh.get(L.F_FOOT), om.get(L.F_FOOT), "foot elements")
    assert_bitwise(mesh.get(L.F_BTU), om.get(L.F_BTU), "backtracked velocity")


@pytest.mark.parametrize("nvrt", [12, 2])
def test_backtrack_lds_staging(ctx, nvrt):
    """The LDS-staged element walk (dry island, node-varying kb, open and land boundaries) gives
    the oracle's feet, velocities and flags bit for bit; at 2 levels a workgroup spans 128 sides,
    more than the table stages, so staged and global element reads mix."""
    spec = cases.CaseSpec(f"island-48x32-{nvrt}", 48, 32, 400.0, nvrt=nvrt, depth=10.0, substeps=6)
    mesh, om, p = _setup(ctx, spec, kb_fn=(lambda d: np.minimum(_lsc_kb(d), nvrt - 2)), depth_fn=_island_depth)
    om.backtrack(p, nthreads=8)
    for staged in (True, False, True):
        print(f"[gpu] backtrack staged={staged} nvrt={nvrt}", flush=True)
        mesh.set_backtrack_staging(staged)
        mesh.backtrack(p)
        assert_bitwise(mesh.get(L.F_FOOT), om.get(L.F_FOOT), f"foot elements (staged={staged})")
        assert_bitwise(mesh.get(L.F_BTU), om.get(L.F_BTU), f"backtracked velocity (staged={staged})")
        assert_bitwise(mesh.get(L.F_BSTATUS), om.get(L.F_BSTATUS), f"backtrack status (staged={staged})")
    st = mesh.get(L.F_BSTATUS)
    assert (st & L.BT_SLIDE).any()
    assert L.lib().gsh_mesh_set_backtrack_staging(mesh.h, 2) == L.GSH_E_ARG


@pytest.mark.parametrize("rpt", [1, 2, 8])
def test_rows_per_thread(ctx, rpt):
    spec = cases.CaseSpec("channel-64x40", 64, 40, 500.0, nvrt=6, depth=None, substeps=4)
    mesh, om, p = _setup(ctx, spec)
    _full_chain(ctx, mesh, om, p, rows_per_thread=rpt)


def _wheel(k=24, r1=1500.0, r2=3000.0):
    """A centre node with k spokes (CSR row of k + 1 entries) inside two rings of k nodes."""
    a1 = 2 * np.pi * np.arange(k) / k
    a2 = a1 + np.pi / k
    xy = np.concatenate([[[0.0, 0.0]], np.c_[r1 * np.cos(a1), r1 * np.sin(a1)],
                         np.c_[r2 * np.cos(a2), r2 * np.sin(a2)]])
    tris = []
    for i in range(k):
        j = (i + 1) % k
        tris += [(0, 1 + i, 1 + j), (1 + i, 1 + k + i, 1 + j), (1 + j, 1 + k + i, 1 + k + j)]
    el = np.array(tris, dtype=np.int32)
    p0, p1, p2 = xy[el[:, 0]], xy[el[:, 1]], xy[el[:, 2]]
    cw = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1]) < 0
    el[cw] = el[cw][:, [0, 2, 1]]
    open_bnd = np.zeros(xy.shape[0], dtype=np.int8)
    open_bnd[1 + k:1 + k + 5] = 1
    return xy, el, open_bnd


def test_chain_high_degree_node(ctx):
    """Rows longer than the assembly's LDS row buffer (16 entries) take the in-place path."""
    spec = cases.CaseSpec("wheel-24", 13, 13, 500.0, nvrt=11, depth=10.0, substeps=5)
    xy, el, ob = _wheel()
    sigma = cases.sigma_levels(spec.nvrt)
    depth = 10.0 + 0.001 * xy[:, 0]
    mesh = core.Mesh(ctx, xy, el, depth, sigma, open_bnd=ob)
    ip = mesh.topology(L.T_CSR_INDPTR)
    assert np.diff(ip).max() == 25
    xym = mesh.topology(L.T_XY)
    eta = 0.1 * np.sin(2 * np.pi * xym[:, 0] / 6000.0)
    uvw = cases.velocity(xym, sigma, 6000.0, 6000.0, 500.0, 3)
    su2 = cases.velocity(cases.side_centres(mesh), sigma, 6000.0, 6000.0, 500.0, 2)
    p = cases.step_params(spec)
    om = oracle_mesh_for(mesh, sigma)
    for m in (mesh, om):
        m.set(L.F_ETA2, eta)
        m.set(L.F_NODE_UVW, uvw)
        m.set(L.F_SIDE_U2, su2)
    mesh.update_levels(p)
    om.levels(p)
    _full_chain(ctx, mesh, om, p, rows_per_thread=1)


def test_three_steps_with_history(ctx):
    """Consecutive steps: extrapolated initial guess 3x0-3x1+x2, eta update, levels update."""
    spec = cases.CaseSpec("channel-60x30", 60, 30, 500.0, nvrt=8, depth=None, substeps=5)
    mesh, om, p = _setup(ctx, spec)
    solver = core.Solver(mesh)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    hist = [np.zeros(mesh.n)] * 3
    for step in range(3):
        print(f"[gpu] step {step}", flush=True)
        mesh.backtrack(p)
        om.backtrack(p)
        mesh.rhs_producers(p)
        om.rhs(p)
        solver.assemble(p)
        val, b = om.assemble(p)
        x, it, st = solver.solve(None, rtol=1e-12, maxiter=1500)
        xo, ito, _, _ = O.pcg(ip, ix, val, b, hist=hist, rtol=1e-12, maxiter=1500)
        assert it == ito
        assert_bitwise(x, xo, f"step {step} solution")
        hist = [xo, hist[0], hist[1]]
        eta_old = mesh.get(L.F_ETA2)
        solver.update_elevation()
        assert_bitwise(mesh.get(L.F_ETA1), eta_old, "eta1 roll")
        assert_bitwise(mesh.get(L.F_ETA2), xo, "eta2 update")
        om.set(L.F_ETA1, eta_old)
        om.set(L.F_ETA2, xo)
        # side velocities of the next step (they feed its vertically integrated velocity)
        mesh.momentum(p)
        om.momentum(p)
        assert_bitwise(mesh.get(L.F_SIDE_U2), om.get(L.F_SIDE_U2), f"step {step} side velocity")
        mesh.vertical_velocity(p)
        om.vertical_velocity()
        assert_bitwise(mesh.get(L.F_ELEM_W), om.get(L.F_ELEM_W), f"step {step} element w")
        mesh.update_levels(p)
        om.levels(p)
        assert_bitwise(mesh.get(L.F_ZC), om.get(L.F_ZC), f"step {step} zc")
        # the next step backtracks through the velocities this step produced
        mesh.nodal_velocity(p)
        om.nodal_velocity(p, nthreads=4)
        assert_bitwise(mesh.get(L.F_NODE_UVW), om.get(L.F_NODE_UVW), f"step {step} nodal velocities")


def test_backtrack_list_matches_full(ctx):
    spec = cases.CaseSpec("channel-50x30", 50, 30, 500.0, nvrt=10, depth=10.0, substeps=5)
    mesh, om, p = _setup(ctx, spec)
    mesh.backtrack(p)
    full_b, full_f = mesh.get(L.F_BTU), mesh.get(L.F_FOOT)
    rng = np.random.default_rng(5)
    sides = rng.integers(0, mesh.ns, 500).astype(np.int32)
    levs = rng.integers(0, mesh.nvrt, 500).astype(np.int32)
    mesh.set(L.F_BTU, np.zeros_like(full_b))
    mesh.backtrack_list(p, sides, levs)
    got_b, got_f = mesh.get(L.F_BTU), mesh.get(L.F_FOOT)
    assert_bitwise(got_b[sides, levs], full_b[sides, levs], "list-mode btu")
    assert_bitwise(got_f[sides, levs], full_f[sides, levs], "list-mode foot")
    om.backtrack(p, sides, levs)
    assert_bitwise(om.get(L.F_FOOT)[sides, levs], full_f[sides, levs], "oracle list-mode foot")


def _laplace_csr(nx, ny, shift):
    n = nx * ny
    rows, cols, vals = [], [], []
    for j in range(ny):
        for i in range(nx):
            k = j * nx + i
            rows.append(k), cols.append(k), vals.append(4.0 + shift)
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                a, bb = i + di, j + dj
                if 0 <= a < nx and 0 <= bb < ny:
                    rows.append(k), cols.append(bb * nx + a), vals.append(-1.0)
    import scipy.sparse as sp

    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return A


def test_standalone_csr_against_scipy_and_oracle(ctx):
    import scipy.sparse.linalg as sla

    A = _laplace_csr(70, 50, 0.05)
    b = cases.normal(11, A.shape[0])
    s = core.Solver(None, ctx=ctx, indptr=A.indptr, indices=A.indices)
    s.set_values(A.data)
    x, it, st = s.solve(b, rtol=1e-12, maxiter=5000)
    xo, ito, _, _ = O.pcg(A.indptr, A.indices, A.data, b, rtol=1e-12, maxiter=5000)
    assert st == 0 and it == ito
    assert_bitwise(x, xo, "stand-alone solution")
    xs = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-9


def test_zero_rhs_and_maxiter(ctx):
    A = _laplace_csr(30, 20, 0.1)
    s = core.Solver(None, ctx=ctx, indptr=A.indptr, indices=A.indices)
    s.set_values(A.data)
    x, it, st = s.solve(np.zeros(A.shape[0]), rtol=1e-12, maxiter=100)
    assert it == 0 and st == 0 and not x.any()
    b = cases.normal(12, A.shape[0])
    s.reset_history()
    x, it, st = s.solve(b, rtol=1e-30, maxiter=25, check_every=10)
    assert it == 25 and st == L.GSH_W_MAXITER
    xo, ito, _, _ = O.pcg(A.indptr, A.indices, A.data, b, rtol=1e-30, maxiter=25, check_every=10)
    assert ito == 25
    assert_bitwise(x, xo, "maxiter solution")


def test_mixed_precision_accuracy(ctx):
    spec = cases.CaseSpec("channel-96x64-lognormal", 96, 64, 1000.0, nvrt=6, depth=None, substeps=5)
    mesh, om, p = _setup(ctx, spec)
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    s64 = core.Solver(mesh, L.PREC_FP64)
    s64.assemble(p)
    x64, it64, _ = s64.solve(None, rtol=1e-12, maxiter=5000)
    smx = core.Solver(mesh, L.PREC_MIXED)
    smx.assemble(p)
    xm, itm, stm = smx.solve(None, rtol=1e-12, maxiter=5000)
    assert stm == 0
    err = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    print(f"[gpu] mixed rel err {err:.3e}, inner iters {itm} vs fp64 {it64}, outer {smx.stats().outer}")
    assert err <= 1e-8


def test_batch_equals_solo(ctx):
    solvers, solo = [], []
    for k, (nx, ny) in enumerate(((40, 30), (64, 20), (33, 47))):
        spec = cases.CaseSpec(f"b{k}", nx, ny, 500.0, nvrt=5, depth=None, substeps=4)
        mesh, om, p = _setup(ctx, spec)
        mesh.backtrack(p)
        mesh.rhs_producers(p)
        s = core.Solver(mesh)
        s.assemble(p)
        x, it, st = s.solve(None, rtol=1e-12, maxiter=3000)
        s.reset_history()
        solo.append((x, it))
        solvers.append((s, mesh))
    its = core.batch_solve([s for s, _ in solvers], rtol=1e-12, maxiter=3000)
    for (s, mesh), (x, it), itb in zip(solvers, solo, its):
        assert itb == it
        s.update_elevation()
        assert_bitwise(mesh.get(L.F_ETA2), x, "batched solution")


def test_depth_field_and_levels(ctx):
    """Setting the bathymetry on the device changes levels and wet flags as the oracle's."""
    spec = cases.CaseSpec("depth-40x30", 40, 30, 500.0, nvrt=6, depth=None, substeps=4)
    mesh, om, p = _setup(ctx, spec)
    d = mesh.topology(L.T_DEPTH) * (1.0 + 0.05 * cases.normal(123, mesh.n)) - 3.0
    mesh.set(L.F_DEPTH, d)
    om.set(L.F_DEPTH, d)
    assert_bitwise(mesh.get(L.F_DEPTH), d, "depth readback")
    assert_bitwise(mesh.topology(L.T_DEPTH), d, "depth topology copy")
    mesh.update_levels(p)
    om.levels(p)
    assert_bitwise(mesh.get(L.F_ZC), om.get(L.F_ZC), "zc after depth change")
    assert_bitwise(mesh.get(L.F_WET), om.get(L.F_WET), "wet flags after depth change")


def test_mixed_batch_equals_solo(ctx):
    """Mixed precision in lock-step: each system's result is the one it gets alone, and within
    1e-8 of the FP64 solution."""
    systems = []
    for k, (nx, ny) in enumerate(((50, 30), (28, 61), (70, 22), (45, 45))):
        spec = cases.CaseSpec(f"m{k}", nx, ny, 800.0, nvrt=4, depth=None, substeps=4)
        mesh, om, p = _setup(ctx, spec)
        mesh.backtrack(p)
        mesh.rhs_producers(p)
        f64 = core.Solver(mesh)
        f64.assemble(p)
        x64, _, _ = f64.solve(None, rtol=1e-12, maxiter=3000)
        mx = core.Solver(mesh, L.PREC_MIXED)
        mx.assemble(p)
        xm, itm, stm = mx.solve(None, rtol=1e-12, maxiter=3000)
        assert stm == 0
        assert np.linalg.norm(xm - x64) / np.linalg.norm(x64) < 1e-8
        mx.reset_history()
        systems.append((mesh, mx, xm, itm))
    print("[gpu] mixed batch of 4", flush=True)
    its = core.batch_solve([m for _, m, _, _ in systems], rtol=1e-12, maxiter=3000)
    for (mesh, mx, xm, itm), itb in zip(systems, its):
        assert itb == itm
        assert_bitwise(mx.get_solution(), xm, "mixed batched solution")


def _timeout_then_same_result(ctx, mesh, solver, val, b, xo, ito):
    """A 1M solve whose host wait gives up at once (timeout 1 us: GSH_E_TIMEOUT while the
    launched kernels still run to their normal end, pcg_loop.cpp:444-456); the next solve of the
    same solver, with the history reset, reproduces the oracle's solution and iteration count.
    The same for the SSOR solver against the oracle's SSOR PCG."""
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    print("[gpu] 1M solve with a 1 us time limit", flush=True)
    solver.reset_history()
    with pytest.raises(L.GshError) as ei:
        solver.solve(None, rtol=1e-12, maxiter=3000, timeout_s=1e-6)
    assert ei.value.code == L.GSH_E_TIMEOUT
    solver.reset_history()
    x, it, st = solver.solve(None, rtol=1e-12, maxiter=3000)
    assert st == 0 and it == ito
    assert_bitwise(x, xo, "1M PCG solution after a timed-out solve")
    ss = core.Solver(mesh, L.PREC_FP64, precond=L.PRECOND_SSOR)
    ss.set_tuning(2)
    ss.set_values(val)
    print("[gpu] 1M SSOR solve with a 1 us time limit", flush=True)
    with pytest.raises(L.GshError) as ei:
        ss.solve(b, rtol=1e-12, maxiter=3000, timeout_s=1e-6)
    assert ei.value.code == L.GSH_E_TIMEOUT
    ss.reset_history()
    x, it, st = ss.solve(b, rtol=1e-12, maxiter=3000)
    xs, its, _, _, _ = O.pcg_ssor(ip, ix, val, b, rtol=1e-12, maxiter=3000, rows_per_thread=2, nthreads=16)
    assert st == 0 and it == its, (it, its)
    assert_bitwise(x, xs, "1M SSOR solution after a timed-out solve")
    ss.close()


def test_full_size_1m_channel(ctx):
    """The 1024 x 1024 log-normal channel (1.05M nodes, 31 levels): every backtrack (all sides x
    levels, 10 substeps) against the oracle; RHS, assembly and the whole FP64 solve bit for bit."""
    spec = cases.spec_gpu_1m()
    data = cases.make_case(spec)
    print(f"[gpu] 1M case {spec.name}: building", flush=True)
    mesh, p, (eta, uvw, su2) = cases.load_mesh(ctx, data)
    om = oracle_mesh_for(mesh, data.sigma)
    om.set(L.F_ETA2, eta)
    om.set(L.F_NODE_UVW, uvw)
    om.set(L.F_SIDE_U2, su2)
    om.levels(p)
    print("[gpu] 1M nodal velocity", flush=True)
    # element w from the nodal field's levels (any smooth values do); the reconstruction from the
    # side velocities then replaces the synthetic nodal field, on both sides
    en = mesh.topology(L.T_ELNODE)
    ew = uvw[en, :, 0].mean(axis=1) * 1e-3
    mesh.set(L.F_ELEM_W, ew)
    om.set(L.F_ELEM_W, ew)
    del uvw, en, ew
    mesh.nodal_velocity(p)
    om.nodal_velocity(p, nthreads=16)
    assert_bitwise(mesh.get(L.F_NODE_UVW), om.get(L.F_NODE_UVW), "1M nodal velocities")
    print("[gpu] 1M backtrack", flush=True)
    mesh.backtrack(p)
    print(f"[gpu] 1M backtrack on the oracle: {mesh.ns} sides x {mesh.nvrt} levels", flush=True)
    om.backtrack(p, nthreads=16)
    foot = mesh.get(L.F_FOOT)
    assert_bitwise(foot, om.get(L.F_FOOT), "1M foot elements (every side and level)")
    assert (foot >= 0).mean() > 0.9
    del foot
    assert_bitwise(mesh.get(L.F_BTU), om.get(L.F_BTU), "1M backtracked velocity")
    assert_bitwise(mesh.get(L.F_BSTATUS), om.get(L.F_BSTATUS), "1M backtracking status")
    print("[gpu] 1M backtrack with the LDS-staged walk", flush=True)
    mesh.set_backtrack_staging(True)
    mesh.backtrack(p)
    mesh.set_backtrack_staging(False)
    assert_bitwise(mesh.get(L.F_FOOT), om.get(L.F_FOOT), "1M foot elements, staged walk")
    assert_bitwise(mesh.get(L.F_BTU), om.get(L.F_BTU), "1M backtracked velocity, staged walk")
    assert_bitwise(mesh.get(L.F_BSTATUS), om.get(L.F_BSTATUS), "1M backtracking status, staged walk")
    print("[gpu] 1M rhs + assembly", flush=True)
    mesh.rhs_producers(p)
    om.rhs(p)
    solver = core.Solver(mesh)
    solver.assemble(p)
    val, b = om.assemble(p, nthreads=16)
    assert_bitwise(solver.get_values(), val, "1M CSR values")
    assert_bitwise(solver.get_rhs(), b, "1M RHS")
    print("[gpu] 1M solve", flush=True)
    x, it, st = solver.solve(None, rtol=1e-12, maxiter=3000)
    xo, ito, _, _ = O.pcg(mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES), val, b, rtol=1e-12,
                          maxiter=3000, nthreads=16)
    assert st == 0 and it == ito, (it, ito, st)
    assert_bitwise(x, xo, "1M PCG solution")
    assert solver.check() < 1e-10
    _timeout_then_same_result(ctx, mesh, solver, val, b, xo, ito)
    print("[gpu] 1M momentum", flush=True)
    solver.update_elevation()
    om.set(L.F_ETA1, eta)
    om.set(L.F_ETA2, xo)
    mesh.momentum(p)
    om.momentum(p)
    assert_bitwise(mesh.get(L.F_SIDE_U2), om.get(L.F_SIDE_U2), "1M side velocity from the momentum step")
    mesh.vertical_velocity(p)
    om.vertical_velocity()
    assert_bitwise(mesh.get(L.F_ELEM_W), om.get(L.F_ELEM_W), "1M element vertical velocity")
    solver.close()
    mesh.close()


@pytest.mark.parametrize("nvrt", [2, 3, 200, 1024])
def test_momentum_many_levels(ctx, nvrt):
    """Two levels (every side 2D) up to deep columns (a workgroup holds fewer sides, down to one at
    1024 levels), random btu, node-varying bottom levels and a new elevation with a dry corner:
    side velocities and element w bit for bit."""
    spec = cases.CaseSpec(f"deep-{nvrt}", 14, 9, 500.0, nvrt=nvrt, depth=10.0, substeps=4)
    mesh, om, p = _setup(ctx, spec, kb_fn=lambda d: (np.arange(d.size) % min(5, nvrt - 1)).astype(np.int32))
    rng = np.random.default_rng(nvrt)
    btu = rng.normal(0.0, 0.4, (mesh.ns, nvrt, 2))
    eta_old = mesh.get(L.F_ETA2)
    xy = mesh.topology(L.T_XY)
    eta_new = eta_old + 0.02 * np.sin(xy[:, 0] / 800.0)
    eta_new[np.argmin(xy[:, 0] + xy[:, 1])] = -11.0  # drier than the depth: its elements leave gradEta2
    print(f"[gpu] momentum nvrt={nvrt} ns={mesh.ns}", flush=True)
    for m in (mesh, om):
        m.set(L.F_BTU, btu)
    mesh.rhs_producers(p)
    om.rhs(p)
    for m in (mesh, om):
        m.set(L.F_ETA1, eta_old)
        m.set(L.F_ETA2, eta_new)
    mesh.momentum(p)
    om.momentum(p)
    assert_bitwise(mesh.get(L.F_SIDE_U2), om.get(L.F_SIDE_U2), f"side velocity, {nvrt} levels")
    mesh.vertical_velocity(p)
    om.vertical_velocity()
    assert_bitwise(mesh.get(L.F_ELEM_W), om.get(L.F_ELEM_W), f"element w, {nvrt} levels")
    mesh.close()


def test_objects_outlive_their_context():
    """A context closed while a mesh and a solver still use it stays alive until the last of
    them is closed (they keep working in between)."""
    spec = cases.CaseSpec("channel-24x16", 24, 16, 500.0, nvrt=4, depth=10.0, substeps=4)
    c2 = core.Context(0)
    mesh, p, _ = cases.load_mesh(c2, cases.make_case(spec))
    s = core.Solver(mesh)
    print("[gpu] closing the context before its mesh and solver", flush=True)
    c2.close()
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    s.assemble(p)
    x, it, st = s.solve(None, rtol=1e-12, maxiter=500, check_every=10)
    assert st == 0 and it > 0 and np.isfinite(x).all()
    s.close()
    mesh.close()


def test_elevation_bc_types_and_tides(ctx, tmp_path):
    """A bctides.in with an elevation-plus-tides boundary (type 5, two constituents) applied through
    the reader, then per-node types mixed over the open nodes (0, 2, 3, 5) at model time t: elbc
    within 1e-14 of the oracle where cosines enter (the device and host cosines may differ in the
    last bit), bit for bit elsewhere; then assembly with per-node essential rows and the solve bit
    for bit from the same elbc."""
    from gschism_amd.bctides import apply_bctides, elevation_forcing, read_bctides

    spec = cases.CaseSpec("tides-60x30", 60, 30, 500.0, nvrt=5, depth=None, substeps=4)
    data = cases.make_case(spec)
    mesh, om, p = _setup(ctx, spec)
    seg_file = np.flatnonzero(data.open_bnd)
    seg_file = seg_file[np.argsort(data.xy[seg_file, 1])]
    k = seg_file.size
    lines = ["01/01/2020 00:00:00", "0 40.", "2", "M2", "1.405189D-04 0.97 90.0", "S2", "1.454441D-04 1.0 0.0",
             "1", f"{k} 5 0 0 0", "M2"]
    lines += [f"{0.4 + 0.01 * j:.3f} {10.0 + j:.1f}" for j in range(k)] + ["S2"]
    lines += [f"{0.15 + 0.005 * j:.4f} {40.0 + 2 * j:.1f}" for j in range(k)]
    path = tmp_path / "bctides.in"
    path.write_text("\n".join(lines) + "\n")
    tide = read_bctides(str(path), 1, k)
    apply_bctides(mesh, tide, [seg_file])
    inv = np.empty_like(mesh.node_perm)
    inv[mesh.node_perm] = np.arange(mesh.n)
    seg = inv[seg_file]
    etype, eth, emo, efa = elevation_forcing(tide, [seg], mesh.n)
    assert np.array_equal(mesh.get(L.F_ETYPE), etype)
    # mix the types along the boundary; a constant elevation for the types that use one
    etype[seg] = np.array([0, 2, 3, 5], dtype=np.int8)[np.arange(k) % 4]
    eth[seg] = 0.05 + 0.01 * np.arange(k)
    for m in (mesh, om):
        m.set(L.F_ETYPE, etype)
        m.set(L.F_ETH, eth)
    om.set_tides(tide.amig, tide.ff, tide.face, emo, efa)
    p.time, p.ramp = 2.5e5, 0.9
    print(f"[gpu] elevation BC types, {k} open nodes", flush=True)
    mesh.backtrack(p)
    om.backtrack(p, nthreads=8)
    mesh.rhs_producers(p)
    om.rhs(p)
    got, ref = mesh.get(L.F_ELBC), om.get(L.F_ELBC)
    harm = np.isin(np.arange(mesh.n), seg[np.isin(etype[seg], (3, 5))])
    assert np.abs(got[harm] - ref[harm]).max() <= 1e-14
    assert_bitwise(got[~harm], ref[~harm], "elbc of the non-tidal nodes")
    assert np.abs(ref[harm]).max() > 0.1
    om.set(L.F_ELBC, got)
    solver = core.Solver(mesh)
    solver.assemble(p)
    val, b = om.assemble(p)
    assert_bitwise(solver.get_values(), val, "CSR values with per-node essential rows")
    assert_bitwise(solver.get_rhs(), b, "RHS with per-node essential rows")
    x, it, st = solver.solve(None, rtol=1e-12, maxiter=2000)
    xo, ito, _, _ = O.pcg(mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES), val, b, rtol=1e-12,
                          maxiter=2000)
    assert st == 0 and it == ito
    assert_bitwise(x, xo, "solution with per-node elevation BCs")
    ess = seg[etype[seg] != 0]
    assert np.abs(x[ess] - got[ess]).max() < 1e-10  # identity rows: x = elbc up to the solve tolerance


def test_momentum_branches(ctx):
    """Momentum branches the default case does not reach: no vertical viscosity (the pure mass
    matrix), sides without drag (u2[kb] = u2[kb+1]) next to sides with drag, clamping active, dry
    sides of an island: side velocities and element w bit for bit."""
    spec = cases.CaseSpec("island-48x32", 48, 32, 400.0, nvrt=9, depth=10.0, substeps=4)
    mesh, om, p = _setup(ctx, spec, kb_fn=_lsc_kb, depth_fn=_island_depth)
    rng = np.random.default_rng(7)
    cd = np.where(np.arange(mesh.n) % 2 == 0, 0.0, 3e-3)
    btu = rng.normal(0.0, 0.4, (mesh.ns, mesh.nvrt, 2))
    eta_old = mesh.get(L.F_ETA2)
    eta_new = eta_old + 0.05 * rng.normal(size=mesh.n)
    p.dfv0, p.rmaxvel = 0.0, 0.3
    for m in (mesh, om):
        m.set(L.F_CD, cd)
        m.set(L.F_BTU, btu)
    mesh.rhs_producers(p)
    om.rhs(p)
    for m in (mesh, om):
        m.set(L.F_ETA1, eta_old)
        m.set(L.F_ETA2, eta_new)
    print(f"[gpu] momentum branches ns={mesh.ns}", flush=True)
    mesh.momentum(p)
    om.momentum(p)
    u2 = mesh.get(L.F_SIDE_U2)
    assert_bitwise(u2, om.get(L.F_SIDE_U2), "side velocity, no viscosity, mixed drag, clamping")
    assert np.abs(u2[..., 0]).max() == 0.3
    mesh.vertical_velocity(p)
    om.vertical_velocity()
    assert_bitwise(mesh.get(L.F_ELEM_W), om.get(L.F_ELEM_W), "element w")
    mesh.close()


def test_step_call_argument_errors(ctx):
    """Bad arguments of the step calls are errors (GSH_E_ARG) raised before any launch, and the
    mesh keeps working afterwards."""
    spec = cases.CaseSpec("args-24x16", 24, 16, 500.0, nvrt=4, depth=10.0, substeps=4)
    mesh, om, p = _setup(ctx, spec)
    n = mesh.n
    bad = [dict(ishapiro=2), dict(dfv0=-1.0), dict(rmaxvel=-1.0), dict(theta=1.5)]
    for kw in bad:
        q = L.default_step_params(**kw)
        with pytest.raises(L.GshError) as ei:
            mesh.momentum(q)
        assert ei.value.code == L.GSH_E_ARG, kw
    with pytest.raises(L.GshError) as ei:
        mesh.set_tides(np.ones(300), np.ones(300), np.zeros(300), np.zeros((n, 300)), np.zeros((n, 300)))
    assert ei.value.code == L.GSH_E_ARG
    with pytest.raises(L.GshError):
        core.Server(ctx, precond=7)
    for kw in (dict(iettype=6), dict(iettype=-1), dict(shapiro0=0.7)):
        with pytest.raises(L.GshError) as ei:
            mesh.rhs_producers(L.default_step_params(**kw))
        assert ei.value.code == L.GSH_E_ARG, kw
    bad_types = np.full(n, -1, dtype=np.int8)
    bad_types[0] = 6
    with pytest.raises(L.GshError) as ei:
        mesh.set(L.F_ETYPE, bad_types)
    assert ei.value.code == L.GSH_E_ARG
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    mesh.momentum(p)
    assert np.isfinite(mesh.get(L.F_SIDE_U2)).all()
    mesh.close()
