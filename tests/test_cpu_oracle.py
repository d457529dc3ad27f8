"""Pins of the CPU oracle: closed forms, invariants and brute force (the reference ships no
tests or golden vectors, so these are what fix the oracle's behaviour)."""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from gschism_amd import _lib as L
from gschism_amd import cases
from oracle import pyoracle as O


def _mesh(nx=24, ny=16, dx=400.0, depth=10.0, nvrt=6, jitter=0.2, kb=None):
    xy, en, ob = cases.channel(nx, ny, dx, jitter)
    d = np.full(xy.shape[0], depth) if np.isscalar(depth) else depth
    return O.OracleMesh(xy, en, d, ob, cases.sigma_levels(nvrt), kb=kb), xy, en, ob


def _params(**kw):
    return L.default_step_params(**kw)


# ---------------------------------------------------------------- geometry
def test_geometry_closed_forms():
    m, xy, en, _ = _mesh()
    area = m.geometry(0)
    p = xy[en]
    ref = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    assert np.allclose(area, ref, rtol=1e-14)
    g = m.geometry(1)
    assert np.abs(g.sum(axis=1)).max() < 1e-15 / 400.0 * 10          # sum of shape gradients is 0
    for i in range(3):                                                # grad phi_i . (x_j - x_i) = delta_ij - 1
        for j in range(3):
            dot = (g[:, i, :] * (p[:, j, :] - p[:, i, :])).sum(1)
            assert np.allclose(dot, 0.0 if i == j else -1.0, atol=1e-12)
    sn = m.topology(L.T_ISIDENODE)
    c = m.geometry(2)
    assert np.allclose(c, 0.5 * (xy[sn[:, 0]] + xy[sn[:, 1]]), rtol=0, atol=1e-9)
    ln = m.geometry(3)
    assert np.allclose(ln, np.linalg.norm(xy[sn[:, 1]] - xy[sn[:, 0]], axis=1), rtol=1e-14)
    nrm = m.geometry(4)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, rtol=1e-14)
    assert np.abs((nrm * (xy[sn[:, 1]] - xy[sn[:, 0]])).sum(1)).max() < 1e-9


# ---------------------------------------------------------------- levels
def test_sigma_levels_closed_form():
    m, xy, en, ob = _mesh(depth=12.0, nvrt=7)
    eta = 0.3 * np.sin(xy[:, 0] / 2000.0)
    m.set(L.F_ETA2, eta)
    m.levels(_params())
    zc = m.get(L.F_ZC)
    sn = m.topology(L.T_ISIDENODE)
    sig = cases.sigma_levels(7)
    zn = lambda i, k: (-12.0 if k == 0 else (eta[i] if k == 6 else (eta[i] + 12.0) * sig[k] + eta[i]))
    for s in range(0, m.ns, 37):
        for k in range(7):
            assert zc[s, k] == 0.5 * (zn(sn[s, 0], k) + zn(sn[s, 1], k))
    # z = eta + (eta + h) sigma (ivcor=1, meshdata.cu:1798-1803) up to rounding
    assert np.allclose(zc[:, 3], 0.5 * ((eta[sn[:, 0]] + 12) * sig[3] + eta[sn[:, 0]] + (eta[sn[:, 1]] + 12) * sig[3] + eta[sn[:, 1]]))


def test_dry_elements_and_kb():
    xy, en, ob = cases.channel(20, 14, 300.0)
    depth = np.where(np.hypot(xy[:, 0] - 3000, xy[:, 1] - 2000) < 900, -2.0, 8.0)
    m = O.OracleMesh(xy, en, depth, ob, cases.sigma_levels(5))
    m.set(L.F_ETA2, np.zeros(xy.shape[0]))
    m.levels(_params())
    wet = m.get(L.F_WET)
    n, ne = xy.shape[0], en.shape[0]
    node_dry_geom = depth + 0.0 <= 0.01
    elem_dry_ref = node_dry_geom[en].any(axis=1)
    assert np.array_equal(wet[n:n + ne] == 0, elem_dry_ref)


# ---------------------------------------------------------------- nodal velocities
def _nodal_brute(m, su2, ew, rmaxvel, wet_elem, node_dry, kb):
    """Nodal velocities element by element, the opposite side found from the side's end nodes."""
    en, es, sn = m.topology(L.T_ELNODE), m.topology(L.T_ELSIDE), m.topology(L.T_ISIDENODE)
    area = m.geometry(0)
    n, nv = m.n, m.nvrt
    usum = np.zeros((n, nv, 2))
    cnt = np.zeros(n)
    wsum = np.zeros((n, nv))
    asum = np.zeros(n)
    lim = np.array([rmaxvel, rmaxvel * 1.013])
    for e in range(m.ne):
        for i in en[e]:
            asum[i] += area[e]
            wsum[i] += ew[e] * area[e]
            if not wet_elem[e] or node_dry[i]:
                continue
            opp = [s for s in es[e] if i not in sn[s]]
            adj = [s for s in es[e] if i in sn[s]]
            assert len(opp) == 1 and len(adj) == 2
            usum[i] += np.clip(su2[adj[0]] + su2[adj[1]] - su2[opp[0]], -lim, lim)
            cnt[i] += 1
    out = np.zeros((n, nv, 3))
    for i in range(n):
        k0 = kb[i]
        out[i, k0:, 2] = wsum[i, k0:] / asum[i]
        if cnt[i] > 0:
            out[i, k0:, :2] = usum[i, k0:] / cnt[i]
    return out


def test_nodal_velocity_reproduces_linear_fields():
    """A triangle's side-centre values extrapolated to a node as u_s2 + u_s3 - u_s1 are exact for
    linear fields (m2 + m3 - m1 is the node), and equal element values of w average to w."""
    m, xy, en, _ = _mesh(nvrt=5)
    n, ne, nv = xy.shape[0], en.shape[0], 5
    p = _params(rmaxvel=100.0)
    m.set(L.F_ETA2, np.zeros(n))
    m.levels(p)
    sc = m.geometry(2)
    lev = np.arange(nv)
    u = 0.3 + 1e-4 * sc[:, :1] - 2e-4 * sc[:, 1:] + 0.01 * lev
    v = -0.2 + 5e-5 * sc[:, :1] + 1e-4 * sc[:, 1:] - 0.02 * lev
    m.set(L.F_SIDE_U2, np.stack([u, v], axis=-1))
    m.set(L.F_ELEM_W, np.tile(1e-3 * (lev + 1), (ne, 1)))
    m.nodal_velocity(p)
    got = m.get(L.F_NODE_UVW)
    xyo = m.topology(L.T_XY)
    eu = 0.3 + 1e-4 * xyo[:, :1] - 2e-4 * xyo[:, 1:] + 0.01 * lev
    ev = -0.2 + 5e-5 * xyo[:, :1] + 1e-4 * xyo[:, 1:] - 0.02 * lev
    assert np.allclose(got[..., 0], eu, rtol=0, atol=1e-12)
    assert np.allclose(got[..., 1], ev, rtol=0, atol=1e-12)
    assert np.allclose(got[..., 2], 1e-3 * (lev + 1), rtol=1e-14, atol=0)


def test_nodal_velocity_matches_element_by_element_sums():
    """Random side velocities and element w on a mesh with dry elements and deep bottom levels:
    the oracle's ball-order sums against element-by-element sums (clamping active)."""
    xy, en, ob = cases.channel(20, 14, 300.0)
    depth = np.where(np.hypot(xy[:, 0] - 3000, xy[:, 1] - 2000) < 900, -2.0, 8.0)
    n, nv = xy.shape[0], 6
    kb = (np.arange(n) % 3).astype(np.int32)
    m = O.OracleMesh(xy, en, depth, ob, cases.sigma_levels(nv), kb=kb)
    p = _params(rmaxvel=1.5)
    m.set(L.F_ETA2, np.zeros(n))
    m.levels(p)
    rng = np.random.default_rng(5)
    su2 = rng.normal(0.0, 1.0, (m.ns, nv, 2))
    ew = rng.normal(0.0, 1e-3, (m.ne, nv))
    m.set(L.F_SIDE_U2, su2)
    m.set(L.F_ELEM_W, ew)
    m.nodal_velocity(p, nthreads=2)
    got = m.get(L.F_NODE_UVW)
    wet = m.get(L.F_WET)
    ref = _nodal_brute(m, su2, ew, 1.5, wet[n:n + m.ne] != 0, wet[:n] == 0, m.topology(L.T_KB))
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-15)
    dry = wet[:n] == 0
    assert dry.any() and np.all(got[dry, :, :2] == 0.0)
    for i in range(n):
        assert np.all(got[i, :kb[i]] == 0.0)
    assert np.abs(got[..., 0]).max() <= 1.5 and np.abs(got[..., 1]).max() <= 1.5 * 1.013
    assert np.abs(got[..., 0]).max() == 1.5  # some node averages only clamped values


# ---------------------------------------------------------------- interpolation
def test_interpolation_reproduces_linear_fields():
    m, xy, en, _ = _mesh(depth=10.0, nvrt=6)
    nv = 6
    m.set(L.F_ETA2, np.zeros(xy.shape[0]))
    m.levels(_params())
    z = cases.sigma_levels(nv) * 10.0                       # flat levels: z_k = -10 + 10 k/5
    a, bx, by, bz = 0.3, 2e-5, -3e-5, 0.01
    uvw = np.empty((xy.shape[0], nv, 3))
    for k in range(nv):
        f = a + bx * xy[:, 0] + by * xy[:, 1] + bz * z[k]
        uvw[:, k, 0] = f
        uvw[:, k, 1] = 2 * f
        uvw[:, k, 2] = -f
    m.set(L.F_NODE_UVW, uvw)
    rng = np.random.default_rng(3)
    for _ in range(200):
        e = int(rng.integers(en.shape[0]))
        w = rng.dirichlet([1, 1, 1])
        px, py = w @ xy[en[e]]
        pz = rng.uniform(-9.9, -0.1)
        out = m.interp(e, px, py, pz)
        f = a + bx * px + by * py + bz * pz
        assert np.allclose(out, [f, 2 * f, -f], rtol=1e-12, atol=1e-14)
    # points below the bed / above the surface clamp to the end levels
    e = 5
    c = xy[en[e]].mean(0)
    out = m.interp(e, c[0], c[1], -50.0)
    assert np.isclose(out[0], a + bx * c[0] + by * c[1] + bz * (-10.0))


# ---------------------------------------------------------------- backtracking
def _uniform_flow(m, xy, nv, u, v):
    uvw = np.zeros((xy.shape[0], nv, 3))
    uvw[:, :, 0], uvw[:, :, 1] = u, v
    su2 = np.zeros((m.ns, nv, 2))
    su2[:, :, 0], su2[:, :, 1] = u, v
    m.set(L.F_NODE_UVW, uvw)
    m.set(L.F_SIDE_U2, su2)


def _contains(xy, tri, p, tol=1e-9):
    a, b, c = xy[tri]
    T = np.array([[b[0] - a[0], c[0] - a[0]], [b[1] - a[1], c[1] - a[1]]])
    l1, l2 = np.linalg.solve(T, p - a)
    return min(l1, l2, 1 - l1 - l2) >= -tol


def test_backtrack_uniform_flow_lands_where_it_should():
    nv = 4
    m, xy, en, _ = _mesh(nx=30, ny=20, dx=400.0, nvrt=nv)
    m.set(L.F_ETA2, np.zeros(xy.shape[0]))
    p = _params(dt=150.0, dtb_min=30.0, dtb_max=30.0)
    m.levels(p)
    u, v = 0.6, -0.35
    _uniform_flow(m, xy, nv, u, v)
    m.backtrack(p)
    foot, btu, st = m.get(L.F_FOOT), m.get(L.F_BTU), m.get(L.F_BSTATUS)
    c = m.geometry(2)
    lx, ly = xy[:, 0].max(), xy[:, 1].max()
    checked = 0
    for s in range(m.ns):
        target = c[s] - 150.0 * np.array([u, v])
        if not (500 < target[0] < lx - 500 and 500 < target[1] < ly - 500):
            continue
        for k in range(nv):
            assert st[s, k] == 0
            assert _contains(xy, en[foot[s, k]], target), (s, k)
            assert np.allclose(btu[s, k], [u, v], rtol=1e-13)
            checked += 1
    assert checked > 1000


def test_backtrack_boundaries_exit_and_slide():
    nv = 3
    m, xy, en, ob = _mesh(nx=20, ny=12, dx=400.0, nvrt=nv)
    m.set(L.F_ETA2, np.zeros(xy.shape[0]))
    p = _params(dt=300.0, dtb_min=60.0, dtb_max=60.0)
    m.levels(p)
    # flow towards +x: the track goes west and leaves through the open west boundary
    _uniform_flow(m, xy, nv, 2.0, 0.0)
    m.backtrack(p)
    st = m.get(L.F_BSTATUS)
    c = m.geometry(2)
    near_west = c[:, 0] < 500.0
    assert (st[near_west, 0] & L.BT_OPEN_EXIT).all()
    # flow towards +y with a small x component: tracks hit the south land boundary and slide
    _uniform_flow(m, xy, nv, 0.1, 2.0)
    m.backtrack(p)
    st = m.get(L.F_BSTATUS)
    btu, foot = m.get(L.F_BTU), m.get(L.F_FOOT)
    near_south = (c[:, 1] < 300.0) & (c[:, 0] > 1500.0)
    assert (st[near_south, 0] & L.BT_SLIDE).all()
    # the slid track ends on the south edge; btu is the velocity interpolated there
    assert (xy[en[foot[near_south, 0]], 1].min(axis=1) == xy[:, 1].min()).all()
    assert np.allclose(btu[near_south, 0], [0.1, 2.0], rtol=1e-13)


def test_backtrack_slow_flow_is_not_tracked():
    nv = 3
    m, xy, en, _ = _mesh(nvrt=nv)
    m.set(L.F_ETA2, np.zeros(xy.shape[0]))
    p = _params()
    m.levels(p)
    _uniform_flow(m, xy, nv, 5e-5, 5e-5)   # below velmin_btrack = 1e-4
    m.backtrack(p)
    assert (m.get(L.F_FOOT) == -1).all()
    assert np.array_equal(m.get(L.F_BTU), np.full((m.ns, nv, 2), 5e-5))


# ---------------------------------------------------------------- RHS producers
def test_rhs_uniform_closed_forms():
    nv = 5
    m, xy, en, _ = _mesh(nvrt=nv, depth=8.0)
    eta = np.full(xy.shape[0], 0.25)
    m.set(L.F_ETA2, eta)
    p = _params(dt=100.0)
    m.levels(p)
    u0 = np.array([0.4, -0.2])
    su2 = np.tile(u0, (m.ns, nv, 1))
    m.set(L.F_SIDE_U2, su2)
    m.set(L.F_BTU, su2)
    m.rhs(p)
    H = 8.25
    U = m.get(L.F_SIDE_U)
    assert np.allclose(U, H * u0, rtol=1e-13)
    chi = m.get(L.F_SIDE_CHI)
    assert np.allclose(chi, 2.5e-3 * np.linalg.norm(u0), rtol=1e-14)
    hh = m.get(L.F_SIDE_HHAT)
    assert np.allclose(hh, H - chi * 100.0, rtol=1e-14)
    g = m.get(L.F_GHAT1)
    assert np.allclose(g, (H - chi[0] * 100.0) * u0, rtol=1e-12)


# ---------------------------------------------------------------- assembly
def _assemble_dense(xy, en, grav, theta, dt, hh_elem):
    n = xy.shape[0]
    A = np.zeros((n, n))
    for e, tri in enumerate(en):
        p = xy[tri]
        area = 0.5 * ((p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1]))
        M = area / 12.0 * (np.ones((3, 3)) + np.eye(3))
        gr = np.array([[p[1, 1] - p[2, 1], p[2, 0] - p[1, 0]], [p[2, 1] - p[0, 1], p[0, 0] - p[2, 0]],
                       [p[0, 1] - p[1, 1], p[1, 0] - p[0, 0]]]) / (2 * area)
        K = area * gr @ gr.T
        A[np.ix_(tri, tri)] += M + grav * theta * theta * dt * dt * hh_elem[e] * K
    return A


def test_assembly_matches_dense_element_assembly():
    nv = 3
    m, xy, en, ob = _mesh(nx=14, ny=9, nvrt=nv, depth=6.0)
    m.set(L.F_ETA2, np.zeros(xy.shape[0]))
    p = _params(dt=120.0, theta=0.55, iettype=0)
    m.levels(p)
    hh = np.linspace(1.0, 6.0, m.ns)
    m.set(L.F_SIDE_HHAT, hh)
    m.set(L.F_SIDE_U, np.zeros((m.ns, 2)))
    m.set(L.F_GHAT1, np.zeros((m.ne, 2)))
    val, b = m.assemble(p)
    ip, ix = m.topology(L.T_CSR_INDPTR), m.topology(L.T_CSR_INDICES)
    A = sp.csr_matrix((val, ix, ip), shape=(m.n, m.n)).toarray()
    es = m.topology(L.T_ELSIDE)
    hh_elem = hh[es].mean(axis=1)
    Ad = _assemble_dense(xy, en, 9.81, 0.55, 120.0, hh_elem)
    assert np.allclose(A, Ad, rtol=1e-12, atol=1e-9 * np.abs(Ad).max())
    # mass row sums: sum_j M_ij = (area of the ball) / 3; stiffness rows sum to zero
    area = m.geometry(0)
    ball = np.zeros(m.n)
    np.add.at(ball, en.reshape(-1), np.repeat(area / 3.0, 3))
    assert np.allclose(A.sum(1), ball, rtol=1e-9)
    # eta = 1, U = G = 0: qel = M . 1 minus the friction term, which vanishes for constant eta
    m.set(L.F_ETA2, np.ones(m.n))
    m.levels(p)
    _, b1 = m.assemble(p)
    assert np.allclose(b1, ball, rtol=1e-10)


def test_essential_bc_rows_and_columns():
    nv = 3
    m, xy, en, ob = _mesh(nx=12, ny=8, nvrt=nv)
    m.set(L.F_ETA2, np.zeros(xy.shape[0]))
    p0 = _params(iettype=0)
    m.levels(p0)
    m.set(L.F_SIDE_HHAT, np.full(m.ns, 5.0))
    eth = np.where(ob == 1, 0.37, 0.0)
    m.set(L.F_ETH, eth)
    m.rhs(p0)
    m.set(L.F_SIDE_HHAT, np.full(m.ns, 5.0))
    val0, b0 = m.assemble(p0)
    p2 = _params(iettype=2)
    m.rhs(p2)
    m.set(L.F_SIDE_HHAT, np.full(m.ns, 5.0))
    val2, b2 = m.assemble(p2)
    ip, ix = m.topology(L.T_CSR_INDPTR), m.topology(L.T_CSR_INDICES)
    A0 = sp.csr_matrix((val0, ix, ip), shape=(m.n, m.n)).toarray()
    A2 = sp.csr_matrix((val2, ix, ip), shape=(m.n, m.n)).toarray()
    bc = ob == 1
    assert np.array_equal(A2[bc], np.eye(m.n)[bc])
    assert np.array_equal(b2[bc], eth[bc])
    assert (A2[np.ix_(~bc, bc)] == 0).all()
    assert np.allclose(A2[np.ix_(~bc, ~bc)], A0[np.ix_(~bc, ~bc)])
    # column elimination moves a_ij * elbc to the right-hand side
    assert np.allclose(b2[~bc], b0[~bc] - A0[np.ix_(~bc, bc)] @ eth[bc], rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- PCG
def _spd(n, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=6.0 / n, random_state=seed, format="csr")
    A = A + A.T
    A = A + sp.diags(np.abs(A).sum(1).A1 + 1.0)
    A = A.tocsr()
    A.sort_indices()
    return A, rng.normal(size=n)


@pytest.mark.parametrize("rpt", [1, 4])
def test_pcg_matches_direct_solve(rpt):
    A, b = _spd(3000, 4)
    x, it, rr, atol2 = O.pcg(A.indptr, A.indices, A.data, b, rtol=1e-12, maxiter=2000, rows_per_thread=rpt)
    xs = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10
    assert it % 10 == 0 and rr <= atol2
    assert np.isclose(atol2, (1e-12 * np.linalg.norm(b)) ** 2, rtol=1e-12)


def test_pcg_stopping_rule_and_special_cases():
    A, b = _spd(500, 5)
    x, it, _, _ = O.pcg(A.indptr, A.indices, A.data, np.zeros(500), rtol=1e-12, maxiter=100)
    assert it == 0 and not x.any()                     # atol == 0 (pcg_solver.py:1089-1095)
    x, it, _, _ = O.pcg(A.indptr, A.indices, A.data, b, rtol=0.0, maxiter=100)
    assert it == 0 and not x.any()                     # rtol = 0 also gives atol == 0
    x, it, _, _ = O.pcg(A.indptr, A.indices, A.data, b, rtol=1e-30, maxiter=17, check_every=10)
    assert it == 17                                    # maxiter returned silently
    x3, it3, _, _ = O.pcg(A.indptr, A.indices, A.data, b, rtol=1e-8, maxiter=200, check_every=3)
    assert it3 % 3 == 0
    # the extrapolated initial guess 3 x0 - 3 x1 + x2 of the exact solution converges at once
    xs = sla.spsolve(A.tocsc(), b)
    x4, it4, _, _ = O.pcg(A.indptr, A.indices, A.data, b, hist=(xs, xs, xs), rtol=1e-6, maxiter=50, check_every=1)
    assert it4 <= 1


def test_canonical_dot_is_accurate():
    rng = np.random.default_rng(9)
    for n in (1, 255, 256, 1025, 100_000):
        a, b = rng.normal(size=n), rng.normal(size=n)
        d = O.dot_canon(a, b)
        exact = math.fsum(a * b)
        bound = n * 2.3e-16 * np.abs(a * b).sum()
        assert abs(d - exact) <= bound + 1e-300


# ---------------------------------------------------------------- multicolour SSOR
def _smallest_last_colouring(indptr, indices):
    """Independent restatement: remove the vertex of least (degree, index) repeatedly, then
    colour first-fit in reverse removal order."""
    import heapq

    n = indptr.size - 1
    nb = [[j for j in indices[indptr[i]:indptr[i + 1]] if j != i] for i in range(n)]
    deg = [len(x) for x in nb]
    heap = [(deg[i], i) for i in range(n)]
    heapq.heapify(heap)
    gone = [False] * n
    order = []
    while heap:
        d, i = heapq.heappop(heap)
        if gone[i] or d != deg[i]:
            continue
        gone[i] = True
        order.append(i)
        for j in nb[i]:
            if not gone[j]:
                deg[j] -= 1
                heapq.heappush(heap, (deg[j], j))
    col = [-1] * n
    for i in reversed(order):
        used = {col[j] for j in nb[i]}
        c = 0
        while c in used:
            c += 1
        col[i] = c
    return np.array(col, dtype=np.int32)


def _ssor_pcg_reference(A, b, omega, rtol, maxiter, check_every, colour):
    """Textbook PCG with z = M^-1 r, M = (D/w + L)(D/w)^-1(D/w + U) in colour-major order."""
    perm = np.lexsort((np.arange(A.shape[0]), colour))
    Ap = A[perm][:, perm].tocsr()
    bp = b[perm]
    d = Ap.diagonal() / omega
    DL = (sp.diags(d) + sp.tril(Ap, -1)).tocsr()
    DU = (sp.diags(d) + sp.triu(Ap, 1)).tocsr()

    def minv(r):
        y = sla.spsolve_triangular(DL, r, lower=True)
        return sla.spsolve_triangular(DU, d * y, lower=False)

    x = np.zeros_like(bp)
    r = bp.copy()
    z = minv(r)
    p = np.zeros_like(bp)
    rho, rho_old = r @ z, 0.0
    atol2 = (rtol * np.linalg.norm(bp)) ** 2
    for k in range(maxiter):
        beta = 0.0 if (k == 0 or rho_old == 0.0) else rho / rho_old
        p = z + beta * p
        q = Ap @ p
        alpha = rho / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = minv(r)
        rho_old, rho = rho, r @ z
        if (k + 1) % check_every == 0 or k == maxiter - 1:
            if r @ r <= atol2:
                break
    out = np.empty_like(x)
    out[perm] = x
    return out, k + 1


def _mesh_system(nx=30, ny=20):
    m, xy, en, ob = _mesh(nx=nx, ny=ny, nvrt=4, depth=20.0)
    m.set(L.F_ETA2, 0.1 * np.sin(xy[:, 0] / 3000.0))
    p = _params(dt=300.0, theta=0.6, iettype=2)
    m.levels(p)
    m.set(L.F_SIDE_HHAT, np.full(m.ns, 20.0))
    m.set(L.F_ELBC, np.zeros(m.n))
    val, b = m.assemble(p)
    return m.topology(L.T_CSR_INDPTR), m.topology(L.T_CSR_INDICES), val, b


def test_colouring_is_valid_and_matches_restatement():
    graphs = [_spd(400, 7)[0], _spd(1500, 8)[0]]
    ip, ix, _, _ = _mesh_system()
    graphs.append(sp.csr_matrix((np.ones(ix.size), ix, ip)))
    for A in graphs:
        col, nc = O.colour(A.indptr, A.indices)
        assert nc == col.max() + 1
        rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
        off = A.indices != rows
        assert not (col[rows[off]] == col[A.indices[off]]).any()
        assert np.array_equal(col, _smallest_last_colouring(A.indptr, A.indices))
    # a planar triangulation is 5-degenerate, so smallest-last needs at most 6 colours
    col, nc = O.colour(ip, ix)
    assert nc <= 6


@pytest.mark.parametrize("omega", [1.0, 1.4])
def test_ssor_pcg_matches_textbook_pcg(omega):
    ip, ix, val, b = _mesh_system()
    A = sp.csr_matrix((val, ix, ip))
    x, it, rr, atol2, nc = O.pcg_ssor(ip, ix, val, b, omega=omega, rtol=1e-10, maxiter=1000, check_every=1)
    col, _ = O.colour(ip, ix)
    xr, itr = _ssor_pcg_reference(A, b, omega, 1e-10, 1000, 1, col)
    assert it == itr
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9
    xs = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    # the stopping test is on the true residual b - A x
    assert rr <= atol2
    assert np.isclose(np.linalg.norm(b - A @ x) ** 2, rr, rtol=1e-3)
    # SSOR needs fewer iterations than Jacobi on the wave-equation system
    _, itj, _, _ = O.pcg(ip, ix, val, b, rtol=1e-10, maxiter=1000, check_every=1)
    assert it < itj


def test_ssor_pcg_special_cases():
    A, b = _spd(700, 12)
    x, it, _, _, _ = O.pcg_ssor(A.indptr, A.indices, A.data, np.zeros(700), rtol=1e-12)
    assert it == 0 and not x.any()
    xs = sla.spsolve(A.tocsc(), b)
    for rpt in (1, 4):
        x, it, _, _, _ = O.pcg_ssor(A.indptr, A.indices, A.data, b, rtol=1e-12, rows_per_thread=rpt)
        assert it % 10 == 0
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10
    # exact guess from the history: converged at the first check
    x, it, _, _, _ = O.pcg_ssor(A.indptr, A.indices, A.data, b, hist=(xs, xs, xs), rtol=1e-6, check_every=1)
    assert it <= 1
    # a pattern without a diagonal entry is rejected
    B = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 2.0]]))
    B.eliminate_zeros()
    with pytest.raises(ValueError):
        O.pcg_ssor(B.indptr, B.indices, B.data, np.ones(2))


def test_bottom_level_range_is_checked():
    """kb in [0, nvrt-2] (at least one layer), rejected otherwise like the library does."""
    xy, en, ob = cases.channel(6, 5, 100.0)
    kb = np.zeros(xy.shape[0], np.int32)
    kb[3] = 4
    with pytest.raises(Exception):
        O.OracleMesh(xy, en, np.full(xy.shape[0], 5.0), ob, cases.sigma_levels(5), kb=kb)
    from gschism_amd import core

    with pytest.raises(L.GshError):
        core.Mesh(None, xy, en, np.full(xy.shape[0], 5.0), cases.sigma_levels(5), open_bnd=ob, kb=kb)
    kb[3] = 3
    O.OracleMesh(xy, en, np.full(xy.shape[0], 5.0), ob, cases.sigma_levels(5), kb=kb)


# ---------------------------------------------------------------- momentum
def _momentum_state(nvrt, kb=None, depth=10.0, cd=0.0, btu=None, seed=0, eta_old=None, eta_new=None):
    """Oracle mesh after levels (old elevation), the RHS producers (chi, H^) and the elevation
    update (eta1 = old, eta2 = new), ready for the momentum step."""
    m, xy, en, ob = _mesh(nx=18, ny=12, nvrt=nvrt, kb=kb, depth=depth)
    n, nv = xy.shape[0], nvrt
    xyo = m.topology(L.T_XY)
    e_old = eta_old(xyo) if eta_old else np.zeros(n)
    e_new = eta_new(xyo) if eta_new else e_old
    p = _params(dfv0=0.05, rmaxvel=100.0, ishapiro=0)
    m.set(L.F_ETA2, e_old)
    m.set(L.F_CD, np.full(n, cd))
    m.set(L.F_GRAV2, np.full(n, 9.81))
    m.levels(p)
    if btu is None:
        rng = np.random.default_rng(seed)
        btu = rng.normal(0.0, 0.3, (m.ns, nv, 2))
    m.set(L.F_BTU, btu)
    m.set(L.F_SIDE_U2, np.zeros((m.ns, nv, 2)))
    m.rhs(p)
    m.set(L.F_ETA1, e_old)
    m.set(L.F_ETA2, e_new)
    return m, p, btu


def _side_info(m):
    n, ne = m.n, m.ne
    wet = m.get(L.F_WET)
    sn = m.topology(L.T_ISIDENODE)
    kbn = m.topology(L.T_KB)
    swet = wet[n + ne:] != 0
    skb = np.where(swet, np.minimum(kbn[sn[:, 0]], kbn[sn[:, 1]]), 0)
    return swet, skb, m.topology(L.T_SIDE_TYPE), m.geometry(4)


def test_momentum_uniform_flow_linear_elevation():
    """Cd = 0, uniform btu = u0 at every level, linear old and new elevations: every wet side,
    3D (any dfv0) or 2D (H^ = H), gets u0 - dt g (theta grad eta2 + (1 - theta) grad eta1) at
    and above its bottom level (u2[kb] = u2[kb+1] without drag), zero below; land sides keep the
    tangential component of that value, open and internal sides all of it."""
    u0 = np.array([0.4, -0.25])
    a1, a2 = np.array([2e-5, -1e-5]), np.array([-3e-5, 4e-5])
    for nvrt, kb in ((6, None), (2, None), (7, "mixed")):
        mm, xy, en, _ = _mesh(nx=18, ny=12, nvrt=nvrt)
        kbv = None if kb is None else (np.arange(xy.shape[0]) % 4).astype(np.int32)
        btu = np.broadcast_to(u0, (mm.ns, nvrt, 2)).copy()
        m, p, _ = _momentum_state(nvrt, kb=kbv, btu=btu, eta_old=lambda q: 0.1 + q @ a1,
                                  eta_new=lambda q: 0.12 + q @ a2)
        m.momentum(p)
        got = m.get(L.F_SIDE_U2)
        swet, skb, stype, snrm = _side_info(m)
        ref = u0 - p.dt * 9.81 * (p.theta * a2 + (1 - p.theta) * a1)
        tang = np.stack([-snrm[:, 1], snrm[:, 0]], axis=1)
        checked = 0
        for s in range(m.ns):
            if not swet[s]:
                assert np.all(got[s] == 0.0)
                continue
            exp = ref if stype[s] != 1 else (ref @ tang[s]) * tang[s]
            assert np.allclose(got[s, skb[s]:], exp, rtol=1e-12, atol=1e-14), (nvrt, s)
            assert np.all(got[s, :skb[s]] == 0.0)
            checked += 1
        assert checked > 0.9 * m.ns
        if kb is not None:
            assert len(set(skb.tolist())) > 2


def test_momentum_matches_galerkin_assembly():
    """3D sides with bottom drag, vertical viscosity and random btu profiles: the oracle's
    tridiagonal rows against the Galerkin system assembled layer by layer (linear elements on
    zc[kb+1..nvrt-1]: mass dz [[1/3, 1/6], [1/6, 1/3]], stiffness [[1, -1], [-1, 1]] / dz), drag
    dt chi on the first unknown, gradient from plane fits of the elevations, solved densely."""
    nv = 8
    mm, xy, en, _ = _mesh(nx=18, ny=12, nvrt=nv)
    kbv = (np.arange(xy.shape[0]) % 3).astype(np.int32)
    cd = 2.5e-3
    m, p, btu = _momentum_state(nv, kb=kbv, cd=cd, seed=4, eta_old=lambda q: 0.05 * np.sin(q[:, 0] / 900.0),
                                eta_new=lambda q: 0.07 * np.cos(q[:, 1] / 700.0))
    m.momentum(p)
    got = m.get(L.F_SIDE_U2)
    swet, skb, stype, _ = _side_info(m)
    zc = m.get(L.F_ZC)
    xyo, eno = m.topology(L.T_XY), m.topology(L.T_ELNODE)
    isdel, sn = m.topology(L.T_ISDEL), m.topology(L.T_ISIDENODE)
    wet = m.get(L.F_WET)
    eta1, eta2, dep = m.get(L.F_ETA1), m.get(L.F_ETA2), m.topology(L.T_DEPTH)
    elem_wet = wet[m.n:m.n + m.ne] != 0

    def plane_grad(e, f):
        A = np.c_[np.ones(3), xyo[eno[e]]]
        return np.linalg.solve(A, f[eno[e]])[1:]

    count = 0
    for s in range(m.ns):
        if not swet[s] or stype[s] == 1 or skb[s] == nv - 2:
            continue
        g1, g2, c1, c2 = np.zeros(2), np.zeros(2), 0, 0
        for e in isdel[s]:
            if e < 0:
                continue
            if elem_wet[e]:
                g1 += plane_grad(e, eta1)
                c1 += 1
            if np.all(eta2[eno[e]] + dep[eno[e]] > p.h0):
                g2 += plane_grad(e, eta2)
                c2 += 1
        g1, g2 = g1 / max(c1, 1), g2 / max(c2, 1)
        G = 9.81 * (p.theta * g2 + (1 - p.theta) * g1)
        k0 = skb[s] + 1
        z = zc[s, k0:]
        nr = z.size
        M, K = np.zeros((nr, nr)), np.zeros((nr, nr))
        for j in range(nr - 1):
            dz = z[j + 1] - z[j]
            M[j:j + 2, j:j + 2] += dz * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
            K[j:j + 2, j:j + 2] += np.array([[1.0, -1.0], [-1.0, 1.0]]) / dz
        chi = cd * np.hypot(*btu[s, k0])
        A = M + p.dt * p.dfv0 * K
        A[0, 0] += p.dt * chi
        rhs = M @ btu[s, k0:] - p.dt * np.outer(M.sum(axis=1), G)
        sol = np.linalg.solve(A, rhs)
        assert np.allclose(got[s, k0:], sol, rtol=1e-10, atol=1e-13), s
        assert np.all(got[s, skb[s]] == 0.0) and np.all(got[s, :skb[s]] == 0.0)
        count += 1
    assert count > 100


def test_momentum_land_sides_have_no_normal_flow():
    m, p, _ = _momentum_state(6, cd=2.5e-3, seed=9, eta_old=lambda q: 1e-5 * q[:, 0],
                              eta_new=lambda q: -2e-5 * q[:, 1])
    m.momentum(p)
    got = m.get(L.F_SIDE_U2)
    swet, skb, stype, snrm = _side_info(m)
    land = np.flatnonzero(stype == 1)
    assert land.size > 10
    un = np.einsum("skc,sc->sk", got[land], snrm[land])
    assert np.abs(un).max() < 1e-15
    assert np.abs(got[land]).max() > 1e-3


# ---------------------------------------------------------------- vertical velocity
def _vv_mesh(nvrt=6, depth=10.0, kb=None, eta=None):
    m, xy, en, ob = _mesh(nx=16, ny=10, nvrt=nvrt, kb=kb, depth=depth)
    xyo = m.topology(L.T_XY)
    m.set(L.F_ETA2, np.zeros(m.n) if eta is None else eta(xyo))
    m.levels(_params())
    return m


def test_vertical_velocity_divergence_closed_form():
    """Flat levels (constant depth and elevation), u = a_k x at level k on every side: the side
    fluxes of a linear field are exact, so w[k+1] = w[k] - 0.5 (a_k + a_k+1) dz and w[0] = 0."""
    nv = 6
    m = _vv_mesh(nvrt=nv)
    sc = m.geometry(2)
    a = 1e-5 * (1.0 + np.arange(nv))
    u2 = np.zeros((m.ns, nv, 2))
    u2[:, :, 0] = sc[:, :1] * a
    m.set(L.F_SIDE_U2, u2)
    m.vertical_velocity()
    w = m.get(L.F_ELEM_W)
    zc = m.get(L.F_ZC)[0]
    exp = np.zeros(nv)
    for k in range(nv - 1):
        exp[k + 1] = exp[k] - 0.5 * (a[k] + a[k + 1]) * (zc[k + 1] - zc[k])
    assert np.allclose(w, exp, rtol=1e-9, atol=1e-15)
    assert np.abs(exp).max() > 1e-5


def test_vertical_velocity_bottom_condition():
    """Uniform u0 over a bed sloping in x: w at each element's bottom level is -u0 . grad(h)."""
    nv = 5
    xy, en, ob = cases.channel(16, 10, 400.0, 0.2)
    b = 2e-3
    m = O.OracleMesh(xy, en, 8.0 + b * xy[:, 0], ob, cases.sigma_levels(nv))
    m.set(L.F_ETA2, np.zeros(m.n))
    m.levels(_params())
    u0 = np.array([0.3, -0.2])
    m.set(L.F_SIDE_U2, np.broadcast_to(u0, (m.ns, nv, 2)).copy())
    m.vertical_velocity()
    w = m.get(L.F_ELEM_W)
    assert np.allclose(w[:, 0], -u0[0] * b, rtol=1e-9, atol=1e-15)


def test_vertical_velocity_matches_prism_flux_balance():
    """Random side velocities, node-varying bottom levels and a non-flat elevation and bed: the
    oracle against the prism flux balance assembled from geometry (outward side normals from the
    element centroid, side faces len x clamped dz, top-face normals from cross products, upward
    flux Q[l+1] = Q[l] - outward side flux, Q[kb] = 0, w = (Q - u_top . n_xy A) / (n_z A))."""
    nv = 7
    xy, en, ob = cases.channel(14, 9, 500.0, 0.25)
    depth = 9.0 + 3e-3 * xy[:, 0] - 2e-3 * xy[:, 1]
    kb = (np.arange(xy.shape[0]) % 3).astype(np.int32)
    m = O.OracleMesh(xy, en, depth, ob, cases.sigma_levels(nv), kb=kb)
    xyo = m.topology(L.T_XY)
    m.set(L.F_ETA2, 0.3 * np.sin(xyo[:, 0] / 1500.0))
    m.levels(_params())
    rng = np.random.default_rng(3)
    u2 = rng.normal(0.0, 0.5, (m.ns, nv, 2))
    m.set(L.F_SIDE_U2, u2)
    m.vertical_velocity()
    w = m.get(L.F_ELEM_W)
    eno, es, sn = m.topology(L.T_ELNODE), m.topology(L.T_ELSIDE), m.topology(L.T_ISIDENODE)
    kbn = m.topology(L.T_KB)
    zc = m.get(L.F_ZC)
    wet = m.get(L.F_WET)
    swet = wet[m.n + m.ne:] != 0
    skb = np.where(swet, np.minimum(kbn[sn[:, 0]], kbn[sn[:, 1]]), 0)
    eta, dep = m.get(L.F_ETA2), m.topology(L.T_DEPTH)
    sig = cases.sigma_levels(nv)

    def node_z(i, k):
        k = max(k, kbn[i])
        return -dep[i] if k <= kbn[i] else (eta[i] if k == nv - 1 else (eta[i] + dep[i]) * sig[k] + eta[i])

    checked = 0
    for e in range(m.ne):
        if not wet[m.n + e]:
            assert np.all(w[e] == 0.0)
            continue
        kbe = kbn[eno[e]].min()
        ce = xyo[eno[e]].mean(axis=0)
        A = np.c_[np.ones(3), xyo[eno[e]]]
        gh = np.linalg.solve(A, dep[eno[e]])[1:]
        ub = np.mean([u2[s, skb[s]] for s in es[e]], axis=0)
        exp = np.zeros(nv)
        exp[kbe] = -gh @ ub
        Q = 0.0
        for l in range(kbe, nv - 1):
            out = 0.0
            for s in es[e]:
                p1, p2 = xyo[sn[s, 0]], xyo[sn[s, 1]]
                t = p2 - p1
                nrm = np.array([t[1], -t[0]]) / np.hypot(*t)
                if nrm @ (0.5 * (p1 + p2) - ce) < 0:
                    nrm = -nrm
                kt, kl = max(l + 1, skb[s]), max(l, skb[s])
                uc = 0.5 * (u2[s, kt] + u2[s, kl])
                out += (uc @ nrm) * np.hypot(*t) * (zc[s, kt] - zc[s, kl])
            Q = Q - out
            pts = np.array([[*xyo[i], node_z(i, l + 1)] for i in eno[e]])
            nvec = np.cross(pts[1] - pts[0], pts[2] - pts[0])
            if nvec[2] < 0:
                nvec = -nvec
            area3 = 0.5 * np.linalg.norm(nvec)
            nu = nvec / np.linalg.norm(nvec)
            ut = np.mean([u2[s, max(l + 1, skb[s])] for s in es[e]], axis=0)
            exp[l + 1] = (Q - (ut @ nu[:2]) * area3) / (nu[2] * area3)
        assert np.allclose(w[e], exp, rtol=1e-9, atol=1e-12), e
        checked += 1
    assert checked > 0.9 * m.ne


def test_shapiro_filter_against_neighbour_average():
    """With ishapiro = 1 the momentum output of every wet internal side becomes
    u + shapiro0 (mean of its four neighbours - u) above its bottom level, the neighbours being
    the other two sides of each of its elements (found here from the element side lists), each
    read at max(k, its kb + 1) when wet; the 2D bottom slot takes the surface value, levels below
    kb - 1 are cleared, boundary and dry sides keep the unfiltered values."""
    nv = 6
    mm, xy, en, _ = _mesh(nx=18, ny=12, nvrt=nv)
    kbv = (np.arange(xy.shape[0]) % 3).astype(np.int32)
    m, p, _ = _momentum_state(nv, kb=kbv, cd=2.5e-3, seed=8, eta_old=lambda q: 1e-5 * q[:, 0],
                              eta_new=lambda q: 0.2 * np.sin(q[:, 1] / 1500.0))
    m.momentum(p)
    raw = m.get(L.F_SIDE_U2)
    m2, p2, _ = _momentum_state(nv, kb=kbv, cd=2.5e-3, seed=8, eta_old=lambda q: 1e-5 * q[:, 0],
                                eta_new=lambda q: 0.2 * np.sin(q[:, 1] / 1500.0))
    p2.ishapiro, p2.shapiro0 = 1, 0.5
    m2.momentum(p2)
    got = m2.get(L.F_SIDE_U2)
    swet, skb, stype, _ = _side_info(m2)
    isdel, es = m2.topology(L.T_ISDEL), m2.topology(L.T_ELSIDE)
    exp = raw.copy()
    for s in range(m2.ns):
        if not swet[s] or isdel[s, 1] < 0:
            continue
        nbs = [t for e in isdel[s] for t in es[e] if t != s]
        assert len(nbs) == 4
        kb = skb[s]
        for k in range(kb + 1, nv):
            acc = sum(raw[n, max(k, skb[n] + 1) if swet[n] else k] for n in nbs)
            exp[s, k] = raw[s, k] + 0.5 * (0.25 * acc - raw[s, k])
        exp[s, :max(kb - 1, 0)] = 0.0
        if kb == nv - 2:
            exp[s, kb] = exp[s, nv - 1]
    assert np.allclose(got, exp, rtol=1e-13, atol=1e-15)
    assert np.abs(got - raw).max() > 1e-3
