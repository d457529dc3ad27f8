"""Pins of the oracle's right-hand-side producers and of the assembly RHS with every term active,
each shown to catch a planted mistake.

The reference ships no tests or golden vectors for this path, so these pins come from the
mathematics:
  - vertically integrated velocity U (meshdata.cu:4459-4476): the trapezoid rule is exact for a
    velocity linear in z, so U = a (zt - zb) + b (zt^2 - zb^2) / 2 over the side's wet levels;
  - chi = Cd |u_b| with u_b the velocity at the first level above the side's bed, and
    H^ = H - chi dt clipped at zero under ihhat = 1, H^2 / (H + chi dt) on two-level sides
    (meshdata.cu:4615-4725);
  - element-average G^ (meshdata.cu:6183-6424): a third of the sum over the element's sides of
    the exact depth integral of a linear backtracked velocity minus chi dt u_b;
  - the assembly RHS (meshdata.cu:3182-3311, 3373-3463, 4400-4446) against an element-by-element
    matrix build: b = M eta - fric K_wet eta + dt area [(1 - theta) grad(phi).U + theta grad(phi).G^]
    + dt/3 vsource, the open-boundary integral -(1 - theta) dt (len / 2) U.n_out at open nodes of
    elevation type 0 (n_out from the element centroid), then the essential rows.
Each planted mistake (OR_PLANT in oracle/gsh_oracle.c, built into its own library) must fail the
pin named for it."""
from __future__ import annotations

import numpy as np
import pytest

from gschism_amd import _lib as L
from gschism_amd import cases
from oracle import build as OB
from oracle import pyoracle as O

G = 9.81


def _island_depth(xy):
    cx, cy = xy[:, 0].mean(), xy[:, 1].mean()
    r2 = ((xy[:, 0] - cx) ** 2 + (xy[:, 1] - cy) ** 2) / (1500.0 ** 2)
    return 9.0 - 13.0 * np.exp(-r2)


def _stretched_sigma(nv):
    s = np.linspace(0.0, 1.0, nv)
    return -1.0 + s ** 1.4


def _side_kb(m, kb):
    sn = m.topology(L.T_ISIDENODE)
    return np.minimum(kb[sn[:, 0]], kb[sn[:, 1]])


# ---------------------------------------------------------------- RHS producers
def pin_rhs_vertical_profiles(library):
    """U, chi, H^ and G^ for velocities linear in z over stretched levels and node-varying kb."""
    nv = 9
    xy, en, ob = cases.channel(22, 14, 350.0)
    depth = 6.0 + 5.0 * np.sin(xy[:, 0] / 1300.0) ** 2
    kb = np.clip((8.0 - depth).astype(np.int32), 0, 3)
    m = O.OracleMesh(xy, en, depth, ob, _stretched_sigma(nv), kb=kb, library=library)
    n = m.n
    eta = 0.3 * np.cos(xy[:, 1] / 700.0)
    m.set(L.F_ETA2, eta)
    rng = np.random.default_rng(11)
    cd = rng.uniform(1e-3, 6e-2, n)  # large enough that H - chi dt < 0 on some sides
    m.set(L.F_CD, cd)
    dt = 240.0
    p = L.default_step_params(dt=dt, ihhat=1)
    m.levels(p)
    zc = m.get(L.F_ZC)
    skb = _side_kb(m, kb)
    ns = m.ns
    # velocities linear in z per side: u = a + b z (side u2), v = c + d z (btu)
    a = rng.normal(0.0, 0.5, (ns, 2))
    b = rng.normal(0.0, 0.08, (ns, 2))
    c = rng.normal(0.0, 0.5, (ns, 2))
    d = rng.normal(0.0, 0.08, (ns, 2))
    su2 = a[:, None, :] + b[:, None, :] * zc[:, :, None]
    btu = c[:, None, :] + d[:, None, :] * zc[:, :, None]
    m.set(L.F_SIDE_U2, su2)
    m.set(L.F_BTU, btu)
    m.rhs(p)
    wet = m.get(L.F_WET)
    side_wet = wet[n + m.ne:] != 0
    idx = np.arange(ns)
    zb, zt = zc[idx, skb], zc[:, nv - 1]

    def integral(p0, p1):  # exact depth integral of p0 + p1 z over [zb, zt]
        return p0 * (zt - zb)[:, None] + 0.5 * p1 * (zt ** 2 - zb ** 2)[:, None]

    U = m.get(L.F_SIDE_U)
    U_exp = np.where(side_wet[:, None], integral(a, b), 0.0)
    assert np.allclose(U, U_exp, rtol=1e-12, atol=1e-12 * np.abs(U_exp).max()), "U"
    assert (skb > 0).any() and (d != 0).all()
    sn = m.topology(L.T_ISIDENODE)
    ub = c + d * zc[idx, skb + 1][:, None]  # the level just above the side's bed
    cd_s = 0.5 * (cd[sn[:, 0]] + cd[sn[:, 1]])
    chi_exp = np.where(side_wet, cd_s * np.hypot(ub[:, 0], ub[:, 1]), 0.0)
    chi = m.get(L.F_SIDE_CHI)
    assert np.allclose(chi, chi_exp, rtol=1e-13, atol=0.0), "chi"
    H = 0.5 * (eta[sn[:, 0]] + eta[sn[:, 1]]) + 0.5 * (depth[sn[:, 0]] + depth[sn[:, 1]])
    two = skb == nv - 2
    hh_exp = np.where(two, H * H / (H + chi_exp * dt), np.maximum(0.0, H - chi_exp * dt))
    hh_exp = np.where(side_wet, hh_exp, 0.0)
    hh = m.get(L.F_SIDE_HHAT)
    assert np.allclose(hh, hh_exp, rtol=1e-12, atol=1e-12), "H^"
    assert ((H - chi_exp * dt < 0) & side_wet & ~two).sum() > 5, "the case must clip some sides"
    # G^: a third of the sum over the element's sides of (int btu dz) - chi dt u_b (3-D sides)
    es = m.topology(L.T_ELSIDE)
    assert not two.any()
    per_side = integral(c, d) - (chi_exp * dt)[:, None] * ub
    elem_wet = wet[n:n + m.ne] != 0
    g_exp = np.where(elem_wet[:, None], per_side[es].sum(axis=1) / 3.0, 0.0)
    g = m.get(L.F_GHAT1)
    assert np.allclose(g, g_exp, rtol=1e-11, atol=1e-11 * np.abs(g_exp).max()), "G^"


# ---------------------------------------------------------------- assembly RHS
def _outward_normals(xy, en, sn, isdel):
    """Unit normals of the sides pointing away from the centroid of each side's first element."""
    p1, p2 = xy[sn[:, 0]], xy[sn[:, 1]]
    t = p2 - p1
    nrm = np.stack([t[:, 1], -t[:, 0]], axis=1) / np.hypot(t[:, 0], t[:, 1])[:, None]
    cen = xy[en[isdel[:, 0]]].mean(axis=1)
    flip = np.einsum("ij,ij->i", nrm, 0.5 * (p1 + p2) - cen) < 0
    nrm[flip] *= -1
    return nrm


def _dense_rhs(m, xy, en, eta, grav, hh, U, gh, vsrc, elem_wet, side_wet, theta, dt, free_open, ess, elbc):
    n = xy.shape[0]
    es = m.topology(L.T_ELSIDE)
    A = np.zeros((n, n))
    b = np.zeros(n)
    for e, tri in enumerate(en):
        p = xy[tri]
        area = 0.5 * ((p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1]))
        M = area / 12.0 * (np.ones((3, 3)) + np.eye(3))
        gr = np.array([[p[1, 1] - p[2, 1], p[2, 0] - p[1, 0]], [p[2, 1] - p[0, 1], p[0, 0] - p[2, 0]],
                       [p[0, 1] - p[1, 1], p[1, 0] - p[0, 0]]]) / (2 * area)
        K = area * gr @ gr.T
        g3 = grav[tri].mean()
        hb = hh[es[e]].mean()
        Ub = U[es[e]].mean(axis=0)
        A[np.ix_(tri, tri)] += M + g3 * theta * theta * dt * dt * hb * K
        fric = g3 * theta * (1.0 - theta) * dt * dt * hb * (1.0 if elem_wet[e] else 0.0)
        b[tri] += M @ eta[tri] - fric * (K @ eta[tri])
        b[tri] += dt * area * ((1.0 - theta) * gr @ Ub + theta * gr @ gh[e])
        b[tri] += dt / 3.0 * vsrc[e]
    sn, isdel = m.topology(L.T_ISIDENODE), m.topology(L.T_ISDEL)
    nout = _outward_normals(xy, en, sn, isdel)
    slen = np.hypot(*(xy[sn[:, 1]] - xy[sn[:, 0]]).T)
    for s in np.flatnonzero(isdel[:, 1] < 0):
        if not side_wet[s]:
            continue
        flux = (1.0 - theta) * dt * 0.5 * slen[s] * (U[s] @ nout[s])
        for i in sn[s]:
            if free_open[i]:
                b[i] -= flux
    # essential rows: identity with elbc; columns eliminated into the other rows
    bfix = b - A[:, ess] @ elbc[ess]
    bfix[ess] = elbc[ess]
    return bfix


def pin_assembly_rhs_dense(library):
    """The assembled RHS with non-zero U and G^, varying eta, gravity and H^, dry elements, a volume
    source, open nodes of type 0 (I_5) and essential open nodes, against the dense build."""
    xy, en, ob = cases.channel(16, 11, 300.0)
    depth = _island_depth(xy)
    m = O.OracleMesh(xy, en, depth, ob, cases.sigma_levels(4), library=library)
    n, ne, ns = m.n, m.ne, m.ns
    rng = np.random.default_rng(5)
    eta = 0.2 * np.sin(xy[:, 0] / 600.0) + 0.05 * rng.normal(size=n)
    theta, dt = 0.63, 180.0
    p = L.default_step_params(dt=dt, theta=theta)
    m.set(L.F_ETA2, eta)
    m.levels(p)
    wet = m.get(L.F_WET)
    elem_wet, side_wet = wet[n:n + ne] != 0, wet[n + ne:] != 0
    assert (~elem_wet).sum() > 3 and (~side_wet).any()
    grav = G + 0.01 * rng.normal(size=n)
    hh = rng.uniform(1.0, 9.0, ns)
    U = rng.normal(0.0, 2.0, (ns, 2))
    gh = rng.normal(0.0, 2.0, (ne, 2))
    vsrc = rng.normal(0.0, 0.5, ne)
    opn = np.flatnonzero(ob)
    etype = np.full(n, -1, dtype=np.int8)
    etype[opn] = np.where(np.arange(opn.size) % 3 == 0, 2, 0)  # every third open node essential
    elbc = np.zeros(n)
    elbc[opn] = rng.normal(0.0, 0.3, opn.size)
    for code, v in ((L.F_GRAV2, grav), (L.F_SIDE_HHAT, hh), (L.F_SIDE_U, U), (L.F_GHAT1, gh),
                    (L.F_VSOURCE, vsrc), (L.F_ETYPE, etype), (L.F_ELBC, elbc)):
        m.set(code, v)
    val, b = m.assemble(p)
    free_open = np.zeros(n, bool)
    free_open[opn[etype[opn] == 0]] = True
    ess = np.zeros(n, bool)
    ess[opn[etype[opn] != 0]] = True
    assert free_open.sum() >= 4
    b_exp = _dense_rhs(m, xy, en, eta, grav, hh, U, gh, vsrc, elem_wet, side_wet, theta, dt, free_open, ess, elbc)
    scale = np.abs(b_exp).max()
    assert np.allclose(b, b_exp, rtol=1e-11, atol=1e-12 * scale), np.abs(b - b_exp).max() / scale


def pin_open_boundary_integral_closed_form(library):
    """A uniform depth-integrated flow U0 through an open boundary of type-0 nodes: the RHS of a
    type-0 node differs from the same node made essential-free without I_5 by
    -(1 - theta) dt (l1 + l2) / 2 U0.n_out over its two boundary sides."""
    xy, en, ob = cases.channel(12, 8, 400.0)
    m = O.OracleMesh(xy, en, np.full(xy.shape[0], 10.0), ob, cases.sigma_levels(3), library=library)
    n = m.n
    theta, dt = 0.55, 150.0
    p = L.default_step_params(dt=dt, theta=theta, iettype=0)
    m.set(L.F_ETA2, np.zeros(n))
    m.levels(p)
    U0 = np.array([1.7, -0.4])
    m.set(L.F_SIDE_U, np.tile(U0, (m.ns, 1)))
    m.set(L.F_GHAT1, np.zeros((m.ne, 2)))
    _, b0 = m.assemble(p)
    # the same RHS with the open flags off: no I_5 anywhere
    m2 = O.OracleMesh(xy, en, np.full(n, 10.0), np.zeros(n, np.int8), cases.sigma_levels(3), library=library)
    m2.set(L.F_ETA2, np.zeros(n))
    m2.levels(p)
    m2.set(L.F_SIDE_U, np.tile(U0, (m.ns, 1)))
    m2.set(L.F_GHAT1, np.zeros((m.ne, 2)))
    _, b1 = m2.assemble(p)
    sn, isdel = m.topology(L.T_ISIDENODE), m.topology(L.T_ISDEL)
    nout = _outward_normals(xy, en, sn, isdel)
    slen = np.hypot(*(xy[sn[:, 1]] - xy[sn[:, 0]]).T)
    exp = np.zeros(n)
    for s in np.flatnonzero(isdel[:, 1] < 0):
        for i in sn[s]:
            if ob[i]:
                exp[i] -= (1.0 - theta) * dt * 0.5 * slen[s] * (U0 @ nout[s])
    diff = b0 - b1
    assert np.allclose(diff, exp, rtol=1e-12, atol=1e-9), np.abs(diff - exp).max()
    assert np.all(diff[~ob.astype(bool)] == 0.0)
    # the west edge is open and U0 points east: inflow, so the RHS grows there
    assert (exp[ob.astype(bool)] > 0).all()


PINS = {
    "rhs_vertical_profiles": pin_rhs_vertical_profiles,
    "assembly_rhs_dense": pin_assembly_rhs_dense,
    "open_boundary_integral": pin_open_boundary_integral_closed_form,
}

# planted mistake -> the pin that must catch it
PLANTED = {
    1: ("theta and 1 - theta swapped in the advective term I_4", "assembly_rhs_dense"),
    2: ("wet flag dropped from the friction term", "assembly_rhs_dense"),
    3: ("sign of the open-boundary integral I_5 flipped", "assembly_rhs_dense"),
    4: ("chi from the bed level kb instead of kb + 1", "rhs_vertical_profiles"),
    5: ("chi from level kb + 2 instead of kb + 1", "rhs_vertical_profiles"),
    6: ("U summed with the lower level's velocity (rectangle, not trapezoid)", "rhs_vertical_profiles"),
    7: ("H^ not clipped at zero under ihhat = 1", "rhs_vertical_profiles"),
    8: ("G^ drag term with u_b from level kb", "rhs_vertical_profiles"),
    9: ("G^ depth integral with the upper level's velocity only", "rhs_vertical_profiles"),
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin_holds(name):
    PINS[name](O.lib())


@pytest.mark.parametrize("k", sorted(PLANTED))
def test_planted_mistake_fails_its_pin(k):
    what, pin = PLANTED[k]
    planted = O.load(OB.build_planted(k))
    with pytest.raises(AssertionError):
        PINS[pin](planted)


def test_open_boundary_integral_pin_catches_a_flipped_sign():
    planted = O.load(OB.build_planted(3))
    with pytest.raises(AssertionError):
        pin_open_boundary_integral_closed_form(planted)
