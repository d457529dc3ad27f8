"""Vertical grids beyond one sigma per level, on the CPU: SZ (ivcor=2) and localized sigma with a
sigma per node and level (ivcor=1, LSC2).

The reference ships no vgrid fixtures for this path, so the oracle's z levels
(Node::compZCoord / updateKb, meshdata.cu:1676-1807) are pinned by what the SZ definition fixes:
  - a pure-S grid (nz = 1) on nodes no deeper than h_c has z = sigma (h + eta) + eta, the same
    levels as the ivcor=1 grid with those sigmas (compared bit for bit with that oracle path);
  - for h_c < h <= h_s the S levels tend to that linear profile as theta_f -> 0 with theta_b = 0
    (C(sigma) = sinh(theta_f sigma) / sinh(theta_f) -> sigma), within O(theta_f^2);
  - every node's levels run from -h at its bed level to eta at the surface, strictly increasing;
    the level nz-1 sits at -min(h, h_s) and the Z levels between the bed and it are ztot exactly;
  - the bottom level is the Z level just below the bed (library, written as a scan, against the
    oracle's bisection) for depths on, between and beyond the Z levels;
  - ivcor=1 with every node carrying the same sigma column equals the one-sigma-per-level path bit
    for bit, and with distinct columns a side's z is the mean of its nodes' (eta + h) sigma + eta.
Planted mistakes (OR_PLANT 10, 11) must each fail their pin. The library's own checks of a grid
(readVGrid's, read_and_process_grid.cpp:1412-1486) are refusals with GSH_E_ARG."""
from __future__ import annotations

import numpy as np
import pytest

from gschism_amd import _lib as L
from gschism_amd import cases, core
from oracle import build as OB
from oracle import pyoracle as O


def _params(**kw):
    p = L.default_step_params(**kw)
    return p


def _mesh_xy(nx=12, ny=8, dx=400.0):
    return cases.channel(nx, ny, dx)


def _sz(nvrt=12, nz=4, h_s=40.0, h_c=10.0, theta_b=0.0, theta_f=3.0):
    ztot = np.concatenate([np.array([-120.0, -80.0, -55.0])[: nz - 1], [-h_s]])
    nsig = nvrt - nz + 1
    s = np.linspace(0.0, 1.0, nsig)
    sigma_s = -1.0 + s ** 1.3
    sigma_s[0], sigma_s[-1] = -1.0, 0.0
    return core.VGrid(2, nz=nz, h_s=h_s, h_c=h_c, theta_b=theta_b, theta_f=theta_f, ztot=ztot, sigma_s=sigma_s)


def _oracle(xy, en, ob, depth, vg, library=None, eta=None):
    om = O.OracleMesh(xy, en, depth, ob, np.zeros(vg.nvrt), library=library)
    om.set_vertical(vg)
    if eta is not None:
        om.set(L.F_ETA2, eta)
    om.levels(_params())
    return om


def _side_nodes(om):
    return om.topology(L.T_ISIDENODE)


# ---------------------------------------------------------------- bottom levels
def test_sz_bottom_levels_library_against_oracle():
    xy, en, ob = _mesh_xy()
    vg = _sz(nvrt=12, nz=4)
    rng = np.random.default_rng(3)
    depth = rng.uniform(1.0, 119.0, xy.shape[0])
    # depths exactly on the Z levels and at h_s
    depth[:5] = [120.0, 80.0, 55.0, 40.0, 39.999]
    m = core.Mesh(None, xy, en, depth, None, ob, renumber=False, vgrid=vg)
    om = _oracle(xy, en, ob, depth, vg)
    kb_lib, kb_or = m.topology(L.T_KB), om.topology(L.T_KB)
    np.testing.assert_array_equal(kb_lib, kb_or)
    assert kb_lib[0] == 0 and kb_lib[1] == 1 and kb_lib[2] == 2 and kb_lib[3] == 3 and kb_lib[4] == 3
    shallow = depth <= vg.h_s
    assert (kb_lib[shallow] == vg.nz - 1).all()


def test_sz_grid_refusals():
    xy, en, ob = _mesh_xy()
    depth = np.full(xy.shape[0], 20.0)
    good = _sz()
    bad = []
    for kw in (dict(h_s=5.0), dict(h_c=4.0), dict(h_c=40.0), dict(theta_b=1.5), dict(theta_f=0.0)):
        args = dict(nvrt=12, nz=4, h_s=40.0, h_c=10.0, theta_b=0.0, theta_f=3.0)
        args.update(kw)
        v = _sz(**args)
        if "h_s" in kw:
            v.ztot[-1] = -kw["h_s"]
        bad.append(v)
    v = _sz()
    v.ztot[1] = v.ztot[0] - 1.0  # inverted Z levels
    bad.append(v)
    v = _sz()
    v.sigma_s[0] = -0.9  # S levels must start at -1
    bad.append(v)
    for v in bad:
        with pytest.raises(L.GshError) as ei:
            core.Mesh(None, xy, en, depth, None, ob, renumber=False, vgrid=v)
        assert ei.value.code == L.GSH_E_ARG
    core.Mesh(None, xy, en, depth, None, ob, renumber=False, vgrid=good)
    deep = depth.copy()
    deep[7] = 500.0  # below the deepest Z level
    with pytest.raises(L.GshError) as ei:
        core.Mesh(None, xy, en, deep, None, ob, renumber=False, vgrid=good)
    assert ei.value.code == L.GSH_E_ARG and "deepest Z level" in str(ei.value)


# ---------------------------------------------------------------- z levels (oracle pins)
def pin_sz_pure_s_equals_sigma(library):
    """nz = 1, every node no deeper than h_c: the SZ levels are the ivcor=1 levels of sigma_s."""
    xy, en, ob = _mesh_xy()
    n = xy.shape[0]
    vg = _sz(nvrt=9, nz=1, h_s=30.0, h_c=12.0, theta_f=4.0)
    depth = np.linspace(2.0, 12.0, n)
    eta = 0.4 * np.sin(xy[:, 0] / 900.0)
    a = _oracle(xy, en, ob, depth, vg, library=library, eta=eta)
    b = O.OracleMesh(xy, en, depth, ob, vg.sigma_s, library=library)
    b.set(L.F_ETA2, eta)
    b.levels(_params())
    np.testing.assert_array_equal(a.get(L.F_ZC), b.get(L.F_ZC))


def pin_sz_small_theta_f_is_linear(library):
    """h_c < h <= h_s, theta_b = 0, theta_f -> 0: z -> sigma (h + eta) + eta within O(theta_f^2)."""
    xy, en, ob = _mesh_xy()
    n = xy.shape[0]
    tf = 1e-3
    vg = _sz(nvrt=10, nz=1, h_s=60.0, h_c=6.0, theta_f=tf)
    depth = np.linspace(15.0, 60.0, n)
    eta = 0.3 * np.cos(xy[:, 1] / 700.0)
    om = _oracle(xy, en, ob, depth, vg, library=library, eta=eta)
    sn = _side_nodes(om)
    zc = om.get(L.F_ZC)
    s = vg.sigma_s
    for side in range(0, sn.shape[0], 7):
        i, j = sn[side]
        zi = s * (depth[i] + eta[i]) + eta[i]
        zj = s * (depth[j] + eta[j]) + eta[j]
        expect = 0.5 * (zi + zj)
        expect[0] = -0.5 * (depth[i] + depth[j])
        expect[-1] = 0.5 * (eta[i] + eta[j])
        assert np.max(np.abs(zc[side] - expect)) < 1e-5 * 60.0


def pin_sz_structure(library):
    """Levels of every side (mean of its nodes' levels above their beds): the bed, ztot between
    the bed and level nz-1, -min(h, h_s) there, eta at the surface, strictly increasing."""
    xy, en, ob = _mesh_xy()
    n = xy.shape[0]
    vg = _sz(nvrt=12, nz=4, h_s=40.0, h_c=10.0, theta_b=0.3, theta_f=3.0)
    rng = np.random.default_rng(5)
    # uniform depth per test mesh so each side's two nodes share it
    for h in (8.0, 25.0, 40.0, 60.0, 100.0):
        depth = np.full(n, h)
        eta = np.full(n, 0.25)
        om = _oracle(xy, en, ob, depth, vg, library=library, eta=eta)
        zc = om.get(L.F_ZC)
        kb = om.topology(L.T_KB)[0]
        iz = vg.nz - 1
        for side in rng.choice(zc.shape[0], 20, replace=False):
            z = zc[side]
            assert z[kb] == -h
            assert z[-1] == 0.25
            assert z[iz] == -min(h, vg.h_s)
            for k in range(kb + 1, iz):
                assert z[k] == vg.ztot[k]
            assert (np.diff(z[kb:]) > 0).all()
            # S levels above nz-1 lie between -min(h, h_s) and eta
            assert (z[iz + 1:-1] > -min(h, vg.h_s)).all() and (z[iz + 1:-1] < 0.25).all()


def pin_lsc2_per_node_sigma(library):
    """ivcor=1 with a sigma per node: all columns equal -> the one-sigma path bit for bit; distinct
    columns -> a side's z is the mean of its nodes' (eta + h) sigma_i + eta above their beds."""
    xy, en, ob = _mesh_xy()
    n = xy.shape[0]
    nv = 8
    sigma = -1.0 + np.linspace(0.0, 1.0, nv) ** 1.2
    depth = np.linspace(6.0, 20.0, n)
    eta = 0.2 * np.sin(xy[:, 0] / 500.0)
    same = core.VGrid(1, sigma_node=np.tile(sigma, (n, 1)))
    a = _oracle(xy, en, ob, depth, same, library=library, eta=eta)
    b = O.OracleMesh(xy, en, depth, ob, sigma, library=library)
    b.set(L.F_ETA2, eta)
    b.levels(_params())
    np.testing.assert_array_equal(a.get(L.F_ZC), b.get(L.F_ZC))
    # distinct columns: node i stretched by its own exponent
    expo = np.linspace(0.8, 2.0, n)
    sn = -1.0 + np.linspace(0.0, 1.0, nv)[None, :] ** expo[:, None]
    om = _oracle(xy, en, ob, depth, core.VGrid(1, sigma_node=sn), library=library, eta=eta)
    zc = om.get(L.F_ZC)
    nodes = _side_nodes(om)
    for side in range(0, nodes.shape[0], 5):
        i, j = nodes[side]
        zi = (eta[i] + depth[i]) * sn[i] + eta[i]
        zj = (eta[j] + depth[j]) * sn[j] + eta[j]
        zi[0], zj[0], zi[-1], zj[-1] = -depth[i], -depth[j], eta[i], eta[j]
        np.testing.assert_array_equal(zc[side], 0.5 * (zi + zj))


PINS = {
    "sz_pure_s_equals_sigma": pin_sz_pure_s_equals_sigma,
    "sz_small_theta_f_is_linear": pin_sz_small_theta_f_is_linear,
    "sz_structure": pin_sz_structure,
    "lsc2_per_node_sigma": pin_lsc2_per_node_sigma,
}

# planted mistake -> the pin that must catch it
PLANTED = {
    10: ("S level index from k - nz instead of k - (nz - 1)", "sz_structure"),
    11: ("h_c and min(h, h_s) swapped in the stretched S formula", "sz_small_theta_f_is_linear"),
    12: ("per-node sigma read from the neighbouring node", "lsc2_per_node_sigma"),
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_vgrid_pin_holds(name):
    PINS[name](O.lib())


@pytest.mark.parametrize("k", sorted(PLANTED))
def test_vgrid_planted_mistake_fails_its_pin(k):
    what, pin = PLANTED[k]
    planted = O.load(OB.build_planted(k))
    with pytest.raises(AssertionError):
        PINS[pin](planted)
