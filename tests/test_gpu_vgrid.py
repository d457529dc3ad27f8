"""GPU parity of the whole step on SZ (ivcor=2) and localized-sigma (ivcor=1, a sigma per node and
level) vertical grids against the CPU oracle, bit for bit: levels, backtracking, RHS producers,
assembly, solve, momentum, vertical velocity and nodal velocities."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_bitwise
from gschism_amd import _lib as L
from gschism_amd import cases, core
from oracle import pyoracle as O
from test_gpu_parity import _full_chain

pytestmark = pytest.mark.gpu


def _basin_depth(xy):
    """A deep trough (beyond h_s, on the Z levels), a shelf between h_c and h_s, shallows below
    h_c and an island above the water."""
    x0, x1 = xy[:, 0].min(), xy[:, 0].max()
    y0, y1 = xy[:, 1].min(), xy[:, 1].max()
    t = (xy[:, 0] - x0) / (x1 - x0)
    trough = 110.0 * np.exp(-((xy[:, 1] - (y0 + 0.3 * (y1 - y0))) / (0.12 * (y1 - y0))) ** 2)
    h = 4.0 + 30.0 * t + trough
    cx, cy = x0 + 0.7 * (x1 - x0), y0 + 0.75 * (y1 - y0)
    r2 = ((xy[:, 0] - cx) ** 2 + (xy[:, 1] - cy) ** 2) / (1500.0 ** 2)
    return h - 45.0 * np.exp(-r2)


def _setup_vgrid(ctx, spec, depth, vg, kb=None):
    data = cases.make_case(spec)
    mesh = core.Mesh(ctx, data.xy, data.elnode, depth, None, open_bnd=data.open_bnd, kb=kb, vgrid=vg)
    xy = mesh.topology(L.T_XY)
    sig = np.linspace(-1.0, 0.0, mesh.nvrt)  # vertical shear profile of the synthetic velocity
    eta, uvw, su2 = cases.initial_state(spec, xy, cases.side_centres(mesh), sig)
    eta = eta + cases.normal(cases.SEED_ETA, spec.n, 0.0, spec.eta_noise)[mesh.node_perm]
    p = cases.step_params(spec)
    for f, v in ((L.F_ETA2, eta), (L.F_NODE_UVW, uvw), (L.F_SIDE_U2, su2)):
        mesh.set(f, v)
    mesh.update_levels(p)
    om = O.OracleMesh(xy, mesh.topology(L.T_ELNODE), mesh.topology(L.T_DEPTH), mesh.topology(L.T_OPEN_BND),
                      np.zeros(mesh.nvrt), kb=mesh.topology(L.T_KB))
    om.set_vertical(mesh.vgrid)
    for f, v in ((L.F_ETA2, eta), (L.F_NODE_UVW, uvw), (L.F_SIDE_U2, su2)):
        om.set(f, v)
    om.levels(p)
    assert_bitwise(mesh.topology(L.T_KB), om.topology(L.T_KB), "bottom levels")
    return mesh, om, p


def test_chain_sz_levels(ctx):
    spec = cases.CaseSpec("sz-basin-48x32", 48, 32, 400.0, nvrt=16, depth=10.0, substeps=6)
    data = cases.make_case(spec)
    depth = _basin_depth(data.xy)
    nz, h_s = 4, 40.0
    nsig = spec.nvrt - nz + 1
    sigma_s = -1.0 + np.linspace(0.0, 1.0, nsig) ** 1.4
    sigma_s[0], sigma_s[-1] = -1.0, 0.0
    vg = core.VGrid(2, nz=nz, h_s=h_s, h_c=10.0, theta_b=0.5, theta_f=4.0,
                    ztot=np.array([-160.0, -110.0, -70.0, -h_s]), sigma_s=sigma_s)
    mesh, om, p = _setup_vgrid(ctx, spec, depth, vg)
    kb = mesh.topology(L.T_KB)
    print(f"[gpu] SZ mesh: kb values {np.unique(kb)}, depth {depth.min():.1f}..{depth.max():.1f}", flush=True)
    assert (kb < nz - 1).any() and (kb == nz - 1).any()
    assert (mesh.get(L.F_WET) == 0).any(), "the island should leave dry entities"
    _full_chain(ctx, mesh, om, p)
    # a new bathymetry moves the bottom levels (Node::updateKb): set, level, compare again
    d2 = mesh.topology(L.T_DEPTH) * 1.1
    mesh.set(L.F_DEPTH, d2)
    om.set(L.F_DEPTH, d2)
    mesh.update_levels(p)
    om.levels(p)
    assert_bitwise(mesh.topology(L.T_KB), om.topology(L.T_KB), "bottom levels after a depth change")
    assert_bitwise(mesh.get(L.F_ZC), om.get(L.F_ZC), "zc after a depth change")
    mesh.backtrack(p)
    om.backtrack(p, nthreads=8)
    assert_bitwise(mesh.get(L.F_FOOT), om.get(L.F_FOOT), "foot elements after a depth change")
    assert_bitwise(mesh.get(L.F_BTU), om.get(L.F_BTU), "backtracked velocity after a depth change")


def test_chain_lsc2_per_node_sigma(ctx):
    spec = cases.CaseSpec("lsc2-48x32", 48, 32, 400.0, nvrt=12, depth=10.0, substeps=6)
    data = cases.make_case(spec)
    n = data.xy.shape[0]
    depth = 3.0 + 25.0 * (data.xy[:, 0] - data.xy[:, 0].min()) / np.ptp(data.xy[:, 0])
    kb = np.clip((6 - depth / 3.0).astype(np.int32), 0, 5)
    rng = np.random.default_rng(8)
    expo = rng.uniform(0.7, 2.2, n)
    sn = -1.0 + np.linspace(0.0, 1.0, spec.nvrt)[None, :] ** expo[:, None]
    vg = core.VGrid(1, sigma_node=sn)
    mesh, om, p = _setup_vgrid(ctx, spec, depth, vg, kb=kb)
    assert (mesh.topology(L.T_KB) > 0).any()
    _full_chain(ctx, mesh, om, p)
