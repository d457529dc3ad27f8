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
    # the LDS-staged element walk on per-node sigma levels
    mesh.set_backtrack_staging(True)
    mesh.backtrack(p)
    om.backtrack(p, nthreads=8)
    assert_bitwise(mesh.get(L.F_FOOT), om.get(L.F_FOOT), "foot elements, staged walk")
    assert_bitwise(mesh.get(L.F_BTU), om.get(L.F_BTU), "backtracked velocity, staged walk")


def test_run_like_case_from_files(ctx, tmp_path):
    """A case shaped like the reference's configured run (constants.h:53, 123-126, 140, 166:
    ivcor=2 with one Z level, nvrt=23, dt=300 s, theta=0.51, 10-30 s ELM substeps): a 129 x 41
    channel at 500 m spacing written as hgrid.gr3, vgrid.in (SZ) and bctides.in (a constant
    elevation on the open boundary), read back through the library's readers and run through a
    whole step against the oracle, bit for bit."""
    from gschism_amd.bctides import apply_bctides, read_bctides
    from gschism_amd.grid_files import read_grid, write_hgrid, write_vgrid_sz
    from test_cpu_grid_files import _boundary_segments

    xy, en, ob = cases.channel(129, 41, 500.0)
    x = (xy[:, 0] - xy[:, 0].min()) / np.ptp(xy[:, 0])
    depth = 4.0 + 46.0 * x + 3.0 * np.sin(xy[:, 1] / 1500.0)
    opens, lands = _boundary_segments(xy, en, ob)
    hpath, vpath, bpath = (str(tmp_path / f) for f in ("hgrid.gr3", "vgrid.in", "bctides.in"))
    write_hgrid(hpath, xy, depth, en, opens, lands, name="run-like channel 129x41")
    nvrt = 23
    sigma_s = -1.0 + np.linspace(0.0, 1.0, nvrt) ** 1.2
    sigma_s[0], sigma_s[-1] = -1.0, 0.0
    write_vgrid_sz(vpath, nvrt, [-100.0], 100.0, 30.0, 0.7, 5.0, sigma_s)
    k = opens[0].size
    (tmp_path / "bctides.in").write_text("\n".join(["01/01/2020 00:00:00", "0 40.", "0", "1",
                                                    f"{k} 2 0 0 0", "0.25"]) + "\n")
    g = read_grid(hpath, vpath)
    assert g.vgrid.ivcor == 2 and g.vgrid.nz == 1 and g.vgrid.nvrt == nvrt
    mesh = g.mesh(ctx)
    apply_bctides(mesh, read_bctides(bpath, 1, k), g.open_segments)
    kb = mesh.topology(L.T_KB)
    assert (kb == 0).all()
    spec = cases.CaseSpec("run-like-129x41", 129, 41, 500.0, nvrt=nvrt, dt=300.0, theta=0.51)
    p = L.default_step_params(dt=300.0, theta=0.51, dtb_min=10.0, dtb_max=30.0)
    xym = mesh.topology(L.T_XY)
    eta, uvw, su2 = cases.initial_state(spec, xym, cases.side_centres(mesh), np.linspace(-1.0, 0.0, nvrt))
    for f, v in ((L.F_ETA2, eta), (L.F_NODE_UVW, uvw), (L.F_SIDE_U2, su2)):
        mesh.set(f, v)
    mesh.update_levels(p)
    om = O.OracleMesh(xym, mesh.topology(L.T_ELNODE), mesh.topology(L.T_DEPTH), mesh.topology(L.T_OPEN_BND),
                      np.zeros(nvrt), kb=kb)
    om.set_vertical(mesh.vgrid)
    for f, v in ((L.F_ETA2, eta), (L.F_NODE_UVW, uvw), (L.F_SIDE_U2, su2), (L.F_ETYPE, mesh.get(L.F_ETYPE)),
                 (L.F_ETH, mesh.get(L.F_ETH))):
        om.set(f, v)
    om.levels(p)
    h = mesh.topology(L.T_DEPTH)
    assert (h > 30.0).any() and (h < 30.0).any()  # both branches of the S levels
    print(f"[gpu] run-like case from files: n={mesh.n} ns={mesh.ns} nvrt={nvrt}", flush=True)
    _full_chain(ctx, mesh, om, p)
    st = mesh.get(L.F_BSTATUS)
    assert (st & L.BT_OPEN_EXIT).any() or (st & L.BT_SLIDE).any()
