"""hgrid.gr3 / vgrid.in readers (gsh_grid_*): a hand-written file with known contents, round trips
of synthetic meshes through the writers, the reference reader's checks as errors, and a mesh built
from the files matching the mesh built from the arrays."""
from __future__ import annotations

import numpy as np
import pytest

from gschism_amd import _lib as L
from gschism_amd import cases, core
from gschism_amd.grid_files import read_grid, write_hgrid, write_vgrid, write_vgrid_lsc2, write_vgrid_sz

HAND = """tiny grid
2 4
1 0.0 0.0 5.0
2 100.0 0.0 6.5
3 100.0 100.0 7.0   ! trailing comment
4 0.0 100.0 8.0
1 3 1 2 3
2 3 1 3 4
1 = Number of open boundaries
2 = Total number of open boundary nodes
2 = Number of nodes for open boundary 1
1
4
1 = Number of land boundaries
3 = Total number of land boundary nodes
3 0 = Number of nodes for land boundary 1
2
3
4
"""


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_hand_written_hgrid(tmp_path):
    g = read_grid(_write(tmp_path, "hgrid.gr3", HAND))
    assert g.name == "tiny grid"
    assert np.array_equal(g.xy, [[0, 0], [100, 0], [100, 100], [0, 100]])
    assert np.array_equal(g.depth, [5.0, 6.5, 7.0, 8.0])
    assert np.array_equal(g.elnode, [[0, 1, 2], [0, 2, 3]])
    assert np.array_equal(g.open_bnd, [1, 0, 0, 1])
    assert [s.tolist() for s in g.open_segments] == [[0, 3]]
    assert [(s.tolist(), isl) for s, isl in g.land_segments] == [([1, 2, 3], 0)]
    assert g.kb is None and g.sigma is None


def _boundary_segments(xy, en, ob):
    """Open segment = the open nodes in order along x = min; one land segment = the other
    boundary nodes (any order: the reader only counts and stores them)."""
    m = core.Mesh(None, xy, en, np.full(xy.shape[0], 10.0), cases.sigma_levels(3), open_bnd=ob, renumber=False)
    st = m.topology(L.T_ISDEL)
    sn = m.topology(L.T_ISIDENODE)
    bnodes = np.unique(sn[st[:, 1] < 0].ravel())
    opens = np.flatnonzero(ob)
    opens = opens[np.argsort(xy[opens, 1])]
    land = np.setdiff1d(bnodes, opens)
    return [opens], [(land, 0)]


def test_round_trip_and_mesh(tmp_path):
    xy, en, ob = cases.channel(30, 20, 250.0)
    depth = cases.bathymetry(30, 20, 5.0, 5.0, 50.0, seed=3)
    opens, lands = _boundary_segments(xy, en, ob)
    path = str(tmp_path / "hgrid.gr3")
    write_hgrid(path, xy, depth, en, opens, lands, name="channel 30x20")
    nv = 7
    sigma = cases.sigma_levels(nv)
    kb = (np.arange(xy.shape[0]) % 3).astype(np.int32)
    vpath = str(tmp_path / "vgrid.in")
    write_vgrid(vpath, sigma, kb)
    g = read_grid(path, vpath)
    assert np.array_equal(g.xy, xy) and np.array_equal(g.depth, depth)
    assert np.array_equal(g.elnode, en) and np.array_equal(g.open_bnd, ob)
    assert np.array_equal(g.kb, kb) and np.array_equal(g.sigma, sigma)
    assert np.array_equal(g.open_segments[0], opens[0])
    a = core.Mesh(None, xy, en, depth, sigma, open_bnd=ob, kb=kb)
    b = core.Mesh(None, g.xy, g.elnode, g.depth, g.sigma, open_bnd=g.open_bnd, kb=g.kb)
    for t in (L.T_XY, L.T_ELNODE, L.T_ELSIDE, L.T_CSR_INDPTR, L.T_CSR_INDICES, L.T_OPEN_BND, L.T_KB):
        assert np.array_equal(a.topology(t), b.topology(t)), t


def _expect_io_error(tmp_path, text, needle, vtext=None):
    h = _write(tmp_path, "h.gr3", text)
    v = None if vtext is None else _write(tmp_path, "v.in", vtext)
    with pytest.raises(L.GshError) as ei:
        read_grid(h, v)
    assert ei.value.code == L.GSH_E_IO and needle in str(ei.value), str(ei.value)


def test_reference_checks_are_errors(tmp_path):
    _expect_io_error(tmp_path, HAND.replace("2 = Total number of open", "3 = Total number of open"), "neta (3)")
    _expect_io_error(tmp_path, HAND.replace("3 = Total number of land", "4 = Total number of land"), "nvel (4)")
    looped = HAND.replace("2 = Number of nodes for open boundary 1\n1\n4\n",
                          "2 = Number of nodes for open boundary 1\n1\n1\n")
    _expect_io_error(tmp_path, looped, "looped open boundary")
    _expect_io_error(tmp_path, HAND.replace("2 3 1 3 4", "2 4 1 3 4 2"), "quadrilateral")
    _expect_io_error(tmp_path, HAND.replace("2 3 1 3 4", "2 3 1 3 9"), "node id 9 out of range")
    _expect_io_error(tmp_path, "name\n2 4\n1 0 0 5\n", "end of file in the node table")
    with pytest.raises(L.GshError) as ei:
        read_grid(str(tmp_path / "missing.gr3"))
    assert ei.value.code == L.GSH_E_IO


def test_vgrid_checks(tmp_path):
    base = "1 !ivcor\n3 !nvrt\n1 1 2 1\n1 -1 -1 -9 -1\n2 -0.5 -0.5 -1 -0.5\n3 0 0 0 0\n"
    h = _write(tmp_path, "h.gr3", HAND)
    g = read_grid(h, _write(tmp_path, "v.in", base))
    assert g.kb.tolist() == [0, 0, 1, 0] and g.sigma.tolist() == [-1.0, -0.5, 0.0] and g.vgrid is None
    # nodes that differ in sigma above their beds: a per-node (LSC2) grid
    lsc = base.replace("2 -0.5 -0.5 -1 -0.5", "2 -0.5 -0.4 -1 -0.5")
    g = read_grid(h, _write(tmp_path, "v2.in", lsc))
    assert g.sigma is None and g.vgrid.ivcor == 1 and g.kb.tolist() == [0, 0, 1, 0]
    assert g.vgrid.sigma_node[:, 1].tolist() == [-0.5, -0.4, -1.0, -0.5]
    # every node's bed above level 1 is accepted, as readVGrid accepts it
    g = read_grid(h, _write(tmp_path, "v3.in", base.replace("1 1 2 1", "2 2 2 2")))
    assert g.kb.tolist() == [1, 1, 1, 1]
    _expect_io_error(tmp_path, HAND, "nvrt kz h_s", "2 !ivcor\n")
    _expect_io_error(tmp_path, HAND, "kbp 3", base.replace("1 1 2 1", "1 1 3 1"))
    _expect_io_error(tmp_path, HAND, "level rows out of order", base.replace("2 -0.5", "5 -0.5"))
    _expect_io_error(tmp_path, HAND, "ivcor must be 1 or 2", "3 !ivcor\n")


SZ_RUN = """2 !ivcor
5 1 1.e6 !nvrt, kz (# of Z-levels); h_s (transition depth between S and Z)
Z levels
1 -1.e6
S levels
30. 0.7 5. !h_c, theta_b, theta_f
   1    -1.000000
   2    -0.700000
   3    -0.400000
   4    -0.150000
   5     0.000000
"""

SZ_KZ3 = """2 !ivcor
6 3 20.0
Z levels
1 -60.0
2 -35.0
3 -20.0
S levels
8.0 0.2 3.0
1 -1.0
2 -0.6
3 -0.3
4 0.0
"""


def test_vgrid_sz(tmp_path):
    h = _write(tmp_path, "h.gr3", HAND)
    g = read_grid(h, _write(tmp_path, "v.in", SZ_RUN))
    v = g.vgrid
    assert g.kb is None and g.sigma is None and v.ivcor == 2 and v.nz == 1 and v.nvrt == 5
    assert (v.h_s, v.h_c, v.theta_b, v.theta_f) == (1e6, 30.0, 0.7, 5.0)
    assert v.ztot.tolist() == [-1e6] and v.sigma_s.tolist() == [-1.0, -0.7, -0.4, -0.15, 0.0]
    g = read_grid(h, _write(tmp_path, "v3.in", SZ_KZ3))
    v = g.vgrid
    assert v.nz == 3 and v.ztot.tolist() == [-60.0, -35.0, -20.0] and v.sigma_s.tolist() == [-1.0, -0.6, -0.3, 0.0]
    # readVGrid's checks
    _expect_io_error(tmp_path, HAND, "h_s needs to be larger", SZ_KZ3.replace("6 3 20.0", "6 3 5.0"))
    _expect_io_error(tmp_path, HAND, "illegal Z level 2", SZ_KZ3.replace("2 -35.0", "2 -10.0"))
    _expect_io_error(tmp_path, HAND, "Z levels inverted at 2", SZ_KZ3.replace("2 -35.0", "2 -70.0"))
    _expect_io_error(tmp_path, HAND, "h_c needs to be larger", SZ_KZ3.replace("8.0 0.2 3.0", "4.0 0.2 3.0"))
    _expect_io_error(tmp_path, HAND, "wrong theta_b", SZ_KZ3.replace("8.0 0.2 3.0", "8.0 1.2 3.0"))
    _expect_io_error(tmp_path, HAND, "wrong theta_f", SZ_KZ3.replace("8.0 0.2 3.0", "8.0 0.2 0.0"))
    _expect_io_error(tmp_path, HAND, "invalid sigma levels at 3", SZ_KZ3.replace("3 -0.3", "3 -0.7"))
    _expect_io_error(tmp_path, HAND, "invalid level number in the S levels", SZ_KZ3.replace("3 -0.3", "4 -0.3"))


def test_vgrid_round_trips_and_meshes(tmp_path):
    xy, en, ob = cases.channel(24, 16, 300.0)
    n = xy.shape[0]
    depth = 4.0 + 50.0 * (xy[:, 0] - xy[:, 0].min()) / np.ptp(xy[:, 0])
    opens, lands = _boundary_segments(xy, en, ob)
    path = str(tmp_path / "hgrid.gr3")
    write_hgrid(path, xy, depth, en, opens, lands)
    # per-node sigma
    nv = 8
    kb = (np.arange(n) % 3).astype(np.int32)
    expo = np.linspace(0.8, 1.6, n)
    sn = -1.0 + np.linspace(0.0, 1.0, nv)[None, :] ** expo[:, None]
    write_vgrid_lsc2(str(tmp_path / "vl.in"), sn, kb)
    g = read_grid(path, str(tmp_path / "vl.in"))
    assert np.array_equal(g.kb, kb)
    above = np.arange(nv)[None, :] > kb[:, None]
    assert np.array_equal(g.vgrid.sigma_node[above], sn[above])
    a = g.mesh(None)
    b = core.Mesh(None, xy, en, depth, None, open_bnd=ob, kb=kb, vgrid=core.VGrid(1, sigma_node=sn))
    assert np.array_equal(a.topology(L.T_KB), b.topology(L.T_KB))
    # SZ
    ss = np.array([-1.0, -0.8, -0.55, -0.3, -0.1, 0.0])
    write_vgrid_sz(str(tmp_path / "vs.in"), 8, [-80.0, -45.0, -25.0], 25.0, 6.0, 0.4, 2.5, ss)
    g = read_grid(path, str(tmp_path / "vs.in"))
    v = g.vgrid
    assert v.nvrt == 8 and v.nz == 3 and np.array_equal(v.sigma_s, ss) and v.ztot.tolist() == [-80.0, -45.0, -25.0]
    m = g.mesh(None, renumber=False)
    kbm = m.topology(L.T_KB)
    assert ((depth <= 25.0) == (kbm == 2)).all() and ((depth > 45.0) == (kbm == 0)).all()
