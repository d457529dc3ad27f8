"""hgrid.gr3 / vgrid.in readers (gsh_grid_*): a hand-written file with known contents, round trips
of synthetic meshes through the writers, the reference reader's checks as errors, and a mesh built
from the files matching the mesh built from the arrays."""
from __future__ import annotations

import numpy as np
import pytest

from gschism_amd import _lib as L
from gschism_amd import cases, core
from gschism_amd.grid_files import read_grid, write_hgrid, write_vgrid

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
    assert g.kb.tolist() == [0, 0, 1, 0] and g.sigma.tolist() == [-1.0, -0.5, 0.0]
    _expect_io_error(tmp_path, HAND, "per-node sigma", base.replace("2 -0.5 -0.5 -1 -0.5", "2 -0.5 -0.4 -1 -0.5"))
    _expect_io_error(tmp_path, HAND, "ivcor=2", "2 !ivcor\n")
    _expect_io_error(tmp_path, HAND, "kbp 3", base.replace("1 1 2 1", "1 1 3 1"))
    _expect_io_error(tmp_path, HAND, "no node uses the bottom level", base.replace("1 1 2 1", "2 2 2 2"))
    _expect_io_error(tmp_path, HAND, "level rows out of order", base.replace("2 -0.5", "5 -0.5"))
