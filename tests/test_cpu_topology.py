"""Topology and geometry: the library's host builder against the oracle's restatement of the
reference grid code, and both against invariants and brute force."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from gschism_amd import _lib as L
from gschism_amd import cases, core
from oracle import pyoracle as O


def _mesh_pair(nx, ny, renumber=True, jitter=0.2, seed=0):
    xy, en, ob = cases.channel(nx, ny, 250.0, jitter, seed)
    depth = cases.bathymetry(nx, ny, 6.0, 5.0, 400.0, seed=3)
    sig = cases.sigma_levels(5)
    m = core.Mesh(None, xy, en, depth, sig, open_bnd=ob, renumber=renumber)
    om = O.OracleMesh(m.topology(L.T_XY), m.topology(L.T_ELNODE), m.topology(L.T_DEPTH),
                      m.topology(L.T_OPEN_BND), sig)
    return m, om


@pytest.mark.parametrize("nx,ny,renumber,jitter", [(12, 9, False, 0.2), (40, 25, True, 0.2), (31, 17, True, 0.0),
                                                   (23, 40, False, 0.3)])
def test_library_topology_equals_oracle(nx, ny, renumber, jitter):
    m, om = _mesh_pair(nx, ny, renumber, jitter)
    assert (m.ns, m.nnz) == (om.ns, om.nnz)
    for code in range(15):
        assert np.array_equal(m.topology(code).reshape(-1), om.topology(code).reshape(-1)), f"array {code}"


def test_csr_pattern_invariants_and_brute_force():
    m, _ = _mesh_pair(35, 22)
    n = m.n
    ip, ix = m.topology(L.T_CSR_INDPTR), m.topology(L.T_CSR_INDICES)
    en = m.topology(L.T_ELNODE)
    nbr = m.topology(L.T_NBR_NODE)
    slot = m.topology(L.T_CSR_SLOT)
    # nnz = sum(numNodes + 1) (main.py:262-279) and indptr[i] = nbrStart[i] + i
    assert ip[-1] == nbr.size + n
    for i in range(n):
        row = ix[ip[i]:ip[i + 1]]
        assert np.all(np.diff(row) > 0), "columns ascending"
        assert i in row, "diagonal present"
        nb = nbr[ip[i] - i:ip[i + 1] - i - 1]
        assert np.array_equal(row[slot[ip[i] - i:ip[i + 1] - i - 1]], nb), "csrSlot maps ball order to CSR slots"
    # brute force: the pattern is node-element incidence squared
    inc = sp.csr_matrix((np.ones(en.size), (np.repeat(np.arange(en.shape[0]), 3), en.reshape(-1))),
                        shape=(en.shape[0], n))
    pat = (inc.T @ inc).tocsr()
    pat.sort_indices()
    assert np.array_equal(pat.indptr, ip) and np.array_equal(pat.indices, ix)
    A = sp.csr_matrix((np.ones(ix.size), ix, ip), shape=(n, n))
    assert (A != A.T).nnz == 0, "symmetric pattern"


def test_balls_are_counter_clockwise_fans():
    m, _ = _mesh_pair(20, 14)
    en = m.topology(L.T_ELNODE)
    bp, be = m.topology(L.T_BALL_PTR), m.topology(L.T_BALL_ELEM)
    isdel = m.topology(L.T_ISDEL)
    bnd_nodes = set(m.topology(L.T_ISIDENODE)[isdel[:, 1] < 0].reshape(-1).tolist())
    for i in range(m.n):
        ball = be[bp[i]:bp[i + 1]]
        for a, b in zip(ball[:-1], ball[1:]):
            # consecutive elements share the edge (i, j) with j the CCW-next node of i in a
            la = list(en[a]).index(i)
            j = en[a][(la + 2) % 3]
            assert j in en[b], "consecutive ball elements share an edge"
        if i not in bnd_nodes:
            last, first = ball[-1], ball[0]
            la = list(en[last]).index(i)
            assert en[last][(la + 2) % 3] in en[first], "interior ball closes"


def test_sides_orientation_and_counts():
    m, _ = _mesh_pair(25, 18)
    isdel, sn = m.topology(L.T_ISDEL), m.topology(L.T_ISIDENODE)
    en, es = m.topology(L.T_ELNODE), m.topology(L.T_ELSIDE)
    # Euler: ns = n + ne - 1 for a simply connected triangulation
    assert m.ns == m.n + m.ne - 1
    assert np.all((isdel[:, 1] < 0) | (isdel[:, 0] < isdel[:, 1])), "elems[0] has the lower index"
    for e in range(m.ne):
        for j in range(3):
            s = es[e, j]
            assert set(sn[s]) == {en[e, (j + 1) % 3], en[e, (j + 2) % 3]}, "side j is opposite node j"
            if isdel[s, 0] == e:
                assert sn[s, 0] == en[e, (j + 1) % 3], "n1 -> n2 follows elems[0]'s CCW order"
    st = m.topology(L.T_SIDE_TYPE)
    ob = m.topology(L.T_OPEN_BND)
    assert np.array_equal(st == 0, isdel[:, 1] >= 0)
    assert np.array_equal(st == 2, (isdel[:, 1] < 0) & (ob[sn[:, 0]] == 1) & (ob[sn[:, 1]] == 1))
