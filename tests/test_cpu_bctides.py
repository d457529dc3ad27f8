"""bctides.in reader (read_bctides, readTides.py:52-293): a hand-written file with known contents,
the reference reader's checks as errors, the per-node elevation forcing of several boundaries,
and the oracle's elevation boundary conditions of every type against their closed forms."""
from __future__ import annotations

import math

import numpy as np
import pytest

from gschism_amd import _lib as L
from gschism_amd import cases
from gschism_amd.bctides import elevation_forcing, read_bctides
from oracle import pyoracle as O

HAND = """01/01/2020 00:00:00 ! start time
2 40. !ntip, tip_dp
M2
2 0.242334 1.405189D-04 1.0 12.5
K1
1 0.141565 7.292117D-05 1.0 30.0
2 !nbfr
M2
1.405189D-04 0.97 90.0
S2
1.454441D-04 1.0 0.0
3 !nope
3 3 0 0 0 !boundary 1: tidal elevation
M2
0.50 10.0
0.55 11.0
0.60 12.0
S2
0.20 40.0
0.21 41.0
0.22 42.0
2 2 2 1 0 !boundary 2: constant elevation, constant flow, T nudged
0.25
-150.0
0.5
2 5 0 2 3 !boundary 3: elevation time series + tides, T given, S nudged
M2
0.30 20.0
0.31 21.0
S2
0.10 50.0
0.11 51.0
12.5
1.0
0.8
"""


def _write(tmp_path, text):
    p = tmp_path / "bctides.in"
    p.write_text(text)
    return str(p)


def test_hand_written_bctides(tmp_path):
    t = read_bctides(_write(tmp_path, HAND), 3, 3)
    assert t.ntip == 2 and t.tip_dp == 40.0 and t.tp_name == ["M2", "K1"]
    assert t.jspc.tolist() == [2, 1] and np.allclose(t.tfreq, [1.405189e-4, 7.292117e-5])
    assert np.allclose(t.tear, np.radians([12.5, 30.0]))
    assert t.nbfr == 2 and np.allclose(t.amig, [1.405189e-4, 1.454441e-4])
    assert np.allclose(t.ff, [0.97, 1.0]) and np.allclose(t.face, np.radians([90.0, 0.0]))
    assert t.iettype.tolist() == [3, 2, 5] and t.ifltype.tolist() == [0, 2, 0]
    assert t.eth.tolist() == [0.0, 0.25, 0.0] and t.qthcon[1] == -150.0
    assert np.allclose(t.emo[0, :, 0], [0.5, 0.55, 0.6]) and np.allclose(t.emo[0, :, 1], [0.2, 0.21, 0.22])
    assert np.allclose(t.efa[0, :, 1], np.radians([40.0, 41.0, 42.0]))
    assert np.allclose(t.emo[2, :2, 0], [0.3, 0.31]) and np.allclose(t.efa[2, :2, 1], np.radians([50.0, 51.0]))
    assert t.itrtype[:2].T.tolist() == [[0, 0], [1, 0], [2, 3]]
    assert t.trobc[0, 1] == 0.5 and t.trobc[0, 2] == 1.0 and t.trobc[1, 2] == 0.8
    assert t.trth[0, 0, 0, 2] == 12.5
    assert (t.nettype, t.nettype2, t.nnode_et) == (0, 1, 2) and t.ntrtype1[0] == 1


def test_reference_checks_are_errors(tmp_path):
    with pytest.raises(ValueError, match="Inconsistent # of open bnds"):
        read_bctides(_write(tmp_path, HAND), 2, 3)
    with pytest.raises(ValueError, match="Illegal tidal species"):
        read_bctides(_write(tmp_path, HAND.replace("2 0.242334", "3 0.242334")), 3, 3)
    with pytest.raises(ValueError, match="Invalid iettype"):
        read_bctides(_write(tmp_path, HAND.replace("2 2 2 1 0 !boundary 2", "2 7 2 1 0 !boundary 2")), 3, 3)
    with pytest.raises(ValueError, match="nudging factor"):
        read_bctides(_write(tmp_path, HAND.replace("-150.0\n0.5\n", "-150.0\n1.5\n")), 3, 3)
    with pytest.raises(NotImplementedError):
        read_bctides(_write(tmp_path, HAND.replace("2 2 2 1 0 !boundary 2", "2 0 -1 1 0 !boundary 2")), 3, 3)
    with pytest.raises(ValueError, match="end of file"):
        read_bctides(_write(tmp_path, HAND[:HAND.index("0.8\n")]), 3, 3)


def test_elevation_forcing_last_boundary_wins(tmp_path):
    t = read_bctides(_write(tmp_path, HAND), 3, 3)
    t.ifltype[:] = 0  # flow forcing is refused (test_flow_forcing_is_refused)
    segs = [np.array([0, 1, 2]), np.array([2, 3]), np.array([3, 4])]
    etype, eth, emo, efa = elevation_forcing(t, segs, 6)
    assert etype.tolist() == [3, 3, 2, 5, 5, 0]
    assert eth.tolist() == [0.0, 0.0, 0.25, 0.0, 0.0, 0.0]
    assert np.allclose(emo[2], 0.0) and np.allclose(emo[3], [0.3, 0.1]) and np.allclose(emo[0], [0.5, 0.2])
    assert np.allclose(efa[4], np.radians([21.0, 51.0]))


def _tidal_mesh():
    xy, en, ob = cases.channel(20, 12, 300.0)
    m = O.OracleMesh(xy, en, np.full(xy.shape[0], 10.0), ob, cases.sigma_levels(4))
    return m, np.flatnonzero(ob)


def test_oracle_elevation_bc_of_every_type():
    """elbc of types 1-5 at time t against ramp eth (1, 2, 4), ramp sum ff emo cos(amig t + face -
    efa) with the angle not reduced (3), their sum (5); type 0 leaves elbc as it was."""
    m, opn = _tidal_mesh()
    n = m.n
    rng = np.random.default_rng(2)
    types = np.array([0, 1, 2, 3, 4, 5])[np.arange(opn.size) % 6]
    etype = np.full(n, -1, dtype=np.int8)
    etype[opn] = types
    m.set(L.F_ETYPE, etype)
    eth = rng.normal(0.0, 0.5, n)
    m.set(L.F_ETH, eth)
    amig = np.array([1.405189e-4, 1.454441e-4, 7.292117e-5])
    ff, face = np.array([0.97, 1.0, 1.1]), np.array([1.2, 0.0, -0.4])
    emo, efa = rng.uniform(0.1, 1.0, (n, 3)), rng.uniform(0.0, 2 * math.pi, (n, 3))
    m.set_tides(amig, ff, face, emo, efa)
    m.set(L.F_ELBC, np.full(n, 7.0))
    m.set(L.F_ETA2, np.zeros(n))
    p = L.default_step_params(time=3.7e6, ramp=0.8)
    m.levels(p)
    m.rhs(p)
    got = m.get(L.F_ELBC)
    harm = p.ramp * (ff * emo * np.cos(amig * p.time + face - efa)).sum(axis=1)
    for i, t in zip(opn, types):
        exp = {0: 7.0, 1: p.ramp * eth[i], 2: p.ramp * eth[i], 3: harm[i], 4: p.ramp * eth[i],
               5: harm[i] + p.ramp * eth[i]}[t]
        assert abs(got[i] - exp) <= 1e-12, (i, t, got[i], exp)
    assert np.all(got[np.setdiff1d(np.arange(n), opn)] == 7.0)


def test_oracle_essential_rows_follow_node_types():
    """Open nodes of type 0 keep their assembled rows and their neighbours' columns; the others
    become identity rows with elbc on the right and are eliminated from their neighbours."""
    m, opn = _tidal_mesh()
    n = m.n
    p = L.default_step_params()
    m.set(L.F_ETA2, 0.1 * np.sin(m.topology(L.T_XY)[:, 0] / 900.0))
    m.levels(p)
    m.set(L.F_ELBC, np.zeros(n))
    m.rhs(p)
    ref_val, ref_b = m.assemble(L.default_step_params(iettype=0))
    etype = np.full(n, -1, dtype=np.int8)
    free = opn[::2]
    etype[opn] = 2
    etype[free] = 0
    m.set(L.F_ETYPE, etype)
    val, b = m.assemble(p)
    ip, ix = m.topology(L.T_CSR_INDPTR), m.topology(L.T_CSR_INDICES)
    ess = np.zeros(n, bool)
    ess[opn] = True
    ess[free] = False
    elbc = m.get(L.F_ELBC)
    for i in range(n):
        lo, hi = ip[i], ip[i + 1]
        if ess[i]:
            assert np.all(val[lo:hi][ix[lo:hi] != i] == 0.0) and val[lo:hi][ix[lo:hi] == i] == 1.0
            assert b[i] == elbc[i]
        else:
            keep = ~ess[ix[lo:hi]]
            assert np.array_equal(val[lo:hi][keep], ref_val[lo:hi][keep])
            assert np.all(val[lo:hi][~keep] == 0.0)
            exp = ref_b[i] - np.sum(ref_val[lo:hi][~keep] * elbc[ix[lo:hi][~keep]])
            assert abs(b[i] - exp) <= 1e-12 * max(1.0, abs(exp))


def test_flow_forcing_is_refused(tmp_path):
    """Flow types (ifltype != 0) are parsed but not applied: the per-node forcing refuses them
    instead of running as if the boundary had none."""
    t = read_bctides(_write(tmp_path, HAND), 3, 3)
    segs = [np.array([0, 1, 2]), np.array([2, 3]), np.array([3, 4])]
    assert t.ifltype.tolist() == [0, 2, 0]
    with pytest.raises(NotImplementedError, match="flow forcing"):
        elevation_forcing(t, segs, 6)
    t.ifltype[1] = 0
    etype, _, _, _ = elevation_forcing(t, segs, 6)
    assert etype.tolist() == [3, 3, 2, 5, 5, 0]

