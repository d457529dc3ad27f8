"""bctides.in: the open-boundary forcing of a SCHISM run, read as read_bctides does
(readTides.py:52-293), and its elevation part handed to the library.

`read_bctides` parses every block the reference parses (tidal potential, boundary forcing
frequencies, per-boundary elevation / flow / tracer types and their values) with the same checks
and the same errors. `elevation_forcing` turns the elevation part into per-node arrays the way
Mesh::populateBoundaryNodes does (meshdata.cu:2485-2530: a node on several open boundaries takes
the values of the last one), and `apply_bctides` sets them on a mesh: the type of each open node
(GSH_F_ETYPE), the constant elevation of type-2 boundaries (GSH_F_ETH) and the tidal
constituents of type-3/5 boundaries (gsh_set_tides). Time series of types 1 and 4 (elev.th,
elev2D.th.nc) are not read: the caller sets GSH_F_ETH each step.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

NATRM = 12  # tracer modules the reference knows, T and S included (schism_consts.py:2)
MNTRACERS = 30  # schism_consts.py:3


@dataclass
class TideData:
    """The fields of readTides.TidalData that the file fills (arrays indexed like the reference's,
    phases and arguments converted to radians as it does)."""
    ntip: int = 0
    tip_dp: float = 0.0
    tp_name: list = field(default_factory=list)
    tamp: np.ndarray | None = None
    tnf: np.ndarray | None = None
    tfreq: np.ndarray | None = None
    jspc: np.ndarray | None = None
    tear: np.ndarray | None = None
    nbfr: int = 0
    amig: np.ndarray | None = None
    ff: np.ndarray | None = None
    face: np.ndarray | None = None
    nope: int = 0
    iettype: np.ndarray | None = None
    ifltype: np.ndarray | None = None
    eth: np.ndarray | None = None
    qthcon: np.ndarray | None = None
    vobc1: np.ndarray | None = None
    vobc2: np.ndarray | None = None
    disch_coef: tuple = (0.0, 0.0, 0.0, 0.0)
    emo: np.ndarray | None = None   # [nope][mnond][nbfr]
    efa: np.ndarray | None = None
    umo: np.ndarray | None = None
    ufa: np.ndarray | None = None
    vmo: np.ndarray | None = None
    vfa: np.ndarray | None = None
    trobc: np.ndarray | None = None
    itrtype: np.ndarray | None = None
    trth: np.ndarray | None = None
    nettype: int = 0
    nfltype: int = 0
    nettype2: int = 0
    nfltype2: int = 0
    nnode_et: int = 0
    nnode_fl: int = 0
    nnode_tr2: np.ndarray | None = None
    ntrtype1: np.ndarray | None = None


def _values(line: str, types) -> tuple:
    """read_values (readTides.py:45-49): Fortran D exponents, as many fields as types."""
    parts = [p.replace("D", "E") for p in line.split()]
    if len(parts) < len(types):
        raise ValueError(f"bctides.in: expected {len(types)} values in line {line.rstrip()!r}")
    return tuple(t(x) for t, x in zip(types, parts))


def read_bctides(path: str, nope_global: int, mnond_global: int, ntrs=None, irange_tr=None) -> TideData:
    """read_bctides (readTides.py:52-293) from the file `path`.

    nope_global / mnond_global: open boundaries of the grid and the most nodes on one of them.
    ntrs: per tracer module, > 0 when active (default T and S only, as main.py:291 passes);
    irange_tr: [first, last] tracer index of each module (default T = 0, S = 1), used for
    type-2 tracer boundaries.
    """
    if ntrs is None:
        ntrs = [1, 1] + [0] * (NATRM - 2)
    if irange_tr is None:
        irange_tr = [[0, 1] + [0] * (NATRM - 2), [0, 1] + [0] * (NATRM - 2)]
    tide = TideData()
    f64 = float
    with open(path, "r") as fh:
        def line() -> str:
            s = fh.readline()
            if not s:
                raise ValueError("bctides.in: unexpected end of file")
            return s

        line()  # start time
        tide.ntip, tide.tip_dp = _values(line(), [int, f64])
        tide.tamp = np.zeros(max(tide.ntip, 0))
        tide.tnf = np.zeros(max(tide.ntip, 0))
        tide.tfreq = np.zeros(max(tide.ntip, 0))
        tide.jspc = np.zeros(max(tide.ntip, 0), dtype=np.int32)
        tide.tear = np.zeros(max(tide.ntip, 0))
        for i in range(tide.ntip):
            tide.tp_name.append(line().strip())
            jspc, tamp, tfreq, tnf, tear = _values(line(), [int] + [f64] * 4)
            if jspc < 0 or jspc > 2:
                raise ValueError(f"Illegal tidal species # {jspc}")
            tide.jspc[i], tide.tamp[i], tide.tfreq[i], tide.tnf[i] = jspc, tamp, tfreq, tnf
            tide.tear[i] = tear * np.pi / 180.0

        tide.nbfr = _values(line(), [int])[0]
        nb = max(tide.nbfr, 0)
        tide.amig, tide.ff, tide.face = np.zeros(nb), np.zeros(nb), np.zeros(nb)
        shape = (nope_global, mnond_global, nb)
        tide.emo, tide.efa = np.zeros(shape), np.zeros(shape)
        tide.umo, tide.ufa, tide.vmo, tide.vfa = (np.zeros(shape) for _ in range(4))
        for i in range(tide.nbfr):
            line()  # constituent name
            amig, ff, face = _values(line(), [f64] * 3)
            tide.amig[i], tide.ff[i], tide.face[i] = amig, ff, face * np.pi / 180.0

        nope1 = _values(line(), [int])[0]
        if nope1 != nope_global:
            raise ValueError(f"Inconsistent # of open bnds {nope1} {nope_global}")
        tide.nope = nope1
        ntrmod = 2 + sum(1 for i in range(2, NATRM) if ntrs[i] > 0)
        tide.iettype = np.zeros(nope1, dtype=np.int32)
        tide.ifltype = np.zeros(nope1, dtype=np.int32)
        tide.eth = np.zeros(nope1)
        tide.qthcon = np.zeros(nope1)
        tide.vobc1 = np.zeros(nope1)
        tide.vobc2 = np.zeros(nope1)
        tide.trobc = np.zeros((NATRM, nope1))
        tide.itrtype = np.full((NATRM, nope1), -99, dtype=np.int32)
        tide.trth = np.zeros((MNTRACERS, 1, 1, nope1))
        tide.nnode_tr2 = np.zeros(NATRM, dtype=np.int32)
        tide.ntrtype1 = np.zeros(NATRM, dtype=np.int32)

        def harmonics(k, ntmp, four: bool):
            for i in range(tide.nbfr):
                line()  # constituent name
                for j in range(ntmp):
                    if four:
                        ua, up, va, vp = _values(line(), [f64] * 4)
                        tide.umo[k, j, i], tide.ufa[k, j, i] = ua, up * np.pi / 180.0
                        tide.vmo[k, j, i], tide.vfa[k, j, i] = va, vp * np.pi / 180.0
                    else:
                        amp, phase = _values(line(), [f64] * 2)
                        tide.emo[k, j, i], tide.efa[k, j, i] = amp, phase * np.pi / 180.0

        for k in range(nope1):
            vals = _values(line(), [int] * (3 + ntrmod))
            ntmp, iettype, ifltype, nwild = vals[0], vals[1], vals[2], vals[3:]
            if ntmp > mnond_global:
                raise ValueError(f"open boundary {k}: {ntmp} nodes, more than mnond = {mnond_global}")
            tide.iettype[k], tide.ifltype[k] = iettype, ifltype
            if iettype == 1:
                tide.nettype += 1
            elif iettype == 2:
                tide.eth[k] = _values(line(), [f64])[0]
            elif iettype == 3:
                harmonics(k, ntmp, False)
            elif iettype in (4, 5):
                tide.nettype2 += 1
                tide.nnode_et += ntmp
                if iettype == 5:
                    harmonics(k, ntmp, False)
            elif iettype != 0:
                raise ValueError("Invalid iettype")

            if ifltype == 1:
                tide.nfltype += 1
            elif ifltype == 2:
                tide.qthcon[k] = _values(line(), [f64])[0]
            elif ifltype == 3:
                harmonics(k, ntmp, True)
            elif abs(ifltype) in (4, 5):
                raise NotImplementedError(f"ifltype={ifltype} needs uv3D.th.nc (open bnd {k})")
            elif ifltype == -1:
                raise NotImplementedError(f"ifltype=-1 Flather needs uthnd (open bnd {k})")
            elif ifltype == -2:
                if iettype != 0:
                    raise ValueError(f"Discharge obc requires iettype=0: {k}")
                tide.disch_coef = _values(line(), [f64] * 4)
            elif ifltype != 0:
                raise ValueError(f"Invalid ifltype: {ifltype}")

            icount = 0
            for i in range(NATRM):
                if ntrs[i] <= 0:
                    continue
                icount += 1
                tide.itrtype[i, k] = nwild[icount - 1]
                if tide.itrtype[i, k] == 2:
                    lo, hi = irange_tr[0][i], irange_tr[1][i]
                    tide.trth[lo:hi + 1, 0, 0, k] = _values(line(), [f64] * (hi - lo + 1))
                    tide.trobc[i, k] = _values(line(), [f64])[0]
                elif tide.itrtype[i, k] in (1, 3, 4):
                    tide.trobc[i, k] = _values(line(), [f64])[0]
                    if tide.itrtype[i, k] == 1:
                        tide.ntrtype1[i] += 1
                    elif tide.itrtype[i, k] == 4:
                        tide.nnode_tr2[i] += ntmp
                elif tide.itrtype[i, k] != 0:
                    raise ValueError(f"Wrong itrtype: {k} {tide.itrtype[i, k]}")
                if not (0 <= tide.trobc[i, k] <= 1):
                    raise ValueError(f"Tr. obc nudging factor wrong: {tide.trobc[i, k]} {k}")
            if icount != ntrmod:
                raise ValueError("INIT:icount!=ntrmod")
    return tide


def elevation_forcing(tide: TideData, open_segments, n: int):
    """Per-node elevation BC arrays from the open boundaries (node ids 0-based, in the order of
    `open_segments`): type (int8 [n], 0 off the open boundary), constant elevation eth [n],
    amplitude emo and phase efa [n][nbfr]. A node on several boundaries takes the last one's
    values (Mesh::populateBoundaryNodes, meshdata.cu:2499-2527)."""
    if len(open_segments) != tide.nope:
        raise ValueError(f"{len(open_segments)} open boundaries in the grid, {tide.nope} in bctides.in")
    flow = [k for k in range(tide.nope) if tide.ifltype[k] != 0]
    if flow:
        # flow forcing needs the I_3 integral, the Flather / discharge matrix terms and the
        # momentum boundary velocities (meshdata.cu:3384-3454, 5357-5446), which are not built
        raise NotImplementedError(f"open boundaries {flow} have flow forcing (ifltype != 0), which is not supported")
    etype = np.zeros(n, dtype=np.int8)
    eth = np.zeros(n)
    emo = np.zeros((n, tide.nbfr))
    efa = np.zeros((n, tide.nbfr))
    for k, seg in enumerate(open_segments):
        seg = np.asarray(seg, dtype=np.int64)
        etype[seg] = tide.iettype[k]
        eth[seg] = tide.eth[k]
        emo[seg] = tide.emo[k, :seg.size]
        efa[seg] = tide.efa[k, :seg.size]
    return etype, eth, emo, efa


def apply_bctides(mesh, tide: TideData, open_segments) -> None:
    """Sets the elevation forcing of `tide` on a core.Mesh. `open_segments` holds file node ids
    (0-based, e.g. GridFiles.open_segments); they are mapped through the mesh's node order."""
    perm = np.asarray(mesh.node_perm)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    segs = [inv[np.asarray(s, dtype=np.int64)] for s in open_segments]
    from . import _lib as L
    etype, eth, emo, efa = elevation_forcing(tide, segs, mesh.n)
    mesh.set(L.F_ETYPE, etype)
    mesh.set(L.F_ETH, eth)
    mesh.set_tides(tide.amig, tide.ff, tide.face, emo, efa)
