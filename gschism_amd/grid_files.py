"""SCHISM grid files: hgrid.gr3 / vgrid.in read through the library (gsh_grid_*), and plain
writers of the same formats for synthetic meshes.

The reader follows readHGrid and readVGrid (read_and_process_grid.cpp:855-1070, 1313-1520);
`GridFiles` holds exactly the arrays `core.Mesh` takes (`GridFiles.mesh` builds it).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import check


@dataclass
class GridFiles:
    name: str
    xy: np.ndarray            # [n, 2]
    depth: np.ndarray         # [n]
    elnode: np.ndarray        # [ne, 3], 0-based
    open_bnd: np.ndarray      # [n] int8, 1 on open boundary segments
    open_segments: list = field(default_factory=list)   # 0-based node ids per open segment
    land_segments: list = field(default_factory=list)   # (0-based node ids, island flag) per land segment
    kb: np.ndarray | None = None      # [n] bottom level (0-based), from vgrid.in (ivcor 1)
    sigma: np.ndarray | None = None   # [nvrt], from vgrid.in (ivcor 1, one sigma column)
    vgrid: object = None              # core.VGrid: per-node sigma (ivcor 1) or SZ (ivcor 2)

    def mesh(self, ctx, renumber: bool = True):
        """The core.Mesh of these arrays (ctx None: host only)."""
        from .core import Mesh

        return Mesh(ctx, self.xy, self.elnode, self.depth, self.sigma, open_bnd=self.open_bnd, kb=self.kb,
                    renumber=renumber, vgrid=self.vgrid)


def read_grid(hgrid: str, vgrid: str | None = None) -> GridFiles:
    """Reads hgrid.gr3 (and vgrid.in, ivcor 1 or 2, when given)."""
    h = C.c_void_p()
    check(L.lib().gsh_grid_read(str(hgrid).encode(), None if vgrid is None else str(vgrid).encode(), C.byref(h)))
    try:
        n, ne, nv, no, nl = (C.c_int32() for _ in range(5))
        check(L.lib().gsh_grid_info(h, C.byref(n), C.byref(ne), C.byref(nv), C.byref(no), C.byref(nl)))
        n, ne, nv = n.value, ne.value, nv.value
        xy = np.empty((n, 2))
        depth = np.empty(n)
        elnode = np.empty((ne, 3), dtype=np.int32)
        ob = np.empty(n, dtype=np.int8)
        pd = C.POINTER(C.c_double)
        pi = C.POINTER(C.c_int32)
        ivcor, per_node, nz = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        vg = None
        if nv:
            check(L.lib().gsh_grid_vgrid(h, C.byref(ivcor), C.byref(per_node), C.byref(nz), None, None, None, None))
        kb = np.empty(n, dtype=np.int32) if nv and ivcor.value == 1 else None
        sigma = np.empty(nv) if nv and ivcor.value == 1 and not per_node.value else None
        check(L.lib().gsh_grid_get(h, xy.ctypes.data_as(pd), depth.ctypes.data_as(pd), elnode.ctypes.data_as(pi),
                                   ob.ctypes.data_as(C.POINTER(C.c_int8)),
                                   kb.ctypes.data_as(pi) if kb is not None else None,
                                   sigma.ctypes.data_as(pd) if sigma is not None else None))
        if nv:
            from .core import VGrid

            if ivcor.value == 2:
                prm = np.empty(4)
                zt = np.empty(nz.value)
                ss = np.empty(nv - nz.value + 1)
                check(L.lib().gsh_grid_vgrid(h, None, None, None, prm.ctypes.data_as(pd), zt.ctypes.data_as(pd),
                                             ss.ctypes.data_as(pd), None))
                vg = VGrid(2, nz=nz.value, h_s=prm[0], h_c=prm[1], theta_b=prm[2], theta_f=prm[3], ztot=zt,
                           sigma_s=ss)
            elif per_node.value:
                sn = np.empty((n, nv))
                check(L.lib().gsh_grid_vgrid(h, None, None, None, None, None, None, sn.ctypes.data_as(pd)))
                vg = VGrid(1, sigma_node=sn)

        def segment(land, s):
            cnt, isl = C.c_int32(), C.c_int32()
            check(L.lib().gsh_grid_segment(h, land, s, None, 0, C.byref(cnt), C.byref(isl)))
            out = np.empty(cnt.value, dtype=np.int32)
            check(L.lib().gsh_grid_segment(h, land, s, out.ctypes.data_as(pi), cnt.value, C.byref(cnt), C.byref(isl)))
            return out, isl.value

        opens = [segment(0, s)[0] for s in range(no.value)]
        lands = [segment(1, s) for s in range(nl.value)]
        with open(hgrid, "r", errors="replace") as f:
            name = f.readline().rstrip("\r\n")
        return GridFiles(name, xy, depth, elnode, ob, opens, lands, kb, sigma, vg)
    finally:
        L.lib().gsh_grid_destroy(h)


def write_hgrid(path: str, xy, depth, elnode, open_segments=(), land_segments=(), name: str = "hgrid") -> None:
    """hgrid.gr3 with 1-based ids; land_segments holds (nodes, island flag) pairs."""
    xy = np.asarray(xy, dtype=np.float64)
    depth = np.asarray(depth, dtype=np.float64)
    en = np.asarray(elnode, dtype=np.int64)
    n, ne = xy.shape[0], en.shape[0]
    with open(path, "w") as f:
        f.write(f"{name}\n{ne} {n}\n")
        ids = np.arange(1, n + 1)
        f.writelines(f"{i} {x!r} {y!r} {d!r}\n" for i, x, y, d in zip(ids, xy[:, 0].tolist(), xy[:, 1].tolist(),
                                                                      depth.tolist()))
        f.writelines(f"{k + 1} 3 {a + 1} {b + 1} {c + 1}\n" for k, (a, b, c) in enumerate(en.tolist()))
        opens = [np.asarray(s) for s in open_segments]
        f.write(f"{len(opens)} = Number of open boundaries\n")
        f.write(f"{sum(len(s) for s in opens)} = Total number of open boundary nodes\n")
        for k, s in enumerate(opens):
            f.write(f"{len(s)} = Number of nodes for open boundary {k + 1}\n")
            f.writelines(f"{i + 1}\n" for i in s.tolist())
        lands = [(np.asarray(s), int(isl)) for s, isl in land_segments]
        f.write(f"{len(lands)} = Number of land boundaries\n")
        f.write(f"{sum(len(s) for s, _ in lands)} = Total number of land boundary nodes\n")
        for k, (s, isl) in enumerate(lands):
            f.write(f"{len(s)} {isl} = Number of nodes for land boundary {k + 1}\n")
            f.writelines(f"{i + 1}\n" for i in s.tolist())


def write_vgrid(path: str, sigma, kb) -> None:
    """vgrid.in, ivcor=1: the per-node sigma rows repeat the one sigma per level (-9 below kb)."""
    sigma = np.asarray(sigma, dtype=np.float64)
    kb = np.asarray(kb, dtype=np.int64)
    nv, n = sigma.size, kb.size
    with open(path, "w") as f:
        f.write("1 !ivcor\n")
        f.write(f"{nv} !nvrt\n")
        f.write(" ".join(str(k + 1) for k in kb.tolist()) + "\n")
        for k in range(nv):
            vals = np.where(kb <= k, sigma[k], -9.0)
            f.write(f"{k + 1} " + " ".join(repr(v) for v in vals.tolist()) + "\n")


def write_vgrid_lsc2(path: str, sigma_node, kb) -> None:
    """vgrid.in, ivcor=1 with a sigma column per node ([n][nvrt]; -9 below each node's kb)."""
    sn = np.asarray(sigma_node, dtype=np.float64)
    kb = np.asarray(kb, dtype=np.int64)
    n, nv = sn.shape
    with open(path, "w") as f:
        f.write("1 !ivcor\n")
        f.write(f"{nv} !nvrt\n")
        f.write(" ".join(str(k + 1) for k in kb.tolist()) + "\n")
        for k in range(nv):
            vals = np.where(kb <= k, sn[:, k], -9.0)
            f.write(f"{k + 1} " + " ".join(repr(v) for v in vals.tolist()) + "\n")


def write_vgrid_sz(path: str, nvrt: int, ztot, h_s: float, h_c: float, theta_b: float, theta_f: float,
                   sigma_s) -> None:
    """vgrid.in, ivcor=2 (SZ): kz = len(ztot) Z levels (the last is -h_s), then the S levels."""
    ztot = np.asarray(ztot, dtype=np.float64)
    ss = np.asarray(sigma_s, dtype=np.float64)
    kz = ztot.size
    assert ss.size == nvrt - kz + 1
    with open(path, "w") as f:
        f.write("2 !ivcor\n")
        f.write(f"{nvrt} {kz} {float(h_s)!r} !nvrt, kz (# of Z-levels); h_s (transition depth between S and Z)\n")
        f.write("Z levels\n")
        for k in range(kz):
            f.write(f"{k + 1} {float(ztot[k])!r}\n")
        f.write("S levels\n")
        f.write(f"{float(h_c)!r} {float(theta_b)!r} {float(theta_f)!r} !h_c, theta_b, theta_f\n")
        for k in range(ss.size):
            f.write(f"{k + 1} {float(ss[k])!r}\n")
