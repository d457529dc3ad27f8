"""Python host API over libgsh: contexts, meshes, solvers and the per-step operations.

The objects own library handles; every method calls the C ABI (include/gsh.h). Arrays
cross the boundary as NumPy host arrays; device pointers are available for zero-copy use.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, dptr

import atexit
import weakref

# Library handles still open at interpreter exit are closed newest first (solvers and meshes
# before their context), before the HIP runtime tears itself down.
_live: list = []


def _register(obj) -> None:
    _live.append(weakref.ref(obj))


@atexit.register
def _close_all() -> None:
    while _live:
        obj = _live.pop()()
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass


class Context:
    """One GPU and one HIP stream (gsh_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(L.lib().gsh_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        _register(self)

    def close(self):
        if getattr(self, "h", None):
            L.lib().gsh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def mem_info(self):
        f, t = C.c_int64(), C.c_int64()
        check(L.lib().gsh_ctx_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def synchronize(self):
        check(L.lib().gsh_ctx_synchronize(self.h))


_FIELD_SPEC = {
    L.F_ETA2: (np.float64, "n", ()), L.F_ETA1: (np.float64, "n", ()),
    L.F_NODE_UVW: (np.float64, "n", ("nvrt", 3)), L.F_SIDE_U2: (np.float64, "ns", ("nvrt", 2)),
    L.F_BTU: (np.float64, "ns", ("nvrt", 2)), L.F_FOOT: (np.int32, "ns", ("nvrt",)),
    L.F_BSTATUS: (np.int32, "ns", ("nvrt",)), L.F_CD: (np.float64, "n", ()),
    L.F_GRAV2: (np.float64, "n", ()), L.F_ETH: (np.float64, "n", ()), L.F_VSOURCE: (np.float64, "ne", ()),
    L.F_SIDE_U: (np.float64, "ns", (2,)), L.F_SIDE_HHAT: (np.float64, "ns", ()),
    L.F_SIDE_CHI: (np.float64, "ns", ()), L.F_GHAT1: (np.float64, "ne", (2,)),
    L.F_ZC: (np.float64, "ns", ("nvrt",)), L.F_ELBC: (np.float64, "n", ()),
    L.F_DEPTH: (np.float64, "n", ()), L.F_ELEM_W: (np.float64, "ne", ("nvrt",)),
    L.F_ETYPE: (np.int8, "n", ()),
}


class VGrid:
    """Vertical grid beyond one sigma per level (gsh_vgrid; readVGrid,
    read_and_process_grid.cpp:1313-1520).

    ivcor=1: `sigma_node` [n][nvrt], a sigma per node and level (LSC2), in the node order given to
    Mesh; each node's bottom level comes from Mesh's kb.
    ivcor=2 (SZ): `nz` Z levels `ztot` (the last one at -h_s), S levels `sigma_s` (-1 .. 0,
    nvrt - nz + 1 of them) and h_s, h_c, theta_b, theta_f; every node's kb follows from its depth.
    """

    def __init__(self, ivcor: int, sigma_node=None, nz: int = 0, h_s: float = 0.0, h_c: float = 0.0,
                 theta_b: float = 0.0, theta_f: float = 0.0, ztot=None, sigma_s=None):
        self.ivcor = int(ivcor)
        self.sigma_node = None if sigma_node is None else np.ascontiguousarray(sigma_node, dtype=np.float64)
        self.nz = int(nz)
        self.h_s, self.h_c, self.theta_b, self.theta_f = float(h_s), float(h_c), float(theta_b), float(theta_f)
        self.ztot = None if ztot is None else np.ascontiguousarray(ztot, dtype=np.float64)
        self.sigma_s = None if sigma_s is None else np.ascontiguousarray(sigma_s, dtype=np.float64)

    @property
    def nvrt(self) -> int:
        if self.ivcor == 1:
            return int(self.sigma_node.shape[1])
        return self.nz - 1 + int(self.sigma_s.size)

    def permuted(self, node_perm) -> "VGrid":
        """The same grid in the node order node_perm[new] = old (per-node sigma rows follow)."""
        sn = None if self.sigma_node is None else self.sigma_node[np.asarray(node_perm)]
        return VGrid(self.ivcor, sn, self.nz, self.h_s, self.h_c, self.theta_b, self.theta_f, self.ztot, self.sigma_s)

    def c_struct(self) -> L.VGrid:
        v = L.VGrid()
        v.ivcor, v.nz = self.ivcor, self.nz
        v.h_s, v.h_c, v.theta_b, v.theta_f = self.h_s, self.h_c, self.theta_b, self.theta_f
        v.sigma_node = dptr(self.sigma_node) if self.sigma_node is not None else None
        v.ztot = dptr(self.ztot) if self.ztot is not None else None
        v.sigma_s = dptr(self.sigma_s) if self.sigma_s is not None else None
        return v


class Mesh:
    """A triangular mesh with its device-resident state (gsh_mesh).

    With renumber=True the nodes are put in reverse Cuthill-McKee order (elements follow);
    every array in and out of the mesh is in that order, and `node_perm[new] = old`.
    ctx=None builds a host-only mesh (topology and geometry, no device state).
    """

    def __init__(self, ctx: Context | None, xy, elnode, depth, sigma, open_bnd=None, kb=None, renumber=True,
                 vgrid: VGrid | None = None):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        elnode = np.ascontiguousarray(elnode, dtype=np.int32).reshape(-1, 3)
        depth = np.ascontiguousarray(depth, dtype=np.float64)
        sigma = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        n, ne = xy.shape[0], elnode.shape[0]
        ob = None if open_bnd is None else np.ascontiguousarray(open_bnd, dtype=np.int8)
        kbv = None if kb is None else np.ascontiguousarray(kb, dtype=np.int32)
        d = L.MeshDesc()
        d.n, d.ne = n, ne
        d.nvrt = vgrid.nvrt if vgrid is not None else sigma.size
        d.xy = dptr(xy)
        d.elnode = dptr(elnode, C.c_int32)
        d.depth = dptr(depth)
        d.open_bnd = ob.ctypes.data_as(C.POINTER(C.c_int8)) if ob is not None else None
        d.sigma = dptr(sigma) if sigma is not None else None
        d.kb = dptr(kbv, C.c_int32) if kbv is not None else None
        d.renumber = 1 if renumber else 0
        self._vg_struct = vgrid.c_struct() if vgrid is not None else None
        d.vgrid = C.pointer(self._vg_struct) if vgrid is not None else None
        self.node_perm = np.empty(n, dtype=np.int32)
        self.elem_perm = np.empty(ne, dtype=np.int32)
        h = C.c_void_p()
        check(L.lib().gsh_mesh_create(ctx.h if ctx is not None else None, C.byref(d), C.byref(h),
                                      dptr(self.node_perm, C.c_int32), dptr(self.elem_perm, C.c_int32)))
        self.h = h
        self.ctx = ctx
        _register(self)
        info = L.MeshInfo()
        check(L.lib().gsh_mesh_get_info(self.h, C.byref(info)))
        self.n, self.ne, self.ns, self.nvrt = info.n, info.ne, info.ns, info.nvrt
        self.nnz, self.max_row, self.n_open = info.nnz, info.max_row, info.n_open
        # the vertical grid in the mesh's own node order (for the oracle of the tests)
        self.vgrid = vgrid.permuted(self.node_perm) if vgrid is not None else None

    def close(self):
        if getattr(self, "h", None):
            L.lib().gsh_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---------------------------------------------------------------- topology
    def topology(self, code: int) -> np.ndarray:
        nbr = self.nnz - self.n
        spec = {
            L.T_XY: (np.float64, (self.n, 2)), L.T_ELNODE: (np.int32, (self.ne, 3)),
            L.T_ELSIDE: (np.int32, (self.ne, 3)), L.T_ISDEL: (np.int32, (self.ns, 2)),
            L.T_ISIDENODE: (np.int32, (self.ns, 2)), L.T_BALL_PTR: (np.int32, (self.n + 1,)),
            L.T_BALL_ELEM: (np.int32, (3 * self.ne,)), L.T_NBR_NODE: (np.int32, (nbr,)),
            L.T_CSR_INDPTR: (np.int32, (self.n + 1,)), L.T_CSR_INDICES: (np.int32, (self.nnz,)),
            L.T_CSR_SLOT: (np.int16, (nbr,)), L.T_DEPTH: (np.float64, (self.n,)),
            L.T_OPEN_BND: (np.int8, (self.n,)), L.T_KB: (np.int32, (self.n,)),
            L.T_SIDE_TYPE: (np.int8, (self.ns,)),
        }
        dt, shp = spec[code]
        out = np.empty(shp, dtype=dt)
        check(L.lib().gsh_mesh_get_topology(self.h, code, out.ctypes.data_as(C.c_void_p)))
        return out

    # ---------------------------------------------------------------- fields
    def _shape(self, code):
        if code == L.F_WET:
            return np.int8, (self.n + self.ne + self.ns,)
        dt, lead, tail = _FIELD_SPEC[code]
        shp = (getattr(self, lead),) + tuple(getattr(self, t) if isinstance(t, str) else t for t in tail)
        return dt, shp

    def set(self, code: int, arr) -> None:
        dt, shp = self._shape(code)
        a = np.ascontiguousarray(arr, dtype=dt).reshape(shp)
        check(L.lib().gsh_mesh_set_field(self.h, code, a.ctypes.data_as(C.c_void_p), a.size))

    def get(self, code: int) -> np.ndarray:
        dt, shp = self._shape(code)
        out = np.empty(shp, dtype=dt)
        check(L.lib().gsh_mesh_get_field(self.h, code, out.ctypes.data_as(C.c_void_p), out.size))
        return out

    def field_ptr(self, code: int) -> int:
        p, cnt = C.c_void_p(), C.c_int64()
        check(L.lib().gsh_mesh_field_ptr(self.h, code, C.byref(p), C.byref(cnt)))
        return p.value

    # ---------------------------------------------------------------- step operations
    def update_levels(self, params) -> None:
        check(L.lib().gsh_update_levels(self.h, C.byref(params)))

    def backtrack(self, params, copy_out: bool = False):
        """backtrackAllNodes: every side x level; returns (btu, foot) when copy_out."""
        if copy_out:
            btu = np.empty((self.ns, self.nvrt, 2), dtype=np.float64)
            foot = np.empty((self.ns, self.nvrt), dtype=np.int32)
            check(L.lib().gsh_backtrack(self.h, C.byref(params), dptr(btu), dptr(foot, C.c_int32)))
            return btu, foot
        check(L.lib().gsh_backtrack(self.h, C.byref(params), None, None))
        return None

    def set_backtrack_staging(self, staged: bool) -> None:
        """Backtracking reads elements through the caches (False, default) or stages each
        workgroup's element patch in LDS first (True); results are identical."""
        check(L.lib().gsh_mesh_set_backtrack_staging(self.h, 1 if staged else 0))

    def backtrack_list(self, params, sides, levels) -> None:
        s = np.ascontiguousarray(sides, dtype=np.int32)
        k = np.ascontiguousarray(levels, dtype=np.int32)
        check(L.lib().gsh_backtrack_list(self.h, C.byref(params), dptr(s, C.c_int32), dptr(k, C.c_int32), s.size))

    def rhs_producers(self, params) -> None:
        check(L.lib().gsh_rhs_producers(self.h, C.byref(params)))

    def nodal_velocity(self, params) -> None:
        """Nodal u2/w2 (GSH_F_NODE_UVW) from side u2 and element w (GSH_F_ELEM_W)."""
        check(L.lib().gsh_nodal_velocity(self.h, C.byref(params)))

    def set_tides(self, amig, ff, face, emo, efa) -> None:
        """Tidal constituents of the open boundary (iettype 3 and 5): angular frequency amig,
        nodal factor ff, equilibrium argument face [nbfr]; amplitude emo and phase efa [n][nbfr]
        (phases and arguments in radians)."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (amig, ff, face)]
        nb = a[0].size
        e = [np.ascontiguousarray(v, dtype=np.float64).reshape(self.n, nb) for v in (emo, efa)]
        pd = C.POINTER(C.c_double)
        check(L.lib().gsh_set_tides(self.h, nb, *[v.ctypes.data_as(pd) for v in a + e]))

    def momentum(self, params) -> None:
        """Side velocities u2 of the next step (GSH_F_SIDE_U2): elevation gradients, the momentum
        solve and the land-boundary slip condition."""
        check(L.lib().gsh_momentum(self.h, C.byref(params)))

    def vertical_velocity(self, params) -> None:
        """Element vertical velocity w (GSH_F_ELEM_W) from the side velocities and the levels of
        the last level update."""
        check(L.lib().gsh_vertical_velocity(self.h, C.byref(params)))


class Solver:
    """PCG on the device (gsh_solver), bound to a mesh or to a caller CSR pattern; Jacobi
    (the reference preconditioner) or multicolour SSOR."""

    def __init__(self, mesh: Mesh | None = None, precision: int = L.PREC_FP64, *, ctx: Context | None = None,
                 indptr=None, indices=None, precond: int = L.PRECOND_JACOBI):
        h = C.c_void_p()
        self.precond = precond
        if mesh is not None:
            check(L.lib().gsh_solver_create(mesh.h, precision, precond, C.byref(h)))
            self.n, self.nnz = mesh.n, mesh.nnz
        else:
            if ctx is None or indptr is None or indices is None:
                raise ValueError("a stand-alone solver needs ctx, indptr and indices")
            ip = np.ascontiguousarray(indptr, dtype=np.int32)
            ix = np.ascontiguousarray(indices, dtype=np.int32)
            self.n, self.nnz = ip.size - 1, ix.size
            check(L.lib().gsh_solver_create_csr(ctx.h, self.n, self.nnz, dptr(ip, C.c_int32), dptr(ix, C.c_int32),
                                                precision, precond, C.byref(h)))
        self.h = h
        self.mesh = mesh
        self.ctx = ctx if ctx is not None else (mesh.ctx if mesh is not None else None)
        self.precision = precision
        _register(self)

    def close(self):
        if getattr(self, "h", None):
            L.lib().gsh_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def csr_ptrs(self):
        d, i, p, nnz = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_int64()
        check(L.lib().gsh_solver_csr(self.h, C.byref(d), C.byref(i), C.byref(p), C.byref(nnz)))
        return d.value, i.value, p.value, nnz.value

    def set_values(self, val) -> None:
        v = np.ascontiguousarray(val, dtype=np.float64)
        if v.size != self.nnz:
            raise ValueError("value array has the wrong length")
        check(L.lib().gsh_solver_set_values(self.h, dptr(v)))

    def get_values(self) -> np.ndarray:
        v = np.empty(self.nnz, dtype=np.float64)
        check(L.lib().gsh_solver_get_values(self.h, dptr(v)))
        return v

    def get_solution(self) -> np.ndarray:
        x = np.empty(self.n, dtype=np.float64)
        check(L.lib().gsh_solver_get_solution(self.h, dptr(x)))
        return x

    def get_rhs(self) -> np.ndarray:
        b = np.empty(self.n, dtype=np.float64)
        check(L.lib().gsh_solver_get_rhs(self.h, dptr(b)))
        return b

    def set_rhs(self, b) -> None:
        v = np.ascontiguousarray(b, dtype=np.float64)
        if v.size != self.n:
            raise ValueError("right-hand side has the wrong length")
        check(L.lib().gsh_solver_set_rhs(self.h, dptr(v)))

    def assemble(self, params, time: float = 0.0) -> None:
        if self.mesh is None:
            raise ValueError("solver is not bound to a mesh")
        check(L.lib().gsh_assemble(self.mesh.h, self.h, C.byref(params), time))

    def solve(self, b=None, rtol: float = 1e-12, maxiter: int = 0, check_every: int = 10, timeout_s: float = 60.0,
              want_x: bool = True):
        """Returns (x or None, iters, status)."""
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.n, dtype=np.float64) if want_x else None
        it, st = C.c_int(), C.c_int()
        check(L.lib().gsh_solver_solve(self.h, dptr(bb) if bb is not None else None,
                                       dptr(x) if x is not None else None, rtol, maxiter, check_every, timeout_s,
                                       C.byref(it), C.byref(st)))
        return x, it.value, st.value

    def check(self) -> float:
        r = C.c_double()
        check(L.lib().gsh_solver_check(self.h, C.byref(r)))
        return r.value

    def reset_history(self) -> None:
        check(L.lib().gsh_solver_reset_history(self.h))

    def stats(self) -> L.SolveStats:
        s = L.SolveStats()
        check(L.lib().gsh_solver_last_stats(self.h, C.byref(s)))
        return s

    def set_tuning(self, rows_per_thread: int = 4, inner_rtol: float = 1e-5) -> None:
        check(L.lib().gsh_solver_set_tuning(self.h, rows_per_thread, inner_rtol))

    def autotune(self, iters: int = 30):
        """Picks rows_per_thread by timing `iters` iterations of each; returns (best, us/iter [4])."""
        best = C.c_int()
        us = np.zeros(4, dtype=np.float64)
        check(L.lib().gsh_solver_autotune(self.h, iters, C.byref(best), dptr(us)))
        return best.value, us

    def set_graphs(self, on: bool) -> None:
        check(L.lib().gsh_solver_set_graphs(self.h, 1 if on else 0))

    def set_omega(self, omega: float) -> None:
        check(L.lib().gsh_solver_set_omega(self.h, omega))

    def ssor_info(self):
        """(ncolours, padded rows, colour of every row) of an SSOR solver."""
        nc, npad = C.c_int32(), C.c_int32()
        col = np.empty(self.n, dtype=np.int32)
        check(L.lib().gsh_solver_ssor_info(self.h, C.byref(nc), C.byref(npad), dptr(col, C.c_int32)))
        return nc.value, npad.value, col

    def profile(self, rtol: float = 1e-12, maxiter: int = 0, check_every: int = 10):
        """One FP64 solve with per-launch timing: (ms per iteration [4], iterations)."""
        ms = np.zeros(4, dtype=np.float64)
        it = C.c_int()
        check(L.lib().gsh_solver_profile(self.h, rtol, maxiter, check_every, dptr(ms), C.byref(it)))
        return ms, it.value

    def update_elevation(self) -> None:
        check(L.lib().gsh_update_elevation(self.mesh.h, self.h))


def colour_csr(indptr, indices):
    """The SSOR solver's colouring of a CSR pattern (host only): (colour [n], ncolours)."""
    ip = np.ascontiguousarray(indptr, dtype=np.int32)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    col = np.empty(ip.size - 1, dtype=np.int32)
    nc = C.c_int32()
    check(L.lib().gsh_colour_csr(ip.size - 1, dptr(ip, C.c_int32), dptr(ix, C.c_int32), dptr(col, C.c_int32),
                                 C.byref(nc)))
    return col, nc.value


def batch_solve(solvers, rtol=1e-12, maxiter=0, check_every=10, timeout_s=60.0):
    arr = (C.c_void_p * len(solvers))(*[s.h.value for s in solvers])
    its = (C.c_int * len(solvers))()
    check(L.lib().gsh_batch_solve(arr, len(solvers), rtol, maxiter, check_every, timeout_s, its))
    return list(its)


class Block:
    """Ensemble members of one mesh solved together on its shared CSR pattern (gsh_block):
    Jacobi PCG per member with the reference's recurrences, member-interleaved on the device."""

    def __init__(self, mesh: Mesh, members: int):
        h = C.c_void_p()
        check(L.lib().gsh_block_create(mesh.h, members, C.byref(h)))
        self.h = h
        self.mesh = mesh
        self.members = members
        _register(self)

    def close(self):
        if getattr(self, "h", None):
            L.lib().gsh_block_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def assemble(self, member: int, params, time: float = 0.0) -> None:
        check(L.lib().gsh_block_assemble(self.h, member, C.byref(params), time))

    def set_member(self, member: int, values=None, b=None) -> None:
        v = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        r = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        check(L.lib().gsh_block_set_member(self.h, member, dptr(v) if v is not None else None,
                                           dptr(r) if r is not None else None))

    def solve(self, rtol: float = 1e-12, maxiter: int = 0, check_every: int = 10, timeout_s: float = 60.0):
        it = np.empty(self.members, dtype=np.int32)
        st = np.empty(self.members, dtype=np.int32)
        check(L.lib().gsh_block_solve(self.h, rtol, maxiter, check_every, timeout_s, dptr(it, C.c_int32),
                                      dptr(st, C.c_int32)))
        return it, st

    def solution(self, member: int) -> np.ndarray:
        x = np.empty(self.mesh.n)
        check(L.lib().gsh_block_get_solution(self.h, member, dptr(x)))
        return x


class Server:
    """Continuous-batching solve service (gsh_server): requests join and leave the lock-step
    PCG launches at check-window boundaries."""

    def __init__(self, ctx: Context, rtol: float = 1e-12, maxiter: int = 3000, check_every: int = 10,
                 max_active: int = 0, hbm_budget_bytes: int = 0, precision: int = L.PREC_FP64,
                 precond: int = L.PRECOND_JACOBI):
        cfg = L.ServerConfig()
        cfg.precond = precond
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.max_active = max_active
        cfg.precision = precision
        cfg.rtol = rtol
        cfg.maxiter = maxiter
        cfg.check_every = check_every
        h = C.c_void_p()
        check(L.lib().gsh_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self.ctx = ctx
        self.meshes: list[Mesh] = []
        _register(self)

    def close(self):
        if getattr(self, "h", None):
            L.lib().gsh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add_mesh(self, mesh: Mesh, params=None) -> int:
        """Registers a mesh; `params` (StepParams) are the step parameters of its requests'
        set-up (default: the library defaults)."""
        i = C.c_int32()
        check(L.lib().gsh_server_add_mesh(self.h, mesh.h, C.byref(i)))
        self.meshes.append(mesh)
        if params is not None:
            self.set_mesh_params(i.value, params)
        return i.value

    def set_mesh_params(self, mesh_id: int, params) -> None:
        check(L.lib().gsh_server_set_mesh_params(self.h, mesh_id, C.byref(params)))

    def submit(self, mesh_id: int, b=None, with_backtrack: bool = False) -> int:
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        t = C.c_int64()
        check(L.lib().gsh_server_submit(self.h, mesh_id, dptr(bb) if bb is not None else None,
                                        1 if with_backtrack else 0, C.byref(t)))
        return t.value

    def step(self):
        a, q = C.c_int32(), C.c_int32()
        check(L.lib().gsh_server_step(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def finished(self, cap: int = 1024) -> list[int]:
        """Tickets retired since the last call."""
        buf = (C.c_int64 * cap)()
        cnt = C.c_int32()
        check(L.lib().gsh_server_finished(self.h, buf, cap, C.byref(cnt)))
        return list(buf[:cnt.value])

    def poll(self, ticket: int, n: int | None = None):
        """(status, x or None, iters, latency_ms); status 0 = done, 1 = pending, else GSH_E_*."""
        x = np.empty(n, dtype=np.float64) if n else None
        it, st, lat = C.c_int32(), C.c_int32(), C.c_double()
        check(L.lib().gsh_server_poll(self.h, ticket, dptr(x) if x is not None else None, C.byref(it), C.byref(st),
                                      C.byref(lat)))
        return st.value, (x if st.value == 0 else None), it.value, lat.value

    def release(self, ticket: int) -> None:
        """Forgets a finished or rejected request, returning its host buffers to the server's pool."""
        check(L.lib().gsh_server_release(self.h, ticket))
