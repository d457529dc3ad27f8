"""ctypes binding of the CPU oracle (test infrastructure: parity checker and CPU baseline).

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"


class OrParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("theta", C.c_double), ("h0", C.c_double), ("ramp", C.c_double),
                ("dtb_min", C.c_double), ("dtb_max", C.c_double), ("velmin_btrack", C.c_double),
                ("ihhat", C.c_int32), ("iettype", C.c_int32), ("rmaxvel", C.c_double),
                ("dfv0", C.c_double), ("time", C.c_double), ("shapiro0", C.c_double),
                ("ishapiro", C.c_int32)]


_P = C.c_void_p
_PD = C.POINTER(C.c_double)
_PI32 = C.POINTER(C.c_int32)
_SIG = {
    "or_error": (C.c_char_p, []),
    "or_mesh_new": (_P, [C.c_int32, C.c_int32, C.c_int32, _PD, _PI32, _PD, C.POINTER(C.c_int8), _PI32, _PD]),
    "or_mesh_free": (None, [_P]),
    "or_mesh_ns": (C.c_int32, [_P]),
    "or_mesh_nnz": (C.c_int64, [_P]),
    "or_mesh_nbr_count": (C.c_int32, [_P]),
    "or_mesh_get": (C.c_int, [_P, C.c_int, _P]),
    "or_set_field": (C.c_int, [_P, C.c_int, _P]),
    "or_get_field": (C.c_int, [_P, C.c_int, _P]),
    "or_levels": (None, [_P, C.POINTER(OrParams)]),
    "or_backtrack": (None, [_P, C.POINTER(OrParams), _PI32, _PI32, C.c_int64, C.c_int]),
    "or_rhs": (None, [_P, C.POINTER(OrParams)]),
    "or_nodal_velocity": (None, [_P, C.POINTER(OrParams), C.c_int]),
    "or_momentum": (None, [_P, C.POINTER(OrParams)]),
    "or_set_tides": (C.c_int, [_P, C.c_int32, _PD, _PD, _PD, _PD, _PD]),
    "or_set_vertical": (C.c_int, [_P, C.c_int32, _PD, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                  _PD, _PD]),
    "or_vertical_velocity": (None, [_P]),
    "or_assemble": (None, [_P, C.POINTER(OrParams), _PD, _PD, C.c_int]),
    "or_pcg": (C.c_int32, [C.c_int32, _PI32, _PI32, _PD, _PD, _PD, _PD, _PD, C.c_double, C.c_int32, C.c_int32,
                           C.c_int32, _PD, _PD, _PD, C.c_int]),
    "or_colour": (C.c_int32, [C.c_int32, _PI32, _PI32, _PI32]),
    "or_pcg_block": (C.c_int, [C.c_int32, _PI32, _PI32, C.c_int32, _PD, _PD, C.c_double, C.c_int32, C.c_int32, _PD,
                               _PI32, _PD, C.c_int]),
    "or_pcg_ssor": (C.c_int32, [C.c_int32, _PI32, _PI32, _PD, _PD, _PD, _PD, _PD, C.c_double, C.c_double,
                                C.c_int32, C.c_int32, C.c_int32, _PD, _PD, _PD, _PI32, C.c_int]),
    "or_ssor_new": (_P, [C.c_int32, _PI32, _PI32, C.c_int32]),
    "or_ssor_ncolours": (C.c_int32, [_P]),
    "or_ssor_solve": (C.c_int32, [_P, _PD, _PD, _PD, _PD, _PD, C.c_double, C.c_double, C.c_int32, C.c_int32, _PD, _PD,
                                  _PD, C.c_int]),
    "or_ssor_free": (None, [_P]),
    "or_geometry": (C.c_int, [_P, C.c_int, _PD]),
    "or_interp": (None, [_P, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, _PD]),
    "or_dot_canon": (C.c_double, [C.c_int32, _PD, _PD, C.c_int32]),
    "or_true_rel": (C.c_double, [C.c_int32, _PI32, _PI32, _PD, _PD, _PD, C.c_int32]),
}

_libs: dict = {}


def load(path) -> C.CDLL:
    """Loads an oracle library (the default build, or one with a planted mistake) once per path."""
    key = str(path)
    if key not in _libs:
        if not Path(path).exists():
            raise RuntimeError(f"{path} is not built: run python oracle/build.py")
        L = C.CDLL(key)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _libs[key] = L
    return _libs[key]


def lib() -> C.CDLL:
    return load(LIB_PATH)


def params(p) -> OrParams:
    """Copies a gschism_amd StepParams (or any object with the same fields) into OrParams."""
    o = OrParams()
    for name, _ in OrParams._fields_:
        setattr(o, name, getattr(p, name))
    return o


def _d(a):
    return None if a is None else a.ctypes.data_as(_PD)


def _i(a):
    return None if a is None else a.ctypes.data_as(_PI32)


# field layouts: code -> (dtype, shape builder)
_FIELDS = {
    0: (np.float64, lambda m: (m.n,)), 1: (np.float64, lambda m: (m.n,)),
    2: (np.float64, lambda m: (m.n, m.nvrt, 3)), 3: (np.float64, lambda m: (m.ns, m.nvrt, 2)),
    4: (np.float64, lambda m: (m.ns, m.nvrt, 2)), 5: (np.int32, lambda m: (m.ns, m.nvrt)),
    6: (np.int32, lambda m: (m.ns, m.nvrt)), 7: (np.float64, lambda m: (m.n,)),
    8: (np.float64, lambda m: (m.n,)), 9: (np.float64, lambda m: (m.n,)),
    10: (np.float64, lambda m: (m.ne,)), 11: (np.float64, lambda m: (m.ns, 2)),
    12: (np.float64, lambda m: (m.ns,)), 13: (np.float64, lambda m: (m.ns,)),
    14: (np.float64, lambda m: (m.ne, 2)), 16: (np.float64, lambda m: (m.ns, m.nvrt)),
    17: (np.float64, lambda m: (m.n,)), 18: (np.int8, lambda m: (m.n + m.ne + m.ns,)),
    19: (np.float64, lambda m: (m.n,)), 20: (np.float64, lambda m: (m.ne, m.nvrt)),
    21: (np.int8, lambda m: (m.n,)),
}

_TOPO = {
    0: (np.float64, lambda m: (m.n, 2)), 1: (np.int32, lambda m: (m.ne, 3)), 2: (np.int32, lambda m: (m.ne, 3)),
    3: (np.int32, lambda m: (m.ns, 2)), 4: (np.int32, lambda m: (m.ns, 2)), 5: (np.int32, lambda m: (m.n + 1,)),
    6: (np.int32, lambda m: (3 * m.ne,)), 7: (np.int32, lambda m: (m.nbr,)), 8: (np.int32, lambda m: (m.n + 1,)),
    9: (np.int32, lambda m: (m.nnz,)), 10: (np.int16, lambda m: (m.nbr,)), 11: (np.float64, lambda m: (m.n,)),
    12: (np.int8, lambda m: (m.n,)), 13: (np.int32, lambda m: (m.n,)), 14: (np.int8, lambda m: (m.ns,)),
}


class OracleMesh:
    """FP64 CPU restatement of one mesh and its per-step state."""

    def __init__(self, xy, elnode, depth, open_bnd, sigma, kb=None, library=None):
        L = self.L = lib() if library is None else library
        self.xy = np.ascontiguousarray(xy, dtype=np.float64)
        self.elnode = np.ascontiguousarray(elnode, dtype=np.int32)
        self.depth = np.ascontiguousarray(depth, dtype=np.float64)
        self.open_bnd = np.ascontiguousarray(open_bnd, dtype=np.int8)
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        self.kb = None if kb is None else np.ascontiguousarray(kb, dtype=np.int32)
        self.n = self.xy.shape[0]
        self.ne = self.elnode.shape[0]
        self.nvrt = self.sigma.shape[0]
        self.h = L.or_mesh_new(self.n, self.ne, self.nvrt, _d(self.xy), _i(self.elnode), _d(self.depth),
                               self.open_bnd.ctypes.data_as(C.POINTER(C.c_int8)), _i(self.kb), _d(self.sigma))
        if not self.h:
            raise RuntimeError("oracle mesh: " + L.or_error().decode())
        self.ns = L.or_mesh_ns(self.h)
        self.nnz = L.or_mesh_nnz(self.h)
        self.nbr = L.or_mesh_nbr_count(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_mesh_free(self.h)
            self.h = None

    def topology(self, code: int) -> np.ndarray:
        dt, shp = _TOPO[code]
        out = np.empty(shp(self), dtype=dt)
        if self.L.or_mesh_get(self.h, code, out.ctypes.data_as(C.c_void_p)):
            raise RuntimeError(self.L.or_error().decode())
        return out

    def set(self, code: int, arr) -> None:
        dt, shp = _FIELDS[code]
        a = np.ascontiguousarray(arr, dtype=dt).reshape(shp(self))
        if self.L.or_set_field(self.h, code, a.ctypes.data_as(C.c_void_p)):
            raise RuntimeError(self.L.or_error().decode())

    def get(self, code: int) -> np.ndarray:
        dt, shp = _FIELDS[code]
        out = np.empty(shp(self), dtype=dt)
        if self.L.or_get_field(self.h, code, out.ctypes.data_as(C.c_void_p)):
            raise RuntimeError(self.L.or_error().decode())
        return out

    def levels(self, p) -> None:
        self.L.or_levels(self.h, C.byref(params(p)))

    def backtrack(self, p, sides=None, levels=None, nthreads: int = 1) -> None:
        if sides is None:
            self.L.or_backtrack(self.h, C.byref(params(p)), None, None, -1, nthreads)
        else:
            s = np.ascontiguousarray(sides, dtype=np.int32)
            k = np.ascontiguousarray(levels, dtype=np.int32)
            self.L.or_backtrack(self.h, C.byref(params(p)), _i(s), _i(k), s.size, nthreads)

    def geometry(self, what: int) -> np.ndarray:
        shp = {0: (self.ne,), 1: (self.ne, 3, 2), 2: (self.ns, 2), 3: (self.ns,), 4: (self.ns, 2)}[what]
        out = np.empty(shp, dtype=np.float64)
        if self.L.or_geometry(self.h, what, _d(out)):
            raise RuntimeError(self.L.or_error().decode())
        return out

    def interp(self, e: int, x: float, y: float, z: float, h0: float = 0.01) -> np.ndarray:
        out = np.empty(3, dtype=np.float64)
        self.L.or_interp(self.h, e, x, y, z, h0, _d(out))
        return out

    def rhs(self, p) -> None:
        self.L.or_rhs(self.h, C.byref(params(p)))

    def nodal_velocity(self, p, nthreads: int = 1) -> None:
        self.L.or_nodal_velocity(self.h, C.byref(params(p)), nthreads)

    def set_vertical(self, vg) -> None:
        """Vertical grid beyond one sigma per level: `vg` has the fields of gsh_vgrid (ivcor,
        sigma_node [n][nvrt] in this mesh's node order, nz, h_s, h_c, theta_b, theta_f, ztot,
        sigma_s), e.g. gschism_amd.core.VGrid."""
        sn = None if vg.sigma_node is None else np.ascontiguousarray(vg.sigma_node, dtype=np.float64)
        zt = None if vg.ztot is None else np.ascontiguousarray(vg.ztot, dtype=np.float64)
        ss = None if vg.sigma_s is None else np.ascontiguousarray(vg.sigma_s, dtype=np.float64)
        self._vg = (sn, zt, ss)
        if self.L.or_set_vertical(self.h, int(vg.ivcor), _d(sn), int(vg.nz), float(vg.h_s), float(vg.h_c),
                                  float(vg.theta_b), float(vg.theta_f), _d(zt), _d(ss)):
            raise RuntimeError("oracle vertical grid: " + self.L.or_error().decode())

    def set_tides(self, amig, ff, face, emo, efa) -> None:
        """Tidal constituents (amig, ff, face [nbfr]) and open-node amplitudes and phases [n][nbfr]."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (amig, ff, face)]
        nb = a[0].size
        e = [np.ascontiguousarray(v, dtype=np.float64).reshape(self.n, nb) for v in (emo, efa)]
        if self.L.or_set_tides(self.h, nb, *[v.ctypes.data_as(_PD) for v in a + e]):
            raise RuntimeError(self.L.or_error().decode())

    def momentum(self, p) -> None:
        self.L.or_momentum(self.h, C.byref(params(p)))

    def vertical_velocity(self) -> None:
        self.L.or_vertical_velocity(self.h)

    def assemble(self, p, nthreads: int = 1):
        val = np.zeros(self.nnz, dtype=np.float64)
        b = np.zeros(self.n, dtype=np.float64)
        self.L.or_assemble(self.h, C.byref(params(p)), _d(val), _d(b), nthreads)
        return val, b


def pcg(indptr, indices, val, b, hist=None, rtol=1e-12, maxiter=1500, check_every=10, rows_per_thread=4,
        nthreads: int = 1):
    """Jacobi PCG with the reference recurrences; returns (x, iters, rr_last, atol2)."""
    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = b.size
    x = np.zeros(n, dtype=np.float64)
    rr = C.c_double()
    at = C.c_double()
    h = [None, None, None] if hist is None else [np.ascontiguousarray(v, dtype=np.float64) for v in hist]
    it = lib().or_pcg(n, _i(indptr), _i(indices), _d(val), _d(b), _d(h[0]), _d(h[1]), _d(h[2]), rtol, maxiter,
                      check_every, rows_per_thread, _d(x), C.byref(rr), C.byref(at), nthreads)
    return x, int(it), rr.value, at.value


def colour(indptr, indices):
    """Smallest-last greedy colouring of a CSR pattern; returns (colour [n], ncolours)."""
    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    n = indptr.size - 1
    col = np.empty(n, dtype=np.int32)
    nc = lib().or_colour(n, _i(indptr), _i(indices), _i(col))
    return col, int(nc)


def pcg_block(indptr, indices, vals, bs, rtol=1e-12, maxiter=1500, check_every=10, nthreads: int = 1):
    """Multi-RHS block Jacobi PCG: vals [members][nnz], bs [members][n] (members padded to a power
    of two with unit values and zero right-hand sides, as the library does). Returns
    (x [members][n], iters [members], rr_last [members])."""
    vals = np.asarray(vals, dtype=np.float64)
    bs = np.asarray(bs, dtype=np.float64)
    k, nnz = vals.shape
    n = bs.shape[1]
    M = 1
    while M < k:
        M *= 2
    vb = np.ones((nnz, M))
    vb[:, :k] = vals.T
    bb = np.zeros((n, M))
    bb[:, :k] = bs.T
    vb = np.ascontiguousarray(vb)
    bb = np.ascontiguousarray(bb)
    x = np.zeros((n, M))
    it = np.zeros(M, dtype=np.int32)
    rr = np.zeros(M)
    ip = np.ascontiguousarray(indptr, dtype=np.int32)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    if lib().or_pcg_block(n, _i(ip), _i(ix), M, _d(vb), _d(bb), rtol, maxiter, check_every, _d(x), _i(it), _d(rr),
                          nthreads):
        raise RuntimeError(lib().or_error().decode())
    return np.ascontiguousarray(x[:, :k].T), it[:k], rr[:k]


def pcg_ssor(indptr, indices, val, b, hist=None, omega=1.0, rtol=1e-12, maxiter=1500, check_every=10,
             rows_per_thread=1, nthreads: int = 1):
    """Multicolour SSOR PCG (Eisenstat form); returns (x, iters, rr_last, atol2, ncolours)."""
    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = b.size
    x = np.zeros(n, dtype=np.float64)
    rr, at, nc = C.c_double(), C.c_double(), C.c_int32()
    h = [None, None, None] if hist is None else [np.ascontiguousarray(v, dtype=np.float64) for v in hist]
    it = lib().or_pcg_ssor(n, _i(indptr), _i(indices), _d(val), _d(b), _d(h[0]), _d(h[1]), _d(h[2]), omega, rtol,
                           maxiter, check_every, rows_per_thread, _d(x), C.byref(rr), C.byref(at), C.byref(nc),
                           nthreads)
    if it < 0:
        raise ValueError(lib().or_error().decode())
    return x, int(it), rr.value, at.value, nc.value


class OracleSsor:
    """SSOR PCG with the colouring and layout built once (as the device solver keeps them)."""

    def __init__(self, indptr, indices, rows_per_thread=1):
        self._ip = np.ascontiguousarray(indptr, dtype=np.int32)
        self._ix = np.ascontiguousarray(indices, dtype=np.int32)
        self.n = self._ip.size - 1
        self.h = lib().or_ssor_new(self.n, _i(self._ip), _i(self._ix), rows_per_thread)
        if not self.h:
            raise ValueError(lib().or_error().decode())
        self.ncolours = int(lib().or_ssor_ncolours(self.h))

    def solve(self, val, b, hist=None, omega=1.0, rtol=1e-12, maxiter=1500, check_every=10, nthreads: int = 1):
        """Returns (x, iters, rr_last, atol2)."""
        val = np.ascontiguousarray(val, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n, dtype=np.float64)
        rr, at = C.c_double(), C.c_double()
        h = [None, None, None] if hist is None else [np.ascontiguousarray(v, dtype=np.float64) for v in hist]
        it = lib().or_ssor_solve(self.h, _d(val), _d(b), _d(h[0]), _d(h[1]), _d(h[2]), omega, rtol, maxiter,
                                 check_every, _d(x), C.byref(rr), C.byref(at), nthreads)
        return x, int(it), rr.value, at.value

    def close(self):
        if getattr(self, "h", None):
            lib().or_ssor_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def dot_canon(a, b, rows_per_thread=4) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().or_dot_canon(a.size, _d(a), _d(b), rows_per_thread)


def true_rel(indptr, indices, val, b, x, rows_per_thread=4) -> float:
    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().or_true_rel(b.size, _i(indptr), _i(indices), _d(val), _d(b), _d(x), rows_per_thread)
