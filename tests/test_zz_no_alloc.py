"""The solve that allocates nothing (gsh_pcg_solve_static) against gsh_solver_solve.

Inputs and output are device buffers the test allocates with the HIP runtime libgsh links; the
entry point keeps every other buffer in the library's static device array. Two calls in a row,
the second on a smaller system, must each equal a fresh FP64 Jacobi solver's gsh_solver_solve
(zero initial guess) bit for bit: iterations, status and solution; sizes beyond the static
array and broken patterns are refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from gschism_amd import _lib as L
from gschism_amd import cases, core
from oracle import pyoracle as O

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    L.lib()  # libgsh first, so the runtime below is the instance it uses
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipSetDevice.argtypes = [C.c_int]
    return h


class DevBuf:
    def __init__(self, hip, arr=None, nbytes=None):
        self.hip = hip
        self.nbytes = int(arr.nbytes if arr is not None else nbytes)
        self.p = C.c_void_p()
        assert hip.hipSetDevice(0) == 0
        assert hip.hipMalloc(C.byref(self.p), max(self.nbytes, 8)) == 0
        if arr is not None:
            a = np.ascontiguousarray(arr)
            assert hip.hipMemcpy(self.p, a.ctypes.data, self.nbytes, H2D) == 0

    def read(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _static_solve(ctx, hip, ip, ix, val, b, rtol=1e-12, maxiter=5000, check_every=10):
    n, nnz = len(ip) - 1, len(ix)
    bufs = [DevBuf(hip, ip.astype(np.int32)), DevBuf(hip, ix.astype(np.int32)), DevBuf(hip, val.astype(np.float64)),
            DevBuf(hip, b.astype(np.float64)), DevBuf(hip, nbytes=8 * n)]
    it, st = C.c_int(-1), C.c_int(-1)
    try:
        L.check(L.lib().gsh_pcg_solve_static(ctx.h, n, nnz, *[q.p for q in bufs], rtol, maxiter, check_every, 60.0,
                                            C.byref(it), C.byref(st)))
        x = bufs[4].read(np.float64, n)
    finally:
        for q in bufs:
            q.free()
    return x, it.value, st.value


def _reference(ctx, ip, ix, val, b, rtol=1e-12, maxiter=5000, check_every=10):
    s = core.Solver(None, ctx=ctx, indptr=ip, indices=ix)
    try:
        s.set_values(val)
        s.reset_history()
        return s.solve(b, rtol=rtol, maxiter=maxiter, check_every=check_every)
    finally:
        s.close()


def _assembled(ctx, nx, ny):
    spec = cases.CaseSpec(f"channel-{nx}x{ny}", nx, ny, 500.0, nvrt=4, depth=None, substeps=5)
    data = cases.make_case(spec)
    mesh, p, _ = cases.load_mesh(ctx, data)
    s = core.Solver(mesh)
    mesh.backtrack(p)
    mesh.rhs_producers(p)
    s.assemble(p)
    ip, ix = mesh.topology(L.T_CSR_INDPTR), mesh.topology(L.T_CSR_INDICES)
    val, b = s.get_values(), s.get_rhs()
    s.close()
    mesh.close()
    return ip, ix, val, b


def test_static_solve_equals_solver_solve_twice(ctx):
    hip = _hip()
    big = _assembled(ctx, 200, 80)
    small = _assembled(ctx, 40, 24)
    for tag, (ip, ix, val, b) in (("200x80", big), ("40x24", small)):
        print(f"[gpu] static solve {tag}: n={len(ip) - 1} nnz={len(ix)}", flush=True)
        x, it, st = _static_solve(ctx, hip, ip, ix, val, b)
        xr, itr, str_ = _reference(ctx, ip, ix, val, b)
        assert (it, st) == (itr, str_), (tag, it, st, itr, str_)
        assert st == 0
        assert_bitwise(x, xr, f"static solve {tag}")
        xo, ito, _, _ = O.pcg(ip, ix, val, b, rtol=1e-12, maxiter=5000, check_every=10)
        assert ito == it
        assert_bitwise(x, xo, f"static solve {tag} vs oracle")


def test_static_solve_maxiter_and_zero_rhs(ctx):
    hip = _hip()
    ip, ix, val, b = _assembled(ctx, 30, 20)
    x, it, st = _static_solve(ctx, hip, ip, ix, val, b, rtol=1e-30, maxiter=25)
    xr, itr, str_ = _reference(ctx, ip, ix, val, b, rtol=1e-30, maxiter=25)
    assert it == itr == 25 and st == str_ == L.GSH_W_MAXITER
    assert_bitwise(x, xr, "static solve at maxiter")
    x, it, st = _static_solve(ctx, hip, ip, ix, val, np.zeros_like(b))
    assert it == 0 and st == 0 and not x.any()


def test_static_solve_refusals(ctx):
    hip = _hip()
    lib = L.lib()
    it, st = C.c_int(), C.c_int()
    # beyond the static array: refused before any device work
    r = lib.gsh_pcg_solve_static(ctx.h, L.STATIC_MAX_ROWS + 1, 10, 1, 1, 1, 1, 1, 1e-12, 10, 10, 60.0, C.byref(it),
                                 C.byref(st))
    assert r == L.GSH_E_ARG and b"GSH_STATIC_MAX_ROWS" in lib.gsh_last_error()
    # a row without its diagonal entry, found by the device check
    ip = np.array([0, 2, 3, 5], dtype=np.int32)
    ix = np.array([0, 1, 0, 1, 2], dtype=np.int32)  # row 1 lacks (1, 1)
    with pytest.raises(L.GshError) as ei:
        _static_solve(ctx, hip, ip, ix, np.ones(5), np.ones(3))
    assert ei.value.code == L.GSH_E_ARG and "diagonal" in str(ei.value)
    ix_bad = np.array([0, 1, 7, 1, 2], dtype=np.int32)
    with pytest.raises(L.GshError) as ei:
        _static_solve(ctx, hip, ip, ix_bad, np.ones(5), np.ones(3))
    assert ei.value.code == L.GSH_E_ARG and "column" in str(ei.value)
    # and the next call is unaffected
    ip, ix, val, b = _assembled(ctx, 30, 20)
    x, it2, st2 = _static_solve(ctx, hip, ip, ix, val, b)
    xr, itr, _ = _reference(ctx, ip, ix, val, b)
    assert it2 == itr
    assert_bitwise(x, xr, "static solve after refusals")
