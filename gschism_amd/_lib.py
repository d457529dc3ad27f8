"""ctypes binding of libgsh.so (the C ABI declared in include/gsh.h).

The library is loaded from this package directory only; a missing library raises instead
of falling back to any other implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().with_name("libgsh.so")

# status codes (include/gsh.h)
GSH_OK = 0
GSH_E_ARG = 1
GSH_E_DEVICE = 2
GSH_E_OOM = 3
GSH_E_TIMEOUT = 4
GSH_E_ADMISSION = 5
GSH_E_TOPOLOGY = 6
GSH_E_STATE = 7
GSH_E_NOTFOUND = 8
GSH_E_IO = 9
GSH_W_MAXITER = 64
PREC_FP64 = 0
PREC_MIXED = 1
PRECOND_JACOBI = 0
PRECOND_SSOR = 1
STATIC_MAX_ROWS = 1048576  # GSH_STATIC_MAX_ROWS: rows of gsh_pcg_solve_static's static workspace

# topology array codes
T_XY, T_ELNODE, T_ELSIDE, T_ISDEL, T_ISIDENODE, T_BALL_PTR, T_BALL_ELEM, T_NBR_NODE = range(8)
T_CSR_INDPTR, T_CSR_INDICES, T_CSR_SLOT, T_DEPTH, T_OPEN_BND, T_KB, T_SIDE_TYPE = range(8, 15)
# field codes
(F_ETA2, F_ETA1, F_NODE_UVW, F_SIDE_U2, F_BTU, F_FOOT, F_BSTATUS, F_CD, F_GRAV2, F_ETH, F_VSOURCE,
 F_SIDE_U, F_SIDE_HHAT, F_SIDE_CHI, F_GHAT1, F_QEL, F_ZC, F_ELBC, F_WET, F_DEPTH, F_ELEM_W, F_ETYPE) = range(22)

BT_WALK_CAP, BT_ITER38, BT_STUCK, BT_OPEN_EXIT, BT_SLIDE = 1, 2, 4, 8, 16


class GshError(RuntimeError):
    """A libgsh call failed; `code` is the GSH_E_* status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[gsh status {code}] {msg}")
        self.code = code


class VGrid(C.Structure):
    _fields_ = [("ivcor", C.c_int32), ("sigma_node", C.POINTER(C.c_double)), ("nz", C.c_int32),
                ("h_s", C.c_double), ("h_c", C.c_double), ("theta_b", C.c_double), ("theta_f", C.c_double),
                ("ztot", C.POINTER(C.c_double)), ("sigma_s", C.POINTER(C.c_double))]


class MeshDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("ne", C.c_int32), ("nvrt", C.c_int32),
                ("xy", C.POINTER(C.c_double)), ("elnode", C.POINTER(C.c_int32)),
                ("depth", C.POINTER(C.c_double)), ("open_bnd", C.POINTER(C.c_int8)),
                ("sigma", C.POINTER(C.c_double)), ("kb", C.POINTER(C.c_int32)), ("renumber", C.c_int32),
                ("vgrid", C.POINTER(VGrid))]


class MeshInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("ne", C.c_int32), ("ns", C.c_int32), ("nvrt", C.c_int32),
                ("nnz", C.c_int64), ("max_row", C.c_int32), ("n_open", C.c_int32)]


class StepParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("theta", C.c_double), ("h0", C.c_double), ("ramp", C.c_double),
                ("dtb_min", C.c_double), ("dtb_max", C.c_double), ("velmin_btrack", C.c_double),
                ("ihhat", C.c_int32), ("iettype", C.c_int32), ("rmaxvel", C.c_double),
                ("dfv0", C.c_double), ("time", C.c_double), ("shapiro0", C.c_double),
                ("ishapiro", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("outer", C.c_int32), ("rr", C.c_double), ("atol2", C.c_double),
                ("true_rel", C.c_double), ("ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active", C.c_int32), ("precision", C.c_int32),
                ("rtol", C.c_double), ("maxiter", C.c_int32), ("check_every", C.c_int32),
                ("precond", C.c_int32)]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_I = C.c_int
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double
_PI32 = C.POINTER(C.c_int32)
_PI64 = C.POINTER(C.c_int64)
_PD = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "gsh_ctx_create": (_I, [_I, _PP]),
    "gsh_ctx_destroy": (_I, [_P]),
    "gsh_last_error": (C.c_char_p, []),
    "gsh_version": (_I, []),
    "gsh_ctx_mem_info": (_I, [_P, _PI64, _PI64]),
    "gsh_ctx_synchronize": (_I, [_P]),
    "gsh_mesh_create": (_I, [_P, C.POINTER(MeshDesc), _PP, _PI32, _PI32]),
    "gsh_mesh_destroy": (_I, [_P]),
    "gsh_mesh_get_info": (_I, [_P, C.POINTER(MeshInfo)]),
    "gsh_mesh_get_topology": (_I, [_P, _I, _P]),
    "gsh_mesh_set_field": (_I, [_P, _I, _P, _I64]),
    "gsh_mesh_get_field": (_I, [_P, _I, _P, _I64]),
    "gsh_mesh_field_ptr": (_I, [_P, _I, _PP, _PI64]),
    "gsh_step_params_default": (None, [C.POINTER(StepParams)]),
    "gsh_update_levels": (_I, [_P, C.POINTER(StepParams)]),
    "gsh_backtrack": (_I, [_P, C.POINTER(StepParams), _PD, _PI32]),
    "gsh_backtrack_list": (_I, [_P, C.POINTER(StepParams), _PI32, _PI32, _I32]),
    "gsh_mesh_set_backtrack_staging": (_I, [_P, _I32]),
    "gsh_rhs_producers": (_I, [_P, C.POINTER(StepParams)]),
    "gsh_nodal_velocity": (_I, [_P, C.POINTER(StepParams)]),
    "gsh_momentum": (_I, [_P, C.POINTER(StepParams)]),
    "gsh_vertical_velocity": (_I, [_P, C.POINTER(StepParams)]),
    "gsh_set_tides": (_I, [_P, _I32, _PD, _PD, _PD, _PD, _PD]),
    "gsh_grid_read": (_I, [C.c_char_p, C.c_char_p, _PP]),
    "gsh_grid_info": (_I, [_P, _PI32, _PI32, _PI32, _PI32, _PI32]),
    "gsh_grid_get": (_I, [_P, _PD, _PD, _PI32, C.POINTER(C.c_int8), _PI32, _PD]),
    "gsh_grid_segment": (_I, [_P, _I, _I32, _PI32, _I32, _PI32, _PI32]),
    "gsh_grid_vgrid": (_I, [_P, _PI32, _PI32, _PI32, _PD, _PD, _PD, _PD]),
    "gsh_grid_destroy": (_I, [_P]),
    "gsh_solver_create": (_I, [_P, _I, _I, _PP]),
    "gsh_solver_create_csr": (_I, [_P, _I32, _I64, _PI32, _PI32, _I, _I, _PP]),
    "gsh_solver_destroy": (_I, [_P]),
    "gsh_solver_csr": (_I, [_P, _PP, _PP, _PP, _PI64]),
    "gsh_solver_set_values": (_I, [_P, _PD]),
    "gsh_solver_get_values": (_I, [_P, _PD]),
    "gsh_solver_vec_ptr": (_I, [_P, _PP, _PP]),
    "gsh_solver_get_rhs": (_I, [_P, _PD]),
    "gsh_solver_set_rhs": (_I, [_P, _PD]),
    "gsh_solver_get_solution": (_I, [_P, _PD]),
    "gsh_assemble": (_I, [_P, _P, C.POINTER(StepParams), _D]),
    "gsh_solver_solve": (_I, [_P, _PD, _PD, _D, _I, _I, _D, C.POINTER(_I), C.POINTER(_I)]),
    "gsh_pcg_solve_static": (_I, [_P, _I32, _I64, _P, _P, _P, _P, _P, _D, _I, _I, _D, C.POINTER(_I), C.POINTER(_I)]),
    "gsh_solver_check": (_I, [_P, _PD]),
    "gsh_solver_reset_history": (_I, [_P]),
    "gsh_solver_last_stats": (_I, [_P, C.POINTER(SolveStats)]),
    "gsh_solver_set_tuning": (_I, [_P, _I, _D]),
    "gsh_update_elevation": (_I, [_P, _P]),
    "gsh_batch_solve": (_I, [_PP, _I, _D, _I, _I, _D, C.POINTER(_I)]),
    "gsh_block_create": (_I, [_P, _I32, _PP]),
    "gsh_block_destroy": (_I, [_P]),
    "gsh_block_assemble": (_I, [_P, _I32, C.POINTER(StepParams), _D]),
    "gsh_block_set_member": (_I, [_P, _I32, _PD, _PD]),
    "gsh_block_solve": (_I, [_P, _D, _I, _I, _D, _PI32, _PI32]),
    "gsh_block_get_solution": (_I, [_P, _I32, _PD]),
    "gsh_server_create": (_I, [_P, C.POINTER(ServerConfig), _PP]),
    "gsh_server_destroy": (_I, [_P]),
    "gsh_server_add_mesh": (_I, [_P, _P, _PI32]),
    "gsh_server_set_mesh_params": (_I, [_P, _I32, C.POINTER(StepParams)]),
    "gsh_server_submit": (_I, [_P, _I32, _PD, _I, _PI64]),
    "gsh_server_step": (_I, [_P, _PI32, _PI32]),
    "gsh_server_finished": (_I, [_P, C.POINTER(C.c_int64), _I32, _PI32]),
    "gsh_server_poll": (_I, [_P, _I64, _PD, _PI32, _PI32, _PD]),
    "gsh_server_release": (_I, [_P, _I64]),
    "gsh_copy_probe": (_I, [_P, _I64, _I, _PD]),
    "gsh_solver_profile": (_I, [_P, _D, _I, _I, _PD, C.POINTER(_I)]),
    "gsh_solver_set_omega": (_I, [_P, _D]),
    "gsh_solver_set_graphs": (_I, [_P, _I]),
    "gsh_solver_autotune": (_I, [_P, _I, C.POINTER(_I), _PD]),
    "gsh_solver_ssor_info": (_I, [_P, _PI32, _PI32, _PI32]),
    "gsh_colour_csr": (_I, [_I32, _PI32, _PI32, _PI32, _PI32]),
    "gsh_gen_channel": (_I, [_I32, _I32, _D, _D, C.c_uint64, _PP]),
    "gsh_gen_sizes": (_I, [_P, _PI32, _PI32]),
    "gsh_gen_get": (_I, [_P, _PD, _PI32, C.POINTER(C.c_int8)]),
    "gsh_gen_destroy": (_I, [_P]),
    "gsh_gen_bathymetry": (_I, [_I32, _I32, _D, _D, _D, C.c_uint64, _PD]),
    "gsh_gen_velocity": (_I, [_I32, _PD, _I32, _PD, _D, _D, _D, _I32, _D, _D, C.c_uint64, _I32, _PD]),
    "gsh_gen_normal": (_I, [C.c_uint64, _I64, _D, _D, _PD]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libgsh.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run gschism_amd.build.build()")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> int:
    if status != GSH_OK:
        msg = lib().gsh_last_error()
        raise GshError(status, msg.decode() if msg else "")
    return status


def dptr(a: np.ndarray, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def default_step_params(**kw) -> StepParams:
    p = StepParams()
    lib().gsh_step_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown step parameter {k}")
        setattr(p, k, v)
    return p
