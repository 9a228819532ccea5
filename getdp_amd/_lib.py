"""ctypes binding of libgdp_hip.so (the C-ABI declared in include/gdp_hip.h).

The library is built in-tree by ``getdp_amd.build``; loading fails loudly if
it is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libgdp_hip.so"

# every symbol declared in include/gdp_hip.h: (name, restype, argtypes)
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_fp = C.POINTER(C.c_float)


class MeshDesc(C.Structure):
    _fields_ = [("nnodes", C.c_int), ("xyz", _dp), ("nelem", C.c_int), ("etype", _ip),
                ("region", _ip), ("conn", _ip), ("edges", _ip), ("nedges", C.c_int)]


class SpaceDesc(C.Structure):
    _fields_ = [("bf", C.c_int), ("numtype", C.c_int), ("nsupport", C.c_int),
                ("support", _ip), ("nfixed", C.c_int), ("fixed_entity", _ip),
                ("fixed_value", _dp)]


class TermDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("equ", C.c_int), ("dof", C.c_int), ("nregions", C.c_int),
                ("regions", _ip), ("npoints", C.c_int), ("jac_sur", C.c_int),
                ("ncoef", C.c_int), ("coef_nreg", C.c_int), ("coef", _dp)]


class PostDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("space", C.c_int), ("nregions", C.c_int), ("regions", _ip),
                ("ncoef", C.c_int), ("coef_nreg", C.c_int), ("coef", _dp), ("npoints", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("its", C.c_int), ("status", C.c_int), ("ilu_row", C.c_int),
                ("ilu_levels", C.c_int), ("outer_its", C.c_int), ("res0", C.c_double),
                ("res_est", C.c_double), ("true_res", C.c_double), ("rel_res", C.c_double),
                ("t_setup_ms", C.c_double), ("t_solve_ms", C.c_double), ("ilu_sweep", C.c_int),
                ("t_factor_ms", C.c_double), ("t_load_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SIGNATURES = [
    ("gdp_init", C.c_int, [C.c_int]),
    ("gdp_finalize", C.c_int, []),
    ("gdp_last_error", C.c_char_p, []),
    ("gdp_device_synchronize", C.c_int, []),
    ("gdp_solver_create", C.c_int, [C.c_char_p, _pp]),
    ("gdp_solver_set", C.c_int, [_vp, C.c_char_p, C.c_char_p]),
    ("gdp_solver_get", C.c_int, [_vp, C.c_char_p, _dp]),
    ("gdp_solver_destroy", C.c_int, [_vp]),
    ("gdp_solver_info", C.c_int, [_vp, C.POINTER(SolveInfo)]),
    ("gdp_vector_create", C.c_int, [C.c_int, _pp]),
    ("gdp_vector_destroy", C.c_int, [_vp]),
    ("gdp_vector_zero", C.c_int, [_vp]),
    ("gdp_vector_size", C.c_int, [_vp, _ip]),
    ("gdp_vector_set", C.c_int, [_vp, _dp, C.c_int]),
    ("gdp_vector_get", C.c_int, [_vp, _dp, C.c_int]),
    ("gdp_vector_add_double", C.c_int, [_vp, C.c_int, C.c_double]),
    ("gdp_vector_assemble", C.c_int, [_vp]),
    ("gdp_vector_norm2", C.c_int, [_vp, _dp]),
    ("gdp_vector_copy", C.c_int, [_vp, _vp]),
    ("gdp_vector_device_ptr", C.c_int, [_vp, C.POINTER(_dp)]),
    ("gdp_matrix_create", C.c_int, [C.c_int, C.c_int, _pp]),
    ("gdp_matrix_destroy", C.c_int, [_vp]),
    ("gdp_matrix_zero", C.c_int, [_vp]),
    ("gdp_matrix_add_double", C.c_int, [_vp, C.c_int, C.c_int, C.c_double]),
    ("gdp_matrix_add_coo", C.c_int, [_vp, C.c_long, _ip, _ip, _dp]),
    ("gdp_matrix_assemble", C.c_int, [_vp]),
    ("gdp_matrix_set_csr", C.c_int, [_vp, C.c_int, _ip, _ip, _dp]),
    ("gdp_matrix_nnz", C.c_int, [_vp, _lp]),
    ("gdp_matrix_get_csr", C.c_int, [_vp, _ip, _ip, _dp]),
    ("gdp_prod_matrix_vector", C.c_int, [_vp, _vp, _vp]),
    ("gdp_bench_spmv", C.c_int, [_vp, _vp, _vp, C.c_int, _dp]),
    ("gdp_matrix_sell_info", C.c_int, [_vp, C.POINTER(C.c_long), _ip]),
    ("gdp_solver_bench_spmv", C.c_int, [_vp, C.c_int, _dp, _ip]),
    ("gdp_solver_prec_apply", C.c_int, [_vp, _vp, _vp, C.c_int, _dp]),
    ("gdp_solve", C.c_int, [_vp, _vp, _vp, _vp]),
    ("gdp_solve_again", C.c_int, [_vp, _vp, _vp, _vp]),
    ("gdp_fe_create", C.c_int, [C.POINTER(MeshDesc), C.c_int, C.POINTER(SpaceDesc), C.c_int,
                                C.POINTER(TermDesc), _pp]),
    ("gdp_fe_destroy", C.c_int, [_vp]),
    ("gdp_fe_num_dofs", C.c_int, [_vp, _ip, _lp]),
    ("gdp_fe_dof_table", C.c_int, [_vp, C.c_int, _ip, C.c_int]),
    ("gdp_fe_pattern", C.c_int, [_vp, _ip, _ip]),
    ("gdp_fe_write_pre", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _ip]),
    ("gdp_fe_generate", C.c_int, [_vp, _vp, _vp]),
    ("gdp_fe_last_timing", C.c_int, [_vp, _dp, _dp, _dp]),
    ("gdp_mesh_read", C.c_int, [C.c_char_p, _pp]),
    ("gdp_mesh_info", C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    ("gdp_mesh_get", C.c_int, [_vp, _dp, _ip, _ip, _ip, _ip, _ip, _ip]),
    ("gdp_mesh_free", C.c_int, [_vp]),
    ("gdp_mesh_write", C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, C.c_int, _ip, _ip, _ip]),
    ("gdp_fe_post", C.c_int, [_vp, _vp, C.POINTER(PostDesc), _ip, _dp, C.c_long, _dp, _dp]),
    ("gdp_number_edges", C.c_int, [C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip, C.c_int]),
    ("gdp_rcmk", C.c_int, [C.c_int, _ip, _ip, _ip, _ip]),
    ("gdp_vector_add", C.c_int, [_vp, _vp, _vp]),
    ("gdp_vector_add_scaled", C.c_int, [_vp, _vp, C.c_double, _vp]),
    ("gdp_vector_sub", C.c_int, [_vp, _vp, _vp]),
    ("gdp_vector_scale", C.c_int, [_vp, C.c_double, _vp]),
    ("gdp_vector_swap", C.c_int, [_vp, _vp]),
    ("gdp_vector_set_all", C.c_int, [_vp, C.c_double]),
    ("gdp_vector_dot", C.c_int, [_vp, _vp, _dp]),
    ("gdp_vector_dot_noalloc", C.c_int, [_vp, _vp, _dp]),
    ("gdp_vector_norm_inf", C.c_int, [_vp, _dp]),
    ("gdp_vector_dummy", C.c_int, [_vp, _ip]),
    ("gdp_matrix_add", C.c_int, [_vp, _vp, _vp]),
    ("gdp_matrix_add_scaled", C.c_int, [_vp, _vp, C.c_double, _vp]),
    ("gdp_matrix_scale", C.c_int, [_vp, C.c_double, _vp]),
    ("gdp_matrix_size", C.c_int, [_vp, _ip, _ip]),
    ("gdp_vector_print", C.c_int, [_vp, _vp]),
    ("gdp_vector_write", C.c_int, [_vp, _vp]),
    ("gdp_vector_scan", C.c_int, [_vp, _vp]),
    ("gdp_vector_read", C.c_int, [_vp, _vp]),
    ("gdp_res_write", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _vp]),
    ("gdp_res_read", C.c_int, [C.c_char_p, C.c_int, C.c_int, _vp, _dp, _dp, _ip, _ip]),
    ("gdp_ilut", C.c_int, [C.c_int, _ip, _ip, _dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                           C.c_long, _ip, _ip, _ip, _dp, _ip, _lp]),
    ("gdp_batch_create", C.c_int, [_vp, C.c_int, _ip, _ip, _pp]),
    ("gdp_batch_destroy", C.c_int, [_vp]),
    ("gdp_batch_run", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("gdp_batch_info", C.c_int, [_vp, _ip, _ip, _dp, _dp, _dp, _dp, _dp]),
    ("gdp_unet_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint, _pp]),
    ("gdp_unet_create_3d", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, _pp]),
    ("gdp_unet_destroy", C.c_int, [_vp]),
    ("gdp_unet_num_params", C.c_int, [_vp, _lp]),
    ("gdp_unet_get_params", C.c_int, [_vp, _fp]),
    ("gdp_unet_set_params", C.c_int, [_vp, _fp]),
    ("gdp_unet_get_grads", C.c_int, [_vp, _fp]),
    ("gdp_unet_forward", C.c_int, [_vp, C.c_int, _fp, _fp, _fp]),
    ("gdp_unet_train_step", C.c_int, [_vp, C.c_int, _fp, _fp, _fp, _fp, C.c_float, _fp]),
    ("gdp_unet_get_activation", C.c_int, [_vp, C.c_int, C.c_int, _fp, _lp]),
    ("gdp_unet_attach", C.c_int, [_vp, _vp, C.c_int, _ip, _fp, _dp]),
    ("gdp_unet_detach", C.c_int, [_vp]),
    ("gdp_unet_set_mode", C.c_int, [_vp, C.c_int]),
]

_LIB = None


class GdpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load libgdp_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `python -m getdp_amd.build` first")
        L = C.CDLL(str(LIB_PATH))
        missing = []
        for name, res, args in SIGNATURES:
            try:
                f = getattr(L, name)
            except AttributeError:
                missing.append(name)
                continue
            f.restype = res
            f.argtypes = args
        L.missing_symbols = missing
        _LIB = L
    return _LIB


def check(code: int, what: str = ""):
    if code < 0:
        msg = lib().gdp_last_error().decode(errors="replace")
        raise GdpError(code, f"{what}: {msg}" if what else msg)
    return code


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_ip)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_fp)
