"""ctypes binding of ``libelemental_amd.so`` (the C ABI in
``include/elemental_amd.h``).  Loading fails loudly when the library has not
been built: there is no fallback path."""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libelemental_amd.so")
_lib = None


class LogicError(RuntimeError):
    """Counterpart of the std::logic_error the reference throws (for example
    "A was not numerically HPD", LVar3.hpp:37-38).  ``column`` is the 0-based
    failing column when known."""

    def __init__(self, msg: str, column: int | None = None):
        super().__init__(msg)
        self.column = column


class SingularMatrixError(RuntimeError):
    """Counterpart of elem::SingularMatrixException (Trsm.hpp:76-82)."""

    def __init__(self):
        super().__init__("Matrix was singular")


class IROpts(ctypes.Structure):
    _fields_ = [
        ("nb", ctypes.c_int),
        ("trailing", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("gmres_restart", ctypes.c_int),
        ("force_gmres", ctypes.c_int),
        ("allow_shift", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("refine", ctypes.c_int),
        ("fallback", ctypes.c_int),
    ]


class IRReportC(ctypes.Structure):
    _fields_ = [
        ("info", ctypes.c_int),
        ("lowprec_info", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("gmres_iterations", ctypes.c_int),
        ("used_gmres", ctypes.c_int),
        ("shifted", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("backward_error", ctypes.c_double),
        ("anorm_inf", ctypes.c_double),
        ("fallback", ctypes.c_int),
    ]


class ServerConfig(ctypes.Structure):  # elemamd_server_config
    _fields_ = [("slots", ctypes.c_int), ("precision", ctypes.c_int), ("nb", ctypes.c_int), ("tol", ctypes.c_double),
                ("memory_budget", ctypes.c_size_t)]


class ServeResult(ctypes.Structure):  # elemamd_serve_result
    _fields_ = [("ticket", ctypes.c_int64), ("status", ctypes.c_int), ("info", ctypes.c_int),
                ("factored", ctypes.c_int), ("iterations", ctypes.c_int), ("gmres_iterations", ctypes.c_int),
                ("slot", ctypes.c_int), ("backward_error", ctypes.c_double), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_done", ctypes.c_double)]


class ServerStats(ctypes.Structure):  # elemamd_server_statistics
    _fields_ = [("completed", ctypes.c_int64), ("reuse_hits", ctypes.c_int64), ("evictions", ctypes.c_int64),
                ("peak_reserved", ctypes.c_size_t), ("reserved", ctypes.c_size_t), ("budget", ctypes.c_size_t),
                ("resident_factors", ctypes.c_int)]


# name -> (restype, argtypes); the symbol list the header declares
_i, _d, _c, _vp, _i64, _u64, _sz = (ctypes.c_int, ctypes.c_double, ctypes.c_char, ctypes.c_void_p,
                                    ctypes.c_int64, ctypes.c_uint64, ctypes.c_size_t)
SIGNATURES = {
    "elemamd_version": (ctypes.c_char_p, []),
    "elemamd_info_read": (_i, [_vp, _vp]),
    "elemamd_potrf_unb_f64": (_i, [_i, _vp, _i64, _vp, _vp]),
    "elemamd_trsm_f64": (_i, [_c, _c, _c, _c, _i, _i, _d, _vp, _i64, _vp, _i64, _vp]),
    "elemamd_herk_f64": (_i, [_c, _c, _i, _i, _d, _vp, _i64, _d, _vp, _i64, _vp]),
    "elemamd_herk_f32": (_i, [_i, _i, _i, ctypes.c_float, _vp, _i64, ctypes.c_float, _vp, _i64, _vp]),
    "elemamd_trrk_f64": (_i, [_c, _c, _c, _i, _i, _d, _vp, _i64, _vp, _i64, _d, _vp, _i64, _vp]),
    "elemamd_trsv_f64": (_i, [_c, _c, _c, _i, _vp, _i64, _vp, _i64, _vp]),
    "elemamd_cholesky_f64": (_i, [_c, _i, _vp, _i64, _i, _vp, _vp]),
    "elemamd_solve_after_f64": (_i, [_c, _c, _i, _i, _vp, _i64, _vp, _i64, _vp]),
    "elemamd_hpd_solve_f64": (_i, [_c, _c, _i, _i, _vp, _i64, _vp, _i64, _i, _vp, _vp]),
    "elemamd_solve_after_f64_noalloc_limit": (_sz, []),
    "elemamd_solve_after_f64_noalloc": (_i, [_c, _c, _i, _i, _vp, _i64, _vp, _i64, _vp]),
    "elemamd_gemm_f64": (_i, [_c, _c, _i, _i, _i, _d, _vp, _i64, _vp, _i64, _d, _vp, _i64, _vp]),
    "elemamd_hpd_inverse_f64": (_i, [_c, _i, _vp, _i64, _i, _vp, _vp]),
    "elemamd_triangular_inverse_f64": (_i, [_c, _c, _i, _vp, _i64, _i, _vp]),
    "elemamd_hermitian_gen_definite_eig_f64": (_i, [_i, _c, _i, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i, _i, _i,
                                                    _d, _d, _vp, _vp, _vp]),
    "elemamd_cholesky_f32": (_i, [_c, _i, _vp, _i64, _i, _vp, _vp]),
    "elemamd_solve_after_f32": (_i, [_c, _c, _i, _i, _vp, _i64, _vp, _i64, _vp]),
    "elemamd_hpd_solve_f32": (_i, [_c, _c, _i, _i, _vp, _i64, _vp, _i64, _i, _vp, _vp]),
    "elemamd_hpd_solve_mixed_workspace": (_sz, [_i, _i, ctypes.POINTER(IROpts)]),
    "elemamd_hpd_solve_mixed_f64": (_i, [_i, _i, _vp, _i64, _vp, _i64, ctypes.POINTER(IROpts), _vp, _sz,
                                         ctypes.POINTER(IRReportC), _vp]),
    "elemamd_fill_normal_f64": (_i, [_vp, _i64, _i64, _i64, _u64, _i64, _vp]),
    "elemamd_sym_residual_f64": (_i, [_i, _i, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp]),
    "elemamd_batch_plan_create": (_i, [_i, _vp, _vp, _i, _i, _vp, _vp, _vp]),
    "elemamd_batch_set_launch": (_i, [_vp, _vp]),
    "elemamd_batch_plan_destroy": (_i, [_vp]),
    "elemamd_batch_plan_layout": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "elemamd_batch_set_refinement": (_i, [_vp, _i, _d]),
    "elemamd_batch_workspace": (_sz, [_vp]),
    "elemamd_batch_hpd_solve": (_i, [_vp, _vp, _vp, _vp, _sz, _i, _vp]),
    "elemamd_batch_results": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "elemamd_batch_backward_error": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "elemamd_server_create": (_i, [ctypes.POINTER(ServerConfig), _vp]),
    "elemamd_server_submit": (_i, [_vp, _i, _i, _i, _i, _vp, _i64, _vp, _i64, _vp, _vp]),
    "elemamd_server_wait": (_i, [_vp, _i64, ctypes.POINTER(ServeResult)]),
    "elemamd_server_evict": (_i, [_vp, _i]),
    "elemamd_server_stats": (_i, [_vp, ctypes.POINTER(ServerStats)]),
    "elemamd_server_destroy": (_i, [_vp]),
}


def load(path: str = LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run `python -m elemental_amd.build` (or __graft_entry__.build()) first; "
            "there is no CPU fallback")
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib
