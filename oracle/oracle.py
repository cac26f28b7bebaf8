"""numpy front end of the CPU oracle (``oracle/elem_oracle.c``).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import this module, and only as the
checker (or the timed CPU baseline) -- never as the product path.

All matrices are column-major (Fortran-ordered) float64 arrays, like
Elemental's ``Matrix<double>`` (``include/elemental/core/matrix.hpp:16-183``,
element (i, j) at ``i + j*ldim``).  Functions work in place where the
reference does and return LAPACK-style ``info`` for the factorization.
"""
from __future__ import annotations

import ctypes
import glob
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libelem_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "-B", "libelem_oracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"oracle library {_LIB_PATH} is not built; run `make -C oracle` "
                "or __graft_entry__.build()")
        L = ctypes.CDLL(_LIB_PATH)
        i, d, c, vp = ctypes.c_int, ctypes.c_double, ctypes.c_char, ctypes.c_void_p
        i64, u64 = ctypes.c_int64, ctypes.c_uint64
        L.eo_splitmix64.argtypes = [ctypes.POINTER(u64)]
        L.eo_splitmix64.restype = u64
        L.eo_fill_normal.argtypes = [vp, i64, u64, i64]
        L.eo_dgemm.argtypes = [c, c, i, i, i, d, vp, i, vp, i, d, vp, i]
        L.eo_dsyrk.argtypes = [c, c, i, i, d, vp, i, d, vp, i]
        L.eo_dtrsm.argtypes = [c, c, c, c, i, i, d, vp, i, vp, i]
        L.eo_dtrsv.argtypes = [c, c, c, i, vp, i, vp]
        L.eo_lvar3unb.argtypes = [i, vp, i]
        L.eo_lvar3unb.restype = i
        L.eo_lvar3.argtypes = [i, vp, i, i]
        L.eo_lvar3.restype = i
        L.eo_lvar3_hostblas.argtypes = [i, vp, i, i]
        L.eo_lvar3_hostblas.restype = i
        L.eo_solve_after.argtypes = [i, i, vp, i, vp, i]
        L.eo_solve_after_hostblas.argtypes = [i, i, vp, i, vp, i]
        L.eo_hpd_solve.argtypes = [i, i, vp, i, vp, i, i]
        L.eo_hpd_solve.restype = i
        L.eo_hpd_solve_hostblas.argtypes = [i, i, vp, i, vp, i, i]
        L.eo_hpd_solve_hostblas.restype = i
        L.eo_scale_trapezoid_lower.argtypes = [d, i, i, vp, i]
        L.eo_axpy_triangle_lower.argtypes = [i, d, vp, i, vp, i]
        L.eo_trrk_tn.argtypes = [i, i, d, vp, i, vp, i, d, vp, i, i]
        L.eo_sym_lower_residual.argtypes = [i, i, vp, i, vp, i, vp, i, vp, i]
        L.eo_uvar3unb.argtypes = [i, vp, i]
        L.eo_uvar3unb.restype = i
        L.eo_uvar3.argtypes = [i, vp, i, i]
        L.eo_uvar3.restype = i
        L.eo_solve_after_upper.argtypes = [i, i, vp, i, vp, i]
        L.eo_hpd_solve_upper.argtypes = [i, i, vp, i, vp, i, i]
        L.eo_hpd_solve_upper.restype = i
        L.eo_trtri_lvar3unb.argtypes = [i, i, vp, i]
        L.eo_trtri_lvar3.argtypes = [i, i, vp, i, i]
        L.eo_trtrmm_lunb.argtypes = [i, vp, i]
        L.eo_hpd_inverse_lvar2.argtypes = [i, vp, i, i]
        L.eo_hpd_inverse_lvar2.restype = i
        L.eo_two_sided_trsm_lunb.argtypes = [i, vp, i, vp, i]
        L.eo_two_sided_trmm_lunb.argtypes = [i, vp, i, vp, i]
        L.eo_bind_blas.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, i]
        L.eo_bind_blas.restype = i
        L.eo_blas_bound.restype = i
        _lib = L
    return _lib


def _f(a: np.ndarray) -> np.ndarray:
    if a.dtype != np.float64 or not a.flags.f_contiguous:
        raise TypeError("oracle expects Fortran-ordered float64 arrays")
    return a


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _ld(a: np.ndarray) -> int:
    return max(1, a.shape[0])


# ---------------------------------------------------------------- generator
def splitmix64_sequence(seed: int, count: int) -> list[int]:
    s = ctypes.c_uint64(seed)
    return [lib().eo_splitmix64(ctypes.byref(s)) for _ in range(count)]


def normal(shape, seed: int, offset: int = 0) -> np.ndarray:
    """Column-major N(0,1) array; entry t (column-major order) is number
    ``offset + t`` of the splitmix64/Box-Muller stream ``seed``."""
    shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
    out = np.empty(shape, dtype=np.float64, order="F")
    lib().eo_fill_normal(_p(out), out.size, seed, offset)
    return out


# ------------------------------------------------------- reference BLAS loops
def dgemm(ta, tb, alpha, A, B, beta, C):
    _f(A), _f(B), _f(C)
    m, n = C.shape
    k = A.shape[0] if ta.upper() == "T" else A.shape[1]
    lib().eo_dgemm(ta.encode(), tb.encode(), m, n, k, alpha, _p(A), _ld(A), _p(B), _ld(B),
                   beta, _p(C), _ld(C))
    return C


def dsyrk(trans, alpha, A, beta, C, uplo="L"):
    _f(A), _f(C)
    n = C.shape[0]
    k = A.shape[1] if trans.upper() == "N" else A.shape[0]
    lib().eo_dsyrk(uplo.encode(), trans.encode(), n, k, alpha, _p(A), _ld(A), beta, _p(C), _ld(C))
    return C


def dtrsm(side, transa, alpha, L, B, diag="N", uplo="L"):
    _f(L), _f(B)
    m, n = B.shape
    lib().eo_dtrsm(side.encode(), uplo.encode(), transa.encode(), diag.encode(), m, n, alpha,
                   _p(L), _ld(L), _p(B), _ld(B))
    return B


def dtrsv(trans, L, x, diag="N", uplo="L"):
    _f(L), _f(x)
    lib().eo_dtrsv(uplo.encode(), trans.encode(), diag.encode(), L.shape[0], _p(L), _ld(L), _p(x))
    return x


# ----------------------------------------------------------- Elemental path
def lvar3unb(A) -> int:
    _f(A)
    return lib().eo_lvar3unb(A.shape[0], _p(A), _ld(A))


def lvar3(A, nb: int = 128, host_blas: bool = False) -> int:
    _f(A)
    fn = lib().eo_lvar3_hostblas if host_blas else lib().eo_lvar3
    if host_blas and not lib().eo_blas_bound():
        raise RuntimeError("host BLAS not bound; call bind_host_blas() first")
    return fn(A.shape[0], _p(A), _ld(A), nb)


def solve_after(L, B, host_blas: bool = False):
    _f(L), _f(B)
    fn = lib().eo_solve_after_hostblas if host_blas else lib().eo_solve_after
    k = B.shape[1] if B.ndim == 2 else 1
    fn(L.shape[0], k, _p(L), _ld(L), _p(B), _ld(B))
    return B


def hpd_solve(A, B, nb: int = 128, host_blas: bool = False) -> int:
    _f(A), _f(B)
    fn = lib().eo_hpd_solve_hostblas if host_blas else lib().eo_hpd_solve
    k = B.shape[1] if B.ndim == 2 else 1
    return fn(A.shape[0], k, _p(A), _ld(A), _p(B), _ld(B), nb)


def uvar3unb(A) -> int:
    _f(A)
    return lib().eo_uvar3unb(A.shape[0], _p(A), _ld(A))


def uvar3(A, nb: int = 128) -> int:
    _f(A)
    return lib().eo_uvar3(A.shape[0], _p(A), _ld(A), nb)


def solve_after_upper(U, B):
    _f(U), _f(B)
    k = B.shape[1] if B.ndim == 2 else 1
    lib().eo_solve_after_upper(U.shape[0], k, _p(U), _ld(U), _p(B), _ld(B))
    return B


def hpd_solve_upper(A, B, nb: int = 128) -> int:
    _f(A), _f(B)
    k = B.shape[1] if B.ndim == 2 else 1
    return lib().eo_hpd_solve_upper(A.shape[0], k, _p(A), _ld(A), _p(B), _ld(B), nb)


def trtri_lvar3unb(L, unit: bool = False):
    """triangular_inverse::LVar3Unb on the lower triangle, in place."""
    _f(L)
    lib().eo_trtri_lvar3unb(int(unit), L.shape[0], _p(L), _ld(L))
    return L


def trtri(L, unit: bool = False, nb: int = 128):
    """triangular_inverse::LVar3 on the lower triangle, in place."""
    _f(L)
    lib().eo_trtri_lvar3(int(unit), L.shape[0], _p(L), _ld(L), nb)
    return L


def trtrmm_lunb(L):
    """TrtrmmLUnblocked(TRANSPOSE): lower triangle of L := L^T L, in place."""
    _f(L)
    lib().eo_trtrmm_lunb(L.shape[0], _p(L), _ld(L))
    return L


def hpd_inverse(A, nb: int = 128) -> int:
    """hpd_inverse::CholeskyLVar2: lower triangle of A := inv(A), in place;
    returns 0 or the 1-based failing column."""
    _f(A)
    return lib().eo_hpd_inverse_lvar2(A.shape[0], _p(A), _ld(A), nb)


def two_sided_trsm(A, L):
    """TwoSidedTrsmLUnb: lower triangle of A := inv(L) A inv(L)^T, in place."""
    _f(A), _f(L)
    lib().eo_two_sided_trsm_lunb(A.shape[0], _p(A), _ld(A), _p(L), _ld(L))
    return A


def two_sided_trmm(A, L):
    """TwoSidedTrmmLUnb: lower triangle of A := L^T A L, in place."""
    _f(A), _f(L)
    lib().eo_two_sided_trmm_lunb(A.shape[0], _p(A), _ld(A), _p(L), _ld(L))
    return A


def hermitian_gen_definite_eig(gtype: int, A, B, nb: int = 128):
    """elem::HermitianGenDefiniteEig(type, LOWER, A, B, w, X)
    (HermitianGenDefiniteEig.hpp:60-86) restated: Cholesky(LOWER, B) (LVar3),
    TwoSidedTrsm (type 1) or TwoSidedTrmm (types 2, 3) in their unblocked
    variant-4 loops, the standard eigenproblem of the reduced lower triangle,
    and the back-transformation (Trsm with L^T, or L Z for type 3).  The
    standard eigenproblem is LAPACK's (numpy.linalg.eigh, dsyevd); the
    reference calls LAPACK's dsyevr there (src/lapack-like/HermitianEig.cpp:
    224-240), which agrees to rounding.  Returns (info, w, X, reduced A)."""
    A = np.array(A, dtype=np.float64, order="F")
    L = np.array(B, dtype=np.float64, order="F")
    info = lvar3(L, nb)
    if info:
        return info, None, None, A
    if gtype == 1:
        two_sided_trsm(A, L)
    else:
        two_sided_trmm(A, L)
    S = np.tril(A) + np.tril(A, -1).T
    w, Z = np.linalg.eigh(S)
    Z = np.asfortranarray(Z)
    Lt = np.tril(L)
    if gtype == 3:
        X = Lt @ Z
    else:
        X = dtrsm("L", "T", 1.0, np.asfortranarray(Lt), Z)
    return 0, w, X, A


def scale_trapezoid_lower(alpha, A):
    _f(A)
    lib().eo_scale_trapezoid_lower(alpha, A.shape[0], A.shape[1], _p(A), _ld(A))
    return A


def axpy_triangle_lower(alpha, X, Y):
    _f(X), _f(Y)
    lib().eo_axpy_triangle_lower(X.shape[0], alpha, _p(X), _ld(X), _p(Y), _ld(Y))
    return Y


def trrk_tn(alpha, A, B, beta, C, bsize: int = 64):
    """Lower trapezoid of C := alpha*A^T*B + beta*C by Elemental's recursive
    halving (TrrkTN / LocalTrrk on a 1x1 grid)."""
    _f(A), _f(B), _f(C)
    k, n = A.shape
    lib().eo_trrk_tn(k, n, alpha, _p(A), _ld(A), _p(B), _ld(B), beta, _p(C), _ld(C), bsize)
    return C


def sym_lower_residual(A, X, B) -> np.ndarray:
    _f(A), _f(X), _f(B)
    R = np.empty_like(B, order="F")
    k = B.shape[1] if B.ndim == 2 else 1
    lib().eo_sym_lower_residual(A.shape[0], k, _p(A), _ld(A), _p(X), _ld(X), _p(B), _ld(B),
                                _p(R), _ld(R))
    return R


# ------------------------------------------------------- host BLAS (baseline)
def find_host_blas():
    """Locate an optimised LP64 host BLAS: scipy's bundled OpenBLAS (symbols
    prefixed ``scipy_``).  Returns (path, prefix, set_threads_symbol, name)."""
    import scipy

    libs = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
    cands = sorted(glob.glob(os.path.join(libs, "libscipy_openblas*.so")))
    if cands:
        return cands[0], "scipy_", "scipy_openblas_set_num_threads", "OpenBLAS (scipy bundle)"
    raise RuntimeError("no host BLAS found")


def bind_host_blas(nthreads: int) -> str:
    path, prefix, setnt, name = find_host_blas()
    rc = lib().eo_bind_blas(path.encode(), prefix.encode(), setnt.encode(), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"binding host BLAS {path} failed (rc={rc})")
    return f"{name}: {os.path.basename(path)}"


def host_blas_info() -> dict:
    """Version string and threading layer of the bound host BLAS (OpenBLAS's
    own get_config / get_parallel), for the CPU baseline's record."""
    path, prefix, _, name = find_host_blas()
    L = ctypes.CDLL(path)
    info = {"blas": name, "library": os.path.basename(path)}
    try:
        f = getattr(L, prefix + "openblas_get_config")
        f.restype = ctypes.c_char_p
        info["config"] = f().decode()
        g = getattr(L, prefix + "openblas_get_parallel")
        info["threading"] = {0: "sequential", 1: "pthreads", 2: "OpenMP"}.get(int(g()), "unknown")
    except AttributeError:
        pass
    return info
