"""elemental_amd -- MI355X (gfx950) HPD Cholesky factor-and-solve.

Python mirror of Elemental's interface for the lower Cholesky / HPDSolve path
(``include/elemental/lapack-like/Cholesky.hpp``, ``HPDSolve.hpp``,
``Cholesky/SolveAfter.hpp``, ``blas-like/level3/Trsm.hpp``, ``Herk.hpp``,
``src/blas-like/Trrk.cpp``).  Every call goes through the C ABI of
``libelemental_amd.so`` (declared in ``include/elemental_amd.h``); there is
no CPU fallback: if the library is missing, importing the compute entry
points raises.

Matrices are :class:`Matrix` objects: column-major device buffers with a
leading dimension, like ``elem::Matrix<F>``.  PyTorch only provides the
device memory and the stream.
"""
from __future__ import annotations

import ctypes
import enum
import os

from . import _lib
from ._lib import LogicError, SingularMatrixError, lib  # noqa: F401
from .matrix import Matrix, View  # noqa: F401

__all__ = [
    "UpperOrLower", "Orientation", "LeftOrRight", "UnitOrNonUnit", "LOWER", "UPPER", "NORMAL",
    "TRANSPOSE", "ADJOINT", "LEFT", "RIGHT", "NON_UNIT", "UNIT", "Matrix", "View", "Blocksize",
    "Cholesky", "LocalCholesky", "HPDSolve", "SolveAfter", "Trsm", "Herk", "Syrk", "Trrk", "Trsv",
    "HPDSolveMixed", "LogicError", "SingularMatrixError", "Gemm", "HPDInverse", "TriangularInverse",
    "HermitianGenDefiniteEig", "HermitianGenDefiniteEigType", "AXBX", "ABX", "BAX",
]


class UpperOrLower(enum.Enum):  # include/elemental/core/types_decl.hpp:152-156
    LOWER = "L"
    UPPER = "U"


class Orientation(enum.Enum):  # types_decl.hpp:129-134
    NORMAL = "N"
    TRANSPOSE = "T"
    ADJOINT = "C"


class LeftOrRight(enum.Enum):  # types_decl.hpp:104-108
    LEFT = "L"
    RIGHT = "R"


class UnitOrNonUnit(enum.Enum):  # types_decl.hpp:141-145
    NON_UNIT = "N"
    UNIT = "U"


LOWER, UPPER = UpperOrLower.LOWER, UpperOrLower.UPPER
NORMAL, TRANSPOSE, ADJOINT = Orientation.NORMAL, Orientation.TRANSPOSE, Orientation.ADJOINT
LEFT, RIGHT = LeftOrRight.LEFT, LeftOrRight.RIGHT
NON_UNIT, UNIT = UnitOrNonUnit.NON_UNIT, UnitOrNonUnit.UNIT

_DEFAULT_BLOCKSIZE = 128  # src/core/global.cpp:194


def Blocksize() -> int:
    """Elemental's default algorithmic block size (the blocksize stack is
    replaced by an explicit ``nb`` argument on every call)."""
    return _DEFAULT_BLOCKSIZE


def _c(x) -> bytes:
    return ctypes.c_char(x.value.encode())


def _stream(stream=None) -> ctypes.c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def _dinfo():
    import torch

    return torch.zeros(1, dtype=torch.int32, device="cuda")


def _raise_info(info: int):
    if info > 0:
        # cholesky::LVar3Unb throws this (LVar3.hpp:37-38) at column info-1
        raise LogicError("A was not numerically HPD", column=info - 1)
    if info < 0:
        raise RuntimeError(f"reading the factorization status failed ({info})")


def _is_f32(A: Matrix) -> bool:
    import torch

    if A.dtype == torch.float32:
        return True
    if A.dtype != torch.float64:
        raise TypeError(f"Matrix<{A.dtype}> is not on this path (float64 and float32 are)")
    return False


def Cholesky(uplo: UpperOrLower, A: Matrix, nb: int | None = None, stream=None) -> None:
    """elem::Cholesky(uplo, Matrix<F>& A) (Cholesky.hpp:49-62): LOWER runs
    cholesky::LVar3 (A's lower triangle becomes L, the upper one is
    untouched), UPPER cholesky::UVar3 (the upper triangle becomes U with
    A = U^T U, the lower one is untouched).  F = float64 or float32.  Raises
    LogicError ("A was not numerically HPD") like the reference."""
    if A.Height() != A.Width():
        raise LogicError("A must be square")
    nb = nb or _DEFAULT_BLOCKSIZE
    d = _dinfo()
    s = _stream(stream)
    fn = lib().elemamd_cholesky_f32 if _is_f32(A) else lib().elemamd_cholesky_f64
    _check(fn(_c(uplo), A.Height(), A.ptr(), A.LDim(), nb, ctypes.c_void_p(d.data_ptr()), s), "elemamd_cholesky")
    _raise_info(lib().elemamd_info_read(ctypes.c_void_p(d.data_ptr()), s))


def LocalCholesky(uplo: UpperOrLower, A: Matrix, nb: int | None = None, stream=None) -> None:
    """elem::LocalCholesky (Cholesky.hpp:29-37): Cholesky on the local
    (replicated) matrix; here the local matrix is the whole Matrix."""
    Cholesky(uplo, A, nb, stream)


def SolveAfter(uplo: UpperOrLower, orientation: Orientation, A: Matrix, B: Matrix, stream=None) -> None:
    """cholesky::SolveAfter (SolveAfter.hpp:20-75): B := A^{-1} B from the
    factor stored in A's uplo triangle."""
    if A.Height() != A.Width():
        raise LogicError("A must be square")
    if A.Height() != B.Height():
        raise LogicError("A and B must be the same height")
    f32 = _is_f32(A)
    if f32 != _is_f32(B):
        raise TypeError("A and B must have the same element type")
    fn = lib().elemamd_solve_after_f32 if f32 else lib().elemamd_solve_after_f64
    _check(fn(_c(uplo), _c(orientation), A.Height(), B.Width(), A.ptr(), A.LDim(), B.ptr(), B.LDim(),
              _stream(stream)), "elemamd_solve_after")


def HPDSolve(uplo: UpperOrLower, orientation: Orientation, A: Matrix, B: Matrix, nb: int | None = None,
             stream=None) -> None:
    """elem::HPDSolve (HPDSolve.hpp:18-28): Cholesky(uplo, A) then
    SolveAfter; A is overwritten by its factor and B by the solution."""
    if A.Height() != A.Width():
        raise LogicError("A must be square")
    if A.Height() != B.Height():
        raise LogicError("A and B must be the same height")
    f32 = _is_f32(A)
    if f32 != _is_f32(B):
        raise TypeError("A and B must have the same element type")
    nb = nb or _DEFAULT_BLOCKSIZE
    d = _dinfo()
    s = _stream(stream)
    fn = lib().elemamd_hpd_solve_f32 if f32 else lib().elemamd_hpd_solve_f64
    _check(fn(_c(uplo), _c(orientation), A.Height(), B.Width(), A.ptr(), A.LDim(), B.ptr(), B.LDim(), nb,
              ctypes.c_void_p(d.data_ptr()), s), "elemamd_hpd_solve")
    _raise_info(lib().elemamd_info_read(ctypes.c_void_p(d.data_ptr()), s))


def Trsm(side: LeftOrRight, uplo: UpperOrLower, orientation: Orientation, diag: UnitOrNonUnit, alpha: float,
         A: Matrix, B: Matrix, checkIfSingular: bool = False, stream=None) -> None:
    """elem::Trsm(Matrix) (Trsm.hpp:49-86)."""
    if A.Height() != A.Width():
        raise LogicError("Triangular matrix must be square")
    if (side is LEFT and A.Height() != B.Height()) or (side is RIGHT and A.Height() != B.Width()):
        raise LogicError("Nonconformal Trsm")
    if checkIfSingular and diag is not UNIT:
        if bool((A.diagonal() == 0).any()):
            raise SingularMatrixError()
    _check(lib().elemamd_trsm_f64(_c(side), _c(uplo), _c(orientation), _c(diag), B.Height(), B.Width(),
                                  float(alpha), A.ptr(), A.LDim(), B.ptr(), B.LDim(), _stream(stream)),
           "elemamd_trsm_f64")


def Herk(uplo: UpperOrLower, orientation: Orientation, alpha: float, A: Matrix, beta: float, C: Matrix,
         stream=None) -> None:
    """elem::Herk(Matrix) (Herk.hpp:17-27): C := alpha op(A) op(A)^H + beta C
    on the uplo triangle (real data: Syrk); op(A) = A (NORMAL, A n x k) or
    A^T (A k x n)."""
    n, k = (A.Height(), A.Width()) if orientation is NORMAL else (A.Width(), A.Height())
    if n != C.Height() or C.Height() != C.Width():
        raise LogicError("Nonconformal Syrk")
    _check(lib().elemamd_herk_f64(_c(uplo), _c(orientation), n, k, float(alpha), A.ptr(), A.LDim(),
                                  float(beta), C.ptr(), C.LDim(), _stream(stream)), "elemamd_herk_f64")


Syrk = Herk  # identical for real data (Syrk.hpp:20-57 with conjugate=false)


def Trrk(uplo: UpperOrLower, orientationOfA: Orientation, orientationOfB: Orientation, alpha: float, A: Matrix,
         B: Matrix, beta: float, C: Matrix, stream=None) -> None:
    """elem::Trrk(Matrix) (src/blas-like/Trrk.cpp:19-38): the uplo triangle
    of C := alpha op(A) op(B) + beta C, op(A) n x k, op(B) k x n."""
    k = A.Height() if orientationOfA is not NORMAL else A.Width()
    _check(lib().elemamd_trrk_f64(_c(uplo), _c(orientationOfA), _c(orientationOfB), C.Height(), k,
                                  float(alpha), A.ptr(), A.LDim(), B.ptr(), B.LDim(), float(beta), C.ptr(), C.LDim(),
                                  _stream(stream)), "elemamd_trrk_f64")


def Trsv(uplo: UpperOrLower, orientation: Orientation, diag: UnitOrNonUnit, A: Matrix, x: Matrix,
         stream=None) -> None:
    """elem::Trsv(Matrix) (level2/Trsv.hpp:20-44) for a column vector x."""
    if x.Width() != 1:
        raise NotImplementedError("x must be a column vector")
    _check(lib().elemamd_trsv_f64(_c(uplo), _c(orientation), _c(diag), A.Height(), A.ptr(), A.LDim(), x.ptr(), 1,
                                  _stream(stream)), "elemamd_trsv_f64")


def _orient(o) -> bytes:
    return ctypes.c_char((o.value if isinstance(o, Orientation) else str(o)).encode())


def Gemm(orientationOfA, orientationOfB, alpha: float, A: Matrix, B: Matrix, beta: float, C: Matrix,
         stream=None) -> None:
    """elem::Gemm(Matrix) (blas-like/level3/Gemm.hpp): C := alpha op(A) op(B)
    + beta C (Orientation values or 'N' / 'T' / 'C')."""
    ta = _orient(orientationOfA).value.decode() != "N"
    k = A.Height() if ta else A.Width()
    _check(lib().elemamd_gemm_f64(_orient(orientationOfA), _orient(orientationOfB), C.Height(), C.Width(), k,
                                  float(alpha), A.ptr(), A.LDim(), B.ptr(), B.LDim(), float(beta), C.ptr(),
                                  C.LDim(), _stream(stream)), "elemamd_gemm_f64")


def HPDInverse(uplo: UpperOrLower, A: Matrix, nb: int | None = None, stream=None) -> None:
    """elem::HPDInverse(uplo, Matrix) (lapack-like/Inverse.hpp:29-40 ->
    hpd_inverse::CholeskyLVar2, Inverse/CholeskyLVar2.hpp:29-75): the uplo
    triangle of A becomes that of inv(A), the other triangle is untouched.
    Raises LogicError ("A was not numerically HPD") like the reference."""
    if A.Height() != A.Width():
        raise LogicError("Nonsquare matrices cannot be triangular")
    nb = nb or _DEFAULT_BLOCKSIZE
    d = _dinfo()
    s = _stream(stream)
    _check(lib().elemamd_hpd_inverse_f64(_c(uplo), A.Height(), A.ptr(), A.LDim(), nb, ctypes.c_void_p(d.data_ptr()),
                                         s), "elemamd_hpd_inverse_f64")
    _raise_info(lib().elemamd_info_read(ctypes.c_void_p(d.data_ptr()), s))


def TriangularInverse(uplo: UpperOrLower, diag: UnitOrNonUnit, A: Matrix, nb: int | None = None,
                      stream=None) -> None:
    """elem::TriangularInverse(uplo, diag, Matrix) (lapack-like/
    TriangularInverse.hpp -> triangular_inverse::LVar3 / UVar3): A := inv(A)
    on the uplo triangle."""
    if A.Height() != A.Width():
        raise LogicError("Nonsquare matrices cannot be triangular")
    _check(lib().elemamd_triangular_inverse_f64(_c(uplo), _c(diag), A.Height(), A.ptr(), A.LDim(),
                                                nb or _DEFAULT_BLOCKSIZE, _stream(stream)),
           "elemamd_triangular_inverse_f64")


class HermitianGenDefiniteEigType(enum.Enum):  # lapack-like_decl.hpp: AXBX, ABX, BAX
    AXBX = 1
    ABX = 2
    BAX = 3


AXBX, ABX, BAX = HermitianGenDefiniteEigType.AXBX, HermitianGenDefiniteEigType.ABX, HermitianGenDefiniteEigType.BAX


def HermitianGenDefiniteEig(type: HermitianGenDefiniteEigType, uplo: UpperOrLower, A: Matrix, B: Matrix,
                            vectors: bool = False, lowerBound=None, upperBound=None, stream=None):
    """elem::HermitianGenDefiniteEig(type, uplo, A, B, w[, X][, a, b])
    (lapack-like/HermitianGenDefiniteEig.hpp): Cholesky(uplo, B), the
    two-sided reduction of A, the standard eigenproblem and the
    back-transformation.  Integer bounds select eigenvalues a..b (0-based,
    inclusive), float bounds the eigenvalues in (a, b].  Returns (w, X): w a
    nev x 1 Matrix of ascending eigenvalues, X the n x nev eigenvectors (None
    unless ``vectors``).  B not numerically HPD raises LogicError."""
    import torch

    n = A.Height()
    if A.Width() != n or B.Height() != B.Width() or B.Height() != n:
        raise LogicError("Hermitian matrices must be square.")
    rng, il, iu, vl, vu = 0, 0, 0, 0.0, 0.0
    if lowerBound is not None:
        if isinstance(lowerBound, int) and isinstance(upperBound, int):
            rng, il, iu = 1, lowerBound, upperBound
        else:
            rng, vl, vu = 2, float(lowerBound), float(upperBound)
    wbuf = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    X = Matrix(n, n) if vectors else None
    nev, info = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().elemamd_hermitian_gen_definite_eig_f64(
        type.value, _c(uplo), n, A.ptr(), A.LDim(), B.ptr(), B.LDim(), ctypes.c_void_p(wbuf.data_ptr()),
        X.ptr() if X is not None else None, X.LDim() if X is not None else 1, rng, il, iu, vl, vu,
        ctypes.byref(nev), ctypes.byref(info), _stream(stream)), "elemamd_hermitian_gen_definite_eig_f64")
    if info.value > 0:
        raise LogicError("A was not numerically HPD", column=info.value - 1)
    k = nev.value
    w = Matrix(k, 1)
    if k > 0:
        w.tensor()[:, 0].copy_(wbuf[:k])
    if X is not None:
        X = View(X, 0, 0, n, k)
    return w, X


from .mixed import HPDSolveMixed, IROptions, IRReport  # noqa: E402,F401
