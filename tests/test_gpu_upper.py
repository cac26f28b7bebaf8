"""UPPER, the remaining BLAS forms and Matrix<float>, through the C ABI.

* cholesky::UVar3 (Cholesky/UVar3.hpp:21-125): the upper factor U with
  A = U^T U overwrites the upper triangle, the strictly lower triangle is
  untouched.  UVar3Unb on A is LVar3Unb on A^T operation by operation, so the
  single-block case is bit-exact against the oracle's UVar3Unb; blocked cases
  within 1e-12 of the oracle's UVar3; integer problems exact at any size.
* SolveAfter / HPDSolve with UPPER (SolveAfter.hpp:33-43, 56-72).
* Trsm, Trsv, Herk and Trrk in every side / uplo / orientation / diagonal
  combination the reference accepts (Trsm.hpp:49-86, Trsv.hpp:20-44,
  Herk.hpp:17-27, Trrk.cpp:19-38), against the oracle's Netlib loops.
* Matrix<float>: Cholesky / SolveAfter / HPDSolve with F = float, against
  the FP64 oracle to float accuracy.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

import elemental_amd as E  # noqa: E402
from elemental_amd import Matrix, inputs, lib  # noqa: E402

EPS = np.finfo(np.float64).eps


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def dev(a, ld=None):
    return Matrix.from_numpy(np.asfortranarray(a), ldim=ld)


def host(M):
    torch.cuda.synchronize()
    return M.to_numpy()


# ------------------------------------------------------------------ UVar3
@pytest.mark.parametrize("n", [1, 2, 7, 64, 127, 128])
def test_upper_single_block_bitexact(oracle, n):
    A = inputs.gram_host(n, seed=n + 100)
    ref = A.copy(order="F")
    assert oracle.uvar3unb(ref) == 0
    M = dev(A, ld=n + 2)
    E.Cholesky(E.UPPER, M)
    assert np.array_equal(host(M), ref)  # upper = U bit for bit, strictly lower untouched


@pytest.mark.parametrize("n,nb", [(129, 128), (300, 64), (1000, 128), (2049, 128), (777, 37)])
def test_upper_cholesky_vs_oracle(oracle, n, nb):
    A = inputs.gram_host(n, seed=3 * n + nb)
    ref = A.copy(order="F")
    assert oracle.uvar3(ref, nb) == 0
    M = dev(A, ld=n + 1)
    E.Cholesky(E.UPPER, M, nb=nb)
    got = host(M)
    assert np.linalg.norm(np.triu(got - ref)) / np.linalg.norm(np.triu(ref)) <= 1e-12
    assert np.array_equal(np.tril(got, -1), np.tril(A, -1))


@pytest.mark.parametrize("n,nb", [(1000, 128), (3000, 128)])
def test_upper_integer_exact(n, nb):
    A, L0 = inputs.integer_factor(n, seed=n)
    M = dev(A)
    E.Cholesky(E.UPPER, M, nb=nb)
    got = host(M)
    assert np.array_equal(np.triu(got), L0.T)
    assert np.array_equal(np.tril(got, -1), np.tril(A, -1))


@pytest.mark.parametrize("col", [0, 127, 128, 300])
def test_upper_non_hpd_raises_like_reference(oracle, col):
    n = 400
    A = inputs.gram_host(n, seed=17)
    A[col, col] = -1.0
    ref = A.copy(order="F")
    info = oracle.uvar3(ref, 128)
    assert info == col + 1
    M = dev(A)
    with pytest.raises(E.LogicError) as ei:
        E.Cholesky(E.UPPER, M)
    assert ei.value.column == col
    got = host(M)
    # the blocks before the failing one are factored as in the reference
    done = (col // 128) * 128
    if done:
        assert np.linalg.norm(np.triu(got - ref)[:done, :]) <= 1e-12 * np.linalg.norm(np.triu(ref)[:done, :])
    assert np.array_equal(np.tril(got, -1), np.tril(A, -1))


@pytest.mark.parametrize("n,k", [(1, 1), (300, 1), (1000, 16), (2048, 33)])
def test_upper_solve_after_and_hpd_solve(oracle, n, k):
    A = inputs.gram_host(n, seed=n + k)
    B = inputs.normal_host(n, k, seed=n + 2 * k)
    U = A.copy(order="F")
    assert oracle.uvar3(U, 128) == 0
    Xr = B.copy(order="F")
    oracle.solve_after_upper(U, Xr)
    # SolveAfter(UPPER) with the oracle's factor; the lower triangle of the
    # stored matrix holds other values that must not be read
    Ustore = np.asfortranarray(np.triu(U) + np.tril(np.full((n, n), 7.0), -1))
    Xd = dev(B)
    E.SolveAfter(E.UPPER, E.NORMAL, dev(Ustore), Xd)
    X = host(Xd)
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-12
    # HPDSolve(UPPER)
    Ad, Bd = dev(A), dev(B)
    E.HPDSolve(E.UPPER, E.NORMAL, Ad, Bd)
    Ug, Xg = host(Ad), host(Bd)
    assert np.linalg.norm(np.triu(Ug - U)) / np.linalg.norm(np.triu(U)) <= 1e-12
    assert np.array_equal(np.tril(Ug, -1), np.tril(A, -1))
    R = oracle.sym_lower_residual(A, Xg, B)
    assert np.linalg.norm(R) / np.linalg.norm(B) <= 1e-13
    assert np.linalg.norm(Xg - Xr) / np.linalg.norm(Xr) <= 1e-12


def test_upper_solve_integer_exact():
    n, k = 2500, 32
    A, L0 = inputs.integer_factor(n, seed=9)
    rng = np.random.default_rng(9)
    X0 = np.asfortranarray(rng.integers(-3, 4, (n, k)).astype(np.float64))
    Bd = dev(A @ X0)
    E.SolveAfter(E.UPPER, E.NORMAL, dev(np.asfortranarray(L0.T)), Bd)
    assert np.array_equal(host(Bd), X0)


# ---------------------------------------------------------- BLAS surface
def _tri(n, upper, seed):
    rng = np.random.default_rng(seed)
    T = rng.standard_normal((n, n)) * 0.3 + np.diag(2.0 + rng.random(n))
    return np.asfortranarray(np.triu(T) if upper else np.tril(T))


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm_every_form(oracle, side, uplo, trans, diag):
    m, n = (300, 7) if side == "L" else (9, 200)
    na = m if side == "L" else n
    T = _tri(na, uplo == "U", seed=na)
    # the other triangle holds values that must not be read; a unit diagonal
    # is not read either
    Tstore = T + (np.tril(np.full((na, na), 5.0), -1) if uplo == "U" else np.triu(np.full((na, na), 5.0), 1))
    if diag == "U":
        Tstore[np.diag_indices(na)] = 1e300
    Tstore = np.asfortranarray(Tstore)
    B = inputs.normal_host(m, n, seed=m + n)
    ref = B.copy(order="F")
    oracle.dtrsm(side, trans, 0.5, np.asfortranarray(Tstore), ref, diag=diag, uplo=uplo)
    Bd = dev(B)
    assert lib().elemamd_trsm_f64(side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, 0.5,
                                  dev(Tstore).ptr(), na, Bd.ptr(), m, S()) == 0
    got = host(Bd)
    assert np.max(np.abs(got - ref)) <= 1e-11 * max(1.0, np.max(np.abs(ref))), (side, uplo, trans, diag)


def test_trsm_rejects_before_touching_b():
    B = inputs.normal_host(40, 3, seed=1)
    Bd = dev(B)
    L = dev(_tri(40, False, 1))
    assert lib().elemamd_trsm_f64(b"X", b"L", b"N", b"N", 40, 3, 2.0, L.ptr(), 40, Bd.ptr(), 40, S()) == -1
    assert lib().elemamd_trsm_f64(b"L", b"Q", b"N", b"N", 40, 3, 2.0, L.ptr(), 40, Bd.ptr(), 40, S()) == -1
    assert lib().elemamd_trsm_f64(b"L", b"L", b"N", b"N", 40, 3, 2.0, L.ptr(), 10, Bd.ptr(), 40, S()) == -1
    assert np.array_equal(host(Bd), B)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsv_every_form(oracle, uplo, trans, diag):
    n = 700
    T = _tri(n, uplo == "U", seed=3)
    x0 = inputs.normal_host(n, 1, seed=4)
    ref = x0.copy(order="F")
    oracle.dtrsv(trans, T, ref, diag=diag, uplo=uplo)
    xd = dev(x0)
    assert lib().elemamd_trsv_f64(uplo.encode(), trans.encode(), diag.encode(), n, dev(T).ptr(), n, xd.ptr(), 1,
                                  S()) == 0
    assert np.max(np.abs(host(xd) - ref)) <= 1e-11 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_herk_every_form(oracle, uplo, trans):
    n, k = 333, 45
    A = oracle.normal((n, k) if trans == "N" else (k, n), seed=7)
    C = oracle.normal((n, n), seed=8)
    ref = C.copy(order="F")
    oracle.dsyrk(trans, -0.5, A, 1.5, ref, uplo=uplo)
    Ad, Cd = dev(A), dev(C, ld=n + 3)
    assert lib().elemamd_herk_f64(uplo.encode(), trans.encode(), n, k, -0.5, Ad.ptr(), Ad.LDim(), 1.5, Cd.ptr(),
                                  Cd.LDim(), S()) == 0
    got = host(Cd)
    tri = np.tril if uplo == "L" else np.triu
    other = (lambda x: np.triu(x, 1)) if uplo == "L" else (lambda x: np.tril(x, -1))
    assert np.max(np.abs(tri(got - ref))) <= 16 * k * EPS * (np.abs(A).max() ** 2 * k + np.abs(C).max())
    assert np.array_equal(other(got), other(C))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("oa,ob", [("T", "N"), ("N", "N"), ("N", "T"), ("T", "T")])
def test_trrk_every_form(uplo, oa, ob):
    n, k = 260, 33
    rng = np.random.default_rng(11)
    A = np.asfortranarray(rng.standard_normal((k, n) if oa == "T" else (n, k)))
    B = np.asfortranarray(rng.standard_normal((n, k) if ob == "T" else (k, n)))
    C = np.asfortranarray(rng.standard_normal((n, n)))
    opA = A.T if oa == "T" else A
    opB = B.T if ob == "T" else B
    ref = 1.5 * opA @ opB - 0.25 * C
    Cd = dev(C)
    E.Trrk(E.LOWER if uplo == "L" else E.UPPER, E.TRANSPOSE if oa == "T" else E.NORMAL,
           E.TRANSPOSE if ob == "T" else E.NORMAL, 1.5, dev(A), dev(B), -0.25, Cd)
    got = host(Cd)
    tri = np.tril if uplo == "L" else np.triu
    other = (lambda x: np.triu(x, 1)) if uplo == "L" else (lambda x: np.tril(x, -1))
    assert np.max(np.abs(tri(got - ref))) <= 1e-12 * k
    assert np.array_equal(other(got), other(C))


# --------------------------------------------------------- Matrix<float>
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [100, 1000, 2049])
def test_float_cholesky_and_hpd_solve(oracle, uplo, n):
    """F = float: the factor against the FP64 oracle's (rounded to float,
    within a few float ulps times the growth of a well-conditioned A), the
    other triangle untouched, and HPDSolve's residual at float level."""
    k = 5
    A = inputs.gram_host(n, seed=n + 1)
    B = inputs.normal_host(n, k, seed=n + 2)
    ref = A.copy(order="F")
    assert (oracle.lvar3(ref, 128) if uplo == "L" else oracle.uvar3(ref, 128)) == 0
    tri = np.tril if uplo == "L" else np.triu
    other = (lambda x: np.triu(x, 1)) if uplo == "L" else (lambda x: np.tril(x, -1))
    A32 = A.astype(np.float32)
    M = Matrix.from_numpy(np.asfortranarray(A32))
    up = E.LOWER if uplo == "L" else E.UPPER
    E.Cholesky(up, M)
    got = host(M).astype(np.float64)
    assert np.linalg.norm(tri(got - ref)) / np.linalg.norm(tri(ref)) <= 1e-5
    assert np.array_equal(other(host(M)), other(A32))
    Ad, Bd = Matrix.from_numpy(np.asfortranarray(A32)), Matrix.from_numpy(np.asfortranarray(B.astype(np.float32)))
    E.HPDSolve(up, E.NORMAL, Ad, Bd)
    X = host(Bd).astype(np.float64)
    R = B - A @ X
    assert np.linalg.norm(R) / (np.linalg.norm(A) * np.linalg.norm(X)) <= 1e-5
    # SolveAfter<float> with the float factor
    Bd2 = Matrix.from_numpy(np.asfortranarray(B.astype(np.float32)))
    E.SolveAfter(up, E.NORMAL, Ad, Bd2)
    assert np.max(np.abs(host(Bd2) - host(Bd))) <= 1e-5 * np.max(np.abs(X))


# ------------------------------------------------------------ test driver
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_cholesky_driver_report(uplo):
    """The harness driver (reference tests/lapack-like/Cholesky.cpp) on the
    GPU: its report's ||A X - L L^H X||_F (or U^H U) is at rounding level."""
    import os
    import re
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "elemental_amd", "bin",
                       "Cholesky")
    assert os.path.exists(exe), "build() compiles the driver"
    r = subprocess.run([exe, "--height", "1000", "--nb", "128", "--uplo", uplo], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Will test Cholesky{uplo}" in r.stdout and "Starting Cholesky factorization...DONE." in r.stdout

    def val(label):
        m = re.search(re.escape(label) + r"\s*=\s*([0-9.eE+-]+)", r.stdout)
        assert m, label
        return float(m.group(1))

    prod = "L L^H" if uplo == "L" else "U^H U"
    err = val(f"||A X - {prod} X||_F")
    assert err <= 1e-13 * val("||A||_F") * val("||X||_F")
