"""Pins of the CPU oracle (oracle/elem_oracle.c) against what fixes the
answer independently of the oracle's own formulas:

* published splitmix64 output sequence (seed 1234567);
* integer-exact known answers: A = L0 L0^T with integer L0 and power-of-two
  diagonal makes every intermediate of LVar3Unb / LVar3 / SolveAfter an exact
  integer, so the factor must be L0 and the solution X0 bit-for-bit; a dropped
  term, wrong index or sign changes the result;
* the Pascal matrix, whose Cholesky factor is the binomial triangle;
* LAPACK dpotrf / dpotrs and the host BLAS (dsyrk, dtrsm, dgemm, dtrsv)
  through scipy, to rounding;
* invariants: L L^T = A, diag(L) > 0, backward error, and the failing column
  of an indefinite matrix.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.linalg.blas as sb

from elemental_amd import inputs


def test_splitmix64_published_sequence(oracle):
    # Vigna's reference splitmix64.c, seed 1234567
    assert oracle.splitmix64_sequence(1234567, 5) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]


def test_host_generator_matches_oracle_stream(oracle):
    # splitmix64_at(seed, i) is output i of the stream; same numbers in numpy
    idx = np.arange(5, dtype=np.uint64)
    assert [int(v) for v in inputs.splitmix64_at(1234567, idx)] == oracle.splitmix64_sequence(1234567, 5)
    a = oracle.normal((37, 11), seed=99, offset=5)
    b = inputs.normal_host(37, 11, seed=99, offset=5)
    assert np.max(np.abs(a - b)) <= 4 * np.finfo(float).eps * np.max(np.abs(a))


def test_normal_statistics(oracle):
    from scipy import stats

    x = oracle.normal(200000, seed=3).ravel()
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    assert stats.kstest(x, "norm").pvalue > 1e-3


@pytest.mark.parametrize("n", [1, 2, 7, 64, 127, 128, 129, 300])
def test_lvar3unb_integer_exact(oracle, n):
    A, L0 = inputs.integer_factor(n, seed=n)
    F = A.copy(order="F")
    assert oracle.lvar3unb(F) == 0
    assert np.array_equal(np.tril(F), L0)
    assert np.array_equal(np.triu(F, 1), np.triu(A, 1))  # strictly upper untouched


@pytest.mark.parametrize("n,nb", [(1, 128), (7, 3), (64, 32), (129, 128), (300, 128), (300, 7), (257, 1)])
def test_lvar3_integer_exact(oracle, n, nb):
    A, L0 = inputs.integer_factor(n, seed=7 * n + nb)
    F = A.copy(order="F")
    assert oracle.lvar3(F, nb) == 0
    assert np.array_equal(np.tril(F), L0)
    assert np.array_equal(np.triu(F, 1), np.triu(A, 1))


def test_pascal_closed_form(oracle):
    P, L = inputs.pascal(20)
    F = P.copy(order="F")
    assert oracle.lvar3(F, 8) == 0
    assert np.array_equal(np.tril(F), L)


def test_diagonal_and_2x2_closed_forms(oracle):
    d = np.array([4.0, 9.0, 0.25, 2.0])
    F = np.asfortranarray(np.diag(d))
    assert oracle.lvar3unb(F) == 0
    assert np.array_equal(np.diag(F), np.sqrt(d))
    a, b, c = 4.0, 2.0, 5.0  # [[a b][b c]] -> [[2 0][1 2]]
    F = np.asfortranarray([[a, b], [b, c]])
    assert oracle.lvar3unb(F) == 0
    assert np.array_equal(np.tril(F), np.array([[2.0, 0.0], [1.0, 2.0]]))


@pytest.mark.parametrize("n,nb", [(2, 128), (100, 32), (300, 128), (513, 128)])
def test_lvar3_vs_lapack_dpotrf(oracle, n, nb):
    A = inputs.gram_host(n, seed=n)
    F = A.copy(order="F")
    assert oracle.lvar3(F, nb) == 0
    Lr = sl.cholesky(A, lower=True)
    assert np.linalg.norm(np.tril(F) - Lr) / np.linalg.norm(Lr) < 1e-14
    assert np.all(np.diag(F) > 0)
    L = np.tril(F)
    assert np.linalg.norm(L @ L.T - A) / np.linalg.norm(A) < 1e-15


def test_lvar3_host_blas_matches_restated_blas(oracle):
    oracle.bind_host_blas(2)
    A = inputs.gram_host(400, seed=5)
    F1, F2 = A.copy(order="F"), A.copy(order="F")
    assert oracle.lvar3(F1, 128) == 0
    assert oracle.lvar3(F2, 128, host_blas=True) == 0
    assert np.linalg.norm(np.tril(F1 - F2)) / np.linalg.norm(np.tril(F1)) < 1e-14


@pytest.mark.parametrize("col", [0, 5, 127, 128, 200])
def test_non_hpd_reports_failing_column(oracle, col):
    n = 260
    A = inputs.gram_host(n, seed=11)
    # make the Schur complement at `col` negative: subtract from a_cc more than it holds
    L = sl.cholesky(A, lower=True)
    A[col, col] -= L[col, col] ** 2 * 1.5
    F = A.copy(order="F")
    assert oracle.lvar3(F, 128) == col + 1
    G = A.copy(order="F")
    assert oracle.lvar3unb(G) == col + 1


def test_nan_pivot_not_detected_like_reference(oracle):
    A = inputs.gram_host(8, seed=1)
    A[3, 3] = np.nan
    F = A.copy(order="F")
    assert oracle.lvar3unb(F) == 0  # NaN <= 0 is false in the reference too
    assert np.isnan(F[3, 3])


@pytest.mark.parametrize("n,k", [(1, 1), (7, 1), (64, 3), (300, 16), (129, 33)])
def test_solve_after_integer_exact(oracle, n, k):
    A, L0 = inputs.integer_factor(n, seed=n + k)
    rng = np.random.default_rng(n)
    X0 = np.asfortranarray(rng.integers(-4, 5, (n, k)).astype(np.float64))
    B = np.asfortranarray(A @ X0)
    X = B.copy(order="F")
    oracle.solve_after(np.asfortranarray(L0), X)
    assert np.array_equal(X, X0)


@pytest.mark.parametrize("n,k", [(1, 1), (50, 1), (300, 16), (513, 4)])
def test_hpd_solve_vs_lapack(oracle, n, k):
    A = inputs.gram_host(n, seed=2 * n)
    B = oracle.normal((n, k), seed=n)
    F, X = A.copy(order="F"), B.copy(order="F")
    assert oracle.hpd_solve(F, X, 128) == 0
    Xr = sl.cho_solve(sl.cho_factor(A, lower=True), B)
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) < 1e-13
    R = oracle.sym_lower_residual(A, X, B)
    assert np.linalg.norm(R) / (np.linalg.norm(A) * np.linalg.norm(X)) < 1e-16


def test_blas_restatements_vs_host_blas(oracle):
    rng = np.random.default_rng(0)
    m, n, k = 37, 29, 13
    A = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.asfortranarray(rng.standard_normal((n, n)))
    C1 = C.copy(order="F")
    oracle.dsyrk("N", -1.0, A, 1.0, C1)
    C2 = sb.dsyrk(-1.0, A, beta=1.0, c=C.copy(order="F"), lower=1)
    assert np.allclose(np.tril(C1), np.tril(C2), rtol=0, atol=1e-13)
    assert np.array_equal(np.triu(C1, 1), np.triu(C, 1))
    L = np.asfortranarray(np.tril(rng.standard_normal((n, n))) + 4 * np.eye(n))
    B = np.asfortranarray(rng.standard_normal((m, n)))
    B1 = B.copy(order="F")
    oracle.dtrsm("R", "T", 1.0, L, B1)
    assert np.allclose(B1 @ L.T, B, atol=1e-12)
    B2 = np.asfortranarray(rng.standard_normal((n, 5)))
    for t in ("N", "T"):
        X = B2.copy(order="F")
        oracle.dtrsm("L", t, 1.0, L, X)
        Lt = L if t == "N" else L.T
        assert np.allclose(Lt @ X, B2, atol=1e-12)
        x = np.asfortranarray(B2[:, :1].copy())
        oracle.dtrsv(t, L, x)
        assert np.allclose(Lt @ x, B2[:, :1], atol=1e-12)
    Bk = np.asfortranarray(rng.standard_normal((k, m)))
    Ck = np.asfortranarray(rng.standard_normal((n, m)))
    C3 = Ck.copy(order="F")
    oracle.dgemm("T", "N", 2.0, A.T.copy(order="F"), Bk, 0.5, C3)
    assert np.allclose(C3, 2.0 * A @ Bk + 0.5 * Ck, atol=1e-12)


@pytest.mark.parametrize("n,k,bsize", [(10, 4, 64), (130, 16, 64), (200, 33, 32), (64, 8, 64)])
def test_trrk_tn_lower_trapezoid(oracle, n, k, bsize):
    rng = np.random.default_rng(n)
    A = np.asfortranarray(rng.standard_normal((k, n)))
    B = np.asfortranarray(rng.standard_normal((k, n)))
    C = np.asfortranarray(rng.standard_normal((n, n)))
    C1 = C.copy(order="F")
    oracle.trrk_tn(-1.5, A, B, 0.5, C1, bsize)
    ref = -1.5 * A.T @ B + 0.5 * C
    assert np.allclose(np.tril(C1), np.tril(ref), atol=1e-12)
    assert np.array_equal(np.triu(C1, 1), np.triu(C, 1))
    # integer exact
    Ai = np.asfortranarray(rng.integers(-3, 4, (k, n)).astype(float))
    Bi = np.asfortranarray(rng.integers(-3, 4, (k, n)).astype(float))
    Ci = np.asfortranarray(rng.integers(-3, 4, (n, n)).astype(float))
    C2 = Ci.copy(order="F")
    oracle.trrk_tn(-1.0, Ai, Bi, 2.0, C2, bsize)
    assert np.array_equal(np.tril(C2), np.tril(-Ai.T @ Bi + 2.0 * Ci))


def test_spectrum_construction_condition():
    A = inputs.spectrum_host(200, 1e6, seed=1)
    w = np.linalg.eigvalsh(A)
    assert abs(w.max() - 1.0) < 1e-10 and abs(w.min() - 1e-6) < 1e-12


@pytest.mark.parametrize("n,k", [(1, 1), (300, 1), (700, 16), (513, 33)])
def test_host_blas_solve_vs_lapack_and_restated_loops(oracle, n, k):
    """The host-BLAS legs (LVar3 + SolveAfter over OpenBLAS, the CPU baseline
    and the reference X of the mixed tests) against LAPACK dpotrf/dpotrs and
    against the restated Netlib loops, to 1e-14 relative; k = 1 takes the
    dtrsv route of SolveAfter."""
    oracle.bind_host_blas(2)
    A = inputs.gram_host(n, seed=3 * n + k)
    B = oracle.normal((n, k), seed=n + 7 * k)
    F1, X1 = A.copy(order="F"), B.copy(order="F")
    assert oracle.hpd_solve(F1, X1, 128, host_blas=True) == 0
    F2, X2 = A.copy(order="F"), B.copy(order="F")
    assert oracle.hpd_solve(F2, X2, 128) == 0
    Xr = sl.cho_solve(sl.cho_factor(A, lower=True), B)
    assert np.linalg.norm(X1 - Xr) / np.linalg.norm(Xr) <= 1e-14
    assert np.linalg.norm(X1 - X2) / np.linalg.norm(X2) <= 1e-14
    assert np.linalg.norm(np.tril(F1 - F2)) / np.linalg.norm(np.tril(F2)) <= 1e-14
    # SolveAfter alone over the host BLAS, with LAPACK's factor
    X3 = B.copy(order="F")
    oracle.solve_after(np.asfortranarray(sl.cholesky(A, lower=True)), X3, host_blas=True)
    assert np.linalg.norm(X3 - Xr) / np.linalg.norm(Xr) <= 1e-14


@pytest.mark.parametrize("n,k", [(1, 1), (64, 3), (300, 16)])
def test_host_blas_solve_integer_exact(oracle, n, k):
    """Integer known answer through the host-BLAS legs: power-of-two pivots
    keep every quotient exact, so X must be X0 bit for bit."""
    oracle.bind_host_blas(2)
    A, L0 = inputs.integer_factor(n, seed=5 * n + k)
    rng = np.random.default_rng(n + k)
    X0 = np.asfortranarray(rng.integers(-4, 5, (n, k)).astype(np.float64))
    F, X = A.copy(order="F"), np.asfortranarray(A @ X0)
    assert oracle.hpd_solve(F, X, 128, host_blas=True) == 0
    assert np.array_equal(np.tril(F), L0)
    assert np.array_equal(X, X0)


@pytest.mark.parametrize("n,k", [(1, 1), (37, 5), (300, 16)])
def test_sym_lower_residual_exact_at_random_integer_x(oracle, n, k):
    """R = B - A X at an X that is NOT the solution, on integers, bit for bit
    against numpy's product with the full symmetric matrix; the strictly upper
    triangle holds garbage that must not be read."""
    rng = np.random.default_rng(n * 31 + k)
    Lw = np.tril(rng.integers(-9, 10, (n, n)).astype(np.float64))
    Asym = Lw + np.tril(Lw, -1).T
    Astored = np.asfortranarray(Lw + np.triu(rng.integers(1000, 2000, (n, n)).astype(np.float64), 1))
    X = np.asfortranarray(rng.integers(-7, 8, (n, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-50, 51, (n, k)).astype(np.float64))
    R = oracle.sym_lower_residual(Astored, X, B)
    assert np.array_equal(R, B - Asym @ X)
    # a mistake planted in the check's input (the upper triangle read as the
    # matrix) is caught by the same comparison
    assert not np.array_equal(oracle.sym_lower_residual(np.asfortranarray(np.triu(Astored)), X, B),
                              B - Asym @ X) or n == 1


# ------------------------------------------------------------------ UPPER
@pytest.mark.parametrize("n", [1, 2, 7, 64, 129, 300])
def test_uvar3unb_integer_exact_and_transpose_identity(oracle, n):
    """A = U0^T U0 with U0 = L0^T: UVar3Unb must give U0 exactly and leave the
    strictly lower triangle untouched; and operation by operation it is
    LVar3Unb on A^T (real data), so on a non-integer matrix the two agree bit
    for bit."""
    A, L0 = inputs.integer_factor(n, seed=3 * n + 1)
    F = A.copy(order="F")
    assert oracle.uvar3unb(F) == 0
    assert np.array_equal(np.triu(F), L0.T)
    assert np.array_equal(np.tril(F, -1), np.tril(A, -1))
    G = inputs.gram_host(n, seed=n + 5)
    G[np.triu_indices(n, 1)] += 0.0  # symmetric already
    Fu, Fl = G.copy(order="F"), np.asfortranarray(G.T)
    assert oracle.uvar3unb(Fu) == 0 and oracle.lvar3unb(Fl) == 0
    assert np.array_equal(np.triu(Fu), np.tril(Fl).T)


@pytest.mark.parametrize("n,nb", [(1, 128), (7, 3), (129, 128), (300, 64), (257, 1)])
def test_uvar3_integer_exact(oracle, n, nb):
    A, L0 = inputs.integer_factor(n, seed=11 * n + nb)
    F = A.copy(order="F")
    assert oracle.uvar3(F, nb) == 0
    assert np.array_equal(np.triu(F), L0.T)
    assert np.array_equal(np.tril(F, -1), np.tril(A, -1))


@pytest.mark.parametrize("n,nb", [(100, 32), (300, 128), (513, 128)])
def test_uvar3_vs_lapack_upper(oracle, n, nb):
    A = inputs.gram_host(n, seed=2 * n)
    F = A.copy(order="F")
    assert oracle.uvar3(F, nb) == 0
    Ur = sl.cholesky(A, lower=False)
    assert np.linalg.norm(np.triu(F) - Ur) / np.linalg.norm(Ur) < 1e-14
    U = np.triu(F)
    assert np.linalg.norm(U.T @ U - A) / np.linalg.norm(A) < 1e-15


@pytest.mark.parametrize("col", [0, 5, 128, 200])
def test_uvar3_non_hpd_reports_column(oracle, col):
    n = 260
    A = inputs.gram_host(n, seed=13)
    L = sl.cholesky(A, lower=True)
    A[col, col] -= L[col, col] ** 2 * 1.5
    F = A.copy(order="F")
    assert oracle.uvar3(F, 128) == col + 1


@pytest.mark.parametrize("n,k", [(1, 1), (9, 1), (64, 3), (300, 16)])
def test_solve_after_upper_integer_exact_and_lapack(oracle, n, k):
    A, L0 = inputs.integer_factor(n, seed=n + 2 * k)
    rng = np.random.default_rng(n + k)
    X0 = np.asfortranarray(rng.integers(-4, 5, (n, k)).astype(np.float64))
    X = np.asfortranarray(A @ X0)
    oracle.solve_after_upper(np.asfortranarray(L0.T), X)
    assert np.array_equal(X, X0)
    G = inputs.gram_host(n, seed=n)
    B = oracle.normal((n, k), seed=n)
    F, Xg = G.copy(order="F"), B.copy(order="F")
    assert oracle.hpd_solve_upper(F, Xg, 128) == 0
    Xr = sl.cho_solve(sl.cho_factor(G, lower=False), B)
    assert np.linalg.norm(Xg - Xr) / np.linalg.norm(Xr) < 1e-13


def test_upper_blas_restatements_vs_host_blas(oracle):
    rng = np.random.default_rng(5)
    m, n, k = 23, 31, 9
    U = np.asfortranarray(np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n))
    for side, trans in (("L", "N"), ("L", "T"), ("R", "N"), ("R", "T")):
        for diag in ("N", "U"):
            B = np.asfortranarray(rng.standard_normal((n, m) if side == "L" else (m, n)))
            X = B.copy(order="F")
            oracle.dtrsm(side, trans, 0.5, U, X, diag=diag, uplo="U")
            T = U.copy() if diag == "N" else np.triu(U, 1) + np.eye(n)
            op = T if trans == "N" else T.T
            lhs = op @ X if side == "L" else X @ op
            assert np.allclose(lhs, 0.5 * B, atol=1e-12), (side, trans, diag)
    for trans in ("N", "T"):
        x0 = rng.standard_normal(n)
        x = np.asfortranarray(x0.copy())
        oracle.dtrsv(trans, U, x, uplo="U")
        op = U if trans == "N" else U.T
        assert np.allclose(op @ x, x0, atol=1e-12)
        A = np.asfortranarray(rng.standard_normal((n, k) if trans == "N" else (k, n)))
        C = np.asfortranarray(rng.standard_normal((n, n)))
        C1 = C.copy(order="F")
        oracle.dsyrk(trans, -0.75, A, 2.0, C1, uplo="U")
        ref = -0.75 * (A @ A.T if trans == "N" else A.T @ A) + 2.0 * C
        assert np.allclose(np.triu(C1), np.triu(ref), atol=1e-12)
        assert np.array_equal(np.tril(C1, -1), np.tril(C, -1))
        C2 = sb.dsyrk(-0.75, A, beta=2.0, c=C.copy(order="F"), trans=0 if trans == "N" else 1, lower=0)
        assert np.allclose(np.triu(C1), np.triu(C2), atol=1e-12)


# ------------------------------------------------ inverses (CholeskyLVar2)
def _bidiag_unit(n, seed):
    """Unit lower bidiagonal L0 with subdiagonal in {-1, 0, 1}: inv(L0) has
    entries in {-1, 0, 1} and A = L0 L0^T is an integer matrix whose inverse
    inv(L0)^T inv(L0) is an integer matrix too (entries <= n), so every
    intermediate of the inverse algorithms is an exact integer."""
    rng = np.random.default_rng(seed)
    L0 = np.eye(n)
    L0[np.arange(1, n), np.arange(n - 1)] = rng.integers(-1, 2, n - 1)
    return L0


def _exact_inverse_of_bidiag(L0):
    n = L0.shape[0]
    X = np.zeros((n, n), dtype=np.int64)  # inv(L0) by forward substitution in integers
    s = np.rint(np.diag(L0, -1)).astype(np.int64)
    for k in range(n):
        X[k, k] = 1
        for i in range(k + 1, n):
            X[i, k] = -s[i - 1] * X[i - 1, k]
    return X


@pytest.mark.parametrize("n", [1, 2, 7, 64, 128])
def test_trtri_unblocked_vs_substitution_and_unit(oracle, n):
    rng = np.random.default_rng(n)
    L = np.tril(rng.standard_normal((n, n))) + n * np.eye(n)
    X = oracle.trtri_lvar3unb(L.copy(order="F"))
    ref = sl.solve_triangular(L, np.eye(n), lower=True)
    assert np.array_equal(np.triu(X, 1), np.triu(L, 1))  # strictly upper untouched
    assert np.max(np.abs(np.tril(X) - ref)) <= 1e-14 * max(1.0, np.max(np.abs(ref)))
    # unit diagonal: the stored diagonal is neither read nor written
    Lu = L.copy(order="F")
    Lu[np.diag_indices(n)] = 7.5
    Xu = oracle.trtri_lvar3unb(Lu, unit=True)
    Lunit = np.tril(L, -1) + np.eye(n)
    refu = sl.solve_triangular(Lunit, np.eye(n), lower=True)
    assert np.all(np.diag(Xu) == 7.5)
    assert np.max(np.abs(np.tril(Xu, -1) - np.tril(refu, -1))) <= 1e-13 * max(1.0, np.max(np.abs(refu)))


@pytest.mark.parametrize("n,nb", [(1, 128), (5, 2), (200, 37), (300, 64), (256, 128)])
def test_trtri_blocked_integer_exact_and_vs_unblocked(oracle, n, nb):
    L0 = _bidiag_unit(n, seed=n + nb)
    X = oracle.trtri(L0.copy(order="F"), nb=nb)
    assert np.array_equal(np.tril(X), _exact_inverse_of_bidiag(L0).astype(float))
    rng = np.random.default_rng(nb)
    L = np.tril(rng.standard_normal((n, n))) + n * np.eye(n)
    Xb = oracle.trtri(L.copy(order="F"), nb=nb)
    ref = sl.solve_triangular(L, np.eye(n), lower=True)
    assert np.max(np.abs(np.tril(Xb) - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
    if nb >= n:  # one block: exactly the unblocked loop
        assert np.array_equal(Xb, oracle.trtri_lvar3unb(L.copy(order="F")))


@pytest.mark.parametrize("n", [1, 3, 50, 128])
def test_trtrmm_integer_exact(oracle, n):
    rng = np.random.default_rng(3 * n)
    X = np.tril(rng.integers(-4, 5, (n, n))).astype(float)
    C = oracle.trtrmm_lunb(X.copy(order="F"))
    assert np.array_equal(np.tril(C), np.tril(X.T @ X))  # exact: small integers
    assert np.array_equal(np.triu(C, 1), np.triu(X, 1))


@pytest.mark.parametrize("n,nb", [(1, 128), (2, 1), (100, 128), (300, 64), (500, 37), (700, 128)])
def test_hpd_inverse_vs_numpy_inverse(oracle, n, nb):
    A = inputs.gram_host(n, seed=5 * n + nb)
    Ai = A.copy(order="F")
    assert oracle.hpd_inverse(Ai, nb) == 0
    ref = np.linalg.inv(A)
    kappa = np.linalg.cond(A)
    err = np.linalg.norm(np.tril(Ai - ref)) / np.linalg.norm(np.tril(ref))
    assert err <= 10 * kappa * n * np.finfo(float).eps, (err, kappa)
    assert np.array_equal(np.triu(Ai, 1), np.triu(A, 1))  # strictly upper untouched


@pytest.mark.parametrize("n,nb", [(60, 128), (300, 64), (513, 128)])
def test_hpd_inverse_integer_exact(oracle, n, nb):
    L0 = _bidiag_unit(n, seed=n)
    A = L0 @ L0.T
    Ai = A.copy(order="F")
    assert oracle.hpd_inverse(Ai, nb) == 0
    X = _exact_inverse_of_bidiag(L0)
    assert np.array_equal(np.tril(Ai), np.tril(X.T @ X).astype(float))


@pytest.mark.parametrize("col", [0, 63, 64, 200])
def test_hpd_inverse_non_hpd_column_and_state(oracle, col):
    n, nb = 300, 64
    A = inputs.gram_host(n, seed=17)
    A[col, col] = -50.0
    Ai = A.copy(order="F")
    assert oracle.hpd_inverse(Ai, nb) == col + 1
    # the steps before the failing block are those of the matrix's leading part
    k0 = (col // nb) * nb
    if k0 > 0:
        lead = A[:k0, :k0].copy(order="F")
        assert oracle.hpd_inverse(lead, nb) == 0
        assert np.array_equal(np.tril(Ai[:k0, :k0]), np.tril(lead))


# -------------------------------------- generalized-definite eigenproblem
@pytest.mark.parametrize("n", [1, 2, 9, 130])
def test_two_sided_reductions_exact_on_dyadic_problems(oracle, n):
    """L unit bidiagonal (entries in {-1, 0, 1}) and A a small-integer
    symmetric matrix: every intermediate of the variant-4 loops is an exact
    dyadic rational (the only divisions are by 1 and 2), so the reductions
    must equal inv(L) A inv(L)^T and L^T A L computed in exact integers."""
    rng = np.random.default_rng(n)
    L0 = _bidiag_unit(n, seed=n + 1)
    M = rng.integers(-3, 4, (n, n))
    A = (M + M.T).astype(float)
    Li = _exact_inverse_of_bidiag(L0)
    A1 = oracle.two_sided_trsm(A.copy(order="F"), L0.copy(order="F"))
    assert np.array_equal(np.tril(A1), np.tril(Li @ A.astype(np.int64) @ Li.T).astype(float))
    assert np.array_equal(np.triu(A1, 1), np.triu(A, 1))
    A2 = oracle.two_sided_trmm(A.copy(order="F"), L0.copy(order="F"))
    Lz = np.rint(L0).astype(np.int64)
    assert np.array_equal(np.tril(A2), np.tril(Lz.T @ A.astype(np.int64) @ Lz).astype(float))


@pytest.mark.parametrize("n", [5, 64, 200])
def test_two_sided_reductions_vs_products(oracle, n):
    rng = np.random.default_rng(2 * n)
    L = np.tril(rng.standard_normal((n, n))) / np.sqrt(n) + 2 * np.eye(n)
    M = rng.standard_normal((n, n))
    A = M + M.T
    A1 = oracle.two_sided_trsm(A.copy(order="F"), L.copy(order="F"))
    Li = np.linalg.inv(L)
    ref1 = Li @ A @ Li.T
    assert np.max(np.abs(np.tril(A1) - np.tril(ref1))) <= 1e-12 * np.max(np.abs(ref1))
    A2 = oracle.two_sided_trmm(A.copy(order="F"), L.copy(order="F"))
    ref2 = L.T @ A @ L
    assert np.max(np.abs(np.tril(A2) - np.tril(ref2))) <= 1e-12 * np.max(np.abs(ref2))


@pytest.mark.parametrize("gtype", [1, 2, 3])
def test_gen_definite_eig_vs_lapack_sygvd(oracle, gtype):
    n = 150
    rng = np.random.default_rng(gtype)
    M = rng.standard_normal((n, n))
    A = M + M.T
    B = inputs.gram_host(n, seed=gtype + 10)
    info, w, X, _ = oracle.hermitian_gen_definite_eig(gtype, A, B)
    assert info == 0
    ref = sl.eigh(A, B, type=gtype, eigvals_only=True)
    assert np.max(np.abs(w - ref)) <= 1e-11 * np.max(np.abs(ref))
    # eigenvector relations of each type, and B-orthonormality (X^T B X = I
    # for AXBX / ABX, X^T inv(B) X = I for BAX)
    if gtype == 1:
        R = A @ X - B @ X * w
    elif gtype == 2:
        R = A @ (B @ X) - X * w
    else:
        R = B @ (A @ X) - X * w
    assert np.max(np.abs(R)) <= 1e-10 * np.max(np.abs(w)) * np.max(np.abs(X))
    G = X.T @ B @ X if gtype != 3 else X.T @ np.linalg.solve(B, X)
    assert np.max(np.abs(G - np.eye(n))) <= 1e-10


def test_gen_definite_eig_b_not_hpd(oracle):
    n = 50
    A = np.eye(n)
    B = inputs.gram_host(n, seed=3)
    B[20, 20] = -1.0
    info, *_ = oracle.hermitian_gen_definite_eig(1, A, B)
    assert info == 21
