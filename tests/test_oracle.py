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
