"""HermitianGenDefiniteEig (HermitianGenDefiniteEig.hpp) through the C ABI:
Cholesky of B with this library's factorization, the two-sided reduction as
products with the factor or its inverse, rocSOLVER's dsyevdx on the reduced
matrix, and the back-transformation; against the oracle (the reference's
Cholesky and unblocked two-sided reductions with LAPACK's eigensolver) and
LAPACK's dsygvd (scipy.linalg.eigh), with each type's eigen-relation and the
B-orthonormality of X as invariants."""
import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

import elemental_amd as E  # noqa: E402
from elemental_amd import Matrix, inputs  # noqa: E402

TYPES = {1: E.AXBX, 2: E.ABX, 3: E.BAX}


def dev(a, ld=None):
    return Matrix.from_numpy(np.asfortranarray(a), ldim=ld)


def host(M):
    torch.cuda.synchronize()
    return M.to_numpy()


def problem(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    return M + M.T, inputs.gram_host(n, seed=seed + 1)


def relation(gtype, A, B, X, w):
    if gtype == 1:
        return A @ X - B @ X * w
    if gtype == 2:
        return A @ (B @ X) - X * w
    return B @ (A @ X) - X * w


@pytest.mark.parametrize("gtype", [1, 2, 3])
@pytest.mark.parametrize("n", [1, 40, 300, 1100])
def test_gen_eig_vs_oracle_and_lapack(oracle, gtype, n):
    A, B = problem(n, seed=100 * gtype + n)
    Ad, Bd = dev(A, ld=n + 3), dev(B, ld=n + 1)
    w, X = E.HermitianGenDefiniteEig(TYPES[gtype], E.LOWER, Ad, Bd, vectors=True)
    w, X = host(w)[:, 0], host(X)
    info, wr, Xr, Ared = oracle.hermitian_gen_definite_eig(gtype, A, B)
    assert info == 0 and len(w) == n
    scale = np.max(np.abs(wr))
    assert np.max(np.abs(w - wr)) <= 1e-10 * scale
    assert np.max(np.abs(w - sl.eigh(A, B, type=gtype, eigvals_only=True))) <= 1e-10 * scale
    R = relation(gtype, A, B, X, w)
    assert np.max(np.abs(R)) <= 1e-9 * scale * max(1.0, np.max(np.abs(X)))
    G = X.T @ B @ X if gtype != 3 else X.T @ np.linalg.solve(B, X)
    assert np.max(np.abs(G - np.eye(n))) <= 1e-9
    # B holds this library's Cholesky factor (bit for bit) and its upper
    # triangle is untouched; A's lower triangle holds the reduced matrix
    Bf = host(Bd)
    Lc = dev(B)
    E.Cholesky(E.LOWER, Lc)
    assert np.array_equal(np.tril(Bf), np.tril(host(Lc)))
    assert np.array_equal(np.triu(Bf, 1), np.triu(B, 1))
    Af = host(Ad)
    assert np.max(np.abs(np.tril(Af) - np.tril(Ared))) <= 1e-10 * max(1.0, np.max(np.abs(Ared)))
    assert np.array_equal(np.triu(Af, 1), np.triu(A, 1))


def test_gen_eig_index_and_value_ranges():
    n = 500
    A, B = problem(n, seed=7)
    wall, Xall = E.HermitianGenDefiniteEig(E.AXBX, E.LOWER, dev(A), dev(B), vectors=True)
    wall, Xall = host(wall)[:, 0], host(Xall)
    w, X = E.HermitianGenDefiniteEig(E.AXBX, E.LOWER, dev(A), dev(B), vectors=True, lowerBound=10, upperBound=20)
    w, X = host(w)[:, 0], host(X)
    assert len(w) == 11 and X.shape == (n, 11)
    assert np.max(np.abs(w - wall[10:21])) <= 1e-10 * np.max(np.abs(wall))
    assert np.max(np.abs(relation(1, A, B, X, w))) <= 1e-9 * np.max(np.abs(wall))
    lo, hi = 0.5 * (wall[99] + wall[100]), 0.5 * (wall[199] + wall[200])
    w, _ = E.HermitianGenDefiniteEig(E.AXBX, E.LOWER, dev(A), dev(B), lowerBound=lo, upperBound=hi)
    w = host(w)[:, 0]
    assert len(w) == 100 and np.max(np.abs(w - wall[100:200])) <= 1e-10 * np.max(np.abs(wall))


@pytest.mark.parametrize("gtype", [1, 3])
def test_gen_eig_upper_matches_lower(gtype):
    n = 400
    A, B = problem(n, seed=11 + gtype)
    wl, _ = E.HermitianGenDefiniteEig(TYPES[gtype], E.LOWER, dev(A), dev(B))
    Au = np.triu(A) + np.tril(np.full((n, n), 5.0), -1)  # only the upper triangles are read
    Bu = np.triu(B) + np.tril(np.full((n, n), -7.0), -1)
    Ad, Bd = dev(Au), dev(Bu)
    wu, Xu = E.HermitianGenDefiniteEig(TYPES[gtype], E.UPPER, Ad, Bd, vectors=True)
    wl, wu, Xu = host(wl)[:, 0], host(wu)[:, 0], host(Xu)
    assert np.max(np.abs(wl - wu)) <= 1e-11 * np.max(np.abs(wl))
    assert np.max(np.abs(relation(gtype, A, B, Xu, wu))) <= 1e-9 * np.max(np.abs(wl))
    Bf = host(Bd)
    assert np.array_equal(np.tril(Bf, -1), np.tril(Bu, -1))
    assert np.max(np.abs(np.triu(Bf).T @ np.triu(Bf) - B)) <= 1e-12 * np.max(np.abs(B))


def test_gen_eig_b_not_hpd_raises_like_reference():
    n = 200
    A, B = problem(n, seed=21)
    B[77, 77] = -5.0
    Ad = dev(A)
    with pytest.raises(E.LogicError) as ei:
        E.HermitianGenDefiniteEig(E.AXBX, E.LOWER, Ad, dev(B), vectors=True)
    assert ei.value.column == 77
    assert np.array_equal(host(Ad), A)  # the reference throws before touching A
