"""Golden fixtures of the factor-and-solve path (tests/golden/cholesky_fixtures.npz,
written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit for bit, and the fixtures
satisfy what does not depend on the oracle (LAPACK's factor, L L^T = A, the
solve residual, the failing column of the non-HPD cases).
GPU: the diagonal-block kernel reproduces the single-block factors bit for
bit; the blocked factorization and HPDSolve agree with the fixtures to
1e-12 relative; the non-HPD cases report the fixture's column and leave the
blocks before the failing one as the fixture has them.
"""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "cholesky_fixtures.npz"))
SIZES = sorted({int(k.split("_")[1]) for k in FIX.files if k.startswith("A_")})
BLOCKS = [32, 128]
BAD = sorted({int(k.split("_")[2]) for k in FIX.files if k.startswith("bad_A_")})


def full_L(n, nb):
    """Fixture factor with the untouched upper triangle of A restored."""
    return np.asfortranarray(FIX[f"L_{n}_{nb}"] + np.triu(FIX[f"A_{n}"], 1))


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("nb", BLOCKS)
def test_oracle_reproduces_fixture(oracle, n, nb):
    A = np.asfortranarray(FIX[f"A_{n}"])
    L = A.copy(order="F")
    assert oracle.lvar3(L, nb) == 0
    assert np.array_equal(L, full_L(n, nb))
    X = np.asfortranarray(FIX[f"B_{n}"]).copy(order="F")
    F = A.copy(order="F")
    assert oracle.hpd_solve(F, X, nb) == 0
    assert np.array_equal(X, FIX[f"X_{n}_{nb}"])


@pytest.mark.parametrize("n", SIZES)
def test_fixture_invariants(n):
    A = FIX[f"A_{n}"]
    Lnp = np.linalg.cholesky(A)
    for nb in BLOCKS:
        L = FIX[f"L_{n}_{nb}"]
        assert np.linalg.norm(L - Lnp) <= 1e-13 * np.linalg.norm(Lnp)
        assert np.linalg.norm(L @ L.T - A) <= 1e-14 * n * np.linalg.norm(A)
        X, B = FIX[f"X_{n}_{nb}"], FIX[f"B_{n}"]
        assert np.abs(B - A @ X).max() <= 1e-13 * (np.abs(A).sum(1).max() * np.abs(X).max() + np.abs(B).max())
    if n <= 32:  # one block for both block sizes: the same unblocked loop
        assert np.array_equal(FIX[f"L_{n}_32"], FIX[f"L_{n}_128"])


@pytest.mark.parametrize("n", BAD)
def test_non_hpd_fixture(oracle, n):
    A = np.asfortranarray(FIX[f"bad_A_{n}"])
    F = A.copy(order="F")
    info = oracle.lvar3(F, 128)
    assert info == int(FIX[f"bad_info_{n}"][0])
    assert np.array_equal(np.tril(F), FIX[f"bad_F_{n}"])
    # the failing pivot is the first non-positive one of LAPACK's factorization
    c = info - 1
    assert np.all(np.linalg.eigvalsh(A[:c, :c]) > 0) if c > 0 else True
    assert np.linalg.eigvalsh(A[: c + 1, : c + 1]).min() <= 0


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("n", [n for n in SIZES if n <= 128])
def test_gpu_potrf_matches_fixture_bitwise(n):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from elemental_amd import Matrix, lib

    M = Matrix.from_numpy(FIX[f"A_{n}"], ldim=n + 2)
    d = torch.zeros(1, dtype=torch.int32, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().elemamd_potrf_unb_f64(n, M.ptr(), M.LDim(), ctypes.c_void_p(d.data_ptr()), s) == 0
    assert lib().elemamd_info_read(ctypes.c_void_p(d.data_ptr()), s) == 0
    torch.cuda.synchronize()
    assert np.array_equal(M.to_numpy(), full_L(n, 128))


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("nb", BLOCKS)
def test_gpu_hpd_solve_matches_fixture(n, nb):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import elemental_amd as E

    A = E.Matrix.from_numpy(FIX[f"A_{n}"])
    X = E.Matrix.from_numpy(FIX[f"B_{n}"])
    E.HPDSolve(E.LOWER, E.NORMAL, A, X, nb=nb)
    torch.cuda.synchronize()
    L, Lf = A.to_numpy(), full_L(n, nb)
    assert np.linalg.norm(np.tril(L) - np.tril(Lf)) <= 1e-12 * np.linalg.norm(np.tril(Lf))
    assert np.array_equal(np.triu(L, 1), np.triu(Lf, 1))
    Xf = FIX[f"X_{n}_{nb}"]
    assert np.abs(X.to_numpy() - Xf).max() <= 1e-12 * np.abs(Xf).max()


@pytest.mark.gpu
@pytest.mark.parametrize("n", BAD)
def test_gpu_non_hpd_matches_fixture(n):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import elemental_amd as E

    M = E.Matrix.from_numpy(FIX[f"bad_A_{n}"])
    with pytest.raises(E.LogicError, match="A was not numerically HPD") as ei:
        E.Cholesky(E.LOWER, M)
    info = int(FIX[f"bad_info_{n}"][0])
    assert ei.value.column == info - 1
    got = np.tril(M.to_numpy())
    ref = FIX[f"bad_F_{n}"]
    b0 = ((info - 1) // 128) * 128  # blocks before the failing one are final
    if b0 > 0:
        assert np.linalg.norm(got[:, :b0] - ref[:, :b0]) <= 1e-12 * np.linalg.norm(ref[:, :b0])
    # inside the failing block the unblocked loop stops where the reference throws
    blk = slice(b0, min(n, b0 + 128))
    assert np.linalg.norm(got[blk, blk] - ref[blk, blk]) <= 1e-10 * np.linalg.norm(ref[blk, blk])
