"""The allocation-free SolveAfter (elemamd_solve_after_f64_noalloc) against
the allocating one (elemamd_solve_after_f64), through the C ABI.

Both run the same chain kernels in the same order, so X must agree bit for
bit.  Two calls in a row share the GPU's fixed scratch array; the second is
smaller, so it runs over what the first left there.  Sizes beyond the array
are rejected before anything is enqueued (B untouched)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("no GPU", allow_module_level=True)

import elemental_amd as E  # noqa: E402
from elemental_amd import Matrix, inputs, lib  # noqa: E402


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def factor(n, seed, uplo):
    A = inputs.gram_host(n, seed=seed)
    M = Matrix.from_numpy(np.asfortranarray(A), ldim=n + 3)
    E.Cholesky(E.UPPER if uplo == b"U" else E.LOWER, M)
    torch.cuda.synchronize()
    return M


def solve_both(M, B, uplo, orient):
    n, k = B.shape
    X0 = Matrix.from_numpy(B, ldim=n + 1)
    X1 = Matrix.from_numpy(B, ldim=n + 1)
    L = lib()
    rc0 = L.elemamd_solve_after_f64(uplo, orient, n, k, M.ptr(), M.LDim(), X0.ptr(), X0.LDim(), S())
    rc1 = L.elemamd_solve_after_f64_noalloc(uplo, orient, n, k, M.ptr(), M.LDim(), X1.ptr(), X1.LDim(), S())
    torch.cuda.synchronize()
    assert rc0 == 0 and rc1 == 0, (rc0, rc1)
    return X0.to_numpy(), X1.to_numpy()


@pytest.mark.parametrize("uplo", [b"L", b"U"])
def test_noalloc_matches_allocating_entry_point_twice(uplo):
    print(f"noalloc solve uplo={uplo} n=3000 k=40 then n=700 k=5", flush=True)
    M = factor(3000, 11, uplo)
    B = inputs.normal_host(3000, 40, seed=12)
    X0, X1 = solve_both(M, B, uplo, b"N")
    assert np.array_equal(X0, X1)
    # second, smaller call: the array still holds the first call's flags and sums
    M2 = factor(700, 13, uplo)
    B2 = inputs.normal_host(700, 5, seed=14)
    X0, X1 = solve_both(M2, B2, uplo, b"T")
    assert np.array_equal(X0, X1)
    # and the answer is the solve: backward error at FP64 level
    A2 = inputs.gram_host(700, seed=13)
    r = B2 - A2 @ X1
    eta = np.abs(r).max() / (np.abs(A2).sum(1).max() * np.abs(X1).max() + np.abs(B2).max())
    assert eta <= 1e-14


def test_noalloc_one_column_and_empty():
    M = factor(513, 21, b"L")
    B = inputs.normal_host(513, 1, seed=22)
    X0, X1 = solve_both(M, B, b"L", b"N")
    assert np.array_equal(X0, X1)
    L = lib()
    assert L.elemamd_solve_after_f64_noalloc(b"L", b"N", 0, 3, M.ptr(), 1, M.ptr(), 1, S()) == 0
    assert L.elemamd_solve_after_f64_noalloc(b"L", b"N", 513, 0, M.ptr(), 513 + 3, M.ptr(), 513, S()) == 0


def test_noalloc_rejects_beyond_its_array():
    L = lib()
    cap = L.elemamd_solve_after_f64_noalloc_limit()
    assert cap == 9 << 20
    n, k = 40000, 32  # 8 n k = 10.2 MB of partial sums > 9 MiB
    B = torch.full((n * k,), 1.5, dtype=torch.float64, device="cuda")
    A = torch.ones(1, dtype=torch.float64, device="cuda")  # never read: rejected first
    rc = L.elemamd_solve_after_f64_noalloc(b"L", b"N", n, k, ctypes.c_void_p(A.data_ptr()), n,
                                           ctypes.c_void_p(B.data_ptr()), n, S())
    torch.cuda.synchronize()
    assert rc == -4  # ELEMAMD_ERR_WORKSPACE
    assert bool((B == 1.5).all())
    assert L.elemamd_solve_after_f64_noalloc(b"X", b"N", 4, 1, ctypes.c_void_p(A.data_ptr()), 4,
                                             ctypes.c_void_p(B.data_ptr()), 4, S()) == -1
