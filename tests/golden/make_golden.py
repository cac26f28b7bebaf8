"""Writes the golden fixtures of the lower Cholesky factor-and-solve path.

The reference's own tests hold no vectors for this path and its library
cannot be built here (it needs a CMake-generated configuration header, MPI
and an external BLAS), so the vectors are the CPU oracle's output, each
cross-checked at generation time against what does not depend on the
oracle: LAPACK's Cholesky factor (numpy), the invariant L L^T = A, and the
solve residual.  Every fixture records the inputs explicitly.

    python tests/golden/make_golden.py      (writes tests/golden/cholesky_fixtures.npz)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from elemental_amd import inputs  # noqa: E402
from oracle import oracle as O  # noqa: E402

SIZES = [1, 2, 7, 64, 127, 128, 129, 160]
BLOCKS = [32, 128]
RHS = {1: 1, 2: 3, 7: 16, 64: 3, 127: 1, 128: 16, 129: 3, 160: 16}


def main(path=os.path.join(HERE, "cholesky_fixtures.npz")):
    out = {}
    for n in SIZES:
        A = inputs.gram_host(n, seed=500 + n)
        B = inputs.normal_host(n, RHS[n], seed=900 + n)
        out[f"A_{n}"] = A
        out[f"B_{n}"] = B
        Lnp = np.linalg.cholesky(A)
        for nb in BLOCKS:
            L = A.copy(order="F")
            assert O.lvar3(L, nb) == 0
            Lt = np.tril(L)
            assert np.linalg.norm(Lt - Lnp) <= 1e-13 * np.linalg.norm(Lnp)  # LAPACK's factor
            assert np.linalg.norm(Lt @ Lt.T - A) <= 1e-14 * n * np.linalg.norm(A)
            assert np.array_equal(np.triu(L, 1), np.triu(A, 1))  # upper triangle untouched
            out[f"L_{n}_{nb}"] = np.tril(L)  # the strictly upper part is A's
            X = B.copy(order="F")
            F = A.copy(order="F")
            assert O.hpd_solve(F, X, nb) == 0
            assert np.array_equal(F, L)
            R = B - A @ X
            assert np.abs(R).max() <= 1e-13 * (np.abs(A).sum(1).max() * np.abs(X).max() + np.abs(B).max())
            out[f"X_{n}_{nb}"] = X
    # non-HPD: the reference throws at the first non-positive pivot
    for n, col in [(7, 3), (129, 128), (160, 150)]:
        A = inputs.gram_host(n, seed=700 + n)
        L = np.linalg.cholesky(A)
        A[col, col] -= 1.5 * L[col, col] ** 2
        F = A.copy(order="F")
        info = O.lvar3(F, 128)
        assert info == col + 1
        out[f"bad_A_{n}"] = A
        out[f"bad_info_{n}"] = np.array([info], dtype=np.int64)
        assert np.array_equal(np.triu(F, 1), np.triu(A, 1))
        out[f"bad_F_{n}"] = np.tril(F)  # lower triangle of the state the reference leaves when it throws
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
