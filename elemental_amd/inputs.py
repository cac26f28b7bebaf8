"""Seeded synthetic inputs of the benchmark configurations.

The generator is a counter-based splitmix64 stream with Box-Muller normals;
the host (numpy) and device (``elemamd_fill_normal_f64``) versions produce
the same numbers up to the last-bit rounding of log/sin/cos.

Constructions (column-major float64):
  * ``gram``     A = G G^T / n + shift*I, G n x n N(0,1)   (shift 1: n=2048 factor-and-solve;
                 shift 0.1: the uneven batch)
  * ``spectrum`` A = Q diag(s) Q^T, Q the Householder-QR factor of an N(0,1)
                 matrix (signs fixed so diag(R) > 0), s log-spaced from 1 to 1/kappa
  * ``integer``  A = L0 L0^T with small-integer L0 and power-of-two diagonal: every
                 intermediate of the factorization and solve is an exact integer,
                 so the factor is L0 and the solution is known bit-for-bit
"""
from __future__ import annotations

import ctypes

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
TWO_PI = 6.283185307179586476925286766559


def splitmix64_at(seed: int, idx: np.ndarray) -> np.ndarray:
    """Output number ``idx`` of the splitmix64 stream started at ``seed``."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def _u01(x: np.ndarray) -> np.ndarray:
    return ((x >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0**-53


def normal_host(rows: int, cols: int, seed: int, offset: int = 0) -> np.ndarray:
    """rows x cols column-major N(0,1) array: entry t (column-major order) is
    normal number offset+t of stream ``seed``."""
    t = np.arange(rows * cols, dtype=np.int64) + offset
    p = (t >> 1).astype(np.uint64)
    u1 = _u01(splitmix64_at(seed, np.uint64(2) * p))
    u2 = _u01(splitmix64_at(seed, np.uint64(2) * p + np.uint64(1)))
    r = np.sqrt(-2.0 * np.log(u1))
    a = TWO_PI * u2
    v = np.where(t & 1, r * np.sin(a), r * np.cos(a))
    return np.asfortranarray(v.reshape(cols, rows).T)


def normal_device(rows: int, cols: int, seed: int, offset: int = 0):
    """Same numbers generated on the GPU; returns a Matrix."""
    import torch

    from ._lib import lib
    from .matrix import Matrix

    M = Matrix(rows, cols)
    rc = lib().elemamd_fill_normal_f64(M.ptr(), rows, cols, M.LDim(), seed, offset,
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc != 0:
        raise RuntimeError(f"elemamd_fill_normal_f64 failed ({rc})")
    return M


# ------------------------------------------------------------------ host
def gram_host(n: int, seed: int, shift: float = 1.0) -> np.ndarray:
    G = normal_host(n, n, seed)
    return np.asfortranarray(G @ G.T / n + shift * np.eye(n))


def spectrum_host(n: int, kappa: float, seed: int) -> np.ndarray:
    G = normal_host(n, n, seed)
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))
    s = np.logspace(0.0, -np.log10(kappa), n)
    A = (Q * s) @ Q.T
    return np.asfortranarray(0.5 * (A + A.T))


def integer_factor(n: int, seed: int, spread: int = 3):
    """(A, L0): L0 lower triangular with off-diagonal integers in
    [-spread, spread] and diagonal in {1, 2, 4}; A = L0 L0^T exactly."""
    rng = np.random.default_rng(seed)
    L0 = np.tril(rng.integers(-spread, spread + 1, (n, n)).astype(np.float64), -1)
    L0 += np.diag(2.0 ** rng.integers(0, 3, n))
    return np.asfortranarray(L0 @ L0.T), np.asfortranarray(L0)


def pascal(n: int):
    """Symmetric Pascal matrix P_ij = C(i+j, i) and its Cholesky factor
    L_ij = C(i, j) (a classical closed form; exact in double for n <= 25)."""
    from math import comb

    P = np.array([[comb(i + j, i) for j in range(n)] for i in range(n)], dtype=np.float64, order="F")
    L = np.array([[comb(i, j) if j <= i else 0 for j in range(n)] for i in range(n)], dtype=np.float64, order="F")
    return P, L


# ---------------------------------------------------------------- device
def gram_device(n: int, seed: int, shift: float = 1.0):
    import torch

    from .matrix import Matrix

    G = normal_device(n, n, seed).tensor()
    A = Matrix(n, n)
    torch.matmul(G, G.t(), out=A.tensor().t())  # symmetric, so the transpose view is the same matrix
    A.tensor().div_(n)
    A.tensor().diagonal().add_(shift)
    return A


def spectrum_device(n: int, kappa: float, seed: int):
    import torch

    from .matrix import Matrix

    G = normal_device(n, n, seed).tensor().clone()
    Q, R = torch.linalg.qr(G)
    Q.mul_(torch.sign(torch.diagonal(R)).unsqueeze(0))
    del R, G
    s = torch.logspace(0.0, -float(np.log10(kappa)), n, dtype=torch.float64, device=Q.device)
    Qs = Q * s.unsqueeze(0)
    A = Matrix(n, n)
    At = A.tensor()
    At.copy_(Qs @ Q.t())
    del Qs
    At.copy_(0.5 * (At + At.t()))
    return A


def integer_device(n: int, seed: int, spread: int = 3):
    """(A, L0) as Matrices for the exact-arithmetic known answer at any size
    (built on the device: A = L0 L0^T is an exact integer product)."""
    import torch

    from .matrix import Matrix

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    L0 = torch.randint(-spread, spread + 1, (n, n), generator=g, device="cuda", dtype=torch.int32).to(torch.float64)
    L0 = torch.tril(L0, -1)
    d = torch.randint(0, 3, (n,), generator=g, device="cuda", dtype=torch.int32).to(torch.float64)
    L0 += torch.diag(torch.pow(2.0, d))
    A = Matrix(n, n)
    A.tensor().copy_(L0 @ L0.t())
    Lm = Matrix(n, n)
    Lm.tensor().copy_(L0)
    return A, Lm
