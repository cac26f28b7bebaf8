"""Batched factor-and-solve of many HPD matrices of uneven size.

No reference counterpart: Elemental factors one ``Matrix<F>`` per call.  Each
matrix goes through the same LVar3 + SolveAfter computation as
``HPDSolve(LOWER, NORMAL, A, B)`` (``include/elemental/lapack-like/HPDSolve.hpp:18-28``),
all of them in one launch sequence that the library captures in a hipGraph.

Matrices are packed: A_i column-major with ld = n_i at element offset
``a_off[i]`` of one float64 device buffer, B_i (n_i x k_i) at ``b_off[i]``.
Matrices are grouped in size buckets, each factored with its own block size;
``autotune`` picks those block sizes by timing and caches them.
"""
from __future__ import annotations

import ctypes
import json
import math
import os

import numpy as np

from ._lib import lib

F64, F32_IR, BF16X3_IR = 0, 1, 2
PRECISIONS = {"f64": F64, "f32": F32_IR, "bf16x3": BF16X3_IR}

# size buckets: (upper bound of n, block size, launch form) used when no tuned
# value is cached; launch form 0 = one launch per block step and kernel,
# g > 0 = one persistent launch of g workgroups
DEFAULT_BUCKETS = [(64, 64, 0), (128, 128, 0), (256, 64, 0), (512, 128, 0), (1 << 30, 128, 0)]
PERSIST_MAX_NB = 128
TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "batch_gfx950.json")


def _stream(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _bucket3(b):
    m, nb = int(b[0]), int(b[1])
    return (m, nb, int(b[2]) if len(b) > 2 else 0)


def tuned_buckets(precision: str = "f64", path: str = TUNING_FILE):
    """(max_n, nb, grid) triples from the autotuner's cache, else the defaults."""
    try:
        with open(path) as f:
            data = json.load(f)
        return [_bucket3(b) for b in data[precision]["buckets"]]
    except (OSError, KeyError, ValueError, IndexError, TypeError):
        return list(DEFAULT_BUCKETS)


class BatchPlan:
    """Plan for matrices of sizes ``n`` with ``k`` right-hand sides each."""

    def __init__(self, n, k, precision: str = "f64", buckets=None):
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.k = np.ascontiguousarray(k, dtype=np.int32)
        if self.n.shape != self.k.shape:
            raise ValueError("n and k must have the same length")
        self.precision = precision
        self.buckets = [_bucket3(b) for b in buckets] if buckets is not None else tuned_buckets(precision)
        bmax = np.array([m for m, _, _ in self.buckets], dtype=np.int32)
        bnb = np.array([nb for _, nb, _ in self.buckets], dtype=np.int32)
        bgrid = np.array([g for _, _, g in self.buckets], dtype=np.int32)
        h = ctypes.c_void_p()
        rc = lib().elemamd_batch_plan_create(len(self.n), self.n.ctypes.data_as(ctypes.c_void_p),
                                             self.k.ctypes.data_as(ctypes.c_void_p), PRECISIONS[precision],
                                             len(bmax), bmax.ctypes.data_as(ctypes.c_void_p),
                                             bnb.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h))
        if rc != 0:
            raise ValueError(f"elemamd_batch_plan_create failed ({rc})")
        self._h = h
        rc = lib().elemamd_batch_set_launch(h, bgrid.ctypes.data_as(ctypes.c_void_p))
        if rc != 0:
            self.close()
            raise ValueError(f"elemamd_batch_set_launch failed ({rc})")
        cnt = len(self.n)
        self.a_off = np.zeros(cnt, dtype=np.int64)
        self.b_off = np.zeros(cnt, dtype=np.int64)
        at, bt = ctypes.c_int64(), ctypes.c_int64()
        lib().elemamd_batch_plan_layout(h, self.a_off.ctypes.data_as(ctypes.c_void_p),
                                        self.b_off.ctypes.data_as(ctypes.c_void_p), ctypes.byref(at), ctypes.byref(bt))
        self.a_total, self.b_total = at.value, bt.value
        self._work = None

    # ------------------------------------------------------------ buffers
    def empty_a(self):
        import torch

        return torch.empty(max(self.a_total, 1), dtype=torch.float64, device="cuda")

    def empty_b(self):
        import torch

        return torch.empty(max(self.b_total, 1), dtype=torch.float64, device="cuda")

    def a_view(self, A, i: int):
        """Column-major n_i x n_i view of A_i (a torch tensor of shape (n, n)
        whose [r, c] element is A_i(r, c))."""
        n = int(self.n[i])
        o = int(self.a_off[i])
        return A[o:o + n * n].view(n, n).t()

    def b_view(self, B, i: int):
        n, k = int(self.n[i]), int(self.k[i])
        o = int(self.b_off[i])
        return B[o:o + n * k].view(k, n).t()

    def workspace_bytes(self) -> int:
        return int(lib().elemamd_batch_workspace(self._h))

    def set_refinement(self, sweeps: int = 4, tol: float = 1e-14):
        rc = lib().elemamd_batch_set_refinement(self._h, sweeps, tol)
        if rc != 0:
            raise ValueError(f"elemamd_batch_set_refinement failed ({rc})")

    # -------------------------------------------------------------- solve
    def solve(self, A, B, use_graph: bool = True, stream=None):
        """Factor and solve every matrix (asynchronous).  Plain FP64
        overwrites A_i with L_i; the refinement modes leave A unchanged.
        B_i is overwritten with X_i."""
        import torch

        wb = self.workspace_bytes()
        if wb and (self._work is None or self._work.numel() < wb):
            self._work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        work = _ptr(self._work) if wb else ctypes.c_void_p()
        rc = lib().elemamd_batch_hpd_solve(self._h, _ptr(A), _ptr(B), work, wb, 1 if use_graph else 0,
                                           _stream(stream))
        if rc != 0:
            raise RuntimeError(f"elemamd_batch_hpd_solve failed ({rc})")

    def results(self, stream=None):
        """(info, backward_error, sweeps) per matrix, caller's order."""
        cnt = len(self.n)
        info = np.zeros(cnt, dtype=np.int32)
        eta = np.zeros(cnt, dtype=np.float64)
        it = np.zeros(cnt, dtype=np.int32)
        rc = lib().elemamd_batch_results(self._h, info.ctypes.data_as(ctypes.c_void_p),
                                         eta.ctypes.data_as(ctypes.c_void_p), it.ctypes.data_as(ctypes.c_void_p),
                                         _stream(stream))
        if rc != 0:
            raise RuntimeError(f"elemamd_batch_results failed ({rc})")
        return info, eta, it

    def backward_error(self, A, X, B, stream=None) -> np.ndarray:
        import torch

        R = torch.empty_like(B)
        eta = np.zeros(len(self.n), dtype=np.float64)
        rc = lib().elemamd_batch_backward_error(self._h, _ptr(A), _ptr(X), _ptr(B), _ptr(R),
                                                eta.ctypes.data_as(ctypes.c_void_p), _stream(stream))
        if rc != 0:
            raise RuntimeError(f"elemamd_batch_backward_error failed ({rc})")
        return eta

    def close(self):
        if getattr(self, "_h", None):
            lib().elemamd_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- inputs
def uneven_sizes(count: int, lo: int, hi: int, kmax: int, seed: int):
    """n_i = floor(exp(U(ln lo, ln hi))) and k_i uniform in 1..kmax, from the
    splitmix64 stream ``seed`` (numbers 2i and 2i+1)."""
    from .inputs import _u01, splitmix64_at

    idx = np.arange(count, dtype=np.uint64)
    u = _u01(splitmix64_at(seed, np.uint64(2) * idx))
    v = _u01(splitmix64_at(seed, np.uint64(2) * idx + np.uint64(1)))
    n = np.floor(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int32)
    n = np.clip(n, lo, hi)
    k = np.minimum((v * kmax).astype(np.int32) + 1, kmax)
    return n, k


def fill_gram(plan: BatchPlan, A, B, seed: int, shift: float = 0.1):
    """A_i = G_i G_i^T / n_i + shift*I with G_i N(0,1) (stream seed + i), B_i
    N(0,1) (stream seed + i, numbers after G_i): the uneven-batch workload."""
    import torch

    from .inputs import normal_device

    for i in range(len(plan.n)):
        n, k = int(plan.n[i]), int(plan.k[i])
        G = normal_device(n, n + k, seed + i).tensor()  # n x (n + k)
        Gi = G[:, :n]
        Ai = plan.a_view(A, i)
        Ai.copy_(Gi @ Gi.t())
        Ai.div_(n)
        Ai.diagonal().add_(shift)
        plan.b_view(B, i).copy_(G[:, n:])


def fill_gram_host(n: int, k: int, seed: int, shift: float = 0.1):
    """Host copy of matrix/right-hand sides i of ``fill_gram`` (seed = seed + i)."""
    from .inputs import normal_host

    G = normal_host(n, n + k, seed)
    Gi = G[:, :n]
    A = Gi @ Gi.T / n + shift * np.eye(n)
    return np.asfortranarray(A), np.asfortranarray(G[:, n:])
