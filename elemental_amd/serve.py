"""Serving loop: continuous batching of HPD factor-and-solve requests.

No reference counterpart (Elemental answers one HPDSolve call at a time);
each request is the computation of ``HPDSolve(LOWER, NORMAL, A, B)``
(``include/elemental/lapack-like/HPDSolve.hpp:18-28``) in plain FP64 or as a
low-precision factor with FP64 refinement.  The native server
(``elemamd_server_*``) runs ``slots`` requests at once on their own HIP
streams, admits queued requests as slots and HBM free up, and keeps factors
resident for requests that reuse them with new right-hand sides.
"""
from __future__ import annotations

import ctypes

from ._lib import ServeResult, ServerConfig, ServerStats, lib

PRECISIONS = {"f64": 0, "f32": 1, "bf16x3": 2}


class Server:
    def __init__(self, slots: int = 4, precision: str = "f64", nb: int = 128, tol: float = 1e-14,
                 memory_budget: int = 0):
        cfg = ServerConfig(slots, PRECISIONS[precision], nb, tol, memory_budget)
        h = ctypes.c_void_p()
        rc = lib().elemamd_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise ValueError(f"elemamd_server_create failed ({rc})")
        self._h = h
        self.precision = precision
        self._keep = {}  # ticket -> (A, B): tensors must outlive their request

    def submit(self, key: int, A, B, reuse: bool = False) -> int:
        """A: (n, n) float64 device tensor holding the matrix column-major (any
        torch layout whose element [i, j] is A(i, j) with unit row stride);
        B: (n, k) likewise, overwritten with X."""
        n, k = B.shape
        if A.stride(0) != 1 or B.stride(0) != 1:
            raise ValueError("A and B must be column-major (unit row stride)")
        import torch

        t = ctypes.c_int64()
        ready = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)  # A and B were produced on it
        rc = lib().elemamd_server_submit(self._h, key, 1 if reuse else 0, n, k, ctypes.c_void_p(A.data_ptr()),
                                         A.stride(1), ctypes.c_void_p(B.data_ptr()), B.stride(1), ready,
                                         ctypes.byref(t))
        if rc != 0:
            raise RuntimeError(f"elemamd_server_submit failed ({rc})")
        self._keep[t.value] = (A, B)
        return t.value

    def wait(self, ticket: int) -> ServeResult:
        r = ServeResult()
        rc = lib().elemamd_server_wait(self._h, ticket, ctypes.byref(r))
        self._keep.pop(ticket, None)
        if rc != 0:
            raise RuntimeError(f"elemamd_server_wait failed ({rc})")
        return r

    def evict(self, key: int):
        rc = lib().elemamd_server_evict(self._h, key)
        if rc != 0:
            raise RuntimeError(f"elemamd_server_evict failed ({rc})")

    def stats(self) -> ServerStats:
        s = ServerStats()
        lib().elemamd_server_stats(self._h, ctypes.byref(s))
        return s

    def close(self):
        if getattr(self, "_h", None):
            lib().elemamd_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def result_dict(r: ServeResult) -> dict:
    return {f: getattr(r, f) for f, _ in ServeResult._fields_}
