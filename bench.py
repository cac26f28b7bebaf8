"""Benchmark: HPD factor-and-solve throughput and latency on one MI355X.

Workload (one "step" = one full solve of A X = B with inputs resident in HBM):
  n = 16384, kappa = 1e6 (A = Q diag(s) Q^T, Q = Householder-QR factor of an
  N(0,1) matrix, s log-spaced from 1 to 1e-6), B 16384 x 32 N(0,1).
  * mixed+IR : low-precision factor (FP64 diagonal potrf, FP32-stored panels,
               FP32 or split-BF16x3 trailing update) refined with FP64 residuals
               to normwise backward error <= 1e-14 (elemamd_hpd_solve_mixed_f64)
  * plain FP64 : Cholesky + SolveAfter in FP64 (elemamd_hpd_solve_f64)
The headline value is mixed+IR solves/s; plain FP64 is reported beside it,
with p50/p99 latency, backward/forward errors, the dominant kernel's roofline
and the oracle's CPU baseline.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "f32": 157.3e12, "bf16": 2.5e15}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--k", type=int, default=32)
    p.add_argument("--kappa", type=float, default=1e6)
    p.add_argument("--nb", type=int, default=128)
    p.add_argument("--trailing", default="bf16x3", choices=["f32", "bf16x3"])
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--cpu-baseline", type=int, default=1, help="0 skips the CPU baseline")
    p.add_argument("--cpu-runs", type=int, default=3)
    return p.parse_args()


def eta_inf(A, X, B, anorm):
    """max_c ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), FP64 on device."""
    import torch

    from elemental_amd import lib

    n, k = B.Height(), B.Width()
    R = torch.empty((k, n), dtype=torch.float64, device="cuda")
    rc = lib().elemamd_sym_residual_f64(n, k, A.ptr(), A.LDim(), X.ptr(), X.LDim(), B.ptr(), B.LDim(),
                                        ctypes.c_void_p(R.data_ptr()), n,
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    rn = R.abs().amax(dim=1)
    xn = X.tensor().abs().amax(dim=0)
    bn = B.tensor().abs().amax(dim=0)
    return float((rn / (anorm * xn + bn)).max())


def sym_inf_norm(A):
    import torch

    T = torch.tril(A.tensor())
    rows = T.abs().sum(dim=1) + T.abs().sum(dim=0) - T.diagonal().abs()
    return float(rows.max())


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def syrk_roofline(n, nb, trailing, stream):
    """Time the trailing-update kernel at every step size of the blocked
    factorization (the same launches the factorization makes), with HIP
    events on the launching stream.  Returns (avg_ms, flops_per_launch)."""
    import torch

    from elemental_amd import lib

    steps = [(n - (j + 1) * nb, min(nb, n - j * nb)) for j in range((n + nb - 1) // nb)]
    steps = [(m, b) for m, b in steps if m > 0]
    if trailing == "f64":
        P = torch.randn(nb, n, dtype=torch.float64, device="cuda")
        C = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    else:
        P = torch.randn(nb, n, dtype=torch.float32, device="cuda")
        C = torch.zeros(n, n, dtype=torch.float32, device="cuda")
    mode = {"f32": 1, "bf16x3": 2}.get(trailing)
    s = ctypes.c_void_p(stream.cuda_stream)

    def launch(m, b):
        if trailing == "f64":
            return lib().elemamd_herk_f64(b"L", b"N", m, b, -1.0, ctypes.c_void_p(P.data_ptr()), n, 1.0,
                                          ctypes.c_void_p(C.data_ptr()), n, s)
        return lib().elemamd_herk_f32(mode, m, b, -1.0, ctypes.c_void_p(P.data_ptr()), n, 1.0,
                                      ctypes.c_void_p(C.data_ptr()), n, s)

    for m, b in steps[:3]:
        assert launch(m, b) == 0
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for m, b in steps:
        assert launch(m, b) == 0
    ev1.record(stream)
    torch.cuda.synchronize()
    total_ms = ev0.elapsed_time(ev1)
    flops = sum(m * (m + 1) * b for m, b in steps)  # lower triangle: m(m+1)/2 entries x 2b
    return total_ms / len(steps), flops / len(steps), len(steps)


def cpu_baseline(A_host, B_host, runs):
    """The oracle's LVar3 (nb=128, scalar LVar3Unb on diagonal blocks, host
    BLAS dtrsm/dsyrk) + SolveAfter, on the same matrix, median of `runs`
    after one warm-up, with a residual self-check per run."""
    from oracle import oracle as O

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    blas = O.bind_host_blas(cores)
    n, k = B_host.shape
    times = []
    for r in range(runs + 1):
        F, X = A_host.copy(order="F"), B_host.copy(order="F")
        t0 = time.perf_counter()
        info = O.hpd_solve(F, X, 128, host_blas=True)
        dt = time.perf_counter() - t0
        assert info == 0
        # self-check on a sample of columns: ||b - A x|| / (||A|| ||x||)
        R = B_host[:, :2] - A_host @ X[:, :2]
        rel = np.linalg.norm(R) / (np.linalg.norm(A_host) * np.linalg.norm(X[:, :2]))
        assert rel < 1e-14, rel
        if r > 0:
            times.append(dt)
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"{runs} full factor+solve runs of the same n={n} k={k} matrix after 1 warm-up, median "
                      f"{med:.3f} s; oracle LVar3(nb=128)+SolveAfter over {blas}"}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "the multi-GPU layout is described in DESIGN.md (sharding section)", file=sys.stderr)
        sys.exit(2)
    import torch

    import elemental_amd as E
    from elemental_amd import inputs

    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    n, k = args.n, args.k
    t0 = time.time()
    A = inputs.spectrum_device(n, args.kappa, seed=args.seed)
    B = inputs.normal_device(n, k, seed=args.seed + 1)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    anorm = sym_inf_norm(A)

    Awork = E.Matrix(n, n)
    X = E.Matrix(n, k)
    opts = E.IROptions(nb=args.nb, trailing=args.trailing)
    from elemental_amd.mixed import workspace_bytes

    ws = torch.empty(workspace_bytes(n, k, opts), dtype=torch.uint8, device=dev)

    def run_plain():
        Awork.tensor().copy_(A.tensor())
        X.tensor().copy_(B.tensor())
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        E.HPDSolve(E.LOWER, E.NORMAL, Awork, X, nb=args.nb)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    def run_mixed():
        X.tensor().copy_(B.tensor())
        torch.cuda.synchronize()
        t = time.perf_counter()
        rep = E.HPDSolveMixed(A, X, opts, workspace=ws)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3, rep

    # ---- plain FP64
    for _ in range(args.warmup):
        run_plain()
    torch.cuda.synchronize()
    plain_ms = [run_plain() for _ in range(args.steps)]
    eta_plain = eta_inf(A, X, B, anorm)
    Xplain = X.tensor().clone()

    # ---- mixed + IR (headline)
    for _ in range(args.warmup):
        run_mixed()
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    mixed = [run_mixed() for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall_mixed = time.perf_counter() - t_all
    mixed_ms = [m for m, _ in mixed]
    rep = mixed[-1][1]
    eta_mixed = eta_inf(A, X, B, anorm)
    fwd = float(torch.linalg.norm(X.tensor() - Xplain) / torch.linalg.norm(Xplain))
    ok = all(r.converged for _, r in mixed) and eta_mixed <= 1e-14

    flops_factor = n ** 3 / 3.0
    mean_mixed = float(np.mean(mixed_ms))
    mean_plain = float(np.mean(plain_ms))
    value = 1e3 / mean_mixed

    # ---- roofline of the dominant kernel (the trailing update of the mixed factor)
    avg_ms, flops_launch, launches = syrk_roofline(n, args.nb, args.trailing, stream)
    peak = PEAK["bf16"] if args.trailing == "bf16x3" else PEAK["f32"]
    achieved = flops_launch / (avg_ms * 1e-3)
    avg64, fl64, _ = syrk_roofline(n, args.nb, "f64", stream)

    # ---- CPU baseline (oracle, host cores), bounded sample
    cpu = None
    if args.cpu_baseline:
        Ah = np.asfortranarray(A.tensor().cpu().numpy())
        Bh = np.asfortranarray(B.tensor().cpu().numpy())
        cpu = cpu_baseline(Ah, Bh, args.cpu_runs)
        del Ah, Bh

    out = {
        "metric": "HPD solves/s + p99 latency at FP64 backward err <=1e-14: mixed+IR vs plain FP64",
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_mixed, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"HPD factor+solve n={n} k={k} kappa={args.kappa:g} (A=Q diag(s) Q^T, log-spaced s)",
            "n": n, "k": k, "kappa": args.kappa, "nb": args.nb, "trailing": args.trailing,
            "parallelism": "single GPU",
        },
        "mixed_ir": {
            "solves_per_s": round(value, 4), "p50_ms": round(pct(mixed_ms, 50), 3),
            "p99_ms": round(pct(mixed_ms, 99), 3), "backward_error_inf": eta_mixed,
            "forward_error_vs_fp64": fwd, "ir_iterations": rep.iterations,
            "gmres_iterations": rep.gmres_iterations, "used_gmres": rep.used_gmres, "converged": ok,
            "gflops_effective": round(flops_factor / (mean_mixed * 1e-3) / 1e9, 1),
        },
        "plain_fp64": {
            "solves_per_s": round(1e3 / mean_plain, 4), "p50_ms": round(pct(plain_ms, 50), 3),
            "p99_ms": round(pct(plain_ms, 99), 3), "backward_error_inf": eta_plain,
            "gflops": round((flops_factor + 2 * n * n * k) / (mean_plain * 1e-3) / 1e9, 1),
        },
        "speedup_mixed_vs_fp64": round(mean_plain / mean_mixed, 3),
        "roofline": {
            "bound": "mfma", "kernel": f"syrk_ln_{args.trailing}", "achieved": achieved / 1e12,
            "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
            "launches": launches, "avg_launch_ms": avg_ms, "flops_per_launch": flops_launch,
            "fp64_syrk": {"achieved": fl64 / (avg64 * 1e-3) / 1e12, "peak": PEAK["f64"] / 1e12,
                          "frac": fl64 / (avg64 * 1e-3) / PEAK["f64"], "avg_launch_ms": avg64},
        },
        "cpu_baseline": cpu,
        "input_generation_s": round(gen_s, 2),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
