"""Benchmark: HPD factor-and-solve throughput and latency on one MI355X.

Workload (one "step" = one full solve of A X = B with inputs resident in HBM):
  n = 16384, kappa = 1e6 (A = Q diag(s) Q^T, Q = Householder-QR factor of an
  N(0,1) matrix, s log-spaced from 1 to 1e-6), B 16384 x 32 N(0,1).
  * mixed+IR : low-precision factor (FP64 diagonal potrf, FP32-stored panels,
               FP32, split-BF16x3 or split-BF16x6 trailing update) refined with FP64 residuals
               to normwise backward error <= 1e-14 (elemamd_hpd_solve_mixed_f64)
  * plain FP64 : Cholesky + SolveAfter in FP64 (elemamd_hpd_solve_f64)
The headline value is mixed+IR solves/s; plain FP64 is reported beside it,
with p50/p99 latency, backward/forward errors, the dominant kernel's roofline
and the oracle's CPU baseline.

Other workloads (each prints its own single JSON line):
  --mode batch : 8192 matrices, n log-uniform in [16, 1024], A_i = G G^T/n + 0.1 I,
                 1-8 right-hand sides each, one captured batch (plain FP64 and FP32+IR)
  --mode sweep : n = 32768, kappa in {1e2, 1e4, 1e6, 1e8}, B 32768 x 8: plain FP64 vs
                 FP32 vs split-BF16x3 vs split-BF16x6 trailing update with FP64 refinement
  --mode serve : Poisson arrivals of factor+solve requests, n log-uniform in [64, 8192],
                 kappa = 1e4, 30% reuse of a resident factor, at 3 arrival rates
  --mode inverse : HPDInverse and TriangularInverse at n (kappa as above) and the
                 generalized-definite eigensolver (AXBX, eigenvectors) at min(n, 4096)
  --mode roofline : only the trailing-update replays of the single workload (the
                 dominant kernel, both precisions) and a 1 GiB device copy; the
                 command profiled with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
                 for the HBM traffic figure (the copy calibrates the counters)
  --mode counters : one warm-up and one measured mixed solve at n, k on a matrix
                 built from elementwise operations only (symmetric random plus a
                 diagonal shift, SPD), so that a counter run profiles no library
                 GEMM: the command of the refinement kernels' traffic summary

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--mode single|batch|sweep|serve|roofline|inverse|counters]
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
# MFMA products per algorithmic product: split-BF16x3 issues 3 bf16 MFMAs per
# term of A A^T, split-BF16x6 6; "achieved" counts algorithmic flops only.
MFMA_TERMS = {"f64": 1, "f32": 1, "bf16x3": 3, "bf16x6": 6}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--k", type=int, default=32)
    p.add_argument("--kappa", type=float, default=1e6)
    p.add_argument("--nb", type=int, default=128)
    p.add_argument("--trailing", default="bf16x3", choices=["f32", "bf16x3", "bf16x6"])
    p.add_argument("--refine", default="cg", choices=["ir", "cg"],
                   help="refinement of the mixed solve: classic IR or preconditioned CG")
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--cpu-baseline", type=int, default=1, help="0 skips the CPU baseline")
    p.add_argument("--cpu-runs", type=int, default=5)
    p.add_argument("--mode", default="single",
                   choices=["single", "batch", "sweep", "serve", "roofline", "inverse", "counters"])
    p.add_argument("--count", type=int, default=8192, help="batch: number of matrices")
    p.add_argument("--requests", type=int, default=160, help="serve: requests per arrival rate")
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


UPDATE_DEPTH = {"f64": 256, "f32": 512, "bf16x3": 512, "bf16x6": 512}  # columns per trailing pass (lapack.hpp)


def trailing_launches(n, nb, trailing):
    """(m, n_cols, k) of every trailing-update launch the blocked factorization
    makes: per group of G = depth/nb block columns, the in-group updates of
    each block column, the next group's block columns and the rest."""
    G = max(1, UPDATE_DEPTH[trailing] // nb)
    gw = min(G * nb, n)
    out = []
    for c0 in range(0, n, gw):
        c1 = min(n, c0 + gw)
        for k in range(c0, c1, nb):
            b = min(nb, n - k)
            if k > c0:
                out.append((n - k, b, k - c0))
        m2 = n - c1
        if m2 <= 0:
            break
        bn = min(gw, m2)
        if m2 > bn:
            out.append((m2 - bn, m2 - bn, c1 - c0))
        out.append((m2, bn, c1 - c0))
    return out


def syrk_roofline(n, nb, trailing, stream):
    """Replay the trailing-update launches of the blocked factorization that
    update the rest of the trailing matrix (same shapes, same kernel; they
    carry all but a few percent of the factorization's flops) on `stream`
    between HIP events.  Returns (average ms per launch, algorithmic flops
    per launch, launches)."""
    import torch

    from elemental_amd import lib

    shapes = trailing_launches(n, nb, trailing)
    kmax = max(k for _, _, k in shapes)
    dt = torch.float64 if trailing == "f64" else torch.float32
    P = torch.randn(kmax, n, dtype=dt, device="cuda")  # column-major n x kmax panel, ld n
    C = torch.zeros(n, n, dtype=dt, device="cuda")
    mode = {"f32": 1, "bf16x3": 2, "bf16x6": 3}.get(trailing)
    s = ctypes.c_void_p(stream.cuda_stream)

    def launch(m, k):
        if trailing == "f64":
            return lib().elemamd_herk_f64(b"L", b"N", m, k, -1.0, ctypes.c_void_p(P.data_ptr()), n, 1.0,
                                          ctypes.c_void_p(C.data_ptr()), n, s)
        return lib().elemamd_herk_f32(mode, m, k, -1.0, ctypes.c_void_p(P.data_ptr()), n, 1.0,
                                      ctypes.c_void_p(C.data_ptr()), n, s)

    # the square launches (rest of the trailing matrix) carry ~all the flops
    rest = [(m, nc, k) for m, nc, k in shapes if m == nc]
    for m, _, k in rest[:2]:
        assert launch(m, k) == 0
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for m, _, k in rest:
        assert launch(m, k) == 0
    ev1.record(stream)
    torch.cuda.synchronize()
    total_ms = ev0.elapsed_time(ev1)
    flops = sum(m * (m + 1) * k for m, _, k in rest)  # lower triangle: m(m+1)/2 entries x 2k
    return total_ms / len(rest), flops / len(rest), len(rest)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "trailing_update_traffic_n16384.md")


def profiled_traffic(n, nb, trailing):
    """HBM bytes per trailing-update launch of the replay, as measured by
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on `bench.py --mode roofline` and
    recorded in the committed profile summary (lines
    `traffic_bytes_per_launch[<trailing>] = <bytes>`); None when the summary
    is absent or was taken on another workload."""
    import re

    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    if f"n={n} nb={nb}" not in text:
        return None
    m = re.search(r"traffic_bytes_per_launch\[" + re.escape(trailing) + r"\] = ([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def bench_roofline(args):
    """Trailing-update replays (the selected precision, then FP64) and a 1 GiB
    device copy, for counter collection."""
    import torch

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    out = {"mode": "roofline", "n": args.n, "nb": args.nb}
    for tr in (args.trailing, "f64"):
        avg_ms, flops, launches = syrk_roofline(args.n, args.nb, tr, stream)
        out[tr] = {"avg_launch_ms": avg_ms, "flops_per_launch": flops, "launches": launches,
                   "shapes": [[m, k] for m, nc, k in trailing_launches(args.n, args.nb, tr) if m == nc]}
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_bytes"] = src.numel() * 8
    return out


def host_cpu_facts():
    """CPU model, online CPUs and the affinity mask of this process."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return model, os.cpu_count() or 1, affinity


def cpu_baseline(A_host, B_host, runs):
    """The oracle's LVar3 (nb=128, scalar LVar3Unb on diagonal blocks, host
    BLAS dtrsm/dsyrk) + SolveAfter, on the same matrix, median of `runs`
    after one warm-up, with a residual self-check per run.  Threads: every
    core of the affinity mask, capped by OMP_NUM_THREADS when it is set (the
    host's CPU share per GPU).  Returns (record, X of the last run)."""
    from oracle import oracle as O

    model, online, affinity = host_cpu_facts()
    cores = affinity
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        cores = min(cores, int(omp))
    O.bind_host_blas(cores)
    blas = O.host_blas_info()
    n, k = B_host.shape
    times = []
    X = None
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
        del F
    med = float(np.median(times))
    rec = {"value": 1.0 / med, "unit": "solves/s", "cores": cores, "kind": "port",
           "sample": f"{runs} full factor+solve runs of the same n={n} k={k} matrix after 1 warm-up, median "
                     f"{med:.3f} s (min {min(times):.3f}, max {max(times):.3f}); oracle LVar3(nb=128)+SolveAfter "
                     f"over {blas.get('config', blas['blas'])} ({blas.get('threading', '?')} threading, "
                     f"{cores} threads); CPU {model}, {online} online, {affinity} in the affinity mask"
                     + (f", OMP_NUM_THREADS={omp}" if omp else ""),
           "cpu_model": model, "cpus_online": online, "cpus_affinity": affinity,
           "blas": blas}
    return rec, X


def vendor_fp64(A, B, steps, warmup, stream):
    """torch.linalg.cholesky + torch.cholesky_solve (ROCm's rocSOLVER /
    hipBLAS) on the same matrix: a reference point beside the plain FP64
    path, not a target."""
    import torch

    At, Bt = A.tensor(), B.tensor()
    ms = []
    for it in range(warmup + max(1, min(steps, 5))):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        Lt = torch.linalg.cholesky(At)
        Xt = torch.cholesky_solve(Bt, Lt)
        e1.record(stream)
        torch.cuda.synchronize()
        if it >= warmup:
            ms.append(e0.elapsed_time(e1))
        del Lt
    m = float(np.median(ms))
    n = At.shape[0]
    del Xt
    torch.cuda.empty_cache()
    return {"what": "torch.linalg.cholesky + torch.cholesky_solve, FP64", "ms": round(m, 3),
            "solves_per_s": round(1e3 / m, 4), "frac_fp64_peak": round(n ** 3 / 3 / (m * 1e-3) / PEAK["f64"], 4)}


# ------------------------------------------------------------ other modes
def _common(metric, unit, value, args, ms, dtype, config, higher=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def bench_batch(args):
    """Uneven batch: every matrix factored and solved in one captured launch
    sequence, plain FP64 and FP32 + FP64 refinement."""
    import torch

    from elemental_amd import batch as EB

    n, k = EB.uneven_sizes(args.count, 16, 1024, 8, args.seed)
    flops = float(np.sum(n.astype(np.float64) ** 3 / 3 + 2.0 * n.astype(np.float64) ** 2 * k))
    res = {}
    for prec in ("f64", "f32"):
        plan = EB.BatchPlan(n, k, prec)
        plan.set_refinement(4)
        A0, B0 = plan.empty_a(), plan.empty_b()
        t0 = time.time()
        EB.fill_gram(plan, A0, B0, args.seed)
        torch.cuda.synchronize()
        gen = time.time() - t0
        A, B = A0.clone(), B0.clone()
        times = []
        for it in range(args.warmup + args.steps):
            A.copy_(A0)
            B.copy_(B0)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            plan.solve(A, B, use_graph=True)
            e1.record()
            torch.cuda.synchronize()
            if it >= args.warmup:
                times.append(e0.elapsed_time(e1))
        info, eta_r, sweeps = plan.results()
        eta = plan.backward_error(A0, B, B0)
        ms = float(np.mean(times))
        res[prec] = {"solves_per_s": round(len(n) / (ms * 1e-3), 1), "ms": round(ms, 3),
                     "p50_ms": round(pct(times, 50), 3), "p99_ms": round(pct(times, 99), 3),
                     "gflops": round(flops / (ms * 1e-3) / 1e9, 1), "max_backward_error": float(eta.max()),
                     "all_le_1e-14": bool((eta <= 1e-14).all()), "failed": int((info != 0).sum()),
                     "buckets": plan.buckets, "input_generation_s": round(gen, 2)}
        if prec != "f64":
            res[prec]["max_sweeps"] = int(sweeps.max())
            res[prec]["mean_sweeps"] = round(float(sweeps.mean()), 2)
        plan.close()
        del A, B, A0, B0
        torch.cuda.empty_cache()
    best = max(res, key=lambda p: res[p]["solves_per_s"] if res[p]["all_le_1e-14"] else -1)
    out = _common("batched HPD solves/s at per-matrix FP64 backward err <=1e-14", "solves/s",
                  res[best]["solves_per_s"], args, res[best]["ms"], "f64",
                  {"workload": f"{len(n)} matrices n~logU[16,1024] A=GG^T/n+0.1I k~U[1,8]",
                   "count": len(n), "sum_n": int(n.sum()), "best": best, "parallelism": "single GPU"})
    out.update(res)
    return out


def bench_sweep(args):
    """Precision sweep at n = 32768: one orthogonal factor Q, A = Q diag(s) Q^T
    for each kappa; plain FP64 vs FP32 vs split-BF16x3 trailing updates."""
    import torch

    import elemental_amd as E
    from elemental_amd import inputs
    from elemental_amd.mixed import workspace_bytes

    n, k = (args.n if args.n != 16384 else 32768), 8
    G = inputs.normal_device(n, n, args.seed).tensor().clone()
    Q, R = torch.linalg.qr(G)
    Q.mul_(torch.sign(torch.diagonal(R)).unsqueeze(0))
    del R, G
    torch.cuda.empty_cache()
    B = inputs.normal_device(n, k, args.seed + 1)
    A = E.Matrix(n, n)
    Aw = E.Matrix(n, n)
    X = E.Matrix(n, k)
    rows = []
    for kappa in (1e2, 1e4, 1e6, 1e8):
        s_ = torch.logspace(0.0, -float(np.log10(kappa)), n, dtype=torch.float64, device="cuda")
        At = A.tensor()
        At.copy_((Q * s_.unsqueeze(0)) @ Q.t())
        At.copy_(0.5 * (At + At.t()))
        anorm = sym_inf_norm(A)
        # plain FP64 (a warm-up solve first: the library's scratch buffers
        # are allocated on first use)
        t64, eta64 = float("nan"), float("nan")
        for rep_ in range(2):
            Aw.tensor().copy_(At)
            X.tensor().copy_(B.tensor())
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            try:
                E.HPDSolve(E.LOWER, E.NORMAL, Aw, X)
                torch.cuda.synchronize()
                t64 = time.perf_counter() - t0
                eta64 = eta_inf(A, X, B, anorm)
            except E.LogicError:
                t64, eta64 = float("nan"), float("nan")
        Xref = X.tensor().clone()
        row = {"kappa": kappa, "f64": {"ms": round(t64 * 1e3, 2), "backward_error": eta64,
                                       "tflops": round(n ** 3 / 3 / t64 / 1e12, 2),
                                       "frac_fp64_peak": round(n ** 3 / 3 / t64 / PEAK["f64"], 3)}}
        for trailing in ("f32", "bf16x3", "bf16x6"):
            opts = E.IROptions(trailing=trailing, refine=args.refine)
            ws = torch.empty(workspace_bytes(n, k, opts), dtype=torch.uint8, device="cuda")
            for rep_ in range(2):  # a warm-up solve, then the timed one
                X.tensor().copy_(B.tensor())
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                rep = E.HPDSolveMixed(A, X, opts, workspace=ws)
                torch.cuda.synchronize()
                t = time.perf_counter() - t0
            eta = eta_inf(A, X, B, anorm)
            fwd = float(torch.linalg.norm(X.tensor() - Xref) / torch.linalg.norm(Xref))
            peak = PEAK["f32"] if trailing == "f32" else PEAK["bf16"]
            row[trailing] = {"ms": round(t * 1e3, 2), "backward_error": eta, "forward_error_vs_fp64": fwd,
                             "ir_iterations": rep.iterations, "gmres_iterations": rep.gmres_iterations,
                             "used_gmres": rep.used_gmres, "shifted": rep.shifted, "converged": rep.converged,
                             "fp64_finish": bool(rep.fallback), "lowprec_info": rep.lowprec_info,
                             "speedup_vs_fp64": round(t64 / t, 3) if t64 == t64 else None,
                             "factor_frac_of_peak": round(n ** 3 / 3 / t / peak, 4)}
            del ws
        rows.append(row)
        print(json.dumps(row), file=sys.stderr, flush=True)
    best = min(rows[2][t]["ms"] for t in ("f32", "bf16x3", "bf16x6"))
    out = _common("precision sweep: mixed+IR vs plain FP64 at n=32768 (ms per solve, kappa=1e6 headline: the "
                  "fastest trailing mode)", "ms", best, args, best, "f64",
                  {"workload": f"HPD factor+solve n={n} k={k} kappa in 1e2..1e8", "n": n, "k": k,
                   "refine": args.refine, "parallelism": "single GPU"}, higher=False)
    out["sweep"] = rows
    return out


def bench_inverse(args):
    """The direct callers of the Cholesky factor: HPDInverse (CholeskyLVar2,
    n^3 flops), TriangularInverse (LVar3, n^3/3 flops) and the
    generalized-definite eigensolver (AXBX with eigenvectors), each timed on
    a resident matrix with its accuracy beside it."""
    import torch

    import elemental_amd as E
    from elemental_amd import inputs

    torch.cuda.init()
    n = args.n
    out = {}
    # HPDInverse at kappa = 1e6 (the single workload's matrix construction)
    A = inputs.spectrum_device(n, args.kappa, seed=args.seed)
    W = E.Matrix(n, n)
    times = []
    for it in range(args.warmup + args.steps):
        W.tensor().copy_(A.tensor())
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        E.HPDInverse(E.LOWER, W)
        e1.record()
        torch.cuda.synchronize()
        if it >= args.warmup:
            times.append(e0.elapsed_time(e1))
    Wi = torch.tril(W.tensor()) + torch.tril(W.tensor(), -1).t()
    # ||inv(A) A - I||_inf / (||A||_inf ||inv(A)||_inf) on a 64-column sample
    cols = torch.arange(0, n, max(1, n // 64), device="cuda")
    At = A.tensor()
    S = torch.tril(At) + torch.tril(At, -1).t()
    R = Wi @ S[:, cols]
    R[cols, torch.arange(len(cols), device="cuda")] -= 1.0
    rel = float(R.abs().max() / (S.abs().sum(1).max() * Wi.abs().sum(1).max()))
    ms = float(np.median(times))
    out["hpd_inverse"] = {"n": n, "ms": round(ms, 3), "tflops": round(n ** 3 / (ms * 1e-3) / 1e12, 2),
                          "frac_fp64_peak": round(n ** 3 / (ms * 1e-3) / PEAK["f64"], 4),
                          "relative_residual_sampled": rel}
    del Wi, R, S
    # TriangularInverse of the HPD factor
    L = E.Matrix(n, n)
    L.tensor().copy_(A.tensor())
    E.Cholesky(E.LOWER, L)
    Lt = torch.tril(L.tensor()).clone()
    times = []
    for it in range(args.warmup + args.steps):
        W.tensor().copy_(Lt)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        E.TriangularInverse(E.LOWER, E.NON_UNIT, W)
        e1.record()
        torch.cuda.synchronize()
        if it >= args.warmup:
            times.append(e0.elapsed_time(e1))
    X = torch.tril(W.tensor())
    R = X @ Lt[:, cols]
    R[cols, torch.arange(len(cols), device="cuda")] -= 1.0
    ms = float(np.median(times))
    out["triangular_inverse"] = {"n": n, "ms": round(ms, 3), "tflops": round(n ** 3 / 3 / (ms * 1e-3) / 1e12, 2),
                                 "frac_fp64_peak": round(n ** 3 / 3 / (ms * 1e-3) / PEAK["f64"], 4),
                                 "max_abs_inv_L_L_minus_I_sampled": float(R.abs().max())}
    del L, Lt, X, R, W, A
    torch.cuda.empty_cache()
    # generalized-definite eigenproblem A x = lambda B x, n = 4096, with vectors
    ng = min(n, 4096)
    G = inputs.normal_device(ng, ng, args.seed + 5).tensor()
    Ag = E.Matrix(ng, ng)
    Ag.tensor().copy_(G + G.t())
    Bg = inputs.spectrum_device(ng, 1e3, seed=args.seed + 6)
    A0, B0 = Ag.tensor().clone(), Bg.tensor().clone()
    times = []
    for it in range(1 + min(args.steps, 3)):
        Ag.tensor().copy_(A0)
        Bg.tensor().copy_(B0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        w, Xv = E.HermitianGenDefiniteEig(E.AXBX, E.LOWER, Ag, Bg, vectors=True)
        torch.cuda.synchronize()
        if it >= 1:
            times.append((time.perf_counter() - t0) * 1e3)
    wv, Xt = w.tensor()[:, 0], Xv.tensor()
    Rg = A0 @ Xt - (B0 @ Xt) * wv.unsqueeze(0)
    out["hermitian_gen_definite_eig"] = {
        "n": ng, "type": "AXBX", "vectors": True, "ms": round(float(np.median(times)), 2),
        "max_abs_AX_minus_BXL": float(Rg.abs().max()),
        "max_abs_XtBX_minus_I": float((Xt.t() @ B0 @ Xt - torch.eye(ng, dtype=torch.float64, device="cuda")).abs().max())}
    res = _common("HPD inverse / triangular inverse / generalized eigensolver (ms, one GPU)", "ms",
                  out["hpd_inverse"]["ms"], args, out["hpd_inverse"]["ms"], "f64",
                  {"workload": f"HPDInverse and TriangularInverse n={n} kappa={args.kappa:g}, "
                               f"HermitianGenDefiniteEig(AXBX) n={ng}", "parallelism": "single GPU"}, higher=False)
    res.update(out)
    return res


def bench_serve(args):
    """Serving loop under Poisson arrivals at three rates, plain FP64 and
    FP32 + FP64 refinement."""
    import threading

    import torch

    from elemental_amd import inputs
    from elemental_amd.serve import Server

    rng = np.random.default_rng(args.seed)
    P = 24
    sizes = np.floor(np.exp(rng.uniform(np.log(64), np.log(8192), P))).astype(int)
    pool = [inputs.spectrum_device(int(nn), 1e4, args.seed + i).tensor() for i, nn in enumerate(sizes)]
    k = 4
    torch.cuda.synchronize()

    def schedule(count, rate):
        arr = np.cumsum(rng.exponential(1.0 / rate, count))
        reqs, seen = [], []
        for j in range(count):
            if seen and rng.random() < 0.3:
                key, i = seen[int(rng.integers(len(seen)))]
                reqs.append((arr[j], key, i, True))
            else:
                i = int(rng.integers(P))
                key = 1000 * j + i
                seen.append((key, i))
                reqs.append((arr[j], key, i, False))
        return reqs

    def run(prec, reqs, slots=4):
        Bs = [inputs.normal_device(pool[i].shape[0], k, 7 + j).tensor() for j, (_, _, i, _) in enumerate(reqs)]
        Xs = [b.clone() for b in Bs]
        torch.cuda.synchronize()
        with Server(slots=slots, precision=prec) as S:
            tickets = [None] * len(reqs)
            t0 = time.perf_counter()

            def submitter():
                for j, (ta, key, i, reuse) in enumerate(reqs):
                    d = ta - (time.perf_counter() - t0)
                    if d > 0:
                        time.sleep(d)
                    tickets[j] = S.submit(key, pool[i], Xs[j], reuse=reuse)

            th = threading.Thread(target=submitter)
            th.start()
            th.join()
            res = [S.wait(t) for t in tickets]
            st = S.stats()
        lat = np.array([(r.t_done - r.t_submit) * 1e3 for r in res])
        span = max(r.t_done for r in res) - min(r.t_submit for r in res)
        ok = all(r.status == 0 for r in res)
        # FP64 backward error of a sample of results
        worst = 0.0
        for j in range(0, len(reqs), max(1, len(reqs) // 12)):
            A = pool[reqs[j][2]]
            Rr = Bs[j] - A @ Xs[j]
            an = A.abs().sum(dim=1).max()
            e = (Rr.abs().amax(dim=0) / (an * Xs[j].abs().amax(dim=0) + Bs[j].abs().amax(dim=0))).max()
            worst = max(worst, float(e))
        return {"solves_per_s": round(len(reqs) / span, 2), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3), "all_ok": ok,
                "max_backward_error_sampled": worst, "reuse_hits": int(st.reuse_hits),
                "evictions": int(st.evictions), "peak_reserved_GB": round(st.peak_reserved / 2**30, 2)}

    # capacity estimate: a closed loop (all requests queued at once) with the same mix
    cal = schedule(args.requests, 1e9)
    cap = {p: run(p, cal)["solves_per_s"] for p in ("f64", "f32")}
    out_rates = []
    for frac in (0.5, 0.8, 1.0):
        rate = frac * max(cap.values())
        reqs = schedule(args.requests, rate)
        row = {"arrival_rate_per_s": round(rate, 2)}
        for p in ("f64", "f32"):
            row[p] = run(p, reqs)
        out_rates.append(row)
        print(json.dumps(row), file=sys.stderr, flush=True)
    mid = out_rates[1]
    best = "f32" if mid["f32"]["all_ok"] and mid["f32"]["solves_per_s"] >= mid["f64"]["solves_per_s"] else "f64"
    out = _common("served HPD solves/s and p99 latency at FP64 backward err <=1e-14", "solves/s",
                  mid[best]["solves_per_s"], args, mid[best]["p50_ms"], "f64",
                  {"workload": f"Poisson arrivals, n~logU[64,8192] from a pool of {P} kappa=1e4 matrices, k={k}, "
                               f"30% reuse of a resident factor, 4 slots", "requests_per_rate": args.requests,
                   "headline_rate": mid["arrival_rate_per_s"], "best": best, "parallelism": "single GPU"})
    out["closed_loop_capacity"] = cap
    out["rates"] = out_rates
    return out


def bench_counters(args):
    """One warm-up and one measured mixed solve (split-BF16x3 factor, CG) at
    n x k on A = (R + R^T)/2 + 2.2 sqrt(n) I, R iid N(0, 1): its spectrum lies
    in about [2.2 - 1.42, 2.2 + 1.42] sqrt(n), so A is SPD with a condition
    number near 4.6.  Only elementwise torch operations build it, so a
    rocprofv3 counter run sees the library's kernels and no vendor GEMM."""
    import torch

    import elemental_amd as E

    n, k = args.n, args.k
    g = torch.Generator(device="cuda")
    g.manual_seed(args.seed)
    A, B = E.Matrix(n, n), E.Matrix(n, k)
    R = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
    At = A.tensor()
    At.copy_(R)
    At.add_(R.t())
    At.mul_(0.5)
    del R
    At.diagonal().add_(2.2 * n ** 0.5)
    B.tensor().copy_(torch.randn(n, k, generator=g, device="cuda", dtype=torch.float64))
    opts = E.IROptions(nb=args.nb, trailing=args.trailing, refine=args.refine)
    runs = []
    for _ in range(2):
        X = E.Matrix(n, k)
        X.tensor().copy_(B.tensor())
        torch.cuda.synchronize()
        t = time.perf_counter()
        rep = E.HPDSolveMixed(A, X, opts)
        torch.cuda.synchronize()
        runs.append({"ms": (time.perf_counter() - t) * 1e3, "iterations": rep.iterations,
                     "backward_error": rep.backward_error, "converged": rep.converged})
    return {"metric": "mixed solve for counter runs (elementwise-built SPD matrix)", "value": runs[-1]["ms"],
            "unit": "ms", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": runs[-1]["ms"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"mixed solve n={n} k={k}, A = (R + R^T)/2 + 2.2 sqrt(n) I", "n": n, "k": k,
                       "trailing": args.trailing, "refine": args.refine, "parallelism": "single GPU"},
            "runs": runs}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "the multi-GPU layout is described in DESIGN.md (sharding section)", file=sys.stderr)
        sys.exit(2)
    if args.mode != "single":
        out = {"batch": bench_batch, "sweep": bench_sweep, "serve": bench_serve,
               "roofline": bench_roofline, "inverse": bench_inverse, "counters": bench_counters}[args.mode](args)
        print(json.dumps(out))
        return
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
    opts = E.IROptions(nb=args.nb, trailing=args.trailing, refine=args.refine)
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
    try:
        vendor = vendor_fp64(A, B, args.steps, args.warmup, stream)
    except Exception as e:  # a reference point only: its failure is reported, not fatal
        vendor = {"error": f"{type(e).__name__}: {e}"[:200]}

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
    peak = PEAK["f32"] if args.trailing == "f32" else PEAK["bf16"]
    achieved = flops_launch / (avg_ms * 1e-3)
    avg64, fl64, _ = syrk_roofline(n, args.nb, "f64", stream)

    # ---- CPU baseline (oracle, host cores), bounded sample; its X is the
    # FP64 oracle solution the GPU results are compared with
    cpu = None
    vs_oracle = None
    if args.cpu_baseline:
        Ah = np.asfortranarray(A.tensor().cpu().numpy())
        Bh = np.asfortranarray(B.tensor().cpu().numpy())
        cpu, Xo = cpu_baseline(Ah, Bh, args.cpu_runs)
        Xm = X.tensor().cpu().numpy()
        Xp = Xplain.cpu().numpy()
        an = float(np.abs(Ah).sum(axis=1).max())

        def eta_host(Xs):
            R = Bh - Ah @ Xs
            return float(max(np.abs(R[:, c]).max() / (an * np.abs(Xs[:, c]).max() + np.abs(Bh[:, c]).max())
                             for c in range(k)))

        no = float(np.linalg.norm(Xo))
        vs_oracle = {"oracle": "CPU LVar3(nb=128)+SolveAfter in FP64 (the cpu_baseline run), same A and B",
                     "forward_error_mixed": float(np.linalg.norm(Xm - Xo) / no),
                     "forward_error_plain_fp64": float(np.linalg.norm(Xp - Xo) / no),
                     "backward_error_mixed_host": eta_host(Xm),
                     "backward_error_plain_fp64_host": eta_host(Xp),
                     "backward_error_oracle_host": eta_host(Xo)}
        del Ah, Bh, Xo

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
            "n": n, "k": k, "kappa": args.kappa, "nb": args.nb, "trailing": args.trailing, "refine": args.refine,
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
            "frac_fp64_peak": round(flops_factor / (mean_plain * 1e-3) / PEAK["f64"], 4),
        },
        "vendor_fp64": vendor,
        "speedup_mixed_vs_fp64": round(mean_plain / mean_mixed, 3),
        "roofline": {
            "bound": "mfma", "kernel": f"syrk_ln_{args.trailing}", "achieved": achieved / 1e12,
            "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
            "mfma_terms": MFMA_TERMS[args.trailing],
            "traffic": profiled_traffic(n, args.nb, args.trailing),
            "launches": launches, "avg_launch_ms": avg_ms, "flops_per_launch": flops_launch,
            "fp64_syrk": {"achieved": fl64 / (avg64 * 1e-3) / 1e12, "peak": PEAK["f64"] / 1e12,
                          "frac": fl64 / (avg64 * 1e-3) / PEAK["f64"], "avg_launch_ms": avg64,
                          "traffic": profiled_traffic(n, args.nb, "f64")},
        },
        "cpu_baseline": cpu,
        "vs_fp64_oracle": vs_oracle,
        "input_generation_s": round(gen_s, 2),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
