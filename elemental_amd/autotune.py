"""Autotuner of the batched factor-and-solve's size buckets.

For each size bucket it times the captured batch on a sample of matrices
drawn log-uniformly inside the bucket, once per candidate shape -- the block
size nb (the diagonal-block / panel width and the depth of the trailing
update's 128 x 128 tiles) and the launch form: one potrf, panel and trailing
launch per block step, or one persistent launch of 256, 512 or 1024
workgroups that each take whole matrices -- and keeps the
fastest.  The result is cached as JSON (``batch.TUNING_FILE``), keyed by
precision, as (max_n, nb, grid) triples that ``batch.tuned_buckets`` reads.

    python -m elemental_amd.autotune [--precisions f64 f32] [--per-bucket 256]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import batch as EB

BOUNDS = [32, 64, 128, 256, 512, 1 << 30]
CANDIDATE_NB = [16, 32, 64, 128]
CANDIDATE_GRID = [0, 256, 512, 1024]  # 0: stepwise launches


def candidates(top):
    """(nb, grid) shapes worth timing for a bucket whose largest n is top."""
    out = []
    for nb in CANDIDATE_NB:
        if nb > 2 * top and nb != CANDIDATE_NB[0]:
            continue
        for g in CANDIDATE_GRID:
            if g > 0 and nb > EB.PERSIST_MAX_NB:
                continue
            out.append((nb, g))
    return out


def _time_plan(n, k, precision, nb, grid, reps=3, seed=99):
    import torch

    plan = EB.BatchPlan(n, k, precision, buckets=[(1 << 30, nb, grid)])
    plan.set_refinement(4)
    A, B = plan.empty_a(), plan.empty_b()
    EB.fill_gram(plan, A, B, seed)
    A0, B0 = A.clone(), B.clone()
    plan.solve(A, B, use_graph=True)  # capture + first replay
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        A.copy_(A0)
        B.copy_(B0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.solve(A, B, use_graph=True)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    info, _, _ = plan.results()
    plan.close()
    if (info != 0).any():
        return float("inf")
    return best


def tune(precisions=("f64", "f32"), per_bucket=256, hi=1024, seed=5, path=EB.TUNING_FILE, log=print):
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    lo = 16
    for prec in precisions:
        buckets, detail = [], []
        prev = lo - 1
        for b, ub in enumerate(BOUNDS):
            top = min(ub, hi)
            if top <= prev:
                break
            n, k = EB.uneven_sizes(per_bucket, prev + 1, top, 8, seed + b)
            times = {}
            for nb, g in candidates(top):
                times[(nb, g)] = _time_plan(n, k, prec, nb, g)
            nb_best, g_best = min(times, key=times.get)
            buckets.append((ub, nb_best, g_best))
            detail.append({"max_n": ub, "sample_n": [int(prev + 1), int(top)],
                           "ms": {f"nb={c} grid={g}": round(t * 1e3, 3) for (c, g), t in times.items()},
                           "nb": nb_best, "grid": g_best})
            log(f"[autotune] {prec} n in [{prev + 1}, {top}]: " +
                ", ".join(f"nb={c}/g={g}: {t * 1e3:.2f} ms" for (c, g), t in times.items()) +
                f" -> nb={nb_best} grid={g_best}")
            prev = top
        out[prec] = {"buckets": buckets, "detail": detail, "per_bucket": per_bucket}
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    return out


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--precisions", nargs="+", default=["f64", "f32"])
    p.add_argument("--per-bucket", type=int, default=256)
    p.add_argument("--out", default=EB.TUNING_FILE)
    a = p.parse_args()
    tune(tuple(a.precisions), a.per_bucket, path=a.out)


if __name__ == "__main__":
    main()
