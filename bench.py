"""Benchmark of the MI355X tiled Cholesky (one GPU).

Workload (default): n = 32768, tile b = 512, A = D (B B^T/n + delta I) D with B iid
N(0,1) (torch generator, seed 2022), delta = 1e-2, D_ii = 10^(2U), U ~ U(0,1)
(seed 2023): the ill-conditioned mixed-precision configuration. A step is one
tiled Cholesky factorization of A on the device (the refreshed working copy of
A is restored from a resident FP64 copy inside the step), through the
persistent DAG executor.

Prints one JSON line. `value` = F(n)/t in TFLOP/s with F(n) = n^3/3 + n^2/2 + n/6.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_PEAK = 78.6  # TFLOP/s, MI355X dense FP64 matrix


def flops(n: int) -> float:
    return n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_matrix(n: int, seed_b: int, seed_d: int, delta: float, scaled: bool):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed_b)
    B = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g)
    A = torch.matmul(B, B.T)
    del B
    A /= n
    A.diagonal().add_(delta)
    if scaled:
        g.manual_seed(seed_d)
        u = torch.rand((n,), dtype=torch.float64, device="cuda", generator=g)
        d = torch.pow(10.0, 2.0 * u)
        A *= d[:, None]
        A *= d[None, :]
    torch.cuda.synchronize()
    return A  # symmetric, so row-major == column-major


def cpu_baseline(n_sample: int, b: int, threads: int):
    """The oracle (Taskflow+Threadpool restatement with sequential MKL per task) on the host."""
    import numpy as np
    from oracle import oracle as orc
    orc.build()
    kind = orc.set_backend("mkl") if orc.have_mkl() else orc.set_backend("textbook")
    A = orc.spd_bbt(n_sample, 2020)
    times = []
    for _ in range(3):
        _, st = orc.cholesky(A, b, n_threads=threads, prio_kind=2)
        times.append(st["total"])
    t = float(np.median(times))
    return {"value": flops(n_sample) / t / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"n={n_sample} b={b} FP64 A=BB^T/n+I, median of 3 runs, {t:.3f} s/run, backend={kind}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--b", type=int, default=512)
    ap.add_argument("--prio", type=int, default=2)
    ap.add_argument("--schedule", default="persistent", choices=["persistent", "stream"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=8192)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch
    import tasktorrent_amd as tt

    n, b = args.n, args.b
    sched = tt.TTAD_SCHED_PERSISTENT if args.schedule == "persistent" else tt.TTAD_SCHED_STREAM
    ctx = tt.Context(0)
    log(f"device {ctx.info()}")
    A = make_matrix(n, 2022, 2023, 1e-2, True)
    sizes = [b] * (n // b) + ([n % b] if n % b else [])
    A64 = tt.TiledMatrix(ctx, sizes, "f64")
    A64.set(A)
    W = tt.TiledMatrix(ctx, sizes, "f64")
    del A
    torch.cuda.empty_cache()
    ctx.synchronize()

    def step():
        W.copy_from(A64)
        info = ctx.potrf_tiled(W, sched, args.prio, "f64")
        if info != 0:
            raise RuntimeError(f"factorization failed, info={info}")
        return ctx.last_elapsed()

    for w in range(args.warmup):
        t = step()
        log(f"warmup {w}: factor {t * 1e3:.2f} ms")
    ctx.synchronize()
    torch.cuda.synchronize()
    dev_times = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        dev_times.append(step())
        log(f"step {s}: factor {dev_times[-1] * 1e3:.2f} ms")
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    t_factor = sum(dev_times) / len(dev_times)
    F = flops(n)
    value = F / (ms_per_step * 1e-3) / 1e12
    achieved = F / t_factor / 1e12
    out = {
        "metric": "Cholesky TFLOP/s n=32768 (mixed+IR vs FP64, <=1e-12 resid); p50/p99 solve ms",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"FP64 tiled Cholesky n={n} b={b} A=D(BB^T/n+1e-2 I)D, {args.schedule} DAG",
                   "n": n, "tile": b, "prio_kind": args.prio, "schedule": args.schedule},
        "factor_ms_device": round(t_factor * 1e3, 3),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK, 4), "traffic": None,
                     "kernel": "k_persistent (whole factorization, one launch)"},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle n={args.cpu_n} b={b} threads={threads}")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, b, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
