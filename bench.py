"""Benchmark of the MI355X tiled Cholesky solver (one GPU).

Workload (default; the mixed-precision configuration the metric is quoted on):
  n = 32768, tile b = 512, A = D (B B^T/n + 1e-2 I) D with B iid N(0,1) (torch
  generator seed 2022), D_ii = 10^(2U), U ~ U(0,1) (seed 2023), kappa ~ 1e4-1e6;
  16 right-hand sides iid N(0,1) (seed 2024). Inputs are resident in HBM before
  timing starts.

A step is one time-to-solution of A X = B to normwise backward error <= 1e-12:
  - fp64 leg: copy A's tiles into the working factor, persistent-DAG FP64 Cholesky,
    FP64 refinement (normally zero or one step);
  - mixed legs: cast A's tiles to FP32, persistent-DAG Cholesky with the trailing
    update in BF16x3 / FP32 / BF16 (FP64 diagonal POTRF), FP64 refinement with
    fallback to FP64 when refinement stalls.
`value` = F(n) / t_step of the headline mixed leg, in TFLOP/s of FP64-equivalent
work, F(n) = n^3/3 + n^2/2 + n/6 (so refinement time counts against it).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}  # TFLOP/s, MI355X dense MFMA
METRIC = "Cholesky TFLOP/s n=32768 (mixed+IR vs FP64, <=1e-12 resid); p50/p99 solve ms"


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
    return A  # symmetric, so row-major storage equals column-major


def measured_traffic(n: int, b: int):
    """HBM bytes per launch of the FP64 factorization kernel, as measured offline with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and scaled by a 1 GiB copy (profiles/), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.txt")
    try:
        for line in open(path):
            key, _, val = line.partition("=")
            if key.strip() == f"cholesky_fp64_n{n}_b{b}_bytes":
                return float(val)
    except OSError:
        pass
    return None


def cpu_baseline(n_sample: int, b: int, threads: int, prio: int = 0):
    """The oracle (Taskflow+Threadpool restatement, sequential MKL per task) on the host cores."""
    import numpy as np
    from oracle import oracle as orc
    orc.build()
    kind = orc.set_backend("mkl") if orc.have_mkl() else orc.set_backend("textbook")
    A = orc.spd_bbt(n_sample, 2020)
    orc.cholesky(A, b, n_threads=threads, prio_kind=prio)  # warm-up, as 2d_cholesky.cpp:54-61
    times = []
    for _ in range(3):
        _, st = orc.cholesky(A, b, n_threads=threads, prio_kind=prio)
        times.append(st["total"])
    t = float(np.median(times))
    return {"value": round(flops(n_sample) / t / 1e12, 4), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"FP64 tiled Cholesky n={n_sample} b={b}, A=BB^T/n+I, seed->join window, median of 3 "
                      f"({t:.3f} s/run), backend={kind}"}


def serving_leg(ctx, rate: float, n_req: int, seed: int = 7, prec: str = "bf16x3", pool: int = 24,
                budget: int = 16 << 30):
    """Poisson arrivals of SPD solves (continuous batching in one resident kernel).

    Sizes n = 64*round(exp(U(ln 64, ln 4096))/64); A = B B^T/n + I from a pool of `pool`
    pre-staged matrices (each request draws one); 1-8 RHS iid N(0,1); about 10% of
    requests solve with the cached factor of an earlier request of the same matrix.
    Latency = arrival -> solution copied back to host memory.
    """
    import numpy as np
    import torch
    import tasktorrent_amd as tt
    rng = np.random.default_rng(seed)
    sizes = []
    while len(sizes) < pool:
        n = int(64 * round(np.exp(rng.uniform(np.log(64), np.log(4096))) / 64))
        sizes.append(max(64, min(4096, n)))
    mats = []
    for q, n in enumerate(sizes):
        g = torch.Generator(device="cuda")
        g.manual_seed(10_000 + q)
        B = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g)
        A = B @ B.T / n
        A.diagonal().add_(1.0)
        b = 256
        M = tt.TiledMatrix(ctx, [b] * (n // b) + ([n % b] if n % b else []), "f64")
        M.set(A)
        mats.append(M)
    torch.cuda.synchronize()
    srv = tt.Server(ctx, hbm_budget=budget, max_tasks=1 << 21, max_requests=2048)
    # warm-up: one request per pool matrix, keeping factors for later reuse
    ids = [srv.submit(M, rng.standard_normal((M.n, 1)), prec=prec, keep_factor=True)[0] for M in mats]
    warm = srv.wait_all(ids, timeout_s=300)
    last_factor = {q: warm[ids[q]]["factor_id"] for q in range(pool)}
    gaps = rng.exponential(1.0 / rate, n_req)
    arrivals = np.cumsum(gaps)
    plan = []
    for i in range(n_req):
        q = int(rng.integers(pool))
        nrhs = int(rng.integers(1, 9))
        reuse = rng.random() < 0.1
        plan.append((q, nrhs, reuse))
    lat, qwait, etas, iters, statuses, reused = [], [], [], [], [], 0
    t0 = time.perf_counter()
    submitted, done = {}, 0
    i = 0
    while done < n_req:
        now = time.perf_counter() - t0
        while i < n_req and arrivals[i] <= now:
            q, nrhs, reuse = plan[i]
            M = mats[q]
            Bv = rng.standard_normal((M.n, nrhs))
            rid, _ = srv.submit(M, Bv, prec=prec, keep_factor=not reuse,
                                reuse_factor=last_factor[q] if reuse else -1)
            submitted[rid] = (arrivals[i], time.perf_counter() - t0, q)
            i += 1
        for r in srv.poll():
            if r["id"] not in submitted:
                continue
            arr, sub, q = submitted.pop(r["id"])
            t_now = time.perf_counter() - t0
            lat.append(t_now - arr)
            qwait.append(r["queue_s"])
            etas.append(r["backward_error"])
            iters.append(r["iterations"])
            statuses.append(r["status"])
            reused += r["reused"]
            if not r["reused"] and r["status"] == 0:
                last_factor[q] = r["factor_id"]
            done += 1
        if time.perf_counter() - t0 > 300:
            raise TimeoutError("serving leg did not finish in 300 s")
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    for M in mats:
        M.close()
    lat_ms = np.array(lat) * 1e3
    return {"rate_per_s": rate, "requests": n_req, "prec": prec, "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3), "mean_ms": round(float(lat_ms.mean()), 3),
            "solves_per_s": round(n_req / wall, 2), "wall_s": round(wall, 3),
            "max_backward_error": float(np.max(etas)), "ok": int(sum(1 for s_ in statuses if s_ == 0)),
            "reused": int(reused), "mean_ir_steps": round(float(np.mean(iters)), 2),
            "p99_queue_ms": round(float(np.percentile(np.array(qwait) * 1e3, 99)), 3),
            "workers": st["workers"], "matrix_pool": pool,
            "sizes": "n = 64*round(exp(U(ln64, ln4096))/64), 1-8 RHS, ~10% cached-factor requests"}


def trailing_update_leg(ctx, mn: int, k: int, reps: int = 10):
    """FP64 C -= A B^T with m = n = mn (ttad_dgemm, the GEMM items the factorization runs, as one plain grid),
    timed with events on the launch stream, beside rocBLAS DGEMM (torch.addmm) on the same operands."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    At = torch.randn((k, mn), dtype=torch.float64, device="cuda", generator=g)  # column-major mn x k
    Bt = torch.randn((k, mn), dtype=torch.float64, device="cuda", generator=g)
    C = torch.randn((mn, mn), dtype=torch.float64, device="cuda", generator=g)
    fl = 2.0 * mn * mn * k

    def timed(fn):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    t_ours = timed(lambda: ctx.gemm(mn, mn, k, -1.0, At, mn, Bt, mn, 1.0, C, mn))
    t_blas = timed(lambda: C.addmm_(Bt.t(), At, alpha=-1.0))
    del At, Bt, C
    torch.cuda.empty_cache()
    return {"shape": f"m=n={mn} k={k}", "ms": round(t_ours * 1e3, 3),
            "tflops": round(fl / t_ours / 1e12, 3), "frac": round(fl / t_ours / 1e12 / PEAK["f64"], 4),
            "rocblas_tflops": round(fl / t_blas / 1e12, 3)}


def panel_leg(ctx, m: int, b: int, reps: int = 10):
    """The panel steps alone, FP64: ttad_dpotrf on one b x b tile (synchronous: the event bracket includes its
    info readback) and ttad_dtrsm on an m x b column of tiles (B := B L^-T, the first panel's TRSMs at m = n - b),
    with the TRSM's HBM rate counted as one read and one write of the m x b panel."""
    import numpy as np
    import torch
    m = max(m, b)
    g = torch.Generator(device="cuda")
    g.manual_seed(13)
    Bm = torch.randn((b, b), dtype=torch.float64, device="cuda", generator=g)
    A = Bm @ Bm.T / b + torch.eye(b, dtype=torch.float64, device="cuda")  # symmetric: row- = column-major
    Ls = torch.empty_like(A)
    t_potrf = []
    for r in range(reps + 2):
        Ls.copy_(A)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        info = ctx.potrf(Ls, b, b)
        e1.record()
        torch.cuda.synchronize()
        if info != 0:
            raise RuntimeError(f"panel leg: potrf info={info}")
        if r >= 2:
            t_potrf.append(e0.elapsed_time(e1) / 1e3)
    P = torch.randn((b, m), dtype=torch.float64, device="cuda", generator=g)  # column-major m x b
    for _ in range(2):
        ctx.trsm(m, b, Ls, b, P, m)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        ctx.trsm(m, b, Ls, b, P, m)
    e1.record()
    torch.cuda.synchronize()
    t_trsm = e0.elapsed_time(e1) / 1e3 / reps
    if not bool(torch.isfinite(P).all()):
        raise RuntimeError("panel leg: TRSM produced non-finite values")
    tp = float(np.median(t_potrf))
    gbps = 2.0 * m * b * 8 / t_trsm / 1e9
    del A, Ls, P, Bm
    torch.cuda.empty_cache()
    return {"potrf_tile": b, "potrf_ms": round(tp * 1e3, 4), "potrf_tflops": round(b ** 3 / 3 / tp / 1e12, 4),
            "trsm_shape": f"m={m} n={b}", "trsm_ms": round(t_trsm * 1e3, 4), "trsm_gbps": round(gbps, 1),
            "trsm_hbm_frac": round(gbps / 8000.0, 4), "trsm_tflops": round(float(m) * b * b / t_trsm / 1e12, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--b", type=int, default=512)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--prec", default="bf16x3", help="headline mixed precision")
    ap.add_argument("--extra-precs", default="f32,bf16", help="other mixed legs (one timed run each)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=8192)
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson request rate of the serving leg (1/s)")
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-update", action="store_true", help="skip the FP64 trailing-update and panel-step legs")
    ap.add_argument("--update-mn", type=int, default=16384, help="m = n of the trailing-update kernel leg")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device copy first (reference kernel for HBM byte counters)")
    ap.add_argument("--prio", type=int, default=2,
                    help="PrioKind of the FP64 leg and the CPU baseline (the mixed legs use TTAD_MIXED_PRIO, default 4: bottom level)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it (--gpus 1)"}))
        sys.exit(2)

    import numpy as np
    import torch
    import tasktorrent_amd as tt

    n, b, nrhs = args.n, args.b, args.nrhs
    F = flops(n)
    ctx = tt.Context(0)
    log(f"device {ctx.info()}; n={n} b={b} nrhs={nrhs}")
    A = make_matrix(n, 2022, 2023, 1e-2, True)
    sizes = [b] * (n // b) + ([n % b] if n % b else [])
    A64 = tt.TiledMatrix(ctx, sizes, "f64")
    A64.set(A)
    del A
    torch.cuda.empty_cache()
    g = torch.Generator(device="cuda")
    g.manual_seed(2024)
    Bt = torch.randn((nrhs, n), dtype=torch.float64, device="cuda", generator=g)  # column-major n x nrhs
    Xt = torch.zeros_like(Bt)
    W64 = tt.TiledMatrix(ctx, sizes, "f64")
    ctx.synchronize()
    torch.cuda.synchronize()
    if args.calib_copy:  # a plain copy of known size: reads 1 GiB, writes 1 GiB
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- FP64 leg: factor + FP64 refinement ----
    def fp64_step():
        t0 = time.perf_counter()
        W64.copy_from(A64)
        info = ctx.potrf_tiled(W64, tt.TTAD_SCHED_PERSISTENT, args.prio, "f64")
        if info != 0:
            raise RuntimeError(f"FP64 factorization failed, info={info}")
        t_fac = ctx.last_elapsed()
        st = ctx.solve_ir(A64, W64, nrhs, Bt, Xt, tol=1e-12, max_iter=10)
        ctx.synchronize()
        return time.perf_counter() - t0, t_fac, st

    res = {}
    for w in range(args.warmup):
        t, tf, st = fp64_step()
        log(f"fp64 warmup {w}: tts {t*1e3:.1f} ms factor {tf*1e3:.1f} ms eta {st['backward_error']:.2e}")
    tts64, fac64 = [], []
    for s in range(args.steps):
        t, tf, st = fp64_step()
        tts64.append(t)
        fac64.append(tf)
        log(f"fp64 step {s}: tts {t*1e3:.1f} ms factor {tf*1e3:.1f} ms ({F/tf/1e12:.2f} TF/s) "
            f"eta {st['backward_error']:.2e} it {st['iterations']}")
    res["f64"] = {"tts_ms": round(1e3 * float(np.median(tts64)), 3), "factor_ms": round(1e3 * float(np.median(fac64)), 3),
                  "factor_tflops": round(F / float(np.median(fac64)) / 1e12, 3),
                  "tts_tflops": round(F / float(np.median(tts64)) / 1e12, 3),
                  "backward_error": st["backward_error"], "ir_steps": st["iterations"]}

    # ---- FP64 trailing-update kernel alone (C -= A B^T, the shape of one panel's update) ----
    if not args.no_update:
        res["f64_trailing_update"] = trailing_update_leg(ctx, args.update_mn, b)
        log(f"trailing update: {res['f64_trailing_update']}")
        res["f64_panel"] = panel_leg(ctx, n - b, b)
        log(f"panel steps: {res['f64_panel']}")

    # ---- mixed legs ----
    def mixed_step(prec):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = ctx.posv_mixed(A64, prec, nrhs, Bt, Xt, tol=1e-12, max_iter=30)
        t = time.perf_counter() - t0
        fac = ctx.last_elapsed() if not st["fell_back"] else float("nan")
        return t, st, fac

    headline = args.prec
    legs = [] if args.fp64_only else [headline] + [p for p in args.extra_precs.split(",") if p and p != headline]
    timed_steps = {}
    for prec in legs:
        reps = args.steps if prec == headline else 1
        for w in range(args.warmup):
            t, st, fac = mixed_step(prec)
            log(f"{prec} warmup {w}: tts {t*1e3:.1f} ms eta {st['backward_error']:.2e} it {st['iterations']} "
                f"fallback {st['fell_back']} rc {st['rc']}")
        ts, facs = [], []
        for s in range(reps):
            t, st, fac = mixed_step(prec)
            ts.append(t)
            facs.append(fac)
            log(f"{prec} step {s}: tts {t*1e3:.1f} ms eta {st['backward_error']:.2e} it {st['iterations']} "
                f"fallback {st['fell_back']} rc {st['rc']}")
        timed_steps[prec] = ts
        tmed = float(np.median(ts))
        res[prec] = {"tts_ms": round(1e3 * tmed, 3), "tts_tflops": round(F / tmed / 1e12, 3),
                     "backward_error": st["backward_error"], "ir_steps": st["iterations"],
                     "fell_back": bool(st["fell_back"]), "rc": st["rc"],
                     "speedup_vs_f64": round(res["f64"]["tts_ms"] / (1e3 * tmed), 3)}
        if not st["fell_back"]:
            res[prec]["last_factor_ms"] = round(1e3 * st["factor_s"], 3)

    if legs and res[headline]["rc"] == 0:
        ts = timed_steps[headline]
        value = F / float(np.median(ts)) / 1e12
        dtype = headline
        steps_ms = float(np.mean(ts)) * 1e3
        workload = f"mixed {headline}+IR time-to-solution"
    else:
        value = res["f64"]["tts_tflops"]
        dtype = "f64"
        steps_ms = float(np.mean(tts64)) * 1e3
        workload = "FP64 time-to-solution"

    t_fac64 = float(np.median(fac64))
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(steps_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"{workload}: Cholesky + FP64 refinement to eta<=1e-12, n={n} b={b} nrhs={nrhs}, "
                               f"A=D(BB^T/n+1e-2 I)D",
                   "n": n, "tile": b, "nrhs": nrhs, "prio_kind": args.prio, "mixed_prio_kind": int(os.environ.get("TTAD_MIXED_PRIO", "4")), "schedule": "persistent"},
        "legs": res,
        "roofline": {"bound": "mfma", "achieved": round(F / t_fac64 / 1e12, 3), "peak": PEAK["f64"],
                     "unit": "TFLOP/s", "frac": round(F / t_fac64 / 1e12 / PEAK["f64"], 4),
                     "traffic": measured_traffic(n, b),
                     "kernel": "k_persistent FP64 factorization (one launch, F(n) flops / event-timed duration)"},
    }
    if not args.no_serve:
        del W64
        torch.cuda.empty_cache()
        log(f"serving leg: rate {args.serve_rate}/s, {args.serve_requests} requests")
        out["serving"] = serving_leg(ctx, args.serve_rate, args.serve_requests)
        log(f"serving: {out['serving']}")
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle n={args.cpu_n} b={b} threads={threads}")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, b, threads, args.prio)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
