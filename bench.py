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


def measured_traffic(prec: str, n: int, b: int):
    """HBM bytes per launch of the factorization kernel of precision `prec`, as measured offline with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and scaled by a 1 GiB copy (profiles/), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.txt")
    try:
        for line in open(path):
            key, _, val = line.partition("=")
            if key.strip() == f"cholesky_{prec}_n{n}_b{b}_bytes":
                return float(val)
    except OSError:
        pass
    return None


def host_threads():
    """Threads of the CPU baseline: the cores this process may run on (sched_getaffinity), capped by
    OMP_NUM_THREADS when the job sets it (the share of the host it was given). Returns (threads, how)."""
    n_aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and 0 < int(env) < n_aff:
        return int(env), f"OMP_NUM_THREADS={env} of {n_aff} usable cores (os.cpu_count()={os.cpu_count()})"
    return n_aff, f"{n_aff} usable cores (os.cpu_count()={os.cpu_count()})"


def cpu_oracle_run(A, b: int, threads: int, prio: int, runs: int):
    """The oracle (Taskflow+Threadpool restatement of 2d_cholesky.cpp, sequential MKL per task): `runs` timed
    factorizations after one warm-up (2d_cholesky.cpp:54-61), each timed seed -> join (:611-619)."""
    import numpy as np
    from oracle import oracle as orc
    orc.cholesky(A[:256, :256].copy(order="F"), 128, n_threads=threads, prio_kind=prio)  # warm-up
    times, L = [], None
    for _ in range(runs):
        L, st = orc.cholesky(A, b, n_threads=threads, prio_kind=prio)
        if st["info"] != 0:
            raise RuntimeError(f"cpu baseline: oracle info={st['info']}")
        times.append(st["total"])
    return float(np.median(times)), L


def cpu_baseline(A_c3, b: int, prio: int, c3: bool):
    """CPU baseline beside the GPU: the oracle on the host cores at n = 2048 (b = 256) and n = 16384 (b = 512) and (c3) on this run's own
    n = 32768 matrix. `value` is the largest configuration run."""
    import numpy as np
    from oracle import oracle as orc
    orc.build()
    kind = orc.set_backend("mkl") if orc.have_mkl() else orc.set_backend("textbook")
    threads, how = host_threads()
    log(f"cpu baseline: {threads} threads ({how}), backend {kind}")
    out = {"unit": "TFLOP/s", "cores": threads, "kind": "port", "threads_from": how, "backend": kind}
    # small case: n = 2048, b = 256, A = B B^T/n + I (seed 2020), median of 5
    A1 = orc.spd_bbt(2048, 2020)
    t1, L1 = cpu_oracle_run(A1, 256, threads, prio, 5)
    out["c1"] = {"n": 2048, "b": 256, "s": round(t1, 5), "tflops": round(flops(2048) / t1 / 1e12, 4),
                 "factor_residual": orc.factor_residual(A1, L1), "runs": 5}
    log(f"cpu c1: {out['c1']}")
    # mid-size case: n = 16384, b = 512, A = B B^T/n + I (torch seed 2021 on the GPU, copied to the host), once
    A2 = make_matrix(16384, 2021, 0, 1.0, False).cpu().numpy()
    t2, _ = cpu_oracle_run(np.asfortranarray(A2), 512, threads, prio, 1)
    del A2
    out["c2"] = {"n": 16384, "b": 512, "s": round(t2, 3), "tflops": round(flops(16384) / t2 / 1e12, 4), "runs": 1}
    log(f"cpu c2: {out['c2']}")
    out.update(value=out["c2"]["tflops"], sample=f"FP64 tiled Cholesky n=16384 b=512, one run, "
                                                  f"{t2:.2f} s seed->join")
    if c3 and A_c3 is not None:
        t3, L3 = cpu_oracle_run(A_c3, b, threads, prio, 1)
        out["c3"] = {"n": A_c3.shape[0], "b": b, "s": round(t3, 3),
                     "tflops": round(flops(A_c3.shape[0]) / t3 / 1e12, 4), "runs": 1}
        log(f"cpu c3: {out['c3']}")
        out.update(value=out["c3"]["tflops"], sample=f"FP64 tiled Cholesky of this run's n={A_c3.shape[0]} "
                                                      f"b={b} matrix, one run, {t3:.2f} s seed->join")
        return out, L3
    return out, None


def _submit_plan(ctx, srv, reqs, mats, rate, rng, prec, max_iter, timeout_s=300.0):
    """Drives one serving run: request q arrives at its Poisson time and is submitted then; results are
    collected as they complete. reqs[q] = (matrix index, nrhs, reuse_of) with reuse_of = index of an earlier
    request whose kept factor this one solves with (-1: factor). Returns per-request records."""
    import numpy as np
    n_req = len(reqs)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    rhs = [rng.standard_normal((mats[m][1].n, nrhs)) for (m, nrhs, _) in reqs]
    factor_of = {}
    recs = [None] * n_req
    waiting_reuse = {}
    reused_later = {r[2] for r in reqs if r[2] >= 0}
    t0 = time.perf_counter()
    submitted, done, i = {}, 0, 0

    def submit(q):
        m, nrhs, reuse_of = reqs[q]
        fid = factor_of.get(reuse_of, -1) if reuse_of >= 0 else -1
        rid, X = srv.submit(mats[m][1], rhs[q], prec=prec, max_iter=max_iter, keep_factor=q in reused_later,
                            reuse_factor=fid)
        submitted[rid] = (q, arrivals[q], X)

    while done < n_req:
        now = time.perf_counter() - t0
        while i < n_req and arrivals[i] <= now:
            m, nrhs, reuse_of = reqs[i]
            if reuse_of >= 0 and reuse_of not in factor_of:
                waiting_reuse.setdefault(reuse_of, []).append(i)  # its factor is not ready yet
            else:
                submit(i)
            i += 1
        for r in srv.poll():
            if r["id"] not in submitted:
                continue
            q, arr, X = submitted.pop(r["id"])
            t_now = time.perf_counter() - t0
            recs[q] = {"latency_s": t_now - arr, "queue_s": r["queue_s"], "eta": r["backward_error"],
                       "iters": r["iterations"], "status": r["status"], "reused": r["reused"],
                       "fell_back": r["fell_back"], "X": X}
            if reqs[q][2] < 0:
                factor_of[q] = r["factor_id"] if r["status"] == 0 else -1
                for w in waiting_reuse.pop(q, []):
                    submit(w)
            done += 1
        if time.perf_counter() - t0 > timeout_s:
            raise TimeoutError(f"serving run did not finish in {timeout_s} s")
    wall = time.perf_counter() - t0
    return recs, rhs, wall


def _host_eta(A, X, B):
    """Normwise backward error in FP64 with torch on the device (A symmetric, full)."""
    import numpy as np
    import torch
    Xt = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    Bt = torch.from_numpy(np.ascontiguousarray(B)).cuda()
    R = Bt - A @ Xt
    an = A.abs().sum(dim=1).max()
    eta = R.abs().amax(dim=0) / (an * Xt.abs().amax(dim=0) + Bt.abs().amax(dim=0))
    return float(eta.max())


def _request_matrices(ctx, sizes, seeds, kappa_inner=None, dexps=None, tile=256):
    """One FP64 tiled matrix per request. kappa_inner None: A = B B^T/n + I with B n x n iid N(0,1) from the
    request's seed. Else A = D (B B^T/n + delta I) D with B n x n/2 (so B B^T/n has n/2 zero eigenvalues and
    cond(B B^T/n + delta I) = (lambda_max + delta) / delta exactly), delta = lambda_max / (kappa_inner - 1)
    (lambda_max by power iteration) and D_ii = 10^(dexp U_i). Returns [(A dense on the device, TiledMatrix)]."""
    import torch
    import tasktorrent_amd as tt
    out = []
    for q, n in enumerate(sizes):
        g = torch.Generator(device="cuda")
        g.manual_seed(int(seeds[q]))
        m = n if kappa_inner is None else max(1, n // 2)
        B = torch.randn((n, m), dtype=torch.float64, device="cuda", generator=g)
        A = B @ B.T / n
        del B
        if kappa_inner is None:
            A.diagonal().add_(1.0)
        else:
            v = torch.randn((n, 1), dtype=torch.float64, device="cuda", generator=g)
            for _ in range(30):
                v = A @ v
                v /= torch.linalg.vector_norm(v)
            lmax = float((v.T @ (A @ v)).item())
            A.diagonal().add_(lmax / (float(kappa_inner[q]) - 1.0))
        if dexps is not None and dexps[q] > 0:
            d = torch.pow(10.0, float(dexps[q]) * torch.rand((n,), dtype=torch.float64, device="cuda", generator=g))
            A *= d[:, None]
            A *= d[None, :]
        M = tt.TiledMatrix(ctx, [tile] * (n // tile) + ([n % tile] if n % tile else []), "f64")
        M.set(A)
        out.append((A, M))
    torch.cuda.synchronize()
    return out


def serving_leg(ctx, rates, n_req: int, seed: int = 7, prec: str = "bf16x3", budget: int = 32 << 30):
    """Serving workload: Poisson arrivals of SPD solves, continuous batching in one resident kernel, at each rate.

    Every request has its own matrix (n = 64*round(exp(U(ln 64, ln 4096))/64), A = B B^T/n + I, B from the
    request's seed) except ~10% that solve new right-hand sides with the kept factor of an earlier request;
    1-8 RHS iid N(0,1). Matrices are staged in HBM before the run (the resident kernel owns the device while it
    runs). Latency = arrival -> solution copied back to host memory. The same request set runs at every rate."""
    import numpy as np
    import tasktorrent_amd as tt
    rng = np.random.default_rng(seed)
    sizes, reqs = [], []
    for q in range(n_req):
        nrhs = int(rng.integers(1, 9))
        if q > 0 and rng.random() < 0.1:
            prev = int(rng.integers(q))
            while reqs[prev][2] >= 0:  # reuse a request that factored
                prev = reqs[prev][2]
            reqs.append((reqs[prev][0], nrhs, prev))
        else:
            n = int(64 * round(np.exp(rng.uniform(np.log(64), np.log(4096))) / 64))
            sizes.append(max(64, min(4096, n)))
            reqs.append((len(sizes) - 1, nrhs, -1))
    mats = _request_matrices(ctx, sizes, [20_000 + m for m in range(len(sizes))])
    out = {"requests": n_req, "prec": prec, "distinct_matrices": len(sizes),
           "sizes": "n = 64*round(exp(U(ln64, ln4096))/64), A = BB^T/n + I per request, 1-8 RHS, "
                    "~10% solve with an earlier request's kept factor", "rates": []}
    for rate in rates:
        srv = tt.Server(ctx, hbm_budget=budget, max_tasks=1 << 21, max_requests=2048)
        try:
            recs, rhs, wall = _submit_plan(ctx, srv, reqs, mats, rate, np.random.default_rng(seed + int(rate)), prec,
                                           max_iter=30)
            st = srv.stats()
        finally:
            srv.close()
        lat = np.array([r["latency_s"] for r in recs]) * 1e3
        etas = [r["eta"] for r in recs]
        # independent FP64 check of every 10th request's solution (torch on the device)
        host = [_host_eta(mats[reqs[q][0]][0], recs[q]["X"], rhs[q]) for q in range(0, n_req, 10)]
        row = {"rate_per_s": rate, "p50_ms": round(float(np.percentile(lat, 50)), 3),
               "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(lat.mean()), 3),
               "solves_per_s": round(n_req / wall, 2), "wall_s": round(wall, 3),
               "ok": int(sum(r["status"] == 0 for r in recs)), "reused": int(sum(r["reused"] for r in recs)),
               "fell_back": int(sum(r["fell_back"] for r in recs)),
               "max_backward_error": float(np.max(etas)), "max_backward_error_fp64_check": float(np.max(host)),
               "mean_ir_steps": round(float(np.mean([r["iters"] for r in recs])), 2),
               "p99_queue_ms": round(float(np.percentile([r["queue_s"] for r in recs], 99)) * 1e3, 3),
               "workers": st["workers"]}
        log(f"serving rate {rate}/s: {row}")
        out["rates"].append(row)
    for A, M in mats:
        M.close()
    return out


def stress_leg(ctx, rate: float, n_req: int, seed: int = 11, prec: str = "bf16x3", budget: int = 200 << 30):
    """Stress serving workload: heavy-tailed sizes n = 256*round(min(16384, Pareto(1.5, x_m = 256))/256), kappa = 10^L with L
    uniform in [2, 10] through A = D (B B^T/n + delta I) D: delta chosen per request so that the inner matrix
    has condition number exactly kappa (B n x n/2), D_ii = 10^(U_i / 2) (the scaling of the formula, at most a
    factor 10 more on cond(A); Cholesky and the refinement are nearly invariant to it). 1-8 RHS, admission by
    HBM footprint under a 200 GB arena, refinement capped at 10 steps and then the automatic FP64
    refactorization. Reports the fallback rate, the backward errors (the server's and an FP64 recomputation of
    every request's), the measured condition numbers of the small requests and the latency percentiles."""
    import numpy as np
    import tasktorrent_amd as tt
    rng = np.random.default_rng(seed)
    sizes, logk, reqs = [], [], []
    for q in range(n_req):
        x = 256.0 * (1.0 - rng.random()) ** (-1.0 / 1.5)
        n = int(256 * round(min(16384.0, x) / 256))
        sizes.append(max(256, n))
        logk.append(float(rng.uniform(2.0, 10.0)))
        reqs.append((q, int(rng.integers(1, 9)), -1))
    kin = [10.0 ** L for L in logk]
    dexps = [0.5 for _ in logk]
    mats = _request_matrices(ctx, sizes, [30_000 + q for q in range(n_req)], kin, dexps)
    # the condition numbers the construction reaches, measured (2-norm) on the requests with n <= 1024
    import torch
    meas = {q: float(torch.linalg.cond(mats[q][0]).item()) for q in range(n_req) if sizes[q] <= 1024}
    srv = tt.Server(ctx, hbm_budget=budget, max_tasks=1 << 22, max_requests=1024)
    try:
        recs, rhs, wall = _submit_plan(ctx, srv, reqs, mats, rate, np.random.default_rng(seed + 1), prec,
                                       max_iter=10, timeout_s=600.0)
    finally:
        srv.close()
    lat = np.array([r["latency_s"] for r in recs]) * 1e3
    host = [_host_eta(mats[q][0], recs[q]["X"], rhs[q]) for q in range(n_req)]
    fb = np.array([r["fell_back"] for r in recs], dtype=bool)
    ok = np.array([r["status"] == 0 for r in recs])
    out = {"rate_per_s": rate, "requests": n_req, "prec": prec, "hbm_budget_gb": budget / 2 ** 30,
           "max_ir_steps_before_fallback": 10,
           "n_range": [int(min(sizes)), int(max(sizes))], "n_mean": round(float(np.mean(sizes)), 1),
           "kappa_log10_range": [round(min(logk), 2), round(max(logk), 2)],
           "fallback_rate": round(float(fb.mean()), 4), "ok": int(ok.sum()),
           "fallback_kappa_log10_min": round(float(min([logk[q] for q in range(n_req) if fb[q]], default=0)), 2),
           "no_fallback_kappa_log10_max": round(float(max([logk[q] for q in range(n_req) if not fb[q]], default=0)), 2),
           "measured_kappa_log10_n_le_1024": [round(float(np.log10(min(meas.values()))), 2),
                                              round(float(np.log10(max(meas.values()))), 2)] if meas else None,
           "fallback_rate_measured_kappa_gt_1e8": round(float(np.mean([fb[q] for q in meas if meas[q] > 1e8])), 3)
           if any(v > 1e8 for v in meas.values()) else None,
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "solves_per_s": round(n_req / wall, 2), "wall_s": round(wall, 3),
           "max_backward_error": float(max(r["eta"] for r in recs)),
           "max_backward_error_fp64_check": float(np.max(host)),
           "ir_steps_mean": round(float(np.mean([r["iters"] for r in recs])), 2),
           "ir_steps_max": int(max(r["iters"] for r in recs)),
           "fallback_rate_by_kappa_decade": {f"1e{d}": round(float(np.mean([fb[q] for q in range(n_req)
                                                                          if int(logk[q]) == d])), 3)
                                             for d in range(2, 10) if any(int(L) == d for L in logk)},
           "ir_steps_mean_by_kappa_decade": {f"1e{d}": round(float(np.mean([recs[q]["iters"] for q in range(n_req)
                                                                            if int(logk[q]) == d])), 2)
                                             for d in range(2, 10) if any(int(L) == d for L in logk)},
           "sizes": "n = 256*round(min(16384, Pareto(alpha=1.5, x_m=256))/256), cond(B B^T/n + delta I) "
                    "log-uniform [1e2, 1e10] through delta, D_ii = 10^(U/2)"}
    for A, M in mats:
        M.close()
    return out


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
    """The panel steps alone, FP64, as the factorization runs them.

    POTRF: one b x b tile factored by the row-block items the factorization uses (b/64 workgroups that wait on
    each other's progress flags), launched alone through the host-sequenced schedule of a one-tile matrix and
    timed with events around that launch. TRSM: ttad_dtrsm on the m x b column of tiles below (the first
    panel's TRSM tasks at m = n - b). GB/s use the per-task algorithmic bytes: POTRF reads and writes its tile
    (2 b^2 8 bytes), a TRSM task reads L_jj and reads and writes its tile (3 b^2 8 bytes)."""
    import numpy as np
    import torch
    import tasktorrent_amd as tt
    m = max(m, b)
    g = torch.Generator(device="cuda")
    g.manual_seed(13)
    Bm = torch.randn((b, b), dtype=torch.float64, device="cuda", generator=g)
    A = Bm @ Bm.T / b + torch.eye(b, dtype=torch.float64, device="cuda")
    T = tt.TiledMatrix(ctx, [b], "f64")
    t_potrf = []
    for r in range(reps + 2):
        T.set(A)
        info = ctx.potrf_tiled(T, tt.TTAD_SCHED_STREAM, 2, "f64")
        if info != 0:
            raise RuntimeError(f"panel leg: potrf info={info}")
        if r >= 2:
            t_potrf.append(ctx.last_elapsed())
    Ls = torch.from_numpy(np.ascontiguousarray(T.get().T)).cuda()  # column-major L (b x b)
    T.close()
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
    n_trsm = m / b
    potrf_gbps = 2.0 * b * b * 8 / tp / 1e9
    trsm_gbps = 3.0 * b * b * 8 * n_trsm / t_trsm / 1e9
    del A, Ls, P, Bm
    torch.cuda.empty_cache()
    return {"potrf_tile": b, "potrf_items": b // 64, "potrf_ms": round(tp * 1e3, 4),
            "potrf_tflops": round(flops(b) / tp / 1e12, 4), "potrf_gbps": round(potrf_gbps, 2),
            "potrf_hbm_frac": round(potrf_gbps / 8000.0, 5),
            "trsm_shape": f"m={m} n={b} ({n_trsm:g} tile tasks)", "trsm_ms": round(t_trsm * 1e3, 4),
            "trsm_gbps": round(trsm_gbps, 1), "trsm_hbm_frac": round(trsm_gbps / 8000.0, 4),
            "trsm_tflops": round(float(m) * b * b / t_trsm / 1e12, 3)}


def update_flops(n: int, b: int) -> float:
    """Flops of the trailing updates (GEMM 2b^3, SYRK b^2(b+1) per task) of the tiled factorization."""
    nt = n // b
    gemm = sum((nt - k - 1) * (nt - k - 2) // 2 for k in range(nt))
    syrk = nt * (nt - 1) // 2
    return gemm * 2.0 * b ** 3 + syrk * float(b * b * (b + 1))


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
    ap.add_argument("--no-cpu-c3", action="store_true", help="CPU baseline at n = 2048 and n = 16384 only")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--serve-rates", default="200,800,1600,3200,6400", help="Poisson request rates of the serving leg (1/s)")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--stress-rate", type=float, default=50.0, help="request rate of the stress serving leg (1/s)")
    ap.add_argument("--stress-requests", type=int, default=400)
    ap.add_argument("--no-stress", action="store_true")
    ap.add_argument("--no-update", action="store_true", help="skip the FP64 trailing-update and panel-step legs")
    ap.add_argument("--no-residual", action="store_true", help="skip the ||LL^T - A||/||A|| reports")
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
    A_host = None
    if not args.no_cpu and not args.no_cpu_c3:
        A_host = np.asfortranarray(A.cpu().numpy())  # symmetric: row-major bytes = column-major
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
                  "factor_frac_of_peak": round(F / float(np.median(fac64)) / 1e12 / PEAK["f64"], 4),
                  "tts_tflops": round(F / float(np.median(tts64)) / 1e12, 3),
                  "backward_error": st["backward_error"], "ir_steps": st["iterations"]}
    if not args.no_residual:  # ||L L^T - A|| / ||A|| of the factor the last step used (untimed)
        res["f64"]["factor_residual"] = ctx.factor_residual(A64, W64)
        log(f"fp64 factor residual {res['f64']['factor_residual']:.3e}")
    L_gpu = None
    if A_host is not None:
        L_gpu = W64.get()

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
    timed_steps, fac_steps = {}, {}
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
            log(f"{prec} step {s}: tts {t*1e3:.1f} ms factor {fac*1e3:.1f} ms eta {st['backward_error']:.2e} "
                f"it {st['iterations']} fallback {st['fell_back']} rc {st['rc']}")
        timed_steps[prec] = ts
        fac_steps[prec] = facs
        tmed = float(np.median(ts))
        res[prec] = {"tts_ms": round(1e3 * tmed, 3), "tts_tflops": round(F / tmed / 1e12, 3),
                     "factor_ms": round(1e3 * float(np.median(facs)), 3),
                     "backward_error": st["backward_error"], "ir_steps": st["iterations"],
                     "fell_back": bool(st["fell_back"]), "rc": st["rc"],
                     "speedup_vs_f64": round(res["f64"]["tts_ms"] / (1e3 * tmed), 3)}
        if not args.no_residual:  # one more (untimed) solve that also reports the factor residual
            ctx.report_factor_residual(True)
            st2 = ctx.posv_mixed(A64, prec, nrhs, Bt, Xt, tol=1e-12, max_iter=30)
            ctx.report_factor_residual(False)
            res[prec]["factor_residual"] = st2["factor_residual"]
            log(f"{prec} factor residual {st2['factor_residual']:.3e}")

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
    roof64 = {"bound": "mfma", "achieved": round(F / t_fac64 / 1e12, 3), "peak": PEAK["f64"], "unit": "TFLOP/s",
              "frac": round(F / t_fac64 / 1e12 / PEAK["f64"], 4), "traffic": measured_traffic("f64", n, b),
              "kernel": "k_persistent FP64 factorization (one launch, F(n) flops / event-timed duration)"}
    if dtype in ("bf16x3", "bf16", "f32"):
        # the headline step's dominant kernel: the mixed factorization. Its MFMA work is the trailing update in
        # the leg's arithmetic (bf16x3: three BF16 products per update flop) against that type's dense peak.
        t_fm = float(np.median(fac_steps[headline]))
        mult = {"bf16x3": 3.0, "bf16": 1.0, "f32": 1.0}[dtype]
        pk = PEAK["f32"] if dtype == "f32" else PEAK["bf16"]
        ach = mult * update_flops(n, b) / t_fm / 1e12
        roofline = {"bound": "mfma", "achieved": round(ach, 3), "peak": pk, "unit": "TFLOP/s",
                    "frac": round(ach / pk, 4), "traffic": measured_traffic(dtype, n, b),
                    "kernel": f"k_persistent {dtype} factorization ({mult:g} x trailing-update flops "
                              f"{update_flops(n, b):.4e} in {'FP32' if dtype == 'f32' else 'BF16'} MFMA / "
                              f"event-timed duration {t_fm*1e3:.2f} ms)"}
    else:
        roofline = roof64
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
        "roofline": roofline,
        "roofline_f64": roof64,
    }
    del W64
    torch.cuda.empty_cache()
    if not args.no_serve:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        log(f"serving leg: rates {rates}/s, {args.serve_requests} requests each")
        out["serving"] = serving_leg(ctx, rates, args.serve_requests)
    if not args.no_stress:
        log(f"stress leg: rate {args.stress_rate}/s, {args.stress_requests} requests")
        out["stress"] = stress_leg(ctx, args.stress_rate, args.stress_requests)
        log(f"stress: {out['stress']}")
    if not args.no_cpu:
        cb, L_cpu = cpu_baseline(A_host, b, args.prio, not args.no_cpu_c3)
        if L_cpu is not None and L_gpu is not None:
            d = float(np.linalg.norm(L_gpu - L_cpu) / np.linalg.norm(L_cpu))
            cb["c3"]["gpu_fp64_L_rel_diff"] = d
            log(f"FP64 L, GPU vs CPU oracle at n={n}: {d:.3e}")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
