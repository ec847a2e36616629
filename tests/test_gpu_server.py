"""Serving loop on the GPU (ttad_server): continuous batching of SPD solves.

- every request reaches a normwise backward error <= 1e-12 (FP64 residual on the host);
- each result is the same alone and in the mix: requests solved one at a time on an
  otherwise idle server and solved all together give bit-identical solutions (each
  request's arithmetic has a fixed order and no floating-point atomics);
- a cached factor solves new right-hand sides without refactoring;
- under an HBM budget that fits only a few requests at once, the rest wait in the
  host queue and are admitted as memory frees up.
"""
import numpy as np
import pytest

import tasktorrent_amd as tt

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    B = np.random.default_rng(seed).standard_normal((n, n))
    A = B @ B.T / n
    A[np.diag_indices(n)] += 1.0
    return np.asfortranarray(A)


def _mat(ctx, A, b=256):
    n = A.shape[0]
    sizes = [b] * (n // b) + ([n % b] if n % b else [])
    M = tt.TiledMatrix(ctx, sizes, "f64")
    M.set(A)
    return M


def eta(A, X, B):
    R = B - A @ X
    an = np.abs(A).sum(axis=1).max()
    return float((np.abs(R).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))).max())


SIZES = [64, 192, 320, 512, 704, 1024, 1600, 2048]


@pytest.fixture(scope="module")
def problems(ctx):
    out = []
    for q, n in enumerate(SIZES):
        A = _spd(n, 100 + q)
        nrhs = 1 + q % 8
        B = np.asfortranarray(np.random.default_rng(200 + q).standard_normal((n, nrhs)))
        out.append((A, _mat(ctx, A), B))
    return out


@pytest.mark.parametrize("prec", ["bf16x3", "f32"])
def test_alone_equals_mix(ctx, problems, prec):
    srv = tt.Server(ctx, hbm_budget=2 << 30, max_tasks=1 << 18, max_requests=64)
    try:
        alone = []
        for A, M, B in problems:
            rid, X = srv.submit(M, B, prec=prec)
            r = srv.wait_all([rid])[rid]
            assert r["status"] == 0, r
            alone.append(r["X"].copy())
        ids = [srv.submit(M, B, prec=prec)[0] for A, M, B in problems]
        res = srv.wait_all(ids)
        for k, rid in enumerate(ids):
            r = res[rid]
            A, M, B = problems[k]
            assert r["status"] == 0, r
            assert r["backward_error"] <= 1e-12, r
            assert eta(A, r["X"], B) <= 2e-12
            assert np.array_equal(r["X"], alone[k]), f"request {k} (n={A.shape[0]}) differs alone vs mix"
    finally:
        srv.close()


def test_factor_reuse(ctx, problems):
    srv = tt.Server(ctx, hbm_budget=1 << 30, max_tasks=1 << 16, max_requests=16)
    try:
        A, M, B = problems[5]
        rid, _ = srv.submit(M, B, prec="bf16x3", keep_factor=True)
        r = srv.wait_all([rid])[rid]
        assert r["status"] == 0 and r["reused"] == 0
        fid = r["factor_id"]
        assert srv.stats()["cached_factors"] == 1
        for seed in range(3):
            B2 = np.asfortranarray(np.random.default_rng(seed).standard_normal((A.shape[0], 3)))
            rid2, _ = srv.submit(M, B2, prec="bf16x3", reuse_factor=fid)
            r2 = srv.wait_all([rid2])[rid2]
            assert r2["status"] == 0 and r2["reused"] == 1
            assert eta(A, r2["X"], B2) <= 2e-12
    finally:
        srv.close()


def test_admission_under_budget(ctx, problems):
    # ~2 requests of n=2048 (FP32 factor 8.4 MB each) fit at once; 6 are submitted together
    srv = tt.Server(ctx, hbm_budget=24 << 20, max_tasks=1 << 16, max_requests=16)
    try:
        A, M, B = problems[-1]
        ids = [srv.submit(M, B, prec="f32")[0] for _ in range(6)]
        st = srv.stats()
        assert st["pending"] > 0  # some had to wait
        res = srv.wait_all(ids)
        Xs = [res[i]["X"] for i in ids]
        for r in res.values():
            assert r["status"] == 0 and r["backward_error"] <= 1e-12
        assert all(np.array_equal(Xs[0], x) for x in Xs[1:])
        assert max(r["queue_s"] for r in res.values()) > 0
    finally:
        srv.close()


def test_fp64_requests(ctx, problems):
    srv = tt.Server(ctx, hbm_budget=1 << 30, max_tasks=1 << 16, max_requests=16)
    try:
        A, M, B = problems[3]
        rid, _ = srv.submit(M, B, prec="f64")
        r = srv.wait_all([rid])[rid]
        assert r["status"] == 0 and eta(A, r["X"], B) <= 1e-15
    finally:
        srv.close()


def test_fallback_forced_stall(ctx, problems):
    # BF16 needs several refinement steps; with max_iter = 1 the low-precision attempt stops short,
    # and the request is re-run with an FP64 factor (one step suffices there) before it is reported
    srv = tt.Server(ctx, hbm_budget=1 << 30, max_tasks=1 << 16, max_requests=16)
    try:
        A, M, B = problems[4]
        rid, _ = srv.submit(M, B, prec="bf16", max_iter=1)
        r = srv.wait_all([rid])[rid]
        assert r["status"] == 0 and r["fell_back"] == 1, r
        assert r["iterations"] == 2, r
        assert eta(A, r["X"], B) <= 1e-12
        # the same request in FP64 from the start gives the same solution bit for bit
        rid2, _ = srv.submit(M, B, prec="f64", max_iter=1)
        r2 = srv.wait_all([rid2])[rid2]
        assert r2["status"] == 0 and r2["fell_back"] == 0
        assert np.array_equal(r["X"], r2["X"])
    finally:
        srv.close()


def test_fallback_ill_conditioned_and_not_spd(ctx, oracle):
    # kappa ~ 1e10 (A = D (B B^T/n + delta I) D): the bf16x3 attempt cannot refine to 1e-12 in 10 steps,
    # the FP64 one does. An indefinite matrix fails both attempts: NOT_SPD, reported once, fell_back = 1.
    n = 768
    Aill = oracle.spd_scaled(n, 31, 32, delta=1e-6, dmax=1e2)
    Abad = _spd(n, 33)
    Abad[np.diag_indices(n)] -= 3.0
    Mill, Mbad = _mat(ctx, Aill), _mat(ctx, Abad)
    B = np.asfortranarray(np.random.default_rng(34).standard_normal((n, 3)))
    srv = tt.Server(ctx, hbm_budget=1 << 30, max_tasks=1 << 16, max_requests=16)
    try:
        ids = [srv.submit(Mill, B, prec="bf16x3", max_iter=10)[0], srv.submit(Mbad, B, prec="bf16x3", max_iter=10)[0]]
        res = srv.wait_all(ids)
        ri, rb = res[ids[0]], res[ids[1]]
        assert ri["status"] == 0 and ri["fell_back"] == 1, ri
        assert eta(Aill, ri["X"], B) <= 2e-12
        assert rb["status"] == tt.TTAD_ERR_NOT_SPD and rb["fell_back"] == 1, rb
        L = np.asfortranarray(Abad.copy())
        assert rb["info"] == oracle.dpotrf(L)
    finally:
        srv.close()
        Mill.close()
        Mbad.close()


def test_context_close_with_live_objects():
    # closing a context first closes the server and matrices still open on it, so
    # that dropping them afterwards never touches the freed context
    c2 = tt.Context(0)
    M = _mat(c2, _spd(256, 5))
    srv = tt.Server(c2, hbm_budget=64 << 20, max_tasks=1 << 12, max_requests=4)
    c2.close()
    assert not M.handle and not srv._h
    del M, srv
