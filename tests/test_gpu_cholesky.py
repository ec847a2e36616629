"""Tiled Cholesky parity on the GPU against the oracle's Taskflow+BLAS restatement
of miniapp/dense_cholesky/2d_cholesky.cpp (one rank, non-accumulate mode).

Bar: FP64 L within 1e-12 relative (Frobenius) of the oracle's L; the persistent
DAG schedule must reproduce the host-sequenced schedule bit for bit (same
per-tile update order k = 0, 1, ..., fixed by the GEMM k-chain of
2d_cholesky.cpp:514-518) and be independent of the priority kind.
"""
import numpy as np
import pytest

import tasktorrent_amd as tt

pytestmark = pytest.mark.gpu


def factor(ctx, A, sizes, schedule, prio=2, prec="f64", storage=None):
    M = tt.TiledMatrix(ctx, sizes, storage or ("f64" if prec == "f64" else "f32"))
    M.set(A)
    info = ctx.potrf_tiled(M, schedule, prio, prec)
    L = M.get()
    M.close()
    return L, info


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("schedule", [tt.TTAD_SCHED_STREAM, tt.TTAD_SCHED_PERSISTENT])
def test_c1_bbt_matches_oracle(ctx, oracle, schedule):
    # configs[0] shape: n=2048, tile 256, A = B B^T/n + I, seed 2020
    n, b = 2048, 256
    A = oracle.spd_bbt(n, 2020)
    Lref, st = oracle.cholesky(A, b, n_threads=4, prio_kind=2)
    L, info = factor(ctx, A, [b] * (n // b), schedule)
    assert info == 0 and st["info"] == 0
    assert rel(L, Lref) < 1e-12
    assert oracle.factor_residual(A, L) < 1e-14


def test_reference_generator(ctx, oracle):
    # the miniapp's own matrix A(i,j) = 1/((i-j)^2+1) (2d_cholesky.cpp:133), ragged tiles
    n, b = 1000, 128
    A = oracle.ref_val_matrix(n)
    sizes = [b] * (n // b) + ([n % b] if n % b else [])
    Lref, _ = oracle.cholesky(A, sizes=sizes, n_threads=2)
    L, info = factor(ctx, A, sizes, tt.TTAD_SCHED_PERSISTENT)
    assert info == 0
    assert rel(L, Lref) < 1e-12
    assert oracle.solve_check(A, L) < 1e-10  # the miniapp's check is 1e-6 (2d_cholesky.cpp:684)


def test_persistent_equals_stream_bitwise(ctx, oracle):
    n, b = 1536, 256
    A = oracle.spd_bbt(n, 7)
    L1, _ = factor(ctx, A, [b] * 6, tt.TTAD_SCHED_STREAM)
    L2, _ = factor(ctx, A, [b] * 6, tt.TTAD_SCHED_PERSISTENT)
    assert np.array_equal(L1, L2)


@pytest.mark.parametrize("prio", [0, 1, 2, 3])
def test_priority_kinds_same_result(ctx, oracle, prio):
    n, b = 1024, 128
    A = oracle.spd_bbt(n, 99)
    L0, _ = factor(ctx, A, [b] * 8, tt.TTAD_SCHED_STREAM)
    L, info = factor(ctx, A, [b] * 8, tt.TTAD_SCHED_PERSISTENT, prio=prio)
    assert info == 0
    assert np.array_equal(L, L0)


def test_nonuniform_blocks(ctx, oracle):
    # upper_block_size > block_size partitions (2d_cholesky.cpp:63-104)
    sizes = oracle.block_sizes(192, 8, 288)
    n = int(sizes.sum())
    A = oracle.spd_bbt(n, 5)
    Lref, _ = oracle.cholesky(A, sizes=sizes, n_threads=4)
    L, info = factor(ctx, A, sizes, tt.TTAD_SCHED_PERSISTENT)
    assert info == 0
    assert rel(L, Lref) < 1e-12


def test_single_tile_and_tiny(ctx, oracle):
    for n, sizes in [(1, [1]), (64, [64]), (70, [64, 6]), (300, [100, 100, 100])]:
        A = oracle.spd_bbt(n, n)
        Lref = np.linalg.cholesky(A)
        L, info = factor(ctx, A, sizes, tt.TTAD_SCHED_PERSISTENT)
        assert info == 0
        assert rel(L, Lref) < 1e-12


def test_not_spd_reports_info(ctx, oracle):
    n, b = 512, 128
    A = oracle.spd_bbt(n, 1)
    A[300, 300] = -5.0
    _, st = oracle.cholesky(A, b, n_threads=2)
    for sched in (tt.TTAD_SCHED_STREAM, tt.TTAD_SCHED_PERSISTENT):
        _, info = factor(ctx, A, [b] * 4, sched)
        assert info == st["info"] > 0


def test_integer_factor(ctx, oracle):
    # unique factor of A = L L^T for a well-conditioned integer L (the oracle matches such
    # factors bitwise; the GPU's inverse-block TRSM is accurate to the blocks' condition)
    n = 384
    Li = oracle.int_lower_factor(n, 4, offmax=1, diag_exp=(3, 4))
    A = np.asfortranarray(Li @ Li.T)
    L, info = factor(ctx, A, [128] * 3, tt.TTAD_SCHED_PERSISTENT)
    assert info == 0
    assert rel(L, Li) < 1e-13


@pytest.mark.parametrize("prec", ["f32", "bf16", "bf16x3"])
def test_mixed_factor_close(ctx, oracle, prec):
    n, b = 1024, 256
    A = oracle.spd_bbt(n, 31)
    Lref = np.linalg.cholesky(A)
    L, info = factor(ctx, A, [b] * 4, tt.TTAD_SCHED_PERSISTENT, prec=prec)
    assert info == 0
    tol = {"f32": 1e-5, "bf16x3": 1e-4, "bf16": 3e-2}[prec]
    assert rel(L, Lref) < tol
    L2, _ = factor(ctx, A, [b] * 4, tt.TTAD_SCHED_STREAM, prec=prec)
    assert np.array_equal(L, L2)


@pytest.mark.parametrize("sizes", [[512] * 4, [512, 320, 448, 96, 384, 100, 256]])
def test_bf16x3_split_operands_bitwise(ctx, oracle, sizes):
    """The persistent bf16x3 factorization reads its update operands split once into bf16 hi/lo
    planes by the TRSM items (tiles whose column count is a multiple of 32, items wholly inside both
    operands); the host-sequenced schedule splits them on the fly. Same factor bit for bit, also for
    ragged tiles that keep the on-the-fly path and for a second run on the same matrix (planes reused)."""
    n = sum(sizes)
    A = oracle.spd_bbt(n, 41)
    M = tt.TiledMatrix(ctx, sizes, "f32")
    try:
        runs = []
        for _ in range(2):
            M.set(A)
            assert ctx.potrf_tiled(M, tt.TTAD_SCHED_PERSISTENT, 2, "bf16x3") == 0
            runs.append(M.get())
    finally:
        M.close()
    L2, info = factor(ctx, A, sizes, tt.TTAD_SCHED_STREAM, prec="bf16x3")
    assert info == 0
    assert np.array_equal(runs[0], L2)
    assert np.array_equal(runs[1], L2)
    assert rel(runs[0], np.linalg.cholesky(A)) < 1e-4


def _longest_path_tasks(indeg_edges, n):
    out_ptr, out_idx = indeg_edges
    pending = np.bincount(out_idx, minlength=n)
    order = [t for t in range(n) if pending[t] == 0]
    for t in order:
        for s in out_idx[out_ptr[t]:out_ptr[t + 1]]:
            pending[s] -= 1
            if pending[s] == 0:
                order.append(s)
    bl = np.zeros(n)
    for t in reversed(order):
        succ = out_idx[out_ptr[t]:out_ptr[t + 1]]
        bl[t] = 1 + (bl[succ].max() if len(succ) else 0)
    return int(bl.max())


@pytest.mark.parametrize("nb,workers", [(8, 16), (12, 0)])
def test_spin_bench_scheduler_overhead(ctx, nb, workers):
    # mini_benchmarks/shared/ttor_cholesky.cpp: the Cholesky DAG with spin bodies. The run can be
    # no shorter than the longest chain of spins nor than the total spin spread over the workers,
    # and the reported efficiency follows the reference's formula.
    spin_us = 200
    tasks, indeg, prio, out_ptr, out_idx = tt.cholesky_dag(nb, 2)
    crit = _longest_path_tasks((out_ptr, out_idx), len(tasks))
    r = tt.cholesky_spin_bench(ctx, nb, spin_us, n_workers=workers, reps=2)
    assert r["tasks"] == len(tasks)
    lower = max(crit * spin_us * 1e-6, len(tasks) * spin_us * 1e-6 / r["workers"])
    assert r["elapsed_s"] >= 0.95 * lower
    assert r["elapsed_s"] <= 2.0 * lower + 5e-3
    assert 0.0 < r["efficiency"] <= 1.0
    assert r["efficiency"] == pytest.approx(len(tasks) * spin_us * 1e-6 / (r["elapsed_s"] * r["workers"]))


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_accumulate_mode_matches_oracle(ctx, oracle, prec):
    # accumulate mode adds the updates of a tile in the order they finish: not bitwise
    # reproducible, but the same factor to rounding
    n, b = 1536, 256
    A = oracle.spd_bbt(n, 17)
    Lref = np.linalg.cholesky(A)
    L, info = factor(ctx, A, [b] * 6, tt.TTAD_SCHED_PERSISTENT, prio=2 | tt.TTAD_ACCUMULATE, prec=prec)
    assert info == 0
    tol = {"f64": 1e-12, "bf16x3": 1e-4}[prec]
    assert rel(L, Lref) < tol
    L2, _ = factor(ctx, A, [b] * 6, tt.TTAD_SCHED_PERSISTENT, prec=prec)
    assert rel(L, L2) < tol


def test_graph_write_deps_matches_oracle(ctx, oracle, tmp_path):
    # the device graph's edges in the reference DepsLogger format (util.cpp:103-112), against the
    # oracle's Taskflow run of the same DAG
    nb = 5
    tasks, indeg, prio, out_ptr, out_idx = tt.cholesky_dag(nb, 2)
    G = tt.Graph(ctx, tasks, indeg, out_ptr, out_idx, prio)
    ours = tmp_path / "ours.csv"
    G.write_deps(str(ours))
    G.close()
    A = oracle.spd_bbt(nb * 8, 1)
    ref = tmp_path / "ref.csv"
    oracle.cholesky(A, 8, n_threads=2, prio_kind=2, deps_path=str(ref))
    lines = ours.read_text().splitlines()
    assert len(lines) == len(set(lines)) == len(out_idx)
    assert set(lines) == {line for line in ref.read_text().splitlines() if line}


def host_threads():
    """Worker threads for the oracle: the host cores this process may use, capped by OMP_NUM_THREADS when set
    (the share of the machine the job was given)."""
    import os
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


def test_fp64_n16384_b512_matches_oracle(ctx, oracle):
    # mid-size case: FP64 tiled Cholesky through the device task DAG at n = 16384, tile 512, A = B B^T/n + I from
    # seed 2021; L against the oracle's Taskflow + sequential-BLAS factor (threads = host cores) within 1e-12
    # relative, and the device's own ||L L^T - A||_F / ||A||_F
    n, b = 16384, 512
    A = oracle.spd_bbt(n, 2021)
    Lref, st = oracle.cholesky(A, b, n_threads=host_threads(), prio_kind=2)
    assert st["info"] == 0
    M = tt.TiledMatrix(ctx, [b] * (n // b), "f64")
    M.set(A)
    W = tt.TiledMatrix(ctx, [b] * (n // b), "f64")
    W.copy_from(M)
    info = ctx.potrf_tiled(W, tt.TTAD_SCHED_PERSISTENT, 2, "f64")
    assert info == 0
    res = ctx.factor_residual(M, W)
    L = W.get()
    W.close()
    M.close()
    err = rel(L, Lref)
    print(f"mid-size case: n={n} b={b} rel(L, L_oracle)={err:.3e} ||LL^T-A||/||A||={res:.3e} "
          f"oracle {st['total']:.2f} s on {host_threads()} threads", flush=True)
    assert err < 1e-12
    assert res < 1e-14


@pytest.mark.parametrize("prec", ["f32", "bf16x3", "bf16"])
def test_factor_residual_matches_host(ctx, oracle, prec):
    # a low-precision factor's residual (~1e-7 .. 1e-3) is far above the FP64 rounding of either computation,
    # so the device and the host must agree closely; ragged last tile
    sizes = [256, 256, 256, 192]
    n = sum(sizes)
    A = oracle.spd_bbt(n, 41)
    M = tt.TiledMatrix(ctx, sizes, "f64")
    M.set(A)
    W = tt.TiledMatrix(ctx, sizes, "f32")
    W.copy_from(M)
    assert ctx.potrf_tiled(W, tt.TTAD_SCHED_PERSISTENT, 2, prec) == 0
    dev = ctx.factor_residual(M, W)
    host = oracle.factor_residual(A, W.get())
    assert host > 1e-9
    assert abs(dev - host) <= 1e-6 * host, (dev, host)
    W.close()
    M.close()


def test_factor_residual_planted_perturbation(ctx, oracle):
    # L' = L + eps e_p e_q^T (p >= q) gives L'L'^T - A = eps (c e_p^T + e_p c^T + eps e_p e_p^T) + (LL^T - A),
    # c = L e_q, whose Frobenius norm has the closed form eps sqrt(2 sum_{a != p} c_a^2 + (2 c_p + eps)^2)
    n, b = 640, 128
    A = oracle.spd_bbt(n, 43)
    M = tt.TiledMatrix(ctx, [b] * 5, "f64")
    M.set(A)
    W = tt.TiledMatrix(ctx, [b] * 5, "f64")
    W.copy_from(M)
    assert ctx.potrf_tiled(W, tt.TTAD_SCHED_PERSISTENT, 2, "f64") == 0
    base = ctx.factor_residual(M, W)
    assert base < 1e-14
    L = W.get()
    eps, p, q = 1e-5, 517, 200
    c = L[:, q].copy()
    Lp = L.copy()
    Lp[p, q] += eps
    W.set(Lp)
    got = ctx.factor_residual(M, W)
    norm_e = eps * np.sqrt(2.0 * (np.sum(c ** 2) - c[p] ** 2) + (2.0 * c[p] + eps) ** 2)
    want = norm_e / np.linalg.norm(A)
    assert abs(got - want) <= 1e-6 * want + 2 * base, (got, want)
    W.close()
    M.close()


def test_graph_trace_readable_by_reference_tools(ctx, oracle, tmp_path):
    # a traced device run, read the way tools/ttor_logging.py reads Logger files: one '>run>' row per task whose
    # name joins the DepsLogger edges, a '>deps>' row per task, times in seconds since the epoch, and every edge
    # respected (a successor starts after its predecessor's last item ended)
    import time

    import trace_util
    nb, b = 6, 128
    tasks, indeg, prio, out_ptr, out_idx = tt.cholesky_dag(nb, 2)
    A = oracle.spd_bbt(nb * b, 12)
    M = tt.TiledMatrix(ctx, [b] * nb, "f64")
    M.set(A)
    G = tt.Graph(ctx, tasks, indeg, out_ptr, out_idx, prio)
    G.bind(0, M)
    G.seed([tt.task_name(t) for t in tasks].index("POTRF_0"))
    t_before = time.time()
    G.launch(trace=True)
    G.wait(60.0)
    t_after = time.time()
    log, deps = tmp_path / "trace.csv", tmp_path / "deps.csv"
    G.write_trace(str(log), "Wk_")
    G.write_deps(str(deps))
    G.close()
    Lref, _ = oracle.cholesky(A, b, n_threads=2)
    assert rel(M.get(), Lref) < 1e-12
    M.close()
    nodes = trace_util.dag_nodes(log).set_index("task")
    edges = trace_util.dag_edges(deps)
    assert nodes.index.is_unique and len(nodes) == len(tasks)
    assert set(nodes.index) == set(edges["start"]) | set(edges["end"])
    assert (nodes["time_end"] >= nodes["time_start"]).all()
    assert nodes["time_start"].min() >= t_before - 1.0 and nodes["time_end"].max() <= t_after + 1.0
    for u, v in zip(edges["start"], edges["end"]):
        assert nodes.loc[v, "time_start"] >= nodes.loc[u, "time_end"], (u, v)
    rows = trace_util.trace_rows(log)
    assert (rows["what"] == "deps").sum() == len(tasks)
    assert all(w.startswith("Wk_") for w in rows["who"])
