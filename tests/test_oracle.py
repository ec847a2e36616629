"""Pins of the CPU oracle (oracle/ttor_oracle.cpp) before it is trusted as the checker.

- golden fixtures: factors from numpy's LAPACK (an independent dpotrf), tests/golden/
- exact known answers: integer GEMM (tests/shared/tests.cpp:361-411 shape) and the
  factor of L L^T for integer L with power-of-two diagonal (exact in FP64)
- tiled = monolithic (tests/shared/tests.cpp:568-583, 1e-13)
- run-once / dependency-order invariants of the runtime on random DAGs
  (tests/shared/tests.cpp:294-358)
- the block partition of 2d_cholesky.cpp:63-104 against an independent Python
  MT19937 + libstdc++ uniform_int_distribution (Lemire) implementation
- planted mistakes in a copy of the oracle's kernel inputs are caught
"""
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla

GOLD = Path(__file__).resolve().parent / "golden"
BACKENDS = ["textbook", "mkl"]


def _backend(oracle, name):
    if name == "mkl" and not oracle.have_mkl():
        pytest.skip("MKL runtime not present")
    oracle.set_backend(name)
    return oracle


@pytest.fixture(params=BACKENDS)
def orc(oracle, request):
    o = _backend(oracle, request.param)
    yield o
    o.set_backend("mkl" if o.have_mkl() else "textbook")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name,b", [("chol_bbt_192", 64), ("chol_bbt_100", 40), ("chol_refval_200", 64)])
@pytest.mark.parametrize("threads", [1, 3])
def test_tiled_cholesky_vs_golden(orc, name, b, threads):
    g = np.load(GOLD / f"{name}.npz")
    A, Lg = np.asfortranarray(g["A"]), g["L"]
    n = A.shape[0]
    sizes = [b] * (n // b) + ([n % b] if n % b else [])
    L, st = orc.cholesky(A, sizes=sizes, n_threads=threads, prio_kind=2)
    assert st["info"] == 0
    assert rel(L, Lg) < 1e-13


@pytest.mark.parametrize("accumulate", [False, True])
@pytest.mark.parametrize("prio", [0, 1, 2, 3])
def test_modes_and_priorities(orc, accumulate, prio):
    g = np.load(GOLD / "chol_bbt_192.npz")
    L, st = orc.cholesky(np.asfortranarray(g["A"]), 48, n_threads=4, prio_kind=prio, accumulate=accumulate)
    assert rel(L, g["L"]) < 1e-13


def test_gemm_integer_kat(orc):
    g = np.load(GOLD / "gemm_int.npz")
    A, B, C = (np.asfortranarray(g[k].astype(np.float64)) for k in ("A", "B", "C"))
    orc.dgemm_nt(-1.0, A, B, 1.0, C)
    assert np.array_equal(C, g["C_minus_ABt"].astype(np.float64))


def test_exact_integer_factor(orc):
    L = orc.int_lower_factor(240, 3)
    A = np.asfortranarray(L @ L.T)
    Lt, st = orc.cholesky(A, 64, n_threads=4)
    assert st["info"] == 0 and np.array_equal(Lt, L)


def test_kernels_vs_scipy(orc):
    rng = np.random.default_rng(0)
    n, m, k = 70, 50, 33
    A = np.asfortranarray(rng.standard_normal((n, n)))
    S = np.asfortranarray(A @ A.T + n * np.eye(n))
    Lp = S.copy(order="F")
    assert orc.dpotrf(Lp) == 0
    assert rel(np.tril(Lp), sla.cholesky(S, lower=True)) < 1e-14
    X = np.asfortranarray(rng.standard_normal((m, n)))
    Lt = np.asfortranarray(np.tril(Lp))
    Y = X.copy(order="F")
    orc.dtrsm_rltn(Lt, Y)
    assert rel(Y, sla.solve_triangular(Lt, X.T, lower=True).T) < 1e-13
    Ak = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.asfortranarray(rng.standard_normal((n, n)))
    C0 = C.copy()
    orc.dsyrk_ln(-1.0, Ak, 1.0, C)
    assert rel(np.tril(C), np.tril(C0 - Ak @ Ak.T)) < 1e-14
    assert np.array_equal(np.triu(C, 1), np.triu(C0, 1))


def test_not_spd_info_matches_lapack(orc):
    g = np.load(GOLD / "chol_bbt_192.npz")
    A = np.asfortranarray(g["A"].copy())
    A[130, 130] = -1.0
    _, info_ref = sla.lapack.dpotrf(A, lower=1)
    _, st = orc.cholesky(A, 64, n_threads=2)
    assert info_ref > 0 and st["info"] == info_ref


def test_bitwise_independent_of_threads_and_priority(oracle):
    o = _backend(oracle, "textbook")
    A = o.spd_bbt(256, 1)
    L0, _ = o.cholesky(A, 64, n_threads=1, prio_kind=0)
    for th, pk in [(2, 1), (4, 2), (8, 3)]:
        L, _ = o.cholesky(A, 64, n_threads=th, prio_kind=pk)
        assert np.array_equal(L, L0)
    o.set_backend("mkl" if o.have_mkl() else "textbook")


def test_planted_mistake_is_caught(oracle):
    # a transposed operand or a dropped update must fail the golden comparison
    g = np.load(GOLD / "chol_bbt_192.npz")
    A = np.asfortranarray(g["A"].copy())
    A[100, 3] += 1e-6  # one perturbed entry below the diagonal
    L, _ = oracle.cholesky(A, 64, n_threads=2)
    assert rel(L, g["L"]) > 1e-10


# ---------------------------------------------------------------------------
# Runtime invariants on random DAGs (tests/shared/tests.cpp:294-358)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,density,threads", [(1, 0.0, 1), (50, 0.1, 2), (200, 0.05, 8), (120, 0.5, 4)])
def test_random_dag_runs_once_in_order(oracle, n, density, threads):
    rng = np.random.default_rng(n * 13 + threads)
    perm = rng.permutation(n)
    edges = [[] for _ in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < density:
                edges[perm[i]].append(perm[j])
    indeg = np.zeros(n, dtype=np.int32)
    for e in edges:
        for v in e:
            indeg[v] += 1
    out_ptr = np.cumsum([0] + [len(e) for e in edges]).astype(np.int32)
    out_idx = np.array([v for e in edges for v in e], dtype=np.int32)
    order, runs = oracle.run_dag(threads, indeg, out_ptr, out_idx, rng.random(n))
    assert np.all(runs == 1)
    pos = np.empty(n, dtype=np.int64)
    pos[order] = np.arange(n)
    for u, e in enumerate(edges):
        for v in e:
            assert pos[u] < pos[v]


# ---------------------------------------------------------------------------
# Block partition (2d_cholesky.cpp:63-104)
# ---------------------------------------------------------------------------
class MT19937:
    """Matsumoto-Nishimura MT19937 with init_genrand seeding (the std::mt19937 engine)."""

    def __init__(self, seed):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.idx = 624

    def __call__(self):
        if self.idx >= 624:
            for i in range(624):
                y = (self.mt[i] & 0x80000000) | (self.mt[(i + 1) % 624] & 0x7FFFFFFF)
                self.mt[i] = self.mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.idx = 0
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def uniform_int_lemire(gen, a, b):
    """libstdc++ (GCC >= 11) uniform_int_distribution for a 32-bit engine: Lemire's nearly divisionless method."""
    rng = b - a + 1
    prod = gen() * rng
    low = prod & 0xFFFFFFFF
    if low < rng:
        thr = ((1 << 32) - rng) % rng
        while low < thr:
            prod = gen() * rng
            low = prod & 0xFFFFFFFF
    return a + (prod >> 32)


def test_mt19937_known_answer():
    # the C++ standard fixes the 10000th output of a default-seeded mt19937 (seed 5489)
    g = MT19937(5489)
    for _ in range(9999):
        g()
    assert g() == 4123659995


@pytest.mark.parametrize("b,nb,ub", [(256, 8, 256), (256, 8, 384), (192, 8, 288), (5, 10, 9), (100, 20, 200)])
def test_block_sizes(oracle, b, nb, ub):
    got = oracle.block_sizes(b, nb, ub)
    g = MT19937(2020)
    n, acc, exp = b * nb, 0, []
    for _ in range(nb - 1):
        s = min(n - acc, uniform_int_lemire(g, 2 * b - ub, ub))
        acc += s
        exp.append(s)
    exp.append(n - acc)
    assert got.tolist() == exp
    assert int(got.sum()) == n


def test_oracle_trace_joins_deps_like_the_reference_tools(oracle, tmp_path):
    # the oracle's Logger / DepsLogger files (restating util.cpp:103-127) read the way tools/ttor_logging.py reads
    # them: every task of the DAG has one '>run>' row whose name joins the DepsLogger edges, and '>deps>' rows exist
    import trace_util
    nb = 5
    A = oracle.spd_bbt(nb * 16, 3)
    log, deps = tmp_path / "log.csv", tmp_path / "deps.csv"
    oracle.cholesky(A, 16, n_threads=3, prio_kind=2, log_path=str(log), deps_path=str(deps))
    nodes = trace_util.dag_nodes(log)
    edges = trace_util.dag_edges(deps)
    names = set(edges["start"]) | set(edges["end"])
    assert nodes["task"].is_unique
    assert set(nodes["task"]) == names
    assert set(nodes["kind"]) == {"POTRF", "TRSM", "GEMM"}
    rows = trace_util.trace_rows(log)
    assert set(rows["what"]) >= {"run", "deps"}
