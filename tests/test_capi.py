"""CPU checks of the C-ABI library (no GPU calls):
- it loads and exports every function include/tasktorrent_amd.h declares;
- the Cholesky DAG compiled through the Taskflow drop-in equals, edge for edge,
  the dependencies the oracle's restated Taskflow/Threadpool records while
  running 2d_cholesky.cpp's PTG (its --depslog edges, 2d_cholesky.cpp:325-329,
  414-418, 509-528), with the same indegrees and PrioKind priorities.
"""
import ctypes

import numpy as np
import pytest

import tasktorrent_amd as tt
from tasktorrent_amd import _capi


def test_library_exports_header_symbols():
    L = tt.lib()
    declared = _capi.declared_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert L.ttad_version() >= 1


def test_bindings_cover_header():
    declared = set(_capi.declared_symbols())
    assert declared <= set(_capi._SIGS), declared - set(_capi._SIGS)


def test_invalid_args_return_codes():
    L = tt.lib()
    assert L.ttad_init(0, None) == _capi.TTAD_ERR_INVALID
    assert b"NULL" in L.ttad_last_error()
    n, e = ctypes.c_int(), ctypes.c_int()
    assert L.ttad_cholesky_dag(0, 2, ctypes.byref(n), ctypes.byref(e), None, None, None, None, None) != 0
    assert L.ttad_cholesky_dag(4, 7, ctypes.byref(n), ctypes.byref(e), None, None, None, None, None) != 0


@pytest.mark.parametrize("nb", [1, 2, 5, 8])
def test_dag_counts(nb):
    tasks, indeg, prio, out_ptr, out_idx = tt.cholesky_dag(nb, 2)
    kinds = [t[0] for t in tasks]
    assert kinds.count(tt.TTAD_POTRF) == nb
    assert kinds.count(tt.TTAD_TRSM) == nb * (nb - 1) // 2
    assert kinds.count(tt.TTAD_SYRK) == nb * (nb - 1) // 2
    assert kinds.count(tt.TTAD_GEMM) == sum((nb - k - 1) * (nb - k - 2) // 2 for k in range(nb))
    # every task receives exactly its indegree: edges + the single host seed of POTRF_0
    incoming = np.bincount(out_idx, minlength=len(tasks))
    seed = np.array([1 if tt.task_name(t) == "POTRF_0" else 0 for t in tasks])
    assert np.array_equal(incoming + seed, indeg)


@pytest.mark.parametrize("nb,prio_kind", [(4, 2), (6, 0), (7, 3), (5, 1)])
def test_dag_equals_oracle_deps(oracle, tmp_path, nb, prio_kind):
    tasks, indeg, prio, out_ptr, out_idx = tt.cholesky_dag(nb, prio_kind)
    names = [tt.task_name(t) for t in tasks]
    ours = {(names[u], names[v]) for u in range(len(tasks)) for v in out_idx[out_ptr[u]:out_ptr[u + 1]]}
    A = oracle.spd_bbt(nb * 8, 1)
    deps = tmp_path / "deps.csv"
    oracle.cholesky(A, 8, n_threads=3, prio_kind=prio_kind, deps_path=str(deps))
    theirs = {tuple(line.split(",")) for line in deps.read_text().splitlines() if line}
    assert ours == theirs
    assert len(out_idx) == len(theirs)  # no duplicated edge


def test_dag_priorities_match_prio_kind():
    nb = 6
    for kind in range(4):
        tasks, indeg, prio, _, _ = tt.cholesky_dag(nb, kind)
        for t, p in zip(tasks, prio):
            k_, _, _, i, j, k = t
            if kind == 2:
                exp = {tt.TTAD_POTRF: 9 * (nb - j) - 1, tt.TTAD_TRSM: 9 * (nb - j) - 2}.get(k_, 9 * nb - 9 * j - 2)
                assert p == exp
            if kind == 0:
                assert p == {tt.TTAD_POTRF: 3.0, tt.TTAD_TRSM: 2.0}.get(k_, 1.0)


@pytest.mark.parametrize("accumulate", [False, True])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("nb", [1, 4, 9])
def test_dag_priorities_equal_oracle(oracle, nb, kind, accumulate):
    """Every task's priority, every PrioKind, chained and accumulate mode, against the oracle's PTG."""
    flag = tt.TTAD_ACCUMULATE if accumulate else 0
    tasks, _, prio, _, _ = tt.cholesky_dag(nb, kind | flag)
    for t, p in zip(tasks, prio):
        name = tt.task_name(t).split("_")[0]
        _, _, _, i, j, k = t
        assert p == oracle.priority(kind, nb, accumulate, name, i, j, k), (kind, accumulate, t, p)


def test_dag_priorities_kind_1_and_3_values():
    """Hand-evaluated PrioKind 1 (row) and 3 (cp_row) values of 2d_cholesky.cpp:157-204 at nb = 4."""
    nb = 4
    for kind, acc, want in [
        (1, False, {"POTRF_1": 9.0, "TRSM_3_1": 2.0, "GEMM_0_2_1": 2.0, "GEMM_1_3_3": 1.0}),
        (3, False, {"POTRF_1": 26.0 + 288.0, "TRSM_3_1": 1 + 4 * 25 + 144.0, "GEMM_0_2_1": 2 + 4 * 31.0,
                    "GEMM_1_3_3": 1 + 4 * 19.0}),
        # accumulate mode drops the k term of the GEMM priority: 9nb - 9j - 2
        (3, True, {"GEMM_0_2_1": 2 + 4 * 25.0, "GEMM_1_3_3": 1 + 4 * 7.0}),
    ]:
        tasks, _, prio, _, _ = tt.cholesky_dag(nb, kind | (tt.TTAD_ACCUMULATE if acc else 0))
        got = {tt.task_name(t): p for t, p in zip(tasks, prio)}
        for name, v in want.items():
            assert got[name] == v, (kind, acc, name, got[name], v)


@pytest.mark.parametrize("nb", [1, 2, 3, 6, 11])
def test_dag_bottom_level_priorities(nb):
    # prio_kind 4: longest path to the end of the DAG with weights POTRF 2, TRSM 2, GEMM/SYRK 1,
    # checked against a dynamic program over the compiled edges in reverse topological order
    tasks, indeg, prio, out_ptr, out_idx = tt.cholesky_dag(nb, 4)
    n = len(tasks)
    weight = {tt.TTAD_POTRF: 2.0, tt.TTAD_TRSM: 2.0, tt.TTAD_GEMM: 1.0, tt.TTAD_SYRK: 1.0}
    pending = np.bincount(out_idx, minlength=n)
    order = [t for t in range(n) if pending[t] == 0]
    for t in order:  # grows while it is walked (Kahn)
        for s in out_idx[out_ptr[t]:out_ptr[t + 1]]:
            pending[s] -= 1
            if pending[s] == 0:
                order.append(s)
    assert len(order) == n
    bl = np.zeros(n)
    for t in reversed(order):
        succ = out_idx[out_ptr[t]:out_ptr[t + 1]]
        bl[t] = weight[tasks[t][0]] + (bl[succ].max() if len(succ) else 0.0)
    assert np.array_equal(bl, prio)
    # the sink is POTRF(nb-1); the source POTRF(0) carries the longest path
    names = [tt.task_name(t) for t in tasks]
    assert prio[names.index(f"POTRF_{nb - 1}")] == 2.0
    assert prio[names.index("POTRF_0")] == prio.max()


@pytest.mark.parametrize("nb", [1, 2, 5, 9])
def test_dag_accumulate_mode(nb):
    # 2d_cholesky.cpp --accumulate (:345-351, 449-455, 532-540), with the ACCU task fused into its GEMM:
    # no GEMM -> GEMM edges; POTRF(j) waits for the j updates of (j, j), TRSM(i, j) for POTRF(j) and j updates.
    tasks, indeg, prio, out_ptr, out_idx = tt.cholesky_dag(nb, 2 | tt.TTAD_ACCUMULATE)
    kinds = [t[0] for t in tasks]
    assert kinds.count(tt.TTAD_POTRF) == nb
    assert kinds.count(tt.TTAD_TRSM) == nb * (nb - 1) // 2
    assert kinds.count(tt.TTAD_SYRK_ACC) == nb * (nb - 1) // 2
    assert kinds.count(tt.TTAD_GEMM_ACC) == sum((nb - k - 1) * (nb - k - 2) // 2 for k in range(nb))
    for t, d in zip(tasks, indeg):
        kind, _, _, i, j, k = t
        expect = {tt.TTAD_POTRF: max(1, j), tt.TTAD_TRSM: 1 + j, tt.TTAD_SYRK_ACC: 1, tt.TTAD_GEMM_ACC: 2}[kind]
        assert d == expect, (t, d)
    incoming = np.bincount(out_idx, minlength=len(tasks))
    seed = np.array([1 if tt.task_name(t) == "POTRF_0" else 0 for t in tasks])
    assert np.array_equal(incoming + seed, indeg)
    upd = {tt.TTAD_GEMM_ACC, tt.TTAD_SYRK_ACC}
    for u in range(len(tasks)):
        if kinds[u] in upd:
            assert not any(kinds[v] in upd for v in out_idx[out_ptr[u]:out_ptr[u + 1]])
