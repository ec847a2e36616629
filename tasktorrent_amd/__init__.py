"""tasktorrent_amd: MI355X-native tiled dense Cholesky (Python front of the C-ABI).

The product is the HIP library libtasktorrent_amd.so (include/tasktorrent_amd.h);
this module binds it with ctypes for tests, benchmarks and Python callers. There
is no CPU fallback: every call goes to the HIP library and fails loudly if the
library is not built or no GPU is present.

PyTorch is used only as plumbing for device buffers and streams.
"""
from __future__ import annotations

import ctypes
import weakref
from pathlib import Path

import numpy as np

from . import _capi
from ._capi import (TTAD_ACCUMULATE, TTAD_BF16, TTAD_BF16X3, TTAD_F32, TTAD_F64, TTAD_GEMM, TTAD_GEMM_ACC,
                    TTAD_POTRF, TTAD_SCHED_PERSISTENT, TTAD_SCHED_STREAM, TTAD_SPIN, TTAD_SYRK, TTAD_SYRK_ACC,
                    TTAD_TRSM, TTAD_ERR_IR_STALLED, TTAD_ERR_NOT_SPD, TtadError, lib, lib_path)

__all__ = [
    "Context", "TiledMatrix", "Graph", "Server", "TtadError", "lib", "lib_path",
    "TTAD_F64", "TTAD_F32", "TTAD_BF16", "TTAD_BF16X3", "TTAD_POTRF", "TTAD_TRSM", "TTAD_SYRK", "TTAD_GEMM", "TTAD_SPIN",
    "TTAD_GEMM_ACC", "TTAD_SYRK_ACC", "TTAD_ACCUMULATE",
    "TTAD_SCHED_STREAM", "TTAD_SCHED_PERSISTENT", "TTAD_ERR_NOT_SPD", "TTAD_ERR_IR_STALLED",
]

PREC = {"f64": TTAD_F64, "f32": TTAD_F32, "bf16": TTAD_BF16, "bf16x3": TTAD_BF16X3}


def _stream_of(torch_stream=None):
    import torch
    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """A device context (ttad_init). Replaces the reference Threadpool construction."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        _capi.check(lib().ttad_init(device, ctypes.byref(self._h)))
        self.device = device
        self._owned = weakref.WeakSet()  # matrices, graphs and servers created on this context

    @property
    def handle(self):
        return self._h

    def _own(self, obj):
        self._owned.add(obj)

    def close(self):
        """Destroys what is still open on this context (servers, then graphs, then
        matrices), then the context itself: a handle must never outlive its context."""
        if self._h:
            order = {Server: 0, Graph: 1, TiledMatrix: 2}
            for obj in sorted(list(self._owned), key=lambda o: order.get(type(o), 3)):
                obj.close()
            lib().ttad_finalize(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        ncu = ctypes.c_int()
        hbm = ctypes.c_longlong()
        name = ctypes.create_string_buffer(64)
        _capi.check(lib().ttad_device_info(self._h, ctypes.byref(ncu), ctypes.byref(hbm), name, 64))
        return {"n_cu": ncu.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}

    def synchronize(self, stream=None):
        _capi.check(lib().ttad_synchronize(self._h, stream))

    # ---- tile kernels on torch device tensors (column-major views, see _colmajor) ----
    def potrf(self, A, n: int, lda: int, stream=None) -> int:
        info = ctypes.c_int()
        fn = lib().ttad_dpotrf if A.dtype.itemsize == 8 else lib().ttad_spotrf
        rc = fn(self._h, n, _ptr(A), lda, ctypes.byref(info), stream or _stream_of())
        if rc not in (0, _capi.TTAD_ERR_NOT_SPD):
            _capi.check(rc)
        return info.value

    def trsm(self, m, n, L, ldl, B, ldb, stream=None):
        fn = lib().ttad_dtrsm if L.dtype.itemsize == 8 else lib().ttad_strsm
        _capi.check(fn(self._h, m, n, _ptr(L), ldl, _ptr(B), ldb, stream or _stream_of()))

    def trsm_noalloc(self, m, n, L, ldl, B, ldb, stream=None):
        """FP64 trsm whose diagonal-block inverses live in a static per-GPU scratch (n <= 4096)."""
        _capi.check(lib().ttad_dtrsm_noalloc(self._h, m, n, _ptr(L), ldl, _ptr(B), ldb, stream or _stream_of()))

    def syrk(self, n, k, alpha, A, lda, beta, C, ldc, stream=None):
        fn = lib().ttad_dsyrk if A.dtype.itemsize == 8 else lib().ttad_ssyrk
        _capi.check(fn(self._h, n, k, alpha, _ptr(A), lda, beta, _ptr(C), ldc, stream or _stream_of()))

    def gemm(self, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
        fn = lib().ttad_dgemm if A.dtype.itemsize == 8 else lib().ttad_sgemm
        _capi.check(fn(self._h, m, n, k, alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(C), ldc,
                       stream or _stream_of()))

    def gemm_bf16(self, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, syrk=False, split=False, stream=None):
        fn = lib().ttad_gemm_bf16x3 if split else lib().ttad_gemm_bf16
        _capi.check(fn(self._h, m, n, k, alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(C), ldc, int(syrk),
                       stream or _stream_of()))

    # ---- tiled Cholesky ----
    def potrf_tiled(self, M: "TiledMatrix", schedule=TTAD_SCHED_PERSISTENT, prio_kind: int = 2,
                    prec: str | int = "f64", stream=None) -> int:
        p = PREC[prec] if isinstance(prec, str) else prec
        info = ctypes.c_int()
        rc = lib().ttad_potrf_tiled(self._h, M.handle, schedule, prio_kind, p, stream, ctypes.byref(info))
        if rc not in (0, _capi.TTAD_ERR_NOT_SPD):
            _capi.check(rc)
        return info.value


    # ---- solve phase (device FP64 column-major n x nrhs arrays: torch (nrhs, n) tensors) ----
    def solve(self, L: "TiledMatrix", nrhs: int, B, X, stream=None):
        _capi.check(lib().ttad_solve(self._h, L.handle, nrhs, _ptr(B), B.shape[-1], _ptr(X), X.shape[-1], stream))
        self.synchronize(stream)

    def residual(self, A: "TiledMatrix", nrhs: int, B, X, R, stream=None) -> float:
        eta = ctypes.c_double()
        _capi.check(lib().ttad_residual(self._h, A.handle, nrhs, _ptr(B), B.shape[-1], _ptr(X), X.shape[-1],
                                        _ptr(R), R.shape[-1], ctypes.byref(eta), stream))
        self.synchronize(stream)
        return eta.value

    @staticmethod
    def _stats(rc, st):
        return {"rc": rc, "iterations": st.iterations, "fell_back": st.fell_back,
                "backward_error": st.backward_error, "factor_residual": st.factor_residual,
                "factor_s": st.factor_s, "solve_s": st.solve_s,
                "total_s": st.total_s, "info": st.info, "prec_used": st.prec_used}

    def solve_ir(self, A: "TiledMatrix", L: "TiledMatrix", nrhs: int, B, X, tol=1e-12, max_iter=30, stream=None):
        st = _capi.SolveStats()
        rc = lib().ttad_solve_ir(self._h, A.handle, L.handle, nrhs, _ptr(B), B.shape[-1], _ptr(X), X.shape[-1],
                                 tol, max_iter, ctypes.byref(st), stream)
        if rc not in (0, _capi.TTAD_ERR_IR_STALLED):
            _capi.check(rc)
        self.synchronize(stream)
        return self._stats(rc, st)

    def posv_mixed(self, A: "TiledMatrix", prec, nrhs: int, B, X, tol=1e-12, max_iter=30, stream=None):
        """Mixed-precision factor + FP64 refinement with FP64 fallback (ttad_posv_mixed)."""
        p = PREC[prec] if isinstance(prec, str) else prec
        st = _capi.SolveStats()
        rc = lib().ttad_posv_mixed(self._h, A.handle, p, nrhs, _ptr(B), B.shape[-1], _ptr(X), X.shape[-1], tol,
                                   max_iter, ctypes.byref(st), stream)
        if rc not in (0, _capi.TTAD_ERR_IR_STALLED, _capi.TTAD_ERR_NOT_SPD):
            _capi.check(rc)
        self.synchronize(stream)
        return self._stats(rc, st)

    def factor_residual(self, A: "TiledMatrix", L: "TiledMatrix", stream=None) -> float:
        """||L L^T - A||_F / ||A||_F on the device (ttad_factor_residual)."""
        r = ctypes.c_double()
        _capi.check(lib().ttad_factor_residual(self._h, A.handle, L.handle, ctypes.byref(r), stream))
        return r.value

    def set_option(self, option: int, value: int) -> None:
        _capi.check(lib().ttad_set_option(self._h, option, value))

    def report_factor_residual(self, on: bool = True) -> None:
        """ttad_solve_ir / ttad_posv_mixed also report ||L L^T - A||_F / ||A||_F (untimed, O(n^3))."""
        self.set_option(_capi.TTAD_OPT_FACTOR_RESIDUAL, int(on))

    def last_elapsed(self) -> float:
        """Device seconds of the last potrf_tiled (HIP events on its stream)."""
        t = ctypes.c_double()
        _capi.check(lib().ttad_last_elapsed(self._h, ctypes.byref(t)))
        return t.value


class TiledMatrix:
    """Lower tiles of a symmetric matrix in HBM (ttad_matrix)."""

    def __init__(self, ctx: Context, block_sizes, storage: str | int = "f64"):
        self.ctx = ctx
        bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
        self.block_sizes = bs
        self.n = int(bs.sum())
        st = PREC[storage] if isinstance(storage, str) else storage
        self.storage = st
        self._h = ctypes.c_void_p()
        _capi.check(lib().ttad_matrix_create(ctx.handle, len(bs), bs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                             st, ctypes.byref(self._h)))
        ctx._own(self)

    @classmethod
    def uniform(cls, ctx, n: int, b: int, storage="f64"):
        nb = (n + b - 1) // b
        sizes = [b] * (nb - 1) + [n - b * (nb - 1)]
        return cls(ctx, sizes, storage)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().ttad_matrix_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, A, stream=None):
        """A: numpy (host, column-major n x n FP64) or a torch CUDA FP64 tensor holding A^T row-major."""
        if isinstance(A, np.ndarray):
            Af = np.asfortranarray(A, dtype=np.float64)
            _capi.check(lib().ttad_matrix_set(self._h, Af.ctypes.data_as(ctypes.c_void_p), self.n, 0, stream))
        else:
            assert A.is_cuda and A.dtype.itemsize == 8 and A.is_contiguous()
            _capi.check(lib().ttad_matrix_set(self._h, _ptr(A), A.shape[1], 1, stream or _stream_of()))

    def get(self, out=None, stream=None):
        """Lower triangle as a host column-major FP64 array (strict upper = 0), or into a torch CUDA tensor."""
        if out is None:
            L = np.zeros((self.n, self.n), dtype=np.float64, order="F")
            _capi.check(lib().ttad_matrix_get(self._h, L.ctypes.data_as(ctypes.c_void_p), self.n, 0, stream))
            return L
        _capi.check(lib().ttad_matrix_get(self._h, _ptr(out), out.shape[1], 1, stream or _stream_of()))
        return out

    def copy_from(self, src: "TiledMatrix", stream=None):
        _capi.check(lib().ttad_matrix_copy(self._h, src.handle, stream))

    def bytes(self) -> int:
        b = ctypes.c_longlong()
        _capi.check(lib().ttad_matrix_info(self._h, None, None, None, ctypes.byref(b)))
        return b.value


class Graph:
    """A compiled device task graph (ttad_graph): the Taskflow/Threadpool drop-in at the C level."""

    def __init__(self, ctx: Context, tasks, indegree, out_ptr, out_idx, priority=None):
        n = len(tasks)
        arr = (_capi.TaskDesc * n)(*[_capi.TaskDesc(*t) for t in tasks])
        self._keep = [np.ascontiguousarray(indegree, dtype=np.int32), np.ascontiguousarray(out_ptr, dtype=np.int32),
                      np.ascontiguousarray(out_idx, dtype=np.int32),
                      np.ascontiguousarray(priority if priority is not None else np.zeros(n), dtype=np.float64)]
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        self._h = ctypes.c_void_p()
        _capi.check(lib().ttad_graph_create(ctx.handle, n, arr, ip(self._keep[0]), ip(self._keep[1]),
                                            ip(self._keep[2]) if len(self._keep[2]) else None,
                                            self._keep[3].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            ctypes.byref(self._h)))
        self.ctx = ctx
        ctx._own(self)

    def bind(self, slot: int, M: TiledMatrix):
        _capi.check(lib().ttad_graph_bind(self._h, slot, M.handle))

    def seed(self, task: int):
        _capi.check(lib().ttad_graph_seed(self._h, task))

    def launch(self, n_workers: int = 0, trace: bool = False, stream=None):
        _capi.check(lib().ttad_graph_launch(self._h, n_workers, int(trace), stream))

    def wait(self, timeout_s: float = 120.0):
        st = _capi.GraphStats()
        rc = lib().ttad_graph_wait(self._h, timeout_s, ctypes.byref(st))
        return rc, {"elapsed_s": st.elapsed_s, "items_done": st.items_done, "items_total": st.items_total,
                    "workers": st.workers, "abort_code": st.abort_code}

    def write_trace(self, path: str, prefix: str = "Wk_"):
        _capi.check(lib().ttad_graph_write_trace(self._h, path.encode(), prefix.encode()))

    def write_deps(self, path: str):
        """Edges as the reference DepsLogger writes them: one "SRC,DST" line per edge."""
        _capi.check(lib().ttad_graph_write_deps(self._h, path.encode()))

    def close(self):
        if self._h:
            lib().ttad_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Device-resident serving loop (ttad_server): continuous batching of SPD solves.

    submit() never blocks; poll() returns completed requests as dicts and copies each
    solution into the numpy array given at submission.
    """

    def __init__(self, ctx: Context, hbm_budget: int = 4 << 30, max_tasks: int = 1 << 20, max_requests: int = 1024,
                 workers: int = 0, idle_timeout_s: int = 600):
        cfg = _capi.ServerConfig(hbm_budget, max_tasks, max_requests, workers, idle_timeout_s)
        self._h = ctypes.c_void_p()
        _capi.check(lib().ttad_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(self._h)))
        self.ctx = ctx
        ctx._own(self)
        self._keep = {}

    def submit(self, A: TiledMatrix, B: np.ndarray, prec="bf16x3", tol=1e-12, max_iter=30, reuse_factor=-1,
               keep_factor=False):
        B = np.asfortranarray(B, dtype=np.float64)
        if B.ndim == 1:
            B = B.reshape(-1, 1, order="F")
        X = np.zeros_like(B, order="F")
        p = PREC[prec] if isinstance(prec, str) else prec
        rid = ctypes.c_longlong()
        _capi.check(lib().ttad_server_submit(self._h, A.handle, p, B.shape[1], B.ctypes.data_as(ctypes.c_void_p),
                                             X.ctypes.data_as(ctypes.c_void_p), tol, max_iter, reuse_factor,
                                             int(keep_factor), ctypes.byref(rid)))
        self._keep[rid.value] = (B, X)
        return rid.value, X

    def poll(self, max_out: int = 256):
        buf = (_capi.ServerResult * max_out)()
        n = ctypes.c_int()
        _capi.check(lib().ttad_server_poll(self._h, buf, max_out, ctypes.byref(n)))
        out = []
        for r in buf[:n.value]:
            d = {f: getattr(r, f) for f, _ in _capi.ServerResult._fields_}
            kept = self._keep.pop(r.id, None)
            d["X"] = kept[1] if kept else None
            out.append(d)
        return out

    def wait_all(self, ids, timeout_s: float = 120.0):
        import time
        want, got, t0 = set(ids), {}, time.time()
        while want - set(got):
            for r in self.poll():
                got[r["id"]] = r
            if time.time() - t0 > timeout_s:
                raise TimeoutError(f"server: {len(want - set(got))} requests not done after {timeout_s} s")
            time.sleep(20e-6)
        return got

    def stats(self):
        st = _capi.ServerStats()
        _capi.check(lib().ttad_server_query(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _capi.ServerStats._fields_}

    def close(self):
        if self._h:
            lib().ttad_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cholesky_dag(num_blocks: int, prio_kind: int = 2):
    """The Cholesky PTG compiled on the host (no device): tasks, indegree, priority, CSR out-edges.

    tasks is a list of (kind, prec, mat, i, j, k); kind uses TTAD_POTRF..TTAD_GEMM (SYRK for i == j).
    """
    nt, ne = ctypes.c_int(), ctypes.c_int()
    _capi.check(lib().ttad_cholesky_dag(num_blocks, prio_kind, ctypes.byref(nt), ctypes.byref(ne),
                                        None, None, None, None, None))
    n, e = nt.value, ne.value
    tasks = (_capi.TaskDesc * n)()
    indeg = np.zeros(n, dtype=np.int32)
    prio = np.zeros(n, dtype=np.float64)
    out_ptr = np.zeros(n + 1, dtype=np.int32)
    out_idx = np.zeros(max(1, e), dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    _capi.check(lib().ttad_cholesky_dag(num_blocks, prio_kind, ctypes.byref(nt), ctypes.byref(ne), tasks, ip(indeg),
                                        prio.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ip(out_ptr),
                                        ip(out_idx)))
    tl = [(t.kind, t.prec, t.mat, t.i, t.j, t.k) for t in tasks]
    return tl, indeg, prio, out_ptr, out_idx[:e]


def cholesky_spin_bench(ctx: Context, num_blocks: int, spin_us: int, prio_kind: int = 2, n_workers: int = 0,
                        reps: int = 3) -> dict:
    """Scheduler-overhead benchmark (mini_benchmarks/shared/ttor_cholesky.cpp:22-108): the Cholesky DAG of
    num_blocks x num_blocks tiles with every task body a spin of spin_us microseconds on the device executor.
    efficiency = tasks * spin / (elapsed * workers), as the reference computes it (:104-108); the best of reps.
    """
    tasks, indeg, prio, out_ptr, out_idx = cholesky_dag(num_blocks, prio_kind)
    spin = [(TTAD_SPIN, 0, 0, t[3], t[4], int(spin_us)) for t in tasks]  # keeps (i, j): queue homing
    G = Graph(ctx, spin, indeg, out_ptr, out_idx, prio)
    try:
        G.seed([task_name(t) for t in tasks].index("POTRF_0"))
        best = None
        for _ in range(reps):
            G.launch(n_workers)
            rc, st = G.wait(120.0)
            _capi.check(rc)
            if best is None or st["elapsed_s"] < best["elapsed_s"]:
                best = st
    finally:
        G.close()
    n = len(tasks)
    eff = n * spin_us * 1e-6 / (best["elapsed_s"] * best["workers"])
    return {"tasks": n, "spin_us": spin_us, "workers": best["workers"], "elapsed_s": best["elapsed_s"],
            "efficiency": eff}


def task_name(t) -> str:
    """Reference task names (2d_cholesky.cpp:206-217): POTRF_j, TRSM_i_j, GEMM_k_i_j."""
    kind, _, _, i, j, k = t
    if kind == TTAD_POTRF:
        return f"POTRF_{j}"
    if kind == TTAD_TRSM:
        return f"TRSM_{i}_{j}"
    return f"GEMM_{k}_{i}_{j}"
