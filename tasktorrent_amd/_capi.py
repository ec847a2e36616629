"""ctypes declarations of include/tasktorrent_amd.h."""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
_LIB = _PKG / "_lib" / "libtasktorrent_amd.so"
HEADER = _PKG.parent / "include" / "tasktorrent_amd.h"

TTAD_OK = 0
TTAD_ERR_INVALID = 1
TTAD_ERR_OOM = 2
TTAD_ERR_NOT_SPD = 3
TTAD_ERR_TIMEOUT = 4
TTAD_ERR_DEVICE = 5
TTAD_ERR_IR_STALLED = 6
TTAD_ERR_GRAPH = 7
TTAD_ERR_BUSY = 8
TTAD_ERR_NOT_FOUND = 9
TTAD_OPT_FACTOR_RESIDUAL = 1
TTAD_OPT_RESIDUAL_ROWS = 2

TTAD_F64, TTAD_F32, TTAD_BF16, TTAD_BF16X3 = 0, 1, 2, 3
TTAD_POTRF, TTAD_TRSM, TTAD_SYRK, TTAD_GEMM = 0, 1, 2, 3
TTAD_SPIN = 5
TTAD_GEMM_ACC, TTAD_SYRK_ACC = 6, 7  # accumulate-mode GEMM / SYRK (product added into C under a sub-tile lock)
TTAD_ACCUMULATE = 0x100  # OR into prio_kind: the accumulate-mode DAG
TTAD_SCHED_STREAM, TTAD_SCHED_PERSISTENT = 0, 1

_NAMES = {0: "OK", 1: "INVALID", 2: "OOM", 3: "NOT_SPD", 4: "TIMEOUT", 5: "DEVICE", 6: "IR_STALLED", 7: "GRAPH",
          8: "BUSY", 9: "NOT_FOUND"}


class TtadError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"TTAD_ERR_{_NAMES.get(code, code)}: {msg}")
        self.code = code


class TaskDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("prec", ctypes.c_int32), ("mat", ctypes.c_int32),
                ("i", ctypes.c_int32), ("j", ctypes.c_int32), ("k", ctypes.c_int32)]


class GraphStats(ctypes.Structure):
    _fields_ = [("elapsed_s", ctypes.c_double), ("items_done", ctypes.c_longlong),
                ("items_total", ctypes.c_longlong), ("workers", ctypes.c_int), ("abort_code", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("fell_back", ctypes.c_int), ("backward_error", ctypes.c_double),
                ("factor_s", ctypes.c_double), ("solve_s", ctypes.c_double), ("total_s", ctypes.c_double),
                ("info", ctypes.c_int), ("prec_used", ctypes.c_int), ("factor_residual", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_longlong), ("max_tasks", ctypes.c_int), ("max_requests", ctypes.c_int),
                ("workers", ctypes.c_int), ("idle_timeout_s", ctypes.c_int)]


class ServerResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("n", ctypes.c_int), ("nrhs", ctypes.c_int),
                ("iterations", ctypes.c_int), ("reused", ctypes.c_int), ("factor_id", ctypes.c_longlong),
                ("backward_error", ctypes.c_double), ("latency_s", ctypes.c_double), ("queue_s", ctypes.c_double),
                ("fell_back", ctypes.c_int), ("info", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("pending", ctypes.c_int), ("active", ctypes.c_int), ("cached_factors", ctypes.c_int),
                ("workers", ctypes.c_int), ("arena_used", ctypes.c_longlong), ("arena_size", ctypes.c_longlong),
                ("items_run", ctypes.c_longlong)]


_lib = None

V = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
F = ctypes.c_float
LL = ctypes.c_longlong
PI = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "ttad_last_error": (ctypes.c_char_p, []),
    "ttad_version": (I, []),
    "ttad_init": (I, [I, ctypes.POINTER(V)]),
    "ttad_finalize": (I, [V]),
    "ttad_device_info": (I, [V, PI, ctypes.POINTER(LL), ctypes.c_char_p, I]),
    "ttad_synchronize": (I, [V, V]),
    "ttad_set_option": (I, [V, I, LL]),
    "ttad_factor_residual": (I, [V, V, V, ctypes.POINTER(D), V]),
    "ttad_dpotrf": (I, [V, I, V, I, PI, V]),
    "ttad_spotrf": (I, [V, I, V, I, PI, V]),
    "ttad_dtrsm": (I, [V, I, I, V, I, V, I, V]),
    "ttad_dtrsm_noalloc": (I, [V, I, I, V, I, V, I, V]),
    "ttad_strsm": (I, [V, I, I, V, I, V, I, V]),
    "ttad_dsyrk": (I, [V, I, I, D, V, I, D, V, I, V]),
    "ttad_ssyrk": (I, [V, I, I, F, V, I, F, V, I, V]),
    "ttad_dgemm": (I, [V, I, I, I, D, V, I, V, I, D, V, I, V]),
    "ttad_sgemm": (I, [V, I, I, I, F, V, I, V, I, F, V, I, V]),
    "ttad_gemm_bf16": (I, [V, I, I, I, F, V, I, V, I, F, V, I, I, V]),
    "ttad_gemm_bf16x3": (I, [V, I, I, I, F, V, I, V, I, F, V, I, I, V]),
    "ttad_matrix_create": (I, [V, I, PI, I, ctypes.POINTER(V)]),
    "ttad_matrix_destroy": (I, [V]),
    "ttad_matrix_set": (I, [V, V, LL, I, V]),
    "ttad_matrix_get": (I, [V, V, LL, I, V]),
    "ttad_matrix_copy": (I, [V, V, V]),
    "ttad_matrix_info": (I, [V, PI, PI, PI, ctypes.POINTER(LL)]),
    "ttad_matrix_tile": (I, [V, I, I, ctypes.POINTER(V), PI]),
    "ttad_potrf_tiled": (I, [V, V, I, I, I, V, PI]),
    "ttad_last_elapsed": (I, [V, ctypes.POINTER(D)]),
    "ttad_graph_create": (I, [V, I, ctypes.POINTER(TaskDesc), ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(V)]),
    "ttad_graph_bind": (I, [V, I, V]),
    "ttad_graph_seed": (I, [V, I]),
    "ttad_graph_clear_seeds": (I, [V]),
    "ttad_graph_launch": (I, [V, I, I, V]),
    "ttad_graph_wait": (I, [V, D, ctypes.POINTER(GraphStats)]),
    "ttad_graph_write_trace": (I, [V, ctypes.c_char_p, ctypes.c_char_p]),
    "ttad_graph_write_deps": (I, [V, ctypes.c_char_p]),
    "ttad_graph_destroy": (I, [V]),
    "ttad_solve": (I, [V, V, I, V, LL, V, LL, V]),
    "ttad_residual": (I, [V, V, I, V, LL, V, LL, V, LL, ctypes.POINTER(D), V]),
    "ttad_solve_ir": (I, [V, V, V, I, V, LL, V, LL, D, I, ctypes.POINTER(SolveStats), V]),
    "ttad_posv_mixed": (I, [V, V, I, I, V, LL, V, LL, D, I, ctypes.POINTER(SolveStats), V]),
    "ttad_server_create": (I, [V, ctypes.POINTER(ServerConfig), ctypes.POINTER(V)]),
    "ttad_server_submit": (I, [V, V, I, I, V, V, D, I, LL, I, ctypes.POINTER(LL)]),
    "ttad_server_poll": (I, [V, ctypes.POINTER(ServerResult), I, PI]),
    "ttad_server_query": (I, [V, ctypes.POINTER(ServerStats)]),
    "ttad_server_destroy": (I, [V]),
    "ttad_cholesky_dag": (I, [I, I, PI, PI, ctypes.POINTER(TaskDesc), ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_int32)]),
}


def lib_path() -> Path:
    return Path(os.environ.get("TTAD_LIB", _LIB))


def lib():
    """Load the HIP library; raises if it was not built (there is no fallback path)."""
    global _lib
    if _lib is None:
        p = lib_path()
        if not p.exists():
            raise RuntimeError(f"tasktorrent_amd: HIP library not built ({p}); run __graft_entry__.build()")
        L = ctypes.CDLL(str(p))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != TTAD_OK:
        raise TtadError(rc, lib().ttad_last_error().decode(errors="replace"))
    return rc


def declared_symbols(header: Path | None = None) -> list[str]:
    """Function names declared in a public header (for the export check)."""
    import re
    text = (header or HEADER).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ttad_[a-z0-9_]+)\s*\(", text)))
