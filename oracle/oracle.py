"""CPU oracle for the tiled dense Cholesky path (ctypes front of ttor_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing
measured on the GPU. The product library (tasktorrent_amd) never imports it.

What it restates (reference file:line):
  * cholesky()        miniapp/dense_cholesky/2d_cholesky.cpp:41-690 on one rank
                      (Taskflow/Threadpool: tasktorrent/src/taskflow.hpp:40-297,
                      threadpool_shared.cpp:17-257)
  * block_sizes()     2d_cholesky.cpp:63-104
  * ref_val_matrix()  2d_cholesky.cpp:133 (A(i,j) = 1/((i-j)^2+1))
  * solve_check()     2d_cholesky.cpp:646-689 (two triangular solves, forward error)
  * the tile kernels  LAPACKE_dpotrf / cblas_dtrsm / cblas_dsyrk / cblas_dgemm as
                      called at 2d_cholesky.cpp:319, 404-405, 490-494
The metrics (factor residual, backward error) are the quantities the solver reports.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libttor_oracle.so"
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (g++ only)."""
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(_HERE), "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        L.orc_set_backend.argtypes = [ctypes.c_int]
        L.orc_backend_name.restype = ctypes.c_char_p
        L.orc_dpotrf.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_dtrsm_rltn.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_dsyrk_ln.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int,
                                   ctypes.c_double, _dp, ctypes.c_int]
        L.orc_dgemm_nt.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int,
                                   _dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int]
        L.orc_block_sizes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip]
        L.orc_cholesky.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp,
                                   ctypes.c_longlong, _dp, ctypes.c_char_p, ctypes.c_char_p]
        L.orc_priority.argtypes = [ctypes.c_int] * 7
        L.orc_priority.restype = ctypes.c_double
        L.orc_run_dag.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _dp, _ip, _ip]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def set_backend(kind: str) -> str:
    """'mkl' (sequential MKL via dlopen, as the reference's BLAS_SEQ_LIBS) or 'textbook'."""
    k = {"textbook": 0, "mkl": 1}[kind]
    if lib().orc_set_backend(k) != 0:
        raise RuntimeError("oracle: MKL runtime (libmkl_rt) not found")
    return backend_name()


def backend_name() -> str:
    return lib().orc_backend_name().decode()


def have_mkl() -> bool:
    cur = backend_name()
    ok = lib().orc_set_backend(1) == 0
    lib().orc_set_backend(0 if cur == "textbook" else 1)
    return ok


# ----------------------------------------------------------------------------- kernels
def dpotrf(A: np.ndarray) -> int:
    """In-place lower Cholesky of a Fortran-ordered square array; returns LAPACK info."""
    assert A.flags.f_contiguous and A.dtype == np.float64
    return lib().orc_dpotrf(A.shape[0], _p(A), max(1, A.shape[0]))


def dtrsm_rltn(L: np.ndarray, B: np.ndarray) -> None:
    """B := B L^{-T} (Right, Lower, Trans, NonUnit), Fortran order, in place."""
    assert L.flags.f_contiguous and B.flags.f_contiguous
    lib().orc_dtrsm_rltn(B.shape[0], B.shape[1], _p(L), max(1, L.shape[0]), _p(B), max(1, B.shape[0]))


def dsyrk_ln(alpha: float, A: np.ndarray, beta: float, C: np.ndarray) -> None:
    """C := beta C + alpha A A^T on the lower triangle."""
    assert A.flags.f_contiguous and C.flags.f_contiguous
    lib().orc_dsyrk_ln(C.shape[0], A.shape[1], alpha, _p(A), max(1, A.shape[0]), beta, _p(C), max(1, C.shape[0]))


def dgemm_nt(alpha: float, A: np.ndarray, B: np.ndarray, beta: float, C: np.ndarray) -> None:
    """C := beta C + alpha A B^T."""
    assert A.flags.f_contiguous and B.flags.f_contiguous and C.flags.f_contiguous
    lib().orc_dgemm_nt(C.shape[0], C.shape[1], A.shape[1], alpha, _p(A), max(1, A.shape[0]), _p(B),
                       max(1, B.shape[0]), beta, _p(C), max(1, C.shape[0]))


# ----------------------------------------------------------------------------- DAG
def block_sizes(block_size: int, num_blocks: int, upper_block_size: int | None = None) -> np.ndarray:
    ub = block_size if upper_block_size is None else upper_block_size
    out = np.zeros(num_blocks, dtype=np.int32)
    if lib().orc_block_sizes(block_size, num_blocks, ub, _pi(out)) != 0:
        raise ValueError("invalid block partition")
    return out


def cholesky(A: np.ndarray, block_size: int | None = None, n_threads: int = 1, prio_kind: int = 2,
             accumulate: bool = False, sizes=None, log_path: str | None = None,
             deps_path: str | None = None):
    """Tiled Cholesky of A's lower triangle through the restated Taskflow/Threadpool DAG.

    Returns (L, stats); L is a new Fortran array with the strict upper triangle zeroed.
    stats = dict(total, potrf, trsm, gemm, accu, steals, info) with times in seconds.
    """
    n = A.shape[0]
    if sizes is None:
        assert block_size is not None and block_size > 0
        sizes = [block_size] * (n // block_size) + ([n % block_size] if n % block_size else [])
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    assert int(sizes.sum()) == n
    W = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    st = np.zeros(7, dtype=np.float64)
    lib().orc_cholesky(n_threads, prio_kind, int(accumulate), len(sizes), _pi(sizes), _p(W), n, _p(st),
                       log_path.encode() if log_path else None, deps_path.encode() if deps_path else None)
    L = np.tril(W)
    names = ["total", "potrf", "trsm", "gemm", "accu", "steals", "info"]
    stats = {k: float(v) for k, v in zip(names, st)}
    stats["info"] = int(st[6])
    return np.asfortranarray(L), stats


def priority(prio_kind: int, nb: int, accumulate: bool, task_kind: str, i: int, j: int, k: int) -> float:
    """Priority the restated PTG gives a task (2d_cholesky.cpp:157-204); task_kind 'POTRF', 'TRSM' or 'GEMM'
    (SYRK is GEMM(k, i, i))."""
    tk = {"POTRF": 0, "TRSM": 1, "GEMM": 2}[task_kind]
    return float(lib().orc_priority(prio_kind, nb, int(accumulate), tk, i, j, k))


def run_dag(n_threads: int, indeg, out_ptr, out_idx, prio=None):
    """Run an arbitrary DAG through the restated runtime; returns (order, run_count)."""
    indeg = np.ascontiguousarray(indeg, dtype=np.int32)
    out_ptr = np.ascontiguousarray(out_ptr, dtype=np.int32)
    out_idx = np.ascontiguousarray(out_idx, dtype=np.int32)
    n = len(indeg)
    order = np.full(n, -1, dtype=np.int32)
    runs = np.zeros(n, dtype=np.int32)
    pr = None if prio is None else np.ascontiguousarray(prio, dtype=np.float64)
    cnt = lib().orc_run_dag(n_threads, n, _pi(indeg), _pi(out_ptr), _pi(out_idx),
                            _p(pr) if pr is not None else None, _pi(order), _pi(runs))
    return order[:cnt], runs


# ----------------------------------------------------------------------------- inputs
def ref_val_matrix(n: int) -> np.ndarray:
    """The reference miniapp's generator A(i,j) = 1/((i-j)^2+1) (2d_cholesky.cpp:133)."""
    i = np.arange(n, dtype=np.float64)
    d = i[:, None] - i[None, :]
    return np.asfortranarray(1.0 / (d * d + 1.0))


def spd_bbt(n: int, seed: int, delta: float = 1.0) -> np.ndarray:
    """A = B B^T / n + delta I, B iid N(0,1) from numpy's PCG64 stream `seed`."""
    B = np.random.default_rng(seed).standard_normal((n, n))
    A = (B @ B.T) / n
    A[np.diag_indices(n)] += delta
    return np.asfortranarray(A)


def spd_scaled(n: int, seed_b: int, seed_d: int, delta: float = 1e-2, dmax: float = 1e2) -> np.ndarray:
    """A = D (B B^T/n + delta I) D with D_ii log-uniform in [1, dmax] (ill-conditioned case)."""
    M = spd_bbt(n, seed_b, delta)
    d = np.exp(np.random.default_rng(seed_d).uniform(0.0, np.log(dmax), n))
    return np.asfortranarray(M * d[:, None] * d[None, :])


def int_lower_factor(n: int, seed: int, offmax: int = 2, diag_exp=(0, 3)) -> np.ndarray:
    """Integer lower-triangular L with power-of-two diagonal and entries in [-offmax, offmax].

    For A = L L^T every step of a substitution-based Cholesky (LAPACK order) is exact
    in FP64 as long as the magnitudes stay below 2^53, so the computed factor must
    equal L bitwise. With the defaults L is extremely ill-conditioned (its inverse
    grows like 3^n); offmax=1, diag_exp=(3, 4) gives a well-conditioned variant.
    """
    rng = np.random.default_rng(seed)
    L = np.tril(rng.integers(-offmax, offmax + 1, size=(n, n)).astype(np.float64), -1)
    L[np.diag_indices(n)] = 2.0 ** rng.integers(diag_exp[0], diag_exp[1], size=n)
    return np.asfortranarray(L)


# ----------------------------------------------------------------------------- metrics
def sym_from_lower(A: np.ndarray) -> np.ndarray:
    return np.tril(A) + np.tril(A, -1).T


def factor_residual(A: np.ndarray, L: np.ndarray) -> float:
    """||L L^T - A||_F / ||A||_F using only the lower triangle of A."""
    S = sym_from_lower(A)
    Lt = np.tril(L)
    return float(np.linalg.norm(Lt @ Lt.T - S) / np.linalg.norm(S))


def backward_error(A: np.ndarray, x: np.ndarray, b: np.ndarray) -> float:
    """Normwise backward error max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    S = sym_from_lower(A)
    x2 = x.reshape(x.shape[0], -1)
    b2 = b.reshape(b.shape[0], -1)
    r = b2 - S @ x2
    an = np.abs(S).sum(axis=1).max()
    eta = np.abs(r).max(axis=0) / (an * np.abs(x2).max(axis=0) + np.abs(b2).max(axis=0))
    return float(eta.max())


def solve_check(A: np.ndarray, L: np.ndarray, seed: int = 0) -> float:
    """2d_cholesky.cpp:676-681: x random, b = A x, two triangular solves, ||x^ - x|| / ||x||."""
    import scipy.linalg as sla
    S = sym_from_lower(A)
    x = np.random.default_rng(seed).uniform(-1.0, 1.0, S.shape[0])
    b = S @ x
    y = sla.solve_triangular(np.tril(L), b, lower=True)
    xh = sla.solve_triangular(np.tril(L).T, y, lower=False)
    return float(np.linalg.norm(xh - x) / np.linalg.norm(x))


def flops_potrf(n: int) -> float:
    """LAPACK dpotrf flop count n^3/3 + n^2/2 + n/6."""
    return n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0
