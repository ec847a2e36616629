"""Writes the golden fixtures of the oracle tests (run from the repo root:
`python tests/golden/make_golden.py`).

Expected factors come from numpy.linalg.cholesky, i.e. the LAPACK dpotrf of the
OpenBLAS build that ships with numpy -- an implementation independent of both
the oracle's textbook loops and MKL. Exact integer cases are computed in int64.
"""
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def spd_bbt(n, seed, delta=1.0):
    B = np.random.default_rng(seed).standard_normal((n, n))
    A = B @ B.T / n
    A[np.diag_indices(n)] += delta
    return A


def main():
    cases = {
        "chol_bbt_192": spd_bbt(192, 11),
        "chol_bbt_100": spd_bbt(100, 12),
        "chol_refval_200": 1.0 / ((np.arange(200)[:, None] - np.arange(200)[None, :]) ** 2 + 1.0),
    }
    for name, A in cases.items():
        L = np.linalg.cholesky(A)
        np.savez_compressed(HERE / f"{name}.npz", A=A, L=L)
    rng = np.random.default_rng(2020)
    A = rng.integers(-4, 5, (40, 30))
    B = rng.integers(-4, 5, (50, 30))
    C = rng.integers(-4, 5, (40, 50))
    np.savez_compressed(HERE / "gemm_int.npz", A=A, B=B, C=C, C_minus_ABt=C - A @ B.T)


if __name__ == "__main__":
    main()
