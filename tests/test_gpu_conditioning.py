"""FP64 factor and solves at condition numbers 1e8 and 1e10.

POTRF and TRSM multiply by explicit inverses of 64 x 64 diagonal blocks where the
reference calls LAPACK's and BLAS's substitutions (LAPACKE_dpotrf and cblas_dtrsm,
miniapp/dense_cholesky/2d_cholesky.cpp:319, 404-405). These cases bound what that
costs. A = B B^T / n + delta I with B n x n/2 (rank n/2) and
delta = lambda_max(B B^T / n) / (kappa - 1), so cond_2(A) = kappa exactly; the
Schur complements around the rank boundary have ill-conditioned diagonal blocks.

Bars, fixed from the literature rather than from measurements:
- factor residual ||L L^T - A||_F / ||A||_F <= 30 n eps, the acceptance ratio of
  LAPACK's own Cholesky tests (dpot01: ||L L^T - A|| / (n ||A|| eps) < 30), for
  the GPU factor and for the oracle's;
- the GPU factor and the oracle's, both backward stable, differ by at most
  2 kappa n eps relative (first-order perturbation bound of the Cholesky factor,
  ||dL||_F / ||L||_2 <= kappa ||dA||_F / ||A||_2, Sun 1991);
- FP64 refinement with the FP64 factor (ttad_solve_ir) reaches the normwise
  backward error 1e-12 of the mixed solver's bar, checked on the host;
- ttad_posv_mixed with bf16x3 returns rc 0 and backward error <= 1e-12, through
  its FP64 fallback where the low-precision factor fails or refinement stalls.
"""
import numpy as np
import pytest

import tasktorrent_amd as tt
from gpu_util import from_dev, to_dev

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _ill_conditioned(n, kappa, seed):
    B = np.random.default_rng(seed).standard_normal((n, n // 2))
    lam_max = float(np.linalg.eigvalsh(B.T @ B / n)[-1])  # nonzero spectrum of B B^T / n
    A = B @ B.T / n
    A[np.diag_indices(n)] += lam_max / (kappa - 1.0)
    return np.asfortranarray(A)


@pytest.mark.parametrize("log_kappa", [8, 10])
def test_fp64_factor_envelope(ctx, oracle, log_kappa):
    n, b, nrhs = 2048, 512, 4
    kappa = 10.0 ** log_kappa
    A = _ill_conditioned(n, kappa, 100 + log_kappa)
    sizes = [b] * (n // b)
    MA = tt.TiledMatrix(ctx, sizes, "f64")
    ML = tt.TiledMatrix(ctx, sizes, "f64")
    try:
        MA.set(A)
        ML.copy_from(MA)
        assert ctx.potrf_tiled(ML, tt.TTAD_SCHED_PERSISTENT, 2, "f64") == 0
        L = np.tril(ML.get())
        fres_dev = ctx.factor_residual(MA, ML)
        Lref, st = oracle.cholesky(A, b, n_threads=4)
        assert st["info"] == 0
        fres = oracle.factor_residual(A, L)
        fres_ref = oracle.factor_residual(A, Lref)
        dL = float(np.linalg.norm(L - Lref) / np.linalg.norm(Lref, 2))
        rhs = np.asfortranarray(np.random.default_rng(log_kappa).standard_normal((n, nrhs)))
        dB, dX = to_dev(rhs), to_dev(np.zeros((n, nrhs)))
        st64 = ctx.solve_ir(MA, ML, nrhs, dB, dX, tol=1e-12, max_iter=10)
        eta64 = oracle.backward_error(A, from_dev(dX), rhs)
        dXm = to_dev(np.zeros((n, nrhs)))
        stm = ctx.posv_mixed(MA, "bf16x3", nrhs, dB, dXm, tol=1e-12, max_iter=10)
        etam = oracle.backward_error(A, from_dev(dXm), rhs)
    finally:
        ML.close()
        MA.close()
    print(f"kappa 1e{log_kappa}: factor residual GPU {fres:.3e} (device {fres_dev:.3e}) oracle {fres_ref:.3e}; "
          f"||L - Lref||_F/||Lref||_2 {dL:.3e}; FP64 IR eta {eta64:.3e} in {st64['iterations']} steps; "
          f"bf16x3 eta {etam:.3e} it {stm['iterations']} fell_back {stm['fell_back']}")
    # (the FP64 evaluation of a residual this small carries rounding of its own size: host and device values
    # are each held to the bar, not to each other)
    assert fres <= 30 * n * EPS and fres_dev <= 30 * n * EPS and fres_ref <= 30 * n * EPS
    assert dL <= 2 * kappa * n * EPS
    assert st64["rc"] == 0 and st64["backward_error"] <= 1e-12 and eta64 <= 1e-12, st64
    assert stm["rc"] == 0 and stm["backward_error"] <= 1e-12 and etam <= 1e-12, stm
