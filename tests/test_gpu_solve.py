"""Solve-phase parity on the GPU.

- ttad_solve with the FP64 factor against the oracle's two triangular solves
  (2d_cholesky.cpp:676-681), relative 1e-12 on well-conditioned A;
- ttad_solve with an FP32 factor (the tile-chained lookahead sweeps) against the host's
  triangular solves with the same factor, and bitwise repeatable;
- ttad_residual against b - A x computed on the host in FP64 (1e-13 relative) and
  the normwise backward error definition;
- mixed precision + iterative refinement reaches eta <= 1e-12 (the bar of the
  mixed solver) on the ill-conditioned A = D (B B^T/n + delta I) D family, and
  matches numpy's FP64 solve to the accuracy its condition number allows;
- the FP64 fallback triggers on a matrix whose low-precision factor breaks down.
"""
import numpy as np
import pytest

import tasktorrent_amd as tt
from gpu_util import from_dev, to_dev

pytestmark = pytest.mark.gpu


def _dev_cols(X):
    """n x k column-major host array -> device tensor (k, n) (same layout)."""
    return to_dev(np.asfortranarray(X))


def _mat(ctx, A, b, storage="f64"):
    n = A.shape[0]
    sizes = [b] * (n // b) + ([n % b] if n % b else [])
    M = tt.TiledMatrix(ctx, sizes, storage)
    M.set(A)
    return M


def backward_error(A, X, B):
    R = B - A @ X
    an = np.abs(A).sum(axis=1).max()
    return float((np.abs(R).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))).max())


@pytest.mark.parametrize("n,b,nrhs", [(1024, 256, 1), (1536, 512, 16), (640, 128, 20), (704, 192, 3)])
def test_solve_fp64_factor(ctx, oracle, n, b, nrhs):
    A = oracle.spd_bbt(n, n + nrhs)
    M = _mat(ctx, A, b)
    assert ctx.potrf_tiled(M, tt.TTAD_SCHED_PERSISTENT, 2, "f64") == 0
    B = np.asfortranarray(np.random.default_rng(nrhs).standard_normal((n, nrhs)))
    dB, dX = _dev_cols(B), _dev_cols(np.zeros((n, nrhs)))
    ctx.solve(M, nrhs, dB, dX)
    X = from_dev(dX)
    Xref = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xref) / np.linalg.norm(Xref) < 1e-12
    assert backward_error(A, X, B) < 1e-15


@pytest.mark.parametrize("sizes,nrhs", [([512] * 6, 16), ([512, 320, 448, 192, 384, 64, 100], 20),
                                        ([128] * 9, 3), ([512] * 3, 1)])
def test_lookahead_sweeps_fp32_factor(ctx, oracle, sizes, nrhs):
    """ttad_solve with an FP32 factor whose tiles have at most 512 rows (the tile-chained lookahead
    sweeps) against the two triangular solves on the host with the same factor in FP64. The sweeps
    multiply by explicit inverses of the diagonal tiles in FP32, so on A = B B^T/n + I (condition
    number below 5) the solution agrees to a small multiple of FP32's unit roundoff times the tile
    size. Tiles of different sizes (multiples of 64 but the last, as the solve requires) exercise
    short far tiles and a partial last block; nrhs = 20 takes two 16-column passes; three tiles
    leave a single far tile. Two calls give the same bits."""
    from scipy.linalg import solve_triangular
    n = sum(sizes)
    A = oracle.spd_bbt(n, n + nrhs)
    W = tt.TiledMatrix(ctx, sizes, "f32")
    W.set(A)
    assert ctx.potrf_tiled(W, tt.TTAD_SCHED_PERSISTENT, 2, "f32") == 0
    L = W.get()
    B = np.asfortranarray(np.random.default_rng(nrhs + n).standard_normal((n, nrhs)))
    dB = _dev_cols(B)
    dX1, dX2 = _dev_cols(np.zeros((n, nrhs))), _dev_cols(np.zeros((n, nrhs)))
    ctx.solve(W, nrhs, dB, dX1)
    ctx.solve(W, nrhs, dB, dX2)
    X1, X2 = from_dev(dX1), from_dev(dX2)
    W.close()
    Xref = solve_triangular(L.T, solve_triangular(L, B, lower=True), lower=False)
    assert np.array_equal(X1, X2)
    err = np.linalg.norm(X1 - Xref) / np.linalg.norm(Xref)
    assert err < 2e-5, err


@pytest.mark.parametrize("nrhs", [1, 16, 40])
def test_residual(ctx, oracle, nrhs):
    n, b = 1024, 256
    A = oracle.spd_scaled(n, 3, 4)
    M = _mat(ctx, A, b)
    rng = np.random.default_rng(nrhs)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    X = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dB, dX, dR = _dev_cols(B), _dev_cols(X), _dev_cols(np.zeros((n, nrhs)))
    eta = ctx.residual(M, nrhs, dB, dX, dR)
    R = from_dev(dR)
    Rref = B - A @ X
    assert np.linalg.norm(R - Rref) / np.linalg.norm(Rref) < 1e-13
    assert abs(eta - backward_error(A, X, B)) <= 1e-12 * backward_error(A, X, B)


@pytest.mark.parametrize("sizes,nrhs", [([512] * 4, 16), ([256] * 6, 1), ([320, 448, 232], 20),
                                        ([512, 512, 100], 7)])
def test_residual_tile_pairs(ctx, oracle, sizes, nrhs):
    """The tile-pair residual (each stored element of A read once) against the host product and
    the row-block kernels (TTAD_OPT_RESIDUAL_ROWS). The diagonal tiles' upper triangles hold
    garbage: both paths must use the stored lower triangle only. Ragged tiles exercise partial
    64-row blocks; nrhs = 20 takes two 16-column passes. Two calls give the same bits (the
    partials are added in a fixed order)."""
    n = sum(sizes)
    A = oracle.spd_scaled(n, 3, 4)
    poisoned = A.copy(order="F")
    poisoned[np.triu_indices(n, 1)] = 1e30
    M = tt.TiledMatrix(ctx, sizes, "f64")
    M.set(poisoned)
    rng = np.random.default_rng(nrhs + n)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    X = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dB, dX = _dev_cols(B), _dev_cols(X)
    dR1, dR1b, dR2 = (_dev_cols(np.zeros((n, nrhs))) for _ in range(3))
    eta1 = ctx.residual(M, nrhs, dB, dX, dR1)
    eta1b = ctx.residual(M, nrhs, dB, dX, dR1b)
    ctx.set_option(tt._capi.TTAD_OPT_RESIDUAL_ROWS, 1)
    try:
        eta2 = ctx.residual(M, nrhs, dB, dX, dR2)
    finally:
        ctx.set_option(tt._capi.TTAD_OPT_RESIDUAL_ROWS, 0)
    M.close()
    R1, R1b, R2 = from_dev(dR1), from_dev(dR1b), from_dev(dR2)
    Rref = B - A @ X
    assert np.linalg.norm(R1 - Rref) / np.linalg.norm(Rref) < 1e-13
    assert np.linalg.norm(R2 - Rref) / np.linalg.norm(Rref) < 1e-13
    assert np.array_equal(R1, R1b) and eta1 == eta1b
    ref_eta = backward_error(A, X, B)
    assert abs(eta1 - ref_eta) <= 1e-12 * ref_eta and abs(eta2 - ref_eta) <= 1e-12 * ref_eta


@pytest.mark.parametrize("prec", ["f32", "bf16x3", "bf16"])
def test_mixed_ir_reaches_1e12(ctx, oracle, prec):
    n, b, nrhs = 2048, 256, 16
    A = oracle.spd_scaled(n, 2022, 2023)  # kappa ~ 1e4 - 1e6
    M = _mat(ctx, A, b)
    B = np.asfortranarray(np.random.default_rng(2024).standard_normal((n, nrhs)))
    dB, dX = _dev_cols(B), _dev_cols(np.zeros((n, nrhs)))
    st = ctx.posv_mixed(M, prec, nrhs, dB, dX, tol=1e-12, max_iter=30)
    X = from_dev(dX)
    eta = backward_error(A, X, B)
    assert st["rc"] == 0, st
    assert st["backward_error"] <= 1e-12 and eta <= 2e-12, (st, eta)
    Xref = np.linalg.solve(A, B)
    kappa = np.linalg.cond(A)
    # forward error of a solution with backward error eta is bounded by about kappa * eta
    fwd = np.linalg.norm(X - Xref) / np.linalg.norm(Xref)
    assert fwd < 10 * kappa * max(st["backward_error"], 1e-16) + 1e-13, (fwd, kappa, st)


@pytest.mark.parametrize("prec", ["bf16x3", "f32", "bf16"])
def test_mixed_tile512_16rhs(ctx, oracle, prec):
    # the headline's shape at a size the host checks quickly: tile 512 (FP32 tiles of 4 x 4 bf16x3 / FP32
    # items, 8 row-block POTRF items in FP64 arithmetic), 16 right-hand sides, A = D (B B^T/n + 1e-2 I) D
    n, b, nrhs = 4096, 512, 16
    A = oracle.spd_scaled(n, 2022, 2023)
    M = _mat(ctx, A, b)
    B = np.asfortranarray(np.random.default_rng(2024).standard_normal((n, nrhs)))
    dB, dX = _dev_cols(B), _dev_cols(np.zeros((n, nrhs)))
    st = ctx.posv_mixed(M, prec, nrhs, dB, dX, tol=1e-12, max_iter=30)
    assert st["rc"] == 0 and st["fell_back"] == 0, st
    X = from_dev(dX)
    assert st["backward_error"] <= 1e-12 and backward_error(A, X, B) <= 2e-12, st


@pytest.mark.parametrize("prec", ["bf16x3", "bf16"])
def test_mixed_refined_to_floor_matches_lapack(ctx, oracle, prec):
    # refined until the backward error stops halving (tol below what FP64 reaches), the mixed solution
    # agrees with LAPACK's FP64 solve to what the condition number allows: both are within ~kappa u of x
    n, b, nrhs = 2048, 256, 4
    A = oracle.spd_scaled(n, 2022, 2023)
    M = _mat(ctx, A, b)
    B = np.asfortranarray(np.random.default_rng(2024).standard_normal((n, nrhs)))
    dB, dX = _dev_cols(B), _dev_cols(np.zeros((n, nrhs)))
    W = tt.TiledMatrix(ctx, M.block_sizes, "f32")
    W.copy_from(M)
    assert ctx.potrf_tiled(W, tt.TTAD_SCHED_PERSISTENT, 4, prec) == 0
    st = ctx.solve_ir(M, W, nrhs, dB, dX, tol=1e-30, max_iter=40)
    assert st["rc"] == tt.TTAD_ERR_IR_STALLED and st["iterations"] < 40, st  # stopped by the stall rule
    X = from_dev(dX)
    assert backward_error(A, X, B) <= 1e-15, st
    Xref = np.linalg.solve(A, B)
    kappa = np.linalg.cond(A)
    fwd = np.linalg.norm(X - Xref) / np.linalg.norm(Xref)
    assert fwd <= 10 * kappa * np.finfo(np.float64).eps, (fwd, kappa)


def test_fallback_on_breakdown(ctx, oracle):
    # kappa ~ 1e12 well beyond what an FP32 factor refines: expect the FP64 fallback
    n, b, nrhs = 1024, 256, 4
    A = oracle.spd_scaled(n, 5, 6, delta=1e-9, dmax=1e3)
    M = _mat(ctx, A, b)
    B = np.asfortranarray(np.random.default_rng(1).standard_normal((n, nrhs)))
    dB, dX = _dev_cols(B), _dev_cols(np.zeros((n, nrhs)))
    st = ctx.posv_mixed(M, "bf16", nrhs, dB, dX, tol=1e-12, max_iter=10)
    assert st["fell_back"] == 1
    assert st["prec_used"] == tt.TTAD_F64
    # the FP64 factor of a kappa ~ 1e12 matrix is backward stable: refinement must reach the target
    assert st["rc"] == 0, st
    assert backward_error(A, from_dev(dX), B) <= 2e-12


def test_not_spd_input_falls_back_then_fails(ctx, oracle):
    # an indefinite symmetric matrix (eigenvalues of B B^T/n + I - 3 I in [-2, 2]): the low-precision
    # factor breaks down, the FP64 one too; the caller gets NOT_SPD with the pivot column
    n, b, nrhs = 512, 256, 2
    A = oracle.spd_bbt(n, 9)
    A[np.diag_indices(n)] -= 3.0
    M = _mat(ctx, A, b)
    B = np.asfortranarray(np.random.default_rng(2).standard_normal((n, nrhs)))
    dB, dX = _dev_cols(B), _dev_cols(np.zeros((n, nrhs)))
    st = ctx.posv_mixed(M, "bf16x3", nrhs, dB, dX, tol=1e-12, max_iter=10)
    assert st["rc"] == tt.TTAD_ERR_NOT_SPD, st
    assert st["fell_back"] == 1 and st["prec_used"] == tt.TTAD_F64
    assert 1 <= st["info"] <= n
    # LAPACK agrees that the leading minor of that order is not positive definite
    L = np.asfortranarray(A.copy())
    assert oracle.dpotrf(L) == st["info"]


def test_factor_reused_for_new_rhs(ctx, oracle):
    n, b = 1024, 256
    A = oracle.spd_bbt(n, 8)
    M = _mat(ctx, A, b)
    W = tt.TiledMatrix(ctx, M.block_sizes, "f32")
    W.copy_from(M)
    assert ctx.potrf_tiled(W, tt.TTAD_SCHED_PERSISTENT, 2, "f32") == 0
    for seed in range(3):
        B = np.asfortranarray(np.random.default_rng(seed).standard_normal((n, 2)))
        dB, dX = _dev_cols(B), _dev_cols(np.zeros((n, 2)))
        st = ctx.solve_ir(M, W, 2, dB, dX, tol=1e-13, max_iter=20)
        assert st["rc"] == 0 and backward_error(A, from_dev(dX), B) <= 2e-13


def test_solve_ir_zero_rhs_and_padded_ld(ctx, oracle):
    # the refinement starts at X = 0, where the residual is B itself: a zero B needs no step,
    # a zero column stays exactly zero, and padding beyond n rows (ld > n) is never read
    import torch
    n, b = 768, 256
    A = oracle.spd_bbt(n, 3)
    M = _mat(ctx, A, b)
    W = tt.TiledMatrix(ctx, M.block_sizes, "f64")
    W.copy_from(M)
    assert ctx.potrf_tiled(W, tt.TTAD_SCHED_PERSISTENT, 2, "f64") == 0

    dZ, dX0 = _dev_cols(np.zeros((n, 2))), _dev_cols(np.ones((n, 2)))
    st = ctx.solve_ir(M, W, 2, dZ, dX0, tol=1e-13, max_iter=20)
    assert st["rc"] == 0 and st["iterations"] == 0 and st["backward_error"] == 0.0
    assert not from_dev(dX0).any()

    B = np.random.default_rng(4).standard_normal((n, 3))
    B[:, 0] = 0.0
    dB = torch.full((3, n + 7), float("nan"), dtype=torch.float64, device="cuda")
    dB[:, :n] = torch.from_numpy(B.T.copy()).cuda()
    dX = torch.zeros((3, n + 5), dtype=torch.float64, device="cuda")
    st = ctx.solve_ir(M, W, 3, dB, dX, tol=1e-13, max_iter=20)
    X = dX[:, :n].cpu().numpy().T
    assert st["rc"] == 0 and st["iterations"] >= 1
    assert not X[:, 0].any()
    assert backward_error(A, X[:, 1:], B[:, 1:]) <= 2e-13


@pytest.mark.parametrize("src,dst", [("f64", "f32"), ("f64", "f64"), ("f32", "f64"), ("f32", "f32")])
def test_matrix_copy_ragged_tiles(ctx, src, dst):
    # tiles of 130, 67 and 3 rows: element counts divisible by four (copied in quads) and not (one by one);
    # every storage pair (one kernel each); FP64 -> FP32 rounds to nearest like numpy's astype, FP32 -> FP64
    # is exact
    rng = np.random.default_rng(3)
    sizes = [130, 67, 3]
    n = sum(sizes)
    A = rng.standard_normal((n, n))
    M = tt.TiledMatrix(ctx, sizes, src)
    M.set(A)
    W = tt.TiledMatrix(ctx, sizes, dst)
    W.copy_from(M)
    exp = np.tril(A)
    if "f32" in (src, dst):
        exp = exp.astype(np.float32).astype(np.float64)
    assert np.array_equal(W.get(), exp)
