"""Tile-kernel parity on the GPU: each C-ABI tile entry point against the oracle's
restatement of the BLAS/LAPACK call it replaces (2d_cholesky.cpp:319, 404-405,
490-491, 493-494), plus the exact integer GEMM known-answer test of
tests/shared/tests.cpp:361-411 (entries in [-4,4], result must be exact).

Tolerances: FP64 kernels 1e-13 relative (Frobenius) against the oracle; FP32
kernels 2e-6; BF16-input kernels are pinned exactly on integer data and to
1e-2 on random data (bf16 has an 8-bit significand).
"""
import numpy as np
import pytest

from gpu_util import from_dev, rel, to_dev

pytestmark = pytest.mark.gpu

SHAPES = [(128, 128, 128), (512, 512, 512), (100, 130, 70), (257, 129, 300), (64, 448, 17), (1, 5, 3)]


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_dgemm_integer_exact(ctx, m, n, k):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = np.asfortranarray(rng.integers(-4, 5, (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-4, 5, (n, k)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-4, 5, (m, n)).astype(np.float64))
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    ctx.gemm(m, n, k, -1.0, dA, m, dB, n, 1.0, dC, m)
    ctx.synchronize()
    exact = C - A @ B.T
    assert np.array_equal(from_dev(dC), exact)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_dgemm_random(ctx, oracle, m, n, k):
    rng = np.random.default_rng(1 + m + n + k)
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    ctx.gemm(m, n, k, -1.0, dA, m, dB, n, 1.0, dC, m)
    ref = C.copy(order="F")
    oracle.dgemm_nt(-1.0, A, B, 1.0, ref)
    assert rel(from_dev(dC), ref) < 1e-13


def test_dgemm_beta0_alpha(ctx, oracle):
    rng = np.random.default_rng(5)
    m, n, k = 200, 150, 90
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.asfortranarray(np.full((m, n), np.nan))
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    ctx.gemm(m, n, k, 0.5, dA, m, dB, n, 0.0, dC, m)  # beta = 0 must not read C (NaN)
    ref = np.zeros((m, n), order="F")
    oracle.dgemm_nt(0.5, A, B, 0.0, ref)
    assert rel(from_dev(dC), ref) < 1e-13


@pytest.mark.parametrize("alpha,beta", [(0.5, -2.0), (-1.0, 0.25), (1.0, 1.0)])
def test_dgemm_general_alpha_beta(ctx, oracle, alpha, beta):
    rng = np.random.default_rng(17)
    m, n, k = 257, 190, 140
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    ctx.gemm(m, n, k, alpha, dA, m, dB, n, beta, dC, m)
    ref = C.copy(order="F")
    oracle.dgemm_nt(alpha, A, B, beta, ref)
    assert rel(from_dev(dC), ref) < 1e-13


@pytest.mark.parametrize("n,k", [(128, 64), (512, 512), (300, 77), (1, 1)])
def test_dsyrk_lower_only(ctx, oracle, n, k):
    rng = np.random.default_rng(n + k)
    A = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.asfortranarray(rng.standard_normal((n, n)))
    dA, dC = to_dev(A), to_dev(C)
    ctx.syrk(n, k, -1.0, dA, n, 1.0, dC, n)
    out = from_dev(dC)
    ref = C.copy(order="F")
    oracle.dsyrk_ln(-1.0, A, 1.0, ref)
    assert rel(np.tril(out), np.tril(ref)) < 1e-13
    assert np.array_equal(np.triu(out, 1), np.triu(C, 1))  # strict upper untouched (dsyrk 'L')


@pytest.mark.parametrize("m,n", [(512, 512), (64, 64), (100, 200), (513, 130), (7, 1)])
def test_dtrsm(ctx, oracle, m, n):
    rng = np.random.default_rng(m * n)
    S = rng.standard_normal((n, n))
    L = np.asfortranarray(np.linalg.cholesky(S @ S.T / n + np.eye(n)))
    B = np.asfortranarray(rng.standard_normal((m, n)))
    dL, dB = to_dev(L), to_dev(B)
    ctx.trsm(m, n, dL, n, dB, m)
    ref = B.copy(order="F")
    oracle.dtrsm_rltn(L, ref)
    assert rel(from_dev(dB), ref) < 1e-13


@pytest.mark.parametrize("n", [512, 256, 64, 100, 333, 1])
def test_dpotrf(ctx, oracle, n):
    A = oracle.spd_bbt(n, 10 + n)
    dA = to_dev(A)
    info = ctx.potrf(dA, n, n)
    assert info == 0
    out = from_dev(dA)
    ref = A.copy(order="F")
    assert oracle.dpotrf(ref) == 0
    assert rel(np.tril(out), np.tril(ref)) < 1e-13
    assert np.array_equal(np.triu(out, 1), np.triu(A, 1))  # upper triangle not referenced


def test_dpotrf_not_spd(ctx, oracle):
    n = 256
    A = oracle.spd_bbt(n, 3)
    A[150, 150] = -1.0
    ref = A.copy(order="F")
    info_ref = oracle.dpotrf(ref)
    info = ctx.potrf(to_dev(A), n, n)
    assert info_ref > 0 and info == info_ref


def test_dpotrf_integer_factor(ctx, oracle):
    # A = L L^T with integer L (diagonal 8, off-diagonal in {-1,0,1}): the unique factor is L.
    # The GPU kernel applies inverses of 64x64 diagonal blocks, so it is not exact like a
    # substitution; its error scales with the condition of those blocks (well-conditioned here).
    n = 192
    L = oracle.int_lower_factor(n, 11, offmax=1, diag_exp=(3, 4))
    A = np.asfortranarray(L @ L.T)
    dA = to_dev(A)
    assert ctx.potrf(dA, n, n) == 0
    assert rel(np.tril(from_dev(dA)), L) < 1e-13


@pytest.mark.parametrize("m,n,k", [(512, 512, 512), (130, 70, 33)])
def test_sgemm(ctx, m, n, k):
    import torch
    rng = np.random.default_rng(m + 2 * n)
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    dA, dB, dC = to_dev(A, torch.float32), to_dev(B, torch.float32), to_dev(C, torch.float32)
    ctx.gemm(m, n, k, -1.0, dA, m, dB, n, 1.0, dC, m)
    A32, B32, C32 = (x.astype(np.float32).astype(np.float64) for x in (A, B, C))
    ref = C32 - A32 @ B32.T
    assert rel(from_dev(dC), ref) < 2e-6


@pytest.mark.parametrize("m,n,k", [(256, 256, 512), (100, 60, 40)])
def test_gemm_bf16_integer_exact(ctx, m, n, k):
    # small integers are exact in bf16 and FP32 accumulation of |sum| < 2^24 is exact
    import torch
    rng = np.random.default_rng(m + n + k)
    A = np.asfortranarray(rng.integers(-4, 5, (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-4, 5, (n, k)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-4, 5, (m, n)).astype(np.float64))
    dA, dB, dC = to_dev(A, torch.float32), to_dev(B, torch.float32), to_dev(C, torch.float32)
    ctx.gemm_bf16(m, n, k, -1.0, dA, m, dB, n, 1.0, dC, m)
    assert np.array_equal(from_dev(dC), C - A @ B.T)


def test_gemm_bf16_random(ctx):
    import torch
    m = n = k = 512
    rng = np.random.default_rng(9)
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k)))
    C = np.zeros((m, n), order="F")
    dA, dB, dC = to_dev(A, torch.float32), to_dev(B, torch.float32), to_dev(C, torch.float32)
    ctx.gemm_bf16(m, n, k, 1.0, dA, m, dB, n, 0.0, dC, m)
    ref = A @ B.T
    assert rel(from_dev(dC), ref) < 1e-2


def test_gemm_bf16x3_accuracy(ctx):
    # the split product keeps ~16 bits per operand: error far below plain bf16, near FP32
    import torch
    m = n = k = 512
    rng = np.random.default_rng(10)
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k)))
    A32, B32 = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    ref = A32 @ B32.T
    dA, dB = to_dev(A, torch.float32), to_dev(B, torch.float32)
    dC = to_dev(np.zeros((m, n), order="F"), torch.float32)
    ctx.gemm_bf16(m, n, k, 1.0, dA, m, dB, n, 0.0, dC, m, split=True)
    e3 = rel(from_dev(dC), ref)
    ctx.gemm_bf16(m, n, k, 1.0, dA, m, dB, n, 0.0, dC, m, split=False)
    e1 = rel(from_dev(dC), ref)
    assert e3 < 1e-4 and e3 < e1 / 50, (e3, e1)


def test_spotrf_strsm_ssyrk(ctx, oracle):
    import torch
    n = 512
    A = oracle.spd_bbt(n, 77)
    dA = to_dev(A, torch.float32)
    assert ctx.potrf(dA, n, n) == 0
    L = np.tril(from_dev(dA))
    ref = np.linalg.cholesky(A.astype(np.float32).astype(np.float64))
    assert rel(L, ref) < 1e-6
    B = np.asfortranarray(np.random.default_rng(1).standard_normal((300, n)))
    dB = to_dev(B, torch.float32)
    dL = to_dev(L, torch.float32)
    ctx.trsm(300, n, dL, n, dB, 300)
    X = from_dev(dB)
    assert rel(X @ L.T, B) < 1e-5
    A2 = np.asfortranarray(np.random.default_rng(3).standard_normal((n, 300)))
    C = np.asfortranarray(np.random.default_rng(2).standard_normal((n, n)))
    dC = to_dev(C, torch.float32)
    ctx.syrk(n, 300, -1.0, to_dev(A2, torch.float32), n, 1.0, dC, n)
    A2f, Cf = A2.astype(np.float32).astype(np.float64), C.astype(np.float32).astype(np.float64)
    assert rel(np.tril(from_dev(dC)), np.tril(Cf - A2f @ A2f.T)) < 2e-6
