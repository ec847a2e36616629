"""ttad_dtrsm_noalloc against ttad_dtrsm through the C-ABI.

The no-allocation variant runs the same kernels with the diagonal-block
inverses in a static per-GPU scratch, so its results must be bitwise equal to
ttad_dtrsm's. Two calls in a row, the second with a smaller input (the scratch
keeps the first call's contents beyond what the second writes), then the size
limit. Both results are also checked against the oracle's cblas_dtrsm
restatement (2d_cholesky.cpp:404-405) to 1e-13 relative.
"""
import numpy as np
import pytest

from gpu_util import from_dev, rel, to_dev

pytestmark = pytest.mark.gpu


def _case(ctx, oracle, m, n, seed):
    rng = np.random.default_rng(seed)
    S = rng.standard_normal((n, n))
    L = np.asfortranarray(np.linalg.cholesky(S @ S.T / n + np.eye(n)))
    B = np.asfortranarray(rng.standard_normal((m, n)))
    dL = to_dev(L)
    dB_alloc, dB_noalloc = to_dev(B), to_dev(B)
    ctx.trsm(m, n, dL, n, dB_alloc, m)
    ctx.trsm_noalloc(m, n, dL, n, dB_noalloc, m)
    ctx.synchronize()
    out_alloc, out_noalloc = from_dev(dB_alloc), from_dev(dB_noalloc)
    assert np.array_equal(out_noalloc, out_alloc), f"m={m} n={n}"
    ref = B.copy(order="F")
    oracle.dtrsm_rltn(L, ref)
    assert rel(out_noalloc, ref) < 1e-13


def test_dtrsm_noalloc_matches_dtrsm(ctx, oracle):
    _case(ctx, oracle, 700, 4096, 11)  # n at the scratch limit
    _case(ctx, oracle, 300, 1000, 12)  # smaller, ragged n: stale scratch beyond 16 blocks


def test_dtrsm_noalloc_limit(ctx):
    import tasktorrent_amd as tt
    import torch
    n = 4160
    dL = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    dB = torch.zeros((n, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(tt.TtadError, match="scratch limit"):
        ctx.trsm_noalloc(8, n, dL, n, dB, 8)
