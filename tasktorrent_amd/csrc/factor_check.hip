// Factor residual ||L L^T - A||_F / ||A||_F on the device (the post-factorization check of
// miniapp/dense_cholesky/2d_cholesky.cpp:646-689 reports the solve error; this reports the factor
// itself, for FP64 and for the low-precision factors of the mixed solver).
//
// One workgroup per 128 x 128 sub-block of a lower tile (i, j): it forms sum_{k <= j} L_ik L_jk^T on
// FP64 MFMA (FP32 factors are widened on their way into LDS), subtracts A's sub-block and writes the
// weighted sums of squares of the difference and of A (off-diagonal elements count twice, as in the
// full symmetric matrix). A second kernel adds the partial sums in a fixed order, so the result does
// not depend on scheduling. L_jj enters through a copy of the diagonal tiles with the strict upper
// triangle zeroed (the stored tile's upper part is not part of L).
#include <algorithm>
#include <cmath>
#include <vector>

#include "runtime.hpp"

namespace ttad {

struct FresItem {
    int32_t i, j, r0, c0;
};

// Copy of diagonal tile j (bsz[j]^2 elements, ld = bsz[j]) with the strict upper triangle zeroed.
template <typename TS>
__global__ void k_lower_copy(const DevMat* M, TS* const* dst) {
    const int j = blockIdx.y;
    const int b = M->bsz[j];
    const TS* src = (const TS*)M->tiles[j + (long)j * M->nt];
    TS* d = dst[j];
    const long tot = (long)b * b;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e % b), c = (int)(e / b);
        d[e] = r >= c ? src[e] : TS(0);
    }
}

template <typename TS>
__global__ __launch_bounds__(NTHREADS, 2) void k_factor_residual(const DevMat* Lm, const DevMat* Am,
                                                                  TS* const* ldiag, const FresItem* items,
                                                                  double* part) {
    extern __shared__ __align__(16) char lds[];
    using G = GemmNT<TS, TS, double, GEMM_BM, GEMM_BM, 16, false, true>;
    const FresItem it = items[blockIdx.x];
    const int i = it.i, j = it.j, r0 = it.r0, c0 = it.c0;
    const int nt = Lm->nt;
    const int m = Lm->bsz[i], n = Lm->bsz[j];
    const int mv = min(GEMM_BM, m - r0), nv = min(GEMM_BM, n - c0);
    typename G::acc_t acc[G::WM][G::WN];
    G::zero(acc);
    for (int k = 0; k <= j; k++) {
        const int K = Lm->bsz[k];
        const TS* Ak = (k == j && i == j) ? ldiag[j] : (const TS*)Lm->tiles[i + (long)k * nt];
        const TS* Bk = (k == j) ? ldiag[j] : (const TS*)Lm->tiles[j + (long)k * nt];
        G::run(acc, Ak + r0, m, mv, Bk + c0, n, nv, K, lds);
    }
    const double* Aij = (const double*)Am->tiles[i + (long)j * nt];
    double se = 0.0, sa = 0.0;
#pragma unroll
    for (int a = 0; a < G::WM; a++)
#pragma unroll
        for (int b = 0; b < G::WN; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = G::row_of(a, r), cc = G::col_of(b, r);
                const int gr = r0 + rr, gc = c0 + cc;
                if (rr < mv && cc < nv && (i != j || gr >= gc)) {
                    const double w = (i != j || gr != gc) ? 2.0 : 1.0;
                    const double av = Aij[gr + (long)gc * m];
                    const double e = acc[a][b][r] - av;
                    se += w * e * e;
                    sa += w * av * av;
                }
            }
    // workgroup sums in a fixed order: wave shuffles, then the four waves in turn
    for (int off = 32; off > 0; off >>= 1) {
        se += __shfl_down(se, off);
        sa += __shfl_down(sa, off);
    }
    __shared__ double s_red[2][NTHREADS / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_red[0][w] = se;
        s_red[1][w] = sa;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double te = 0.0, ta = 0.0;
        for (int q = 0; q < NTHREADS / 64; q++) {
            te += s_red[0][q];
            ta += s_red[1][q];
        }
        part[2 * (long)blockIdx.x] = te;
        part[2 * (long)blockIdx.x + 1] = ta;
    }
}

// out[0] = sum of part[2q], out[1] = sum of part[2q+1], in a fixed order (one workgroup).
__global__ __launch_bounds__(NTHREADS) void k_fres_reduce(const double* part, int n_items, double* out) {
    double se = 0.0, sa = 0.0;
    for (int q = threadIdx.x; q < n_items; q += NTHREADS) {
        se += part[2 * (long)q];
        sa += part[2 * (long)q + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        se += __shfl_down(se, off);
        sa += __shfl_down(sa, off);
    }
    __shared__ double s_red[2][NTHREADS / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_red[0][w] = se;
        s_red[1][w] = sa;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double te = 0.0, ta = 0.0;
        for (int q = 0; q < NTHREADS / 64; q++) {
            te += s_red[0][q];
            ta += s_red[1][q];
        }
        out[0] = te;
        out[1] = ta;
    }
}

template <typename TS>
static constexpr int fres_lds() {
    return GemmNT<TS, TS, double, GEMM_BM, GEMM_BM, 16, false, true>::LDS_BYTES;
}

int factor_residual(ttad_context* ctx, ttad_matrix* A, ttad_matrix* L, double* rel, hipStream_t s) {
    if (!ctx || !A || !L || !rel || A->storage != P_F64 || A->nt != L->nt || A->bsz != L->bsz)
        return set_error(TTAD_ERR_INVALID, "factor_residual: A must be FP64 and L of the same partition");
    const int nt = L->nt;
    // sub-blocks of the lower tiles, the longest sums (large j) first
    std::vector<FresItem> items;
    for (int j = nt - 1; j >= 0; j--)
        for (int i = j; i < nt; i++) {
            const int m = L->bsz[i], n = L->bsz[j];
            for (int c0 = 0; c0 < n; c0 += GEMM_BM)
                for (int r0 = (i == j ? c0 : 0); r0 < m; r0 += GEMM_BM) items.push_back(FresItem{i, j, r0, c0});
        }
    const int n_items = (int)items.size();
    DevBuf d_items, d_part, d_out, d_diag, d_diagtab;
    size_t diag_elems = 0;
    std::vector<size_t> doff(nt);
    for (int j = 0; j < nt; j++) {
        doff[j] = diag_elems;
        diag_elems += ((size_t)L->bsz[j] * L->bsz[j] + 63) / 64 * 64;
    }
    TTAD_HIP(d_items.alloc(std::max<size_t>(1, items.size()) * sizeof(FresItem)));
    TTAD_HIP(d_part.alloc((size_t)std::max(1, n_items) * 2 * sizeof(double)));
    TTAD_HIP(d_out.alloc(2 * sizeof(double)));
    TTAD_HIP(d_diag.alloc(std::max<size_t>(1, diag_elems) * L->elem));
    TTAD_HIP(d_diagtab.alloc(nt * sizeof(void*)));
    std::vector<void*> dp(nt);
    for (int j = 0; j < nt; j++) dp[j] = (char*)d_diag.p + doff[j] * L->elem;
    TTAD_HIP(hipMemcpyAsync(d_items.p, items.data(), items.size() * sizeof(FresItem), hipMemcpyHostToDevice, s));
    TTAD_HIP(hipMemcpyAsync(d_diagtab.p, dp.data(), nt * sizeof(void*), hipMemcpyHostToDevice, s));
    const DevMat* Ld = L->desc.as<DevMat>();
    const DevMat* Ad = A->desc.as<DevMat>();
    if (n_items > 0) {
        if (L->storage == P_F64) {
            hipLaunchKernelGGL(k_lower_copy<double>, dim3(64, nt), dim3(NTHREADS), 0, s, Ld, d_diagtab.as<double*>());
            hipLaunchKernelGGL(k_factor_residual<double>, dim3(n_items), dim3(NTHREADS), fres_lds<double>(), s, Ld,
                               Ad, d_diagtab.as<double*>(), d_items.as<FresItem>(), d_part.as<double>());
        } else {
            hipLaunchKernelGGL(k_lower_copy<float>, dim3(64, nt), dim3(NTHREADS), 0, s, Ld, d_diagtab.as<float*>());
            hipLaunchKernelGGL(k_factor_residual<float>, dim3(n_items), dim3(NTHREADS), fres_lds<float>(), s, Ld,
                               Ad, d_diagtab.as<float*>(), d_items.as<FresItem>(), d_part.as<double>());
        }
    }
    hipLaunchKernelGGL(k_fres_reduce, dim3(1), dim3(NTHREADS), 0, s, d_part.as<double>(), n_items,
                       d_out.as<double>());
    TTAD_HIP(hipGetLastError());
    double h[2] = {0.0, 0.0};
    TTAD_HIP(hipMemcpyAsync(h, d_out.p, sizeof(h), hipMemcpyDeviceToHost, s));
    TTAD_HIP(hipStreamSynchronize(s));
    *rel = h[1] > 0.0 ? std::sqrt(h[0] / h[1]) : (h[0] > 0.0 ? INFINITY : 0.0);
    return TTAD_OK;
}

hipError_t factor_check_init() {
    hipError_t e = hipFuncSetAttribute((const void*)k_factor_residual<double>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, fres_lds<double>());
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_factor_residual<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                fres_lds<float>());
    return e;
}

}  // namespace ttad

using namespace ttad;

extern "C" int ttad_factor_residual(ttad_context* ctx, ttad_matrix* A, ttad_matrix* L, double* rel, void* stream) {
    return factor_residual(ctx, A, L, rel, stream ? (hipStream_t)stream : (ctx ? ctx->stream : nullptr));
}
