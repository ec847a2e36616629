// Device tile operations of the Cholesky DAG, written for gfx950 (CDNA4):
// one workgroup = 256 threads = 4 wave64; all products on MFMA 16x16 tiles.
//
// Replaces the CPU BLAS/LAPACK task bodies of miniapp/dense_cholesky/2d_cholesky.cpp:
//   LAPACKE_dpotrf('L')                       :319      -> item_potrf
//   cblas_dtrsm(Right, Lower, Trans, NonUnit)  :404-405  -> item_trsm
//   cblas_dsyrk(Lower, NoTrans, -1, ., beta)   :490-491  -> item_gemm(syrk=true)
//   cblas_dgemm(NoTrans, Trans, -1, ., beta)   :493-494  -> item_gemm
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include "ttad_types.hpp"

namespace ttad {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef short s8_t __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// MFMA 16x16 primitives. Operand layouts (lane l of the wave):
//   f64/f32 16x16x4 : A[m = l%16][k = l/16], B[k = l/16][n = l%16]
//   bf16   16x16x32 : A[m = l%16][k = 8(l/16)+e], B[k = 8(l/16)+e][n = l%16], e = 0..7
//   result f32/bf16 : D[m = 4(l/16)+r][n = l%16], r = 0..3
//   result f64      : D[m = (l/16)+4r][n = l%16], r = 0..3  (measured on gfx950)
// ---------------------------------------------------------------------------
template <typename TC> struct Mfma;
template <> struct Mfma<double> {
    typedef d4_t acc_t;
    static constexpr int KSTEP = 4;
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef f4_t acc_t;
    static constexpr int KSTEP = 4;
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ unsigned short f2bf_bits(float x) {
    // round to nearest even; NaN kept quiet
    unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <typename T> __device__ __forceinline__ T zero_val() { return T(0); }

// ---------------------------------------------------------------------------
// acc (BM x BN) += A(BM x K) * B(BN x K)^T with A, B column-major (lda, ldb),
// rows >= mv of A and rows >= nv of B read as zero. TA/TB are storage types,
// TC the MFMA type; conversion happens on the global -> LDS path.
// 2x2 waves, each wave owns a (BM/2) x (BN/2) block of 16x16 MFMA tiles.
// K is staged through LDS in chunks of BK with a one-chunk register prefetch.
// ---------------------------------------------------------------------------
template <typename TA, typename TB, typename TC, int BM, int BN, int BK>
struct GemmNT {
    static constexpr int PA = BM + 2;
    static constexpr int PB = BN + 2;
    static constexpr int WM = BM / 32;
    static constexpr int WN = BN / 32;
    static constexpr int EA = BM * BK / NTHREADS;
    static constexpr int EB = BN * BK / NTHREADS;
    static constexpr int LDS_BYTES = BK * (PA + PB) * (int)sizeof(TC);
    typedef typename Mfma<TC>::acc_t acc_t;
    static_assert(EA * NTHREADS == BM * BK && EB * NTHREADS == BN * BK, "tile/thread mismatch");

    static __device__ __forceinline__ void zero(acc_t (&acc)[WM][WN]) {
#pragma unroll
        for (int a = 0; a < WM; a++)
#pragma unroll
            for (int b = 0; b < WN; b++) acc[a][b] = acc_t{0, 0, 0, 0};
    }
    // Element (row, col) of the output block held by this lane in acc[a][b][r].
    static __device__ __forceinline__ int row_of(int a, int r) {
        const int lane = threadIdx.x & 63, wr = (threadIdx.x >> 6) >> 1;
        return wr * (WM * 16) + a * 16 + Mfma<TC>::row(lane, r);
    }
    static __device__ __forceinline__ int col_of(int b) {
        const int lane = threadIdx.x & 63, wc = (threadIdx.x >> 6) & 1;
        return wc * (WN * 16) + b * 16 + (lane & 15);
    }

    // Each thread stages a fixed row (tid % BM) and every (NTHREADS/BM)-th k of a chunk.
    static constexpr int KS_A = NTHREADS / BM;
    static constexpr int KS_B = NTHREADS / BN;
    static_assert(NTHREADS % BM == 0 && NTHREADS % BN == 0, "BM, BN must divide the workgroup");

    template <typename T, int E, int KS>
    static __device__ __forceinline__ void load_chunk(TC (&reg)[E], const T* __restrict__ base, long ld, bool row_ok,
                                                      int kfirst, int K) {
#pragma unroll
        for (int s = 0; s < E; s++) {
            const int kk = kfirst + s * KS;
            reg[s] = (row_ok && kk < K) ? (TC)base[(long)kk * ld] : zero_val<TC>();
        }
    }

    static __device__ __forceinline__ void run(acc_t (&acc)[WM][WN], const TA* __restrict__ A, long lda, int mv,
                               const TB* __restrict__ B, long ldb, int nv, int K, char* lds) {
        TC* As = reinterpret_cast<TC*>(lds);
        TC* Bs = As + BK * PA;
        const int tid = threadIdx.x;
        const int lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
        const int ra_row = tid % BM, ra_k = tid / BM;
        const int rb_row = tid % BN, rb_k = tid / BN;
        const bool a_ok = ra_row < mv, b_ok = rb_row < nv;
        const TA* Ab = A + ra_row;
        const TB* Bb = B + rb_row;
        TC ra[EA], rb[EB];
        const int nchunks = (K + BK - 1) / BK;
        if (nchunks > 0) {
            load_chunk<TA, EA, KS_A>(ra, Ab, lda, a_ok, ra_k, K);
            load_chunk<TB, EB, KS_B>(rb, Bb, ldb, b_ok, rb_k, K);
        }
        for (int c = 0; c < nchunks; c++) {
            __syncthreads();
#pragma unroll
            for (int s = 0; s < EA; s++) As[(ra_k + s * KS_A) * PA + ra_row] = ra[s];
#pragma unroll
            for (int s = 0; s < EB; s++) Bs[(rb_k + s * KS_B) * PB + rb_row] = rb[s];
            __syncthreads();
            if (c + 1 < nchunks) {
                const int k0 = (c + 1) * BK;
                load_chunk<TA, EA, KS_A>(ra, Ab, lda, a_ok, k0 + ra_k, K);
                load_chunk<TB, EB, KS_B>(rb, Bb, ldb, b_ok, k0 + rb_k, K);
            }
#pragma unroll 2
            for (int ks = 0; ks < BK / 4; ks++) {
                const int kr = ks * 4 + (lane >> 4);
                TC av[WM], bv[WN];
#pragma unroll
                for (int a = 0; a < WM; a++) av[a] = As[kr * PA + wr * (WM * 16) + a * 16 + (lane & 15)];
#pragma unroll
                for (int b = 0; b < WN; b++) bv[b] = Bs[kr * PB + wc * (WN * 16) + b * 16 + (lane & 15)];
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) acc[a][b] = Mfma<TC>::mma(av[a], bv[b], acc[a][b]);
            }
        }
        __syncthreads();
    }
};

// ---------------------------------------------------------------------------
// BF16 inputs, FP32 accumulation: acc (BM x BN) += bf16(A) * bf16(B)^T.
// LDS holds bf16 in [row][k] order so that one lane reads its 8 consecutive k
// values with a single 16-byte load.
// ---------------------------------------------------------------------------
// SPLIT = true: each FP32 operand x is split into hi = bf16(x) and lo = bf16(x - hi)
// and the product accumulates hi*hi + hi*lo + lo*hi (three MFMAs), which keeps about
// 16 significant bits of every operand (Henry, Tang, Heinecke, "Leveraging the
// bfloat16 artificial intelligence datatype for higher-precision computations", 2019).
template <typename TA, int BM, int BN, bool SPLIT = false>
struct GemmNTbf16 {
    static constexpr int BK = 32;
    static constexpr int PK = BK + 8;  // bf16 elements per LDS row (16-byte aligned, skewed)
    static constexpr int WM = BM / 32;
    static constexpr int WN = BN / 32;
    static constexpr int NP = SPLIT ? 2 : 1;  // hi (and lo) planes
    static constexpr int LDS_BYTES = NP * (BM + BN) * PK * 2;
    typedef f4_t acc_t;
    // each thread converts 8 consecutive k of one row per step
    static constexpr int GA = BM * (BK / 8) / NTHREADS;
    static constexpr int GB = BN * (BK / 8) / NTHREADS;
    static_assert(GA * NTHREADS == BM * (BK / 8) && GB * NTHREADS == BN * (BK / 8), "tile/thread mismatch");

    static __device__ __forceinline__ void zero(acc_t (&acc)[WM][WN]) {
#pragma unroll
        for (int a = 0; a < WM; a++)
#pragma unroll
            for (int b = 0; b < WN; b++) acc[a][b] = acc_t{0, 0, 0, 0};
    }
    static __device__ __forceinline__ int row_of(int a, int r) {
        const int lane = threadIdx.x & 63, wr = (threadIdx.x >> 6) >> 1;
        return wr * (WM * 16) + a * 16 + 4 * (lane >> 4) + r;
    }
    static __device__ __forceinline__ int col_of(int b) {
        const int lane = threadIdx.x & 63, wc = (threadIdx.x >> 6) & 1;
        return wc * (WN * 16) + b * 16 + (lane & 15);
    }

    template <int G, int GS>
    static __device__ __forceinline__ void load8(s8_t (&hi)[G], s8_t (&lo)[G], const TA* __restrict__ base, long ld,
                                                 bool row_ok, int kfirst, int K) {
#pragma unroll
        for (int s = 0; s < G; s++) {
            const int k = kfirst + s * GS * 8;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const float x = (row_ok && k + q < K) ? (float)base[(long)(k + q) * ld] : 0.0f;
                const unsigned short h = f2bf_bits(x);
                hi[s][q] = (short)h;
                if constexpr (SPLIT) lo[s][q] = (short)f2bf_bits(x - __uint_as_float((unsigned)h << 16));
            }
        }
    }

    static __device__ __forceinline__ void run(acc_t (&acc)[WM][WN], const TA* __restrict__ A, long lda, int mv,
                               const TA* __restrict__ B, long ldb, int nv, int K, char* lds) {
        unsigned short* As = reinterpret_cast<unsigned short*>(lds);  // hi plane of A
        unsigned short* Bs = As + BM * PK;                            // hi plane of B
        unsigned short* Al = Bs + BN * PK;                            // lo planes (SPLIT)
        unsigned short* Bl = Al + BM * PK;
        const int tid = threadIdx.x;
        const int lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
        // thread -> fixed row (tid % BM) and k-groups of 8 (tid / BM + s * NTHREADS / BM)
        const int a_row = tid % BM, a_g = tid / BM, b_row = tid % BN, b_g = tid / BN;
        const bool a_ok = a_row < mv, b_ok = b_row < nv;
        const TA* Ab = A + a_row;
        const TA* Bb = B + b_row;
        s8_t ra[GA], rb[GB], la[GA], lb[GB];
        const int nchunks = (K + BK - 1) / BK;
        if (nchunks > 0) {
            load8<GA, NTHREADS / BM>(ra, la, Ab, lda, a_ok, a_g * 8, K);
            load8<GB, NTHREADS / BN>(rb, lb, Bb, ldb, b_ok, b_g * 8, K);
        }
        for (int c = 0; c < nchunks; c++) {
            __syncthreads();
#pragma unroll
            for (int s = 0; s < GA; s++) {
                const int off = a_row * PK + (a_g + s * (NTHREADS / BM)) * 8;
                *reinterpret_cast<s8_t*>(As + off) = ra[s];
                if constexpr (SPLIT) *reinterpret_cast<s8_t*>(Al + off) = la[s];
            }
#pragma unroll
            for (int s = 0; s < GB; s++) {
                const int off = b_row * PK + (b_g + s * (NTHREADS / BN)) * 8;
                *reinterpret_cast<s8_t*>(Bs + off) = rb[s];
                if constexpr (SPLIT) *reinterpret_cast<s8_t*>(Bl + off) = lb[s];
            }
            __syncthreads();
            if (c + 1 < nchunks) {
                const int k0 = (c + 1) * BK;
                load8<GA, NTHREADS / BM>(ra, la, Ab, lda, a_ok, k0 + a_g * 8, K);
                load8<GB, NTHREADS / BN>(rb, lb, Bb, ldb, b_ok, k0 + b_g * 8, K);
            }
            const int kg = (lane >> 4) * 8;
            s8_t av[WM], bv[WN];
#pragma unroll
            for (int a = 0; a < WM; a++)
                av[a] = *reinterpret_cast<const s8_t*>(As + (wr * (WM * 16) + a * 16 + (lane & 15)) * PK + kg);
#pragma unroll
            for (int b = 0; b < WN; b++)
                bv[b] = *reinterpret_cast<const s8_t*>(Bs + (wc * (WN * 16) + b * 16 + (lane & 15)) * PK + kg);
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
            if constexpr (SPLIT) {
                s8_t avl[WM], bvl[WN];
#pragma unroll
                for (int a = 0; a < WM; a++)
                    avl[a] = *reinterpret_cast<const s8_t*>(Al + (wr * (WM * 16) + a * 16 + (lane & 15)) * PK + kg);
#pragma unroll
                for (int b = 0; b < WN; b++)
                    bvl[b] = *reinterpret_cast<const s8_t*>(Bl + (wc * (WN * 16) + b * 16 + (lane & 15)) * PK + kg);
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) {
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bvl[b], acc[a][b], 0, 0, 0);
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(avl[a], bv[b], acc[a][b], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }
};

// ---------------------------------------------------------------------------
// GEMM / SYRK work item: C[r0:r0+128, c0:c0+128] = beta*C + alpha * A[r0:,:] B[c0:,:]^T
// over the valid part of an m x n tile, K = inner size. With syrk, only the
// lower triangle (global row >= global col) is written (dsyrk 'L').
// beta == 0 overwrites C without reading it (accumulate-mode temporaries).
// ---------------------------------------------------------------------------
template <typename TS, int PREC>
__device__ __forceinline__ void item_gemm(const TS* __restrict__ A, long lda, const TS* __restrict__ B, long ldb, TS* __restrict__ C,
                          long ldc, int m, int n, int K, int r0, int c0, bool syrk, double alpha, double beta,
                          char* lds) {
    const int mv = min(GEMM_BM, m - r0), nv = min(GEMM_BM, n - c0);
    if (mv <= 0 || nv <= 0) return;
    if constexpr (PREC == P_BF16 || PREC == P_BF16X3) {
        using G = GemmNTbf16<TS, GEMM_BM, GEMM_BM, PREC == P_BF16X3>;
        typename G::acc_t acc[G::WM][G::WN];
        G::zero(acc);
        G::run(acc, A + r0, lda, mv, B + c0, ldb, nv, K, lds);
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G::row_of(a, r), cc = G::col_of(b);
                    if (rr < mv && cc < nv && (!syrk || r0 + rr >= c0 + cc)) {
                        TS* p = C + (r0 + rr) + (long)(c0 + cc) * ldc;
                        const double v = alpha * (double)acc[a][b][r];
                        *p = (TS)(beta == 0.0 ? v : beta * (double)(*p) + v);
                    }
                }
    } else {
        typedef TS TC;  // F64 tiles on FP64 MFMA, F32 tiles on FP32 MFMA
        using G = GemmNT<TS, TS, TC, GEMM_BM, GEMM_BM, 16>;
        typename G::acc_t acc[G::WM][G::WN];
        G::zero(acc);
        G::run(acc, A + r0, lda, mv, B + c0, ldb, nv, K, lds);
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G::row_of(a, r), cc = G::col_of(b);
                    if (rr < mv && cc < nv && (!syrk || r0 + rr >= c0 + cc)) {
                        TS* p = C + (r0 + rr) + (long)(c0 + cc) * ldc;
                        const TC v = (TC)alpha * acc[a][b][r];
                        *p = (beta == 0.0) ? v : (TC)beta * (*p) + v;
                    }
                }
    }
}

// Decode the (row block, col block) of GEMM/SYRK item `it` in a tile of tr x tc sub-tiles.
__device__ __forceinline__ void decode_gemm_item(int it, int tr, bool syrk, int& rb, int& cb) {
    if (!syrk) {
        rb = it % tr;
        cb = it / tr;
        return;
    }
    // lower-triangular enumeration, column-major: (c, r) with r >= c
    int c = 0, rem = it;
    while (rem >= tr - c) { rem -= tr - c; c++; }
    cb = c;
    rb = c + rem;
}

// ---------------------------------------------------------------------------
// 64x64 diagonal block in LDS (column-major, ld DP): factor (lower Cholesky) and
// invert in place. FP64 throughout. Padding rows/cols beyond `w` hold identity.
// ---------------------------------------------------------------------------
constexpr int DP = DB + 1;

// Right-looking Cholesky of D (w x w valid); one barrier per column. Returns the
// first failing local column + 1 (0 if positive definite) to every thread.
__device__ __forceinline__ int lds_potrf64(double* D, int w) {
    const int tid = threadIdx.x;
    __shared__ int s_bad;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int k = 0; k < w; k++) {
        const double dkk = D[k + k * DP];
        if (tid == 0 && !(dkk > 0.0) && s_bad == 0) s_bad = k + 1;
        const double inv = 1.0 / dkk;
        // trailing lower part: rows i in (k, w), cols j in (k, i]
        const int t = w - k - 1;
        for (int e = tid; e < t * t; e += NTHREADS) {
            const int i = k + 1 + e % t, j = k + 1 + e / t;
            if (i >= j) D[i + j * DP] -= D[i + k * DP] * D[j + k * DP] * inv;
        }
        __syncthreads();
    }
    // scale: L(:,k) = D(:,k) / sqrt(D(k,k))
    __shared__ double s_sq[DB];
    for (int k = tid; k < w; k += NTHREADS) s_sq[k] = sqrt(D[k + k * DP]);
    __syncthreads();
    for (int e = tid; e < w * w; e += NTHREADS) {
        const int i = e % w, k = e / w;
        if (i > k) D[i + k * DP] /= s_sq[k];
        else if (i == k) D[k + k * DP] = s_sq[k];
    }
    __syncthreads();
    return s_bad;
}

// In-place inverse of the lower-triangular w x w L held in D (LAPACK dtrti2 order:
// columns from last to first, each a triangular matrix-vector product with the
// already inverted trailing block). Upper part is zeroed; padding becomes identity.
__device__ __forceinline__ void lds_trtri64(double* D, int w) {
    const int tid = threadIdx.x;
    for (int e = tid; e < DB * DB; e += NTHREADS) {
        const int i = e % DB, j = e / DB;
        if (i < j) D[i + j * DP] = 0.0;
        else if (i >= w || j >= w) D[i + j * DP] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int j = w - 1; j >= 0; j--) {
        const double ajj = 1.0 / D[j + j * DP];
        // x_i = sum_{q=j+1}^{i} Xinv(i,q) * L(q,j) for i > j
        double x = 0.0;
        const int i = j + 1 + tid;
        if (tid < w - j - 1) {
            for (int q = j + 1; q <= i; q++) x += D[i + q * DP] * D[q + j * DP];
        }
        __syncthreads();
        if (tid < w - j - 1) D[i + j * DP] = -ajj * x;
        if (tid == 0) D[j + j * DP] = ajj;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// POTRF work item: the whole n x n diagonal tile in one workgroup, left-looking
// over 64-column blocks; leaves inverses of the 64x64 diagonal blocks of L in dinv.
// Arithmetic is FP64 whatever the tile storage (the mixed variants keep the
// diagonal factorization in FP64).
// ---------------------------------------------------------------------------
template <typename TS>
__device__ __forceinline__ void item_potrf(TS* __restrict__ A, long lda, int n, double* __restrict__ dinv, int* info, int col_off,
                           char* lds) {
    using G = GemmNT<TS, TS, double, 64, 64, 32>;
    using G2 = GemmNT<TS, double, double, 64, 64, 32>;
    static_assert(G::LDS_BYTES == G2::LDS_BYTES, "staging size");
    double* D = reinterpret_cast<double*>(lds + G::LDS_BYTES);
    const int tid = threadIdx.x;
    const int nblk = (n + DB - 1) / DB;
    for (int pb = 0; pb < nblk; pb++) {
        const int p0 = pb * DB, pw = min(DB, n - p0);
        // D = A(p,p) - L(p, 0:p0) L(p, 0:p0)^T
        typename G::acc_t acc[G::WM][G::WN];
        G::zero(acc);
        if (p0 > 0) G::run(acc, A + p0, lda, pw, A + p0, lda, pw, p0, lds);
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G::row_of(a, r), cc = G::col_of(b);
                    double v;
                    if (rr < pw && cc < pw) v = (rr >= cc) ? (double)A[(p0 + rr) + (long)(p0 + cc) * lda] - acc[a][b][r] : 0.0;
                    else v = (rr == cc) ? 1.0 : 0.0;
                    D[rr + cc * DP] = v;
                }
        __syncthreads();
        const int bad = lds_potrf64(D, pw);
        if (bad && tid == 0) atomicCAS(info, 0, col_off + p0 + bad);
        // write L(p,p) (lower) back to the tile
        for (int e = tid; e < pw * pw; e += NTHREADS) {
            const int i = e % pw, j = e / pw;
            if (i >= j) A[(p0 + i) + (long)(p0 + j) * lda] = (TS)D[i + j * DP];
        }
        __syncthreads();
        lds_trtri64(D, pw);
        double* Xinv = dinv + (long)pb * DB * DB;
        for (int e = tid; e < DB * DB; e += NTHREADS) Xinv[e] = D[(e % DB) + (e / DB) * DP];
        __syncthreads();
        // panel below: L(q,p) = (A(q,p) - L(q,0:p0) L(p,0:p0)^T) Xinv^T
        for (int qb = pb + 1; qb < nblk; qb++) {
            const int q0 = qb * DB, qw = min(DB, n - q0);
            TS* T = A + q0 + (long)p0 * lda;
            G::zero(acc);
            if (p0 > 0) G::run(acc, A + q0, lda, qw, A + p0, lda, pw, p0, lds);
#pragma unroll
            for (int a = 0; a < G::WM; a++)
#pragma unroll
                for (int b = 0; b < G::WN; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int rr = G::row_of(a, r), cc = G::col_of(b);
                        if (rr < qw && cc < pw) {
                            TS* p = T + rr + (long)cc * lda;
                            *p = (TS)((double)(*p) - acc[a][b][r]);
                        }
                    }
            __syncthreads();
            G2::zero(acc);
            G2::run(acc, T, lda, qw, Xinv, DB, pw, pw, lds);
#pragma unroll
            for (int a = 0; a < G::WM; a++)
#pragma unroll
                for (int b = 0; b < G::WN; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int rr = G::row_of(a, r), cc = G::col_of(b);
                        if (rr < qw && cc < pw) T[rr + (long)cc * lda] = (TS)acc[a][b][r];
                    }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// TRSM work item: rows [r0, r0+64) of X (m x n) := X L^{-T} (Right, Lower,
// Trans, NonUnit), in place, by 64-column blocks: subtract the already solved
// columns times L's off-diagonal blocks, then multiply by the inverse of L's
// 64x64 diagonal block left by POTRF. TC is the MFMA type (FP64, or FP32 for the
// mixed variants).
// ---------------------------------------------------------------------------
template <typename TS, typename TC>
__device__ __forceinline__ void item_trsm(const TS* __restrict__ L, long ldl, const double* __restrict__ dinv, TS* __restrict__ X,
                          long ldx, int m, int n, int r0, char* lds) {
    using G = GemmNT<TS, TS, TC, 64, 64, 32>;
    using G2 = GemmNT<TS, double, TC, 64, 64, 32>;
    const int mv = min(TRSM_BM, m - r0);
    if (mv <= 0) return;
    TS* Xr = X + r0;
    const int nblk = (n + DB - 1) / DB;
    typename G::acc_t acc[G::WM][G::WN];
    for (int cb = 0; cb < nblk; cb++) {
        const int c0 = cb * DB, cw = min(DB, n - c0);
        TS* T = Xr + (long)c0 * ldx;
        G::zero(acc);
        if (c0 > 0) G::run(acc, Xr, ldx, mv, L + c0, ldl, cw, c0, lds);
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G::row_of(a, r), cc = G::col_of(b);
                    if (rr < mv && cc < cw) {
                        TS* p = T + rr + (long)cc * ldx;
                        *p = (TS)((TC)(*p) - acc[a][b][r]);
                    }
                }
        __syncthreads();
        G2::zero(acc);
        G2::run(acc, T, ldx, mv, dinv + (long)cb * DB * DB, DB, cw, cw, lds);
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G::row_of(a, r), cc = G::col_of(b);
                    if (rr < mv && cc < cw) T[rr + (long)cc * ldx] = (TS)acc[a][b][r];
                }
        __syncthreads();
    }
}

// LDS bytes each item kind needs (the executor reserves the maximum).
constexpr int LDS_GEMM_F64 = GemmNT<double, double, double, GEMM_BM, GEMM_BM, 16>::LDS_BYTES;
constexpr int LDS_GEMM_F32 = GemmNT<float, float, float, GEMM_BM, GEMM_BM, 16>::LDS_BYTES;
constexpr int LDS_GEMM_BF16 = GemmNTbf16<float, GEMM_BM, GEMM_BM, true>::LDS_BYTES;
constexpr int LDS_POTRF = GemmNT<double, double, double, 64, 64, 32>::LDS_BYTES + DB * DP * 8;
constexpr int LDS_TRSM = GemmNT<double, double, double, 64, 64, 32>::LDS_BYTES;
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int LDS_ITEM_MAX = cmax(cmax(cmax(LDS_GEMM_F64, LDS_GEMM_F32), cmax(LDS_GEMM_BF16, LDS_POTRF)), LDS_TRSM);

}  // namespace ttad
