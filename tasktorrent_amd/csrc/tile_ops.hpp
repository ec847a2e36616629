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
#include <type_traits>
#include "ttad_types.hpp"

namespace ttad {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef short s8_t __attribute__((ext_vector_type(8)));

// Device memory accesses of the item code go through address-space-1 pointers. Tile
// pointers reach it through tables in memory, so the compiler sees generic pointers and
// emits flat instructions, which count against the LDS wait counter as well: every LDS
// wait of a GEMM loop would then also wait for the prefetched loads of the next chunk.
#define TTAD_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ const TTAD_GLOBAL T* gmem(const T* p) {
    return (const TTAD_GLOBAL T*)p;
}
template <typename T> __device__ __forceinline__ TTAD_GLOBAL T* gmem(T* p) { return (TTAD_GLOBAL T*)p; }
// A value the whole wave holds, moved to scalar registers: values loaded from task records in global
// memory are vector loads (the executors write global memory), so the compiler cannot tell they are
// uniform; address arithmetic on them would otherwise be 64-bit vector work per lane.
__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <typename T> __device__ __forceinline__ T* wave_uniform(T* p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (T*)(((unsigned long long)hi << 32) | lo);
}

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

template <typename T> __device__ __forceinline__ T zero_val() { return T(0); }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// 16-byte register vectors of a storage type (aligned: LDS side; u: global side, element-aligned).
template <typename T> struct V16;
template <> struct V16<double> {
    static constexpr int N = 2;
    typedef double t __attribute__((ext_vector_type(2)));
    typedef double u __attribute__((ext_vector_type(2), aligned(8)));
};
template <> struct V16<float> {
    static constexpr int N = 4;
    typedef float t __attribute__((ext_vector_type(4)));
    typedef float u __attribute__((ext_vector_type(4), aligned(4)));
};

// ---------------------------------------------------------------------------
// acc (BM x BN) += A(BM x K) * B(BN x K)^T with A, B column-major (lda, ldb),
// rows >= mv of A and rows >= nv of B read as zero. TA/TB are storage types,
// TC the MFMA type; conversion happens on the register -> LDS path.
// 2x2 waves, each wave owns a (BM/2) x (BN/2) block of 16x16 MFMA tiles.
// K is staged through two LDS buffers in chunks of BK: while the MFMAs consume
// chunk c, chunk c+1 goes from registers to the other buffer and the global
// loads of chunk c+2 are in flight (one barrier per chunk).
// VEC: each lane loads 16 bytes (2 FP64 / 4 FP32 consecutive rows of a column) and
//      stores them to LDS in one instruction (TA = TB = TC).
// SWAP: the MFMA takes B's block as its first operand, so that a lane's results lie
//      along a column of the output (row = lane % 16): C is then read and written in
//      16-row segments instead of 16 scattered columns.
// ---------------------------------------------------------------------------
// WR: waves stacked along the rows (2: 2x2 waves; 4: each wave owns BM/4 full rows).
template <typename TA, typename TB, typename TC, int BM, int BN, int BK, bool VEC = false, bool SWAP = false, int WR = 2>
struct GemmNT {
    static constexpr int NVA = VEC ? V16<TA>::N : 1;  // rows per lane load
    static constexpr int NVB = VEC ? V16<TB>::N : 1;
    static constexpr int PA = BM + (NVA > 2 ? NVA : 2);  // LDS row pitch (16-byte multiple when VEC)
    static constexpr int PB = BN + (NVB > 2 ? NVB : 2);
    static constexpr int WC = 4 / WR;             // waves along the columns
    static constexpr int WM = BM / (16 * WR);     // 16x16 blocks per wave
    static constexpr int WN = BN / (16 * WC);
    static constexpr int LA = BM / NVA;          // lanes per column of the A chunk
    static constexpr int LB = BN / NVB;
    static constexpr int KS_A = NTHREADS / LA;   // columns per load round
    static constexpr int KS_B = NTHREADS / LB;
    static constexpr int EA = BK / KS_A;         // loads per thread per chunk
    static constexpr int EB = BK / KS_B;
    static constexpr int STAGE = BK * (PA + PB);  // TC elements of one LDS buffer
    static constexpr int LDS_BYTES = 2 * STAGE * (int)sizeof(TC);
    typedef typename Mfma<TC>::acc_t acc_t;
    typedef std::conditional_t<VEC, typename V16<TA>::t, TA> RA;
    typedef std::conditional_t<VEC, typename V16<TB>::t, TB> RB;
    static_assert(NTHREADS % LA == 0 && NTHREADS % LB == 0 && BK % KS_A == 0 && BK % KS_B == 0, "tile/thread mismatch");
    static_assert(!VEC || (std::is_same<TA, TC>::value && std::is_same<TB, TC>::value), "VEC stages without conversion");
    static_assert(BK % 4 == 0, "BK is a multiple of the MFMA k-step");

    static __device__ __forceinline__ void zero(acc_t (&acc)[WM][WN]) {
#pragma unroll
        for (int a = 0; a < WM; a++)
#pragma unroll
            for (int b = 0; b < WN; b++) acc[a][b] = acc_t{0, 0, 0, 0};
    }
    // Element (row, col) of the output block held by this lane in acc[a][b][r].
    static __device__ __forceinline__ int row_of(int a, int r) {
        const int lane = threadIdx.x & 63, wr = (threadIdx.x >> 6) / WC;
        return wr * (WM * 16) + a * 16 + (SWAP ? (lane & 15) : Mfma<TC>::row(lane, r));
    }
    static __device__ __forceinline__ int col_of(int b, int r) {
        const int lane = threadIdx.x & 63, wc = (threadIdx.x >> 6) % WC;
        return wc * (WN * 16) + b * 16 + (SWAP ? Mfma<TC>::row(lane, r) : (lane & 15));
    }

    // Thread tid stages rows NV*(tid % L) .. +NV-1 of every (NTHREADS/L)-th column of a chunk:
    // one wave instruction reads whole columns (coalesced).
    template <typename T, typename R, int E, int NV, int L, int KS>
    static __device__ __forceinline__ void load_chunk(R (&reg)[E], const T* __restrict__ base, long ld, int nvalid,
                                                      int k0, int K) {
        const int row = NV * ((int)threadIdx.x % L), kc = (int)threadIdx.x / L;
        if (nvalid >= L * NV && k0 + KS * E <= K) {  // whole chunk inside (uniform): no guards
#pragma unroll
            for (int s = 0; s < E; s++) {
                const TTAD_GLOBAL T* p = gmem(base) + row + (long)(k0 + kc + s * KS) * ld;
                if constexpr (NV == 1) reg[s] = *p;
                else reg[s] = *reinterpret_cast<const TTAD_GLOBAL typename V16<T>::u*>(p);
            }
            return;
        }
#pragma unroll
        for (int s = 0; s < E; s++) {
            const int kk = k0 + kc + s * KS;
            const TTAD_GLOBAL T* p = gmem(base) + row + (long)kk * ld;
            if constexpr (NV == 1) {
                reg[s] = (row < nvalid && kk < K) ? *p : T(0);
            } else {
#pragma unroll
                for (int q = 0; q < NV; q++) reg[s][q] = (kk < K && row + q < nvalid) ? p[q] : T(0);
            }
        }
    }
    template <typename R, int E, int NV, int L, int KS, int P>
    static __device__ __forceinline__ void store_chunk(TC* __restrict__ S, const R (&reg)[E]) {
        const int row = NV * ((int)threadIdx.x % L), kc = (int)threadIdx.x / L;
#pragma unroll
        for (int s = 0; s < E; s++) {
            TC* q = S + (kc + s * KS) * P + row;
            if constexpr (NV == 1) *q = (TC)reg[s];
            else *reinterpret_cast<R*>(q) = reg[s];
        }
    }
    static __device__ __forceinline__ void load(RA (&ra)[EA], RB (&rb)[EB], const TA* A, long lda, int mv, const TB* B,
                                                long ldb, int nv, int k0, int K) {
        load_chunk<TA, RA, EA, NVA, LA, KS_A>(ra, A, lda, mv, k0, K);
        load_chunk<TB, RB, EB, NVB, LB, KS_B>(rb, B, ldb, nv, k0, K);
    }
    static __device__ __forceinline__ void store(TC* __restrict__ S, const RA (&ra)[EA], const RB (&rb)[EB]) {
        store_chunk<RA, EA, NVA, LA, KS_A, PA>(S, ra);
        store_chunk<RB, EB, NVB, LB, KS_B, PB>(S + BK * PA, rb);
    }

    static __device__ __forceinline__ void compute(acc_t (&acc)[WM][WN], const TC* __restrict__ S, int lane, int wr,
                                                   int wc) {
        const TC* As = S + wr * (WM * 16) + (lane & 15);
        const TC* Bs = S + BK * PA + wc * (WN * 16) + (lane & 15);
#pragma unroll 4
        for (int ks = 0; ks < BK / 4; ks++) {
            const int kr = ks * 4 + (lane >> 4);
            TC av[WM], bv[WN];
#pragma unroll
            for (int a = 0; a < WM; a++) av[a] = As[kr * PA + a * 16];
#pragma unroll
            for (int b = 0; b < WN; b++) bv[b] = Bs[kr * PB + b * 16];
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++)
                    acc[a][b] = SWAP ? Mfma<TC>::mma(bv[b], av[a], acc[a][b]) : Mfma<TC>::mma(av[a], bv[b], acc[a][b]);
        }
    }

    // Starts and ends with a workgroup barrier: the caller may use `lds` before and after.
    static __device__ __forceinline__ void run(acc_t (&acc)[WM][WN], const TA* __restrict__ A, long lda, int mv,
                                               const TB* __restrict__ B, long ldb, int nv, int K, char* lds) {
        TC* S0 = reinterpret_cast<TC*>(lds);
        const int tid = threadIdx.x;
        const int lane = tid & 63, w = tid >> 6, wr = w / WC, wc = w % WC;
        const int nchunks = (K + BK - 1) / BK;
        __syncthreads();
        if (nchunks == 0) return;
        RA ra[EA];
        RB rb[EB];
        load(ra, rb, A, lda, mv, B, ldb, nv, 0, K);
        store(S0, ra, rb);
        if (nchunks > 1) load(ra, rb, A, lda, mv, B, ldb, nv, BK, K);
        __syncthreads();
        for (int c = 0; c < nchunks; c++) {
            compute(acc, S0 + (c & 1) * STAGE, lane, wr, wc);
            if (c + 1 < nchunks) {
                // the other buffer was last read in chunk c-1, before the previous barrier
                store(S0 + ((c + 1) & 1) * STAGE, ra, rb);
                if (c + 2 < nchunks) load(ra, rb, A, lda, mv, B, ldb, nv, (c + 2) * BK, K);
            }
            __syncthreads();
        }
    }
};

// ---------------------------------------------------------------------------
// BF16 inputs, FP32 accumulation: acc (BM x BN) += bf16(A) * bf16(B)^T.
// LDS holds bf16 in [row][k] order so that one lane reads its 8 consecutive k
// values with a single 16-byte load. Raw FP32 values are prefetched one chunk
// ahead in registers and converted (v_cvt_pk_bf16_f32, round to nearest even)
// on their way into the LDS buffer not being read; two buffers, one barrier per chunk.
// ---------------------------------------------------------------------------
// SPLIT = true: each FP32 operand x is split into hi = bf16(x) and lo = bf16(x - hi)
// and the product accumulates hi*hi + hi*lo + lo*hi (three MFMAs), which keeps about
// 16 significant bits of every operand (Henry, Tang, Heinecke, "Leveraging the
// bfloat16 artificial intelligence datatype for higher-precision computations", 2019).
typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));
typedef unsigned u4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned pack_bf16x2(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f2_t){a, b}, bf2_t));
}

template <typename TA, int BM, int BN, bool SPLIT = false, bool SWAP = false>
struct GemmNTbf16 {
    // K-chunk: 64 (one plane), 32 with the hi/lo planes, so that both fit two workgroups per
    // CU (72 KB of LDS). Rows of 72 elements (144 B) are read 16 bytes at a time; rows of 36
    // (72 B, the split planes) 8 bytes at a time; both pitches spread 16 rows over distinct banks.
    static constexpr int BK = SPLIT ? 32 : 64;
    static constexpr int PK = SPLIT ? BK + 4 : BK + 8;  // bf16 elements per LDS row
    static constexpr bool NARROW = (PK * 2) % 16 != 0;  // LDS accesses in 8-byte halves
    static constexpr int WM = BM / 32;
    static constexpr int WN = BN / 32;
    static constexpr int NP = SPLIT ? 2 : 1;  // hi (and lo) planes
    static constexpr int STAGE = NP * (BM + BN) * PK;  // bf16 elements of one LDS buffer
    static constexpr int LDS_BYTES = 2 * STAGE * 2;
    typedef f4_t acc_t;
    // each thread owns groups of 8 consecutive k of one row
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
        return wr * (WM * 16) + a * 16 + (SWAP ? (lane & 15) : 4 * (lane >> 4) + r);
    }
    static __device__ __forceinline__ int col_of(int b, int r) {
        const int lane = threadIdx.x & 63, wc = (threadIdx.x >> 6) & 1;
        return wc * (WN * 16) + b * 16 + (SWAP ? 4 * (lane >> 4) + r : (lane & 15));
    }
    static __device__ __forceinline__ f4_t mma(s8_t a, s8_t b, f4_t c) {
        return SWAP ? __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a, c, 0, 0, 0)
                    : __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }

    // full: the whole chunk is inside the operand (uniform), so the loads need no guards; the
    // guarded form, one exec-masked branch per load, is kept for edge chunks. base is the
    // operand's (uniform) first row and kfirst is uniform over the wave, so each column's address
    // is a scalar base plus this lane's 32-bit row offset: no per-lane 64-bit address is kept per
    // column (those spilled, and a reload from scratch waits for every load in flight, which made
    // the one-chunk prefetch synchronous).
    template <int G, int GS>
    static __device__ __forceinline__ void load_raw(float (&r)[G][8], const TA* __restrict__ base, long ld, int row,
                                                    bool row_ok, int kfirst, int K, bool full) {
        const TTAD_GLOBAL TA* gb = gmem(base);
        if (full) {
#pragma unroll
            for (int s = 0; s < G; s++) {
                const int k = kfirst + s * GS * 8;
#pragma unroll
                for (int q = 0; q < 8; q++) r[s][q] = (float)(gb + (long)(k + q) * ld)[row];
            }
            return;
        }
#pragma unroll
        for (int s = 0; s < G; s++) {
            const int k = kfirst + s * GS * 8;
#pragma unroll
            for (int q = 0; q < 8; q++)
                r[s][q] = (row_ok && k + q < K) ? (float)(gb + (long)(k + q) * ld)[row] : 0.0f;
        }
    }

    template <int G, int GS>
    static __device__ __forceinline__ void store_planes(unsigned short* __restrict__ hi_plane,
                                                        unsigned short* __restrict__ lo_plane, const float (&r)[G][8],
                                                        int row, int g0) {
#pragma unroll
        for (int s = 0; s < G; s++) {
            u4_t h, l;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float x0 = r[s][2 * q], x1 = r[s][2 * q + 1];
                const unsigned hp = pack_bf16x2(x0, x1);
                h[q] = hp;
                if constexpr (SPLIT)
                    l[q] = pack_bf16x2(x0 - __uint_as_float(hp << 16), x1 - __uint_as_float(hp & 0xffff0000u));
            }
            const int off = row * PK + (g0 + s * GS) * 8;
            if constexpr (NARROW) {
                typedef unsigned u2_t __attribute__((ext_vector_type(2)));
                *reinterpret_cast<u2_t*>(hi_plane + off) = u2_t{h[0], h[1]};
                *reinterpret_cast<u2_t*>(hi_plane + off + 4) = u2_t{h[2], h[3]};
                if constexpr (SPLIT) {
                    *reinterpret_cast<u2_t*>(lo_plane + off) = u2_t{l[0], l[1]};
                    *reinterpret_cast<u2_t*>(lo_plane + off + 4) = u2_t{l[2], l[3]};
                }
            } else {
                *reinterpret_cast<u4_t*>(hi_plane + off) = h;
                if constexpr (SPLIT) *reinterpret_cast<u4_t*>(lo_plane + off) = l;
            }
        }
    }

    // 8 consecutive bf16 (one lane's k-group, already split) into LDS
    static __device__ __forceinline__ void put8(unsigned short* p, u4_t v) {
        if constexpr (NARROW) {
            typedef unsigned u2_t __attribute__((ext_vector_type(2)));
            *reinterpret_cast<u2_t*>(p) = u2_t{v[0], v[1]};
            *reinterpret_cast<u2_t*>(p + 4) = u2_t{v[2], v[3]};
        } else {
            *reinterpret_cast<u4_t*>(p) = v;
        }
    }

    // 8 consecutive bf16 (one lane's k-group) from LDS
    static __device__ __forceinline__ s8_t ld8(const unsigned short* p) {
        if constexpr (NARROW) {
            typedef short s4_t __attribute__((ext_vector_type(4)));
            const s4_t a = *reinterpret_cast<const s4_t*>(p), b = *reinterpret_cast<const s4_t*>(p + 4);
            return s8_t{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
        } else {
            return *reinterpret_cast<const s8_t*>(p);
        }
    }
    static __device__ __forceinline__ void compute(acc_t (&acc)[WM][WN], const unsigned short* __restrict__ S,
                                                   int lane, int wr, int wc) {
        const unsigned short* As = S;
        const unsigned short* Bs = As + BM * PK;
        const unsigned short* Al = Bs + BN * PK;
        const unsigned short* Bl = Al + BM * PK;
#pragma unroll 1  // one k-step's hi/lo operands live at a time (register budget of the executors)
        for (int ks = 0; ks < BK / 32; ks++) {
            const int kg = ks * 32 + (lane >> 4) * 8;
            s8_t av[WM], bv[WN];
#pragma unroll
            for (int a = 0; a < WM; a++)
                av[a] = ld8(As + (wr * (WM * 16) + a * 16 + (lane & 15)) * PK + kg);
#pragma unroll
            for (int b = 0; b < WN; b++)
                bv[b] = ld8(Bs + (wc * (WN * 16) + b * 16 + (lane & 15)) * PK + kg);
#pragma unroll
            for (int a = 0; a < WM; a++)
#pragma unroll
                for (int b = 0; b < WN; b++)
                    acc[a][b] = mma(av[a], bv[b], acc[a][b]);
            if constexpr (SPLIT) {
                s8_t avl[WM], bvl[WN];
#pragma unroll
                for (int a = 0; a < WM; a++)
                    avl[a] = ld8(Al + (wr * (WM * 16) + a * 16 + (lane & 15)) * PK + kg);
#pragma unroll
                for (int b = 0; b < WN; b++)
                    bvl[b] = ld8(Bl + (wc * (WN * 16) + b * 16 + (lane & 15)) * PK + kg);
#pragma unroll
                for (int a = 0; a < WM; a++)
#pragma unroll
                    for (int b = 0; b < WN; b++) {
                        acc[a][b] = mma(av[a], bvl[b], acc[a][b]);
                        acc[a][b] = mma(avl[a], bv[b], acc[a][b]);
                    }
            }
        }
    }

    // Starts and ends with a workgroup barrier: the caller may use `lds` before and after.
    static __device__ __forceinline__ void run(acc_t (&acc)[WM][WN], const TA* __restrict__ A, long lda, int mv,
                               const TA* __restrict__ B, long ldb, int nv, int K, char* lds) {
        unsigned short* S0 = reinterpret_cast<unsigned short*>(lds);
        const int tid = threadIdx.x;
        const int lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
        // thread -> fixed row (tid % BM) and k-groups of 8 (tid / BM + s * NTHREADS / BM); BM and BN are
        // multiples of the wave size, so a_g and b_g are the same over a wave
        static_assert(BM % 64 == 0 && BN % 64 == 0, "k-group of a thread must be uniform over its wave");
        const int a_row = tid % BM, b_row = tid % BN;
        const int a_g = __builtin_amdgcn_readfirstlane(tid / BM), b_g = __builtin_amdgcn_readfirstlane(tid / BN);
        const bool a_ok = a_row < mv, b_ok = b_row < nv;
        const int nchunks = (K + BK - 1) / BK;
        __syncthreads();
        if (nchunks == 0) return;
        float ra[GA][8], rb[GB][8];
        auto stage = [&](unsigned short* S) {
            store_planes<GA, NTHREADS / BM>(S, S + (BM + BN) * PK, ra, a_row, a_g);
            store_planes<GB, NTHREADS / BN>(S + BM * PK, S + (BM + BN) * PK + BM * PK, rb, b_row, b_g);
        };
        // a chunk is full when both operands have all their rows and the chunk lies inside K
        const bool rows_full = __builtin_amdgcn_readfirstlane((int)(mv >= BM && nv >= BN)) != 0;
        auto full = [&](int k0) { return rows_full && k0 + BK <= K; };
        load_raw<GA, NTHREADS / BM>(ra, A, lda, a_row, a_ok, a_g * 8, K, full(0));
        load_raw<GB, NTHREADS / BN>(rb, B, ldb, b_row, b_ok, b_g * 8, K, full(0));
        stage(S0);
        if (nchunks > 1) {
            load_raw<GA, NTHREADS / BM>(ra, A, lda, a_row, a_ok, BK + a_g * 8, K, full(BK));
            load_raw<GB, NTHREADS / BN>(rb, B, ldb, b_row, b_ok, BK + b_g * 8, K, full(BK));
        }
        __syncthreads();
        for (int c = 0; c < nchunks; c++) {
            compute(acc, S0 + (c & 1) * STAGE, lane, wr, wc);
            if (c + 1 < nchunks) {
                stage(S0 + ((c + 1) & 1) * STAGE);
                if (c + 2 < nchunks) {
                    const int k0 = (c + 2) * BK;
                    load_raw<GA, NTHREADS / BM>(ra, A, lda, a_row, a_ok, k0 + a_g * 8, K, full(k0));
                    load_raw<GB, NTHREADS / BN>(rb, B, ldb, b_row, b_ok, k0 + b_g * 8, K, full(k0));
                }
            }
            __syncthreads();
        }
    }

    // As run() for SPLIT with BM x BN fully inside both operands and K % BK == 0, from operands
    // already split into bf16 planes (plane_granule layout; PA and PB point at the granule of the
    // item's first row, arows / brows are the tiles' row counts). Each thread moves whole 16-byte
    // granules, so a chunk costs 4 loads per operand and no conversion; the LDS buffers receive
    // exactly what run() would store from the FP32 values the planes were split from.
    static __device__ __forceinline__ void run_planes(acc_t (&acc)[WM][WN], const u4_t* __restrict__ PA, int arows,
                                                      const u4_t* __restrict__ PB, int brows, int K, char* lds) {
        static_assert(SPLIT && BK == 32, "plane chunks are 32 columns of hi and lo");
        unsigned short* S0 = reinterpret_cast<unsigned short*>(lds);
        const int tid = threadIdx.x;
        const int lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
        const int a_row = tid % BM, b_row = tid % BN;
        const int a_g = __builtin_amdgcn_readfirstlane(tid / BM), b_g = __builtin_amdgcn_readfirstlane(tid / BN);
        const int nchunks = wave_uniform(K / BK);
        arows = wave_uniform(arows);
        brows = wave_uniform(brows);
        __syncthreads();
        if (nchunks == 0) return;
        u4_t ra[GA][2], rb[GB][2];
        // a uniform granule base per (chunk, plane, group) plus this lane's 32-bit row offset
        const TTAD_GLOBAL u4_t* ga = gmem(wave_uniform(PA));
        const TTAD_GLOBAL u4_t* gb = gmem(wave_uniform(PB));
        auto fetch = [&](int c) {
#pragma unroll
            for (int s = 0; s < GA; s++)
#pragma unroll
                for (int p = 0; p < 2; p++)
                    ra[s][p] = (ga + plane_granule(c, p, a_g + s * (NTHREADS / BM), 0, arows))[a_row];
#pragma unroll
            for (int s = 0; s < GB; s++)
#pragma unroll
                for (int p = 0; p < 2; p++)
                    rb[s][p] = (gb + plane_granule(c, p, b_g + s * (NTHREADS / BN), 0, brows))[b_row];
        };
        auto stage = [&](unsigned short* S) {
            unsigned short* Ah = S;
            unsigned short* Bh = S + BM * PK;
            unsigned short* Al = S + (BM + BN) * PK;
            unsigned short* Bl = Al + BM * PK;
#pragma unroll
            for (int s = 0; s < GA; s++) {
                const int off = a_row * PK + (a_g + s * (NTHREADS / BM)) * 8;
                put8(Ah + off, ra[s][0]);
                put8(Al + off, ra[s][1]);
            }
#pragma unroll
            for (int s = 0; s < GB; s++) {
                const int off = b_row * PK + (b_g + s * (NTHREADS / BN)) * 8;
                put8(Bh + off, rb[s][0]);
                put8(Bl + off, rb[s][1]);
            }
        };
        fetch(0);
        stage(S0);
        if (nchunks > 1) fetch(1);
        __syncthreads();
        for (int c = 0; c < nchunks; c++) {
            compute(acc, S0 + (c & 1) * STAGE, lane, wr, wc);
            if (c + 1 < nchunks) {
                stage(S0 + ((c + 1) & 1) * STAGE);
                if (c + 2 < nchunks) fetch(c + 2);
            }
            __syncthreads();
        }
    }
};

// Splits 4 consecutive FP32 values of a row (columns k .. k+3, k % 4 == 0) into the bf16 hi/lo planes
// of a tile of m rows (plane_granule layout), with the rounding GemmNTbf16 applies on its way to LDS.
__device__ __forceinline__ void store_split4(unsigned short* __restrict__ planes, int m, int row, int k, f4_t x) {
    typedef unsigned u2_t __attribute__((ext_vector_type(2)));
    const unsigned h01 = pack_bf16x2(x[0], x[1]), h23 = pack_bf16x2(x[2], x[3]);
    const unsigned l01 = pack_bf16x2(x[0] - __uint_as_float(h01 << 16), x[1] - __uint_as_float(h01 & 0xffff0000u));
    const unsigned l23 = pack_bf16x2(x[2] - __uint_as_float(h23 << 16), x[3] - __uint_as_float(h23 & 0xffff0000u));
    TTAD_GLOBAL unsigned short* g = gmem(planes);
    const long hi = plane_granule(k / 32, 0, (k % 32) / 8, row, m) * 8 + (k % 8);
    const long lo = plane_granule(k / 32, 1, (k % 32) / 8, row, m) * 8 + (k % 8);
    *reinterpret_cast<TTAD_GLOBAL u2_t*>(g + hi) = u2_t{h01, h23};
    *reinterpret_cast<TTAD_GLOBAL u2_t*>(g + lo) = u2_t{l01, l23};
}

// K-chunk of the 128x128 FP64/FP32 GEMM items: 16 FP64 / 32 FP32 columns, 65 KB of LDS for the
// two buffers, so that two workgroups share a CU: one's barriers, loads and C traffic overlap
// the other's MFMAs (FP64, m = n = 16384, K = 512: 43.5 TF/s with one 130 KB workgroup per CU,
// 53 TF/s with two; 66 TF/s with the 16-byte loads, coalesced C and C read up front).
template <typename TC> constexpr int gemm_bk() { return sizeof(TC) == 8 ? 16 : 32; }
// K-chunk of a TRSM item's update from the solved columns: each chunk waits for its loads (issued one chunk
// ahead), so fewer chunks cost fewer round trips. 64 FP32 columns fit the LDS of the FP32 executors; the FP64
// ones allow 32.
template <typename TS> constexpr int trsm_bk() { return sizeof(TS) == 4 ? 64 : 32; }

// ---------------------------------------------------------------------------
// GEMM / SYRK work item: C[r0:r0+128, c0:c0+128] = beta*C + alpha * A[r0:,:] B[c0:,:]^T
// over the valid part of an m x n tile, K = inner size. With syrk, only the
// lower triangle (global row >= global col) is written (dsyrk 'L').
// beta == 0 overwrites C without reading it (accumulate-mode temporaries).
// ---------------------------------------------------------------------------
// CPRE (the Cholesky update, alpha = -1, beta = 1): the accumulators start from -C, read
// together with the first K-chunk instead of after the last one, and C = -acc is stored.
// PA / PB (bf16x3 only; null = none): the bf16 split planes of the operand tiles (A and B with lda and ldb
// rows, plane_granule layout), used by items that lie wholly inside both with K % 32 == 0.
template <typename TS, int PREC, bool CPRE = false>
__device__ __forceinline__ void item_gemm(const TS* __restrict__ A, long lda, const TS* __restrict__ B, long ldb, TS* __restrict__ C,
                          long ldc, int m, int n, int K, int r0, int c0, bool syrk, double alpha, double beta,
                          char* lds, const void* PA = nullptr, const void* PB = nullptr) {
    const int mv = min(GEMM_BM, m - r0), nv = min(GEMM_BM, n - c0);
    if (mv <= 0 || nv <= 0) return;
    auto body = [&](auto* gsel) {
        using G = typename std::remove_pointer<decltype(gsel)>::type;
        typedef typename G::acc_t acc_t;
        acc_t acc[G::WM][G::WN];
        auto owned = [&](int rr, int cc) { return rr < mv && cc < nv && (!syrk || r0 + rr >= c0 + cc); };
        if constexpr (CPRE) {
#pragma unroll
            for (int a = 0; a < G::WM; a++)
#pragma unroll
                for (int b = 0; b < G::WN; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int rr = G::row_of(a, r), cc = G::col_of(b, r);
                        acc[a][b][r] = owned(rr, cc) ? -gmem(C)[(r0 + rr) + (long)(c0 + cc) * ldc] : TS(0);
                    }
        } else {
            G::zero(acc);
        }
        bool planes = false;
        if constexpr (PREC == P_BF16X3)
            planes = PA && PB && mv == GEMM_BM && nv == GEMM_BM && K % G::BK == 0;
        if constexpr (PREC == P_BF16X3) {
            if (planes)
                G::run_planes(acc, static_cast<const u4_t*>(PA) + r0, (int)lda, static_cast<const u4_t*>(PB) + c0,
                              (int)ldb, K, lds);
        }
        if (!planes) G::run(acc, A + r0, lda, mv, B + c0, ldb, nv, K, lds);
        if constexpr (CPRE) {
#pragma unroll
            for (int a = 0; a < G::WM; a++)
#pragma unroll
                for (int b = 0; b < G::WN; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int rr = G::row_of(a, r), cc = G::col_of(b, r);
                        if (owned(rr, cc)) gmem(C)[(r0 + rr) + (long)(c0 + cc) * ldc] = (TS)(-acc[a][b][r]);
                    }
        } else {
            // C is read after the last K-chunk. Every load of a group of fragment rows is issued before any
            // store of the group: as far as the compiler knows a store may alias a later load, so a load and
            // a store per element in turn would cost one memory round trip per element (64 per item). Groups:
            // all of the item for FP32 tiles, one fragment row for FP64 (registers). Loads come from clamped addresses
            // (C[0] for elements not owned), so none waits under a branch; interior items (all of the
            // 128 x 128 owned) store without a condition.
            const bool full = __builtin_amdgcn_readfirstlane(
                                  (int)(mv == GEMM_BM && nv == GEMM_BM && (!syrk || r0 >= c0 + GEMM_BM))) != 0;
            if constexpr (sizeof(TS) == 4) {
                // Interior FP32 items go through LDS (free after the K loop): a lane's results lie along a row,
                // C is column-major, so the accumulators are transposed in LDS and every thread then reads,
                // updates and writes 16 contiguous bytes of C per access (16 loads and 16 stores per thread
                // instead of 64 each). C is read first, so its latency overlaps the transposition.
                if (full) {
                    constexpr int P = GEMM_BM + 4;  // LDS column pitch: the 4 column groups of a wave hit distinct banks
                    static_assert(GEMM_BM * P * 4 <= G::LDS_BYTES, "transposed item fits the item's LDS");
                    typedef typename V16<TS>::u f4u;
                    TS* T = reinterpret_cast<TS*>(lds);
                    const int rg = threadIdx.x % 32, cl = threadIdx.x / 32;  // rows 4rg .. 4rg+3 of columns cl + 8j
                    const long ldu = wave_uniform((int)ldc);
                    TTAD_GLOBAL TS* base = gmem(C) + (r0 + 4 * rg) + (long)(c0 + cl) * ldu;
                    constexpr int NJ = GEMM_BM / 8;
                    f4_t cv[NJ];
                    if (beta != 0.0) {
#pragma unroll
                        for (int j = 0; j < NJ; j++) cv[j] = *reinterpret_cast<TTAD_GLOBAL f4u*>(base + 8 * j * ldu);
                    }
#pragma unroll
                    for (int a = 0; a < G::WM; a++)
#pragma unroll
                        for (int b = 0; b < G::WN; b++)
#pragma unroll
                            for (int r = 0; r < 4; r++) T[G::col_of(b, r) * P + G::row_of(a, r)] = acc[a][b][r];
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < NJ; j++) {
                        const f4_t v = *reinterpret_cast<const f4_t*>(T + (cl + 8 * j) * P + 4 * rg);
                        f4_t o;
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const double x = alpha * (double)v[r];
                            o[r] = (TS)(beta == 0.0 ? x : beta * (double)cv[j][r] + x);
                        }
                        *reinterpret_cast<TTAD_GLOBAL f4u*>(base + 8 * j * ldu) = o;
                    }
                    __syncthreads();  // the caller may use lds again
                    return;
                }
            }
            constexpr int AG = sizeof(TS) == 8 ? 1 : G::WM;
#pragma unroll
            for (int a0 = 0; a0 < G::WM; a0 += AG) {
                TS cv[AG][G::WN][4];
                if (beta != 0.0) {
#pragma unroll
                    for (int a = 0; a < AG; a++)
#pragma unroll
                        for (int b = 0; b < G::WN; b++)
#pragma unroll
                            for (int r = 0; r < 4; r++) {
                                const int rr = G::row_of(a0 + a, r), cc = G::col_of(b, r);
                                const long off = (r0 + rr) + (long)(c0 + cc) * ldc;
                                cv[a][b][r] = gmem(C)[(full || owned(rr, cc)) ? off : 0];
                            }
                }
#pragma unroll
                for (int a = 0; a < AG; a++)
#pragma unroll
                    for (int b = 0; b < G::WN; b++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int rr = G::row_of(a0 + a, r), cc = G::col_of(b, r);
                            const double v = alpha * (double)acc[a0 + a][b][r];
                            const TS o = (TS)(beta == 0.0 ? v : beta * (double)cv[a][b][r] + v);
                            if (full || owned(rr, cc)) gmem(C)[(r0 + rr) + (long)(c0 + cc) * ldc] = o;
                        }
            }
        }
    };
    if constexpr (PREC == P_BF16 || PREC == P_BF16X3) {
        body((GemmNTbf16<TS, GEMM_BM, GEMM_BM, PREC == P_BF16X3, true>*)nullptr);
    } else {  // F64 tiles on FP64 MFMA, F32 tiles on FP32 MFMA
        body((GemmNT<TS, TS, TS, GEMM_BM, GEMM_BM, gemm_bk<TS>(), true, true>*)nullptr);
    }
}

// Accumulate-mode GEMM / SYRK item (2d_cholesky.cpp:479-501 with beta = 0, then the ACCU task
// `*blocks[i+j*nb] += *Atmp`, :570, fused): the product is formed from zero and added into C while
// holding `lock`, the sub-tile's lock word (agent-scope acquire / release), so contributions of
// different k land in any order but one at a time. A lock wait gives up after spin_limit ticks or
// once *err is nonzero (and sets *err = 1 when it times out).
template <typename TS, int PREC>
__device__ __forceinline__ void item_gemm_acc(const TS* __restrict__ A, long lda, const TS* __restrict__ B, long ldb,
                                              TS* __restrict__ C, long ldc, int m, int n, int K, int r0, int c0,
                                              bool syrk, double alpha, int32_t* lock, int32_t* err,
                                              int64_t spin_limit, char* lds) {
    const int mv = min(GEMM_BM, m - r0), nv = min(GEMM_BM, n - c0);
    if (mv <= 0 || nv <= 0) return;
    auto body = [&](auto* gsel) {
        using G = typename std::remove_pointer<decltype(gsel)>::type;
        typedef typename G::acc_t acc_t;
        acc_t acc[G::WM][G::WN];
        G::zero(acc);
        G::run(acc, A + r0, lda, mv, B + c0, ldb, nv, K, lds);
        if (threadIdx.x == 0) {
            const int64_t t0 = (int64_t)wall_clock64();
            int expected = 0;
            while (!__hip_atomic_compare_exchange_strong(lock, &expected, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) {
                expected = 0;
                if ((int64_t)wall_clock64() - t0 > spin_limit ||
                    __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the previous holder's additions
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G::row_of(a, r), cc = G::col_of(b, r);
                    if (rr < mv && cc < nv && (!syrk || r0 + rr >= c0 + cc)) {
                        TTAD_GLOBAL TS* p = gmem(C) + (r0 + rr) + (long)(c0 + cc) * ldc;
                        *p = (TS)((double)(*p) + alpha * (double)acc[a][b][r]);
                    }
                }
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(lock, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    };
    if constexpr (PREC == P_BF16 || PREC == P_BF16X3) {
        body((GemmNTbf16<TS, GEMM_BM, GEMM_BM, PREC == P_BF16X3, true>*)nullptr);
    } else {
        body((GemmNT<TS, TS, TS, GEMM_BM, GEMM_BM, gemm_bk<TS>(), true, true>*)nullptr);
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
// 64x64 diagonal block in LDS (column-major, ld DP), blocked by 16: factor
// (lower Cholesky) and inverse. FP64 throughout. Padding rows/cols beyond `w`
// hold identity. Diagonal 16x16 blocks are handled element-wise (thread t owns
// element (t%16, t/16), one barrier per column); everything else is 16x16x16
// products on FP64 MFMA, one block per wave.
// ---------------------------------------------------------------------------
constexpr int DP = DB + 1;

// acc += P Q (q_trans = false) or P Q^T (q_trans = true); P, Q 16x16 blocks in LDS, ld DP.
__device__ __forceinline__ d4_t mma16(d4_t acc, const double* P, const double* Q, bool q_trans) {
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int k = 4 * s + kq;
        const double b = q_trans ? Q[m + k * DP] : Q[k + m * DP];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(P[m + k * DP], b, acc, 0, 0, 0);
    }
    return acc;
}
// R = sign * acc (acc in the FP64 MFMA result layout), or R += sign * acc.
__device__ __forceinline__ void put16(double* R, d4_t acc, double sign, bool add) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        double* p = R + ((lane >> 4) + 4 * r) + (lane & 15) * DP;
        *p = add ? *p + sign * acc[r] : sign * acc[r];
    }
}
// out = P T with T held in MFMA result registers (its row 4s + lane/16 is acc[s]: the B operand layout).
__device__ __forceinline__ d4_t mma16_reg(const double* P, d4_t T) {
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    d4_t out = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; s++) out = __builtin_amdgcn_mfma_f64_16x16x4f64(P[m + (4 * s + kq) * DP], T[s], out, 0, 0, 0);
    return out;
}

// 1/x to about one ulp without the FP64 division sequence: the hardware reciprocal estimate refined by two
// Newton-Raphson steps r <- r + r (1 - x r), each of which doubles the number of correct bits. The POTRF
// diagonal blocks divide by every pivot in turn, so the division's latency is on the factorization's
// critical path.
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// X_KK = L_KK^{-1} for the 16x16 lower-triangular diagonal block at Lk (X block at Xk,
// upper part left untouched). Row-by-row substitution with the partial sums in registers:
// one barrier per row. All threads participate.
__device__ __forceinline__ void tri_inv16(const double* Lk, double* Xk) {
    const int ti = threadIdx.x & 15, tc = threadIdx.x >> 4;
    const double rd = rcp_nr(Lk[ti + ti * DP]);  // 1 / L(ti, ti), ahead of the row-by-row chain
    double acc = 0.0;
    for (int q = 0; q < 16; q++) {
        if (ti == q && tc <= q) Xk[q + tc * DP] = ((q == tc ? 1.0 : 0.0) - acc) * rd;
        __syncthreads();
        if (ti > q && tc <= q) acc += Lk[ti + q * DP] * Xk[q + tc * DP];
    }
    __syncthreads();
}

// Off-diagonal blocks of X = L^{-1} from L (D) and the diagonal inverses already in X:
// X_IJ = -X_II sum_{Q=J}^{I-1} L_IQ X_QJ, in order of I - J (one wave per block).
__device__ __forceinline__ void tri_inv_offdiag(const double* D, double* X) {
    const int wv = threadIdx.x >> 6;
    for (int d = 1; d < 4; d++) {
        const int J = wv, I = wv + d;
        if (I < 4) {
            d4_t T = {0, 0, 0, 0};
            for (int Q = J; Q < I; Q++) T = mma16(T, D + 16 * I + 16 * Q * DP, X + 16 * Q + 16 * J * DP, false);
            put16(X + 16 * I + 16 * J * DP, mma16_reg(X + 16 * I + 16 * I * DP, T), -1.0, false);
        }
        __syncthreads();
    }
}

// Factor D in place (lower; the upper part is left undefined) and write X = L^{-1}
// (lower, upper part zero). Returns the first failing local column + 1 (0 if
// positive definite) to every thread.
__device__ __forceinline__ int lds_potrf_inv64(double* D, double* X, int w) {
    const int tid = threadIdx.x, wv = tid >> 6;
    const int ti = tid & 15, tj = tid >> 4;
    __shared__ int s_bad;
    if (tid == 0) s_bad = 0;
    for (int e = tid; e < DB * DP; e += NTHREADS) X[e] = 0.0;
    __syncthreads();
    for (int K = 0; K < 4; K++) {
        const int k0 = 16 * K;
        double* Dk = D + k0 + k0 * DP;
        // diagonal block: unscaled right-looking elimination, then column scaling
        for (int k = 0; k < 16; k++) {
            const double dkk = Dk[k + k * DP];
            if (tid == 0 && !(dkk > 0.0) && s_bad == 0 && k0 + k < w) s_bad = k0 + k + 1;
            if (ti > k && tj > k && ti >= tj) Dk[ti + tj * DP] -= Dk[ti + k * DP] * Dk[tj + k * DP] * rcp_nr(dkk);
            __syncthreads();
        }
        const double sj = sqrt(Dk[tj + tj * DP]);
        const double v = Dk[ti + tj * DP];
        __syncthreads();
        if (ti > tj) Dk[ti + tj * DP] = v * rcp_nr(sj);
        else if (ti == tj) Dk[ti + tj * DP] = sj;
        __syncthreads();
        tri_inv16(Dk, X + k0 + k0 * DP);
        if (K == 3) break;
        // panel: L_IK = D_IK X_KK^T (one wave per block below)
        {
            const int I = K + 1 + wv;
            if (I < 4) {
                double* Dik = D + 16 * I + k0 * DP;
                const d4_t acc = mma16(d4_t{0, 0, 0, 0}, Dik, X + k0 + k0 * DP, true);
                put16(Dik, acc, 1.0, false);
            }
        }
        __syncthreads();
        // trailing update: D_IJ -= L_IK L_JK^T for K < J <= I < 4
        {
            const int nblk = (3 - K) * (4 - K) / 2;
            for (int q = wv; q < nblk; q += 4) {
                int J = K + 1, rem = q;
                while (rem >= 4 - J) { rem -= 4 - J; J++; }
                const int I = J + rem;
                const d4_t acc = mma16(d4_t{0, 0, 0, 0}, D + 16 * I + k0 * DP, D + 16 * J + k0 * DP, true);
                put16(D + 16 * I + 16 * J * DP, acc, -1.0, true);
            }
        }
        __syncthreads();
    }
    tri_inv_offdiag(D, X);
    return s_bad;
}

// In-place style inverse of a lower-triangular 64x64 L held in D (w valid, identity
// padding set here): X = L^{-1}, lower, upper part zero.
__device__ __forceinline__ void lds_trtri64(double* D, double* X, int w) {
    const int tid = threadIdx.x;
    for (int e = tid; e < DB * DB; e += NTHREADS) {
        const int i = e % DB, j = e / DB;
        if (i >= w || j >= w) D[i + j * DP] = (i == j) ? 1.0 : 0.0;
    }
    for (int e = tid; e < DB * DP; e += NTHREADS) X[e] = 0.0;
    __syncthreads();
    for (int K = 0; K < 4; K++) tri_inv16(D + 16 * K + 16 * K * DP, X + 16 * K + 16 * K * DP);
    tri_inv_offdiag(D, X);
}

// ---------------------------------------------------------------------------
// POTRF work item: the whole n x n diagonal tile in one workgroup, left-looking
// over 64-column panels. Per panel: one tall GEMM updates the panel rows p0..n
// (256 rows per call), the 64x64 diagonal block is factored and inverted in LDS
// (lds_potrf_inv64), and the rows below are multiplied by the inverse's transpose.
// Leaves the inverses of the 64x64 diagonal blocks of L in dinv. Arithmetic is
// FP64 whatever the tile storage (the mixed variants keep the diagonal
// factorization in FP64).
// ---------------------------------------------------------------------------
constexpr int POTRF_BM = 128;  // rows per panel GEMM call
template <typename TS>
struct PotrfGemm {
    using G = GemmNT<TS, TS, double, POTRF_BM, DB, 16>;        // panel update
    using G2 = GemmNT<TS, double, double, POTRF_BM, DB, 16>;   // times the inverse
};

template <typename TS>
__device__ __forceinline__ void item_potrf(TS* __restrict__ A, long lda, int n, double* __restrict__ dinv, int* info, int col_off,
                           char* lds) {
    using G = typename PotrfGemm<TS>::G;
    using G2 = typename PotrfGemm<TS>::G2;
    double* D = reinterpret_cast<double*>(lds);  // shares the GEMM staging area (used between GEMM calls)
    double* X = D + DB * DP;
    const int tid = threadIdx.x;
    const int nblk = (n + DB - 1) / DB;
    for (int pb = 0; pb < nblk; pb++) {
        const int p0 = pb * DB, pw = min(DB, n - p0), rows = n - p0;
        typename G::acc_t acc[G::WM][G::WN];
        // panel update: A(p0:n, p) -= L(p0:n, 0:p0) L(p, 0:p0)^T
        if (p0 > 0) {
            for (int r0 = 0; r0 < rows; r0 += POTRF_BM) {
                const int mv = min(POTRF_BM, rows - r0);
                TTAD_GLOBAL TS* T = gmem(A) + (p0 + r0) + (long)p0 * lda;
                G::zero(acc);
                G::run(acc, A + p0 + r0, lda, mv, A + p0, lda, pw, p0, lds);
#pragma unroll
                for (int a = 0; a < G::WM; a++)
#pragma unroll
                    for (int b = 0; b < G::WN; b++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int rr = G::row_of(a, r), cc = G::col_of(b, r);
                            if (rr < mv && cc < pw && (r0 + rr >= cc)) {
                                TTAD_GLOBAL TS* q = T + rr + (long)cc * lda;
                                *q = (TS)((double)(*q) - acc[a][b][r]);
                            }
                        }
            }
            __syncthreads();
        }
        // diagonal block into LDS (identity padding), every load in flight before the
        // first LDS store; then factor + inverse
        {
            constexpr int PER = DB * DB / NTHREADS;
            const int i = tid % DB, j0 = tid / DB;
            TTAD_GLOBAL const TS* Ac = gmem(A) + (p0 + i) + (long)p0 * lda;
            TS v[PER];
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int j = j0 + q * (NTHREADS / DB);
                v[q] = (i < pw && j < pw && i >= j) ? Ac[(long)j * lda] : TS(0);
            }
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int j = j0 + q * (NTHREADS / DB);
                D[i + j * DP] = (i < pw && j < pw) ? (double)v[q] : (i == j) ? 1.0 : 0.0;
            }
        }
        __syncthreads();
        const int bad = lds_potrf_inv64(D, X, pw);
        if (bad && tid == 0) atomicCAS(info, 0, col_off + p0 + bad);
        double* Xinv = dinv + (long)pb * DB * DB;
        {
            constexpr int PER = DB * DB / NTHREADS;
            const int i = tid % DB, j0 = tid / DB;
            TTAD_GLOBAL TS* Ac = gmem(A) + (p0 + i) + (long)p0 * lda;
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int j = j0 + q * (NTHREADS / DB);
                const double dv = D[i + j * DP], xv = X[i + j * DP];
                if (i < pw && j < pw && i >= j) Ac[(long)j * lda] = (TS)dv;
                gmem(Xinv)[i + j * DB] = xv;
            }
        }
        __syncthreads();
        // rows below: L(q, p) = A(q, p) Xinv^T
        for (int r0 = DB; r0 < rows; r0 += POTRF_BM) {
            const int mv = min(POTRF_BM, rows - r0);
            TS* T = A + (p0 + r0) + (long)p0 * lda;
            TTAD_GLOBAL TS* Tg = gmem(T);
            G2::zero(acc);
            G2::run(acc, T, lda, mv, Xinv, DB, pw, pw, lds);
#pragma unroll
            for (int a = 0; a < G::WM; a++)
#pragma unroll
                for (int b = 0; b < G::WN; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int rr = G::row_of(a, r), cc = G::col_of(b, r);
                        if (rr < mv && cc < pw) Tg[rr + (long)cc * lda] = (TS)acc[a][b][r];
                    }
        }
        __syncthreads();
    }
}

// Waits until *count >= want and returns the value seen. The polls are relaxed (an acquire
// load invalidates the L2 on every iteration, which with hundreds of waiting workgroups
// evicts everyone's data); one agent-scope acquire follows. Gives up after `limit` ticks
// (sets *err = 1) or as soon as *err is nonzero (an earlier wait timed out, or the owner
// asked the kernel to stop), so that a failed run drains in one time limit. `slow` polls
// about once a microsecond instead of back to back.
__device__ __forceinline__ int wait_count(const int32_t* count, int want, int32_t* err, int64_t limit,
                                          bool slow = false) {
    const int64_t t0 = (int64_t)wall_clock64();
    int v;
    while ((v = __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < want) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) { v = want; break; }
        if ((int64_t)wall_clock64() - t0 > limit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            v = want;
            break;
        }
        if (slow) __builtin_amdgcn_s_sleep(32);
        else __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return v;
}


// ---------------------------------------------------------------------------
// POTRF split into row-block items (used by the persistent executor): item rb of the
// task owns rows [64 rb, 64 rb + 64) of the n x n tile and computes L(rb, 0..rb) by a
// right-looking pass over the 64-column panels j < rb:
//   L(rb,j)   = A(rb,j) Xinv(j)^T          once panel j is published
//   A(rb,rb) -= L(rb,j) L(rb,j)^T          kept in registers
//   A(rb,q)  -= L(rb,j) L(q,j)^T, j<q<rb  once L(q,j) is published
// and then factors and inverts its diagonal block (lds_potrf_inv64) and publishes
// Xinv(rb). prog[rb] counts the panels whose part of row block rb is final (rb + 1:
// the whole row block, diagonal block and its inverse included). An item waits only on
// lower-numbered items of its task, which are claimed before it, so every wait ends.
// From one diagonal block to the next the chain is one 64^3 product, one 64^3 update
// and the 64x64 factorization, where the single-workgroup item_potrf runs every panel
// GEMM of the tile in sequence. Same arithmetic (FP64, operands rounded to the tile's
// storage type as stored) and outputs (L in the tile, dinv) as item_potrf.
// ---------------------------------------------------------------------------
constexpr int POTRF_FLAGS = 128;  // ints of progress flags per diagonal tile
constexpr int POTRF_FIN = POTRF_FLAGS - 1;  // completion count (clears the flags for the next run)
__host__ __device__ inline int potrf_row_items(int n) { return (n + DB - 1) / DB; }

// Dst[k * P + m] = (double) T[m + k * ld] for m < rows, k < cols (else 0); 64 x 64, every
// global load in flight before the first LDS store.
template <typename TS, int P>
__device__ __forceinline__ void stage64(double* Dst, const TTAD_GLOBAL TS* T, long ld, int rows, int cols) {
    constexpr int PER = DB * DB / NTHREADS;
    const int m = threadIdx.x % DB, k0 = threadIdx.x / DB;
    TS v[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int k = k0 + q * (NTHREADS / DB);
        v[q] = (m < rows && k < cols) ? T[m + (long)k * ld] : TS(0);
    }
#pragma unroll
    for (int q = 0; q < PER; q++) Dst[(k0 + q * (NTHREADS / DB)) * P + m] = (double)v[q];
}

template <typename TS>
__device__ __forceinline__ void item_potrf_rows(TS* __restrict__ A, long lda, int n, double* __restrict__ dinv,
                                                int* info, int col_off, int rb, int32_t* prog, int32_t* err,
                                                int64_t spin_limit, char* lds) {
    using G3 = GemmNT<double, double, double, DB, DB, DB, false, true>;  // 64x64x64, both operands in LDS
    using acc_t = typename G3::acc_t;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w / G3::WC, wc = w % G3::WC;
    const int nrb = potrf_row_items(n);
    const int p0 = rb * DB, pw = min(DB, n - p0);
    double* At = reinterpret_cast<double*>(lds);  // At[k * PA + m] = first operand (m, k)
    double* Bt = At + DB * G3::PA;                // Bt[k * PB + c] = second operand (c, k)
    TTAD_GLOBAL TS* Ag = gmem(A);
    TTAD_GLOBAL TS* Arow = Ag + p0;  // row block rb
    acc_t accd[G3::WM][G3::WN], acc[G3::WM][G3::WN];
    // accd = -A(rb, rb), lower part
#pragma unroll
    for (int a = 0; a < G3::WM; a++)
#pragma unroll
        for (int b = 0; b < G3::WN; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = G3::row_of(a, r), cc = G3::col_of(b, r);
                accd[a][b][r] = (rr < pw && cc < pw && rr >= cc) ? -(double)Arow[rr + (long)(p0 + cc) * lda] : 0.0;
            }
    for (int j = 0; j < rb; j++) {
        const int c0 = j * DB;  // panel j < rb is a full 64 columns
        if (tid == 0) wait_count(prog + j, j + 1, err, spin_limit);
        __syncthreads();  // Xinv(j) and L(j, 0..j) final; At / Bt free
        stage64<TS, G3::PA>(At, Arow + (long)c0 * lda, lda, pw, DB);
        stage64<double, G3::PB>(Bt, gmem(dinv) + (long)j * DB * DB, DB, DB, DB);
        __syncthreads();
        G3::zero(acc);
        G3::compute(acc, At, lane, wr, wc);
        __syncthreads();
        // L(rb, j) into the tile and, as stored, into both operand buffers
#pragma unroll
        for (int a = 0; a < G3::WM; a++)
#pragma unroll
            for (int b = 0; b < G3::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G3::row_of(a, r), cc = G3::col_of(b, r);
                    const TS v = (TS)acc[a][b][r];
                    if (rr < pw) Arow[rr + (long)(c0 + cc) * lda] = v;
                    At[cc * G3::PA + rr] = (double)v;
                    Bt[cc * G3::PB + rr] = (double)v;
                }
        __syncthreads();
        if (tid == 0) __hip_atomic_store(prog + rb, j + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        G3::compute(accd, At, lane, wr, wc);  // -A(rb,rb) + L(rb,j) L(rb,j)^T
        // blocks of this row to the right of panel j: A(rb,q) -= L(rb,j) L(q,j)^T
        for (int q = j + 1; q < rb; q++) {
            if (tid == 0) wait_count(prog + q, j + 1, err, spin_limit);
            __syncthreads();  // L(q, j) final; Bt free
            stage64<TS, G3::PB>(Bt, Ag + q * DB + (long)c0 * lda, lda, DB, DB);
            TTAD_GLOBAL TS* C = Arow + (long)q * DB * lda;
#pragma unroll
            for (int a = 0; a < G3::WM; a++)
#pragma unroll
                for (int b = 0; b < G3::WN; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int rr = G3::row_of(a, r), cc = G3::col_of(b, r);
                        acc[a][b][r] = rr < pw ? -(double)C[rr + (long)cc * lda] : 0.0;
                    }
            __syncthreads();
            G3::compute(acc, At, lane, wr, wc);
#pragma unroll
            for (int a = 0; a < G3::WM; a++)
#pragma unroll
                for (int b = 0; b < G3::WN; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int rr = G3::row_of(a, r), cc = G3::col_of(b, r);
                        if (rr < pw) C[rr + (long)cc * lda] = (TS)(-acc[a][b][r]);
                    }
        }
    }
    // diagonal block (identity padding): factor and invert in LDS
    double* D = At;
    double* X = D + DB * DP;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < G3::WM; a++)
#pragma unroll
        for (int b = 0; b < G3::WN; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = G3::row_of(a, r), cc = G3::col_of(b, r);
                D[rr + cc * DP] = (rr < pw && cc < pw) ? (rr >= cc ? -accd[a][b][r] : 0.0) : (rr == cc ? 1.0 : 0.0);
            }
    __syncthreads();
    const int bad = lds_potrf_inv64(D, X, pw);
    if (bad && tid == 0) atomicCAS(info, 0, col_off + p0 + bad);
    {
        constexpr int PER = DB * DB / NTHREADS;
        const int i = tid % DB, j0 = tid / DB;
        TTAD_GLOBAL double* Xg = gmem(dinv) + (long)rb * DB * DB;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int j = j0 + q * (NTHREADS / DB);
            const double dv = D[i + j * DP], xv = X[i + j * DP];
            if (i < pw && j < pw && i >= j) Arow[i + (long)(p0 + j) * lda] = (TS)dv;
            Xg[i + j * DB] = xv;
        }
    }
    __syncthreads();
    if (tid == 0) {
        __hip_atomic_store(prog + rb, rb + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        // the last item of the task to finish clears the flags for the next run on this tile
        if (__hip_atomic_fetch_add(prog + POTRF_FIN, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == nrb - 1) {
            for (int i = 0; i < nrb; i++) __hip_atomic_store(prog + i, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(prog + POTRF_FIN, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// TRSM work item: rows [r0, r0+trsm_rows(sizeof(TS))) of X (m x n) := X L^{-T} (Right, Lower,
// Trans, NonUnit), in place, by 64-column blocks: the block's right-hand side goes
// into the accumulators, the already solved columns times L's off-diagonal blocks
// are subtracted (one GEMM), and the result is multiplied by the inverse of L's
// 64x64 diagonal block left by POTRF straight from LDS (no round trip through
// memory between the two products). TC is the MFMA type (FP64, or FP32 for the
// mixed variants).
// ---------------------------------------------------------------------------
// planes (FP32 tiles; null = none): the bf16 split planes of X (ldx rows, plane_granule layout), written
// beside X for the bf16x3 updates that read it; needs n % 32 == 0.
template <typename TS, typename TC>
__device__ __forceinline__ void item_trsm(const TS* __restrict__ L, long ldl, const double* __restrict__ dinv, TS* __restrict__ X,
                          long ldx, int m, int n, int r0, char* lds, void* planes = nullptr) {
    constexpr bool VEC = std::is_same<TS, TC>::value;
    constexpr int BMT = trsm_rows(sizeof(TS));
    using G = GemmNT<TS, TS, TC, BMT, DB, trsm_bk<TS>(), VEC, true>;  // update from the solved columns
    using G3 = GemmNT<TC, TC, TC, BMT, DB, DB, false, true>;  // times Dinv^T, both operands in LDS
    const int mv = min(BMT, m - r0);
    if (mv <= 0) return;
    TS* Xr = X + r0;
    const int nblk = (n + DB - 1) / DB;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    TC* At = reinterpret_cast<TC*>(lds);  // G3 staging: At[k][m] = t(m, k), Bt[k][j] = Dinv(j, k)
    TC* Bt = At + DB * G3::PA;
    typename G::acc_t acc[G::WM][G::WN];
    for (int cb = 0; cb < nblk; cb++) {
        const int c0 = cb * DB, cw = min(DB, n - c0);
        TTAD_GLOBAL TS* T = gmem(Xr) + (long)c0 * ldx;
        // acc = -(right-hand side block) + X[:, :c0] L[c0:c0+cw, :c0]^T = -t
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G::row_of(a, r), cc = G::col_of(b, r);
                    acc[a][b][r] = (rr < mv && cc < cw) ? -(TC)T[rr + (long)cc * ldx] : TC(0);
                }
        if (c0 > 0) G::run(acc, Xr, ldx, mv, L + c0, ldl, cw, c0, lds);
        else __syncthreads();  // the previous block's products have left the staging area
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) At[G::col_of(b, r) * G3::PA + G::row_of(a, r)] = -acc[a][b][r];
        // Dinv block into Bt, every global load in flight before the first LDS store (a
        // load-store loop would wait for each load in turn)
        {
            constexpr int PER = DB * DB / (2 * NTHREADS);
            using d2 = V16<double>::t;
            const TTAD_GLOBAL d2* Dv = reinterpret_cast<const TTAD_GLOBAL d2*>(gmem(dinv) + (long)cb * DB * DB);
            d2 v[PER];  // Dv[i + j*DB] = Dinv(i, j), two rows of a column per element
#pragma unroll
            for (int q = 0; q < PER; q++) v[q] = Dv[threadIdx.x + q * NTHREADS];
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int e = 2 * (threadIdx.x + q * NTHREADS);
                TC* d = Bt + (e / DB) * G3::PB + (e % DB);
                d[0] = (TC)v[q].x;
                d[1] = (TC)v[q].y;
            }
        }
        __syncthreads();
        G3::zero(acc);
        G3::compute(acc, At, lane, wr, wc);
#pragma unroll
        for (int a = 0; a < G::WM; a++)
#pragma unroll
            for (int b = 0; b < G::WN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = G3::row_of(a, r), cc = G3::col_of(b, r);
                    if (rr < mv && cc < cw) T[rr + (long)cc * ldx] = (TS)acc[a][b][r];
                }
        if constexpr (std::is_same<TS, float>::value && std::is_same<TC, float>::value) {
            // a lane holds 4 consecutive columns of one row per fragment (cw is a multiple of 32 here)
            if (planes) {
#pragma unroll
                for (int a = 0; a < G::WM; a++)
#pragma unroll
                    for (int b = 0; b < G::WN; b++) {
                        const int rr = G3::row_of(a, 0), cc = G3::col_of(b, 0);
                        if (rr < mv && cc < cw)
                            store_split4(static_cast<unsigned short*>(planes), (int)ldx, r0 + rr, c0 + cc, acc[a][b]);
                    }
            }
        }
        __syncthreads();
    }
}

// LDS bytes each item kind needs (the executor reserves the maximum).
constexpr int LDS_GEMM_F64 = GemmNT<double, double, double, GEMM_BM, GEMM_BM, gemm_bk<double>(), true, true>::LDS_BYTES;
constexpr int LDS_GEMM_F32 = GemmNT<float, float, float, GEMM_BM, GEMM_BM, gemm_bk<float>(), true, true>::LDS_BYTES;
constexpr int LDS_GEMM_BF16 = GemmNTbf16<float, GEMM_BM, GEMM_BM, true, true>::LDS_BYTES;
constexpr int LDS_GEMM_BF16_1 = GemmNTbf16<float, GEMM_BM, GEMM_BM, false, true>::LDS_BYTES;
// LDS of one GEMM item of the given storage / precision (the standalone kernels launch with it).
template <typename TS, int PREC> constexpr int gemm_item_lds() {
    return PREC == P_BF16X3 ? LDS_GEMM_BF16 : PREC == P_BF16 ? LDS_GEMM_BF16_1 : sizeof(TS) == 8 ? LDS_GEMM_F64 : LDS_GEMM_F32;
}
constexpr int LDS_POTRF = PotrfGemm<double>::G::LDS_BYTES > 2 * DB * DP * 8 ? PotrfGemm<double>::G::LDS_BYTES
                                                                             : 2 * DB * DP * 8;
constexpr int LDS_POTRF_ROWS = GemmNT<double, double, double, DB, DB, DB, false, true>::STAGE * 8;
constexpr int LDS_TRSM = cmax(GemmNT<double, double, double, TRSM_BM, DB, 32, true, true>::LDS_BYTES,
                              GemmNT<double, double, double, TRSM_BM, DB, DB, false, true>::STAGE * 8);
constexpr int LDS_TRSM_F32 = cmax(GemmNT<float, float, float, TRSM_BM_F32, DB, trsm_bk<float>(), true, true>::LDS_BYTES,
                                  GemmNT<float, float, float, TRSM_BM_F32, DB, DB, false, true>::STAGE * 4);
constexpr int LDS_ITEM_MAX =
    cmax(cmax(cmax(LDS_GEMM_F64, LDS_GEMM_F32), cmax(LDS_GEMM_BF16, LDS_POTRF)), cmax(LDS_TRSM, LDS_POTRF_ROWS));

}  // namespace ttad
