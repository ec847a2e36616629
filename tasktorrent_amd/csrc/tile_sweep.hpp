// Triangular sweeps with a factor in FP32 tiles, chained over tiles with a one-tile lookahead.
//
// The 64-row block sweeps (item_sweep) chain every block on its predecessor: at n = 32768 that is 512
// dependent steps per sweep, each a counter round trip, a load of the predecessor's block and two small
// products (about 7 us). Here the chain runs over tiles (64 steps at b = 512), and each step is a single
// counter round trip followed by one product whose matrix operand is already in LDS.
//
// Forward sweep L y = r, with Tinv_t = L_tt^{-1} and r'_t = r_t - sum_{u<t} L_tu y_u, so that y_t = Tinv_t r'_t.
// The term of the previous tile is written through r'_{t-1}: L_{t,t-1} y_{t-1} = M_t r'_{t-1} with
// M_t = L_{t,t-1} Tinv_{t-1}, formed once per factor. Workgroup (t, s) owns the 64 rows s of tile t:
//   A. acc = sum_{u <= t-2} L_tu[s, :] y_u, each tile u as soon as all its blocks of y are final;
//   B. acc += M_t[s, :] r'_{t-1} once every block of r'_{t-1} is published (M_t[s, :] staged in LDS ahead);
//   2. r'_t[s] = r_t[s] - acc is published (tile counter of r');
//   C. once all of r'_t is published, y_t[s] = Tinv_t[s, :] r'_t (Tinv_t[s, :] loaded at the start, by the
//      workgroup's other four waves).
// So the chain r'_{t-1} -> r'_t costs one counter round trip and one 64 x b x 16 product; y_t follows off
// the chain. The backward sweep L^T x = y runs the same way from the last tile, with U_t = Tinv_t^T,
// KT_t = (Tinv_{t+1} L_{t+1,t})^T, x_t = U_t r''_t and X += x.
// Products with explicit inverses are backward stable for well-conditioned diagonal tiles and lose accuracy
// like cond(L_tt) otherwise; the FP64 refinement that follows every solve with an FP32 factor absorbs it.
//
// Right-hand sides: 16 per sweep; Y, D and r' are row-major n x 16 FP32 (one 64-byte row per matrix row),
// so that the rows of a tile load as whole 256-byte lines. Tiles of at most 512 rows.
#pragma once
#include "solve_ops.hpp"

namespace ttad {

constexpr int LS_NR = 16;               // right-hand sides per sweep
constexpr int LS_BMAX = 512;            // largest tile
constexpr int LS_KS = LS_BMAX / 4;      // MFMA k-steps of a tile-wide product
constexpr int LS_LDK = LS_BMAX + 4;     // LDS pitch of a staged 64 x b operand (row-major, k contiguous)
constexpr int LSWEEP_LDS = SB * LS_LDK * 4;
constexpr int TILE_INV_LDS = 3 * DB * DP * (int)sizeof(double);

struct LSweepTab {
    int32_t P;                // 64-row blocks
    int32_t nt;               // tiles
    const int32_t* blk_tile;  // [P]
    const int32_t* blk_off;   // [P]
    const int32_t* tfirst;    // [nt + 1]: first block of each tile
    float* const* tinv;       // [nt] Tinv_t (b_t x b_t, column-major, upper part zero)
    float* const* uinv;       // [nt] U_t = Tinv_t^T
    float* const* mfac;       // [nt] M_t = L_{t,t-1} Tinv_{t-1} (b_t x b_{t-1}); t >= 1
    float* const* kfac;       // [nt] KT_t = (Tinv_{t+1} L_{t+1,t})^T (b_t x b_{t+1}); t <= nt - 2
    float* const* tfrag;      // [nt] Tinv_t in MFMA A-fragment order (fragment_index), 512 x 512 zero-padded
    float* const* ufrag;      // [nt] U_t likewise
};

// Order of k in the tile-wide products: MFMA k-step ks of lane group kl (lane / 16) takes k = kperm(ks, kl), so
// that the four k-steps 4 j .. 4 j + 3 of a lane cover four consecutive k (16 j + 4 kl + 0..3). A product sums
// over all k, so any order that both operands follow gives it; this one lets a lane load four k-steps of a
// k-contiguous operand (a transposed factor block, LDS rows) with one 16-byte access.
__host__ __device__ inline int kperm(int ks, int kl) { return ((ks >> 2) << 4) + (kl << 2) + (ks & 3); }

// Position of element (r, c) of a 512 x 512 operand stored in MFMA A-fragment order: for the 64-row block r / 64,
// wave (r % 64) / 16 and lane 16 kl + r % 16 with c = 16 j + 4 kl + q, the k-steps 4 j + q (q = 0..3) together
// (16 bytes).
__host__ __device__ inline long fragment_index(int r, int c) {
    const int j = c >> 4, kl = (c >> 2) & 3, q = c & 3, lane = (kl << 4) + (r & 15);
    return ((((long)(r >> 6) * 4 + ((r >> 4) & 3)) * (LS_KS / 4) + j) * 64 + lane) * 4 + q;
}

// ---------------------------------------------------------------------------
// Once per factor: Tinv_t and U_t = Tinv_t^T of every diagonal tile (workgroup (t, J): block column J of
// Tinv_t, top down: X_JJ = inverse of the 64 x 64 diagonal block that POTRF left, X_IJ = -X_II sum_{Q=J}^{I-1}
// L_IQ X_QJ), with zeros above (Tinv) / below (U) the diagonal blocks. FP64 arithmetic, stored in TL.
// ---------------------------------------------------------------------------
template <typename TL>
__global__ __launch_bounds__(NTHREADS) void k_tile_inv(const DevMat* Mp, TL* const* tinv, TL* const* uinv,
                                                       TL* const* tfrag, TL* const* ufrag) {
    extern __shared__ __align__(16) char lds[];
    const DevMat& M = *Mp;
    const int t = blockIdx.x, J = blockIdx.y;
    const int bt = M.bsz[t];
    const int nbt = (bt + SB - 1) / SB;
    if (J >= nbt) return;
    double* La = reinterpret_cast<double*>(lds);
    double* Xa = La + DB * DP;
    double* Ta = Xa + DB * DP;
    const TTAD_GLOBAL TL* L = gmem((const TL*)M.tiles[t + (long)t * M.nt]);
    const TTAD_GLOBAL double* dv = gmem((const double*)M.dinv[t]);
    TTAD_GLOBAL TL* X = gmem(tinv[t]);
    TTAD_GLOBAL TL* U = gmem(uinv[t]);
    TTAD_GLOBAL TL* XF = gmem(tfrag[t]);
    TTAD_GLOBAL TL* UF = gmem(ufrag[t]);
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    auto rows = [&](int I) { return min(SB, bt - I * SB); };
    const int wJ = rows(J);
    auto put = [&](int r, int c, double v) {  // Tinv(r, c) and U(c, r), column-major and in fragment order
        X[r + (long)c * bt] = (TL)v;
        U[c + (long)r * bt] = (TL)v;
        XF[fragment_index(r, c)] = (TL)v;
        UF[fragment_index(c, r)] = (TL)v;
    };
    for (int e = tid; e < J * SB * wJ; e += NTHREADS) put(e % (J * SB), J * SB + e / (J * SB), 0.0);
    for (int e = tid; e < DB * DB; e += NTHREADS) {
        const int i = e % DB, j = e / DB;
        if (i < wJ && j < wJ) put(J * SB + i, J * SB + j, dv[(long)J * DB * DB + i + j * DB]);
    }
    // each wave's stores reach the L2 before other waves read them back (vmcnt(0) expcnt(7) lgkmcnt(15))
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    for (int I = J + 1; I < nbt; I++) {
        const int wI = rows(I);
        d4_t acc[4];
#pragma unroll
        for (int c = 0; c < 4; c++) acc[c] = d4_t{0, 0, 0, 0};
        for (int Q = J; Q < I; Q++) {
            const int wQ = rows(Q);
            for (int e = tid; e < DB * DB; e += NTHREADS) {
                const int i = e % DB, j = e / DB;
                La[i + j * DP] = (i < wI && j < wQ) ? (double)L[(I * SB + i) + (long)(Q * SB + j) * bt] : 0.0;
                Xa[i + j * DP] = (i < wQ && j < wJ) ? (double)X[(Q * SB + i) + (long)(J * SB + j) * bt] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int q = wv * 4 + c, bi = q & 3, bj = q >> 2;
#pragma unroll
                for (int K = 0; K < 4; K++) acc[c] = mma16(acc[c], La + 16 * bi + 16 * K * DP, Xa + 16 * K + 16 * bj * DP, false);
            }
            __syncthreads();
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int q = wv * 4 + c, bi = q & 3, bj = q >> 2;
            put16(Ta + 16 * bi + 16 * bj * DP, acc[c], 1.0, false);
        }
        for (int e = tid; e < DB * DB; e += NTHREADS) {
            const int i = e % DB, j = e / DB;
            La[i + j * DP] = (i < wI && j < wI) ? dv[(long)I * DB * DB + i + j * DB] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int q = wv * 4 + c, bi = q & 3, bj = q >> 2;
            d4_t o = {0, 0, 0, 0};
#pragma unroll
            for (int K = 0; K < 4; K++) o = mma16(o, La + 16 * bi + 16 * K * DP, Ta + 16 * K + 16 * bj * DP, false);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * bi + (lane >> 4) + 4 * r, col = 16 * bj + (lane & 15);
                if (row < wI && col < wJ) put(I * SB + row, J * SB + col, -o[r]);
            }
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
    }
}

// LT_t = L_{t+1,t}^T (b_t x b_{t+1}, column-major, ld b_t) for t = 0 .. nt-2: workgroup (t, 64 x 64 block).
template <typename TL>
__global__ __launch_bounds__(NTHREADS) void k_sub_transpose(const DevMat* Mp, TL* const* lt) {
    __shared__ TL S[SB][SB + 1];
    const DevMat& M = *Mp;
    const int t = blockIdx.x;
    const int m = M.bsz[t + 1], n = M.bsz[t];  // L_{t+1,t}: m x n, ld m
    const int bi = blockIdx.y % ((m + SB - 1) / SB), bj = blockIdx.y / ((m + SB - 1) / SB);
    if (bj * SB >= n) return;
    const TTAD_GLOBAL TL* L = gmem((const TL*)M.tiles[(t + 1) + (long)t * M.nt]);
    TTAD_GLOBAL TL* O = gmem(lt[t]);
    for (int e = threadIdx.x; e < SB * SB; e += NTHREADS) {
        const int i = e % SB, j = e / SB, r = bi * SB + i, c = bj * SB + j;
        S[j][i] = (r < m && c < n) ? L[r + (long)c * m] : TL(0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < SB * SB; e += NTHREADS) {
        const int i = e % SB, j = e / SB, r = bj * SB + i, c = bi * SB + j;  // O(r, c) = L(c, r)
        if (r < n && c < m) O[r + (long)c * n] = S[i][j];
    }
}

// Batched C_q = A_q B_q^T (beta = 0) over 128 x 128 output items: the M_t and KT_t products.
struct PairGemm {
    const float* A;
    const float* B;
    float* C;
    int32_t lda, ldb, ldc, m, n, K;
    int32_t item0;  // first item of this product
};
__global__ __launch_bounds__(NTHREADS, 2) void k_pair_gemm(const PairGemm* prods, int n_prods) {
    extern __shared__ __align__(16) char lds[];
    int lo = 0, hi = n_prods - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prods[mid].item0 <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const PairGemm g = prods[lo];
    const int it = blockIdx.x - g.item0, tr = (g.m + GEMM_BM - 1) / GEMM_BM;
    item_gemm<float, P_F32>(g.A, g.lda, g.B, g.ldb, g.C, g.ldc, g.m, g.n, g.K, (it % tr) * GEMM_BM,
                            (it / tr) * GEMM_BM, false, 1.0, 0.0, lds);
}

// ---------------------------------------------------------------------------
// Lookahead sweep
// ---------------------------------------------------------------------------
// A value the whole wave holds alike, read from memory: passed through a scalar register so that branches
// on it are scalar and loads under them are not exec-masked.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Layout of the sweeps' row vectors (Y, D, r'; n x 16 FP32): the 16-row groups of each tile hold, per MFMA lane,
// four consecutive k-steps together, so that a lane loads the B fragments of 4 k-steps with one 16-byte load.
// Element (k, n) of the tile that starts at row g0 (k tile-local, k = 16 j + 4 kl + q = kperm(4 j + q, kl)) lies
// at 16 g0 + ((64 j + 16 kl + n) * 4 + q); tiles start at multiples of 64 rows.
__device__ __forceinline__ long swz(int g0, int k, int n) {
    const int j = k >> 4, kl = (k >> 2) & 3, q = k & 3;
    return (long)g0 * LS_NR + (((j << 6) + (kl << 4) + n) << 2) + q;
}

// B fragments of a tile-wide product from a swizzled row vector: the tile at row g0 with bu rows; rows at or
// beyond bu read as 0. 32 loads of 16 bytes per lane, all in flight together.
__device__ __forceinline__ void load_rows16(float (&b)[LS_KS], const float* Y, int g0, int bu) {
    const int lane = threadIdx.x & 63;
    typedef float f4a_t __attribute__((ext_vector_type(4)));
    const TTAD_GLOBAL f4a_t* p = reinterpret_cast<const TTAD_GLOBAL f4a_t*>(gmem(Y) + (long)g0 * LS_NR) + lane;
    if (bu == LS_BMAX) {  // whole tile (uniform): no condition at all
#pragma unroll
        for (int j = 0; j < LS_KS / 4; j++) {
            const f4a_t v = p[j * 64];
#pragma unroll
            for (int q = 0; q < 4; q++) b[4 * j + q] = v[q];
        }
        return;
    }
    const int kl = lane >> 4;
#pragma unroll
    for (int j = 0; j < LS_KS / 4; j++) {  // the buffer is padded past the last tile: these loads stay in it
        const f4a_t v = p[j * 64];
#pragma unroll
        for (int q = 0; q < 4; q++) b[4 * j + q] = 16 * j + 4 * kl + q < bu ? v[q] : 0.0f;
    }
}

// As[m][k] = A(m, k) for the w rows and bk columns of a column-major operand (ld), zero elsewhere in the
// 64 x LS_BMAX window. Thread tid loads rows 4 (tid % 16) .. +3 of every 16th column, all loads first, from
// clamped addresses (values outside selected to 0).
__device__ __forceinline__ void stage_rows(float* As, const float* A, long ld, int w, int bk) {
    const int r4 = (threadIdx.x & 15) * 4, c0 = threadIdx.x >> 4;
    typedef float f4u_t __attribute__((ext_vector_type(4), aligned(4)));
    f4_t v[LS_BMAX / 16];
    const TTAD_GLOBAL float* Ag = gmem(A);
    if (w == SB && bk == LS_BMAX) {  // whole 64 x 512 window (uniform): one 16-byte load per column
#pragma unroll
        for (int q = 0; q < LS_BMAX / 16; q++) {
            const f4u_t x = *reinterpret_cast<const TTAD_GLOBAL f4u_t*>(Ag + r4 + (long)(c0 + 16 * q) * ld);
            v[q] = f4_t{x[0], x[1], x[2], x[3]};
        }
    } else {
#pragma unroll
        for (int q = 0; q < LS_BMAX / 16; q++) {
            const int c = c0 + 16 * q;
            const long cc = (long)min(c, bk - 1) * ld;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float x = Ag[min(r4 + i, w - 1) + cc];
                v[q][i] = (c < bk && r4 + i < w) ? x : 0.0f;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < LS_BMAX / 16; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) As[(r4 + i) * LS_LDK + c0 + 16 * q] = v[q][i];
}

// acc += As(64 x LS_BMAX) * B(LS_BMAX x 16), A from LDS (zero beyond the operand), B fragments in registers
// (zero beyond the tile); wave w owns rows 16w .. 16w + 15. Four k-steps per 16-byte LDS read (kperm); four
// accumulators shorten the chain of dependent MFMAs.
__device__ __forceinline__ void lds_tile_mma(f4_t& acc, const float* As, const float (&b)[LS_KS]) {
    const int lane = threadIdx.x & 63;
    const float* a = As + ((threadIdx.x >> 6) * 16 + (lane & 15)) * LS_LDK + 4 * (lane >> 4);
    f4_t c[4] = {f4_t{0, 0, 0, 0}, f4_t{0, 0, 0, 0}, f4_t{0, 0, 0, 0}, f4_t{0, 0, 0, 0}};
#pragma unroll
    for (int j = 0; j < LS_KS / 4; j++) {
        const f4_t v = *reinterpret_cast<const f4_t*>(a + 16 * j);
#pragma unroll
        for (int q = 0; q < 4; q++) c[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(v[q], b[4 * j + q], c[q], 0, 0, 0);
    }
    acc += (c[0] + c[1]) + (c[2] + c[3]);
}

// B fragments of block c (64 k) of a tile-wide product from a swizzled row vector: the tile at row g0 with bu
// rows; rows at or beyond bu read as 0. Four loads of 16 bytes per lane (k-steps 16 c .. 16 c + 15).
__device__ __forceinline__ void load_rows_blk(float (&b)[KS], const float* Y, int g0, int bu, int c) {
    const int lane = threadIdx.x & 63;
    typedef float f4a_t __attribute__((ext_vector_type(4)));
    const TTAD_GLOBAL f4a_t* p = reinterpret_cast<const TTAD_GLOBAL f4a_t*>(gmem(Y) + (long)g0 * LS_NR) + lane;
    if (bu == LS_BMAX) {  // whole tile (uniform): no condition at all
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const f4a_t v = p[(4 * c + jj) * 64];
#pragma unroll
            for (int q = 0; q < 4; q++) b[4 * jj + q] = v[q];
        }
        return;
    }
    const int kl = lane >> 4;
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {  // the buffer is padded past the last tile: these loads stay in it
        const int j = 4 * c + jj;
        const f4a_t v = p[j * 64];
#pragma unroll
        for (int q = 0; q < 4; q++) b[4 * jj + q] = 16 * j + 4 * kl + q < bu ? v[q] : 0.0f;
    }
}

// The far-tile stream's loads: whole 64 x 64 blocks and whole 64-row blocks of the right-hand sides, with no
// condition at all, so that the stream is straight-line code and each product waits for its own slot only.
// Row m of a mode-0 block is clamped by the caller (rows beyond the workgroup's block are computed and dropped).
__device__ __forceinline__ void load_a_whole(float (&a)[KS], const float* Tp, long ld, int row, int oq, int mode) {
    const int kg = (threadIdx.x & 63) >> 4;
    const TTAD_GLOBAL float* T = gmem(Tp);
    if (mode == 0) {
        const TTAD_GLOBAL float* P = T + row + (long)oq * ld;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) a[ks] = P[(long)kperm(ks, kg) * ld];
    } else {
        typedef float f4u_t __attribute__((ext_vector_type(4), aligned(4)));
        const TTAD_GLOBAL float* P = T + oq + (long)row * ld;
#pragma unroll
        for (int j = 0; j < KS / 4; j++) {
            const f4u_t v = *reinterpret_cast<const TTAD_GLOBAL f4u_t*>(P + 16 * j + 4 * kg);
#pragma unroll
            for (int q = 0; q < 4; q++) a[4 * j + q] = v[q];
        }
    }
}
__device__ __forceinline__ void load_rows_whole(float (&b)[KS], const float* Y, int g0, int c) {
    typedef float f4a_t __attribute__((ext_vector_type(4)));
    const TTAD_GLOBAL f4a_t* p =
        reinterpret_cast<const TTAD_GLOBAL f4a_t*>(gmem(Y) + (long)g0 * LS_NR) + (threadIdx.x & 63);
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const f4a_t v = p[(4 * c + jj) * 64];
#pragma unroll
        for (int q = 0; q < 4; q++) b[4 * jj + q] = v[q];
    }
}

// A fragments of wave w for one 64 x 64 block of a factor tile T (column-major, ld) with k = kperm(ks, lane / 16):
// mode 0 reads rows op + m and columns oq + k, mode 1 the transposed block (rows oq + k, columns op + m; k
// contiguous in memory, four k-steps per 16-byte load). Zero beyond wp rows (m) and wq columns (k), from clamped
// addresses.
__device__ __forceinline__ void load_a_blk(float (&a)[KS], const float* Tp, long ld, int op, int oq, int wp, int wq,
                                           int mode) {
    const int lane = threadIdx.x & 63;
    const int m = (threadIdx.x >> 6) * 16 + (lane & 15), kg = lane >> 4;
    const TTAD_GLOBAL float* T = gmem(Tp);
    if (wp == SB && wq == SB) {  // whole block (uniform): no condition at all
        if (mode == 0) {
            const TTAD_GLOBAL float* P = T + (op + m) + (long)oq * ld;
#pragma unroll
            for (int ks = 0; ks < KS; ks++) a[ks] = P[(long)kperm(ks, kg) * ld];
        } else {
            typedef float f4u_t __attribute__((ext_vector_type(4), aligned(4)));
            const TTAD_GLOBAL float* P = T + oq + (long)(op + m) * ld;
#pragma unroll
            for (int j = 0; j < KS / 4; j++) {
                const f4u_t v = *reinterpret_cast<const TTAD_GLOBAL f4u_t*>(P + 16 * j + 4 * kg);
#pragma unroll
                for (int q = 0; q < 4; q++) a[4 * j + q] = v[q];
            }
        }
        return;
    }
    const int mc = min(m, wp - 1);
    if (mode == 0) {
        const TTAD_GLOBAL float* P = T + (op + mc) + (long)oq * ld;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int k = kperm(ks, kg);
            const float v = P[(long)min(k, wq - 1) * ld];
            a[ks] = (m < wp && k < wq) ? v : 0.0f;
        }
    } else {
        const TTAD_GLOBAL float* P = T + oq + (long)(op + mc) * ld;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int k = kperm(ks, kg);
            const float v = P[min(k, wq - 1)];
            a[ks] = (m < wp && k < wq) ? v : 0.0f;
        }
    }
}

// A far tile of one workgroup, looked up once per sweep: its factor tile (L_tu forward, L_ut backward), its
// rows and its first row.
struct FarTile {
    const float* T;
    int32_t rows;
    int32_t row0;
};
constexpr int LS_MAXT = 1024;  // most tiles of a lookahead sweep (the far-tile table: 16 KB of LDS)

// Waits (thread 0 polls, the barrier passes it on) until *cnt >= want.
__device__ __forceinline__ void ls_wait(const int32_t* cnt, int want, int32_t* err, int64_t spin_limit, bool slow) {
    if (threadIdx.x == 0) (void)wait_count(cnt, want, err, spin_limit, slow);
    __syncthreads();
}

// One 64-row block p of a lookahead sweep (see the top of this file). flags[t]: blocks of tile t whose r'
// is published; flags[nt + t]: blocks of tile t whose output (y or x) is final.
//
// The far products stream 64 x 64 blocks of the factor and 64-row blocks of the right-hand sides through a
// ring of LS_RING register slots: the loads of stream block i + LS_RING go out right after the products of
// block i, so LS_RING - 1 blocks are in flight behind the products, across far tiles too, and no load in the
// stream depends on a condition (a wait before a product covers that product's slot only). Each wave polls the
// far tiles' counters itself, up to 64 tiles at a time, and acquires once per batch of final tiles. The last
// far tile (the one the chain waits for) comes after the stream, with its factor blocks loaded before its wait.
// All sums run in the same order as a plain loop over far tiles, blocks and k-steps, so the results do not
// depend on timing.
//
// The workgroup has eight waves. Waves 0-3 run steps A, B and 2, then wait for all of r'_t (thread 0 polls
// with no load of its own outstanding) and end; waves 4-7 load step C's operand (Tinv_t or U_t rows, 128 KB)
// at the start, pass the same barriers and run step C. So the operand's loads are off the chain and the poll
// that starts step C is not queued behind them (a wave's loads return in order). The two groups take the same
// barriers in the same order; step C's last one is taken by waves 4-7 alone (a barrier waits for the waves of
// the workgroup that have not ended).
constexpr int LS_RING = 4;
constexpr int LS_THREADS = 2 * NTHREADS;
template <int FWD>
__device__ __forceinline__ void item_lsweep(const DevMat& M, const LSweepTab& lt, int p, const double* R, long ldr,
                                            float* Y, float* D, double* X, long ldx, float* Rp, int nrhs,
                                            int32_t* flags, int32_t* err, int64_t spin_limit, FarTile* ft, float* As) {
    constexpr bool forward = FWD != 0;
    constexpr int mode = forward ? 0 : 1;
    constexpr int NBT = LS_BMAX / SB;
    static_assert(NBT % LS_RING == 0, "a tile's blocks fill the ring a whole number of times");
    const int nt = lt.nt;
    const int t = uni(lt.blk_tile[p]), boff_p = uni(lt.blk_off[p]);
    const int bt = uni(M.bsz[t]), nbt = uni(lt.tfirst[t + 1]) - uni(lt.tfirst[t]);
    const int tg0 = uni(M.boff[t]), g0 = tg0 + boff_p, w = min(SB, bt - boff_p);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const float* Src = forward ? Y : D;
    int32_t* rcnt = flags;
    int32_t* ocnt = flags + nt;
    const bool has_near = forward ? t >= 1 : t + 1 < nt;
    const int tn = forward ? t - 1 : t + 1;

    if (threadIdx.x >= NTHREADS) {  // waves 4-7: step C
        const int wc = wv - 4;      // the wave's 16 rows of the block
        float at[NBT][KS], rb[LS_RING][KS];
        {
            // C's operand: Tinv_t[s, :] (forward) or U_t[s, :] (backward), stored in fragment order (k_tile_inv)
            typedef float f4a_t __attribute__((ext_vector_type(4)));
            const TTAD_GLOBAL f4a_t* T =
                reinterpret_cast<const TTAD_GLOBAL f4a_t*>(gmem(forward ? lt.tfrag[t] : lt.ufrag[t])) +
                (long)(boff_p / SB * 4 + wc) * (LS_KS / 4) * 64 + lane;
#pragma unroll
            for (int j = 0; j < LS_KS / 4; j++) {
                const f4a_t v = T[j * 64];
#pragma unroll
                for (int q = 0; q < 4; q++) at[j / 4][(j % 4) * 4 + q] = v[q];
            }
        }
        __syncthreads();                 // the far-tile table
        if (has_near) __syncthreads();  // step B's wait
        __syncthreads();                 // step 2
        __syncthreads();                 // all of r'_t published (waves 0-3 poll)
#pragma unroll
        for (int c = 0; c < LS_RING; c++) load_rows_blk(rb[c], Rp, tg0, bt, c);
        f4_t acc = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < NBT; c++) {
#pragma unroll
            for (int k = 0; k < KS; k++)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(at[c][k], rb[c % LS_RING][k], acc, 0, 0, 0);
            if (c + LS_RING < NBT) load_rows_blk(rb[c % LS_RING], Rp, tg0, bt, c + LS_RING);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = wc * 16 + Mfma<float>::row(lane, r), n = lane & 15;
            if (m < w) {
                const long gi = g0 + m;
                if (forward) {
                    gmem(Y)[swz(tg0, boff_p + m, n)] = acc[r];
                } else {
                    gmem(D)[swz(tg0, boff_p + m, n)] = acc[r];
                    if (n < nrhs) gmem(X)[gi + (long)n * ldx] += (double)acc[r];
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();  // waves 4-7 only: waves 0-3 have ended
        if (threadIdx.x == NTHREADS) __hip_atomic_fetch_add(ocnt + t, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }

    // this block's right-hand side (MFMA result layout), off the chain
    float rhs[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int m = wv * 16 + Mfma<float>::row(lane, r), n = lane & 15;
        rhs[r] = 0.0f;
        if (m < w && n < nrhs) rhs[r] = forward ? (float)gmem(R)[(g0 + m) + (long)n * ldr] : gmem(Y)[swz(tg0, boff_p + m, n)];
    }
    // the lookahead operand, staged while the far tiles are processed
    const int btn = has_near ? uni(M.bsz[tn]) : 0, nbn = has_near ? uni(lt.tfirst[tn + 1]) - uni(lt.tfirst[tn]) : 0;
    if (has_near) stage_rows(As, (forward ? lt.mfac[t] : lt.kfac[t]) + boff_p, bt, w, btn);
    f4_t acc = {0, 0, 0, 0};
    float ra[LS_RING][KS], rb[LS_RING][KS];  // the ring: factor blocks and right-hand-side blocks

    // A. far tiles (t - 2 and before in sweep order), each once all its outputs are final. Their geometry goes
    // to LDS first, so that the stream reads no table from global memory (such a load's wait would drain it).
    const int nfar = forward ? max(0, t - 1) : max(0, nt - t - 2);
    auto far_u = [&](int f) { return forward ? f : nt - 1 - f; };
    for (int f = threadIdx.x; f < nfar; f += NTHREADS) {  // waves 0-3
        const int u = far_u(f);
        ft[f] = FarTile{(const float*)M.tiles[forward ? t + (long)u * nt : u + (long)t * nt], M.bsz[u], M.boff[u]};
    }
    __syncthreads();
    int ready = 0;  // far tiles f < ready are final and acquired by this wave
    auto make_ready = [&](int f) {
        if (f < ready) return;
        const int fi = f + lane;
        bool fin = false;
        if (fi < nfar)
            fin = __hip_atomic_load(ocnt + far_u(fi), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                  (ft[fi].rows + SB - 1) / SB;
        const unsigned long long notfin = ~(unsigned long long)__ballot(fin);
        const int run = notfin == 0 ? 64 : (int)__builtin_ctzll(notfin);  // final tiles from f on, in order
        if (run > 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            ready = f + run;
        } else {
            (void)wait_count(ocnt + far_u(f), (uni(ft[f].rows) + SB - 1) / SB, err, spin_limit, f + 1 < nfar);
            ready = f + 1;
        }
    };
    const int arow = boff_p + min(wv * 16 + (lane & 15), w - 1);  // this lane's row of the workgroup's block
    // Stream block (far tile f, block c). Past the end of a tile or of the stream it loads a block of the stream
    // again (the products skip it): every slot is refilled by the same loads on every path, so that the wait
    // before a product covers that slot's loads only.
    // The stream covers far tiles f0 .. f_end - 1; the last far tile (t - 2 in sweep order, the one on the chain)
    // comes after it, with its factor blocks loaded before its wait.
    const int f_end = nfar - 1;
    auto issue = [&](float (&a)[KS], float (&b)[KS], int f, int c) {
        const int fc = min(f, f_end - 1);
        const int rows = uni(ft[fc].rows);
        const int cc = min(c, rows / SB - 1);
        load_a_whole(a, wave_uniform(ft[fc].T), forward ? bt : rows, arow, SB * cc, mode);
        load_rows_whole(b, Src, uni(ft[fc].row0), cc);
    };
    int f0 = 0;
    if (nfar > 0 && (uni(ft[0].rows) & (SB - 1)) != 0) {
        // the last tile of the matrix with a partial last block (the first far tile of a backward sweep): block
        // by block, zero beyond its rows, before the stream
        make_ready(0);
        const int rows = uni(ft[0].rows), row0 = uni(ft[0].row0);
        const float* T = wave_uniform(ft[0].T);
        for (int c = 0; c * SB < rows; c++) {
            load_a_blk(ra[0], T, forward ? bt : rows, boff_p, SB * c, w, min(SB, rows - SB * c), mode);
            load_rows_blk(rb[0], Src, row0, rows, c);
#pragma unroll
            for (int k = 0; k < KS; k++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ra[0][k], rb[0][k], acc, 0, 0, 0);
        }
        f0 = 1;
    }
    if (f0 < f_end) {
        make_ready(f0);
#pragma unroll
        for (int c = 0; c < LS_RING; c++) issue(ra[c], rb[c], f0, c);
    }
    for (int f = f0; f < f_end; f++) {
        const int rows = uni(ft[f].rows);
#pragma unroll
        for (int c = 0; c < NBT; c++) {
            if (c * SB < rows) {
#pragma unroll
                for (int k = 0; k < KS; k++)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ra[c % LS_RING][k], rb[c % LS_RING][k], acc, 0, 0, 0);
            }
            if (c + LS_RING < NBT) {
                issue(ra[c % LS_RING], rb[c % LS_RING], f, c + LS_RING);
            } else {
                if (c + LS_RING == NBT && f + 1 < f_end) make_ready(f + 1);
                issue(ra[c % LS_RING], rb[c % LS_RING], f + 1, c + LS_RING - NBT);
            }
        }
    }
    if (f_end >= f0) {  // a whole tile (a partial one is the first far tile, done above)
        const int rows = uni(ft[f_end].rows), row0 = uni(ft[f_end].row0);
        const float* T = wave_uniform(ft[f_end].T);
        float al[NBT][KS];
#pragma unroll
        for (int c = 0; c < NBT; c++)
            if (c * SB < rows) load_a_whole(al[c], T, forward ? bt : rows, arow, SB * c, mode);
        make_ready(f_end);
#pragma unroll
        for (int c = 0; c < LS_RING; c++)
            if (c * SB < rows) load_rows_whole(rb[c], Src, row0, c);
#pragma unroll
        for (int c = 0; c < NBT; c++) {
            if (c * SB < rows) {
#pragma unroll
                for (int k = 0; k < KS; k++)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(al[c][k], rb[c % LS_RING][k], acc, 0, 0, 0);
            }
            if (c + LS_RING < NBT && (c + LS_RING) * SB < rows) load_rows_whole(rb[c % LS_RING], Src, row0, c + LS_RING);
        }
    }
    // B. the neighbour tile through its published r' (the chain step)
    if (has_near) {
        float b[LS_KS];
        ls_wait(rcnt + tn, nbn, err, spin_limit, false);  // also publishes As
        load_rows16(b, Rp, uni(M.boff[tn]), btn);
        lds_tile_mma(acc, As, b);
    }
    // 2. publish r'[p]
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int m = wv * 16 + Mfma<float>::row(lane, r), n = lane & 15;
        if (m < w) gmem(Rp)[swz(tg0, boff_p + m, n)] = rhs[r] - acc[r];
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the stores are in the L2 before the barrier
    __syncthreads();                     // (also: every wave is done reading As)
    if (threadIdx.x == 0) __hip_atomic_fetch_add(rcnt + t, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    // all of r'_t published: waves 4-7 take step C
    ls_wait(rcnt + t, nbt, err, spin_limit, false);
}

// One workgroup per 64-row block, claimed through a ticket counter in sweep order: every block a workgroup
// waits on (earlier tiles, and the blocks of its own tile) holds a smaller ticket or belongs to its own tile,
// whose blocks hold consecutive tickets, so it is running or done. FWD: forward sweep (else backward).
template <int FWD>
__global__ __launch_bounds__(LS_THREADS) void k_lsweep(const DevMat* Mp, LSweepTab tab, const double* R, long ldr,
                                                     float* Y, float* D, double* X, long ldx, float* Rp, int nrhs,
                                                     int32_t* flags, int32_t* ticket, int32_t* err,
                                                     int64_t spin_limit) {
    extern __shared__ __align__(16) char lds[];
    __shared__ int s_t;
    if (threadIdx.x == 0) s_t = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int q = s_t;
    if (q >= tab.P) return;
    const int p = FWD ? q : tab.P - 1 - q;
    __shared__ FarTile ft[LS_MAXT];
    item_lsweep<FWD>(*Mp, tab, p, R, ldr, Y, D, X, ldx, Rp, nrhs, flags, err, spin_limit, ft,
                     reinterpret_cast<float*>(lds));
}

}  // namespace ttad
