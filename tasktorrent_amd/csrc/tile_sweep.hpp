// Triangular sweeps blocked by factor tile, with the inverses of the diagonal tiles.
//
// The 64-row block sweeps (item_sweep) chain every block on its predecessor: at n = 32768 that is 512
// dependent steps per sweep, each a counter round trip plus a 64 x 64 product. Here the dependency
// chain runs over tiles instead. Forward sweep L y = r, for the 64-row block s of tile t (one
// workgroup each):
//   1. acc = sum over the blocks q of earlier tiles of L[p, q] y[q], as each earlier tile completes;
//   2. r'[p] = r[p] - acc is published, and the tile's counter of published blocks is incremented;
//   3. once every block of tile t has published, y[p] = Tinv_t[s, :] r'_t (Tinv_t = L_tt^{-1}).
// The backward sweep L^T x = y runs the same way over transposed blocks, from the last tile to the
// first, with x[p] = (Tinv_t^T)[s, :] r'_t and X += x. So one tile costs two counter round trips
// instead of one per 64-row block. Tinv_t is formed once per factor (k_tile_inv) from the inverses
// of the 64 x 64 diagonal blocks that POTRF leaves: X_IJ = -X_II sum_{Q=J}^{I-1} L_IQ X_QJ.
// Products with an explicit inverse are backward stable for a well-conditioned diagonal tile and
// lose accuracy like cond(L_tt) otherwise; the refinement that follows every solve absorbs it.
#pragma once
#include "solve_ops.hpp"

namespace ttad {

struct TileSweepTab {
    int32_t P;                // 64-row blocks
    int32_t nt;               // tiles
    const int32_t* blk_tile;  // [P]
    const int32_t* blk_off;   // [P]
    const int32_t* tfirst;    // [nt + 1]: first block of each tile
};

constexpr int TILE_INV_LDS = 3 * DB * DP * (int)sizeof(double);

// Inverse of the diagonal tile t of a factor: Tinv[t] (b_t x b_t, column-major, ld b_t, lower triangular; the
// strict upper part is not written). Workgroup (t, J) forms block column J from the top down. Arithmetic in
// FP64; results stored in the factor's precision and read back for the next blocks of the column.
template <typename TL>
__global__ __launch_bounds__(NTHREADS) void k_tile_inv(const DevMat* Mp, TL* const* tinv) {
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
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    auto rows = [&](int I) { return min(SB, bt - I * SB); };
    const int wJ = rows(J);
    for (int e = tid; e < DB * DB; e += NTHREADS) {
        const int i = e % DB, j = e / DB;
        if (i < wJ && j < wJ) X[(J * SB + i) + (long)(J * SB + j) * bt] = (TL)dv[(long)J * DB * DB + i + j * DB];
    }
    // each wave's stores reach the L2 before the other waves read them back
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
                if (row < wI && col < wJ) X[(I * SB + row) + (long)(J * SB + col) * bt] = (TL)(-o[r]);
            }
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
    }
}

// A operand block of Tinv_t for wave w (rows 16w.. of the 64-row block): trans == false: rows of block s,
// columns of block c; trans == true: A(m, k) = Tinv_t(64c + k, 64s + m). Outside the tile reads as 0.
template <typename TL>
__device__ __forceinline__ void load_tinv(TL (&a)[KS], const TL* T, int bt, int s, int c, bool trans) {
    const int lane = threadIdx.x & 63;
    const int m = (threadIdx.x >> 6) * 16 + (lane & 15), kg = lane >> 4;
    const int r = SB * s + m;
    const TTAD_GLOBAL TL* Tg = gmem(T);
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        const int col = SB * c + 4 * ks + kg;
        const bool ok = r < bt && col < bt;
        a[ks] = ok ? (trans ? Tg[col + (long)r * bt] : Tg[r + (long)col * bt]) : TL(0);
    }
}

// One 64-row block of a tile-blocked sweep (see the top of this file). forward: Y[p] = rows of L^{-1} R;
// backward: D[p] = rows of L^{-T} Y, X[p] += D[p]. Rp holds r' (ld ldp). flags[0] counts finished
// blocks (whole tiles finish together), flags[1 + t] the blocks of tile t that published r'.
template <typename TL, typename TC, int NR>
__device__ __forceinline__ void item_tsweep(const DevMat& M, const TileSweepTab& tt_, TL* const* tinv, int forward,
                                            int p, const double* R, long ldr, TC* Y, long ldy, TC* D, long ldd,
                                            double* X, long ldx, TC* Rp, long ldp, int nrhs, int32_t* flags,
                                            int32_t* err, int64_t spin_limit) {
    constexpr int NB = NR / 16;
    const SweepTab tab{tt_.P, tt_.blk_tile, tt_.blk_off};
    const int P = tt_.P;
    const int t = tt_.blk_tile[p], s = tt_.blk_off[p] / SB;
    const int bt = M.bsz[t], t0 = tt_.tfirst[t], nbt = tt_.tfirst[t + 1] - t0;
    const int g0 = M.boff[t] + tt_.blk_off[p], w = min(SB, bt - tt_.blk_off[p]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const TC* Src = forward ? Y : D;
    const long lds_ = forward ? ldy : ldd;
    const int mode = forward ? 0 : 1;
    typedef typename Mfma<TC>::acc_t acc_t;
    acc_t acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = acc_t{0, 0, 0, 0};

    // this block's right-hand side, off the critical path
    TC rhs[NB][4];
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<TC>::row(lane, r), n = nb * 16 + (lane & 15);
            rhs[nb][r] = TC(0);
            if (m < w && n < nrhs)
                rhs[nb][r] = forward ? (TC)gmem(R)[(g0 + m) + (long)n * ldr] : gmem(Y)[(g0 + m) + (long)n * ldy];
        }

    // ---- phase 1: contributions of the earlier tiles (in sweep order), each once it is complete ----
    const int need = forward ? t0 : P - (t0 + nbt);
    auto qof = [&](int i) { return forward ? i : P - 1 - i; };
    // finished-block count at which predecessor i's tile is complete
    auto want_of = [&](int i) {
        const int u = tt_.blk_tile[qof(i)];
        return forward ? tt_.tfirst[u + 1] : P - tt_.tfirst[u];
    };
    __shared__ int s_seen[2];
    int nwait = 0, seen = 0;
    // (blocks of one tile finish in any order: only whole-tile counts tell that a block is final)
    auto wait_for = [&](int i) {
        const int want = want_of(i);
        if (want <= seen) return;  // uniform
        // blocks of tiles before the previous one wait with slow polls
        const bool far = want < need;
        if (threadIdx.x == 0) s_seen[nwait & 1] = wait_count(flags, want, err, spin_limit, far);
        __syncthreads();
        seen = s_seen[nwait & 1];
        nwait++;
    };
    {
        TL a0[KS], a1[KS];
        TC b0[NB][KS], b1[NB][KS];
        bool h0 = false, h1 = false;
        if (need > 0) load_a<TL>(a0, M, tab, p, qof(0), mode);
        if (need > 1) load_a<TL>(a1, M, tab, p, qof(1), mode);
        auto step = [&](int i, TL (&a)[KS], TC (&b)[NB][KS], bool& h) {
            const int q = qof(i);
            wait_for(i);
            TC bc[NB][KS];
            if (h) {
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
#pragma unroll
                    for (int ks = 0; ks < KS; ks++) bc[nb][ks] = b[nb][ks];
            } else {
                load_b<TC, TC, NR>(bc, Src, lds_, blk_g0(M, tab, q), blk_rows(M, tab, q), nrhs);
            }
            TL ac[KS];
#pragma unroll
            for (int ks = 0; ks < KS; ks++) ac[ks] = a[ks];
            h = false;
            if (i + 2 < need) {
                const int qn = qof(i + 2);
                load_a<TL>(a, M, tab, p, qn, mode);
                if (want_of(i + 2) <= seen) {
                    load_b<TC, TC, NR>(b, Src, lds_, blk_g0(M, tab, qn), blk_rows(M, tab, qn), nrhs);
                    h = true;
                }
            }
            reg_mma<TC, TL, NB>(acc, ac, bc);
        };
        for (int i = 0; i < need; i += 2) {
            step(i, a0, b0, h0);
            if (i + 1 < need) step(i + 1, a1, b1, h1);
        }
    }

    // ---- phase 2: publish r'[p] ----
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<TC>::row(lane, r), n = nb * 16 + (lane & 15);
            if (m < w && n < nrhs) gmem(Rp)[(g0 + m) + (long)n * ldp] = rhs[nb][r] - acc[nb][r];
        }
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(flags + 1 + t, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);

    // ---- phase 3: y[p] = Tinv_t[s, :] r'_t (forward, blocks c <= s) or (Tinv_t^T)[s, :] r'_t (c >= s) ----
    const TL* T = tinv[t];
    const int c_lo = forward ? 0 : s, c_hi = forward ? s : nbt - 1;
    const int nc = c_hi - c_lo + 1;
    TL ta0[KS], ta1[KS];
    load_tinv<TL>(ta0, T, bt, s, c_lo, !forward);  // constant operands: loaded before the wait
    if (nc > 1) load_tinv<TL>(ta1, T, bt, s, c_lo + 1, !forward);
    if (threadIdx.x == 0) (void)wait_count(flags + 1 + t, nbt, err, spin_limit, false);
    __syncthreads();
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = acc_t{0, 0, 0, 0};
    const int tg0 = M.boff[t];
    auto step3 = [&](int c, TL (&a)[KS]) {
        TC bc[NB][KS];
        load_b<TC, TC, NR>(bc, Rp, ldp, tg0 + SB * c, min(SB, bt - SB * c), nrhs);
        TL ac[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ks++) ac[ks] = a[ks];
        if (c + 2 <= c_hi) load_tinv<TL>(a, T, bt, s, c + 2, !forward);
        reg_mma<TC, TL, NB>(acc, ac, bc);
    };
    for (int c = c_lo; c <= c_hi; c += 2) {
        step3(c, ta0);
        if (c + 1 <= c_hi) step3(c + 1, ta1);
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<TC>::row(lane, r), n = nb * 16 + (lane & 15);
            if (m < w && n < nrhs) {
                const long gi = g0 + m;
                if (forward) {
                    gmem(Y)[gi + (long)n * ldy] = acc[nb][r];
                } else {
                    gmem(D)[gi + (long)n * ldd] = acc[nb][r];
                    gmem(X)[gi + (long)n * ldx] += (double)acc[nb][r];
                }
            }
        }
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(flags, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// One workgroup per 64-row block, claimed through a ticket counter in sweep order: every block a
// workgroup waits on (earlier tiles, and the other blocks of its own tile in phase 3) holds a smaller
// ticket or belongs to its own tile, whose blocks hold consecutive tickets, so it is running or done.
template <typename TL, typename TC, int NR>
__global__ __launch_bounds__(NTHREADS) void k_tsweep(const DevMat* Mp, TileSweepTab tab, TL* const* tinv, int forward,
                                                     const double* R, long ldr, TC* Y, long ldy, TC* D, long ldd,
                                                     double* X, long ldx, TC* Rp, long ldp, int nrhs, int32_t* flags,
                                                     int32_t* ticket, int32_t* err, int64_t spin_limit) {
    __shared__ int s_t;
    if (threadIdx.x == 0) s_t = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int q = s_t;
    if (q >= tab.P) return;
    const int p = forward ? q : tab.P - 1 - q;
    item_tsweep<TL, TC, NR>(*Mp, tab, tinv, forward, p, R, ldr, Y, ldy, D, ldd, X, ldx, Rp, ldp, nrhs, flags, err,
                            spin_limit);
}

}  // namespace ttad
