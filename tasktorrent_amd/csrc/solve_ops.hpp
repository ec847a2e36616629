// Device kernels of the solve phase on a tiled factor: forward/backward
// triangular sweeps over 64-row blocks and the FP64 residual r = b - A x.
//
// Replaces the reference's check solve (L.solveInPlace / L^T.solveInPlace / A*x,
// miniapp/dense_cholesky/2d_cholesky.cpp:676-681) and provides the refinement
// steps of the mixed-precision solver.
//
// Matrix views: block p (64 rows) of a tiled matrix lies in tile blk_tile[p] at
// local row blk_off[p]; every tile size is a multiple of 64 except possibly the
// last one. Right-hand sides are column-major n x nrhs arrays (ld >= n).
//
// Sweeps run one workgroup per block. Blocks are claimed in dependency order
// through a ticket counter (so every block a workgroup waits on is owned by a
// workgroup that is already running), and published through a release counter.
#pragma once
#include "tile_ops.hpp"

namespace ttad {

constexpr int SB = 64;        // rows per sweep block
constexpr int SPAD = 4;       // LDS row padding (elements)
constexpr int KS = SB / 4;    // MFMA k-steps of one 64-wide block product

// Sweep waiting policy: a block with at most `near` predecessors left waits for each
// one as it finishes; a block further from the front waits until `batch` more (up to
// `near` short of its own turn) are done, then takes them in one go with one acquire,
// and polls slowly meanwhile. Measured at n = 32768 (512 blocks, FP64, 16 rhs): batches
// make the far blocks fall behind the front (batch 8: 13.4 ms for both sweeps, batch 32:
// 31 ms), so the default takes every predecessor as soon as it is final (8.1 ms), with
// only the slow polling of the blocks that are not next in line.
constexpr int SWEEP_NEAR = 1;
constexpr int SWEEP_BATCH = 1;

struct SweepTab {
    int32_t P;                // number of 64-row blocks
    const int32_t* blk_tile;  // [P]
    const int32_t* blk_off;   // [P]
    int32_t near = SWEEP_NEAR;
    int32_t batch = SWEEP_BATCH;
};

__device__ __forceinline__ int blk_rows(const DevMat& M, const SweepTab& t, int p) {
    return min(SB, M.bsz[t.blk_tile[p]] - t.blk_off[p]);
}
__device__ __forceinline__ int blk_g0(const DevMat& M, const SweepTab& t, int p) {
    return M.boff[t.blk_tile[p]] + t.blk_off[p];
}

// Block products run on 16x16x4 MFMAs with operands read straight from memory into
// registers (no LDS, no barrier): wave w owns rows 16w..16w+15 of the 64-row block, and
// in k-step ks lane l supplies A[m = 16w + l%16][k = 4ks + l/16] and B[k][n = 16nb + l%16].

// A operand of wave w: rows of block p x columns of block q of the symmetric/triangular
// matrix held in the lower tiles of M.
//   mode 0: plain block L[p,q] (requires tile(p) >= tile(q))
//   mode 1: transposed block L[q,p]^T (requires tile(q) >= tile(p))
//   mode 2: symmetric diagonal block (p == q, lower part stored)
// Out-of-range rows/cols read as 0.
template <typename TL>
__device__ __forceinline__ void load_a(TL (&a)[KS], const DevMat& M, const SweepTab& tab, int p, int q, int mode) {
    const int tp = tab.blk_tile[p], tq = tab.blk_tile[q];
    const int op = tab.blk_off[p], oq = tab.blk_off[q];
    const int wp = min(SB, M.bsz[tp] - op), wq = min(SB, M.bsz[tq] - oq);
    const int lane = threadIdx.x & 63;
    const int m = (threadIdx.x >> 6) * 16 + (lane & 15), kg = lane >> 4;
    const bool mok = m < wp;
    if (mode == 0) {
        const long ld = M.bsz[tp];
        const TTAD_GLOBAL TL* T = gmem((const TL*)M.tiles[tp + (long)tq * M.nt]) + (op + m) + (long)oq * ld;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int k = 4 * ks + kg;
            a[ks] = (mok && k < wq) ? T[(long)k * ld] : TL(0);
        }
    } else if (mode == 1) {
        const long ld = M.bsz[tq];
        const TTAD_GLOBAL TL* T = gmem((const TL*)M.tiles[tq + (long)tp * M.nt]) + oq + (long)(op + m) * ld;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int k = 4 * ks + kg;
            a[ks] = (mok && k < wq) ? T[k] : TL(0);
        }
    } else {
        const long ld = M.bsz[tp];
        const TTAD_GLOBAL TL* T = gmem((const TL*)M.tiles[tp + (long)tp * M.nt]) + op + (long)op * ld;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int k = 4 * ks + kg;
            a[ks] = (mok && k < wp) ? (m >= k ? T[m + (long)k * ld] : T[k + (long)m * ld]) : TL(0);
        }
    }
}

// B operand: rows [g0, g0 + w) of a column-major n x nrhs array (k = row within the block).
template <typename TS, typename TC, int NR>
__device__ __forceinline__ void load_b(TC (&b)[NR / 16][KS], const TS* Y, long ldy, int g0, int w, int nrhs) {
    const int lane = threadIdx.x & 63, kg = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < NR / 16; nb++) {
        const int n = nb * 16 + (lane & 15);
        const TTAD_GLOBAL TS* Yc = gmem(Y) + g0 + (long)n * ldy;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int k = 4 * ks + kg;
            b[nb][ks] = (k < w && n < nrhs) ? (TC)Yc[k] : TC(0);
        }
    }
}

template <typename TC, typename TL, int NB>
__device__ __forceinline__ void reg_mma(typename Mfma<TC>::acc_t (&acc)[NB], const TL (&a)[KS], const TC (&b)[NB][KS]) {
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[nb] = Mfma<TC>::mma((TC)a[ks], b[nb][ks], acc[nb]);
}

// Fill Ls[k][m] (TC, k-major, ld 64+SPAD) with the 64x64 block (rows of block p, cols of block q)
// of the symmetric/triangular matrix held in the lower tiles of M.
//   mode 0: plain block L[p,q] (requires tile(p) >= tile(q)): Ls[k][m] = L[p0+m][q0+k]
//   mode 1: transposed block L[q,p]^T (requires tile(q) >= tile(p)): Ls[k][m] = L[q0+k][p0+m]
//   mode 2: symmetric diagonal block (p == q, lower part stored)
// abs != 0 takes |values| (row norms). Out-of-range rows/cols read as 0.
template <typename TL, typename TC>
__device__ __forceinline__ void fill_block(TC* Ls, const DevMat& M, const SweepTab& tab, int p, int q, int mode,
                                           bool absval) {
    const int tp = tab.blk_tile[p], tq = tab.blk_tile[q];
    const int op = tab.blk_off[p], oq = tab.blk_off[q];
    const int wp = min(SB, M.bsz[tp] - op), wq = min(SB, M.bsz[tq] - oq);
    const int nt = M.nt;
    constexpr int LD = SB + SPAD;
    if (mode == 0) {
        const TTAD_GLOBAL TL* T = gmem((const TL*)M.tiles[tp + (long)tq * nt]);
        const long ld = M.bsz[tp];
        for (int e = threadIdx.x; e < SB * SB; e += NTHREADS) {
            const int m = e % SB, k = e / SB;
            TC v = (m < wp && k < wq) ? (TC)T[(op + m) + (long)(oq + k) * ld] : TC(0);
            Ls[k * LD + m] = absval ? (TC)fabs((double)v) : v;
        }
    } else if (mode == 1) {
        const TTAD_GLOBAL TL* T = gmem((const TL*)M.tiles[tq + (long)tp * nt]);
        const long ld = M.bsz[tq];
        for (int e = threadIdx.x; e < SB * SB; e += NTHREADS) {
            const int k = e % SB, m = e / SB;  // contiguous in k
            TC v = (m < wp && k < wq) ? (TC)T[(oq + k) + (long)(op + m) * ld] : TC(0);
            Ls[k * LD + m] = absval ? (TC)fabs((double)v) : v;
        }
    } else {
        const TTAD_GLOBAL TL* T = gmem((const TL*)M.tiles[tp + (long)tp * nt]);
        const long ld = M.bsz[tp];
        for (int e = threadIdx.x; e < SB * SB; e += NTHREADS) {
            const int m = e % SB, k = e / SB;
            TC v = TC(0);
            if (m < wp && k < wp) v = (m >= k) ? (TC)T[(op + m) + (long)(op + k) * ld] : (TC)T[(op + k) + (long)(op + m) * ld];
            Ls[k * LD + m] = absval ? (TC)fabs((double)v) : v;
        }
    }
}

// Fill Ys[k][n] (k-major, ld NR+SPAD) with rows [g0, g0+w) of a column-major n x nrhs array.
template <typename TS, typename TC, int NR>
__device__ __forceinline__ void fill_rhs(TC* Ys, const TS* Y, long ldy, int g0, int w, int nrhs) {
    constexpr int LD = NR + SPAD;
    for (int e = threadIdx.x; e < SB * NR; e += NTHREADS) {
        const int k = e % SB, n = e / SB;
        Ys[k * LD + n] = (k < w && n < nrhs) ? (TC)gmem(Y)[(g0 + k) + (long)n * ldy] : TC(0);
    }
}

// acc[nb] (wave w owns rows 16w..16w+15, 16-column block nb) += Ls(64x64) * Ys(64 x NR)
template <typename TC, int NR>
__device__ __forceinline__ void block_mma(typename Mfma<TC>::acc_t (&acc)[NR / 16], const TC* Ls, const TC* Ys) {
    constexpr int LD = SB + SPAD, LDY = NR + SPAD;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll 4
    for (int ks = 0; ks < SB / 4; ks++) {
        const int kr = ks * 4 + (lane >> 4);
        const TC a = Ls[kr * LD + w * 16 + (lane & 15)];
#pragma unroll
        for (int nb = 0; nb < NR / 16; nb++) acc[nb] = Mfma<TC>::mma(a, Ys[kr * LDY + nb * 16 + (lane & 15)], acc[nb]);
    }
}

// One 64-row block p of a triangular sweep with the factor held in M.
// forward != 0: Y[p] = L[p,p]^{-1} (R[p] - sum_{q<p} L[p,q] Y[q])          (R FP64, Y in TC)
// forward == 0: D[p] = L[p,p]^{-T} (Y[p] - sum_{q>p} L[q,p]^T D[q]), X[p] += D[p] (FP64)
// Blocks finish in sweep order (block p needs every earlier block), so progress is one
// counter, flags[0] = number of finished blocks: a block waits only when it reaches a
// predecessor beyond the count it last saw (policy in SweepTab), and one acquire by the
// polling thread, passed on by the barrier, then covers all of them. The next factor
// block, and the next right-hand side block once it is known to be final, are loaded
// while the current product runs. The counter is incremented (release) when the block
// is done. The caller guarantees that every block this one waits on belongs to a
// running workgroup.
template <typename TL, typename TC, int NR, bool PF = true>
__device__ __forceinline__ void item_sweep(const DevMat& M, const SweepTab& tab, int forward, int p, const double* R,
                                           long ldr, TC* Y, long ldy, TC* D, long ldd, double* X, long ldx, int nrhs,
                                           int32_t* flags, int32_t* err, int64_t spin_limit, char* lds) {
    constexpr int NB = NR / 16;
    constexpr int LD = SB + SPAD, LDY = NR + SPAD;
    TC* Ds = reinterpret_cast<TC*>(lds);  // diagonal inverse, Ds[k][m] = A operand (m, k)
    TC* Ys = Ds + SB * LD;                // t = rhs - sum, B operand of the diagonal product
    const int P = tab.P;
    const int g0 = blk_g0(M, tab, p), w = blk_rows(M, tab, p);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const TC* Src = forward ? Y : D;  // right-hand side blocks of the predecessors
    const long lds_ = forward ? ldy : ldd;
    typedef typename Mfma<TC>::acc_t acc_t;
    acc_t acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = acc_t{0, 0, 0, 0};

    // Off the critical path: this block's right-hand side and diagonal inverse.
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
    {
        const TTAD_GLOBAL double* Dv = gmem(M.dinv[tab.blk_tile[p]]) + (long)(tab.blk_off[p] / SB) * SB * SB;
        // every load in flight before the first LDS store
        constexpr int PER = SB * SB / NTHREADS;
        const int a = threadIdx.x % SB, b0 = threadIdx.x / SB;  // Dv[a + b*64] = Dinv(a, b)
        double dv[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) dv[q] = Dv[a + (b0 + q * (NTHREADS / SB)) * SB];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int b = b0 + q * (NTHREADS / SB);
            if (forward) Ds[b * LD + a] = (TC)dv[q];   // (m, k) = Dinv(m, k)
            else Ds[a * LD + b] = (TC)dv[q];           // (m, k) = Dinv(k, m)
        }
    }

    // The count is passed through LDS slot (wait number & 1): a wave cannot be two
    // barriers behind the polling thread, so a slot is never rewritten before it is read.
    __shared__ int s_seen[2];
    int nwait = 0;
    const int need = forward ? p : P - 1 - p;  // predecessors, in sweep order
    int seen = 0;                               // predecessors known to be finished
    const int mode = forward ? 0 : 1;
    auto qof = [&](int i) { return forward ? i : P - 1 - i; };
    // Waits (thread 0 polls, the barrier passes the count on) until predecessor i is final.
    auto wait_for = [&](int i) {
        if (i < seen) return;  // uniform
        const int left = need - i;
        const bool far = left > tab.near;
        const int want = far ? min(i + tab.batch, need - tab.near) : i + 1;
        if (threadIdx.x == 0) s_seen[nwait & 1] = wait_count(flags, want, err, spin_limit, far);
        __syncthreads();
        seen = s_seen[nwait & 1];
        nwait++;
    };
    if constexpr (PF) {
        // Two predecessors in flight: slot i & 1 holds predecessor i's factor block and, when it
        // was known final at load time, its right-hand side block; the loads for i + 2 are issued
        // before the product of i. A block behind the front is bound by these loads' latency,
        // one round trip per two predecessors instead of one per predecessor.
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
                if (i + 2 < seen) {
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
    } else {  // LDS-staged operands: few live registers (kernels with other paths)
        for (int i = 0; i < need; i++) {
            const int q = qof(i);
            wait_for(i);
            TC* Ls = Ys + SB * LDY;
            TC* Bs = Ls + SB * LD;
            fill_block<TL, TC>(Ls, M, tab, p, q, mode, false);
            fill_rhs<TC, TC, NR>(Bs, Src, lds_, blk_g0(M, tab, q), blk_rows(M, tab, q), nrhs);
            __syncthreads();
            block_mma<TC, NR>(acc, Ls, Bs);
            __syncthreads();
        }
    }
    // t = rhs block - acc, staged in Ys as the B operand of the diagonal product
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<TC>::row(lane, r), n = nb * 16 + (lane & 15);
            Ys[m * LDY + n] = rhs[nb][r] - acc[nb][r];
        }
    __syncthreads();
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        const int kr = ks * 4 + (lane >> 4);
        const TC a = Ds[kr * LD + wv * 16 + (lane & 15)];
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[nb] = Mfma<TC>::mma(a, Ys[kr * LDY + nb * 16 + (lane & 15)], acc[nb]);
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
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(flags, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS of item_sweep<.., PF> (and of item_resid<NR, false> for TC = double).
template <typename TC, int NR, bool PF = true>
constexpr int sweep_lds_bytes() {
    return (PF ? 1 : 2) * (SB * (SB + SPAD) * (int)sizeof(TC) + SB * (NR + SPAD) * (int)sizeof(TC));
}

// Stand-alone sweep launch: one workgroup per block, blocks claimed through a ticket
// counter in dependency order.
template <typename TL, typename TC, int NR>
__global__ __launch_bounds__(NTHREADS) void k_sweep(const DevMat* Mp, SweepTab tab, int forward, const double* R,
                                                    long ldr, TC* Y, long ldy, TC* D, long ldd, double* X, long ldx,
                                                    int nrhs, int32_t* flags, int32_t* ticket, int32_t* err,
                                                    int64_t spin_limit) {
    __shared__ __align__(16) char lds[sweep_lds_bytes<TC, NR>()];
    __shared__ int s_t;
    if (threadIdx.x == 0) s_t = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int t = s_t;
    if (t >= tab.P) return;
    const int p = forward ? t : tab.P - 1 - t;
    item_sweep<TL, TC, NR>(*Mp, tab, forward, p, R, ldr, Y, ldy, D, ldd, X, ldx, nrhs, flags, err, spin_limit, lds);
}

// Residual r = b - A x (FP64, A symmetric from its lower tiles) for one 64-row block p,
// with per-rhs maxima of |r|, |x|, |b| (bit patterns of non-negative doubles are ordered
// like the values, so an integer atomicMax reduces them). rownorm != 0 computes the row
// sums of |A| instead (x ignored) and reduces their maximum into norms[0]. The blocks of
// the next column block are loaded while the current product runs.
template <int NR, bool PF = true>
__device__ __forceinline__ void item_resid(const DevMat& M, const SweepTab& tab, int p, const double* B, long ldb,
                                           const double* X, long ldx, double* R, long ldr, int nrhs,
                                           unsigned long long* norms, int rownorm, char* lds) {
    constexpr int NB = NR / 16;
    const int g0 = blk_g0(M, tab, p), w = blk_rows(M, tab, p);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tp = tab.blk_tile[p];
    auto mode_of = [&](int q) {
        const int tq = tab.blk_tile[q];
        if (q == p) return 2;
        if (tp > tq || (tp == tq && tab.blk_off[p] > tab.blk_off[q])) return 0;
        return 1;
    };
    d4_t acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = d4_t{0, 0, 0, 0};
    double an[KS], bn[NB][KS];
    auto load_next = [&](int q) {
        load_a<double>(an, M, tab, p, q, mode_of(q));
        if (rownorm) {
            const int wq = blk_rows(M, tab, q), kg = lane >> 4;
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
#pragma unroll
                for (int ks = 0; ks < KS; ks++)
                    bn[nb][ks] = (nb == 0 && (lane & 15) == 0 && 4 * ks + kg < wq) ? 1.0 : 0.0;
        } else {
            load_b<double, double, NR>(bn, X, ldx, blk_g0(M, tab, q), blk_rows(M, tab, q), nrhs);
        }
    };
    if constexpr (!PF) {  // LDS-staged operands: few live registers (kernels with other paths)
        constexpr int LDY = NR + SPAD;
        double* Ls = reinterpret_cast<double*>(lds);
        double* Ys = Ls + SB * (SB + SPAD);
        for (int q = 0; q < tab.P; q++) {
            fill_block<double, double>(Ls, M, tab, p, q, mode_of(q), rownorm != 0);
            const int gq = blk_g0(M, tab, q), wq = blk_rows(M, tab, q);
            if (rownorm) {
                for (int e = threadIdx.x; e < SB * NR; e += NTHREADS) {
                    const int k = e % SB, n = e / SB;
                    Ys[k * LDY + n] = (k < wq && n == 0) ? 1.0 : 0.0;
                }
            } else {
                fill_rhs<double, double, NR>(Ys, X, ldx, gq, wq, nrhs);
            }
            __syncthreads();
            block_mma<double, NR>(acc, Ls, Ys);
            __syncthreads();
        }
    } else {
    load_next(0);
    for (int q = 0; q < tab.P; q++) {
        double ac[KS], bc[NB][KS];
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            ac[ks] = rownorm ? fabs(an[ks]) : an[ks];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) bc[nb][ks] = bn[nb][ks];
        }
        if (q + 1 < tab.P) load_next(q + 1);
        reg_mma<double, double, NB>(acc, ac, bc);
    }
    }
    for (int nb = 0; nb < NB; nb++)
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<double>::row(lane, r), n = nb * 16 + (lane & 15);
            if (m >= w || n >= nrhs) continue;
            const long gi = g0 + m;
            if (rownorm) {
                if (n == 0) atomicMax(norms, (unsigned long long)__double_as_longlong(acc[nb][r]));
                continue;
            }
            const double b = gmem(B)[gi + (long)n * ldb];
            const double rv = b - acc[nb][r];
            gmem(R)[gi + (long)n * ldr] = rv;
            atomicMax(norms + 3 * n + 0, (unsigned long long)__double_as_longlong(fabs(rv)));
            atomicMax(norms + 3 * n + 1, (unsigned long long)__double_as_longlong(fabs(gmem(X)[gi + (long)n * ldx])));
            atomicMax(norms + 3 * n + 2, (unsigned long long)__double_as_longlong(fabs(b)));
        }
}

// Split residual: workgroup (p, seg) sums A[p, q] x_q over the column blocks q of segment
// seg and writes the 64 x nrhs partial to part[seg] (ld n); k_resid_reduce forms
// r = b - sum of the partials and the per-rhs maxima. Many more workgroups than row blocks
// keep enough loads in flight to stream A at HBM rate (one workgroup per row block, each
// reading a whole block row of A, left most of the bandwidth unused).
template <int NR>
__global__ __launch_bounds__(NTHREADS) void k_resid_part(const DevMat* Mp, SweepTab tab, const double* X, long ldx,
                                                         double* part, long n, int nseg, int nrhs) {
    constexpr int NB = NR / 16;
    const DevMat& M = *Mp;
    const int P = tab.P;
    const int p = blockIdx.x % P, seg = blockIdx.x / P;
    const int q0 = (int)((long)P * seg / nseg), q1 = (int)((long)P * (seg + 1) / nseg);
    const int g0 = blk_g0(M, tab, p), w = blk_rows(M, tab, p);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tp = tab.blk_tile[p];
    auto mode_of = [&](int q) {
        const int tq = tab.blk_tile[q];
        if (q == p) return 2;
        if (tp > tq || (tp == tq && tab.blk_off[p] > tab.blk_off[q])) return 0;
        return 1;
    };
    d4_t acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = d4_t{0, 0, 0, 0};
    double an[KS], bn[NB][KS];
    if (q0 < q1) {
        load_a<double>(an, M, tab, p, q0, mode_of(q0));
        load_b<double, double, NR>(bn, X, ldx, blk_g0(M, tab, q0), blk_rows(M, tab, q0), nrhs);
    }
    for (int q = q0; q < q1; q++) {
        double ac[KS], bc[NB][KS];
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            ac[ks] = an[ks];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) bc[nb][ks] = bn[nb][ks];
        }
        if (q + 1 < q1) {
            load_a<double>(an, M, tab, p, q + 1, mode_of(q + 1));
            load_b<double, double, NR>(bn, X, ldx, blk_g0(M, tab, q + 1), blk_rows(M, tab, q + 1), nrhs);
        }
        reg_mma<double, double, NB>(acc, ac, bc);
    }
    TTAD_GLOBAL double* out = gmem(part) + (long)seg * n * NR;
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<double>::row(lane, r), c = nb * 16 + (lane & 15);
            if (m < w && c < nrhs) out[(g0 + m) + (long)c * n] = acc[nb][r];
        }
}

// r = b - sum over the segments' partials for one 64-row block, and the per-rhs maxima of
// |r|, |x|, |b| (reduced in the workgroup, then one atomicMax per rhs and quantity).
template <int NR>
__global__ __launch_bounds__(NTHREADS) void k_resid_reduce(const DevMat* Mp, SweepTab tab, const double* part, long n,
                                                           int nseg, const double* B, long ldb, const double* X,
                                                           long ldx, double* R, long ldr, int nrhs,
                                                           unsigned long long* norms) {
    const int p = blockIdx.x;
    const int g0 = blk_g0(*Mp, tab, p), w = blk_rows(*Mp, tab, p);
    __shared__ unsigned long long s_max[3 * NR];
    for (int e = threadIdx.x; e < 3 * NR; e += NTHREADS) s_max[e] = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < SB * NR; e += NTHREADS) {
        const int m = e % SB, c = e / SB;
        if (m >= w || c >= nrhs) continue;
        const long gi = g0 + m;
        double sum = 0.0;
        for (int sg = 0; sg < nseg; sg++) sum += gmem(part)[(long)sg * n * NR + gi + (long)c * n];
        const double b = gmem(B)[gi + (long)c * ldb];
        const double rv = b - sum;
        gmem(R)[gi + (long)c * ldr] = rv;
        atomicMax(s_max + 3 * c + 0, (unsigned long long)__double_as_longlong(fabs(rv)));
        atomicMax(s_max + 3 * c + 1, (unsigned long long)__double_as_longlong(fabs(gmem(X)[gi + (long)c * ldx])));
        atomicMax(s_max + 3 * c + 2, (unsigned long long)__double_as_longlong(fabs(b)));
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * nrhs; e += NTHREADS)
        if (s_max[e]) atomicMax(norms + e, s_max[e]);
}

template <int NR>
__global__ __launch_bounds__(NTHREADS) void k_residual(const DevMat* Mp, SweepTab tab, const double* B, long ldb,
                                                       const double* X, long ldx, double* R, long ldr, int nrhs,
                                                       unsigned long long* norms, int rownorm) {
    if ((int)blockIdx.x >= tab.P) return;
    item_resid<NR>(*Mp, tab, blockIdx.x, B, ldb, X, ldx, R, ldr, nrhs, norms, rownorm, nullptr);
}

}  // namespace ttad
