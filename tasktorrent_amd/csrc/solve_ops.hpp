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
// workgroup that is already running), and published with a release flag.
#pragma once
#include "tile_ops.hpp"

namespace ttad {

constexpr int SB = 64;        // rows per sweep block
constexpr int SPAD = 4;       // LDS row padding (elements)

struct SweepTab {
    int32_t P;                // number of 64-row blocks
    const int32_t* blk_tile;  // [P]
    const int32_t* blk_off;   // [P]
};

__device__ __forceinline__ int blk_rows(const DevMat& M, const SweepTab& t, int p) {
    return min(SB, M.bsz[t.blk_tile[p]] - t.blk_off[p]);
}
__device__ __forceinline__ int blk_g0(const DevMat& M, const SweepTab& t, int p) {
    return M.boff[t.blk_tile[p]] + t.blk_off[p];
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
        const TL* T = (const TL*)M.tiles[tp + (long)tq * nt];
        const long ld = M.bsz[tp];
        for (int e = threadIdx.x; e < SB * SB; e += NTHREADS) {
            const int m = e % SB, k = e / SB;
            TC v = (m < wp && k < wq) ? (TC)T[(op + m) + (long)(oq + k) * ld] : TC(0);
            Ls[k * LD + m] = absval ? (TC)fabs((double)v) : v;
        }
    } else if (mode == 1) {
        const TL* T = (const TL*)M.tiles[tq + (long)tp * nt];
        const long ld = M.bsz[tq];
        for (int e = threadIdx.x; e < SB * SB; e += NTHREADS) {
            const int k = e % SB, m = e / SB;  // contiguous in k
            TC v = (m < wp && k < wq) ? (TC)T[(oq + k) + (long)(op + m) * ld] : TC(0);
            Ls[k * LD + m] = absval ? (TC)fabs((double)v) : v;
        }
    } else {
        const TL* T = (const TL*)M.tiles[tp + (long)tp * nt];
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
        Ys[k * LD + n] = (k < w && n < nrhs) ? (TC)Y[(g0 + k) + (long)n * ldy] : TC(0);
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

// Waits until *count >= want and returns the value seen. The polls are relaxed (an acquire
// load invalidates the L2 on every iteration, which with hundreds of waiting workgroups
// evicts everyone's data); one agent-scope acquire follows. Gives up after `limit` ticks
// (sets *err = 1) or as soon as *err is nonzero (an earlier wait timed out, or the owner
// asked the kernel to stop), so that a failed run drains in one time limit.
__device__ __forceinline__ int wait_count(const int32_t* count, int want, int32_t* err, int64_t limit) {
    const int64_t t0 = (int64_t)wall_clock64();
    int v;
    while ((v = __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < want) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) { v = want; break; }
        if ((int64_t)wall_clock64() - t0 > limit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            v = want;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return v;
}

// One 64-row block p of a triangular sweep with the factor held in M.
// forward != 0: Y[p] = L[p,p]^{-1} (R[p] - sum_{q<p} L[p,q] Y[q])          (R FP64, Y in TC)
// forward == 0: D[p] = L[p,p]^{-T} (Y[p] - sum_{q>p} L[q,p]^T D[q]), X[p] += D[p] (FP64)
// Blocks finish in sweep order (block p needs every earlier block), so progress is one
// counter, flags[0] = number of finished blocks: a block waits only when it reaches a
// predecessor beyond the count it last saw, and one acquire then covers all of them.
// The counter is incremented (release) when the block is done. The caller guarantees
// that every block this one waits on belongs to a running workgroup.
template <typename TL, typename TC, int NR>
__device__ __forceinline__ void item_sweep(const DevMat& M, const SweepTab& tab, int forward, int p, const double* R,
                                           long ldr, TC* Y, long ldy, TC* D, long ldd, double* X, long ldx, int nrhs,
                                           int32_t* flags, int32_t* err, int64_t spin_limit, char* lds) {
    constexpr int NB = NR / 16;
    constexpr int LD = SB + SPAD, LDY = NR + SPAD;
    TC* Ls = reinterpret_cast<TC*>(lds);
    TC* Ys = Ls + SB * LD;
    const int P = tab.P;
    const int g0 = blk_g0(M, tab, p), w = blk_rows(M, tab, p);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    typename Mfma<TC>::acc_t acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = typename Mfma<TC>::acc_t{0, 0, 0, 0};
    __shared__ int s_seen;
    const int need = forward ? p : P - 1 - p;  // predecessors, in sweep order
    int seen = 0;                               // predecessors known to be finished
    for (int i = 0; i < need; i++) {
        const int q = forward ? i : P - 1 - i;
        // the factor block does not depend on the counter: fetch it first
        if (forward) fill_block<TL, TC>(Ls, M, tab, p, q, 0, false);
        else fill_block<TL, TC>(Ls, M, tab, p, q, 1, false);
        if (i >= seen) {  // uniform
            if (threadIdx.x == 0) s_seen = wait_count(flags, i + 1, err, spin_limit);
            __syncthreads();
            seen = s_seen;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        const int gq = blk_g0(M, tab, q), wq = blk_rows(M, tab, q);
        if (forward) fill_rhs<TC, TC, NR>(Ys, Y, ldy, gq, wq, nrhs);
        else fill_rhs<TC, TC, NR>(Ys, D, ldd, gq, wq, nrhs);
        __syncthreads();
        block_mma<TC, NR>(acc, Ls, Ys);
        __syncthreads();
    }
    // t = rhs block - acc, staged in Ys as the B operand of the diagonal product
    for (int nb = 0; nb < NB; nb++)
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<TC>::row(lane, r), n = nb * 16 + (lane & 15);
            TC v = TC(0);
            if (m < w && n < nrhs) {
                const TC rhs = forward ? (TC)R[(g0 + m) + (long)n * ldr] : Y[(g0 + m) + (long)n * ldy];
                v = rhs - acc[nb][r];
            }
            Ys[m * LDY + n] = v;
        }
    // Dinv of the 64x64 diagonal block: forward uses it, backward its transpose
    const int tp = tab.blk_tile[p];
    const double* Dv = M.dinv[tp] + (long)(tab.blk_off[p] / SB) * SB * SB;
    for (int e = threadIdx.x; e < SB * SB; e += NTHREADS) {
        const int a = e % SB, b = e / SB;  // Dv[a + b*64] = Dinv(a, b)
        if (forward) Ls[b * LD + a] = (TC)Dv[e];   // Ls[k][m] = Dinv(m, k)
        else Ls[a * LD + b] = (TC)Dv[e];           // Ls[k][m] = Dinv(k, m) = Dinv^T(m, k)
    }
    __syncthreads();
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = typename Mfma<TC>::acc_t{0, 0, 0, 0};
    block_mma<TC, NR>(acc, Ls, Ys);
    for (int nb = 0; nb < NB; nb++)
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<TC>::row(lane, r), n = nb * 16 + (lane & 15);
            if (m < w && n < nrhs) {
                const long gi = g0 + m;
                if (forward) {
                    Y[gi + (long)n * ldy] = acc[nb][r];
                } else {
                    D[gi + (long)n * ldd] = acc[nb][r];
                    X[gi + (long)n * ldx] += (double)acc[nb][r];
                }
            }
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(flags, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename TC, int NR>
constexpr int sweep_lds_bytes() {
    return SB * (SB + SPAD) * (int)sizeof(TC) + SB * (NR + SPAD) * (int)sizeof(TC);
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
// sums of |A| instead (x ignored) and reduces their maximum into norms[0].
template <int NR>
__device__ __forceinline__ void item_resid(const DevMat& M, const SweepTab& tab, int p, const double* B, long ldb,
                                           const double* X, long ldx, double* R, long ldr, int nrhs,
                                           unsigned long long* norms, int rownorm, char* lds) {
    constexpr int NB = NR / 16;
    constexpr int LD = SB + SPAD, LDY = NR + SPAD;
    double* Ls = reinterpret_cast<double*>(lds);
    double* Ys = Ls + SB * LD;
    const int g0 = blk_g0(M, tab, p), w = blk_rows(M, tab, p);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tp = tab.blk_tile[p];
    d4_t acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = d4_t{0, 0, 0, 0};
    for (int q = 0; q < tab.P; q++) {
        const int tq = tab.blk_tile[q];
        int mode;
        if (q == p) mode = 2;
        else if (tp > tq || (tp == tq && tab.blk_off[p] > tab.blk_off[q])) mode = 0;
        else mode = 1;
        fill_block<double, double>(Ls, M, tab, p, q, mode, rownorm != 0);
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
    for (int nb = 0; nb < NB; nb++)
        for (int r = 0; r < 4; r++) {
            const int m = wv * 16 + Mfma<double>::row(lane, r), n = nb * 16 + (lane & 15);
            if (m >= w || n >= nrhs) continue;
            const long gi = g0 + m;
            if (rownorm) {
                if (n == 0) atomicMax(norms, (unsigned long long)__double_as_longlong(acc[nb][r]));
                continue;
            }
            const double b = B[gi + (long)n * ldb];
            const double rv = b - acc[nb][r];
            R[gi + (long)n * ldr] = rv;
            atomicMax(norms + 3 * n + 0, (unsigned long long)__double_as_longlong(fabs(rv)));
            atomicMax(norms + 3 * n + 1, (unsigned long long)__double_as_longlong(fabs(X[gi + (long)n * ldx])));
            atomicMax(norms + 3 * n + 2, (unsigned long long)__double_as_longlong(fabs(b)));
        }
}

template <int NR>
__global__ __launch_bounds__(NTHREADS) void k_residual(const DevMat* Mp, SweepTab tab, const double* B, long ldb,
                                                       const double* X, long ldx, double* R, long ldr, int nrhs,
                                                       unsigned long long* norms, int rownorm) {
    __shared__ __align__(16) char lds[sweep_lds_bytes<double, NR>()];
    if ((int)blockIdx.x >= tab.P) return;
    item_resid<NR>(*Mp, tab, blockIdx.x, B, ldb, X, ldx, R, ldr, nrhs, norms, rownorm, lds);
}

}  // namespace ttad
