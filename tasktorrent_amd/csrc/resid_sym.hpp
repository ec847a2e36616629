// FP64 residual r = b - A x of a symmetric matrix held as its lower tiles, reading each
// stored element of A once.
//
// The row-block residual (k_resid_part) reads every off-diagonal 64 x 64 block twice: as
// A[p,q] for the rows of block p and, transposed, for the rows of block q. Here one
// workgroup takes one lower tile pair (i, j), i >= j, and passes its 64 x 64 blocks through
// LDS once; both products take their operands from the staged block:
//   row part  y_i[a] += A_ij[a, b] x_j[b]      (rows of tile i)
//   col part  y_j[b] += A_ij[a, b]^T x_i[a]    (rows of tile j; i > j or a > b)
// In a diagonal tile both parts belong to the tile's own rows; its diagonal blocks are made
// symmetric in LDS from their stored lower triangle and contribute once.
//
// Partials: part[s] (n x 16, row-major) for s <= nt. For the rows of tile t, slot s < nt
// holds the col part of tile pair (s, t) when s >= t and the row part of tile pair (t, s)
// when s < t; slot nt holds the row part of the diagonal tile pair (t, t). Every (t, s) has
// exactly one writer, and k_resid_sym_reduce adds the slots in slot order, so r is the same
// bit for bit from run to run. The partials add 2 * 8 * (nt + 1) * n * 16 bytes of traffic
// (written, then read), 12 % of A's bytes at 512-row tiles.
//
// Workgroup: 8 waves over one staged block; waves 0-3 form the row part (16 rows each),
// waves 4-7 the col part (16 columns each), so each keeps only its own accumulators.
//
// Per call: A's stored lower tiles (8 bytes per element) are read once; 4 flops per stored
// element and right-hand side on FP64 MFMA (16 x 16 x 4), 64 flops per 8 bytes at 16 rhs.
#pragma once
#include "solve_ops.hpp"

namespace ttad {

constexpr int RSY_NR = 16;                         // right-hand sides per pass
constexpr int RSY_LD = SB + 1;                     // LDS pitch (doubles) of a staged A block
constexpr int RSY_A = SB * RSY_LD;                 // doubles of a staged A block
constexpr int RSY_STAGE = RSY_A + SB * RSY_NR;     // A block + the x_j block
constexpr int LDS_RESID_SYM = 2 * RSY_STAGE * 8;   // 81 KB (two stages)
constexpr int RSY_MAX_TILE = 512;                  // tile rows the register accumulators cover
constexpr int RSY_THREADS = 512;

// A block in LDS: element (r, c) at r * 65 + c. Sixteen lanes that read one column of 16
// rows, or 16 columns of one row, hit distinct banks, and the k-steps of a product differ by
// constant offsets.
__device__ __forceinline__ int rsy_idx(int r, int c) { return r * RSY_LD + c; }

// Columns [0, nrhs) of a column-major n x nrhs array as a row-major (n_pad x 16) array,
// zero in the columns >= nrhs and the rows >= n: the x operand layout of k_resid_sym.
__global__ void k_rhs_rows16(const double* X, long ldx, long n, long n_pad, int nrhs, double* Xr) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_pad * RSY_NR) return;
    const long g = e / RSY_NR;
    const int c = (int)(e % RSY_NR);
    gmem(Xr)[e] = (g < n && c < nrhs) ? gmem(X)[g + (long)c * ldx] : 0.0;
}

// One workgroup per lower tile pair, workgroup x -> (i, j) with x = i (i + 1) / 2 + j.
// Requires every tile to have at most RSY_MAX_TILE rows; Xr has n + 64 rows.
__global__ __launch_bounds__(RSY_THREADS, 1) void k_resid_sym(const DevMat* Mp, const double* Xr, double* part,
                                                              long n) {
    extern __shared__ __align__(16) char lds[];
    double* S = reinterpret_cast<double*>(lds);
    const DevMat& M = *Mp;
    const int x = blockIdx.x;
    int ti = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= x) ti++;
    while (ti * (ti + 1) / 2 > x) ti--;
    const int tj = x - ti * (ti + 1) / 2;
    const bool diag = ti == tj;
    const int bi = M.bsz[ti], bj = M.bsz[tj];
    const int nbi = (bi + SB - 1) / SB, nbj = (bj + SB - 1) / SB;
    const long ld = bi;
    const TTAD_GLOBAL double* T = gmem((const double*)M.tiles[ti + (long)tj * M.nt]);
    const int gi0 = M.boff[ti], gj0 = M.boff[tj];
    const TTAD_GLOBAL double* X = gmem(Xr);
    const int tid = threadIdx.x, lane = tid & 63, lm = lane & 15, lk = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool colw = wv >= 4;      // col-part wave
    const int w = wv & 3;           // its 16-row (row part) or 16-column (col part) slice
    const int sr = 2 * (tid & 31), sc = tid >> 5;   // A staging: rows sr, sr + 1 of columns sc + 16 s
    const int xr = tid >> 3, xc = 2 * (tid & 7);    // x staging: row xr, columns xc, xc + 1
    typedef V16<double>::u vu;
    typedef V16<double>::t vt;

    double ra[8], rx[2], xi[KS];
    d4_t accr = d4_t{0, 0, 0, 0}, accc[RSY_MAX_TILE / SB];
#pragma unroll
    for (int b = 0; b < RSY_MAX_TILE / SB; b++) accc[b] = d4_t{0, 0, 0, 0};

    // block (a, b) of the tile and rows b of x_j into registers
    auto load_step = [&](int a, int b) {
        const int wr = min(SB, bi - a * SB), wc = min(SB, bj - b * SB);
        const TTAD_GLOBAL double* src = T + a * SB + (long)b * SB * ld;
        if (wr == SB && wc == SB) {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const vu v = *reinterpret_cast<const TTAD_GLOBAL vu*>(src + sr + (long)(sc + 16 * s) * ld);
                ra[2 * s] = v[0];
                ra[2 * s + 1] = v[1];
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int c = sc + 16 * s;
                const long o = sr + (long)c * ld;
                ra[2 * s] = (sr < wr && c < wc) ? src[o] : 0.0;
                ra[2 * s + 1] = (sr + 1 < wr && c < wc) ? src[o + 1] : 0.0;
            }
        }
        const vu v = *reinterpret_cast<const TTAD_GLOBAL vu*>(X + (long)(gj0 + b * SB + xr) * RSY_NR + xc);
        rx[0] = v[0];
        rx[1] = v[1];
    };
    // registers -> LDS; sym: a diagonal block of a diagonal tile, mirrored from its lower triangle
    auto stage = [&](double* St, bool sym) {
        if (!sym) {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int c = sc + 16 * s;
                St[rsy_idx(sr, c)] = ra[2 * s];
                St[rsy_idx(sr + 1, c)] = ra[2 * s + 1];
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int c = sc + 16 * s;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int r = sr + h;
                    if (r >= c) St[rsy_idx(r, c)] = ra[2 * s + h];
                    if (r > c) St[rsy_idx(c, r)] = ra[2 * s + h];
                }
            }
        }
        *reinterpret_cast<vt*>(St + RSY_A + xr * RSY_NR + xc) = vt{rx[0], rx[1]};
    };
    // rows a of x_i in the B operand layout of the col part
    auto load_xi = [&](int a) {
        const TTAD_GLOBAL double* xs = X + (long)(gi0 + a * SB + lk) * RSY_NR + lm;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) xi[ks] = xs[(long)ks * 4 * RSY_NR];
    };
    auto store = [&](const d4_t& acc, int slot, int g0, int rows) {
        TTAD_GLOBAL double* out = gmem(part) + (long)slot * n * RSY_NR;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = 16 * w + Mfma<double>::row(lane, r);
            if (m < rows) out[(long)(g0 + m) * RSY_NR + lm] = acc[r];
        }
    };
    auto last_b = [&](int a) { return diag ? a : nbj - 1; };

    // step (0, 0) staged, the step after it in flight
    load_step(0, 0);
    stage(S, diag);
    {
        int a1 = 0, b1 = 1;
        if (b1 > last_b(0)) {
            a1 = 1;
            b1 = 0;
        }
        if (a1 < nbi) load_step(a1, b1);
    }
    __syncthreads();
    int e = 0;
    for (int a = 0; a < nbi; a++) {
        const int blast = last_b(a);
        if (colw) load_xi(a);
#pragma unroll
        for (int b = 0; b < RSY_MAX_TILE / SB; b++) {
            if (b <= blast) {
                const double* cur = S + (e & 1) * RSY_STAGE;
                double* nxt = S + ((e + 1) & 1) * RSY_STAGE;
                // stage the next step (its buffer was last read before the previous barrier), then
                // start the loads of the step after it
                int a1 = a, b1 = b + 1;
                if (b1 > blast) {
                    a1 = a + 1;
                    b1 = 0;
                }
                if (a1 < nbi) {
                    stage(nxt, diag && a1 == b1);
                    int a2 = a1, b2 = b1 + 1;
                    if (b2 > last_b(a1)) {
                        a2 = a1 + 1;
                        b2 = 0;
                    }
                    if (a2 < nbi) load_step(a2, b2);
                }
                if (!colw) {
                    const double* Sx = cur + RSY_A;
#pragma unroll
                    for (int ks = 0; ks < KS; ks++) {
                        const int k = 4 * ks + lk;
                        accr = Mfma<double>::mma(cur[rsy_idx(16 * w + lm, k)], Sx[k * RSY_NR + lm], accr);
                    }
                } else if (!diag || b != a) {  // a diagonal block of a diagonal tile has no col part
#pragma unroll
                    for (int ks = 0; ks < KS; ks++) {
                        const int k = 4 * ks + lk;
                        accc[b] = Mfma<double>::mma(cur[rsy_idx(k, 16 * w + lm)], xi[ks], accc[b]);
                    }
                }
                __syncthreads();
                e++;
            }
        }
        if (!colw) {
            store(accr, diag ? M.nt : tj, gi0 + a * SB, min(SB, bi - a * SB));
            accr = d4_t{0, 0, 0, 0};
        }
    }
    if (colw) {
#pragma unroll
        for (int b = 0; b < RSY_MAX_TILE / SB; b++)
            if (b < nbj) store(accc[b], ti, gj0 + b * SB, min(SB, bj - b * SB));
    }
}

// r = b - (sum of the nslot partials, in slot order) for the first nrhs columns, and the
// per-rhs maxima of |r|, |x|, |b| (reduced in the workgroup, then one atomicMax each).
__global__ __launch_bounds__(NTHREADS) void k_resid_sym_reduce(const double* part, long n, int nslot, const double* B,
                                                               long ldb, const double* X, long ldx, double* R,
                                                               long ldr, int nrhs, unsigned long long* norms) {
    __shared__ unsigned long long s_max[3 * RSY_NR];
    for (int e = threadIdx.x; e < 3 * RSY_NR; e += NTHREADS) s_max[e] = 0;
    __syncthreads();
    const long e = (long)blockIdx.x * NTHREADS + threadIdx.x;
    const long g = e / RSY_NR;
    const int c = (int)(e % RSY_NR);
    if (g < n && c < nrhs) {
        const TTAD_GLOBAL double* p = gmem(part) + e;
        double sum = 0.0;
        for (int s = 0; s < nslot; s++) sum += p[(long)s * n * RSY_NR];
        const double b = gmem(B)[g + (long)c * ldb];
        const double rv = b - sum;
        gmem(R)[g + (long)c * ldr] = rv;
        atomicMax(s_max + 3 * c + 0, (unsigned long long)__double_as_longlong(fabs(rv)));
        atomicMax(s_max + 3 * c + 1, (unsigned long long)__double_as_longlong(fabs(gmem(X)[g + (long)c * ldx])));
        atomicMax(s_max + 3 * c + 2, (unsigned long long)__double_as_longlong(fabs(b)));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * nrhs; i += NTHREADS)
        if (s_max[i]) atomicMax(norms + i, s_max[i]);
}

}  // namespace ttad
