// Solve phase: triangular solves with a tiled factor, FP64 residual, iterative
// refinement and the mixed-precision solver with FP64 fallback.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "runtime.hpp"
#include "resid_sym.hpp"
#include "solve_ops.hpp"
#include "tile_sweep.hpp"

namespace ttad {
// Integer tuning override from the environment, clamped to [lo, hi].
static int env_int(const char* name, int dflt, int lo, int hi) {
    const char* v = getenv(name);
    if (!v || !*v) return dflt;
    return std::min(hi, std::max(lo, atoi(v)));
}

// Priority kind of the factorizations inside ttad_posv_mixed: 4, the bottom level of each
// task in the DAG (its longest remaining path). The bf16x3 factorization is short enough for
// its tail, where the serial update chains of the last tiles meet the panel chain, to matter.
// Measured at n = 32768: time to solution 131 ms (kind 4) vs 133 ms (kind 3) and 140 ms (kind 2,
// the reference's critical-path priorities).
constexpr int MIXED_PRIO_KIND_DEFAULT = 4;
static int mixed_prio_kind() {  // TTAD_MIXED_PRIO overrides it (tuning)
    static const int k = env_int("TTAD_MIXED_PRIO", MIXED_PRIO_KIND_DEFAULT, 0, 4);
    return k;
}


// Per-matrix 64-row block table (built on first use).
static int ensure_sweep_tab(ttad_matrix* M) {
    if (M->sweep_P > 0) return TTAD_OK;
    std::vector<int32_t> bt, bo, tf;
    for (int i = 0; i < M->nt; i++) {
        if (i < M->nt - 1 && M->bsz[i] % SB != 0)
            return set_error(TTAD_ERR_INVALID, "solve: tile %d has size %d, not a multiple of %d", i, M->bsz[i], SB);
        tf.push_back((int32_t)bt.size());
        for (int o = 0; o < M->bsz[i]; o += SB) {
            bt.push_back(i);
            bo.push_back(o);
        }
    }
    tf.push_back((int32_t)bt.size());
    TTAD_HIP(M->tab_blk_tile.alloc(std::max<size_t>(1, bt.size()) * 4));
    TTAD_HIP(M->tab_blk_off.alloc(std::max<size_t>(1, bo.size()) * 4));
    TTAD_HIP(M->tab_tfirst.alloc(tf.size() * 4));
    TTAD_HIP(hipMemcpy(M->tab_blk_tile.p, bt.data(), bt.size() * 4, hipMemcpyHostToDevice));
    TTAD_HIP(hipMemcpy(M->tab_blk_off.p, bo.data(), bo.size() * 4, hipMemcpyHostToDevice));
    TTAD_HIP(hipMemcpy(M->tab_tfirst.p, tf.data(), tf.size() * 4, hipMemcpyHostToDevice));
    M->sweep_P = (int)bt.size();
    return TTAD_OK;
}

// Sweep waiting policy; TTAD_SWEEP_NEAR / TTAD_SWEEP_BATCH override the defaults (tuning).
static SweepTab tab_of(const ttad_matrix* M) {
    static const int near = env_int("TTAD_SWEEP_NEAR", SWEEP_NEAR, 1, 1 << 20);
    static const int batch = env_int("TTAD_SWEEP_BATCH", SWEEP_BATCH, 1, 1 << 20);
    return SweepTab{M->sweep_P, M->tab_blk_tile.as<int32_t>(), M->tab_blk_off.as<int32_t>(), near, batch};
}

static int ensure(DevBuf& b, size_t bytes) {
    if (b.bytes < bytes) {  // new buffers start zeroed (the sweep error flag relies on it)
        TTAD_HIP(b.alloc(bytes));
        TTAD_HIP(hipMemset(b.p, 0, b.bytes));
    }
    return TTAD_OK;
}

__global__ void k_zero_i32(int32_t* p, int n) {
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x) p[x] = 0;
}
__global__ void k_zero_u64(unsigned long long* p, int n) {
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x) p[x] = 0;
}
__global__ void k_zero_cols(double* X, long ld, int n, int nrhs) {
    const long tot = (long)n * nrhs;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < tot; x += (long)gridDim.x * blockDim.x)
        X[(x % n) + (x / n) * ld] = 0.0;
}
// Largest |B| over a column-major n x nrhs array (bit pattern of a non-negative double, so
// integer order is numeric order and a NaN ranks above every number), folded into *out.
__global__ void k_absmax_cols(const double* B, long ldb, int n, int nrhs, unsigned long long* out) {
    unsigned long long m = 0;
    const long tot = (long)n * nrhs;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < tot; x += (long)gridDim.x * blockDim.x) {
        const unsigned long long v = (unsigned long long)__double_as_longlong(fabs(B[(x % n) + (x / n) * ldb]));
        m = v > m ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}
__global__ void k_copy_cols(const double* S, long lds_, double* D, long ldd, int n, int nrhs) {
    const long tot = (long)n * nrhs;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < tot; x += (long)gridDim.x * blockDim.x)
        D[(x % n) + (x / n) * ldd] = S[(x % n) + (x / n) * lds_];
}

static int64_t spin_ticks(ttad_context* ctx) {
    int rate_khz = 100000;
    (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device);
    return (int64_t)rate_khz * 1000 * 5;
}

template <typename TL, typename TC>
static void launch_sweep(ttad_matrix* L, int forward, const double* R, long ldr, TC* Y, long ldy, TC* D, long ldd,
                         double* X, long ldx, int nrhs, int32_t* flags, int32_t* ticket, int32_t* err, int64_t spin,
                         hipStream_t s) {
    const SweepTab t = tab_of(L);
    const int P = t.P;
    hipLaunchKernelGGL(k_zero_i32, dim3((P + 255) / 256 + 1), dim3(256), 0, s, flags, P);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, ticket, 1);
    const DevMat* Md = L->desc.as<DevMat>();
    if (nrhs <= 16)
        hipLaunchKernelGGL((k_sweep<TL, TC, 16>), dim3(P), dim3(NTHREADS), 0, s, Md, t, forward, R, ldr, Y, ldy, D,
                           ldd, X, ldx, nrhs, flags, ticket, err, spin);
    else
        hipLaunchKernelGGL((k_sweep<TL, TC, 32>), dim3(P), dim3(NTHREADS), 0, s, Md, t, forward, R, ldr, Y, ldy, D,
                           ldd, X, ldx, nrhs, flags, ticket, err, spin);
}

// Lookahead sweeps (tile_sweep.hpp) for factors in FP32 tiles of at most 512 rows (at most LS_MAXT tiles),
// 16 right-hand sides at a time; the chain of 64-row blocks otherwise, and always with TTAD_SWEEP=block (for comparisons).
static bool lookahead_enabled() {
    static const bool on = [] {
        const char* v = getenv("TTAD_SWEEP");
        return !(v && strcmp(v, "block") == 0);
    }();
    return on;
}
static bool lookahead_ok(const ttad_matrix* L) {
    if (!lookahead_enabled() || L->storage != P_F32 || L->nt > LS_MAXT) return false;
    for (int b : L->bsz)
        if (b > LS_BMAX || b <= 0) return false;
    return true;
}

// Operands of the lookahead sweeps for the factor L (FP32 tiles): Tinv_t, U_t = Tinv_t^T, M_t, KT_t, formed
// from L's tiles and POTRF's diagonal-block inverses unless L->tinv_fresh (every solve call clears it, so a
// factor rewritten in between is never used stale).
static int ensure_lookahead(ttad_matrix* L, hipStream_t s) {
    if (L->tinv_fresh) return TTAD_OK;
    const int nt = L->nt;
    auto up = [](size_t x) { return (x + 63) / 64 * 64; };  // 256-byte aligned sections
    std::vector<size_t> o_inv(nt), o_u(nt), o_lt(nt, 0), o_m(nt, 0), o_k(nt, 0), o_tf(nt), o_uf(nt);
    size_t tot = 0;
    int maxnb = 1;
    for (int t = 0; t < nt; t++) {
        const size_t bt = L->bsz[t];
        maxnb = std::max(maxnb, (L->bsz[t] + SB - 1) / SB);
        o_inv[t] = tot;
        tot += up(bt * bt);
        o_u[t] = tot;
        tot += up(bt * bt);
        o_tf[t] = tot;  // fragment-order copies: 512 x 512, zero-padded
        tot += (size_t)LS_BMAX * LS_BMAX;
        o_uf[t] = tot;
        tot += (size_t)LS_BMAX * LS_BMAX;
        if (t + 1 < nt) {
            o_lt[t] = tot;
            tot += up(bt * L->bsz[t + 1]);
            o_k[t] = tot;
            tot += up(bt * L->bsz[t + 1]);
        }
        if (t >= 1) {
            o_m[t] = tot;
            tot += up(bt * L->bsz[t - 1]);
        }
    }
    if (!L->tinv.p || L->tinv.bytes < std::max<size_t>(1, tot) * 4) {
        TTAD_HIP(L->tinv.alloc(std::max<size_t>(1, tot) * 4));
        TTAD_HIP(hipMemset(L->tinv.p, 0, L->tinv.bytes));  // the padding of the fragment-order copies stays zero
        float* base = L->tinv.as<float>();
        std::vector<float*> ptr((size_t)7 * nt, nullptr);
        for (int t = 0; t < nt; t++) {
            ptr[t] = base + o_inv[t];
            ptr[nt + t] = base + o_u[t];
            ptr[2 * nt + t] = t >= 1 ? base + o_m[t] : nullptr;
            ptr[3 * nt + t] = t + 1 < nt ? base + o_k[t] : nullptr;
            ptr[4 * nt + t] = t + 1 < nt ? base + o_lt[t] : nullptr;
            ptr[5 * nt + t] = base + o_tf[t];
            ptr[6 * nt + t] = base + o_uf[t];
        }
        // products of the lookahead: M_t = L_{t,t-1} U_{t-1}^T and KT_t = LT_t Tinv_{t+1}^T, as A B^T
        std::vector<PairGemm> pg;
        int items = 0;
        auto tile = [&](int i, int j) { return (const float*)((char*)L->tiles.p + L->tile_off[i + (size_t)j * nt] * 4); };
        for (int t = 1; t < nt; t++) {
            const int m = L->bsz[t], k = L->bsz[t - 1];
            pg.push_back(PairGemm{tile(t, t - 1), base + o_u[t - 1], base + o_m[t], m, k, m, m, k, k, items});
            items += ((m + GEMM_BM - 1) / GEMM_BM) * ((k + GEMM_BM - 1) / GEMM_BM);
        }
        for (int t = 0; t + 1 < nt; t++) {
            const int m = L->bsz[t], k = L->bsz[t + 1];
            pg.push_back(PairGemm{base + o_lt[t], base + o_inv[t + 1], base + o_k[t], m, k, m, m, k, k, items});
            items += ((m + GEMM_BM - 1) / GEMM_BM) * ((k + GEMM_BM - 1) / GEMM_BM);
        }
        L->la_items = items;
        L->la_prods = (int)pg.size();
        TTAD_HIP(L->tab_tinv.alloc(ptr.size() * sizeof(float*)));
        TTAD_HIP(hipMemcpy(L->tab_tinv.p, ptr.data(), ptr.size() * sizeof(float*), hipMemcpyHostToDevice));
        TTAD_HIP(L->tab_pairs.alloc(std::max<size_t>(1, pg.size()) * sizeof(PairGemm)));
        if (!pg.empty())
            TTAD_HIP(hipMemcpy(L->tab_pairs.p, pg.data(), pg.size() * sizeof(PairGemm), hipMemcpyHostToDevice));
    }
    float* const* tab = L->tab_tinv.as<float*>();
    const DevMat* Md = L->desc.as<DevMat>();
    hipLaunchKernelGGL(k_tile_inv<float>, dim3(nt, maxnb), dim3(NTHREADS), TILE_INV_LDS, s, Md, tab, tab + nt,
                       tab + 5 * nt, tab + 6 * nt);
    if (nt > 1) {
        hipLaunchKernelGGL(k_sub_transpose<float>, dim3(nt - 1, maxnb * maxnb), dim3(NTHREADS), 0, s, Md, tab + 4 * nt);
        constexpr int lds_pg = gemm_item_lds<float, P_F32>();
        hipLaunchKernelGGL(k_pair_gemm, dim3(L->la_items), dim3(NTHREADS), lds_pg, s, L->tab_pairs.as<PairGemm>(),
                           L->la_prods);
    }
    TTAD_HIP(hipGetLastError());
    L->tinv_fresh = true;
    return TTAD_OK;
}

hipError_t solve_init() {
    hipError_t e = hipFuncSetAttribute((const void*)k_tile_inv<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       TILE_INV_LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_lsweep<1>, hipFuncAttributeMaxDynamicSharedMemorySize, LSWEEP_LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_lsweep<0>, hipFuncAttributeMaxDynamicSharedMemorySize, LSWEEP_LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_pair_gemm, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (gemm_item_lds<float, P_F32>()));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_resid_sym, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_RESID_SYM);
    return e;
}

static void launch_lsweep(ttad_matrix* L, int forward, const double* R, long ldr, float* Y, float* D, double* X,
                          long ldx, float* Rp, int nrhs, int32_t* flags, int32_t* ticket, int32_t* err, int64_t spin,
                          hipStream_t s) {
    const int P = L->sweep_P, nt = L->nt;
    hipLaunchKernelGGL(k_zero_i32, dim3((2 * nt + 255) / 256), dim3(256), 0, s, flags, 2 * nt);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, ticket, 1);
    float* const* tab = L->tab_tinv.as<float*>();
    const LSweepTab lt{P, nt, L->tab_blk_tile.as<int32_t>(), L->tab_blk_off.as<int32_t>(), L->tab_tfirst.as<int32_t>(),
                       tab, tab + nt, tab + 2 * nt, tab + 3 * nt, tab + 5 * nt, tab + 6 * nt};
    if (forward)
        hipLaunchKernelGGL(k_lsweep<1>, dim3(P), dim3(LS_THREADS), LSWEEP_LDS, s, L->desc.as<DevMat>(), lt, R, ldr, Y, D,
                           X, ldx, Rp, nrhs, flags, ticket, err, spin);
    else
        hipLaunchKernelGGL(k_lsweep<0>, dim3(P), dim3(LS_THREADS), LSWEEP_LDS, s, L->desc.as<DevMat>(), lt, R, ldr, Y, D,
                           X, ldx, Rp, nrhs, flags, ticket, err, spin);
}

// X += (L L^T)^{-1} R in the work precision of L: lookahead sweeps (16 right-hand sides at a time) where they
// apply, else the 64-row block chain (at most 32 at a time).
static int solve_add(ttad_context* ctx, ttad_matrix* L, int nrhs, const double* R, long ldr, double* X, long ldx,
                     hipStream_t s) {
    int rc = ensure_sweep_tab(L);
    if (rc) return rc;
    const bool la = lookahead_ok(L);
    if (la && (rc = ensure_lookahead(L, s))) return rc;
    const int n = L->n, P = L->sweep_P;
    const size_t es = L->storage == P_F64 ? 8 : 4;
    const int chunk = la ? std::min(nrhs, LS_NR) : std::min(nrhs, 32);
    // the lookahead sweeps' row vectors are padded by a tile: their loads of the last tile stay inside
    const size_t rows = la ? (size_t)n + LS_BMAX : (size_t)n, cols = la ? LS_NR : chunk;
    if ((rc = ensure(ctx->sw_Y, rows * cols * es))) return rc;
    if ((rc = ensure(ctx->sw_D, rows * cols * es))) return rc;
    if (la && (rc = ensure(ctx->sw_P, rows * LS_NR * 4))) return rc;
    if ((rc = ensure(ctx->sw_flags, (size_t)2 * (P + 1) * 4))) return rc;
    if ((rc = ensure(ctx->sw_ticket, 16))) return rc;
    if ((rc = ensure(ctx->sw_err, 16))) return rc;
    const int64_t spin = spin_ticks(ctx);
    int32_t* flags = ctx->sw_flags.as<int32_t>();
    int32_t* ticket = ctx->sw_ticket.as<int32_t>();
    int32_t* err = ctx->sw_err.as<int32_t>();
    for (int c0 = 0; c0 < nrhs; c0 += chunk) {
        const int nc = std::min(chunk, nrhs - c0);
        const double* Rc = R + (long)c0 * ldr;
        double* Xc = X + (long)c0 * ldx;
        if (la) {
            float* Y = ctx->sw_Y.as<float>();
            float* D = ctx->sw_D.as<float>();
            float* Rp = ctx->sw_P.as<float>();
            launch_lsweep(L, 1, Rc, ldr, Y, D, Xc, ldx, Rp, nc, flags, ticket, err, spin, s);
            launch_lsweep(L, 0, Rc, ldr, Y, D, Xc, ldx, Rp, nc, flags, ticket, err, spin, s);
        } else if (L->storage == P_F64) {
            double* Y = ctx->sw_Y.as<double>();
            double* D = ctx->sw_D.as<double>();
            launch_sweep<double, double>(L, 1, Rc, ldr, Y, n, D, n, Xc, ldx, nc, flags, ticket, err, spin, s);
            launch_sweep<double, double>(L, 0, Rc, ldr, Y, n, D, n, Xc, ldx, nc, flags, ticket, err, spin, s);
        } else {
            float* Y = ctx->sw_Y.as<float>();
            float* D = ctx->sw_D.as<float>();
            launch_sweep<float, float>(L, 1, Rc, ldr, Y, n, D, n, Xc, ldx, nc, flags, ticket, err, spin, s);
            launch_sweep<float, float>(L, 0, Rc, ldr, Y, n, D, n, Xc, ldx, nc, flags, ticket, err, spin, s);
        }
    }
    TTAD_HIP(hipGetLastError());
    return TTAD_OK;
}

// r = b - A x and the normwise backward error per column.
static int residual(ttad_context* ctx, ttad_matrix* A, int nrhs, const double* B, long ldb, const double* X,
                    long ldx, double* R, long ldr, double* eta_max, hipStream_t s) {
    int rc = ensure_sweep_tab(A);
    if (rc) return rc;
    if (A->storage != P_F64) return set_error(TTAD_ERR_INVALID, "residual: A must have FP64 tiles");
    const SweepTab t = tab_of(A);
    const DevMat* Md = A->desc.as<DevMat>();
    if ((rc = ensure(ctx->sw_norms, (size_t)(3 * 32 + 8) * 8))) return rc;
    unsigned long long* norms = ctx->sw_norms.as<unsigned long long>();
    if (!A->anorm_valid) {
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, s, norms, 3 * 32 + 8);
        hipLaunchKernelGGL((k_residual<16>), dim3(t.P), dim3(NTHREADS), 0, s, Md, t, B, ldb, X, ldx, R, ldr, 1,
                           norms, 1);
        unsigned long long h = 0;
        TTAD_HIP(hipMemcpyAsync(&h, norms, 8, hipMemcpyDeviceToHost, s));
        TTAD_HIP(hipStreamSynchronize(s));
        double v;
        memcpy(&v, &h, 8);
        A->anorm = v;
        A->anorm_valid = true;
    }
    double worst = 0.0;
    const long n = A->n;
    const int nt = A->nt;
    // tile-pair pass (A read once, 16 right-hand sides per pass) when the tiles fit its accumulators
    // and its partial sums, written and read once, cost less than the second read of A
    const bool sym = !ctx->residual_rows &&
                     *std::max_element(A->bsz.begin(), A->bsz.end()) <= RSY_MAX_TILE &&
                     (double)(nt + 1) * n * RSY_NR * 8 <= 0.5 * (double)A->bytes;
    // otherwise column segments per row block: about 4096 workgroups in all
    const int nseg = std::max(1, std::min(16, 4096 / std::max(1, t.P)));
    const long n_pad = n + SB;
    const size_t part_doubles = sym ? (size_t)(nt + 1) * n * RSY_NR + (size_t)n_pad * RSY_NR : (size_t)nseg * n * 32;
    if ((rc = ensure(ctx->sw_part, part_doubles * 8))) return rc;
    double* part = ctx->sw_part.as<double>();
    const int width = sym ? RSY_NR : 32;
    for (int c0 = 0; c0 < nrhs; c0 += width) {
        const int nc = std::min(width, nrhs - c0);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, s, norms, 3 * 32 + 8);
        const double* Bc = B + (long)c0 * ldb;
        const double* Xc = X + (long)c0 * ldx;
        double* Rc = R + (long)c0 * ldr;
        if (sym) {
            double* Xr = part + (size_t)(nt + 1) * n * RSY_NR;
            hipLaunchKernelGGL(k_rhs_rows16, dim3((unsigned)((n_pad * RSY_NR + 255) / 256)), dim3(256), 0, s, Xc, ldx,
                               n, n_pad, nc, Xr);
            hipLaunchKernelGGL(k_resid_sym, dim3(nt * (nt + 1) / 2), dim3(RSY_THREADS), LDS_RESID_SYM, s, Md, Xr,
                               part, n);
            hipLaunchKernelGGL(k_resid_sym_reduce, dim3((unsigned)((n * RSY_NR + NTHREADS - 1) / NTHREADS)),
                               dim3(NTHREADS), 0, s, part, n, nt + 1, Bc, ldb, Xc, ldx, Rc, ldr, nc, norms);
        } else if (nc <= 16) {
            hipLaunchKernelGGL((k_resid_part<16>), dim3(t.P * nseg), dim3(NTHREADS), 0, s, Md, t, Xc, ldx, part, n,
                               nseg, nc);
            hipLaunchKernelGGL((k_resid_reduce<16>), dim3(t.P), dim3(NTHREADS), 0, s, Md, t, part, n, nseg, Bc, ldb,
                               Xc, ldx, Rc, ldr, nc, norms);
        } else {
            hipLaunchKernelGGL((k_resid_part<32>), dim3(t.P * nseg), dim3(NTHREADS), 0, s, Md, t, Xc, ldx, part, n,
                               nseg, nc);
            hipLaunchKernelGGL((k_resid_reduce<32>), dim3(t.P), dim3(NTHREADS), 0, s, Md, t, part, n, nseg, Bc, ldb,
                               Xc, ldx, Rc, ldr, nc, norms);
        }
        TTAD_HIP(hipGetLastError());
        unsigned long long h[3 * 32];
        TTAD_HIP(hipMemcpyAsync(h, norms, sizeof(unsigned long long) * 3 * nc, hipMemcpyDeviceToHost, s));
        TTAD_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < nc; j++) {
            double rn, xn, bn;
            memcpy(&rn, &h[3 * j], 8);
            memcpy(&xn, &h[3 * j + 1], 8);
            memcpy(&bn, &h[3 * j + 2], 8);
            const double den = A->anorm * xn + bn;
            const double eta = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
            worst = std::max(worst, std::isnan(eta) ? INFINITY : eta);
        }
    }
    if (eta_max) *eta_max = worst;
    return TTAD_OK;
}

static int check_err_flag(ttad_context* ctx, hipStream_t s) {
    int32_t e = 0;
    TTAD_HIP(hipMemcpyAsync(&e, ctx->sw_err.p, 4, hipMemcpyDeviceToHost, s));
    TTAD_HIP(hipStreamSynchronize(s));
    if (e) {  // cleared for the next call: a set flag also makes every device wait give up at once
        TTAD_HIP(hipMemsetAsync(ctx->sw_err.p, 0, 4, s));
        TTAD_HIP(hipStreamSynchronize(s));
        return set_error(TTAD_ERR_TIMEOUT, "solve: a sweep block waited past the device watchdog");
    }
    return TTAD_OK;
}

// Refinement loop: X := X + L^{-T}L^{-1}(B - A X) until eta <= tol; stops early when
// eta no longer halves (stall). Returns TTAD_ERR_IR_STALLED if tol is not reached.
// X = 0 on entry, so the first residual is B itself and needs no product with A: its
// backward error is max|b| / (||A|| * 0 + max|b|) = 1 for any nonzero B (0 for B = 0), and
// the first correction is solved from B directly.
static int refine(ttad_context* ctx, ttad_matrix* A, ttad_matrix* L, int nrhs, const double* B, long ldb, double* X,
                  long ldx, double tol, int max_iter, int* iters, double* eta_out, hipStream_t s) {
    const int n = A->n;
    L->tinv_fresh = false;  // the factor may have changed since the last solve: its tile inverses are formed again
    int rc = ensure(ctx->sw_R, (size_t)n * nrhs * 8);
    if (rc) return rc;
    if ((rc = ensure(ctx->sw_norms, (size_t)(3 * 32 + 8) * 8))) return rc;
    double* R = ctx->sw_R.as<double>();
    double eta = INFINITY, prev = INFINITY;
    int it = 0, stall = 0;
    {
        unsigned long long* norms = ctx->sw_norms.as<unsigned long long>();
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, s, norms, 1);
        hipLaunchKernelGGL(k_absmax_cols, dim3(256), dim3(256), 0, s, B, ldb, n, nrhs, norms);
        TTAD_HIP(hipGetLastError());
        unsigned long long h = 0;
        TTAD_HIP(hipMemcpyAsync(&h, norms, 8, hipMemcpyDeviceToHost, s));
        TTAD_HIP(hipStreamSynchronize(s));
        double bmax;
        memcpy(&bmax, &h, 8);
        eta = std::isnan(bmax) ? INFINITY : (bmax > 0 ? 1.0 : 0.0);
    }
    while (eta > tol && it < max_iter) {
        if (it == 0)
            rc = solve_add(ctx, L, nrhs, B, ldb, X, ldx, s);
        else
            rc = solve_add(ctx, L, nrhs, R, n, X, ldx, s);
        if (rc) return rc;
        it++;
        prev = eta;
        if ((rc = residual(ctx, A, nrhs, B, ldb, X, ldx, R, n, &eta, s))) return rc;
        if (!(eta <= 0.5 * prev)) {
            if (++stall >= 2) break;
        } else {
            stall = 0;
        }
    }
    if ((rc = check_err_flag(ctx, s))) return rc;
    if (iters) *iters = it;
    if (eta_out) *eta_out = eta;
    return eta <= tol ? TTAD_OK : set_error(TTAD_ERR_IR_STALLED, "refinement reached eta=%.3e > tol=%.1e after %d steps",
                                            eta, tol, it);
}

// Working copy of A (storage st) for the factorization, cached per (A, storage).
static ttad_matrix* working(ttad_context* ctx, ttad_matrix* A, int st, int& rc) {
    auto key = std::make_pair(A->id, st);
    auto it = ctx->work_mats.find(key);
    if (it != ctx->work_mats.end()) return it->second;
    ttad_matrix* W = nullptr;
    rc = ttad_matrix_create(ctx, A->nt, A->bsz.data(), st, &W);
    if (rc != TTAD_OK) return nullptr;
    ctx->work_mats[key] = W;
    return W;
}

}  // namespace ttad

using namespace ttad;

extern "C" {

int ttad_solve(ttad_context* ctx, ttad_matrix* L, int nrhs, const double* B, long long ldb, double* X, long long ldx,
               void* stream) {
    if (!ctx || !L || nrhs <= 0 || !B || !X || ldb < L->n || ldx < L->n)
        return set_error(TTAD_ERR_INVALID, "solve: bad args");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipLaunchKernelGGL(k_zero_cols, dim3(256), dim3(256), 0, s, X, ldx, L->n, nrhs);
    L->tinv_fresh = false;
    int rc = solve_add(ctx, L, nrhs, B, ldb, X, ldx, s);
    if (rc) return rc;
    return check_err_flag(ctx, s);
}

int ttad_residual(ttad_context* ctx, ttad_matrix* A, int nrhs, const double* B, long long ldb, const double* X,
                  long long ldx, double* R, long long ldr, double* eta, void* stream) {
    if (!ctx || !A || nrhs <= 0 || !B || !X || !R || ldb < A->n || ldx < A->n || ldr < A->n)
        return set_error(TTAD_ERR_INVALID, "residual: bad args");
    return residual(ctx, A, nrhs, B, ldb, X, ldx, R, ldr, eta, stream ? (hipStream_t)stream : ctx->stream);
}

int ttad_solve_ir(ttad_context* ctx, ttad_matrix* A, ttad_matrix* L, int nrhs, const double* B, long long ldb,
                  double* X, long long ldx, double tol, int max_iter, ttad_solve_stats* st, void* stream) {
    if (!ctx || !A || !L || A->nt != L->nt || A->bsz != L->bsz || nrhs <= 0 || !B || !X || ldb < A->n || ldx < A->n)
        return set_error(TTAD_ERR_INVALID, "solve_ir: bad args");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipLaunchKernelGGL(k_zero_cols, dim3(256), dim3(256), 0, s, X, ldx, A->n, nrhs);
    int iters = 0;
    double eta = 0;
    const int rc = refine(ctx, A, L, nrhs, B, ldb, X, ldx, tol, max_iter, &iters, &eta, s);
    if (st) {
        memset(st, 0, sizeof(*st));
        st->iterations = iters;
        st->backward_error = eta;
        st->prec_used = L->storage;
        st->factor_residual = -1.0;
        if (ctx->report_factor_residual && (rc == TTAD_OK || rc == TTAD_ERR_IR_STALLED)) {
            const int r2 = factor_residual(ctx, A, L, &st->factor_residual, s);
            if (r2 != TTAD_OK) return r2;
        }
    }
    return rc;
}

int ttad_posv_mixed(ttad_context* ctx, ttad_matrix* A, int prec, int nrhs, const double* B, long long ldb, double* X,
                    long long ldx, double tol, int max_iter, ttad_solve_stats* st, void* stream) {
    if (!ctx || !A || A->storage != P_F64 || prec < 0 || prec > 3 || nrhs <= 0 || !B || !X || ldb < A->n ||
        ldx < A->n)
        return set_error(TTAD_ERR_INVALID, "posv_mixed: bad args");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    auto secs = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    int rc = TTAD_OK;
    ttad_solve_stats loc{};
    loc.prec_used = prec;
    loc.factor_residual = -1.0;
    ttad_matrix* W = working(ctx, A, prec == P_F64 ? P_F64 : P_F32, rc);
    if (!W) return rc;
    if ((rc = ttad_matrix_copy(W, A, s))) return rc;
    int info = 0;
    rc = ttad_potrf_tiled(ctx, W, TTAD_SCHED_PERSISTENT, mixed_prio_kind(), prec, s, &info);
    loc.factor_s = secs();
    loc.info = info;
    int iters = 0;
    double eta = INFINITY;
    if (rc == TTAD_OK) {
        hipLaunchKernelGGL(k_zero_cols, dim3(256), dim3(256), 0, s, X, ldx, A->n, nrhs);
        rc = refine(ctx, A, W, nrhs, B, ldb, X, ldx, tol, max_iter, &iters, &eta, s);
    }
    loc.iterations = iters;
    loc.backward_error = eta;
    if ((rc == TTAD_ERR_NOT_SPD || rc == TTAD_ERR_IR_STALLED) && prec != P_F64) {
        // automatic fallback: FP64 factorization, then refinement with it
        loc.fell_back = 1;
        ttad_matrix* W64 = working(ctx, A, P_F64, rc);
        if (!W64) return rc;
        if ((rc = ttad_matrix_copy(W64, A, s))) return rc;
        rc = ttad_potrf_tiled(ctx, W64, TTAD_SCHED_PERSISTENT, mixed_prio_kind(), P_F64, s, &info);
        loc.info = info;
        if (rc == TTAD_OK) {
            hipLaunchKernelGGL(k_zero_cols, dim3(256), dim3(256), 0, s, X, ldx, A->n, nrhs);
            int it2 = 0;
            rc = refine(ctx, A, W64, nrhs, B, ldb, X, ldx, tol, max_iter, &it2, &eta, s);
            loc.iterations += it2;
            loc.backward_error = eta;
        }
        loc.prec_used = P_F64;
        W = W64;
    }
    TTAD_HIP(hipStreamSynchronize(s));
    loc.total_s = secs();
    loc.solve_s = loc.total_s - loc.factor_s;
    // the factor residual is reported after the solution is complete and is not part of its times
    if (ctx->report_factor_residual && (rc == TTAD_OK || rc == TTAD_ERR_IR_STALLED)) {
        const int r2 = factor_residual(ctx, A, W, &loc.factor_residual, s);
        if (r2 != TTAD_OK) return r2;
    }
    if (st) *st = loc;
    return rc;
}

}  // extern "C"
