// C-ABI of tasktorrent_amd (declared in include/tasktorrent_amd.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <thread>

#include "runtime.hpp"

namespace ttad {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

static hipStream_t pick(ttad_context* c, void* s) { return s ? (hipStream_t)s : c->stream; }

// Scratch of ttad_dtrsm_noalloc: the inverses of the 64 x 64 diagonal blocks of L (FP64, 64 x 64 per 64 columns),
// for n up to NOALLOC_TRSM_MAX_N: 2 MiB, one copy per GPU, held until the process ends. Host code uses the
// address that hipGetSymbolAddress returns for the current device (ttad_context::noalloc_dinv), never the name.
constexpr int NOALLOC_TRSM_MAX_N = 4096;
static __device__ __align__(256) double scratch[(NOALLOC_TRSM_MAX_N / DB) * DB * DB];

// ---------------------------------------------------------------------------
// Host-sequenced schedule: per panel step k one launch each for POTRF(k), all
// TRSM(i,k) and all SYRK/GEMM(k,i,j); stream order carries the dependencies.
// ---------------------------------------------------------------------------
struct StreamPlan {
    struct Batch {
        int task_off, prefix_off, n_tasks, n_items;
    };
    std::vector<Batch> batches;
    DevBuf d_tasks, d_prefix, d_err;
    int vset = N_VSETS - 1;  // executor instantiation covering the plan's item variants
};

// POTRF tasks (graphs and stream plans) run as one item per 64-row block of the tile (item_potrf_rows,
// default); TTAD_POTRF_ROWS=0 runs each as one single-workgroup item (for comparisons).
static bool potrf_rows_enabled() {
    static const bool on = [] {
        const char* v = getenv("TTAD_POTRF_ROWS");
        return !(v && v[0] == '0');
    }();
    return on;
}

static std::shared_ptr<StreamPlan> build_stream_plan(ttad_matrix* M, int prec, int& rc) {
    auto plan = std::make_shared<StreamPlan>();
    std::vector<DevTask> tasks;
    std::vector<int32_t> prefix;
    const int nt = M->nt;
    auto add_batch = [&](const std::vector<DevTask>& ts) {
        if (ts.empty()) return;
        StreamPlan::Batch b{(int)tasks.size(), (int)prefix.size(), (int)ts.size(), 0};
        for (auto& t : ts) {
            prefix.push_back(b.n_items);
            b.n_items += t.n_items;
            tasks.push_back(t);
        }
        plan->batches.push_back(b);
    };
    auto mk = [&](int kind, int i, int j, int k) {
        DevTask t{};
        t.kind = kind;
        t.prec = (kind == K_GEMM || kind == K_SYRK) ? prec : M->storage;
        t.mat = 0;
        t.i = i;
        t.j = j;
        t.k = k;
        t.n_items = kind == K_POTRF && potrf_rows_enabled() ? potrf_row_items(M->bsz[j])
                                                            : std::max(1, items_of(kind, M->bsz[i], M->bsz[j], (int)M->elem));
        t.bucket = 0;
        t.crit = 0;
        return t;
    };
    for (int k = 0; k < nt; k++) {
        add_batch({mk(K_POTRF, k, k, k)});
        std::vector<DevTask> tr;
        for (int i = k + 1; i < nt; i++) tr.push_back(mk(K_TRSM, i, k, k));
        add_batch(tr);
        std::vector<DevTask> up;
        for (int j = k + 1; j < nt; j++)
            for (int i = j; i < nt; i++) up.push_back(mk(i == j ? K_SYRK : K_GEMM, i, j, k));
        add_batch(up);
    }
    rc = TTAD_OK;
    plan->vset = vset_index(item_variant(K_POTRF, prec, M->storage) | item_variant(K_GEMM, prec, M->storage));
    if (plan->d_tasks.alloc(std::max<size_t>(1, tasks.size()) * sizeof(DevTask)) != hipSuccess ||
        plan->d_prefix.alloc(std::max<size_t>(1, prefix.size()) * sizeof(int32_t)) != hipSuccess ||
        plan->d_err.alloc(sizeof(int32_t)) != hipSuccess || hipMemset(plan->d_err.p, 0, sizeof(int32_t)) != hipSuccess) {
        rc = set_error(TTAD_ERR_OOM, "stream plan allocation failed");
        return nullptr;
    }
    if (hipMemcpy(plan->d_tasks.p, tasks.data(), tasks.size() * sizeof(DevTask), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(plan->d_prefix.p, prefix.data(), prefix.size() * sizeof(int32_t), hipMemcpyHostToDevice) !=
            hipSuccess) {
        rc = set_error(TTAD_ERR_DEVICE, "stream plan upload failed");
        return nullptr;
    }
    return plan;
}

int cholesky_persistent(ttad_context* ctx, ttad_matrix* M, int prio_kind, int prec, hipStream_t s, int* info);

}  // namespace ttad

using namespace ttad;

extern "C" {

const char* ttad_last_error(void) { return g_err.c_str(); }
int ttad_version(void) { return 1; }

int ttad_init(int device, ttad_context** out) {
    if (!out) return set_error(TTAD_ERR_INVALID, "out is NULL");
    int ndev = 0;
    TTAD_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(TTAD_ERR_INVALID, "device %d not present (%d)", device, ndev);
    TTAD_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    TTAD_HIP(hipGetDeviceProperties(&p, device));
    auto* c = new ttad_context();
    c->device = device;
    c->n_cu = p.multiProcessorCount;
    c->hbm_bytes = (long long)p.totalGlobalMem;
    c->name = p.gcnArchName;
    hipError_t e = hipStreamCreate(&c->stream);
    if (e == hipSuccess) e = kernels_init();
    if (e == hipSuccess) e = factor_check_init();
    if (e == hipSuccess) e = solve_init();
    if (e == hipSuccess) e = c->scratch_info.alloc(sizeof(int));
    // this device's copy of the no-allocation scratch (the device is current here)
    if (e == hipSuccess) e = hipGetSymbolAddress((void**)&c->noalloc_dinv, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        delete c;
        return set_error(TTAD_ERR_DEVICE, "init: %s", hipGetErrorString(e));
    }
    for (int q = 0; q < N_VSETS; q++) c->vset_workers_per_cu[q] = persistent_blocks_per_cu(q);
    c->workers_per_cu = c->vset_workers_per_cu[0];
    if (hipEventCreate(&c->ev_a) != hipSuccess || hipEventCreate(&c->ev_b) != hipSuccess) {
        delete c;
        return set_error(TTAD_ERR_DEVICE, "init: event creation failed");
    }
    *out = c;
    return TTAD_OK;
}

int ttad_finalize(ttad_context* ctx) {
    if (!ctx) return TTAD_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->chol_graphs) ttad_graph_destroy(kv.second);
    ctx->chol_graphs.clear();
    {
        auto wm = ctx->work_mats;
        ctx->work_mats.clear();
        for (auto& kv : wm) ttad_matrix_destroy(kv.second);
    }
    ctx->stream_plans.clear();
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->ev_a) (void)hipEventDestroy(ctx->ev_a);
    if (ctx->ev_b) (void)hipEventDestroy(ctx->ev_b);
    delete ctx;
    return TTAD_OK;
}

int ttad_device_info(ttad_context* ctx, int* n_cu, long long* hbm, char* name, int name_len) {
    if (!ctx) return set_error(TTAD_ERR_INVALID, "ctx is NULL");
    if (n_cu) *n_cu = ctx->n_cu;
    if (hbm) *hbm = ctx->hbm_bytes;
    if (name && name_len > 0) {
        strncpy(name, ctx->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    return TTAD_OK;
}

int ttad_synchronize(ttad_context* ctx, void* stream) {
    if (!ctx) return set_error(TTAD_ERR_INVALID, "ctx is NULL");
    TTAD_HIP(hipStreamSynchronize(pick(ctx, stream)));
    return TTAD_OK;
}

int ttad_set_option(ttad_context* ctx, int option, long long value) {
    if (!ctx) return set_error(TTAD_ERR_INVALID, "ctx is NULL");
    switch (option) {
        case TTAD_OPT_FACTOR_RESIDUAL:
            ctx->report_factor_residual = value != 0;
            return TTAD_OK;
        case TTAD_OPT_RESIDUAL_ROWS:
            ctx->residual_rows = value != 0;
            return TTAD_OK;
        default:
            return set_error(TTAD_ERR_INVALID, "set_option: unknown option %d", option);
    }
}

// ---------------------------------------------------------------------------
// Tile kernels
// ---------------------------------------------------------------------------
static int ensure_dinv(ttad_context* ctx, int n) {
    const size_t need = (size_t)((n + DB - 1) / DB) * DB * DB * sizeof(double);
    if (ctx->scratch_dinv.bytes < need) TTAD_HIP(ctx->scratch_dinv.alloc(need));
    return TTAD_OK;
}

static int potrf_common(ttad_context* ctx, int prec, int n, void* A, int lda, int* info, void* stream) {
    if (!ctx || (!A && n > 0) || n < 0 || lda < std::max(1, n)) return set_error(TTAD_ERR_INVALID, "potrf: bad args");
    if (info) *info = 0;
    if (n == 0) return TTAD_OK;
    int rc = ensure_dinv(ctx, n);
    if (rc) return rc;
    hipStream_t s = pick(ctx, stream);
    TTAD_HIP(hipMemsetAsync(ctx->scratch_info.p, 0, sizeof(int), s));
    launch_raw_potrf(prec, A, lda, n, ctx->scratch_dinv.as<double>(), ctx->scratch_info.as<int>(), s);
    TTAD_HIP(hipGetLastError());
    int h = 0;
    TTAD_HIP(hipMemcpyAsync(&h, ctx->scratch_info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TTAD_HIP(hipStreamSynchronize(s));
    if (info) *info = h;
    return h ? set_error(TTAD_ERR_NOT_SPD, "potrf: non-positive pivot at column %d", h) : TTAD_OK;
}

static int trsm_common(ttad_context* ctx, int prec, int m, int n, const void* L, int ldl, void* B, int ldb,
                       void* stream) {
    if (!ctx || m < 0 || n < 0 || ldl < std::max(1, n) || ldb < std::max(1, m))
        return set_error(TTAD_ERR_INVALID, "trsm: bad args");
    if (m == 0 || n == 0) return TTAD_OK;
    int rc = ensure_dinv(ctx, n);
    if (rc) return rc;
    hipStream_t s = pick(ctx, stream);
    launch_trtri_diag(prec, L, ldl, n, ctx->scratch_dinv.as<double>(), s);
    launch_raw_trsm(prec, L, ldl, ctx->scratch_dinv.as<double>(), B, ldb, m, n, s);
    TTAD_HIP(hipGetLastError());
    return TTAD_OK;
}

static int gemm_common(ttad_context* ctx, int prec, int m, int n, int k, double alpha, const void* A, int lda,
                       const void* B, int ldb, double beta, void* C, int ldc, int syrk, void* stream) {
    if (!ctx || m < 0 || n < 0 || k < 0 || lda < std::max(1, m) || ldb < std::max(1, n) || ldc < std::max(1, m))
        return set_error(TTAD_ERR_INVALID, "gemm/syrk: bad args");
    if (m == 0 || n == 0) return TTAD_OK;
    launch_raw_gemm(prec, A, lda, B, ldb, C, ldc, m, n, k, syrk, alpha, beta, pick(ctx, stream));
    TTAD_HIP(hipGetLastError());
    return TTAD_OK;
}

int ttad_dpotrf(ttad_context* ctx, int n, double* A, int lda, int* info, void* stream) {
    return potrf_common(ctx, P_F64, n, A, lda, info, stream);
}
int ttad_spotrf(ttad_context* ctx, int n, float* A, int lda, int* info, void* stream) {
    return potrf_common(ctx, P_F32, n, A, lda, info, stream);
}
int ttad_dtrsm(ttad_context* ctx, int m, int n, const double* L, int ldl, double* B, int ldb, void* stream) {
    return trsm_common(ctx, P_F64, m, n, L, ldl, B, ldb, stream);
}
// The kernels of ttad_dtrsm with the file-scope scratch: no allocation, no synchronization. Every call on one
// GPU shares that GPU's scratch, so calls that overlap on one GPU (other streams, other contexts) race.
int ttad_dtrsm_noalloc(ttad_context* ctx, int m, int n, const double* L, int ldl, double* B, int ldb, void* stream) {
    if (!ctx || m < 0 || n < 0 || ldl < std::max(1, n) || ldb < std::max(1, m))
        return set_error(TTAD_ERR_INVALID, "dtrsm_noalloc: bad args");
    if (n > NOALLOC_TRSM_MAX_N)
        return set_error(TTAD_ERR_INVALID, "dtrsm_noalloc: n = %d exceeds its scratch limit of %d", n,
                         NOALLOC_TRSM_MAX_N);
    if (!ctx->noalloc_dinv) return set_error(TTAD_ERR_DEVICE, "dtrsm_noalloc: no scratch address for this device");
    if (m == 0 || n == 0) return TTAD_OK;
    hipStream_t s = pick(ctx, stream);
    launch_trtri_diag(P_F64, L, ldl, n, ctx->noalloc_dinv, s);  // writes every element it later reads
    launch_raw_trsm(P_F64, L, ldl, ctx->noalloc_dinv, B, ldb, m, n, s);
    TTAD_HIP(hipGetLastError());
    return TTAD_OK;
}
int ttad_strsm(ttad_context* ctx, int m, int n, const float* L, int ldl, float* B, int ldb, void* stream) {
    return trsm_common(ctx, P_F32, m, n, L, ldl, B, ldb, stream);
}
int ttad_dsyrk(ttad_context* ctx, int n, int k, double alpha, const double* A, int lda, double beta, double* C,
               int ldc, void* stream) {
    return gemm_common(ctx, P_F64, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1, stream);
}
int ttad_ssyrk(ttad_context* ctx, int n, int k, float alpha, const float* A, int lda, float beta, float* C, int ldc,
               void* stream) {
    return gemm_common(ctx, P_F32, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1, stream);
}
int ttad_dgemm(ttad_context* ctx, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
               int ldb, double beta, double* C, int ldc, void* stream) {
    return gemm_common(ctx, P_F64, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, stream);
}
int ttad_sgemm(ttad_context* ctx, int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
               float beta, float* C, int ldc, void* stream) {
    return gemm_common(ctx, P_F32, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, stream);
}
int ttad_gemm_bf16(ttad_context* ctx, int m, int n, int k, float alpha, const float* A, int lda, const float* B,
                   int ldb, float beta, float* C, int ldc, int syrk, void* stream) {
    if (syrk && m != n) return set_error(TTAD_ERR_INVALID, "gemm_bf16: syrk needs m == n");
    return gemm_common(ctx, P_BF16, m, n, k, alpha, A, lda, syrk ? A : B, syrk ? lda : ldb, beta, C, ldc, syrk,
                       stream);
}

int ttad_gemm_bf16x3(ttad_context* ctx, int m, int n, int k, float alpha, const float* A, int lda, const float* B,
                     int ldb, float beta, float* C, int ldc, int syrk, void* stream) {
    if (syrk && m != n) return set_error(TTAD_ERR_INVALID, "gemm_bf16x3: syrk needs m == n");
    return gemm_common(ctx, P_BF16X3, m, n, k, alpha, A, lda, syrk ? A : B, syrk ? lda : ldb, beta, C, ldc, syrk,
                       stream);
}

// ---------------------------------------------------------------------------
// Tiled matrices
// ---------------------------------------------------------------------------
int ttad_matrix_create(ttad_context* ctx, int num_blocks, const int* block_sizes, int storage, ttad_matrix** out) {
    if (!ctx || !out || num_blocks <= 0 || !block_sizes || (storage != TTAD_F64 && storage != TTAD_F32))
        return set_error(TTAD_ERR_INVALID, "matrix_create: bad args");
    auto M = std::make_unique<ttad_matrix>();
    M->ctx = ctx;
    M->id = ctx->next_matrix_id++;
    M->nt = num_blocks;
    M->storage = storage;
    M->elem = storage == TTAD_F64 ? 8 : 4;
    M->bsz.assign(block_sizes, block_sizes + num_blocks);
    M->boff.assign(num_blocks + 1, 0);
    for (int i = 0; i < num_blocks; i++) {
        if (block_sizes[i] < 0 || block_sizes[i] > 4096)
            return set_error(TTAD_ERR_INVALID, "matrix_create: block size %d out of [0, 4096]", block_sizes[i]);
        M->boff[i + 1] = M->boff[i] + block_sizes[i];
    }
    M->n = M->boff[num_blocks];
    const int nt = num_blocks;
    const size_t align = 256 / M->elem;
    M->tile_off.assign((size_t)nt * nt, 0);
    size_t off = 0;
    for (int j = 0; j < nt; j++)
        for (int i = j; i < nt; i++) {
            M->tile_off[i + (size_t)j * nt] = off;
            off += ((size_t)M->bsz[i] * M->bsz[j] + align - 1) / align * align;
        }
    size_t dinv_per = 0;
    std::vector<size_t> dinv_off(nt);
    for (int j = 0; j < nt; j++) {
        dinv_off[j] = dinv_per;
        dinv_per += (size_t)((M->bsz[j] + DB - 1) / DB) * DB * DB;
    }
    TTAD_HIP(hipSetDevice(ctx->device));
    if (M->tiles.alloc(off * M->elem) != hipSuccess || M->dinv.alloc(std::max<size_t>(1, dinv_per) * 8) != hipSuccess)
        return set_error(TTAD_ERR_OOM, "matrix_create: %zu bytes of tiles do not fit", off * M->elem);
    M->bytes = (long long)(off * M->elem + dinv_per * 8);
    std::vector<void*> tp((size_t)nt * nt, nullptr);
    std::vector<double*> dp(nt);
    for (int j = 0; j < nt; j++) {
        dp[j] = M->dinv.as<double>() + dinv_off[j];
        for (int i = j; i < nt; i++)
            tp[i + (size_t)j * nt] = (char*)M->tiles.p + M->tile_off[i + (size_t)j * nt] * M->elem;
    }
    TTAD_HIP(M->tab_bsz.alloc(nt * 4));
    TTAD_HIP(M->tab_boff.alloc((nt + 1) * 4));
    TTAD_HIP(M->tab_tiles.alloc((size_t)nt * nt * sizeof(void*)));
    TTAD_HIP(M->tab_dinv.alloc(nt * sizeof(double*)));
    TTAD_HIP(M->info.alloc(sizeof(int)));
    TTAD_HIP(M->pflags.alloc((size_t)nt * POTRF_FLAGS * sizeof(int32_t)));
    TTAD_HIP(M->locks.alloc((size_t)nt * nt * 16 * sizeof(int32_t)));
    TTAD_HIP(M->desc.alloc(sizeof(DevMat)));
    TTAD_HIP(hipMemcpy(M->tab_bsz.p, M->bsz.data(), nt * 4, hipMemcpyHostToDevice));
    TTAD_HIP(hipMemcpy(M->tab_boff.p, M->boff.data(), (nt + 1) * 4, hipMemcpyHostToDevice));
    TTAD_HIP(hipMemcpy(M->tab_tiles.p, tp.data(), tp.size() * sizeof(void*), hipMemcpyHostToDevice));
    TTAD_HIP(hipMemcpy(M->tab_dinv.p, dp.data(), dp.size() * sizeof(double*), hipMemcpyHostToDevice));
    TTAD_HIP(hipMemset(M->info.p, 0, sizeof(int)));
    TTAD_HIP(hipMemset(M->pflags.p, 0, (size_t)nt * POTRF_FLAGS * sizeof(int32_t)));
    TTAD_HIP(hipMemset(M->locks.p, 0, (size_t)nt * nt * 16 * sizeof(int32_t)));
    TTAD_HIP(hipMemset(M->tiles.p, 0, off * M->elem));
    DevMat d{};
    d.nt = nt;
    d.storage = storage;
    d.bsz = M->tab_bsz.as<int32_t>();
    d.boff = M->tab_boff.as<int32_t>();
    d.tiles = M->tab_tiles.as<void*>();
    d.dinv = M->tab_dinv.as<double*>();
    d.info = M->info.as<int32_t>();
    d.pflags = M->pflags.as<int32_t>();
    d.locks = M->locks.as<int32_t>();
    M->host_desc = d;
    TTAD_HIP(hipMemcpy(M->desc.p, &d, sizeof(DevMat), hipMemcpyHostToDevice));
    *out = M.release();
    return TTAD_OK;
}

int ttad_matrix_destroy(ttad_matrix* M) {
    if (!M) return TTAD_OK;
    ttad_context* ctx = M->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto it = ctx->work_mats.begin(); it != ctx->work_mats.end();) {
        if (it->first.first == M->id) {  // working copies derived from M go with it
            ttad_matrix* W = it->second;
            it = ctx->work_mats.erase(it);
            ttad_matrix_destroy(W);
        } else if (it->second == M) {
            it = ctx->work_mats.erase(it);
        } else {
            ++it;
        }
    }
    for (auto it = ctx->stream_plans.begin(); it != ctx->stream_plans.end();)
        it = (it->first.first == M->id) ? ctx->stream_plans.erase(it) : std::next(it);
    for (auto it = ctx->chol_graphs.begin(); it != ctx->chol_graphs.end();) {
        if (std::get<0>(it->first) == M->id) {
            ttad_graph_destroy(it->second);
            it = ctx->chol_graphs.erase(it);
        } else {
            ++it;
        }
    }
    delete M;
    return TTAD_OK;
}

int ttad_matrix_set(ttad_matrix* M, const double* A, long long lda, int src_device, void* stream) {
    if (!M || !A || lda < std::max(1, M->n)) return set_error(TTAD_ERR_INVALID, "matrix_set: bad args");
    M->anorm_valid = false;
    hipStream_t s = pick(M->ctx, stream);
    if (src_device) {
        launch_pack(M->storage, A, lda, M->desc.as<DevMat>(), M->nt, s);
        TTAD_HIP(hipGetLastError());
        return TTAD_OK;
    }
    DevBuf tmp;
    if (tmp.alloc((size_t)M->n * M->n * 8) != hipSuccess) return set_error(TTAD_ERR_OOM, "matrix_set: staging");
    TTAD_HIP(hipMemcpy2DAsync(tmp.p, (size_t)M->n * 8, A, (size_t)lda * 8, (size_t)M->n * 8, M->n,
                              hipMemcpyHostToDevice, s));
    launch_pack(M->storage, tmp.as<double>(), M->n, M->desc.as<DevMat>(), M->nt, s);
    TTAD_HIP(hipGetLastError());
    TTAD_HIP(hipStreamSynchronize(s));
    return TTAD_OK;
}

int ttad_matrix_get(ttad_matrix* M, double* A, long long lda, int dst_device, void* stream) {
    if (!M || !A || lda < std::max(1, M->n)) return set_error(TTAD_ERR_INVALID, "matrix_get: bad args");
    hipStream_t s = pick(M->ctx, stream);
    if (dst_device) {
        launch_unpack(M->storage, A, lda, M->desc.as<DevMat>(), M->nt, 1, s);
        TTAD_HIP(hipGetLastError());
        return TTAD_OK;
    }
    DevBuf tmp;
    if (tmp.alloc((size_t)M->n * M->n * 8) != hipSuccess) return set_error(TTAD_ERR_OOM, "matrix_get: staging");
    launch_unpack(M->storage, tmp.as<double>(), M->n, M->desc.as<DevMat>(), M->nt, 1, s);
    TTAD_HIP(hipGetLastError());
    TTAD_HIP(hipMemcpy2DAsync(A, (size_t)lda * 8, tmp.p, (size_t)M->n * 8, (size_t)M->n * 8, M->n,
                              hipMemcpyDeviceToHost, s));
    TTAD_HIP(hipStreamSynchronize(s));
    return TTAD_OK;
}

int ttad_matrix_copy(ttad_matrix* dst, const ttad_matrix* src, void* stream) {
    if (!dst || !src || dst->nt != src->nt || dst->bsz != src->bsz)
        return set_error(TTAD_ERR_INVALID, "matrix_copy: partitions differ");
    dst->anorm_valid = false;
    launch_copy_tiles(dst->storage, src->storage, src->desc.as<DevMat>(), dst->desc.as<DevMat>(), dst->nt,
                      pick(dst->ctx, stream));
    TTAD_HIP(hipGetLastError());
    return TTAD_OK;
}

int ttad_matrix_info(const ttad_matrix* M, int* n, int* num_blocks, int* storage, long long* bytes) {
    if (!M) return set_error(TTAD_ERR_INVALID, "matrix is NULL");
    if (n) *n = M->n;
    if (num_blocks) *num_blocks = M->nt;
    if (storage) *storage = M->storage;
    if (bytes) *bytes = M->bytes;
    return TTAD_OK;
}

int ttad_matrix_tile(const ttad_matrix* M, int i, int j, void** ptr, int* ld) {
    if (!M || i < 0 || j < 0 || i >= M->nt || j > i) return set_error(TTAD_ERR_INVALID, "matrix_tile: (%d,%d)", i, j);
    if (ptr) *ptr = (char*)M->tiles.p + M->tile_off[i + (size_t)j * M->nt] * M->elem;
    if (ld) *ld = std::max(1, M->bsz[i]);
    return TTAD_OK;
}

// ---------------------------------------------------------------------------
// Tiled Cholesky
// ---------------------------------------------------------------------------
int ttad_potrf_tiled(ttad_context* ctx, ttad_matrix* M, int schedule, int prio_kind, int prec, void* stream,
                     int* info) {
    if (!ctx || !M || (prio_kind & ~TTAD_ACCUMULATE) < 0 || (prio_kind & ~TTAD_ACCUMULATE) > 4)
        return set_error(TTAD_ERR_INVALID, "potrf_tiled: bad args");
    if ((prec == TTAD_F64) != (M->storage == TTAD_F64) || prec < 0 || prec > 3)
        return set_error(TTAD_ERR_INVALID, "potrf_tiled: precision %d does not match the tile storage", prec);
    if (info) *info = 0;
    hipStream_t s = pick(ctx, stream);
    TTAD_HIP(hipMemsetAsync(M->info.p, 0, sizeof(int), s));
    TTAD_HIP(hipMemsetAsync(M->pflags.p, 0, (size_t)M->nt * POTRF_FLAGS * sizeof(int32_t), s));
    int rc = TTAD_OK;
    ctx->ev_valid = false;
    ctx->kernel_s = -1.0;
    TTAD_HIP(hipEventRecord(ctx->ev_a, s));
    if (schedule == TTAD_SCHED_STREAM) {
        auto key = std::make_pair(M->id, prec);
        auto it = ctx->stream_plans.find(key);
        std::shared_ptr<StreamPlan> plan;
        if (it == ctx->stream_plans.end()) {
            plan = build_stream_plan(M, prec, rc);
            if (!plan) return rc;
            ctx->stream_plans[key] = plan;
        } else {
            plan = it->second;
        }
        for (auto& b : plan->batches)
            launch_items(plan->vset, M->desc.as<DevMat>(), plan->d_tasks.as<DevTask>() + b.task_off,
                         plan->d_prefix.as<int32_t>() + b.prefix_off, b.n_tasks, b.n_items, plan->d_err.as<int32_t>(),
                         s);
        TTAD_HIP(hipGetLastError());
        int err = 0;
        TTAD_HIP(hipMemcpyAsync(&err, plan->d_err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        TTAD_HIP(hipStreamSynchronize(s));
        if (err) {
            (void)hipMemset(plan->d_err.p, 0, sizeof(int32_t));
            if (err == 1) return set_error(TTAD_ERR_TIMEOUT, "potrf_tiled: a row-block POTRF item waited too long");
            return set_error(TTAD_ERR_DEVICE, "potrf_tiled: item variant missing from the executor (%d)", err);
        }
    } else if (schedule == TTAD_SCHED_PERSISTENT) {
        rc = cholesky_persistent(ctx, M, prio_kind, prec, s, info);
        if (rc != TTAD_OK && rc != TTAD_ERR_NOT_SPD) return rc;
    } else {
        return set_error(TTAD_ERR_INVALID, "potrf_tiled: unknown schedule %d", schedule);
    }
    TTAD_HIP(hipEventRecord(ctx->ev_b, s));
    ctx->ev_valid = true;
    int h = 0;
    TTAD_HIP(hipMemcpyAsync(&h, M->info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TTAD_HIP(hipStreamSynchronize(s));
    if (info) *info = h;
    if (h) return set_error(TTAD_ERR_NOT_SPD, "potrf_tiled: non-positive pivot at column %d", h);
    return TTAD_OK;
}

int ttad_last_elapsed(ttad_context* ctx, double* seconds) {
    if (!ctx || !seconds) return set_error(TTAD_ERR_INVALID, "last_elapsed: bad args");
    if (!ctx->ev_valid) return set_error(TTAD_ERR_INVALID, "last_elapsed: nothing timed yet");
    if (ctx->kernel_s >= 0.0) {  // persistent schedule: the executor launch itself
        *seconds = ctx->kernel_s;
        return TTAD_OK;
    }
    TTAD_HIP(hipEventSynchronize(ctx->ev_b));
    float ms = 0.f;
    TTAD_HIP(hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b));
    *seconds = ms * 1e-3;
    return TTAD_OK;
}

// ---------------------------------------------------------------------------
// Device task graphs
// ---------------------------------------------------------------------------
// Marks the tasks on a longest path of the DAG, weighting each kind by the latency of its items
// (POTRF 8, TRSM 3, GEMM/SYRK 1: one workgroup factors a whole tile, a TRSM item solves 64 rows
// block column by block column). Those items run with no other item started on their CU.
// TTAD_CRITICAL=0 turns it off (for comparisons).
static void mark_critical(ttad_graph& g) {
    static const bool on = [] {
        const char* v = getenv("TTAD_CRITICAL");
        return !(v && v[0] == '0');
    }();
    const int n = g.n_tasks;
    for (auto& t : g.tasks) t.crit = 0;
    if (!on || n == 0) return;
    std::vector<int> w(n), indeg(n, 0), order;
    order.reserve(n);
    for (int t = 0; t < n; t++) {
        const int kd = g.tasks[t].kind;
        w[t] = kd == K_POTRF ? 8 : kd == K_TRSM ? 3 : 1;
    }
    for (int v : g.out_idx) indeg[v]++;
    for (int t = 0; t < n; t++)
        if (indeg[t] == 0) order.push_back(t);
    for (size_t q = 0; q < order.size(); q++) {
        const int t = order[q];
        for (int e = g.out_ptr[t]; e < g.out_ptr[t + 1]; e++)
            if (--indeg[g.out_idx[e]] == 0) order.push_back(g.out_idx[e]);
    }
    if ((int)order.size() != n) return;  // not a DAG: no marking (the graph would not complete anyway)
    std::vector<long long> top(n, 0), bot(n, 0);
    for (int t : order) {
        top[t] += w[t];
        for (int e = g.out_ptr[t]; e < g.out_ptr[t + 1]; e++)
            top[g.out_idx[e]] = std::max(top[g.out_idx[e]], top[t]);
    }
    long long len = 0;
    for (int q = n - 1; q >= 0; q--) {
        const int t = order[q];
        long long b = 0;
        for (int e = g.out_ptr[t]; e < g.out_ptr[t + 1]; e++) b = std::max(b, bot[g.out_idx[e]]);
        bot[t] = b + w[t];
        len = std::max(len, top[t]);
    }
    for (int t = 0; t < n; t++) g.tasks[t].crit = top[t] + bot[t] - w[t] == len;
}

// Which GEMM/SYRK work queues are homed per XCD. By default the FP32 / BF16 / bf16x3 updates are
// (measured at n = 32768: bf16x3 time to solution 128.9 ms homed, 144.2 ms shared) and the FP64 ones share
// one queue per priority bucket (FP64 factorization 223.7 ms shared, 229.4 ms homed); spin tasks are
// homed. TTAD_XCD_AFFINITY=0 shares every queue, TTAD_XCD_AFFINITY=1 homes FP64 updates too.
static bool xcd_homed(int kind, int prec) {
    static const int mode = [] {
        const char* v = getenv("TTAD_XCD_AFFINITY");
        return v ? (v[0] == '0' ? 0 : 1) : -1;
    }();
    if (kind == K_NOP) return mode != 0;
    if (kind != K_GEMM && kind != K_SYRK) return false;
    return mode == 1 || (mode < 0 && prec != P_F64);
}

int ttad_graph_create(ttad_context* ctx, int n_tasks, const ttad_task_desc* tasks, const int32_t* indegree,
                      const int32_t* out_ptr, const int32_t* out_idx, const double* priority, ttad_graph** out) {
    if (!ctx || !out || n_tasks <= 0 || !tasks || !indegree || !out_ptr)
        return set_error(TTAD_ERR_INVALID, "graph_create: bad args");
    if (n_tasks >= (1 << 23)) return set_error(TTAD_ERR_INVALID, "graph_create: too many tasks");
    auto g = std::make_unique<ttad_graph>();
    g->ctx = ctx;
    g->n_tasks = n_tasks;
    g->tasks.resize(n_tasks);
    g->spin_us.assign(n_tasks, 0);
    g->indegree.assign(indegree, indegree + n_tasks);
    g->out_ptr.assign(out_ptr, out_ptr + n_tasks + 1);
    if (g->out_ptr[0] != 0) return set_error(TTAD_ERR_INVALID, "graph_create: out_ptr[0] != 0");
    const int n_edges = g->out_ptr[n_tasks];
    if (n_edges < 0 || (n_edges > 0 && !out_idx)) return set_error(TTAD_ERR_INVALID, "graph_create: bad edges");
    if (n_edges > 0) g->out_idx.assign(out_idx, out_idx + n_edges);
    for (int t = 0; t < n_tasks; t++) {
        if (g->out_ptr[t + 1] < g->out_ptr[t]) return set_error(TTAD_ERR_INVALID, "graph_create: out_ptr not sorted");
        if (indegree[t] < 1) return set_error(TTAD_ERR_GRAPH, "graph_create: task %d has indegree < 1", t);
        const ttad_task_desc& d = tasks[t];
        const bool kind_ok = (d.kind >= TTAD_POTRF && d.kind <= TTAD_GEMM) || d.kind == TTAD_SPIN ||
                             d.kind == TTAD_GEMM_ACC || d.kind == TTAD_SYRK_ACC;
        if (!kind_ok || d.prec < 0 || d.prec > 3 ||
            d.mat < 0 || d.mat > 65535 || (d.kind == TTAD_SPIN && (d.k < 0 || d.k > 10000000)))
            return set_error(TTAD_ERR_INVALID, "graph_create: bad descriptor for task %d", t);
        DevTask& x = g->tasks[t];
        x.kind = d.kind == TTAD_GEMM_ACC ? K_GEMM : d.kind == TTAD_SYRK_ACC ? K_SYRK : d.kind;
        x.accum = d.kind == TTAD_GEMM_ACC || d.kind == TTAD_SYRK_ACC;
        x.prec = d.prec;
        x.mat = d.mat;
        x.i = d.i;
        x.j = d.j;
        x.k = d.k;
        x.n_items = 1;
        x.bucket = 0;
        x.crit = 0;
        g->spin_us[t] = d.kind == TTAD_SPIN ? d.k : 0;
    }
    for (int e = 0; e < n_edges; e++)
        if (g->out_idx[e] < 0 || g->out_idx[e] >= n_tasks)
            return set_error(TTAD_ERR_INVALID, "graph_create: edge target %d out of range", g->out_idx[e]);
    mark_critical(*g);
    // priority -> bucket: distinct priorities ranked highest first, quantized to at most 64
    // buckets (fewer buckets hold more items each: fewer contended claims, measured faster than 256)
    std::vector<double> pr(n_tasks, 0.0);
    if (priority) pr.assign(priority, priority + n_tasks);
    std::vector<double> uniq = pr;
    std::sort(uniq.begin(), uniq.end(), std::greater<double>());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    const int nu = (int)uniq.size();
    const int cap = QUEUE_BUCKETS;
    const int n_prio = std::min(nu, cap);
    std::vector<int> pb(n_tasks), home(n_tasks, -1);
    for (int t = 0; t < n_tasks; t++) {
        const int rank = (int)(std::lower_bound(uniq.begin(), uniq.end(), pr[t], std::greater<double>()) - uniq.begin());
        pb[t] = (nu <= cap) ? rank : (int)((long long)rank * cap / nu);
        const DevTask& x = g->tasks[t];
        // GEMM/SYRK items are homed by output tile row: the items of one update share their
        // A and B panels, and the updates of one row share A, in one XCD's L2
        // spin tasks are homed like GEMM/SYRK (by their row index), so that tiny tasks spread their claims
        if (xcd_homed(x.kind, x.prec)) home[t] = x.i % NXCD;
    }
    // work queues: bucket b's shared queue is shared_queue(b), its queue homed on XCD x homed_queue(x, b)
    // (the executors scan the shared and own-XCD queues by bucket); the other ids stay empty
    static_assert(QUEUE_BUCKETS >= 64, "a bucket per priority level");
    g->q_home.assign(MAX_QUEUES, -1);
    for (int x = 0; x < NXCD; x++)
        for (int b = 0; b < QUEUE_BUCKETS; b++) g->q_home[homed_queue(x, b)] = x;
    g->n_buckets = MAX_QUEUES;
    for (int t = 0; t < n_tasks; t++) g->tasks[t].bucket = home[t] < 0 ? shared_queue(pb[t]) : homed_queue(home[t], pb[t]);
    g->seeds.assign(n_tasks, 0);
    TTAD_HIP(hipEventCreate(&g->ev_start));
    TTAD_HIP(hipEventCreate(&g->ev_stop));
    *out = g.release();
    return TTAD_OK;
}

int ttad_graph_bind(ttad_graph* g, int slot, ttad_matrix* M) {
    if (!g || slot < 0 || slot > 65535 || !M) return set_error(TTAD_ERR_INVALID, "graph_bind: bad args");
    if (g->running) return set_error(TTAD_ERR_BUSY, "graph_bind: graph is running");
    if ((int)g->mats.size() <= slot) g->mats.resize(slot + 1, nullptr);
    if (g->mats[slot] != M) g->mats_dirty = true;
    g->mats[slot] = M;
    return TTAD_OK;
}

int ttad_graph_seed(ttad_graph* g, int task) {
    if (!g || task < 0 || task >= g->n_tasks) return set_error(TTAD_ERR_INVALID, "graph_seed: bad task");
    if (g->running) return set_error(TTAD_ERR_BUSY, "graph_seed: graph is running");
    g->seeds[task]++;
    g->seeds_dirty = true;
    return TTAD_OK;
}

int ttad_graph_clear_seeds(ttad_graph* g) {
    if (!g) return set_error(TTAD_ERR_INVALID, "graph is NULL");
    std::fill(g->seeds.begin(), g->seeds.end(), 0);
    g->seeds_dirty = true;
    return TTAD_OK;
}

static int graph_prepare(ttad_graph* g) {
    const int n = g->n_tasks;
    if (g->mats_dirty) {
        int rate_khz = 100000;
        (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, g->ctx->device);
        for (int t = 0; t < n; t++) {
            DevTask& x = g->tasks[t];
            if (x.kind == K_NOP) {  // spin task: k microseconds, stored as device clock ticks
                x.n_items = 1;
                x.mat = 0;
                x.k = (int32_t)std::min<long long>((long long)g->spin_us[t] * rate_khz / 1000, INT32_MAX);
                continue;
            }
            if (x.mat >= (int)g->mats.size() || !g->mats[x.mat])
                return set_error(TTAD_ERR_INVALID, "graph: task %d uses unbound matrix slot %d", t, x.mat);
            ttad_matrix* M = g->mats[x.mat];
            if (x.i < 0 || x.j < 0 || x.k < 0 || x.i >= M->nt || x.j >= M->nt || x.k >= M->nt || x.j > x.i)
                return set_error(TTAD_ERR_INVALID, "graph: task %d tile (%d,%d,%d) outside the matrix", t, x.i, x.j,
                                 x.k);
            if ((x.kind == K_GEMM || x.kind == K_SYRK) && ((x.prec == P_F64) != (M->storage == P_F64)))
                return set_error(TTAD_ERR_INVALID, "graph: task %d precision does not match tile storage", t);
            x.n_items = x.kind == K_POTRF && potrf_rows_enabled() ? potrf_row_items(M->bsz[x.j])
                                                                  : std::max(1, items_of(x.kind, M->bsz[x.i], M->bsz[x.j], (int)M->elem));
        }
        int needed = 0;
        for (auto& x : g->tasks)
            if (x.kind != K_NOP) needed |= item_variant(x.kind, x.prec, g->mats[x.mat]->storage) | (x.accum ? V_ACC : 0);
        g->vset = vset_index(needed);
        g->bucket_items.assign(g->n_buckets, 0);
        // queue entries hold (task << 20) | item in 64 bits and slots are 32-bit indices
        long long total = 0;
        for (auto& x : g->tasks) {
            if (x.n_items >= (1 << 20))
                return set_error(TTAD_ERR_INVALID, "graph: a task has %d work items (limit 2^20)", x.n_items);
            total += x.n_items;
        }
        if (total >= (1LL << 31) - 1)
            return set_error(TTAD_ERR_INVALID, "graph: %lld work items in all (limit 2^31 - 1)", total);
        g->total_items = (int)total;
        for (auto& x : g->tasks) g->bucket_items[x.bucket] += x.n_items;
        g->q_base.assign(g->n_buckets + 1, 0);
        for (int b = 0; b < g->n_buckets; b++) g->q_base[b + 1] = g->q_base[b] + g->bucket_items[b];
        std::vector<DevMat> dm(g->mats.size());
        for (size_t s = 0; s < g->mats.size(); s++)
            if (g->mats[s]) {
                dm[s] = g->mats[s]->host_desc;
                if (g->with_planes && g->mats[s]->tab_planes.p) dm[s].planes = g->mats[s]->tab_planes.as<void*>();
            }
        const int ne = std::max(1, g->out_ptr[n]);
        TTAD_HIP(g->d_tasks.alloc(n * sizeof(DevTask)));
        TTAD_HIP(g->d_out_ptr.alloc((n + 1) * 4));
        TTAD_HIP(g->d_out_idx.alloc(ne * 4));
        TTAD_HIP(g->d_deps0.alloc(n * 4));
        TTAD_HIP(g->d_deps.alloc(n * 4));
        TTAD_HIP(g->d_items_left.alloc(n * 4));
        TTAD_HIP(g->d_gops.alloc((size_t)n * sizeof(GemmOperands)));
        TTAD_HIP(g->d_tail0.alloc(g->n_buckets * 4));
        TTAD_HIP(g->d_q_head.alloc(g->n_buckets * 4));
        TTAD_HIP(g->d_q_tail.alloc(g->n_buckets * 4));
        TTAD_HIP(g->d_q_count.alloc(g->n_buckets * 4));
        TTAD_HIP(g->d_xcd_run.alloc((size_t)n * NXCD * 4));
        TTAD_HIP(g->d_q_base.alloc((g->n_buckets + 1) * 4));
        TTAD_HIP(g->d_q_home.alloc(g->n_buckets * 4));
        TTAD_HIP(g->d_q_items.alloc((size_t)g->total_items * 8));
        TTAD_HIP(g->d_done.alloc(4));
        TTAD_HIP(g->d_abort.alloc(4));
        TTAD_HIP(g->d_trace_cursor.alloc(4));
        TTAD_HIP(g->d_cu_slots.alloc((2 * CU_KEYS + NXCD * 64) * 4));  // slots, critical counts, idle primaries
        TTAD_HIP(g->d_mats.alloc(dm.size() * sizeof(DevMat)));
        TTAD_HIP(hipMemcpy(g->d_tasks.p, g->tasks.data(), n * sizeof(DevTask), hipMemcpyHostToDevice));
        TTAD_HIP(hipMemcpy(g->d_out_ptr.p, g->out_ptr.data(), (n + 1) * 4, hipMemcpyHostToDevice));
        if (!g->out_idx.empty())
            TTAD_HIP(hipMemcpy(g->d_out_idx.p, g->out_idx.data(), g->out_idx.size() * 4, hipMemcpyHostToDevice));
        TTAD_HIP(hipMemcpy(g->d_q_base.p, g->q_base.data(), (g->n_buckets + 1) * 4, hipMemcpyHostToDevice));
        TTAD_HIP(hipMemcpy(g->d_q_home.p, g->q_home.data(), g->n_buckets * 4, hipMemcpyHostToDevice));
        TTAD_HIP(hipMemcpy(g->d_mats.p, dm.data(), dm.size() * sizeof(DevMat), hipMemcpyHostToDevice));
        g->mats_dirty = false;
        g->seeds_dirty = true;
    }
    if (g->seeds_dirty) {
        // every task must receive exactly indegree promises: edges + host seeds
        std::vector<int> incoming(n, 0);
        for (int v : g->out_idx) incoming[v]++;
        for (int t = 0; t < n; t++)
            if (incoming[t] + g->seeds[t] != g->indegree[t])
                return set_error(TTAD_ERR_GRAPH, "graph: task %d has indegree %d but receives %d edges + %d seeds", t,
                                 g->indegree[t], incoming[t], g->seeds[t]);
        std::vector<int32_t> deps0(n), tail0(g->n_buckets, 0);
        std::vector<int64_t> init;  // queue entries (task << 20) | item, as the device pushes them
        std::vector<int32_t> init_slot;
        for (int t = 0; t < n; t++) {
            deps0[t] = g->indegree[t] - g->seeds[t];
            if (deps0[t] == 0) {
                const DevTask& x = g->tasks[t];
                for (int r = 0; r < x.n_items; r++) {
                    init_slot.push_back(g->q_base[x.bucket] + tail0[x.bucket]++);
                    init.push_back(((int64_t)t << 20) | r);
                }
            }
        }
        if (init.empty()) return set_error(TTAD_ERR_GRAPH, "graph: no task is ready (seed one with ttad_graph_seed)");
        g->n_init = (int)init.size();
        TTAD_HIP(g->d_init.alloc(init.size() * 8));
        TTAD_HIP(g->d_init_slot.alloc(init_slot.size() * 4));
        TTAD_HIP(hipMemcpy(g->d_init.p, init.data(), init.size() * 8, hipMemcpyHostToDevice));
        TTAD_HIP(hipMemcpy(g->d_init_slot.p, init_slot.data(), init_slot.size() * 4, hipMemcpyHostToDevice));
        TTAD_HIP(hipMemcpy(g->d_deps0.p, deps0.data(), n * 4, hipMemcpyHostToDevice));
        TTAD_HIP(hipMemcpy(g->d_tail0.p, tail0.data(), g->n_buckets * 4, hipMemcpyHostToDevice));
        g->seeds_dirty = false;
    }
    return TTAD_OK;
}

static DevGraph graph_image(ttad_graph* g, bool trace, int64_t spin_limit) {
    DevGraph d{};
    d.n_tasks = g->n_tasks;
    d.n_buckets = g->n_buckets;
    d.total_items = g->total_items;
    d.n_init = g->n_init;
    d.tasks = g->d_tasks.as<DevTask>();
    d.out_ptr = g->d_out_ptr.as<int32_t>();
    d.out_idx = g->d_out_idx.as<int32_t>();
    d.deps0 = g->d_deps0.as<int32_t>();
    d.tail0 = g->d_tail0.as<int32_t>();
    d.init_items = g->d_init.as<int64_t>();
    d.init_slots = g->d_init_slot.as<int32_t>();
    d.deps = g->d_deps.as<int32_t>();
    d.items_left = g->d_items_left.as<int32_t>();
    d.gops = g->d_gops.as<GemmOperands>();
    d.q_head = g->d_q_head.as<int32_t>();
    d.q_tail = g->d_q_tail.as<int32_t>();
    d.q_count = g->d_q_count.as<int32_t>();
    d.xcd_run = g->d_xcd_run.as<int32_t>();
    d.q_base = g->d_q_base.as<int32_t>();
    d.q_home = g->d_q_home.as<int32_t>();
    d.q_items = g->d_q_items.as<int64_t>();
    d.done_items = g->d_done.as<int32_t>();
    d.abort_flag = g->d_abort.as<int32_t>();
    d.mats = g->d_mats.as<DevMat>();
    d.trace = trace ? g->d_trace.as<uint64_t>() : nullptr;
    d.trace_cursor = g->d_trace_cursor.as<int32_t>();
    d.cu_slots = g->d_cu_slots.as<int32_t>();
    d.cu_crit = d.cu_slots + CU_KEYS;
    d.idle_primary = d.cu_slots + 2 * CU_KEYS;
    d.spin_limit = spin_limit;
    return d;
}

int ttad_graph_launch(ttad_graph* g, int n_workers, int trace, void* stream) {
    if (!g) return set_error(TTAD_ERR_INVALID, "graph is NULL");
    if (g->running) return set_error(TTAD_ERR_BUSY, "graph_launch: already running");
    ttad_context* ctx = g->ctx;
    int rc = graph_prepare(g);
    if (rc) return rc;
    if (trace) TTAD_HIP(g->d_trace.alloc((size_t)g->total_items * TRACE_WORDS * 8));
    g->traced = trace != 0;
    int rate_khz = 100000;
    (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device);
    const int64_t spin_limit = (int64_t)rate_khz * 1000 * 5;  // 5 s without finding work
    const DevGraph d = graph_image(g, trace != 0, spin_limit);
    static const int env_workers = [] {  // TTAD_WORKERS: default worker count override (tuning)
        const char* v = getenv("TTAD_WORKERS");
        return v ? atoi(v) : 0;
    }();
    const int dflt = env_workers > 0 ? std::min(env_workers, ctx->n_cu * ctx->vset_workers_per_cu[g->vset])
                                     : ctx->n_cu * ctx->vset_workers_per_cu[g->vset];
    g->workers = n_workers > 0 ? n_workers : dflt;
    g->stream = pick(ctx, stream);
    launch_graph_reset(d, g->total_items, g->stream);
    for (auto* M : g->mats)
        if (M) {
            TTAD_HIP(hipMemsetAsync(M->info.p, 0, sizeof(int), g->stream));
            // row-block POTRF flags: cleared by every completed task, and here after a stopped run
            TTAD_HIP(hipMemsetAsync(M->pflags.p, 0, (size_t)M->nt * POTRF_FLAGS * sizeof(int32_t), g->stream));
            TTAD_HIP(hipMemsetAsync(M->locks.p, 0, (size_t)M->nt * M->nt * 16 * sizeof(int32_t), g->stream));
        }
    TTAD_HIP(hipEventRecord(g->ev_start, g->stream));
    launch_persistent(g->vset, d, g->workers, g->stream);
    TTAD_HIP(hipGetLastError());
    TTAD_HIP(hipEventRecord(g->ev_stop, g->stream));
    g->running = true;
    return TTAD_OK;
}

int ttad_graph_wait(ttad_graph* g, double timeout_s, ttad_graph_stats* stats) {
    if (!g) return set_error(TTAD_ERR_INVALID, "graph is NULL");
    if (!g->running) return set_error(TTAD_ERR_INVALID, "graph_wait: not launched");
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        hipError_t e = hipEventQuery(g->ev_stop);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) {
            g->running = false;
            return set_error(TTAD_ERR_DEVICE, "graph_wait: %s", hipGetErrorString(e));
        }
        if (timeout_s > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > timeout_s)
            return set_error(TTAD_ERR_TIMEOUT, "graph_wait: timed out after %.1f s (graph still running)", timeout_s);
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    g->running = false;
    int done = 0, abort_code = 0;
    TTAD_HIP(hipMemcpy(&done, g->d_done.p, 4, hipMemcpyDeviceToHost));
    TTAD_HIP(hipMemcpy(&abort_code, g->d_abort.p, 4, hipMemcpyDeviceToHost));
    float ms = 0.f;
    TTAD_HIP(hipEventElapsedTime(&ms, g->ev_start, g->ev_stop));
    if (stats) {
        stats->elapsed_s = ms * 1e-3;
        stats->items_done = done;
        stats->items_total = g->total_items;
        stats->workers = g->workers;
        stats->abort_code = abort_code;
    }
    if (abort_code)
        return set_error(abort_code == 1 ? TTAD_ERR_TIMEOUT : TTAD_ERR_DEVICE,
                         "graph aborted on the device (code %d, %d of %d items done)", abort_code, done,
                         g->total_items);
    if (done != g->total_items)
        return set_error(TTAD_ERR_DEVICE, "graph finished with %d of %d items", done, g->total_items);
    for (auto* M : g->mats) {
        if (!M) continue;
        int h = 0;
        TTAD_HIP(hipMemcpy(&h, M->info.p, 4, hipMemcpyDeviceToHost));
        if (h) return set_error(TTAD_ERR_NOT_SPD, "graph: non-positive pivot at column %d", h);
    }
    return TTAD_OK;
}

static std::string task_label(const DevTask& t) {
    char b[96];
    if (t.kind == K_POTRF) snprintf(b, sizeof(b), "POTRF_%d", t.j);
    else if (t.kind == K_TRSM) snprintf(b, sizeof(b), "TRSM_%d_%d", t.i, t.j);
    else if (t.kind == K_NOP) snprintf(b, sizeof(b), "SPIN_%d_%d", t.i, t.j);
    else snprintf(b, sizeof(b), "GEMM_%d_%d_%d", t.k, t.i, t.j);
    return b;
}

int ttad_graph_write_trace(ttad_graph* g, const char* path, const char* prefix) {
    if (!g || !path) return set_error(TTAD_ERR_INVALID, "graph_write_trace: bad args");
    if (!g->traced || g->running) return set_error(TTAD_ERR_INVALID, "graph_write_trace: no finished traced run");
    constexpr int W = TRACE_WORDS;
    std::vector<uint64_t> rec((size_t)g->total_items * W);
    int cnt = 0;
    TTAD_HIP(hipMemcpy(&cnt, g->d_trace_cursor.p, 4, hipMemcpyDeviceToHost));
    cnt = std::min(cnt, g->total_items);
    TTAD_HIP(hipMemcpy(rec.data(), g->d_trace.p, (size_t)cnt * W * 8, hipMemcpyDeviceToHost));
    int rate_khz = 100000;
    (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, g->ctx->device);
    const double tick = 1.0 / ((double)rate_khz * 1000.0);
    // Device clock -> seconds since the epoch, as the reference's Logger writes times (util.cpp:18-20, 117-127):
    // read the device clock between two host clock readings and take the midpoint.
    DevBuf clk;
    TTAD_HIP(clk.alloc(8));
    TTAD_HIP(hipStreamSynchronize(g->ctx->stream));
    const double h0 = std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch()).count();
    launch_read_clock(clk.as<uint64_t>(), g->ctx->stream);
    TTAD_HIP(hipStreamSynchronize(g->ctx->stream));
    const double h1 = std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch()).count();
    uint64_t dclk = 0;
    TTAD_HIP(hipMemcpy(&dclk, clk.p, 8, hipMemcpyDeviceToHost));
    const double origin = 0.5 * (h0 + h1);  // epoch seconds at device clock dclk
    auto epoch = [&](uint64_t c) { return origin + ((double)(int64_t)(c - dclk)) * tick; };
    // One ">run>" row per task, as a Threadpool worker logs it (threadpool_shared.cpp:36-57): from the start
    // of its first work item to the end of its last, on the workgroup that ran its first item; one ">deps>" row
    // for the release of its successors (the fulfill step), on the workgroup that ran its last item.
    struct Span {
        uint64_t start = ~0ull, end = 0, deps_start = 0, deps_end = 0;
        uint64_t who_run = 0, who_deps = 0;
        int items = 0;
    };
    std::vector<Span> span(g->n_tasks);
    for (int r = 0; r < cnt; r++) {
        const uint64_t* x = rec.data() + (size_t)r * W;
        const size_t t = (size_t)(x[0] >> 20);
        if (t >= span.size()) continue;
        Span& s = span[t];
        if (x[2] < s.start) {
            s.start = x[2];
            s.who_run = x[1];
        }
        if (x[3] > s.end) s.end = x[3];
        if (x[4]) {
            s.deps_start = x[3];
            s.deps_end = x[4];
            s.who_deps = x[1];
        }
        s.items++;
    }
    std::ofstream f(path);
    if (!f) return set_error(TTAD_ERR_INVALID, "graph_write_trace: cannot open %s", path);
    f.precision(17);
    const std::string pre = prefix ? prefix : "Wk_";
    for (int t = 0; t < g->n_tasks; t++) {
        const Span& s = span[t];
        if (s.items == 0) continue;
        const std::string name = task_label(g->tasks[t]);
        f << pre << s.who_run << ">run>" << name << "," << epoch(s.start) << "," << epoch(s.end) << "\n";
        if (s.deps_end)
            f << pre << s.who_deps << ">deps>" << name << "," << epoch(s.deps_start) << "," << epoch(s.deps_end)
              << "\n";
    }
    return f ? TTAD_OK : set_error(TTAD_ERR_INVALID, "graph_write_trace: write failed");
}

int ttad_graph_write_deps(ttad_graph* g, const char* path) {
    if (!g || !path) return set_error(TTAD_ERR_INVALID, "graph_write_deps: bad args");
    std::ofstream f(path);
    if (!f) return set_error(TTAD_ERR_INVALID, "graph_write_deps: cannot open %s", path);
    for (int t = 0; t < g->n_tasks; t++) {
        const std::string src = task_label(g->tasks[t]);
        for (int e = g->out_ptr[t]; e < g->out_ptr[t + 1]; e++) f << src << "," << task_label(g->tasks[g->out_idx[e]]) << "\n";
    }
    return f ? TTAD_OK : set_error(TTAD_ERR_INVALID, "graph_write_deps: write failed");
}

int ttad_graph_destroy(ttad_graph* g) {
    if (!g) return TTAD_OK;
    if (g->running) (void)hipEventSynchronize(g->ev_stop);
    if (g->ev_start) (void)hipEventDestroy(g->ev_start);
    if (g->ev_stop) (void)hipEventDestroy(g->ev_stop);
    delete g;
    return TTAD_OK;
}

}  // extern "C"
