// Device-side task execution: item dispatch, the host-sequenced item kernel and
// the persistent DAG kernel that replaces the Threadpool worker loop
// (tasktorrent/src/threadpool_shared.cpp:104-195) and the per-thread dependency
// maps of Taskflow::fulfill_promise (tasktorrent/src/taskflow.hpp:241-296).
#pragma once
#include "tile_ops.hpp"

namespace ttad {

// Device image of a task graph (built on the host from the Taskflow callbacks).
struct DevGraph {
    int32_t n_tasks;
    int32_t n_buckets;
    int32_t total_items;
    int32_t n_init;             // number of initially ready entries (prefilled queue slots)
    const DevTask* tasks;
    const int32_t* out_ptr;     // CSR successors
    const int32_t* out_idx;
    const int32_t* deps0;       // initial dependency counters (indegree - host seeds)
    const int32_t* tail0;       // initial queue tails per bucket
    const int64_t* init_items;  // [n_init] (bucket << 40) | entry for prefilled slots
    int32_t* deps;
    int32_t* items_left;
    int32_t* q_head;            // [n_buckets]
    int32_t* q_tail;            // [n_buckets]
    const int32_t* q_base;      // [n_buckets+1]
    int64_t* q_items;           // entries (task << 20) | item; -1 = not yet written
    int32_t* done_items;
    int32_t* abort_flag;        // 1 = watchdog, 2 = queue overflow
    const DevMat* mats;
    uint64_t* trace;            // optional: 4 x u64 per item {task<<20|item, wg, t_start, t_end}
    int32_t* trace_cursor;
    int64_t spin_limit;         // watchdog in wall_clock64 ticks
};

__device__ __forceinline__ int ld_relaxed(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t ld_acquire64(const int64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// Execute work item `it` of task t. All 256 threads participate.
__device__ __forceinline__ void exec_item(const DevTask& t, int it, const DevMat* mats, char* lds) {
    const DevMat& M = mats[t.mat];
    const int nt = M.nt;
    auto tile = [&](int i, int j) { return M.tiles[i + (long)j * nt]; };
    switch (t.kind) {
        case K_POTRF: {
            const int j = t.j, n = M.bsz[j];
            if (M.storage == P_F64) item_potrf<double>((double*)tile(j, j), n, n, M.dinv[j], M.info, M.boff[j], lds);
            else item_potrf<float>((float*)tile(j, j), n, n, M.dinv[j], M.info, M.boff[j], lds);
            break;
        }
        case K_TRSM: {
            const int i = t.i, j = t.j, m = M.bsz[i], n = M.bsz[j];
            if (M.storage == P_F64)
                item_trsm<double, double>((const double*)tile(j, j), n, M.dinv[j], (double*)tile(i, j), m, m, n,
                                          it * TRSM_BM, lds);
            else
                item_trsm<float, float>((const float*)tile(j, j), n, M.dinv[j], (float*)tile(i, j), m, m, n,
                                        it * TRSM_BM, lds);
            break;
        }
        case K_SYRK:
        case K_GEMM: {
            const int k = t.k, i = t.i, j = t.j;
            const int m = M.bsz[i], n = M.bsz[j], K = M.bsz[k];
            const bool syrk = (t.kind == K_SYRK);
            int rb, cb;
            decode_gemm_item(it, (m + GEMM_BM - 1) / GEMM_BM, syrk, rb, cb);
            if (t.prec == P_F64)
                item_gemm<double, P_F64>((const double*)tile(i, k), m, (const double*)tile(j, k), n,
                                         (double*)tile(i, j), m, m, n, K, rb * GEMM_BM, cb * GEMM_BM, syrk, -1.0,
                                         1.0, lds);
            else if (t.prec == P_F32)
                item_gemm<float, P_F32>((const float*)tile(i, k), m, (const float*)tile(j, k), n,
                                        (float*)tile(i, j), m, m, n, K, rb * GEMM_BM, cb * GEMM_BM, syrk, -1.0, 1.0,
                                        lds);
            else if (t.prec == P_BF16)
                item_gemm<float, P_BF16>((const float*)tile(i, k), m, (const float*)tile(j, k), n,
                                         (float*)tile(i, j), m, m, n, K, rb * GEMM_BM, cb * GEMM_BM, syrk, -1.0,
                                         1.0, lds);
            else
                item_gemm<float, P_BF16X3>((const float*)tile(i, k), m, (const float*)tile(j, k), n,
                                           (float*)tile(i, j), m, m, n, K, rb * GEMM_BM, cb * GEMM_BM, syrk, -1.0,
                                           1.0, lds);
            break;
        }
        default:
            break;
    }
}

}  // namespace ttad
