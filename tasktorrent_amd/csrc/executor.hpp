// Device-side task execution: item dispatch, the host-sequenced item kernel and
// the persistent DAG kernel that replaces the Threadpool worker loop
// (tasktorrent/src/threadpool_shared.cpp:104-195) and the per-thread dependency
// maps of Taskflow::fulfill_promise (tasktorrent/src/taskflow.hpp:241-296).
#pragma once
#include "tile_ops.hpp"

namespace ttad {

// The operands of a GEMM/SYRK task, C(i,j) -= A(i,k) B(j,k)^T, resolved from the matrix's tile tables
// before a run, so that an item finds them in one load with its task record instead of through the
// matrix descriptor and its pointer tables (two more dependent memory round trips).
struct alignas(16) GemmOperands {
    const void* A;
    const void* B;
    void* C;
    const void* PA;  // bf16 planes of A and B (null: none)
    const void* PB;
    int32_t m, n, K, pad;
};

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
    const int64_t* init_items;  // [n_init] queue entries (task << 20) | item of the prefilled slots
    const int32_t* init_slots;  // [n_init] their slots in q_items
    int32_t* deps;
    int32_t* items_left;
    int32_t* q_head;            // [n_buckets]
    int32_t* q_tail;            // [n_buckets]
    int32_t* q_count;           // [n_buckets] entries pushed and not yet claimed (dips below zero while
                                // a failed claim gives its credit back)
    const int32_t* q_base;      // [n_buckets+1]
    const int32_t* q_home;      // [n_buckets] -1 = any worker, else the XCD served first
    int64_t* q_items;           // entries (task << 20) | item; -1 = not yet written
    int32_t* done_items;
    int32_t* abort_flag;        // 1 = watchdog, 2 = queue overflow
    const DevMat* mats;
    GemmOperands* gops;         // [n_tasks] resolved operands of the GEMM/SYRK tasks (filled before each run)
    uint64_t* trace;            // optional: TRACE_WORDS x u64 per item {task<<20|item, wg, t_start, t_end,
                                // t_end of the successor release (the task's last item; else 0), 0}
    int32_t* trace_cursor;
    int32_t* cu_slots;          // [CU_KEYS] workers started per CU (slot 0 = primary)
    int32_t* idle_primary;      // [NXCD * 64] primary workers currently looking for work, per XCD (entry 64 * x)
    int32_t* cu_crit;           // [CU_KEYS] critical items running per CU (its other worker waits)
    int32_t* xcd_run;           // [n_tasks * NXCD] items of a task running on each XCD (release elision)
    int64_t spin_limit;         // watchdog in wall_clock64 ticks
};

constexpr int TRACE_WORDS = 6;  // u64 per trace record

// XCDs of the device (MI355X: 8 XCDs of 32 CUs, each with its own L2).
constexpr int NXCD = 8;
// Work queues of a graph: 64 priority buckets, each with a shared queue and one queue homed on each
// XCD. Bucket b's shared queue is queue b * QUEUE_SHARED_STRIDE (every worker claims from the shared
// queues: their counters get a cache line each), its queue homed on XCD x is
// QUEUE_HOMED + x * QUEUE_BUCKETS + b (only that XCD's workers and stealers claim there: the 64 counters
// of one XCD share 256 bytes and no line with another XCD's). The ids in between are empty queues.
// A scan reads one bucket per lane of a wave.
constexpr int QUEUE_BUCKETS = 64;
constexpr int QUEUE_SHARED_STRIDE = 32;  // 128-byte lines of 4-byte counters
constexpr int QUEUE_HOMED = QUEUE_BUCKETS * QUEUE_SHARED_STRIDE;
constexpr int MAX_QUEUES = QUEUE_HOMED + QUEUE_BUCKETS * NXCD;
__host__ __device__ constexpr int shared_queue(int b) { return b * QUEUE_SHARED_STRIDE; }
__host__ __device__ constexpr int homed_queue(int x, int b) { return QUEUE_HOMED + x * QUEUE_BUCKETS + b; }
constexpr int CLAIM_TRIES = 4;  // queues a claimer tries from one scan
// Hardware id of the XCD (XCC) this wave runs on.
__device__ __forceinline__ int xcc_id() {
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return (int)(v & 0xF) % NXCD;
}

// Identifier of the CU this wave runs on: XCC, shader engine, shader array and CU fields of
// the hardware id registers (fewer than CU_KEYS values).
constexpr int CU_KEYS = 2048;
__device__ __forceinline__ int cu_key() {
    unsigned hw;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw));
    const int cu = (hw >> 8) & 15, sh = (hw >> 12) & 1, se = (hw >> 13) & 7;
    return ((xcc_id() * 8 + se) * 2 + sh) * 16 + cu;
}

__device__ __forceinline__ int ld_relaxed(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t ld_acquire64(const int64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t ld_relaxed64(const int64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Item variants a kernel is compiled with (bit mask). A kernel holding every
// variant allocates registers for the largest one; the executors are therefore
// instantiated per precision and the host picks the smallest set that covers a graph.
enum ItemVariant : int {
    V_F64S = 1,      // FP64 tiles: POTRF, TRSM, GEMM/SYRK on FP64 MFMA
    V_F32S = 2,      // FP32 tiles: POTRF (FP64 arithmetic), TRSM on FP32 MFMA
    V_G32 = 4,       // GEMM/SYRK FP32
    V_GBF16 = 8,     // GEMM/SYRK BF16
    V_GBF16X3 = 16,  // GEMM/SYRK BF16x3
    V_ALL = 31,
    V_ACC = 32       // accumulate-mode GEMM/SYRK (only in the last set, so the others keep their registers)
};
// The compiled sets, smallest first.
constexpr int N_VSETS = 6;
constexpr int VSETS[N_VSETS] = {V_F64S, V_F32S | V_G32, V_F32S | V_GBF16, V_F32S | V_GBF16X3, V_ALL, V_ALL | V_ACC};

__host__ __device__ inline int item_variant(int kind, int prec, int storage) {
    if (kind == K_GEMM || kind == K_SYRK)
        return prec == P_F64 ? V_F64S : prec == P_F32 ? V_G32 : prec == P_BF16 ? V_GBF16 : V_GBF16X3;
    return storage == P_F64 ? V_F64S : V_F32S;
}
// LDS of an executor instantiation: the largest item of its variants.
constexpr int vset_lds(int vm) {
    int b = 0;
    if (vm & V_F64S) b = cmax(b, cmax(cmax(LDS_GEMM_F64, cmax(LDS_POTRF, LDS_POTRF_ROWS)), LDS_TRSM));
    if (vm & V_F32S) b = cmax(b, cmax(cmax(LDS_POTRF, LDS_POTRF_ROWS), LDS_TRSM_F32));
    if (vm & V_G32) b = cmax(b, LDS_GEMM_F32);
    if (vm & V_GBF16) b = cmax(b, LDS_GEMM_BF16_1);
    if (vm & V_GBF16X3) b = cmax(b, LDS_GEMM_BF16);
    return b;
}
// Workgroups per CU an instantiation is compiled for: two when two of them fit the CU's
// 160 KB of LDS (registers are then capped at 256 per lane, VGPRs and AGPRs together).
constexpr int vset_occupancy(int vm) { return 2 * (vset_lds(vm) + 1024) <= 160 * 1024 ? 2 : 1; }
inline int vset_index(int needed) {
    for (int q = 0; q < N_VSETS; q++)
        if ((needed & ~VSETS[q]) == 0) return q;
    return N_VSETS - 1;
}

// POTRF and TRSM items are called out of line from the executors: a kernel holding them
// inline is register-allocated for their largest live set and spills in the GEMM loop,
// which runs far more often. The call saves the callee's registers once per item.
// One instance per executor instantiation (VM): a callee shared by kernels compiled for
// different occupancies would be register-allocated for the lowest of them.
// POTRF and TRSM carry the critical path: their waves run at raised issue priority, so
// that a GEMM item sharing the CU (two workgroups per CU) does not halve their MFMA share.
template <typename TS, int VM>
__device__ __attribute__((noinline)) void potrf_call(TS* A, long lda, int n, double* dinv, int* info, int col_off,
                                                     char* lds) {
    __builtin_amdgcn_s_setprio(3);
    item_potrf<TS>(A, lda, n, dinv, info, col_off, lds);
    __builtin_amdgcn_s_setprio(0);
}
template <typename TS, int VM>
__device__ __attribute__((noinline)) void potrf_rows_call(TS* A, long lda, int n, double* dinv, int* info, int col_off,
                                                          int rb, int32_t* prog, int32_t* err, int64_t spin_limit,
                                                          char* lds) {
    __builtin_amdgcn_s_setprio(3);
    item_potrf_rows<TS>(A, lda, n, dinv, info, col_off, rb, prog, err, spin_limit, lds);
    __builtin_amdgcn_s_setprio(0);
}
// Accumulate-mode GEMM items are called out of line too, so the default GEMM loop keeps its registers.
template <typename TS, int PREC, int VM>
__device__ __attribute__((noinline)) void gemm_acc_call(const TS* A, long lda, const TS* B, long ldb, TS* C, long ldc,
                                                        int m, int n, int K, int r0, int c0, bool syrk, int32_t* lock,
                                                        int32_t* err, int64_t spin_limit, char* lds) {
    item_gemm_acc<TS, PREC>(A, lda, B, ldb, C, ldc, m, n, K, r0, c0, syrk, -1.0, lock, err, spin_limit, lds);
}
template <typename TS, typename TC, int VM>
__device__ __attribute__((noinline)) void trsm_call(const TS* L, long ldl, const double* dinv, TS* X, long ldx, int m,
                                                    int n, int r0, char* lds, void* planes) {
    __builtin_amdgcn_s_setprio(3);
    item_trsm<TS, TC>(L, ldl, dinv, X, ldx, m, n, r0, lds, planes);
    __builtin_amdgcn_s_setprio(0);
}

// Execute work item `it` of task t. All 256 threads participate. Returns false
// (doing nothing) for an item variant outside VM. A POTRF task of more than one item
// runs as row-block items that wait on each other (item_potrf_rows): the caller must
// have every lower-numbered item of the task running or done; their waits give up
// after spin_limit ticks or once *err is nonzero (and set *err = 1 when they time out).
template <int VM>
__device__ __forceinline__ bool exec_item(const DevTask& t, int it, const DevMat* mats, char* lds, int32_t* err,
                                          int64_t spin_limit, const GemmOperands* go = nullptr) {
    if (t.kind == K_NOP) {  // spin task (scheduler-overhead benchmark): k device clock ticks, no tile access
        if (threadIdx.x == 0) {
            const uint64_t t0 = wall_clock64();
            while (wall_clock64() - t0 < (uint64_t)t.k) __builtin_amdgcn_s_sleep(1);
        }
        __syncthreads();
        return true;
    }
    const DevMat& M = mats[t.mat];
    const int nt = M.nt;
    auto tile = [&](int i, int j) { return M.tiles[i + (long)j * nt]; };
    auto planes = [&](int i, int j) { return M.planes ? M.planes[i + (long)j * nt] : nullptr; };
    const bool f64s = M.storage == P_F64;
    switch (t.kind) {
        case K_POTRF: {
            const int j = t.j, n = M.bsz[j];
            if (t.n_items > 1) {
                int32_t* prog = M.pflags + (long)j * POTRF_FLAGS;
                if constexpr ((VM & V_F64S) != 0)
                    if (f64s) {
                        potrf_rows_call<double, VM>((double*)tile(j, j), n, n, M.dinv[j], M.info, M.boff[j], it, prog,
                                                    err, spin_limit, lds);
                        return true;
                    }
                if constexpr ((VM & V_F32S) != 0)
                    if (!f64s) {
                        potrf_rows_call<float, VM>((float*)tile(j, j), n, n, M.dinv[j], M.info, M.boff[j], it, prog,
                                                   err, spin_limit, lds);
                        return true;
                    }
                return false;
            }
            if constexpr ((VM & V_F64S) != 0)
                if (f64s) {
                    potrf_call<double, VM>((double*)tile(j, j), n, n, M.dinv[j], M.info, M.boff[j], lds);
                    return true;
                }
            if constexpr ((VM & V_F32S) != 0)
                if (!f64s) {
                    potrf_call<float, VM>((float*)tile(j, j), n, n, M.dinv[j], M.info, M.boff[j], lds);
                    return true;
                }
            return false;
        }
        case K_TRSM: {
            const int i = t.i, j = t.j, m = M.bsz[i], n = M.bsz[j];
            if constexpr ((VM & V_F64S) != 0)
                if (f64s) {
                    trsm_call<double, double, VM>((const double*)tile(j, j), n, M.dinv[j], (double*)tile(i, j), m, m, n,
                                              it * trsm_rows(8), lds, nullptr);
                    return true;
                }
            if constexpr ((VM & V_F32S) != 0)
                if (!f64s) {
                    trsm_call<float, float, VM>((const float*)tile(j, j), n, M.dinv[j], (float*)tile(i, j), m, m, n,
                                            it * trsm_rows(4), lds, planes(i, j));
                    return true;
                }
            return false;
        }
        case K_SYRK:
        case K_GEMM: {
            const int k = t.k, i = t.i, j = t.j;
            const bool syrk = (t.kind == K_SYRK);
            const void *opA, *opB, *opPA, *opPB;
            void* opC;
            int m, n, K;
            // resolved before the run (persistent executor); the FP64 instantiation keeps the table lookups,
            // measured faster there (the FP64 items are long, the lookup latency is a small part of them)
            constexpr bool use_go = (VM & (V_G32 | V_GBF16 | V_GBF16X3)) != 0;
            if (use_go && go) {
                opA = wave_uniform(go->A);
                opB = wave_uniform(go->B);
                opC = wave_uniform(go->C);
                opPA = wave_uniform(go->PA);
                opPB = wave_uniform(go->PB);
                m = wave_uniform(go->m);
                n = wave_uniform(go->n);
                K = wave_uniform(go->K);
            } else {
                opA = tile(i, k);
                opB = tile(j, k);
                opC = tile(i, j);
                opPA = planes(i, k);
                opPB = planes(j, k);
                m = M.bsz[i];
                n = M.bsz[j];
                K = M.bsz[k];
            }
            int rb, cb;
            decode_gemm_item(it, (m + GEMM_BM - 1) / GEMM_BM, syrk, rb, cb);
            const int r0 = rb * GEMM_BM, c0 = cb * GEMM_BM;
            // accumulate mode: one lock word per 128x128 sub-tile (16 per tile, shared beyond 4x4 sub-tiles)
            int32_t* lock = nullptr;
            if constexpr ((VM & V_ACC) != 0) {
                if (t.accum) lock = M.locks + ((long)(i + (long)j * nt) * 16 + (it & 15));
            } else if (t.accum) {
                return false;
            }
#define TTAD_GEMM_ITEM(TS_, PREC_, CPRE_)                                                                         \
    if constexpr ((VM & V_ACC) != 0) {                                                                             \
        if (lock) {                                                                                                \
            gemm_acc_call<TS_, PREC_, VM>((const TS_*)opA, m, (const TS_*)opB, n, (TS_*)opC, m, m, n, K, r0, c0,  \
                                          syrk, lock, err, spin_limit, lds);                                       \
            return true;                                                                                           \
        }                                                                                                          \
    }                                                                                                              \
    item_gemm<TS_, PREC_, CPRE_>((const TS_*)opA, m, (const TS_*)opB, n, (TS_*)opC, m, m, n, K, r0, c0, syrk,      \
                                 -1.0, 1.0, lds, opPA, opPB);                                                      \
    return true;
            // Products read C after the last K-chunk: measured faster than reading it with the first chunk
            // for FP64 (factorization at n = 32768: 246.7 -> 229.4 ms) and for bf16x3 items (11 %).
            if constexpr ((VM & V_F64S) != 0)
                if (t.prec == P_F64) { TTAD_GEMM_ITEM(double, P_F64, false) }
            if constexpr ((VM & V_G32) != 0)
                if (t.prec == P_F32) { TTAD_GEMM_ITEM(float, P_F32, false) }
            if constexpr ((VM & V_GBF16) != 0)
                if (t.prec == P_BF16) { TTAD_GEMM_ITEM(float, P_BF16, false) }
            if constexpr ((VM & V_GBF16X3) != 0)
                if (t.prec == P_BF16X3) { TTAD_GEMM_ITEM(float, P_BF16X3, false) }
#undef TTAD_GEMM_ITEM
            return false;
        }
        default:
            return false;
    }
}

}  // namespace ttad
