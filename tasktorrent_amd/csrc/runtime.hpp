// Host-side runtime objects behind the C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/tasktorrent_amd.h"
#include "executor.hpp"
#include "kernels.hpp"

namespace ttad {

int set_error(int code, const char* fmt, ...);
hipError_t factor_check_init();
hipError_t solve_init();
// ||L L^T - A||_F / ||A||_F (factor_check.hip); synchronizes s.
int factor_residual(ttad_context* ctx, ttad_matrix* A, ttad_matrix* L, double* rel, hipStream_t s);

#define TTAD_HIP(call)                                                                                   \
    do {                                                                                                 \
        hipError_t _e = (call);                                                                          \
        if (_e != hipSuccess) return ::ttad::set_error(TTAD_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(_e)); \
    } while (0)

// A device buffer owned by a handle.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t b) {
        release();
        if (b == 0) b = 16;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        else p = nullptr;
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct StreamPlan;  // cached host-sequenced schedule

}  // namespace ttad

struct ttad_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    long long hbm_bytes = 0;
    std::string name;
    int workers_per_cu = 1;            // of the FP64 executor
    int vset_workers_per_cu[8] = {1, 1, 1, 1, 1, 1, 1, 1};  // per executor instantiation
    ttad::DevBuf scratch_dinv;  // inverse diagonal blocks for the raw tile entry points
    ttad::DevBuf scratch_info;  // one int
    double* noalloc_dinv = nullptr;  // this device's address of the static scratch of ttad_dtrsm_noalloc
    // host-sequenced schedules, keyed by (matrix id, prec)
    std::map<std::pair<long long, int>, std::shared_ptr<ttad::StreamPlan>> stream_plans;
    // persistent graphs for ttad_potrf_tiled, keyed by (matrix id, prio kind, prec)
    std::map<std::tuple<long long, int, int>, ttad_graph*> chol_graphs;
    long long next_matrix_id = 1;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;  // bracket the last ttad_potrf_tiled on its stream
    bool ev_valid = false;
    double kernel_s = -1.0;  // event time of the last persistent factorization kernel (< 0: stream schedule)
    // solve-phase workspace (grown on demand)
    ttad::DevBuf sw_Y, sw_D, sw_P, sw_R, sw_flags, sw_ticket, sw_err, sw_norms, sw_part;
    // options (ttad_set_option)
    bool report_factor_residual = false;
    bool residual_rows = false;  // FP64 residuals on the row-block kernels (A read twice)
    // working copies of FP64 matrices for ttad_posv_mixed, keyed by (matrix id, storage)
    std::map<std::pair<long long, int>, ttad_matrix*> work_mats;
};

struct ttad_matrix {
    ttad_context* ctx = nullptr;
    long long id = 0;
    int nt = 0, n = 0, storage = TTAD_F64;
    std::vector<int> bsz, boff;
    std::vector<size_t> tile_off;  // element offset of tile (i,j) (i >= j), index i + j*nt
    ttad::DevBuf tiles, dinv, tab_bsz, tab_boff, tab_tiles, tab_dinv, info, desc, pflags, locks;
    // bf16 hi/lo planes of the off-diagonal tiles (FP32 storage; allocated by the first bf16x3 factorization)
    ttad::DevBuf planes, tab_planes;
    size_t elem = 8;
    long long bytes = 0;
    ttad::DevMat host_desc{};
    // 64-row block table of the solve phase (built on first use)
    int sweep_P = 0;
    ttad::DevBuf tab_blk_tile, tab_blk_off, tab_tfirst;
    // operands of the lookahead sweeps of an FP32 factor (inverses of the diagonal tiles, their transposes,
    // the lookahead products), formed once per solve call
    ttad::DevBuf tinv, tab_tinv, tab_pairs;
    int la_items = 0, la_prods = 0;
    bool tinv_fresh = false;
    double anorm = 0.0;  // ||A||_inf, valid until the next ttad_matrix_set/copy into this matrix
    bool anorm_valid = false;
};

struct ttad_graph {
    ttad_context* ctx = nullptr;
    int n_tasks = 0, n_buckets = 0, total_items = 0;  // n_buckets counts work queues
    std::vector<ttad::DevTask> tasks;
    std::vector<int32_t> indegree, seeds, out_ptr, out_idx, q_base, bucket_items;
    std::vector<int32_t> q_home;  // per queue: -1 shared, else the XCD whose workers it feeds first
    std::vector<ttad_matrix*> mats;
    bool with_planes = false;  // hand the matrices' bf16 planes to the items (a full Cholesky PTG only)
    std::vector<int32_t> spin_us;  // per task: spin time of a TTAD_SPIN task (microseconds)
    bool seeds_dirty = true, mats_dirty = true, running = false;
    ttad::DevBuf d_tasks, d_out_ptr, d_out_idx, d_deps0, d_tail0, d_init, d_init_slot, d_deps, d_items_left, d_gops, d_q_head, d_q_tail, d_q_count, d_xcd_run,
        d_q_base, d_q_home, d_q_items, d_done, d_abort, d_mats, d_trace, d_trace_cursor, d_cu_slots;
    int n_init = 0;
    int workers = 0;
    int vset = 4;  // executor instantiation (index into the item-variant sets) covering the graph
    bool traced = false;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipStream_t stream = nullptr;
    std::vector<std::string> names;  // task names for traces (optional)
};
