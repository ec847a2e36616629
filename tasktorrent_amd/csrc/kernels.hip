// HIP kernels of the tiled Cholesky library (gfx950).
#include <algorithm>

#include "kernels.hpp"
#include "executor.hpp"

namespace ttad {

// ---------------------------------------------------------------------------
// Host-sequenced executor: one block per work item of a batch of tasks.
// prefix[t] = first global item of task t (exclusive scan of n_items).
// ---------------------------------------------------------------------------
template <int VM>
__global__ __launch_bounds__(NTHREADS, vset_occupancy(VM)) void k_items(const DevMat* mats, const DevTask* tasks, const int32_t* prefix,
                                                    int n_tasks, int32_t* err) {
    extern __shared__ __align__(16) char lds[];
    const int g = blockIdx.x;
    int lo = 0, hi = n_tasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const DevTask t = tasks[lo];
    // row-block POTRF items wait only on lower-numbered blocks of the same launch, which
    // are dispatched first
    if (!exec_item<VM>(t, g - prefix[lo], mats, lds, err, (int64_t)500000000) && threadIdx.x == 0) atomicCAS(err, 0, 3);
}

// ---------------------------------------------------------------------------
// Persistent DAG executor (replaces the Threadpool worker loop). Work queues are ordered
// by priority; a queue is either shared or homed on one XCD (the GEMM/SYRK items of a
// priority bucket are split by output tile row, so that the items of one tile update,
// and the updates that share its A operand, meet in one XCD's L2). Every workgroup loops:
//   wave 0 reads the heads/tails of the queues 64 at a time (one per lane) and claims
//   the head entry of the first non-empty queue that is shared or homed on its own XCD;
//   when there is none it takes the first non-empty queue of another XCD (stealing).
//   A claim takes one credit from the queue's count of unclaimed entries (fetch-add -1;
//   a claimer that finds none gives it back) and then its slot (fetch-add on the head):
//   the credits never exceed the pushed entries, so a slot never passes the tail, and
//   no claim retries a compare-and-swap (CAS claims near the tail, with hundreds of
//   workers on few queues, cost more than half the executor's time in measurements).
//   The workgroup runs that work item, then thread 0 decrements the
//   task's remaining-items counter; the last item of a task decrements its
//   successors' dependency counters and pushes every item of a successor that
//   reaches zero into its queue. A worker exits when every queue is fully claimed.
// Memory ordering: item outputs are released (agent scope) before the counter
// RMWs; queue entries are published with release stores and read with acquire
// loads, so a consumer on another XCD sees the producer's tiles.
// The loop also ends when the abort flag is raised (watchdog: no work found for
// spin_limit ticks; queue overflow; an item variant missing from this instantiation).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void push_task(const DevGraph& g, int s) {
    const DevTask t = g.tasks[s];
    const int b = t.bucket;
    const int pos = __hip_atomic_fetch_add(g.q_tail + b, t.n_items, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int cap = g.q_base[b + 1] - g.q_base[b];
    if (pos + t.n_items > cap) {
        __hip_atomic_store(g.abort_flag, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    // the caller's release fence orders these entries after the outputs they announce
    for (int r = 0; r < t.n_items; r++)
        __hip_atomic_store(g.q_items + g.q_base[b] + pos + r, ((int64_t)s << 20) | r, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    // claim credits (a claimer that gets a slot whose entry is not written yet polls it)
    __hip_atomic_fetch_add(g.q_count + b, t.n_items, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int VM>
__global__ __launch_bounds__(NTHREADS, vset_occupancy(VM)) void k_persistent(DevGraph g) {
    extern __shared__ __align__(16) char lds[];
    __shared__ int64_t s_item;
    const int tid = threadIdx.x;
    const int my_xcd = xcc_id();
    // The first worker started on a CU is its primary. A secondary (the second workgroup on that
    // CU) only takes items from a queue that holds more than there are idle primaries, so that
    // scarce work spreads one item per CU instead of pairing up on some CUs while others idle.
    __shared__ int s_primary, s_key;
    if (tid == 0) {
        s_key = cu_key();
        s_primary = __hip_atomic_fetch_add(g.cu_slots + s_key, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    }
    __syncthreads();
    const bool primary = s_primary != 0;
    const int my_key = s_key;
    // idle primaries are counted per XCD (counters 256 bytes apart): the queues a worker mostly claims
    // from are those of its XCD
    int32_t* idle_primary = g.idle_primary + my_xcd * 64;
    if (primary && tid == 0) __hip_atomic_fetch_add(idle_primary, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t rng = blockIdx.x * 2654435761u + 0x9e3779b9u;
    const int own_q = homed_queue(my_xcd, 0);  // first queue homed on this XCD
    for (;;) {
        if (tid < 64) {
            // wave 0, lane = priority bucket (highest first): the bucket's shared queue and the one homed on
            // this XCD. Each scan reads the two (their counters lie in 256-byte blocks that only the workers
            // of one XCD hammer) in one memory round trip; the other XCDs' queues only when both are empty.
            int64_t got = -1;
            const int64_t t_start = (int64_t)wall_clock64();
            for (;;) {
                int cand[CLAIM_TRIES];  // queues to try, best first
                int nc = 0;
                bool any_open = false;
                // a secondary leaves queues that the idle primaries can drain; nothing is taken while
                // the other worker of this CU runs an item of the critical path
                // the lane index, opaque to the compiler here: addresses derived from it are formed per scan
                // instead of being hoisted out of the worker loop, where they would hold registers through
                // every item
                int lane = tid;
                asm volatile("" : "+v"(lane));
                // (idle primaries of this XCD for its own queues, of all XCDs for shared queues and stealing)
                int idle[NXCD];
#pragma unroll
                for (int x = 0; x < NXCD; x++) idle[x] = primary ? 0 : ld_relaxed(g.idle_primary + x * 64);
                const bool blocked = ld_relaxed(g.cu_crit + my_key) != 0;
                const int cs = ld_relaxed(g.q_count + shared_queue(lane)), co = ld_relaxed(g.q_count + own_q + lane);
                int reserve_all = 0, reserve_own = 0;
#pragma unroll
                for (int x = 0; x < NXCD; x++) {
                    reserve_all += idle[x];
                    if (x == my_xcd) reserve_own = idle[x];
                }
                const unsigned long long ms = __ballot(!blocked && cs > reserve_all);
                const unsigned long long mo = __ballot(!blocked && co > reserve_own);
                // within a bucket the shared queue (POTRF, TRSM, FP64 updates) comes first
                for (unsigned long long m = ms | mo; m != 0 && nc < CLAIM_TRIES; m &= m - 1) {
                    const int b = __ffsll((long long)m) - 1;
                    if ((ms >> b) & 1) cand[nc++] = shared_queue(b);
                    if (nc < CLAIM_TRIES && ((mo >> b) & 1)) cand[nc++] = own_q + b;
                }
                if (nc == 0 && !blocked) {  // nothing on this XCD: steal from the best bucket of another one
                    int cx[NXCD];
#pragma unroll
                    for (int x = 0; x < NXCD; x++) cx[x] = ld_relaxed(g.q_count + homed_queue(x, lane));
                    unsigned long long mx[NXCD], any = 0;
#pragma unroll
                    for (int x = 0; x < NXCD; x++) {
                        mx[x] = x == my_xcd ? 0ull : __ballot(cx[x] > reserve_all);
                        any |= mx[x];
                    }
                    if (any) {
                        const int b = __ffsll((long long)any) - 1;
#pragma unroll
                        for (int d = 1; d < NXCD; d++) {  // from the next XCD on, so that stealers spread
                            const int x = (my_xcd + d) % NXCD;
                            if (nc < CLAIM_TRIES && ((mx[x] >> b) & 1)) cand[nc++] = homed_queue(x, b);
                        }
                    }
                }
                if (nc == 0) {  // nothing claimable: is anything left to claim at all?
                    int qs[NXCD + 1], hs[NXCD + 1];  // this lane's bucket: the shared queue, the homed ones
#pragma unroll
                    for (int x = 0; x <= NXCD; x++) {
                        const int q = x == NXCD ? shared_queue(lane) : homed_queue(x, lane);
                        hs[x] = ld_relaxed(g.q_head + q);
                        qs[x] = g.q_base[q + 1] - g.q_base[q];
                    }
                    bool open = false;
#pragma unroll
                    for (int x = 0; x <= NXCD; x++) open |= hs[x] < qs[x];
                    any_open = __ballot(open) != 0;
                }
                const bool all_claimed = nc == 0 && !any_open;
                if (nc > 0) {
                    int64_t e = -1;
                    if (tid == 0) {
                        int slot = -1, fb = -1;
                        // a credit lost to another worker: the next candidate at once (one atomic) rather than
                        // a new scan
                        for (int c = 0; c < nc; c++) {
                            const int b = cand[c];
                            const int reserve = b >= own_q && b < own_q + QUEUE_BUCKETS ? reserve_own : reserve_all;
                            if (__hip_atomic_fetch_add(g.q_count + b, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >
                                reserve) {
                                slot = __hip_atomic_fetch_add(g.q_head + b, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                fb = b;
                                break;
                            }
                            // no credit left for this worker: give it back
                            __hip_atomic_fetch_add(g.q_count + b, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        if (slot >= 0) {
                            const int64_t* sp = g.q_items + g.q_base[fb] + slot;
                            // claimed before its producer wrote it: poll relaxed (an acquire load would
                            // invalidate the L2 on every poll); the acquire fence follows the broadcast
                            while ((e = ld_relaxed64(sp)) < 0) {
                                if ((int64_t)wall_clock64() - t_start > g.spin_limit) {
                                    __hip_atomic_store(g.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                    e = -2;
                                    break;
                                }
                                __builtin_amdgcn_s_sleep(1);
                            }
                        }
                    }
                    e = __shfl(e, 0);
                    if (e >= 0 || e == -2) {
                        got = e;
                        break;
                    }
                    // lost the race for the last credits: back off a random 0-15 x 64 cycles so the losers do
                    // not retry in lockstep
                    rng = rng * 1664525u + 1013904223u;
                    for (uint32_t q = rng >> 28; q > 0; q--) __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                if (all_claimed || ld_relaxed(g.abort_flag) != 0) {  // every item is taken: this worker is done
                    got = -2;
                    break;
                }
                if ((int64_t)wall_clock64() - t_start > g.spin_limit) {
                    if (tid == 0) __hip_atomic_store(g.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    got = -2;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            if (tid == 0) {
                s_item = got;
                if (primary && got >= 0)
                    __hip_atomic_fetch_add(idle_primary, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
        const int64_t e = s_item;
        if (e < 0) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        // the entry is the same for the whole workgroup (read from LDS): scalar registers, so that the task
        // record, its tiles and every size derived from them are scalar and their branches uniform
        const int task = __builtin_amdgcn_readfirstlane((int)(e >> 20));
        const int item = __builtin_amdgcn_readfirstlane((int)(e & 0xFFFFF));
        const DevTask t = g.tasks[task];
        const GemmOperands go = g.gops[task];  // read with the task record (GEMM/SYRK items use it)
        if (tid == 0) {
            if (t.crit) __hip_atomic_fetch_add(g.cu_crit + my_key, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(g.xcd_run + (long)task * NXCD + my_xcd, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        uint64_t t0 = 0;
        if (g.trace && tid == 0) t0 = wall_clock64();
        const bool ran = exec_item<VM>(t, item, g.mats, lds, g.abort_flag, g.spin_limit, &go);
        // Every wave waits until the L2 has acknowledged its own stores (gfx9: vmcnt counts stores too;
        // encoding vmcnt(0) expcnt(7) lgkmcnt(15)) before the barrier: after it, thread 0's decrement
        // below is issued only once all of this item's outputs sit in this XCD's L2.
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        if (tid == 0) {
            // Release elision: an agent-scope release (buffer_wbl2) writes back every dirty line of this
            // XCD's L2, so the item that leaves no other item of its task running on this XCD releases for
            // all of them; the others skip it. This relies on the hardware, not on the memory model: an
            // item's stores reached the L2 before its relaxed decrement was issued (the wait above), the
            // decrements of one counter are serialized at the L2, and the last one's write-back is issued
            // after its decrement returned, so it covers every earlier item's lines. Each item decrements
            // the task's remaining count after that, so the count reaches zero only after every XCD that
            // ran items of the task has released.
            if (__hip_atomic_fetch_add(g.xcd_run + (long)task * NXCD + my_xcd, -1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT) == 1)
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (!ran) __hip_atomic_store(g.abort_flag, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t t1 = g.trace ? wall_clock64() : 0;
            // Counters are relaxed: the release fence above publishes this item's outputs. The thread
            // that takes a counter to zero acquires what the other decrements released and releases it again
            // before its own next counter update or queue entries (one acq_rel fence per task, and per
            // successor it makes ready).
            const int left =
                __hip_atomic_fetch_add(g.items_left + task, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - 1;
            if (left == 0) {  // last item of the task: release its successors
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                for (int q = g.out_ptr[task]; q < g.out_ptr[task + 1]; q++) {
                    const int s = g.out_idx[q];
                    const int d =
                        __hip_atomic_fetch_add(g.deps + s, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - 1;
                    if (d == 0) {
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                        push_task(g, s);
                    }
                }
                __hip_atomic_fetch_add(g.done_items, t.n_items, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (g.trace) {  // {entry, workgroup, start, end, end of the successor release (last item only), 0}
                const int slot = __hip_atomic_fetch_add(g.trace_cursor, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (slot < g.total_items) {
                    uint64_t* rec = g.trace + TRACE_WORDS * (long)slot;
                    rec[0] = (uint64_t)e;
                    rec[1] = blockIdx.x;
                    rec[2] = t0;
                    rec[3] = t1;
                    rec[4] = left == 0 ? wall_clock64() : 0;
                    rec[5] = 0;
                }
            }
            if (primary) __hip_atomic_fetch_add(idle_primary, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t.crit) __hip_atomic_fetch_add(g.cu_crit + my_key, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Restores the counters and queues of a graph so that it can run again.
__global__ void k_graph_reset_a(DevGraph g, int n_slots) {
    const int stride = gridDim.x * blockDim.x;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < n_slots; x += stride) g.q_items[x] = -1;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < g.n_tasks; x += stride) {
        g.deps[x] = g.deps0[x];
        g.items_left[x] = g.tasks[x].n_items;
        const DevTask t = g.tasks[x];
        if (t.kind == K_GEMM || t.kind == K_SYRK) {
            const DevMat& M = g.mats[t.mat];
            const long nt = M.nt;
            GemmOperands o;
            o.A = M.tiles[t.i + t.k * nt];
            o.B = M.tiles[t.j + t.k * nt];
            o.C = M.tiles[t.i + t.j * nt];
            o.PA = M.planes ? M.planes[t.i + t.k * nt] : nullptr;
            o.PB = M.planes ? M.planes[t.j + t.k * nt] : nullptr;
            o.m = M.bsz[t.i];
            o.n = M.bsz[t.j];
            o.K = M.bsz[t.k];
            o.pad = 0;
            g.gops[x] = o;
        }
    }
    for (long x = blockIdx.x * blockDim.x + threadIdx.x; x < (long)g.n_tasks * NXCD; x += stride) g.xcd_run[x] = 0;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < g.n_buckets; x += stride) {
        g.q_head[x] = 0;
        g.q_tail[x] = g.tail0[x];
        g.q_count[x] = g.tail0[x];
    }
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < CU_KEYS; x += stride) {
        g.cu_slots[x] = 0;
        g.cu_crit[x] = 0;
    }
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < NXCD * 64; x += stride) g.idle_primary[x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *g.done_items = 0;
        *g.abort_flag = 0;
        if (g.trace_cursor) *g.trace_cursor = 0;
    }
}
__global__ void k_graph_reset_b(DevGraph g) {
    const int stride = gridDim.x * blockDim.x;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < g.n_init; x += stride) {
        g.q_items[g.init_slots[x]] = g.init_items[x];
    }
}

// ---------------------------------------------------------------------------
// Raw tile kernels behind the BLAS-shaped C-ABI entry points.
// ---------------------------------------------------------------------------
template <typename TS, int PREC>
__global__ __launch_bounds__(NTHREADS, 2 * (gemm_item_lds<TS, PREC>() + 1024) <= 160 * 1024 ? 2 : 1) void k_raw_gemm(const TS* A, long lda, const TS* B, long ldb, TS* C, long ldc,
                                                       int m, int n, int K, int syrk, double alpha, double beta) {
    extern __shared__ __align__(16) char lds[];
    int rb, cb;
    decode_gemm_item(blockIdx.x, (m + GEMM_BM - 1) / GEMM_BM, syrk != 0, rb, cb);
    item_gemm<TS, PREC>(A, lda, B, ldb, C, ldc, m, n, K, rb * GEMM_BM, cb * GEMM_BM, syrk != 0, alpha, beta, lds);
}

template <typename TS>
__global__ __launch_bounds__(NTHREADS) void k_raw_potrf(TS* A, long lda, int n, double* dinv, int* info) {
    extern __shared__ __align__(16) char lds[];
    item_potrf<TS>(A, lda, n, dinv, info, 0, lds);
}

template <typename TS, typename TC>
__global__ __launch_bounds__(NTHREADS) void k_raw_trsm(const TS* L, long ldl, const double* dinv, TS* X, long ldx,
                                                       int m, int n) {
    extern __shared__ __align__(16) char lds[];
    item_trsm<TS, TC>(L, ldl, dinv, X, ldx, m, n, blockIdx.x * trsm_rows(sizeof(TS)), lds);
}

// Inverses of the 64x64 diagonal blocks of a lower-triangular n x n L (one block per workgroup).
template <typename TS>
__global__ __launch_bounds__(NTHREADS) void k_trtri_diag(const TS* L, long ldl, int n, double* dinv) {
    __shared__ double D[DB * DP];
    __shared__ double Xs[DB * DP];
    const int pb = blockIdx.x, p0 = pb * DB, pw = min(DB, n - p0);
    for (int e = threadIdx.x; e < DB * DB; e += NTHREADS) {
        const int i = e % DB, j = e / DB;
        D[i + j * DP] = (i < pw && j < pw && i >= j) ? (double)L[(p0 + i) + (long)(p0 + j) * ldl] : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    lds_trtri64(D, Xs, pw);
    double* X = dinv + (long)pb * DB * DB;
    for (int e = threadIdx.x; e < DB * DB; e += NTHREADS) X[e] = Xs[(e % DB) + (e / DB) * DP];
}

// ---------------------------------------------------------------------------
// Tile layout conversion: full column-major n x n (FP64) <-> lower tiles.
// One block per (i, j) tile of the lower triangle, j = blockIdx.y, i = blockIdx.x.
// ---------------------------------------------------------------------------
template <typename TS>
__global__ void k_pack(const double* __restrict__ src, long lds_, const DevMat* M, int nt) {
    const int i = blockIdx.x, j = blockIdx.y;
    if (i < j) return;
    const DevMat& m = *M;
    TS* t = (TS*)m.tiles[i + (long)j * nt];
    const int r = m.bsz[i], c = m.bsz[j], ri = m.boff[i], cj = m.boff[j];
    const long cnt = (long)r * c;
    for (long e = threadIdx.x; e < cnt; e += blockDim.x) {
        const int p = (int)(e % r), q = (int)(e / r);
        t[e] = (TS)src[(ri + p) + (long)(cj + q) * lds_];
    }
}
template <typename TS>
__global__ void k_unpack(double* __restrict__ dst, long ldd, const DevMat* M, int nt, int zero_upper) {
    const int i = blockIdx.x, j = blockIdx.y;
    const DevMat& m = *M;
    const int ri = m.boff[i], cj = m.boff[j], r = m.bsz[i], c = m.bsz[j];
    const long cnt = (long)r * c;
    if (i < j) {
        if (zero_upper)
            for (long e = threadIdx.x; e < cnt; e += blockDim.x)
                dst[(ri + (int)(e % r)) + (long)(cj + (int)(e / r)) * ldd] = 0.0;
        return;
    }
    const TS* t = (const TS*)m.tiles[i + (long)j * nt];
    for (long e = threadIdx.x; e < cnt; e += blockDim.x) {
        const int p = (int)(e % r), q = (int)(e / r);
        if (i == j && p < q) {
            if (zero_upper) dst[(ri + p) + (long)(cj + q) * ldd] = 0.0;
            continue;
        }
        dst[(ri + p) + (long)(cj + q) * ldd] = (double)t[e];
    }
}
// Tile-wise copy with conversion between two matrices of the same partition, one workgroup of 256 per lower
// tile (workgroup x -> tile (i, j), x = i (i + 1) / 2 + j). Four consecutive elements move as one vector access
// (two 16-byte loads of FP64, one 16-byte store of FP32), four quads in flight per thread; a tile whose pointers
// or element count do not allow quads is copied element by element.
template <typename TD, typename TSRC>
__global__ __launch_bounds__(256) void k_copy_tiles(const DevMat* S, const DevMat* D, int nt) {
    typedef TSRC s4 __attribute__((ext_vector_type(4)));
    typedef TD d4 __attribute__((ext_vector_type(4)));
    const int x = blockIdx.x;
    int i = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= x) i++;
    while (i * (i + 1) / 2 > x) i--;
    const int j = x - i * (i + 1) / 2;
    const TSRC* s = (const TSRC*)S->tiles[i + (long)j * nt];
    TD* d = (TD*)D->tiles[i + (long)j * nt];
    const long cnt = (long)S->bsz[i] * S->bsz[j];
    const bool quads = (cnt % 4) == 0 && ((uintptr_t)s % sizeof(s4)) == 0 && ((uintptr_t)d % sizeof(d4)) == 0;
    if (!quads) {
        for (long e = threadIdx.x; e < cnt; e += 256) d[e] = (TD)s[e];
        return;
    }
    const TTAD_GLOBAL s4* sp = gmem((const s4*)s);
    TTAD_GLOBAL d4* dp = gmem((d4*)d);
    const long nq = cnt / 4;
    long e = threadIdx.x;
    for (; e + 3 * 256 < nq; e += 4 * 256) {
        const s4 v0 = sp[e], v1 = sp[e + 256], v2 = sp[e + 512], v3 = sp[e + 768];
        dp[e] = __builtin_convertvector(v0, d4);
        dp[e + 256] = __builtin_convertvector(v1, d4);
        dp[e + 512] = __builtin_convertvector(v2, d4);
        dp[e + 768] = __builtin_convertvector(v3, d4);
    }
    for (; e < nq; e += 256) dp[e] = __builtin_convertvector(sp[e], d4);
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side, C++ linkage; called by the C-ABI layer).
// ---------------------------------------------------------------------------
static hipError_t set_lds(const void* f, int bytes) {
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

hipError_t kernels_init() {
    hipError_t e = hipSuccess;
#define TTAD_SETLDS2(fn, bytes)                      \
    do {                                             \
        e = set_lds((const void*)(fn), (bytes));     \
        if (e != hipSuccess) return e;               \
    } while (0)
    TTAD_SETLDS2(k_items<VSETS[0]>, vset_lds(VSETS[0]));
    TTAD_SETLDS2(k_items<VSETS[1]>, vset_lds(VSETS[1]));
    TTAD_SETLDS2(k_items<VSETS[2]>, vset_lds(VSETS[2]));
    TTAD_SETLDS2(k_items<VSETS[3]>, vset_lds(VSETS[3]));
    TTAD_SETLDS2(k_items<VSETS[4]>, vset_lds(VSETS[4]));
    TTAD_SETLDS2(k_persistent<VSETS[0]>, vset_lds(VSETS[0]));
    TTAD_SETLDS2(k_persistent<VSETS[1]>, vset_lds(VSETS[1]));
    TTAD_SETLDS2(k_persistent<VSETS[2]>, vset_lds(VSETS[2]));
    TTAD_SETLDS2(k_persistent<VSETS[3]>, vset_lds(VSETS[3]));
    TTAD_SETLDS2(k_persistent<VSETS[4]>, vset_lds(VSETS[4]));
    TTAD_SETLDS2(k_persistent<VSETS[5]>, vset_lds(VSETS[5]));
    TTAD_SETLDS2((k_raw_gemm<double, P_F64>), (gemm_item_lds<double, P_F64>()));
    TTAD_SETLDS2((k_raw_gemm<float, P_F32>), (gemm_item_lds<float, P_F32>()));
    TTAD_SETLDS2((k_raw_gemm<float, P_BF16>), (gemm_item_lds<float, P_BF16>()));
    TTAD_SETLDS2((k_raw_gemm<float, P_BF16X3>), (gemm_item_lds<float, P_BF16X3>()));
    TTAD_SETLDS2((k_raw_potrf<double>), LDS_POTRF);
    TTAD_SETLDS2((k_raw_potrf<float>), LDS_POTRF);
    TTAD_SETLDS2((k_raw_trsm<double, double>), LDS_TRSM);
    TTAD_SETLDS2((k_raw_trsm<float, float>), LDS_TRSM_F32);
#undef TTAD_SETLDS2
    return e;
}

static const void* persistent_fn(int vset) {
    switch (vset) {
        case 0: return (const void*)k_persistent<VSETS[0]>;
        case 1: return (const void*)k_persistent<VSETS[1]>;
        case 2: return (const void*)k_persistent<VSETS[2]>;
        case 3: return (const void*)k_persistent<VSETS[3]>;
        case 4: return (const void*)k_persistent<VSETS[4]>;
        default: return (const void*)k_persistent<VSETS[5]>;
    }
}

static int vset_lds_of(int vset) {
    switch (vset) {
        case 0: return vset_lds(VSETS[0]);
        case 1: return vset_lds(VSETS[1]);
        case 2: return vset_lds(VSETS[2]);
        case 3: return vset_lds(VSETS[3]);
        case 4: return vset_lds(VSETS[4]);
        default: return vset_lds(VSETS[5]);
    }
}

int persistent_blocks_per_cu(int vset) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, persistent_fn(vset), NTHREADS, vset_lds_of(vset)) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

void launch_items(int vset, const DevMat* mats, const DevTask* tasks, const int32_t* prefix, int n_tasks,
                  int n_items, int32_t* err, hipStream_t s) {
    if (n_items <= 0) return;
    const dim3 grid(n_items), block(NTHREADS);
    // (no accumulate-mode variant here: host-sequenced plans hold no accumulate tasks)
    const int lds = vset_lds_of(std::min(vset, N_VSETS - 2));
    switch (vset) {
        case 0: hipLaunchKernelGGL(k_items<VSETS[0]>, grid, block, lds, s, mats, tasks, prefix, n_tasks, err); break;
        case 1: hipLaunchKernelGGL(k_items<VSETS[1]>, grid, block, lds, s, mats, tasks, prefix, n_tasks, err); break;
        case 2: hipLaunchKernelGGL(k_items<VSETS[2]>, grid, block, lds, s, mats, tasks, prefix, n_tasks, err); break;
        case 3: hipLaunchKernelGGL(k_items<VSETS[3]>, grid, block, lds, s, mats, tasks, prefix, n_tasks, err); break;
        default: hipLaunchKernelGGL(k_items<VSETS[4]>, grid, block, lds, s, mats, tasks, prefix, n_tasks, err); break;
    }
}

__global__ void k_read_clock(uint64_t* out) {
    if (threadIdx.x == 0) *out = wall_clock64();
}
void launch_read_clock(uint64_t* out, hipStream_t s) { hipLaunchKernelGGL(k_read_clock, dim3(1), dim3(64), 0, s, out); }

void launch_graph_reset(const DevGraph& g, int n_slots, hipStream_t s) {
    hipLaunchKernelGGL(k_graph_reset_a, dim3(256), dim3(256), 0, s, g, n_slots);
    hipLaunchKernelGGL(k_graph_reset_b, dim3(64), dim3(256), 0, s, g);
}

void launch_persistent(int vset, const DevGraph& g, int n_workers, hipStream_t s) {
    const dim3 grid(n_workers), block(NTHREADS);
    const int lds = vset_lds_of(vset);
    switch (vset) {
        case 0: hipLaunchKernelGGL(k_persistent<VSETS[0]>, grid, block, lds, s, g); break;
        case 1: hipLaunchKernelGGL(k_persistent<VSETS[1]>, grid, block, lds, s, g); break;
        case 2: hipLaunchKernelGGL(k_persistent<VSETS[2]>, grid, block, lds, s, g); break;
        case 3: hipLaunchKernelGGL(k_persistent<VSETS[3]>, grid, block, lds, s, g); break;
        case 4: hipLaunchKernelGGL(k_persistent<VSETS[4]>, grid, block, lds, s, g); break;
        default: hipLaunchKernelGGL(k_persistent<VSETS[5]>, grid, block, lds, s, g); break;
    }
}

void launch_raw_gemm(int prec, const void* A, long lda, const void* B, long ldb, void* C, long ldc, int m, int n,
                     int K, int syrk, double alpha, double beta, hipStream_t s) {
    const int tr = (m + GEMM_BM - 1) / GEMM_BM, tc = (n + GEMM_BM - 1) / GEMM_BM;
    const int items = syrk ? tr * (tr + 1) / 2 : tr * tc;
    if (items <= 0) return;
    if (prec == P_F64)
        hipLaunchKernelGGL((k_raw_gemm<double, P_F64>), dim3(items), dim3(NTHREADS), (gemm_item_lds<double, P_F64>()), s,
                           (const double*)A, lda, (const double*)B, ldb, (double*)C, ldc, m, n, K, syrk, alpha, beta);
    else if (prec == P_F32)
        hipLaunchKernelGGL((k_raw_gemm<float, P_F32>), dim3(items), dim3(NTHREADS), (gemm_item_lds<float, P_F32>()), s, (const float*)A,
                           lda, (const float*)B, ldb, (float*)C, ldc, m, n, K, syrk, alpha, beta);
    else if (prec == P_BF16)
        hipLaunchKernelGGL((k_raw_gemm<float, P_BF16>), dim3(items), dim3(NTHREADS), (gemm_item_lds<float, P_BF16>()), s,
                           (const float*)A, lda, (const float*)B, ldb, (float*)C, ldc, m, n, K, syrk, alpha, beta);
    else
        hipLaunchKernelGGL((k_raw_gemm<float, P_BF16X3>), dim3(items), dim3(NTHREADS), (gemm_item_lds<float, P_BF16X3>()), s,
                           (const float*)A, lda, (const float*)B, ldb, (float*)C, ldc, m, n, K, syrk, alpha, beta);
}

void launch_raw_potrf(int prec, void* A, long lda, int n, double* dinv, int* info, hipStream_t s) {
    if (prec == P_F64)
        hipLaunchKernelGGL((k_raw_potrf<double>), dim3(1), dim3(NTHREADS), LDS_POTRF, s, (double*)A, lda, n, dinv,
                           info);
    else
        hipLaunchKernelGGL((k_raw_potrf<float>), dim3(1), dim3(NTHREADS), LDS_POTRF, s, (float*)A, lda, n, dinv,
                           info);
}

void launch_trtri_diag(int prec, const void* L, long ldl, int n, double* dinv, hipStream_t s) {
    const int nb = (n + DB - 1) / DB;
    if (prec == P_F64)
        hipLaunchKernelGGL((k_trtri_diag<double>), dim3(nb), dim3(NTHREADS), 0, s, (const double*)L, ldl, n, dinv);
    else
        hipLaunchKernelGGL((k_trtri_diag<float>), dim3(nb), dim3(NTHREADS), 0, s, (const float*)L, ldl, n, dinv);
}

void launch_raw_trsm(int prec, const void* L, long ldl, const double* dinv, void* X, long ldx, int m, int n,
                     hipStream_t s) {
    const int bm = trsm_rows(prec == P_F64 ? 8 : 4);
    const int items = (m + bm - 1) / bm;
    if (items <= 0) return;
    if (prec == P_F64)
        hipLaunchKernelGGL((k_raw_trsm<double, double>), dim3(items), dim3(NTHREADS), LDS_TRSM, s,
                           (const double*)L, ldl, dinv, (double*)X, ldx, m, n);
    else
        hipLaunchKernelGGL((k_raw_trsm<float, float>), dim3(items), dim3(NTHREADS), LDS_TRSM_F32, s, (const float*)L,
                           ldl, dinv, (float*)X, ldx, m, n);
}

void launch_pack(int storage, const double* src, long lda, const DevMat* M, int nt, hipStream_t s) {
    if (storage == P_F64) hipLaunchKernelGGL((k_pack<double>), dim3(nt, nt), dim3(256), 0, s, src, lda, M, nt);
    else hipLaunchKernelGGL((k_pack<float>), dim3(nt, nt), dim3(256), 0, s, src, lda, M, nt);
}

void launch_unpack(int storage, double* dst, long ldd, const DevMat* M, int nt, int zero_upper, hipStream_t s) {
    if (storage == P_F64)
        hipLaunchKernelGGL((k_unpack<double>), dim3(nt, nt), dim3(256), 0, s, dst, ldd, M, nt, zero_upper);
    else hipLaunchKernelGGL((k_unpack<float>), dim3(nt, nt), dim3(256), 0, s, dst, ldd, M, nt, zero_upper);
}

void launch_copy_tiles(int dst_storage, int src_storage, const DevMat* S, const DevMat* D, int nt, hipStream_t s) {
    const unsigned nlow = (unsigned)nt * (nt + 1) / 2;
    if (dst_storage == P_F32 && src_storage == P_F64)
        hipLaunchKernelGGL((k_copy_tiles<float, double>), dim3(nlow), dim3(256), 0, s, S, D, nt);
    else if (dst_storage == P_F64 && src_storage == P_F32)
        hipLaunchKernelGGL((k_copy_tiles<double, float>), dim3(nlow), dim3(256), 0, s, S, D, nt);
    else if (dst_storage == P_F64)
        hipLaunchKernelGGL((k_copy_tiles<double, double>), dim3(nlow), dim3(256), 0, s, S, D, nt);
    else hipLaunchKernelGGL((k_copy_tiles<float, float>), dim3(nlow), dim3(256), 0, s, S, D, nt);
}

}  // namespace ttad
