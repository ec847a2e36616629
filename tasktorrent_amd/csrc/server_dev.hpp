// Device side of the serving loop: a long-running persistent executor that
// admits requests posted by the host while it runs (continuous batching).
//
// Each request is a task graph in a recycled task pool: INIT (zero the solve
// state), CONVERT (cast the FP64 tiles of A into the working factor), ANORM
// (||A||_inf), the Cholesky PTG on the factor (2d_cholesky.cpp:316-551), and a
// refinement chain FWD -> BWD -> RESID -> CHECK. CHECK either completes the
// request or re-arms the chain (an iteration of refinement) inside the DAG.
// Requests of any size share the same workgroups; each request's arithmetic is
// independent of what else runs (fixed update order, no floating-point atomics),
// so its result does not depend on the mix.
#pragma once
#include "executor.hpp"
#include "solve_ops.hpp"

namespace ttad {

enum SrvKind : int32_t { S_INIT = 16, S_CONVERT = 17, S_ANORM = 18, S_FWD = 19, S_BWD = 20, S_RESID = 21, S_CHECK = 22 };
enum SrvStatus : int32_t { SRV_RUNNING = 0, SRV_OK = 1, SRV_STALLED = 2, SRV_NOT_SPD = 3 };

struct SrvTask {        // 48 bytes
    int32_t kind, prec, mat, n_items;
    int32_t i, j, k, req;
    int32_t edge_off, edge_cnt, bucket, pad;
};

// Per-request state in HBM (written by the host through DMA before admission).
struct SrvReq {
    int32_t mat_a, mat_w;       // DevMat slots: FP64 input, working factor
    int32_t n, nrhs;
    int32_t P;                  // 64-row blocks
    int32_t max_iter;
    int32_t t_fwd, t_bwd, t_res, t_chk;
    int32_t done_slot;          // index of the host-visible completion record
    int32_t iter, stall, pad;
    double tol, eta_prev;
    const double* B;            // n x nrhs FP64
    double* X;                  // solution
    double* R;                  // residual
    void* Y;                    // sweep intermediates (factor precision)
    void* D;
    int32_t* flags_f;           // [P]
    int32_t* flags_b;           // [P]
    unsigned long long* norms;  // [3*32]: |r|, |x|, |b| maxima per rhs
    unsigned long long* anorm;  // [1]
    const int32_t* blk_tile;    // 64-row block table of the partition
    const int32_t* blk_off;
    double* x_host;             // pinned host buffer that receives X (n x nrhs) on completion
};

// Admission mail of one request slot (pinned host memory, written by the host, read by
// the workgroup that admits the request): a list of copies host -> HBM, then the seed.
struct SrvCopy {
    unsigned long long dst;     // device address (4-byte aligned)
    unsigned long long src;     // device address of pinned host memory
    unsigned long long bytes;   // exact payload size, a multiple of 4: nothing past it is written
};
struct SrvMail {
    int32_t state;              // 0 free, 1 posted by the host, 2 copied by the device
    int32_t n_copies;
    int32_t seed;               // task to push once the copies are done
    int32_t pad;
    unsigned long long copies;  // device address of SrvCopy[n_copies]
};

// Host-visible completion record (pinned host memory).
struct SrvDone {
    int32_t state;              // SrvStatus, written last with a system-scope release
    int32_t iterations;
    double eta;
    uint64_t t_done;            // device wall clock
    int32_t info;
    int32_t pad;
};

struct SrvRing {                // per bucket ring of work-item entries
    unsigned long long* head;   // claimed
    unsigned long long* tail;   // reserved
    unsigned long long* slots;  // entry = ((lap+1) & 0xFFFFFF) << 40 | task << 16 | item
    long long* count;           // entries pushed and not yet claimed (claim credits; dips below zero while a
                                // failed claim gives its credit back)
};

struct DevServer {
    int32_t n_buckets;
    int32_t ring_log2;          // capacity 2^ring_log2 per bucket
    SrvRing rings[4];
    SrvTask* tasks;
    int32_t* edges;             // successor task ids
    int32_t* deps;
    int32_t* items_left;
    SrvReq* reqs;
    DevMat* mats;
    SrvDone* done;              // host pinned
    const volatile int32_t* db_entries;  // host pinned: request slots posted for admission
    const volatile int32_t* db_count;    // host pinned: number of posted slots
    const volatile int32_t* shutdown;    // host pinned
    SrvMail* mail;              // host pinned, one per request slot
    int32_t db_cap;
    int32_t* db_taken;          // device: doorbell entries consumed
    int32_t* active;            // device: requests in flight
    int32_t* err;               // host pinned: watchdog codes; nonzero also ends every device wait early
    int64_t spin_limit;
    int64_t idle_limit;
};

__device__ __forceinline__ void srv_push(const DevServer& s, int task) {
    const SrvTask t = s.tasks[task];
    const SrvRing& R = s.rings[t.bucket];
    const unsigned long long cap = 1ull << s.ring_log2;
    const unsigned long long pos =
        __hip_atomic_fetch_add(R.tail, (unsigned long long)t.n_items, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int r = 0; r < t.n_items; r++) {
        const unsigned long long idx = pos + r;
        const unsigned long long lap = ((idx >> s.ring_log2) + 1) & 0xFFFFFFull;
        const unsigned long long e = (lap << 40) | ((unsigned long long)task << 16) | (unsigned long long)r;
        __hip_atomic_store(R.slots + (idx & (cap - 1)), e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_fetch_add(R.count, (long long)t.n_items, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A task finished: decrement successors, push those that become ready.
__device__ __forceinline__ void srv_complete(const DevServer& s, int task) {
    const SrvTask t = s.tasks[task];
    for (int q = 0; q < t.edge_cnt; q++) {
        const int succ = s.edges[t.edge_off + q];
        const int d = __hip_atomic_fetch_add(s.deps + succ, -1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) - 1;
        if (d == 0) srv_push(s, succ);
    }
}

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }

// Admit the request posted in mail slot `slot`: the whole workgroup copies its
// descriptors and right-hand sides from pinned host memory into HBM, then thread 0
// releases them and pushes the seed task. No runtime copy is needed while the
// persistent kernel owns the device.
__device__ __forceinline__ void srv_admit(const DevServer& s, int slot) {
    SrvMail* m = s.mail + slot;
    __shared__ int s_nc, s_seed;
    __shared__ unsigned long long s_copies;
    if (threadIdx.x == 0) {
        s_nc = __hip_atomic_load(&m->n_copies, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        s_seed = __hip_atomic_load(&m->seed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        s_copies = __hip_atomic_load(&m->copies, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __syncthreads();
    // every wave reads the host's payloads after thread 0's acquire of the doorbell
    __atomic_thread_fence(__ATOMIC_ACQUIRE);  // system scope
    const SrvCopy* cl = (const SrvCopy*)s_copies;
    for (int c = 0; c < s_nc; c++) {  // 4-byte words: task and edge ranges start at any entry
        const uint32_t* src = (const uint32_t*)cl[c].src;
        uint32_t* dst = (uint32_t*)cl[c].dst;
        const long words = (long)(cl[c].bytes >> 2);
        for (long w = threadIdx.x; w < words; w += NTHREADS) dst[w] = src[w];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // each wave publishes its own copies
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&m->state, 2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_fetch_add(s.active, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        srv_push(s, s_seed);
    }
}

// Serving task bodies. All 256 threads of the workgroup participate.
__device__ __forceinline__ void srv_exec(const DevServer& s, const SrvTask& t, int task, int it, char* lds) {
    SrvReq& q = s.reqs[t.req];
    switch (t.kind) {
        case S_INIT: {  // X = 0, R = B, flags = 0, norms = 0
            const long tot = (long)q.n * q.nrhs;
            for (long e = threadIdx.x; e < tot; e += NTHREADS) {
                q.X[e] = 0.0;
                q.R[e] = q.B[e];
            }
            for (int e = threadIdx.x; e < q.P; e += NTHREADS) {
                q.flags_f[e] = 0;
                q.flags_b[e] = 0;
            }
            for (int e = threadIdx.x; e < 3 * 32; e += NTHREADS) q.norms[e] = 0;
            if (threadIdx.x == 0) {
                *q.anorm = 0;
                q.iter = 0;
                q.stall = 0;
                q.eta_prev = INFINITY;
            }
            break;
        }
        case S_CONVERT: {  // one lower tile of A (FP64) into the working factor
            const DevMat& A = s.mats[q.mat_a];
            const DevMat& W = s.mats[q.mat_w];
            int j = 0, r = it;
            while (r >= A.nt - j) { r -= A.nt - j; j++; }
            const int i = j + r;
            const double* src = (const double*)A.tiles[i + (long)j * A.nt];
            const long cnt = (long)A.bsz[i] * A.bsz[j];
            if (W.storage == P_F64) {
                double* dst = (double*)W.tiles[i + (long)j * W.nt];
                for (long e = threadIdx.x; e < cnt; e += NTHREADS) dst[e] = src[e];
            } else {
                float* dst = (float*)W.tiles[i + (long)j * W.nt];
                for (long e = threadIdx.x; e < cnt; e += NTHREADS) dst[e] = (float)src[e];
            }
            if (it == 0 && threadIdx.x == 0) *W.info = 0;
            break;
        }
        case S_ANORM: {
            const SweepTab tab{q.P, q.blk_tile, q.blk_off};
            item_resid<16, false>(s.mats[q.mat_a], tab, it, q.B, q.n, q.X, q.n, q.R, q.n, q.nrhs, q.anorm, 1, lds);
            break;
        }
        case S_FWD:
        case S_BWD: {
            const SweepTab tab{q.P, q.blk_tile, q.blk_off};
            const DevMat& W = s.mats[q.mat_w];
            const int fwd = t.kind == S_FWD;
            const int p = fwd ? it : q.P - 1 - it;
            int32_t* fl = fwd ? q.flags_f : q.flags_b;
            if (W.storage == P_F64) {
                if (q.nrhs <= 16)
                    item_sweep<double, double, 16, false>(W, tab, fwd, p, q.R, q.n, (double*)q.Y, q.n, (double*)q.D, q.n, q.X,
                                                   q.n, q.nrhs, fl, s.err, s.spin_limit, lds);
                else
                    item_sweep<double, double, 32, false>(W, tab, fwd, p, q.R, q.n, (double*)q.Y, q.n, (double*)q.D, q.n, q.X,
                                                   q.n, q.nrhs, fl, s.err, s.spin_limit, lds);
            } else {
                if (q.nrhs <= 16)
                    item_sweep<float, float, 16, false>(W, tab, fwd, p, q.R, q.n, (float*)q.Y, q.n, (float*)q.D, q.n, q.X,
                                                 q.n, q.nrhs, fl, s.err, s.spin_limit, lds);
                else
                    item_sweep<float, float, 32, false>(W, tab, fwd, p, q.R, q.n, (float*)q.Y, q.n, (float*)q.D, q.n, q.X,
                                                 q.n, q.nrhs, fl, s.err, s.spin_limit, lds);
            }
            break;
        }
        case S_RESID: {
            const SweepTab tab{q.P, q.blk_tile, q.blk_off};
            if (q.nrhs <= 16)
                item_resid<16, false>(s.mats[q.mat_a], tab, it, q.B, q.n, q.X, q.n, q.R, q.n, q.nrhs, q.norms, 0, lds);
            else
                item_resid<32, false>(s.mats[q.mat_a], tab, it, q.B, q.n, q.X, q.n, q.R, q.n, q.nrhs, q.norms, 0, lds);
            break;
        }
        case S_CHECK: {
            __shared__ int s_state;  // SRV_RUNNING = re-arm
            __shared__ double s_eta;
            if (threadIdx.x == 0) {
                const double an = bits_to_double(*q.anorm);
                double eta = 0.0;
                for (int j = 0; j < q.nrhs; j++) {
                    const double rn = bits_to_double(q.norms[3 * j]), xn = bits_to_double(q.norms[3 * j + 1]),
                                 bn = bits_to_double(q.norms[3 * j + 2]);
                    const double den = an * xn + bn;
                    double e = den > 0.0 ? rn / den : (rn > 0.0 ? INFINITY : 0.0);
                    if (e != e) e = INFINITY;
                    eta = fmax(eta, e);
                }
                q.iter++;
                const int info = *s.mats[q.mat_w].info;
                int state = SRV_RUNNING;
                if (info != 0) state = SRV_NOT_SPD;
                else if (eta <= q.tol) state = SRV_OK;
                else {
                    if (!(eta <= 0.5 * q.eta_prev)) q.stall++;
                    else q.stall = 0;
                    if (q.stall >= 2 || q.iter >= q.max_iter) state = SRV_STALLED;
                }
                q.eta_prev = eta;
                s_state = state;
                s_eta = eta;
            }
            __syncthreads();
            const int state = s_state;
            if (state != SRV_RUNNING) {  // done: solution to the caller's pinned buffer, then the record
                const long tot = (long)q.n * q.nrhs;
                for (long e = threadIdx.x; e < tot; e += NTHREADS) q.x_host[e] = q.X[e];
                __atomic_thread_fence(__ATOMIC_RELEASE);  // system scope: each wave's part of X reaches the host
                __syncthreads();
                if (threadIdx.x == 0) {
                    SrvDone& d = s.done[q.done_slot];
                    d.iterations = q.iter;
                    d.eta = s_eta;
                    d.info = *s.mats[q.mat_w].info;
                    d.t_done = wall_clock64();
                    __hip_atomic_store(&d.state, state, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                    __hip_atomic_fetch_add(s.active, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {  // another refinement step: reset the chain and start it again
                for (int e = threadIdx.x; e < q.P; e += NTHREADS) {
                    q.flags_f[e] = 0;
                    q.flags_b[e] = 0;
                }
                for (int e = threadIdx.x; e < 3 * 32; e += NTHREADS) q.norms[e] = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __syncthreads();
                if (threadIdx.x == 0) {
                    const int ids[4] = {q.t_fwd, q.t_bwd, q.t_res, q.t_chk};
                    for (int c = 0; c < 4; c++) s.items_left[ids[c]] = s.tasks[ids[c]].n_items;
                    s.deps[q.t_bwd] = 1;
                    s.deps[q.t_res] = 1;
                    s.deps[q.t_chk] = 1;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                    srv_push(s, q.t_fwd);
                }
            }
            break;
        }
        default: {  // factorization tasks
            DevTask d{};
            d.kind = t.kind;
            d.prec = t.prec;
            d.mat = t.mat;
            d.n_items = t.n_items;
            d.i = t.i;
            d.j = t.j;
            d.k = t.k;
            (void)exec_item<V_ALL>(d, it, s.mats, lds, s.err, s.spin_limit);  // request graphs only hold supported kinds
            break;
        }
    }
}

}  // namespace ttad
