// Serving loop: device-resident persistent executor with continuous batching of
// SPD solve requests (host side + kernel).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <thread>
#include <unordered_map>

#include "../../include/tasktorrent_amd/taskflow.hpp"
#include "runtime.hpp"
#include "server_dev.hpp"

namespace ttad {

// ---------------------------------------------------------------------------
// Device loop
// ---------------------------------------------------------------------------
constexpr int SRV_LDS = cmax(LDS_ITEM_MAX, sweep_lds_bytes<double, 32, false>());

__device__ __forceinline__ int64_t srv_try_pop(const DevServer& s, uint32_t& rng) {
    // lane 0 only: claim the head of the first non-empty ring in priority order. A claim takes
    // one credit from the ring's count of unclaimed entries (a claimer that finds none gives it
    // back and rescans after a random back-off), then its slot with a fetch-add on the head:
    // credits never exceed the pushed entries, so the slot never passes the tail.
    for (int b = 0; b < s.n_buckets; b++) {
        const SrvRing& R = s.rings[b];
        if (__hip_atomic_load(R.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= 0) continue;
        if (__hip_atomic_fetch_add(R.count, -1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= 0) {
            __hip_atomic_fetch_add(R.count, 1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            rng = rng * 1664525u + 1013904223u;
            for (uint32_t q = rng >> 28; q > 0; q--) __builtin_amdgcn_s_sleep(1);
            return -1;  // lost the race for the last credits: caller rescans
        }
        const unsigned long long h = __hip_atomic_fetch_add(R.head, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long cap = 1ull << s.ring_log2;
        const unsigned long long lap = ((h >> s.ring_log2) + 1) & 0xFFFFFFull;
        const unsigned long long* slot = R.slots + (h & (cap - 1));
        const int64_t t0 = (int64_t)wall_clock64();
        for (;;) {
            // relaxed: an acquire load would invalidate the L2 on every poll; the caller fences once
            const unsigned long long e = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((e >> 40) == lap) return (int64_t)(e & 0xFFFFFFFFFFull);
            if ((int64_t)wall_clock64() - t0 > s.spin_limit ||
                __hip_atomic_load(s.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) {
                __hip_atomic_store(s.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                return -2;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return -1;
}

__global__ __launch_bounds__(NTHREADS) void k_server(DevServer s) {
    extern __shared__ __align__(16) char lds[];
    __shared__ int64_t s_item;
    const int tid = threadIdx.x;
    int64_t idle_since = (int64_t)wall_clock64();
    uint32_t rng = blockIdx.x * 2654435761u + 0x9e3779b9u;
    for (;;) {
        if (tid == 0) {
            int64_t got = -1;
            for (;;) {
                got = srv_try_pop(s, rng);
                if (got >= 0 || got == -2) break;
                const int shut = __hip_atomic_load((const int32_t*)s.shutdown, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_SYSTEM);
                if (shut >= 2) {  // forced stop
                    got = -2;
                    break;
                }
                // doorbell (host memory): polled by one workgroup in eight to limit bus traffic
                const bool poller = (blockIdx.x & 7) == 0;
                // (relaxed polls: an acquire load invalidates caches every time; one fence once posted)
                const int posted = poller ? __hip_atomic_load((const int32_t*)s.db_count, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_SYSTEM)
                                          : 0;
                int taken = __hip_atomic_load(s.db_taken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (poller && taken < posted) {
                    if (__hip_atomic_compare_exchange_strong(s.db_taken, &taken, taken + 1, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        __atomic_thread_fence(__ATOMIC_ACQUIRE);  // system scope: the entry was written before the count
                        const int slot = __hip_atomic_load((const int32_t*)(s.db_entries + (taken % s.db_cap)),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                        got = -100 - (int64_t)slot;  // admission, done by the whole workgroup
                        break;
                    }
                    continue;
                }
                if (shut && __hip_atomic_load(s.active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
                    __hip_atomic_load(s.db_taken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                        __hip_atomic_load((const int32_t*)s.db_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) {
                    got = -2;
                    break;
                }
                if ((int64_t)wall_clock64() - idle_since > s.idle_limit) {
                    __hip_atomic_store(s.err, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    got = -2;
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
            }
            s_item = got;
        }
        __syncthreads();
        const int64_t e = s_item;
        if (e <= -100) {
            srv_admit(s, (int)(-100 - e));
            idle_since = (int64_t)wall_clock64();
            __syncthreads();
            continue;
        }
        if (e < 0) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const int task = (int)(e >> 16), item = (int)(e & 0xFFFF);
        const SrvTask t = s.tasks[task];
        srv_exec(s, t, task, item, lds);
        __syncthreads();
        if (tid == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (t.kind != S_CHECK) {
                const int left =
                    __hip_atomic_fetch_add(s.items_left + task, -1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) - 1;
                if (left == 0) srv_complete(s, task);
            }
        }
        idle_since = (int64_t)wall_clock64();
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// First-fit range allocator with coalescing (offsets in units).
class RangeAlloc {
  public:
    explicit RangeAlloc(long long size = 0) { reset(size); }
    void reset(long long size) {
        free_.clear();
        if (size > 0) free_[0] = size;
        size_ = size;
        used_ = 0;
    }
    long long alloc(long long n, long long align = 1) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            const long long a = (it->first + align - 1) / align * align;
            const long long end = it->first + it->second;
            if (a + n <= end) {
                const long long b0 = it->first, b1 = end;
                free_.erase(it);
                if (a > b0) free_[b0] = a - b0;
                if (a + n < b1) free_[a + n] = b1 - (a + n);
                used_ += n;
                return a;
            }
        }
        return -1;
    }
    void release(long long off, long long n) {
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    long long used() const { return used_; }
    long long size() const { return size_; }

  private:
    std::map<long long, long long> free_;
    long long size_ = 0, used_ = 0;
};

struct Template {  // request task graph, task ids relative to the request's range
    std::vector<SrvTask> tasks;
    std::vector<int32_t> deps0;
    std::vector<int32_t> edges;
    int t_init = 0, t_fwd = 0, t_bwd = 0, t_res = 0, t_chk = 0;
    int items = 0;
};

struct Block {  // arena allocation
    long long off = -1, bytes = 0;
};

struct Factor {  // a working factor in the arena (reusable for new right-hand sides)
    long long id = -1;
    int slot = -1;
    Block mem;
    long long a_id = -1;  // matrix it factors
    int storage = P_F32;
    int users = 0;
    bool ready = false;
    uint64_t last_use = 0;
};

}  // namespace ttad

struct ttad_server {
    ttad_context* ctx = nullptr;
    ttad_server_config cfg{};
    hipStream_t s_run = nullptr;
    bool running = false;
    // device pools
    ttad::DevBuf arena, tasks, edges, deps, items_left, reqs, mats, ring_hdr, ring_slots, misc;
    int ring_log2 = 0;
    ttad::RangeAlloc arena_alloc, task_alloc, edge_alloc;
    std::vector<int> free_req_slots, free_mat_slots;
    // host pinned (device-visible): doorbell, mail, completion records, mailbox and result pools
    ttad::SrvDone* done = nullptr;
    ttad::SrvMail* mail = nullptr;
    int32_t* db_entries = nullptr;
    int32_t* db_count = nullptr;
    int32_t* shutdown = nullptr;
    int32_t* err = nullptr;
    char* mbox = nullptr;          // mailbox pool (copy lists + payloads)
    char* rbuf = nullptr;          // result pool (solutions)
    ttad::RangeAlloc mbox_alloc, rbuf_alloc;
    long long mbox_bytes = 0, rbuf_bytes = 0;
    int db_posted = 0;
    int workers = 0;
    uint64_t clock_rate_hz = 100000000;
    // bookkeeping
    struct Pending {
        long long id;
        ttad_matrix* A;
        int prec, nrhs, max_iter;
        double tol;
        const double* B;
        double* X;
        long long reuse;   // factor id to reuse, -1 = factor A
        int keep;          // keep the factor after completion
        double t_submit;   // host seconds
        int fell_back = 0;    // this is the FP64 re-run of a request whose low-precision attempt failed
        int prior_iters = 0;  // refinement steps of that attempt
        double t_admit0 = -1; // first admission (host seconds)
    };
    struct Active {
        Pending p;
        int slot = -1;
        long long t0 = -1, ntasks = 0, e0 = -1, nedges = 0;
        ttad::Block vec;
        long long factor = -1;
        long long items = 0;
        long long mbox_off = -1, mbox_len = 0, rbuf_off = -1, rbuf_len = 0;
        double t_admit = 0;
    };
    long long items_cap = 0, items_admitted = 0;
    std::deque<Pending> pending;
    std::unordered_map<long long, Active> active;
    std::unordered_map<long long, ttad::Factor> factors;
    std::unordered_map<long long, std::pair<int, int>> a_slots;  // matrix id -> (DevMat slot of A, active requests)
    std::map<std::pair<std::vector<int>, int>, std::shared_ptr<ttad::Template>> templates;
    long long next_id = 1;
    uint64_t use_clock = 0;
    std::deque<ttad_server_result> results;
    std::chrono::steady_clock::time_point epoch = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - epoch).count();
    }
    // device addresses of the pinned allocations (base host pointer, base device pointer, bytes)
    std::vector<std::tuple<const char*, unsigned long long, size_t>> pinned_maps;
    unsigned long long dev_of(const void* host) const {
        const char* h = (const char*)host;
        for (const auto& m : pinned_maps)
            if (h >= std::get<0>(m) && h < std::get<0>(m) + std::get<2>(m))
                return std::get<1>(m) + (unsigned long long)(h - std::get<0>(m));
        return 0;
    }
    void* pin(size_t bytes) {
        void* p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return nullptr;
        memset(p, 0, bytes);
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
            (void)hipHostFree(p);
            return nullptr;
        }
        pinned_maps.emplace_back((const char*)p, (unsigned long long)d, bytes);
        return p;
    }
};

namespace ttad {

static constexpr int RING_BUCKETS = 4;
static constexpr int MAX_RHS = 32;

static int bucket_of(int kind) {
    switch (kind) {
        case S_INIT:
        case S_CHECK:
        case S_FWD:
        case S_BWD:
        case S_RESID:
            return 0;
        case K_POTRF:
        case K_TRSM:
            return 1;
        case K_GEMM:
        case K_SYRK:
            return 2;
        default:
            return 3;
    }
}

int cholesky_dag_compile(int nb, int prio_kind, std::vector<ttad_task_desc>& desc, std::vector<int32_t>& indeg,
                         std::vector<int32_t>& out_ptr, std::vector<int32_t>& out_idx);

// Request graph: INIT -> {CONVERT -> factor PTG -> FWD, ANORM -> RESID}, FWD -> BWD -> RESID -> CHECK.
static std::shared_ptr<Template> build_template(const std::vector<int>& bsz, int prec, bool factor, int& rc) {
    auto T = std::make_shared<Template>();
    const int nt = (int)bsz.size();
    int P = 0;
    for (int b : bsz) P += (b + SB - 1) / SB;
    auto add = [&](int kind, int items, int i, int j, int k, int d0) {
        SrvTask t{};
        t.kind = kind;
        t.prec = prec;
        t.n_items = std::max(1, items);
        t.i = i;
        t.j = j;
        t.k = k;
        t.bucket = bucket_of(kind);
        T->tasks.push_back(t);
        T->deps0.push_back(d0);
        return (int)T->tasks.size() - 1;
    };
    std::vector<std::vector<int>> out;
    auto edge = [&](int a, int b) {
        if ((int)out.size() <= a) out.resize(a + 1);
        out[a].push_back(b);
    };
    const int t_init = add(S_INIT, 1, 0, 0, 0, 0);
    const int t_anorm = add(S_ANORM, P, 0, 0, 0, 1);
    edge(t_init, t_anorm);
    int t_last = t_init;
    if (factor) {
        const int t_conv = add(S_CONVERT, nt * (nt + 1) / 2, 0, 0, 0, 1);
        edge(t_init, t_conv);
        std::vector<ttad_task_desc> desc;
        std::vector<int32_t> indeg, op, oi;
        rc = cholesky_dag_compile(nt, 2, desc, indeg, op, oi);
        if (rc != TTAD_OK) return nullptr;
        const int base = (int)T->tasks.size();
        int potrf0 = -1, sink = -1;
        for (size_t q = 0; q < desc.size(); q++) {
            const ttad_task_desc& d = desc[q];
            const int kind = d.kind;
            const int items =
                kind == K_POTRF ? potrf_row_items(bsz[d.j]) : items_of(kind, bsz[d.i], bsz[d.j], prec == P_F64 ? 8 : 4);
            const int dp = (kind == K_GEMM || kind == K_SYRK) ? prec : (prec == P_F64 ? P_F64 : P_F32);
            const int id = add(kind, items, d.i, d.j, d.k, indeg[q]);
            T->tasks[id].prec = dp;
            if (kind == K_POTRF && d.j == 0) potrf0 = id;
            if (kind == K_POTRF && d.j == nt - 1) sink = id;
        }
        for (size_t q = 0; q < desc.size(); q++)
            for (int e = op[q]; e < op[q + 1]; e++) edge(base + (int)q, base + oi[e]);
        edge(t_conv, potrf0);  // replaces the host seed of POTRF(0) (indegree stays 1)
        t_last = sink;
    }
    const int t_fwd = add(S_FWD, P, 0, 0, 0, 1);
    const int t_bwd = add(S_BWD, P, 0, 0, 0, 1);
    const int t_res = add(S_RESID, P, 0, 0, 0, 2);
    const int t_chk = add(S_CHECK, 1, 0, 0, 0, 1);
    edge(t_last, t_fwd);
    edge(t_fwd, t_bwd);
    edge(t_bwd, t_res);
    edge(t_anorm, t_res);
    edge(t_res, t_chk);
    out.resize(T->tasks.size());
    for (size_t t = 0; t < T->tasks.size(); t++) {
        T->tasks[t].edge_off = (int)T->edges.size();
        T->tasks[t].edge_cnt = (int)out[t].size();
        for (int v : out[t]) T->edges.push_back(v);
        T->items += T->tasks[t].n_items;
    }
    T->t_init = t_init;
    T->t_fwd = t_fwd;
    T->t_bwd = t_bwd;
    T->t_res = t_res;
    T->t_chk = t_chk;
    rc = TTAD_OK;
    return T;
}

static size_t tile_bytes_of(const std::vector<int>& bsz, int storage, std::vector<size_t>& off) {
    const int nt = (int)bsz.size();
    const size_t es = storage == P_F64 ? 8 : 4, align = 256 / es;
    off.assign((size_t)nt * nt, 0);
    size_t o = 0;
    for (int j = 0; j < nt; j++)
        for (int i = j; i < nt; i++) {
            off[i + (size_t)j * nt] = o;
            o += ((size_t)bsz[i] * bsz[j] + align - 1) / align * align;
        }
    return o * es;
}

static long long align_up(long long x, long long a) { return (x + a - 1) / a * a; }

struct FactorLayout {
    long long total = 0;    // bytes of the factor's arena block
    long long tab_off = 0;  // offset of its tables (bsz, boff, pointers, info) inside the block
};

// Builds the DevMat of a factor living in the arena at `base` (tiles, dinv, tables).
static FactorLayout factor_layout(const std::vector<int>& bsz, int storage, std::vector<char>& host_tables,
                                  long long base_dev_off, char* arena_dev, DevMat& d) {
    const int nt = (int)bsz.size();
    std::vector<size_t> toff;
    const size_t tbytes = tile_bytes_of(bsz, storage, toff);
    size_t dinv_elems = 0;
    std::vector<size_t> doff(nt);
    for (int j = 0; j < nt; j++) {
        doff[j] = dinv_elems;
        dinv_elems += (size_t)((bsz[j] + DB - 1) / DB) * DB * DB;
    }
    long long o = 0;
    const long long o_tiles = o;
    o = align_up(o + (long long)tbytes, 256);
    const long long o_dinv = o;
    o = align_up(o + (long long)dinv_elems * 8, 256);
    const long long o_bsz = o;
    o = align_up(o + nt * 4, 256);
    const long long o_boff = o;
    o = align_up(o + (nt + 1) * 4, 256);
    const long long o_tab = o;
    o = align_up(o + (long long)nt * nt * 8, 256);
    const long long o_dtab = o;
    o = align_up(o + nt * 8, 256);
    const long long o_info = o;
    o = align_up(o + 8, 256);
    const long long o_pflags = o;  // row-block POTRF progress flags (uploaded as zeros with the tables)
    o = align_up(o + (long long)nt * POTRF_FLAGS * 4, 256);
    const long long total = o;
    // host image of the tables (bsz, boff, tile pointers, dinv pointers, info)
    host_tables.assign((size_t)(total - o_bsz), 0);
    char* h = host_tables.data();
    auto dev = [&](long long off) { return arena_dev + base_dev_off + off; };
    int32_t* hb = (int32_t*)(h + (o_bsz - o_bsz));
    int32_t* ho = (int32_t*)(h + (o_boff - o_bsz));
    void** ht = (void**)(h + (o_tab - o_bsz));
    double** hd = (double**)(h + (o_dtab - o_bsz));
    int acc = 0;
    for (int i = 0; i < nt; i++) {
        hb[i] = bsz[i];
        ho[i] = acc;
        acc += bsz[i];
    }
    ho[nt] = acc;
    const size_t es = storage == P_F64 ? 8 : 4;
    for (int j = 0; j < nt; j++) {
        hd[j] = (double*)dev(o_dinv) + doff[j];
        for (int i = j; i < nt; i++) ht[i + (size_t)j * nt] = dev(o_tiles) + toff[i + (size_t)j * nt] * es;
    }
    d = DevMat{};
    d.nt = nt;
    d.storage = storage;
    d.bsz = (const int32_t*)dev(o_bsz);
    d.boff = (const int32_t*)dev(o_boff);
    d.tiles = (void* const*)dev(o_tab);
    d.dinv = (double* const*)dev(o_dtab);
    d.info = (int32_t*)dev(o_info);
    d.pflags = (int32_t*)dev(o_pflags);
    return FactorLayout{total, o_bsz};
}

}  // namespace ttad

using namespace ttad;

extern "C" {

int ttad_server_create(ttad_context* ctx, const ttad_server_config* cfg_in, ttad_server** out) {
    if (!ctx || !out) return set_error(TTAD_ERR_INVALID, "server_create: bad args");
    ttad_server_config cfg{};
    if (cfg_in) cfg = *cfg_in;
    if (cfg.hbm_budget <= 0) cfg.hbm_budget = 4LL << 30;
    if (cfg.max_tasks <= 0) cfg.max_tasks = 1 << 20;
    if (cfg.max_requests <= 0) cfg.max_requests = 1024;
    if (cfg.max_tasks > (1 << 23)) return set_error(TTAD_ERR_INVALID, "server_create: max_tasks > 2^23");
    auto S = std::make_unique<ttad_server>();
    S->ctx = ctx;
    S->cfg = cfg;
    TTAD_HIP(hipSetDevice(ctx->device));
    TTAD_HIP(hipStreamCreateWithFlags(&S->s_run, hipStreamNonBlocking));
    if (S->arena.alloc((size_t)cfg.hbm_budget) != hipSuccess)
        return set_error(TTAD_ERR_OOM, "server_create: arena of %lld bytes does not fit", cfg.hbm_budget);
    S->arena_alloc.reset(cfg.hbm_budget);
    const int T = cfg.max_tasks;
    TTAD_HIP(S->tasks.alloc((size_t)T * sizeof(SrvTask)));
    TTAD_HIP(S->edges.alloc((size_t)T * 4 * 4));
    TTAD_HIP(S->deps.alloc((size_t)T * 4));
    TTAD_HIP(S->items_left.alloc((size_t)T * 4));
    TTAD_HIP(S->reqs.alloc((size_t)cfg.max_requests * sizeof(SrvReq)));
    const int max_mats = 2 * cfg.max_requests + 64;
    TTAD_HIP(S->mats.alloc((size_t)max_mats * sizeof(DevMat)));
    S->task_alloc.reset(T);
    S->edge_alloc.reset((long long)T * 4);
    const int lg = 21;  // ring capacity per bucket; admission keeps admitted items below it
    S->ring_log2 = lg;
    S->items_cap = 1LL << lg;
    TTAD_HIP(S->ring_hdr.alloc(RING_BUCKETS * 3 * 8));
    TTAD_HIP(S->ring_slots.alloc((size_t)RING_BUCKETS << lg << 3));
    TTAD_HIP(hipMemset(S->ring_hdr.p, 0, RING_BUCKETS * 3 * 8));
    TTAD_HIP(hipMemset(S->ring_slots.p, 0, (size_t)RING_BUCKETS << lg << 3));
    TTAD_HIP(S->misc.alloc(256));
    TTAD_HIP(hipMemset(S->misc.p, 0, 256));
    for (int r = cfg.max_requests - 1; r >= 0; r--) S->free_req_slots.push_back(r);
    for (int m = max_mats - 1; m >= 0; m--) S->free_mat_slots.push_back(m);
    const int db_cap = 1 << 16;
    S->done = (SrvDone*)S->pin(sizeof(SrvDone) * cfg.max_requests);
    S->mail = (SrvMail*)S->pin(sizeof(SrvMail) * cfg.max_requests);
    S->db_entries = (int32_t*)S->pin(sizeof(int32_t) * (db_cap + 64));
    S->mbox_bytes = 256LL << 20;
    S->rbuf_bytes = 64LL << 20;
    S->mbox = (char*)S->pin((size_t)S->mbox_bytes);
    S->rbuf = (char*)S->pin((size_t)S->rbuf_bytes);
    if (!S->done || !S->mail || !S->db_entries || !S->mbox || !S->rbuf)
        return set_error(TTAD_ERR_OOM, "server_create: pinned host memory");
    S->mbox_alloc.reset(S->mbox_bytes);
    S->rbuf_alloc.reset(S->rbuf_bytes);
    S->db_count = S->db_entries + db_cap;
    S->shutdown = S->db_entries + db_cap + 16;
    S->err = S->db_entries + db_cap + 32;
    int rate_khz = 100000;
    (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device);
    S->clock_rate_hz = (uint64_t)rate_khz * 1000ull;
    // all set-up work is finished before the device loop starts: while it runs, the host
    // only touches pinned memory (a runtime copy or memset could need the CUs it holds)
    TTAD_HIP(hipDeviceSynchronize());
    DevServer d{};
    d.n_buckets = RING_BUCKETS;
    d.ring_log2 = lg;
    unsigned long long* hdr = S->ring_hdr.as<unsigned long long>();
    for (int b = 0; b < RING_BUCKETS; b++) {
        d.rings[b].head = hdr + 3 * b;
        d.rings[b].tail = hdr + 3 * b + 1;
        d.rings[b].count = (long long*)(hdr + 3 * b + 2);
        d.rings[b].slots = S->ring_slots.as<unsigned long long>() + ((size_t)b << lg);
    }
    d.tasks = S->tasks.as<SrvTask>();
    d.edges = S->edges.as<int32_t>();
    d.deps = S->deps.as<int32_t>();
    d.items_left = S->items_left.as<int32_t>();
    d.reqs = S->reqs.as<SrvReq>();
    d.mats = S->mats.as<DevMat>();
    d.done = (SrvDone*)S->dev_of(S->done);
    d.mail = (SrvMail*)S->dev_of(S->mail);
    const int32_t* dbe = (const int32_t*)S->dev_of(S->db_entries);
    d.db_entries = dbe;
    d.db_count = dbe + db_cap;
    d.shutdown = dbe + db_cap + 16;
    d.err = (int32_t*)(dbe + db_cap + 32);
    d.db_cap = db_cap;
    int32_t* misc = S->misc.as<int32_t>();
    d.db_taken = misc;
    d.active = misc + 16;
    d.spin_limit = (int64_t)S->clock_rate_hz * 10;
    d.idle_limit = (int64_t)S->clock_rate_hz * (cfg.idle_timeout_s > 0 ? cfg.idle_timeout_s : 600);
    if (!d.done || !d.mail || !dbe) return set_error(TTAD_ERR_DEVICE, "server_create: pinned memory not mapped");
    TTAD_HIP(hipFuncSetAttribute((const void*)k_server, hipFuncAttributeMaxDynamicSharedMemorySize, SRV_LDS));
    int per_cu = 1;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_server, NTHREADS, SRV_LDS);
    S->workers = cfg.workers > 0 ? cfg.workers : ctx->n_cu * std::max(1, per_cu);
    hipLaunchKernelGGL(k_server, dim3(S->workers), dim3(NTHREADS), SRV_LDS, S->s_run, d);
    TTAD_HIP(hipGetLastError());
    S->running = true;
    *out = S.release();
    return TTAD_OK;
}

int ttad_server_destroy(ttad_server* S) {
    if (!S) return TTAD_OK;
    int rc = TTAD_OK;
    if (S->running) {
        // graceful: finish the requests in flight; forced after 30 s (err ends every device
        // wait, workers exit at their next poll)
        __atomic_store_n(S->shutdown, 1, __ATOMIC_RELEASE);
        auto t0 = std::chrono::steady_clock::now();
        while (hipStreamQuery(S->s_run) == hipErrorNotReady) {
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > 30.0) {
                if (__atomic_load_n(S->err, __ATOMIC_ACQUIRE) == 0) __atomic_store_n(S->err, 4, __ATOMIC_RELEASE);
                __atomic_store_n(S->shutdown, 2, __ATOMIC_RELEASE);
                rc = set_error(TTAD_ERR_TIMEOUT, "server_destroy: requests still in flight after 30 s; forced stop");
                break;
            }
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
        (void)hipStreamSynchronize(S->s_run);
        S->running = false;
    }
    for (void* p : {(void*)S->done, (void*)S->mail, (void*)S->db_entries, (void*)S->mbox, (void*)S->rbuf})
        if (p) (void)hipHostFree(p);
    (void)hipStreamDestroy(S->s_run);
    delete S;
    return rc;
}

}  // extern "C"

namespace ttad {

// Builds the admission mail of one request: payloads and the list of copies the
// admitting workgroup performs (host pinned -> HBM).
struct MailBuilder {
    ttad_server* S;
    long long off, cap;  // mailbox range
    long long used = 0;
    std::vector<SrvCopy> copies;
    bool ok = true;
    // Exact sizes: the device writes `bytes` and nothing after them (the next words of a
    // task or edge range may belong to a request that is running).
    void add(void* dst_dev, const void* src, size_t bytes) {
        used = (used + 255) / 256 * 256;
        if (bytes % 4 || (unsigned long long)dst_dev % 4 || used + (long long)bytes > cap) {
            ok = false;
            return;
        }
        char* h = S->mbox + off + used;
        memcpy(h, src, bytes);
        copies.push_back(SrvCopy{(unsigned long long)dst_dev, S->dev_of(h), (unsigned long long)bytes});
        used += (long long)bytes;
    }
    unsigned long long finish() {  // the copy list itself, at the end of the range
        used = (used + 255) / 256 * 256;
        const size_t lb = copies.size() * sizeof(SrvCopy);
        if (used + (long long)lb > cap) {
            ok = false;
            return 0;
        }
        char* h = S->mbox + off + used;
        memcpy(h, copies.data(), lb);
        used += (long long)lb;
        return S->dev_of(h);
    }
};

static long long mail_bytes_estimate(const Template& T, int n, int nrhs, size_t ftab, int P) {
    return (long long)(T.tasks.size() * sizeof(SrvTask) + T.edges.size() * 4 + T.tasks.size() * 8 + sizeof(SrvReq) +
                       2 * sizeof(DevMat) + ftab + (size_t)P * 8 + (size_t)n * nrhs * 8) +
           64 * 256 + 16 * (long long)sizeof(SrvCopy);
}

// Tries to admit one pending request; TTAD_OK if posted, TTAD_ERR_BUSY if resources are short.
static int server_admit(ttad_server* S, const ttad_server::Pending& p) {
    ttad_matrix* A = p.A;
    const int n = A->n, nrhs = p.nrhs;
    const bool factor = p.reuse < 0;
    const int storage = p.prec == P_F64 ? P_F64 : P_F32;
    std::vector<int32_t> bt, bo;
    for (int i = 0; i < A->nt; i++)
        for (int o = 0; o < A->bsz[i]; o += SB) {
            bt.push_back(i);
            bo.push_back(o);
        }
    const int P = (int)bt.size();
    auto key = std::make_pair(A->bsz, (factor ? 0 : 100) + p.prec);
    std::shared_ptr<Template> T;
    auto it = S->templates.find(key);
    if (it == S->templates.end()) {
        int rc = TTAD_OK;
        T = build_template(A->bsz, p.prec, factor, rc);
        if (!T) return rc;
        S->templates[key] = T;
    } else {
        T = it->second;
    }
    if (S->free_req_slots.empty()) return TTAD_ERR_BUSY;
    if (S->items_admitted + T->items > S->items_cap) return TTAD_ERR_BUSY;
    Factor* F = nullptr;
    if (!factor) {
        auto fit = S->factors.find(p.reuse);
        if (fit == S->factors.end() || !fit->second.ready)
            return set_error(TTAD_ERR_NOT_FOUND, "server: factor %lld not available", p.reuse);
        F = &fit->second;
    }
    // vectors in the arena: B, X, R (FP64), Y, D (factor precision), flags, norms, anorm, block tables
    const size_t es = storage == P_F64 ? 8 : 4;
    long long vb = 0;
    auto take = [&](long long bytes) {
        const long long o = vb;
        vb = align_up(vb + bytes, 256);
        return o;
    };
    const long long oB = take((long long)n * nrhs * 8), oX = take((long long)n * nrhs * 8),
                    oR = take((long long)n * nrhs * 8), oY = take((long long)n * nrhs * es),
                    oD = take((long long)n * nrhs * es), oFf = take(P * 4), oFb = take(P * 4),
                    oN = take(3 * 32 * 8), oAn = take(8), oBt = take(P * 4), oBo = take(P * 4);
    std::vector<char> ftab;
    DevMat fdesc{};
    long long fbytes = 0, ftab_off = 0;
    if (factor) {
        const FactorLayout fl = factor_layout(A->bsz, storage, ftab, 0, nullptr, fdesc);
        fbytes = fl.total;
        ftab_off = fl.tab_off;
    }
    const long long mneed = mail_bytes_estimate(*T, n, nrhs, ftab.size(), P);
    const long long rneed = (long long)n * nrhs * 8;
    const long long moff = S->mbox_alloc.alloc(mneed, 256);
    if (moff < 0) return TTAD_ERR_BUSY;
    const long long roff = S->rbuf_alloc.alloc(rneed, 256);
    if (roff < 0) {
        S->mbox_alloc.release(moff, mneed);
        return TTAD_ERR_BUSY;
    }
    long long voff = S->arena_alloc.alloc(vb, 256);
    long long foff = -1;
    auto undo = [&]() {
        S->mbox_alloc.release(moff, mneed);
        S->rbuf_alloc.release(roff, rneed);
        if (voff >= 0) S->arena_alloc.release(voff, vb);
        if (foff >= 0) S->arena_alloc.release(foff, fbytes);
    };
    if (voff < 0) {
        undo();
        return TTAD_ERR_BUSY;
    }
    if (factor) {
        foff = S->arena_alloc.alloc(fbytes, 256);
        while (foff < 0) {  // evict the least recently used idle cached factor
            long long victim = -1;
            uint64_t oldest = ~0ull;
            for (auto& kv : S->factors)
                if (kv.second.users == 0 && kv.second.ready && kv.second.last_use < oldest) {
                    oldest = kv.second.last_use;
                    victim = kv.first;
                }
            if (victim < 0) break;
            Factor& v = S->factors[victim];
            S->arena_alloc.release(v.mem.off, v.mem.bytes);
            S->free_mat_slots.push_back(v.slot);
            S->factors.erase(victim);
            foff = S->arena_alloc.alloc(fbytes, 256);
        }
        if (foff < 0) {
            undo();
            return TTAD_ERR_BUSY;
        }
    }
    const size_t nt = T->tasks.size(), ne = T->edges.size();
    const long long t0 = S->task_alloc.alloc((long long)nt);
    const long long e0 = S->edge_alloc.alloc(std::max<long long>(1, (long long)ne));
    const bool need_a_slot = S->a_slots.find(A->id) == S->a_slots.end();
    const size_t slots_needed = (factor ? 1 : 0) + (need_a_slot ? 1 : 0);
    if (t0 < 0 || e0 < 0 || S->free_mat_slots.size() < slots_needed) {
        if (t0 >= 0) S->task_alloc.release(t0, (long long)nt);
        if (e0 >= 0) S->edge_alloc.release(e0, std::max<long long>(1, (long long)ne));
        undo();
        return TTAD_ERR_BUSY;
    }
    char* arena = (char*)S->arena.p;
    MailBuilder mb{S, moff, mneed};
    int a_slot;
    if (need_a_slot) {
        a_slot = S->free_mat_slots.back();
        S->free_mat_slots.pop_back();
        S->a_slots[A->id] = {a_slot, 1};
        mb.add(S->mats.as<DevMat>() + a_slot, &A->host_desc, sizeof(DevMat));
    } else {
        auto& as = S->a_slots[A->id];
        a_slot = as.first;
        as.second++;
    }
    int w_slot;
    if (factor) {
        w_slot = S->free_mat_slots.back();
        S->free_mat_slots.pop_back();
        (void)factor_layout(A->bsz, storage, ftab, foff, arena, fdesc);
        Factor nf;
        nf.id = p.id;
        nf.slot = w_slot;
        nf.mem = Block{foff, fbytes};
        nf.a_id = A->id;
        nf.storage = storage;
        nf.users = 1;
        S->factors[p.id] = nf;
        mb.add(arena + foff + ftab_off, ftab.data(), ftab.size());
        mb.add(S->mats.as<DevMat>() + w_slot, &fdesc, sizeof(fdesc));
    } else {
        w_slot = F->slot;
        F->users++;
        F->last_use = ++S->use_clock;
    }
    const int slot = S->free_req_slots.back();
    S->free_req_slots.pop_back();
    std::vector<SrvTask> tk(T->tasks);
    for (auto& t : tk) {
        t.edge_off += (int)e0;
        t.req = slot;
        t.mat = w_slot;
    }
    std::vector<int32_t> ed(T->edges);
    for (auto& v : ed) v += (int)t0;
    std::vector<int32_t> il(nt);
    for (size_t q = 0; q < nt; q++) il[q] = tk[q].n_items;
    SrvReq rq{};
    rq.mat_a = a_slot;
    rq.mat_w = w_slot;
    rq.n = n;
    rq.nrhs = nrhs;
    rq.P = P;
    rq.max_iter = p.max_iter;
    rq.t_fwd = (int)t0 + T->t_fwd;
    rq.t_bwd = (int)t0 + T->t_bwd;
    rq.t_res = (int)t0 + T->t_res;
    rq.t_chk = (int)t0 + T->t_chk;
    rq.done_slot = slot;
    rq.tol = p.tol;
    rq.eta_prev = INFINITY;
    auto dv = [&](long long off) { return arena + voff + off; };
    rq.B = (const double*)dv(oB);
    rq.X = (double*)dv(oX);
    rq.R = (double*)dv(oR);
    rq.Y = dv(oY);
    rq.D = dv(oD);
    rq.flags_f = (int32_t*)dv(oFf);
    rq.flags_b = (int32_t*)dv(oFb);
    rq.norms = (unsigned long long*)dv(oN);
    rq.anorm = (unsigned long long*)dv(oAn);
    rq.blk_tile = (const int32_t*)dv(oBt);
    rq.blk_off = (const int32_t*)dv(oBo);
    rq.x_host = (double*)S->dev_of(S->rbuf + roff);
    mb.add(S->tasks.as<SrvTask>() + t0, tk.data(), nt * sizeof(SrvTask));
    if (ne) mb.add(S->edges.as<int32_t>() + e0, ed.data(), ne * 4);
    mb.add(S->deps.as<int32_t>() + t0, T->deps0.data(), nt * 4);
    mb.add(S->items_left.as<int32_t>() + t0, il.data(), nt * 4);
    mb.add(S->reqs.as<SrvReq>() + slot, &rq, sizeof(rq));
    mb.add(dv(oBt), bt.data(), P * 4);
    mb.add(dv(oBo), bo.data(), P * 4);
    mb.add(dv(oB), p.B, (size_t)n * nrhs * 8);
    const unsigned long long cl = mb.finish();
    if (!mb.ok) return set_error(TTAD_ERR_INVALID, "server: admission mail larger than its estimate");
    memset(&S->done[slot], 0, sizeof(SrvDone));
    SrvMail& m = S->mail[slot];
    m.n_copies = (int)mb.copies.size();
    m.seed = (int)t0 + T->t_init;
    m.copies = cl;
    __atomic_store_n(&m.state, 1, __ATOMIC_RELEASE);
    ttad_server::Active a;
    a.p = p;
    a.slot = slot;
    a.t0 = t0;
    a.ntasks = (long long)nt;
    a.e0 = e0;
    a.nedges = std::max<long long>(1, (long long)ne);
    a.vec = Block{voff, vb};
    a.factor = factor ? p.id : p.reuse;
    a.items = T->items;
    a.mbox_off = moff;
    a.mbox_len = mneed;
    a.rbuf_off = roff;
    a.rbuf_len = rneed;
    S->items_admitted += T->items;
    a.t_admit = S->now();
    S->active[p.id] = a;
    // ring the doorbell with the slot
    const int db_cap = 1 << 16;
    S->db_entries[S->db_posted % db_cap] = slot;
    __atomic_thread_fence(__ATOMIC_SEQ_CST);
    S->db_posted++;
    __atomic_store_n(S->db_count, S->db_posted, __ATOMIC_RELEASE);
    return TTAD_OK;
}

static int server_pump(ttad_server* S) {
    // mailboxes the device has copied can be reused
    for (auto& kv : S->active) {
        auto& a = kv.second;
        if (a.mbox_off >= 0 && __atomic_load_n(&S->mail[a.slot].state, __ATOMIC_ACQUIRE) == 2) {
            S->mbox_alloc.release(a.mbox_off, a.mbox_len);
            a.mbox_off = -1;
        }
    }
    while (!S->pending.empty()) {
        const int rc = server_admit(S, S->pending.front());
        if (rc == TTAD_ERR_BUSY) break;
        if (rc != TTAD_OK) {
            ttad_server_result r{};
            r.id = S->pending.front().id;
            r.status = rc;
            r.fell_back = S->pending.front().fell_back;
            r.iterations = S->pending.front().prior_iters;
            S->results.push_back(r);
            S->pending.pop_front();
            continue;
        }
        S->pending.pop_front();
    }
    return TTAD_OK;
}

static void server_release(ttad_server* S, ttad_server::Active& a, bool keep_factor) {
    S->items_admitted -= a.items;
    S->task_alloc.release(a.t0, a.ntasks);
    S->edge_alloc.release(a.e0, a.nedges);
    S->arena_alloc.release(a.vec.off, a.vec.bytes);
    if (a.mbox_off >= 0) S->mbox_alloc.release(a.mbox_off, a.mbox_len);
    S->rbuf_alloc.release(a.rbuf_off, a.rbuf_len);
    S->mail[a.slot].state = 0;
    S->free_req_slots.push_back(a.slot);
    auto ait = S->a_slots.find(a.p.A->id);
    if (ait != S->a_slots.end() && --ait->second.second == 0) {  // no request in flight reads this A
        S->free_mat_slots.push_back(ait->second.first);
        S->a_slots.erase(ait);
    }
    auto fit = S->factors.find(a.factor);
    if (fit != S->factors.end()) {
        Factor& F = fit->second;
        F.users--;
        if (a.p.reuse < 0) F.ready = keep_factor;
        F.last_use = ++S->use_clock;
        if (!F.ready && F.users == 0) {
            S->arena_alloc.release(F.mem.off, F.mem.bytes);
            S->free_mat_slots.push_back(F.slot);
            S->factors.erase(fit);
        }
    }
}

}  // namespace ttad

extern "C" {

int ttad_server_submit(ttad_server* S, ttad_matrix* A, int prec, int nrhs, const double* B, double* X, double tol,
                       int max_iter, long long reuse_factor, int keep_factor, long long* id) {
    if (!S || !A || !B || !X || nrhs <= 0 || nrhs > MAX_RHS || prec < 0 || prec > 3 || A->storage != P_F64 ||
        A->ctx != S->ctx)
        return set_error(TTAD_ERR_INVALID, "server_submit: bad args (nrhs <= %d, FP64 tiles of this context)",
                         MAX_RHS);
    for (int i = 0; i + 1 < A->nt; i++)
        if (A->bsz[i] % SB) return set_error(TTAD_ERR_INVALID, "server_submit: tile sizes must be multiples of %d", SB);
    ttad_server::Pending p{};
    p.id = S->next_id++;
    p.A = A;
    p.prec = prec;
    p.nrhs = nrhs;
    p.B = B;
    p.X = X;
    p.tol = tol > 0 ? tol : 1e-12;
    p.max_iter = max_iter > 0 ? max_iter : 30;
    p.reuse = reuse_factor;
    p.keep = keep_factor;
    p.t_submit = S->now();
    S->pending.push_back(p);
    if (id) *id = p.id;
    return server_pump(S);
}

int ttad_server_poll(ttad_server* S, ttad_server_result* out, int max_out, int* n_out) {
    if (!S || (!out && max_out > 0) || !n_out) return set_error(TTAD_ERR_INVALID, "server_poll: bad args");
    std::vector<long long> finished;
    for (auto& kv : S->active) {
        const int state = __atomic_load_n(&S->done[kv.second.slot].state, __ATOMIC_ACQUIRE);
        if (state != SRV_RUNNING) finished.push_back(kv.first);
    }
    for (long long id : finished) {
        auto& a = S->active[id];
        const SrvDone d = S->done[a.slot];
        if (d.state != SRV_OK && a.p.prec != P_F64 && !a.p.fell_back) {
            // Automatic FP64 fallback: the low-precision factor was not positive definite or refinement
            // stalled. The request goes back to the head of the admission queue with an FP64 factor of A
            // (a reused low-precision factor is dropped for this request) and is reported when that ends.
            ttad_server::Pending p = a.p;
            p.prec = P_F64;
            p.reuse = -1;
            p.fell_back = 1;
            p.prior_iters = d.iterations;
            p.t_admit0 = a.t_admit;
            server_release(S, a, false);
            S->active.erase(id);
            S->pending.push_front(p);
            continue;
        }
        ttad_server_result r{};
        r.id = id;
        r.n = a.p.A->n;
        r.nrhs = a.p.nrhs;
        r.iterations = a.p.prior_iters + d.iterations;
        r.fell_back = a.p.fell_back;
        r.info = d.info;
        r.backward_error = d.eta;
        r.factor_id = a.factor;
        r.reused = a.p.reuse >= 0;
        r.status = d.state == SRV_OK ? TTAD_OK : (d.state == SRV_NOT_SPD ? TTAD_ERR_NOT_SPD : TTAD_ERR_IR_STALLED);
        memcpy(a.p.X, S->rbuf + a.rbuf_off, (size_t)r.n * r.nrhs * 8);  // solution written by the device
        r.latency_s = S->now() - a.p.t_submit;
        r.queue_s = (a.p.t_admit0 >= 0 ? a.p.t_admit0 : a.t_admit) - a.p.t_submit;
        const bool keep = a.p.keep && r.status == TTAD_OK;
        server_release(S, a, keep);
        S->active.erase(id);
        S->results.push_back(r);
    }
    server_pump(S);
    int k = 0;
    while (k < max_out && !S->results.empty()) {
        out[k++] = S->results.front();
        S->results.pop_front();
    }
    *n_out = k;
    const int err = __atomic_load_n(S->err, __ATOMIC_ACQUIRE);
    if (err) return set_error(TTAD_ERR_TIMEOUT, "server: device watchdog code %d", err);
    return TTAD_OK;
}

int ttad_server_query(ttad_server* S, ttad_server_stats* st) {
    if (!S || !st) return set_error(TTAD_ERR_INVALID, "server_query: bad args");
    memset(st, 0, sizeof(*st));
    st->pending = (int)S->pending.size();
    st->active = (int)S->active.size();
    for (auto& kv : S->factors)
        if (kv.second.ready) st->cached_factors++;
    st->arena_used = S->arena_alloc.used();
    st->arena_size = S->arena_alloc.size();
    st->workers = S->workers;
    st->items_run = 0;
    return TTAD_OK;
}

}  // extern "C"
