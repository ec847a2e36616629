// CPU oracle for the tiled dense Cholesky path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker and the timed CPU
// baseline; it is never linked into, loaded by or called from the product
// library (tasktorrent_amd/). Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may use it.
//
// It restates, on the CPU, the algorithm of the reference's hot path:
//   * Threadpool_shared worker loop     tasktorrent/src/threadpool_shared.cpp:17-257
//   * Task + less_pTask                 tasktorrent/src/tasks.hpp:13-29, tasks.cpp:6-16
//   * Taskflow<K> + fulfill_promise     tasktorrent/src/taskflow.hpp:40-297
//   * hash_int_N (FNV-1a over 4N bytes) tasktorrent/src/hashes.hpp:27-57
//   * Logger/DepsLogger CSV formats     tasktorrent/src/util.cpp:51-127
//   * the 2D Cholesky PTG on one rank   miniapp/dense_cholesky/2d_cholesky.cpp:41-690
//     (POTRF 316-361, TRSM 398-464, GEMM/SYRK 469-551, ACCU 556-608,
//      priorities 157-204, block sizes 63-104, timing window 610-628)
//
// Tile arithmetic: the reference calls sequential CBLAS/LAPACKE
// (LAPACKE_dpotrf, cblas_dtrsm, cblas_dsyrk, cblas_dgemm; 2d_cholesky.cpp:319,
// 404-405, 490-494; Makefile.conf.armstrong:8 = sequential MKL). The oracle
// loads the same routines at run time from MKL's single dynamic library
// (libmkl_rt, sequential threading layer) through dlopen with hand-declared
// Fortran-ABI prototypes, or uses its own textbook loops (backend 0).
//
// Parity pins (every function below has at least one; see tests/test_oracle.py):
//   textbook kernels  - exact integer KATs (tests/shared/tests.cpp:361-411 shape),
//                       golden LAPACK fixtures (tests/golden/*.npz), identities
//   tiled DAG         - golden fixtures, exact factor of integer L with power-of-2
//                       diagonal, uniqueness of the Cholesky factor vs numpy LAPACK,
//                       bitwise equality across thread counts / priority kinds
//   block sizes       - the libstdc++ std::mt19937 + uniform_int_distribution call
//                       sequence itself (2d_cholesky.cpp:64-79)
//   scheduler         - run-once / counters-reach-zero invariants of
//                       tests/shared/tests.cpp:294-358 on random DAGs
#include <algorithm>
#include <array>
#include <atomic>
#include <cassert>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <queue>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Timers and logs (util.cpp:13-24, 51-127)
// ---------------------------------------------------------------------------
using clk = std::chrono::high_resolution_clock;
using timer = clk::time_point;
static inline timer wctime() { return clk::now(); }
static inline double elapsed(timer a, timer b) { return std::chrono::duration<double>(b - a).count(); }
static inline double epoch(timer t) { return std::chrono::duration<double>(t.time_since_epoch()).count(); }

struct Event {
    std::string name;
    timer start, end;
};

// Preallocated, lock-free append; drops events once full (util.cpp:51-63).
struct Logger {
    std::vector<Event> ev;
    std::atomic<int> n{0};
    std::atomic<bool> full{false};
    explicit Logger(int cap) : ev(cap) {}
    void record(Event e) {
        int id = n++;
        if (id < (int)ev.size()) ev[id] = std::move(e);
        else full = true;
    }
    void write(const char* path) const {
        std::ofstream f(path);
        f.precision(17);
        int m = std::min((int)ev.size(), n.load());
        for (int i = 0; i < m; i++) f << ev[i].name << "," << epoch(ev[i].start) << "," << epoch(ev[i].end) << "\n";
    }
};

struct DepsLogger {
    std::mutex mtx;
    std::vector<std::pair<std::string, std::string>> edges;
    void add(const std::string& a, const std::string& b) {
        std::lock_guard<std::mutex> l(mtx);
        edges.emplace_back(a, b);
    }
    void write(const char* path) {
        std::ofstream f(path);
        for (auto& e : edges) f << e.first << "," << e.second << "\n";
    }
};

// ---------------------------------------------------------------------------
// Task (tasks.hpp:13-29): run + fulfill closures, priority, name.
// ---------------------------------------------------------------------------
struct Task {
    std::function<void()> run = [] {};
    std::function<void()> fulfill = [] {};
    double priority = 0.0;
    std::string name = "_";
};
struct LessTask {  // max-heap by priority (tasks.cpp:13-16)
    bool operator()(const Task* a, const Task* b) const { return a->priority < b->priority; }
};
using TaskQueue = std::priority_queue<Task*, std::vector<Task*>, LessTask>;

// Name of the task currently running on this thread; fulfill_promise calls made
// from its fulfill() are recorded as dependency edges (2d_cholesky.cpp --depslog).
static thread_local const std::string* tl_running = nullptr;

// ---------------------------------------------------------------------------
// Threadpool (threadpool_shared.cpp:17-257)
// ---------------------------------------------------------------------------
class Threadpool {
  public:
    explicit Threadpool(int n) : ready_(n), ready_mtx_(n), bound_(n), bound_mtx_(n), threads_(n) { start(); }
    ~Threadpool() { if (!joined_) join(); }
    int size() const { return (int)threads_.size(); }
    void set_logger(Logger* l) { logger_ = l; }

    // threadpool_shared.cpp:200-224
    void insert(Task* t, int where, bool binding) {
        if (where < 0 || where >= size()) {
            fprintf(stderr, "oracle: insert where=%d out of range for %s\n", where, t->name.c_str());
            abort();
        }
        ++in_flight_;
        if (!binding) {
            std::lock_guard<std::mutex> l(ready_mtx_[where]);
            ready_[where].push(t);
        } else {
            std::lock_guard<std::mutex> l(bound_mtx_[where]);
            bound_[where].push(t);
        }
    }

    // threadpool_shared.cpp:234-241: release the start token, wait for workers.
    void join() {
        --in_flight_;
        for (auto& t : threads_) t.join();
        joined_ = true;
    }

    long long n_steals() const { return steals_.load(); }

  private:
    std::vector<TaskQueue> ready_;
    std::vector<std::mutex> ready_mtx_;
    std::vector<TaskQueue> bound_;
    std::vector<std::mutex> bound_mtx_;
    std::vector<std::thread> threads_;
    std::atomic<int> in_flight_{0};
    std::atomic<bool> done_{false};
    std::atomic<long long> steals_{0};
    Logger* logger_ = nullptr;
    bool joined_ = false;

    void consume(int self, Task* t) {
        timer t0 = wctime();
        tl_running = &t->name;
        t->run();
        timer t1 = wctime();
        t->fulfill();
        tl_running = nullptr;
        timer t2 = wctime();
        if (logger_) {
            logger_->record({"Wk_" + std::to_string(self) + ">run>" + t->name, t0, t1});
            logger_->record({"Wk_" + std::to_string(self) + ">deps>" + t->name, t1, t2});
        }
        delete t;
        --in_flight_;
    }

    // threadpool_shared.cpp:81-198
    void start() {
        ++in_flight_;  // start token, released by join()
        const int n = size();
        for (int self = 0; self < n; self++) {
            threads_[self] = std::thread([this, self, n]() {
                for (;;) {
                    Task* t = nullptr;
                    {
                        std::lock_guard<std::mutex> lr(ready_mtx_[self]);
                        std::lock_guard<std::mutex> lb(bound_mtx_[self]);
                        TaskQueue& r = ready_[self];
                        TaskQueue& b = bound_[self];
                        if (!r.empty() && !b.empty()) {
                            if (b.top()->priority >= r.top()->priority) { t = b.top(); b.pop(); }
                            else { t = r.top(); r.pop(); }
                        } else if (!r.empty()) { t = r.top(); r.pop(); }
                        else if (!b.empty()) { t = b.top(); b.pop(); }
                    }
                    if (!t) {  // steal round-robin from the other threads' ready queues
                        for (int j = 1; j < n && !t; j++) {
                            int other = (self + j) % n;
                            std::lock_guard<std::mutex> l(ready_mtx_[other]);
                            if (!ready_[other].empty()) {
                                t = ready_[other].top();
                                ready_[other].pop();
                                ++steals_;
                            }
                        }
                    }
                    if (t) {
                        consume(self, t);
                    } else {
                        if (self == 0 && in_flight_.load() == 0) done_.store(true);
                        if (done_.load()) return;
                        std::this_thread::sleep_for(std::chrono::microseconds(10));
                    }
                }
            });
        }
    }
};

// ---------------------------------------------------------------------------
// hash_int_N for std::array<int,N>: FNV-1a over the 4N bytes (hashes.hpp:27-57)
// ---------------------------------------------------------------------------
template <class K> struct HashK { size_t operator()(const K& k) const { return std::hash<K>{}(k); } };
template <size_t N> struct HashK<std::array<int, N>> {
    size_t operator()(const std::array<int, N>& t) const {
        uint32_t h = 0x811C9DC5u;
        const unsigned char* p = reinterpret_cast<const unsigned char*>(t.data());
        for (size_t b = 0; b < 4 * N; b++) h = (p[b] ^ h) * 0x01000193u;
        return h;
    }
};

// ---------------------------------------------------------------------------
// Taskflow<K> (taskflow.hpp:40-297)
// ---------------------------------------------------------------------------
template <class K> class Taskflow {
  public:
    Taskflow(Threadpool* tp, DepsLogger* dl = nullptr) : tp_(tp), dl_(dl), dep_map_(tp->size()) {}
    Taskflow& set_task(std::function<void(K)> f) { f_run_ = std::move(f); return *this; }
    Taskflow& set_fulfill(std::function<void(K)> f) { f_fulfill_ = std::move(f); return *this; }
    Taskflow& set_mapping(std::function<int(K)> f) { f_mapping_ = std::move(f); return *this; }
    Taskflow& set_binding(std::function<bool(K)> f) { f_binding_ = std::move(f); return *this; }
    Taskflow& set_indegree(std::function<int(K)> f) { f_indegree_ = std::move(f); return *this; }
    Taskflow& set_name(std::function<std::string(K)> f) { f_name_ = std::move(f); return *this; }
    Taskflow& set_priority(std::function<double(K)> f) { f_prio_ = std::move(f); return *this; }
    std::string name(K k) const { return f_name_(k); }

    // taskflow.hpp:241-296. Indegree-1 tasks are inserted directly; others go
    // through an internal max-priority task bound to the mapped thread, which
    // owns that thread's dependency map, so counters need no atomics.
    void fulfill_promise(K k) {
        if (dl_ && tl_running) dl_->add(*tl_running, f_name_(k));
        if (f_indegree_(k) == 1) { insert(k); return; }
        const int where = f_mapping_(k);
        Task* t = new Task();
        t->name = "dep_map_intern_" + std::to_string(where);
        t->priority = std::numeric_limits<double>::max();
        t->run = [this, where, k]() {
            auto& m = dep_map_[where];
            auto it = m.find(k);
            if (it == m.end()) it = m.emplace(k, f_indegree_(k)).first;
            const int count = --it->second;
            if (count < 0) {
                fprintf(stderr, "oracle: count < 0 for %s\n", f_name_(k).c_str());
                abort();
            }
            if (count == 0) { m.erase(it); insert(k); }
        };
        tp_->insert(t, where, true);
    }

  private:
    Threadpool* tp_;
    DepsLogger* dl_;
    std::vector<std::unordered_map<K, int, HashK<K>>> dep_map_;
    std::function<void(K)> f_run_ = [](K) {};
    std::function<void(K)> f_fulfill_ = [](K) {};
    std::function<int(K)> f_mapping_ = [](K) { return 0; };
    std::function<bool(K)> f_binding_ = [](K) { return false; };
    std::function<int(K)> f_indegree_ = [](K) { return 1; };
    std::function<std::string(K)> f_name_ = [](K) { return std::string("_"); };
    std::function<double(K)> f_prio_ = [](K) { return 0.0; };

    void insert(K k) {  // make_task (taskflow.hpp:105-115) + Threadpool::insert
        Task* t = new Task();
        t->run = [this, k]() { f_run_(k); };
        t->fulfill = [this, k]() { f_fulfill_(k); };
        t->name = f_name_(k);
        t->priority = f_prio_(k);
        tp_->insert(t, f_mapping_(k), f_binding_(k));
    }
};

// ---------------------------------------------------------------------------
// Tile arithmetic backends
// ---------------------------------------------------------------------------
// Fortran-ABI LAPACK/BLAS symbols exported by libmkl_rt (hand-declared: no MKL headers here).
typedef void (*dgemm_t)(const char*, const char*, const int*, const int*, const int*, const double*, const double*,
                        const int*, const double*, const int*, const double*, double*, const int*);
typedef void (*dsyrk_t)(const char*, const char*, const int*, const int*, const double*, const double*, const int*,
                        const double*, double*, const int*);
typedef void (*dtrsm_t)(const char*, const char*, const char*, const char*, const int*, const int*, const double*,
                        const double*, const int*, double*, const int*);
typedef void (*dpotrf_t)(const char*, const int*, double*, const int*, int*);
typedef int (*mkl_set_layer_t)(int);

struct Blas {
    int kind = 0;  // 0 textbook, 1 MKL
    std::string name = "textbook";
    void* handle = nullptr;
    dgemm_t dgemm = nullptr;
    dsyrk_t dsyrk = nullptr;
    dtrsm_t dtrsm = nullptr;
    dpotrf_t dpotrf = nullptr;
};
static Blas g_blas;

static bool load_mkl() {
    if (g_blas.handle) return true;
    std::vector<std::string> paths;
    if (const char* e = getenv("TTAD_ORACLE_MKL")) paths.push_back(e);
    paths.push_back("libmkl_rt.so.1");
    paths.push_back("libmkl_rt.so");
    paths.push_back("/opt/conda/lib/libmkl_rt.so.1");
    paths.push_back("/opt/conda/lib/libmkl_rt.so");
    for (auto& p : paths) {
        void* h = dlopen(p.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        auto layer = (mkl_set_layer_t)dlsym(h, "MKL_Set_Threading_Layer");
        if (layer) layer(1);  // MKL_THREADING_SEQUENTIAL: one BLAS thread per task, as Makefile.conf.armstrong:8
        g_blas.dgemm = (dgemm_t)dlsym(h, "dgemm_");
        g_blas.dsyrk = (dsyrk_t)dlsym(h, "dsyrk_");
        g_blas.dtrsm = (dtrsm_t)dlsym(h, "dtrsm_");
        g_blas.dpotrf = (dpotrf_t)dlsym(h, "dpotrf_");
        if (g_blas.dgemm && g_blas.dsyrk && g_blas.dtrsm && g_blas.dpotrf) {
            g_blas.handle = h;
            g_blas.name = "mkl_rt(sequential):" + p;
            return true;
        }
        dlclose(h);
    }
    return false;
}

// Textbook column-major kernels. Loop orders are chosen for unit-stride inner loops.
// dpotrf('L'): A = L L^T, lower triangle overwritten; returns LAPACK-style info.
static int tb_potrf(int n, double* A, int lda) {
    for (int j = 0; j < n; j++) {
        double d = A[j + (size_t)j * lda];
        for (int k = 0; k < j; k++) d -= A[j + (size_t)k * lda] * A[j + (size_t)k * lda];
        if (!(d > 0.0)) return j + 1;
        d = std::sqrt(d);
        A[j + (size_t)j * lda] = d;
        for (int k = 0; k < j; k++) {
            const double ljk = A[j + (size_t)k * lda];
            double* cj = A + (size_t)j * lda;
            const double* ck = A + (size_t)k * lda;
            for (int i = j + 1; i < n; i++) cj[i] -= ck[i] * ljk;
        }
        const double inv = 1.0 / d;
        for (int i = j + 1; i < n; i++) A[i + (size_t)j * lda] *= inv;
    }
    return 0;
}
// dtrsm(Right, Lower, Trans, NonUnit, alpha=1): B := B L^{-T}, B is m x n, L is n x n.
static void tb_trsm_rltn(int m, int n, const double* L, int ldl, double* B, int ldb) {
    for (int j = 0; j < n; j++) {
        double* bj = B + (size_t)j * ldb;
        for (int k = 0; k < j; k++) {
            const double ljk = L[j + (size_t)k * ldl];
            const double* bk = B + (size_t)k * ldb;
            for (int i = 0; i < m; i++) bj[i] -= bk[i] * ljk;
        }
        const double d = L[j + (size_t)j * ldl];
        for (int i = 0; i < m; i++) bj[i] /= d;
    }
}
// dsyrk(Lower, NoTrans): C := beta C + alpha A A^T on the lower triangle, A is n x k.
static void tb_syrk_ln(int n, int k, double alpha, const double* A, int lda, double beta, double* C, int ldc) {
    for (int j = 0; j < n; j++) {
        double* cj = C + (size_t)j * ldc;
        for (int i = j; i < n; i++) cj[i] = (beta == 0.0) ? 0.0 : beta * cj[i];
        for (int p = 0; p < k; p++) {
            const double a = alpha * A[j + (size_t)p * lda];
            const double* ap = A + (size_t)p * lda;
            for (int i = j; i < n; i++) cj[i] += ap[i] * a;
        }
    }
}
// dgemm(NoTrans, Trans): C := beta C + alpha A B^T, A m x k, B n x k.
static void tb_gemm_nt(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
                       double beta, double* C, int ldc) {
    for (int j = 0; j < n; j++) {
        double* cj = C + (size_t)j * ldc;
        for (int i = 0; i < m; i++) cj[i] = (beta == 0.0) ? 0.0 : beta * cj[i];
        for (int p = 0; p < k; p++) {
            const double b = alpha * B[j + (size_t)p * ldb];
            const double* ap = A + (size_t)p * lda;
            for (int i = 0; i < m; i++) cj[i] += ap[i] * b;
        }
    }
}

static int k_potrf(int n, double* A, int lda) {
    if (n == 0) return 0;
    if (g_blas.kind == 1) {
        int info = 0;
        g_blas.dpotrf("L", &n, A, &lda, &info);
        return info;
    }
    return tb_potrf(n, A, lda);
}
static void k_trsm(int m, int n, const double* L, int ldl, double* B, int ldb) {
    if (m == 0 || n == 0) return;
    if (g_blas.kind == 1) {
        const double one = 1.0;
        g_blas.dtrsm("R", "L", "T", "N", &m, &n, &one, L, &ldl, B, &ldb);
        return;
    }
    tb_trsm_rltn(m, n, L, ldl, B, ldb);
}
static void k_syrk(int n, int k, double alpha, const double* A, int lda, double beta, double* C, int ldc) {
    if (n == 0) return;
    if (g_blas.kind == 1) {
        g_blas.dsyrk("L", "N", &n, &k, &alpha, A, &lda, &beta, C, &ldc);
        return;
    }
    tb_syrk_ln(n, k, alpha, A, lda, beta, C, ldc);
}
static void k_gemm(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
                   double beta, double* C, int ldc) {
    if (m == 0 || n == 0) return;
    if (g_blas.kind == 1) {
        g_blas.dgemm("N", "T", &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
        return;
    }
    tb_gemm_nt(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// ---------------------------------------------------------------------------
// The tiled Cholesky PTG on one rank (2d_cholesky.cpp:41-690, nprows=npcols=1)
// ---------------------------------------------------------------------------
using int2 = std::array<int, 2>;
using int3 = std::array<int, 3>;

struct CholStats {
    double total = 0, potrf = 0, trsm = 0, gemm = 0, accu = 0;
    long long steals = 0;
    int info = 0;  // first non-positive pivot: 1-based global column (0 = none)
};

// 2d_cholesky.cpp:157-204
static double prio_potrf(int kind, int nb, int j) {
    if (kind == 3) return (double)(9 * (nb - j) - 1) + 18.0 * nb * nb;
    if (kind == 2) return (double)(9 * (nb - j) - 1);
    if (kind == 1) return 3.0 * (double)(nb - j);
    return 3.0;
}
static double prio_trsm(int kind, int nb, int i, int j) {
    if (kind == 3) return (double)((nb - i) + nb * (9.0 * nb - 9.0 * j - 2.0) + 9.0 * nb * nb);
    if (kind == 2) return (double)(9 * (nb - j) - 2);
    if (kind == 1) return 2.0 * (double)(nb - i);
    return 2.0;
}
static double prio_gemm(int kind, int nb, int accumulate, int k, int i, int j) {
    if (kind == 3) {
        if (accumulate) return (double)(nb - i) + nb * (9.0 * nb - 9.0 * j - 2.0);
        return (double)(nb - i) + nb * (9.0 * nb - 3.0 * j - 6.0 * k - 2.0);
    }
    if (kind == 2) return (double)(9 * nb - 9 * j - 2);
    if (kind == 1) return (double)(nb - i);
    return 1.0;
}

static CholStats run_cholesky(int n_threads, int prio_kind, int accumulate, int nb, const int* bs,
                              std::vector<std::vector<double>>& blocks, const char* log_path,
                              const char* deps_path) {
    CholStats st;
    std::vector<int> lda(nb);
    for (int i = 0; i < nb; i++) lda[i] = std::max(1, bs[i]);
    auto blk = [&](int i, int j) -> double* { return blocks[i + (size_t)j * nb].data(); };
    auto block_2_thread = [&](int i, int j) { return (i + j * nb) % n_threads; };  // 2d_cholesky.cpp:124-129

    std::atomic<long long> potrf_us(0), trsm_us(0), gemm_us(0), accu_us(0);
    std::atomic<int> first_bad(0);

    // Parallel-accumulation temporaries, one map per output tile (2d_cholesky.cpp:149-155)
    struct Acc {
        std::map<int, std::vector<double>> pending;
        std::mutex mtx;
    };
    std::vector<Acc> acc(accumulate ? (size_t)nb * nb : 0);

    std::unique_ptr<Logger> logger(log_path ? new Logger(4 * 1000000) : nullptr);
    std::unique_ptr<DepsLogger> dlog(deps_path ? new DepsLogger() : nullptr);

    Threadpool tp(n_threads);
    if (logger) tp.set_logger(logger.get());
    Taskflow<int> potrf(&tp, dlog.get());
    Taskflow<int2> trsm(&tp, dlog.get());
    Taskflow<int3> gemm(&tp, dlog.get());
    Taskflow<int3> accu(&tp, dlog.get());

    potrf.set_task([&](int j) {
            timer a = wctime();
            int info = k_potrf(bs[j], blk(j, j), lda[j]);
            timer b = wctime();
            potrf_us += (long long)(1e6 * elapsed(a, b));
            if (info > 0) {
                int col = 0;
                for (int q = 0; q < j; q++) col += bs[q];
                int expect = 0;
                first_bad.compare_exchange_strong(expect, col + info);
            }
        })
        .set_fulfill([&](int j) {
            for (int i = j + 1; i < nb; i++) trsm.fulfill_promise({i, j});
        })
        .set_indegree([&](int j) { return accumulate ? (j == 0 ? 1 : j) : 1; })
        .set_priority([&](int j) { return prio_potrf(prio_kind, nb, j); })
        .set_mapping([&](int j) { return block_2_thread(j, j); })
        .set_name([](int j) { return "POTRF_" + std::to_string(j); });

    trsm.set_task([&](int2 ij) {
            const int i = ij[0], j = ij[1];
            timer a = wctime();
            k_trsm(bs[i], bs[j], blk(j, j), lda[j], blk(i, j), lda[i]);
            timer b = wctime();
            trsm_us += (long long)(1e6 * elapsed(a, b));
        })
        .set_fulfill([&](int2 ij) {
            const int i = ij[0], j = ij[1];
            for (int k = j + 1; k < i; k++) gemm.fulfill_promise({j, i, k});
            for (int k = i; k < nb; k++) gemm.fulfill_promise({j, k, i});
        })
        .set_indegree([&](int2 ij) { return accumulate ? 1 + ij[1] : 1 + (ij[1] == 0 ? 0 : 1); })
        .set_priority([&](int2 ij) { return prio_trsm(prio_kind, nb, ij[0], ij[1]); })
        .set_mapping([&](int2 ij) { return block_2_thread(ij[0], ij[1]); })
        .set_name([](int2 ij) { return "TRSM_" + std::to_string(ij[0]) + "_" + std::to_string(ij[1]); });

    gemm.set_task([&](int3 kij) {
            const int k = kij[0], i = kij[1], j = kij[2];
            std::vector<double> tmp;
            double* C = blk(i, j);
            double beta = 1.0;
            if (accumulate) {
                tmp.assign((size_t)bs[i] * bs[j], 0.0);
                C = tmp.data();
                beta = 0.0;
            }
            timer a = wctime();
            if (i == j) k_syrk(bs[i], bs[k], -1.0, blk(i, k), lda[i], beta, C, lda[i]);
            else k_gemm(bs[i], bs[j], bs[k], -1.0, blk(i, k), lda[i], blk(j, k), lda[j], beta, C, lda[i]);
            timer b = wctime();
            gemm_us += (long long)(1e6 * elapsed(a, b));
            if (accumulate) {
                std::lock_guard<std::mutex> l(acc[i + (size_t)j * nb].mtx);
                acc[i + (size_t)j * nb].pending[k] = std::move(tmp);
            }
        })
        .set_fulfill([&](int3 kij) {
            const int k = kij[0], i = kij[1], j = kij[2];
            if (accumulate) { accu.fulfill_promise(kij); return; }
            if (k < j - 1) gemm.fulfill_promise({k + 1, i, j});
            else if (i == j) potrf.fulfill_promise(i);
            else trsm.fulfill_promise({i, j});
        })
        .set_indegree([&](int3 kij) {
            if (accumulate) return kij[1] == kij[2] ? 1 : 2;
            return (kij[1] == kij[2] ? 1 : 2) + (kij[0] == 0 ? 0 : 1);
        })
        .set_priority([&](int3 kij) { return prio_gemm(prio_kind, nb, accumulate, kij[0], kij[1], kij[2]); })
        .set_mapping([&](int3 kij) { return block_2_thread(kij[1], kij[2]); })
        .set_binding([](int3) { return false; })
        .set_name([](int3 kij) {
            return "GEMM_" + std::to_string(kij[0]) + "_" + std::to_string(kij[1]) + "_" + std::to_string(kij[2]);
        });

    accu.set_task([&](int3 kij) {
            const int k = kij[0], i = kij[1], j = kij[2];
            std::vector<double> tmp;
            {
                std::lock_guard<std::mutex> l(acc[i + (size_t)j * nb].mtx);
                tmp = std::move(acc[i + (size_t)j * nb].pending[k]);
                acc[i + (size_t)j * nb].pending.erase(k);
            }
            timer a = wctime();
            double* C = blk(i, j);
            const size_t cnt = (size_t)bs[i] * bs[j];
            for (size_t e = 0; e < cnt; e++) C[e] += tmp[e];
            timer b = wctime();
            accu_us += (long long)(1e6 * elapsed(a, b));
        })
        .set_fulfill([&](int3 kij) {
            const int i = kij[1], j = kij[2];
            if (i == j) potrf.fulfill_promise(i);
            else trsm.fulfill_promise({i, j});
        })
        .set_indegree([](int3) { return 1; })
        .set_mapping([&](int3 kij) { return block_2_thread(kij[1], kij[2]); })
        .set_priority([&](int3 kij) { return prio_gemm(prio_kind, nb, accumulate, kij[0], kij[1], kij[2]); })
        .set_binding([](int3) { return true; })
        .set_name([](int3 kij) {
            return "ACCU_" + std::to_string(kij[0]) + "_" + std::to_string(kij[1]) + "_" + std::to_string(kij[2]);
        });

    // Timing window: seed POTRF(0) -> join (2d_cholesky.cpp:610-619)
    timer t0 = wctime();
    potrf.fulfill_promise(0);
    tp.join();
    timer t1 = wctime();

    st.total = elapsed(t0, t1);
    st.potrf = potrf_us.load() * 1e-6;
    st.trsm = trsm_us.load() * 1e-6;
    st.gemm = gemm_us.load() * 1e-6;
    st.accu = accu_us.load() * 1e-6;
    st.steals = tp.n_steals();
    st.info = first_bad.load();
    if (logger) logger->write(log_path);
    if (dlog) dlog->write(deps_path);
    return st;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI for tests / bench (ctypes)
// ---------------------------------------------------------------------------
extern "C" {

// 0 = textbook loops, 1 = MKL (sequential). Returns 0 on success, -1 if MKL is unavailable.
int orc_set_backend(int kind) {
    if (kind == 1) {
        if (!orc::load_mkl()) return -1;
        orc::g_blas.kind = 1;
        return 0;
    }
    orc::g_blas.kind = 0;
    return 0;
}
const char* orc_backend_name(void) { return orc::g_blas.kind == 1 ? orc::g_blas.name.c_str() : "textbook"; }

// Tile kernels, column-major; same argument meaning as the BLAS calls they restate.
int orc_dpotrf(int n, double* A, int lda) { return orc::k_potrf(n, A, lda); }
void orc_dtrsm_rltn(int m, int n, const double* L, int ldl, double* B, int ldb) { orc::k_trsm(m, n, L, ldl, B, ldb); }
void orc_dsyrk_ln(int n, int k, double alpha, const double* A, int lda, double beta, double* C, int ldc) {
    orc::k_syrk(n, k, alpha, A, lda, beta, C, ldc);
}
void orc_dgemm_nt(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb, double beta,
                  double* C, int ldc) {
    orc::k_gemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Block partition of 2d_cholesky.cpp:63-104: std::mt19937(2020) + uniform_int_distribution
// over [2b-ub, ub] for all blocks but the last, which takes the remainder.
int orc_block_sizes(int block_size, int num_blocks, int upper_block_size, int* out) {
    if (upper_block_size < block_size || upper_block_size > 2 * block_size || num_blocks <= 0) return -1;
    const int n = block_size * num_blocks;
    std::mt19937 gen(2020);
    std::uniform_int_distribution<> distrib(2 * block_size - upper_block_size, upper_block_size);
    int acc = 0;
    for (int i = 0; i < num_blocks - 1; i++) {
        int b = std::min(n - acc, distrib(gen));
        acc += b;
        out[i] = b;
    }
    if (n - acc < 0) return -1;
    out[num_blocks - 1] = n - acc;
    return 0;
}

// Runs the tiled Cholesky DAG on n_threads CPU threads. A is n x n column-major
// (lda >= n); its lower triangle is read and overwritten with L. Copies into and
// out of per-tile storage are outside the timed window. stats receives
// [total, potrf, trsm, gemm, accu] seconds, steals, info.
int orc_cholesky(int n_threads, int prio_kind, int accumulate, int num_blocks, const int* block_sizes, double* A,
                 long long lda, double* stats, const char* log_path, const char* deps_path) {
    if (n_threads <= 0 || num_blocks <= 0 || prio_kind < 0 || prio_kind > 3) return -1;
    std::vector<int> displ(num_blocks + 1, 0);
    for (int i = 0; i < num_blocks; i++) displ[i + 1] = displ[i] + block_sizes[i];
    const int nb = num_blocks;
    std::vector<std::vector<double>> blocks((size_t)nb * nb);
    for (int jj = 0; jj < nb; jj++)
        for (int ii = jj; ii < nb; ii++) {
            auto& t = blocks[ii + (size_t)jj * nb];
            const int r = block_sizes[ii], c = block_sizes[jj];
            t.resize((size_t)std::max(1, r) * std::max(1, c));
            for (int q = 0; q < c; q++)
                for (int p = 0; p < r; p++)
                    t[p + (size_t)q * std::max(1, r)] = A[(displ[ii] + p) + (size_t)(displ[jj] + q) * lda];
        }
    orc::CholStats st = orc::run_cholesky(n_threads, prio_kind, accumulate, nb, block_sizes, blocks, log_path, deps_path);
    for (int jj = 0; jj < nb; jj++)
        for (int ii = jj; ii < nb; ii++) {
            auto& t = blocks[ii + (size_t)jj * nb];
            const int r = block_sizes[ii], c = block_sizes[jj];
            for (int q = 0; q < c; q++)
                for (int p = 0; p < r; p++) {
                    if (ii == jj && p < q) continue;  // keep the strict upper triangle of A untouched
                    A[(displ[ii] + p) + (size_t)(displ[jj] + q) * lda] = t[p + (size_t)q * std::max(1, r)];
                }
        }
    if (stats) {
        stats[0] = st.total; stats[1] = st.potrf; stats[2] = st.trsm; stats[3] = st.gemm; stats[4] = st.accu;
        stats[5] = (double)st.steals; stats[6] = (double)st.info;
    }
    return st.info;
}

// The priority run_cholesky gives a task (2d_cholesky.cpp:157-204): task_kind 0 POTRF(j), 1 TRSM(i, j),
// 2 GEMM/SYRK(k, i, j) (SYRK when i == j), prio_kind 0..3, accumulate as the --accumulate flag.
double orc_priority(int prio_kind, int nb, int accumulate, int task_kind, int i, int j, int k) {
    if (task_kind == 0) return orc::prio_potrf(prio_kind, nb, j);
    if (task_kind == 1) return orc::prio_trsm(prio_kind, nb, i, j);
    return orc::prio_gemm(prio_kind, nb, accumulate, k, i, j);
}

// Scheduler check on an arbitrary DAG given in CSR form (tests/shared/tests.cpp:294-358):
// every task runs once, after all its predecessors; order[] receives the completion order.
// Task t has indegree indeg[t]; tasks with indeg 0 are seeded by the main thread.
int orc_run_dag(int n_threads, int n_tasks, const int* indeg, const int* out_ptr, const int* out_idx,
                const double* prio, int* order, int* run_count) {
    using namespace orc;
    std::atomic<int> pos(0);
    std::vector<std::atomic<int>> runs(n_tasks);
    for (auto& r : runs) r = 0;
    {
        Threadpool tp(n_threads);
        Taskflow<int> tf(&tp);
        tf.set_task([&](int k) { runs[k]++; order[pos++] = k; })
            .set_fulfill([&](int k) {
                for (int e = out_ptr[k]; e < out_ptr[k + 1]; e++) tf.fulfill_promise(out_idx[e]);
            })
            .set_indegree([&](int k) { return std::max(1, indeg[k]); })
            .set_mapping([&](int k) { return k % n_threads; })
            .set_priority([&](int k) { return prio ? prio[k] : 0.0; })
            .set_name([](int k) { return "T_" + std::to_string(k); });
        for (int t = 0; t < n_tasks; t++)
            if (indeg[t] == 0) tf.fulfill_promise(t);
        tp.join();
    }
    for (int t = 0; t < n_tasks; t++) run_count[t] = runs[t].load();
    return pos.load();
}

}  // extern "C"
