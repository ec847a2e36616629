// The tiled dense Cholesky PTG of miniapp/dense_cholesky/2d_cholesky.cpp:41-690
// (one rank, non-accumulate mode), written against the Taskflow<K>/Threadpool
// drop-in API and compiled into a persistent device graph.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <array>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tasktorrent_amd/taskflow.hpp"
#include "runtime.hpp"

namespace ttad {

// TTAD_PLANES=0 turns the pre-split bf16x3 operands off (for comparisons).
static bool planes_enabled() {
    static const bool on = [] {
        const char* v = getenv("TTAD_PLANES");
        return !(v && v[0] == '0');
    }();
    return on;
}

using int2 = std::array<int, 2>;
using int3 = std::array<int, 3>;

// Priorities of 2d_cholesky.cpp:157-204 (PrioKind 0 none, 1 row, 2 critical path, 3 cp + row).
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
// accumulate: the accumulate-mode formula of kind 3 (2d_cholesky.cpp:187-189), which ranks the
// updates of one tile alike since they no longer form a chain in k.
static double prio_gemm(int kind, int nb, int k, int i, int j, bool accumulate) {
    if (kind == 3 && accumulate) return (double)(nb - i) + nb * (9.0 * nb - 9.0 * j - 2.0);
    if (kind == 3) return (double)(nb - i) + nb * (9.0 * nb - 3.0 * j - 6.0 * k - 2.0);
    if (kind == 2) return (double)(9 * nb - 9 * j - 2);
    if (kind == 1) return (double)(nb - i);
    return 1.0;
}

// Priority kind 4 (this library's, not the reference's): the length of the longest path from a
// task to the end of the DAG ("bottom level"), with task weights POTRF 2, TRSM 2, GEMM/SYRK 1
// (their measured latencies on the critical path, in units of one GEMM item). Unlike kind 2,
// which ranks every update of column j alike, it also counts the serial chain of updates
// GEMM(i,j,0..j-1) of each tile, so the early updates of far tiles are not left to the end.
// Closed form of the bottom levels of the Cholesky PTG (dependencies of 2d_cholesky.cpp):
//   GEMM(i,j,k): (j-k) wG + (i == j ? P(j) : T(i,j))     (the rest of the tile's chain)
//   T(i,k) = wT + max((i-k) wG + P(i), wG + T(i,k+1))     (via column i, or via row i)
//   P(k)   = wP + max_{i>k} T(i,k),  P(nb-1) = wP
struct BottomLevels {
    static constexpr double wP = 2, wT = 2, wG = 1;
    int nb;
    std::vector<double> P, T;  // T[i + k*nb], i > k
    explicit BottomLevels(int n) : nb(n), P(n, 0.0), T((size_t)n * n, 0.0) {
        for (int k = nb - 1; k >= 0; k--) {
            // T(i,k) for i > k needs P(i) (i > k, done) and T(i,k+1) (done: column k+1 first)
            double best = 0.0;
            for (int i = nb - 1; i > k; i--) {
                double v = (i - k) * wG + P[i];
                if (k + 1 < i) v = std::max(v, wG + T[i + (size_t)(k + 1) * nb]);
                T[i + (size_t)k * nb] = wT + v;
                best = std::max(best, T[i + (size_t)k * nb]);
            }
            P[k] = wP + best;
        }
    }
    double potrf(int j) const { return P[j]; }
    double trsm(int i, int j) const { return T[i + (size_t)j * nb]; }
    double gemm(int k, int i, int j) const { return (j - k) * wG + (i == j ? P[j] : T[i + (size_t)j * nb]); }
};

// Builds the DAG through the Taskflow API exactly as the reference miniapp declares it.
// ctx == nullptr: dry compile into `dry` (no device).
// prio_kind may carry TTAD_ACCUMULATE: the accumulate-mode DAG (2d_cholesky.cpp:149-155, 479-608), with
// the ACCU task fused into its GEMM (the product is added into the tile under a sub-tile lock), so a GEMM
// releases TRSM(i,j) / POTRF(j) directly, which wait for 1 + j / j contributions.
static int build_cholesky_graph(ttad_context* ctx, int nb, int storage, int prec, int prio_kind, ttad_graph** out,
                                Threadpool** dry = nullptr) {
    const bool acc = (prio_kind & TTAD_ACCUMULATE) != 0;
    prio_kind &= ~TTAD_ACCUMULATE;
    std::unique_ptr<BottomLevels> bl;
    if (prio_kind == 4) bl = std::make_unique<BottomLevels>(nb);
    auto tp_owned = std::make_unique<Threadpool>(ctx);
    Threadpool& tp = *tp_owned;
    Taskflow<int> potrf(&tp);
    Taskflow<int2> trsm(&tp);
    Taskflow<int3> gemm(&tp);
    auto block_2_thread = [&](int i, int j) { return (i + j * nb) % 1; };

    potrf.set_task([&](int j) { return DeviceTask{TTAD_POTRF, storage, 0, j, j, j}; })
        .set_fulfill([&](int j) {  // 2d_cholesky.cpp:323-334
            for (int i = j + 1; i < nb; i++) trsm.fulfill_promise({i, j});
        })
        .set_indegree([&](int j) { return acc && j > 0 ? j : 1; })  // 2d_cholesky.cpp:345-351
        .set_priority([&](int j) { return bl ? bl->potrf(j) : prio_potrf(prio_kind, nb, j); })
        .set_mapping([&](int j) { return block_2_thread(j, j); })
        .set_name([](int j) { return "POTRF_" + std::to_string(j); });

    trsm.set_task([&](int2 ij) { return DeviceTask{TTAD_TRSM, storage, 0, ij[0], ij[1], ij[1]}; })
        .set_fulfill([&](int2 ij) {  // 2d_cholesky.cpp:409-431
            const int i = ij[0], j = ij[1];
            for (int k = j + 1; k < i; k++) gemm.fulfill_promise({j, i, k});
            for (int k = i; k < nb; k++) gemm.fulfill_promise({j, k, i});
        })
        .set_indegree([&](int2 ij) { return acc ? 1 + ij[1] : 1 + (ij[1] == 0 ? 0 : 1); })  // :449-455
        .set_priority([&](int2 ij) { return bl ? bl->trsm(ij[0], ij[1]) : prio_trsm(prio_kind, nb, ij[0], ij[1]); })
        .set_mapping([&](int2 ij) { return block_2_thread(ij[0], ij[1]); })
        .set_name([](int2 ij) { return "TRSM_" + std::to_string(ij[0]) + "_" + std::to_string(ij[1]); });

    gemm.set_task([&](int3 kij) {
            const int kind = kij[1] == kij[2] ? (acc ? TTAD_SYRK_ACC : TTAD_SYRK) : (acc ? TTAD_GEMM_ACC : TTAD_GEMM);
            return DeviceTask{kind, prec, 0, kij[1], kij[2], kij[0]};
        })
        .set_fulfill([&](int3 kij) {  // 2d_cholesky.cpp:502-531 (accumulate: through ACCU, :575-595)
            const int k = kij[0], i = kij[1], j = kij[2];
            if (!acc && k < j - 1) gemm.fulfill_promise({k + 1, i, j});
            else if (i == j) potrf.fulfill_promise(i);
            else trsm.fulfill_promise({i, j});
        })
        .set_indegree([&](int3 kij) {  // :532-540
            return (kij[1] == kij[2] ? 1 : 2) + (acc || kij[0] == 0 ? 0 : 1);
        })
        .set_priority([&](int3 kij) {
            return bl ? bl->gemm(kij[0], kij[1], kij[2]) : prio_gemm(prio_kind, nb, kij[0], kij[1], kij[2], acc);
        })
        .set_mapping([&](int3 kij) { return block_2_thread(kij[1], kij[2]); })
        .set_binding([](int3) { return false; })
        .set_name([](int3 kij) {
            return "GEMM_" + std::to_string(kij[0]) + "_" + std::to_string(kij[1]) + "_" + std::to_string(kij[2]);
        });

    potrf.fulfill_promise(0);  // 2d_cholesky.cpp:614
    const int rc = tp.compile();
    if (rc != TTAD_OK) return set_error(rc, "cholesky graph: %s", tp.error().c_str());
    if (out) *out = tp.release_graph();
    if (dry) *dry = tp_owned.release();
    return TTAD_OK;
}

// The bf16 hi/lo planes of M's off-diagonal tiles (FP32 storage): every tile (i, j), i > j, whose
// column count bsz[j] is a multiple of 32 gets bsz[i] * bsz[j] * 4 bytes in one slab; the pointer table
// is null elsewhere. Allocated once per matrix. Returns false (and allocates nothing) when the slab
// does not fit; the factorization then splits operands on the fly.
static bool ensure_planes(ttad_matrix* M) {
    if (M->tab_planes.p) return true;
    const int nt = M->nt;
    std::vector<size_t> off((size_t)nt * nt, SIZE_MAX);
    size_t total = 0;
    for (int j = 0; j < nt; j++)
        for (int i = j + 1; i < nt; i++)
            if (M->bsz[j] > 0 && M->bsz[j] % 32 == 0 && M->bsz[i] > 0) {
                off[i + (size_t)j * nt] = total;
                total += ((size_t)M->bsz[i] * M->bsz[j] * 4 + 255) / 256 * 256;
            }
    if (total == 0) return false;
    if (M->planes.alloc(total) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    std::vector<void*> tab((size_t)nt * nt, nullptr);
    for (size_t q = 0; q < tab.size(); q++)
        if (off[q] != SIZE_MAX) tab[q] = (char*)M->planes.p + off[q];
    if (M->tab_planes.alloc(tab.size() * sizeof(void*)) != hipSuccess ||
        hipMemcpy(M->tab_planes.p, tab.data(), tab.size() * sizeof(void*), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        M->tab_planes.release();
        M->planes.release();
        return false;
    }
    return true;
}

int cholesky_persistent(ttad_context* ctx, ttad_matrix* M, int prio_kind, int prec, hipStream_t s, int* info) {
    (void)info;
    const auto key = std::make_tuple(M->id, prio_kind, prec);
    ttad_graph* g = nullptr;
    auto it = ctx->chol_graphs.find(key);
    if (it == ctx->chol_graphs.end()) {
        const int rc = build_cholesky_graph(ctx, M->nt, M->storage, prec, prio_kind, &g);
        if (rc != TTAD_OK) return rc;
        ctx->chol_graphs[key] = g;
        // bf16x3 updates read their operands pre-split (written by the TRSM items of the same run,
        // which precede every update that reads them in the PTG); accumulate mode keeps the FP32 path
        if (prec == TTAD_BF16X3 && M->storage == TTAD_F32 && !(prio_kind & TTAD_ACCUMULATE) && planes_enabled() &&
            ensure_planes(M)) {
            g->with_planes = true;
            g->mats_dirty = true;
        }
    } else {
        g = it->second;
    }
    int rc = ttad_graph_bind(g, 0, M);
    if (rc != TTAD_OK) return rc;
    rc = ttad_graph_launch(g, 0, 0, s);
    if (rc != TTAD_OK) return rc;
    ttad_graph_stats st{};
    rc = ttad_graph_wait(g, 120.0, &st);
    if (rc == TTAD_OK) ctx->kernel_s = st.elapsed_s;
    return rc;
}

// The PTG compiled on the host (no device), for graphs that embed it (serving).
int cholesky_dag_compile(int nb, int prio_kind, std::vector<ttad_task_desc>& desc, std::vector<int32_t>& indeg,
                         std::vector<int32_t>& out_ptr, std::vector<int32_t>& out_idx) {
    Threadpool* tp = nullptr;
    const int rc = build_cholesky_graph(nullptr, nb, TTAD_F64, TTAD_F64, prio_kind, nullptr, &tp);
    if (rc != TTAD_OK) return rc;
    std::unique_ptr<Threadpool> hold(tp);
    desc = tp->task_descs();
    indeg = tp->indegrees();
    out_ptr = tp->out_ptr();
    out_idx = tp->out_idx();
    return TTAD_OK;
}

}  // namespace ttad

extern "C" int ttad_cholesky_dag(int num_blocks, int prio_kind, int* n_tasks, int* n_edges, ttad_task_desc* tasks,
                                 int32_t* indegree, double* priority, int32_t* out_ptr, int32_t* out_idx) {
    using namespace ttad;
    if (num_blocks <= 0 || num_blocks > 4096 || (prio_kind & ~TTAD_ACCUMULATE) < 0 ||
        (prio_kind & ~TTAD_ACCUMULATE) > 4 || !n_tasks || !n_edges)
        return set_error(TTAD_ERR_INVALID, "cholesky_dag: bad args");
    Threadpool* tp = nullptr;
    int rc = build_cholesky_graph(nullptr, num_blocks, TTAD_F64, TTAD_F64, prio_kind, nullptr, &tp);
    if (rc != TTAD_OK) return rc;
    std::unique_ptr<Threadpool> hold(tp);
    const int n = tp->n_tasks(), e = (int)tp->out_idx().size();
    const bool fill = tasks && indegree && out_ptr && (e == 0 || out_idx);
    if (fill) {
        if (*n_tasks < n || *n_edges < e) return set_error(TTAD_ERR_INVALID, "cholesky_dag: arrays too small");
        for (int t = 0; t < n; t++) {
            tasks[t] = tp->task_descs()[t];
            indegree[t] = tp->indegrees()[t];
            if (priority) priority[t] = tp->priorities()[t];
        }
        for (int t = 0; t <= n; t++) out_ptr[t] = tp->out_ptr()[t];
        for (int q = 0; q < e; q++) out_idx[q] = tp->out_idx()[q];
    }
    *n_tasks = n;
    *n_edges = e;
    return TTAD_OK;
}
