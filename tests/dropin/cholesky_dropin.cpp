// The dense Cholesky PTG of miniapp/dense_cholesky/2d_cholesky.cpp:281-616 (one rank, non-accumulate mode)
// written against the Taskflow/Threadpool drop-in (include/tasktorrent_amd/taskflow.hpp), as a caller of
// the reference would port it: the same set_indegree / set_fulfill / set_priority / set_mapping /
// set_binding / set_name callbacks, with device task descriptors as bodies.
//
//   cholesky_dropin dry NB PRIO        compile the graph without a device; print "SRC,DST" edges
//   cholesky_dropin binding NB         a bound TRSM task: print the status join/compile returns
//   cholesky_dropin gpu N B IN OUT     factor the n x n column-major FP64 matrix in file IN on GPU 0,
//                                      write L (column-major, strict upper zero) to OUT
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <tasktorrent_amd/taskflow.hpp>

using int2 = std::array<int, 2>;
using int3 = std::array<int, 3>;

static int run(ttad_context* ctx, ttad_matrix* M, int nb, int prio_kind, bool bind_trsm, bool print_edges) {
    ttad::Threadpool tp(ctx);
    if (M) tp.bind(0, M);
    ttad::Taskflow<int> potrf(&tp);
    ttad::Taskflow<int2> trsm(&tp);
    ttad::Taskflow<int3> gemm(&tp);
    auto block_2_thread = [&](int i, int j) { return (i + j * nb) % 4; };
    // 2d_cholesky.cpp:157-204, PrioKind cp (2) or none (0)
    auto p_potrf = [&](int j) { return prio_kind == 2 ? 9.0 * (nb - j) - 1 : 3.0; };
    auto p_trsm = [&](int2 ij) { return prio_kind == 2 ? 9.0 * (nb - ij[1]) - 2 : 2.0; };
    auto p_gemm = [&](int3 kij) { return prio_kind == 2 ? 9.0 * nb - 9.0 * kij[2] - 2 : 1.0; };

    potrf.set_task([&](int j) { return ttad::DeviceTask{TTAD_POTRF, TTAD_F64, 0, j, j, j}; })
        .set_fulfill([&](int j) {
            for (int i = j + 1; i < nb; i++) trsm.fulfill_promise({i, j});
        })
        .set_indegree([](int j) { return 1; })
        .set_priority(p_potrf)
        .set_mapping([&](int j) { return block_2_thread(j, j); })
        .set_binding([](int) { return false; })
        .set_name([](int j) { return "POTRF_" + std::to_string(j); });
    trsm.set_task([&](int2 ij) { return ttad::DeviceTask{TTAD_TRSM, TTAD_F64, 0, ij[0], ij[1], ij[1]}; })
        .set_fulfill([&](int2 ij) {
            const int i = ij[0], j = ij[1];
            for (int k = j + 1; k < i; k++) gemm.fulfill_promise({j, i, k});
            for (int k = i; k < nb; k++) gemm.fulfill_promise({j, k, i});
        })
        .set_indegree([](int2 ij) { return ij[1] == 0 ? 1 : 2; })
        .set_priority(p_trsm)
        .set_mapping([&](int2 ij) { return block_2_thread(ij[0], ij[1]); })
        .set_binding([&](int2) { return bind_trsm; })
        .set_name([](int2 ij) { return "TRSM_" + std::to_string(ij[0]) + "_" + std::to_string(ij[1]); });
    gemm.set_task([&](int3 kij) {
            return ttad::DeviceTask{kij[1] == kij[2] ? TTAD_SYRK : TTAD_GEMM, TTAD_F64, 0, kij[1], kij[2], kij[0]};
        })
        .set_fulfill([&](int3 kij) {
            const int k = kij[0], i = kij[1], j = kij[2];
            if (k < j - 1) gemm.fulfill_promise({k + 1, i, j});
            else if (i == j) potrf.fulfill_promise(i);
            else trsm.fulfill_promise({i, j});
        })
        .set_indegree([](int3 kij) { return (kij[1] == kij[2] ? 1 : 2) + (kij[0] == 0 ? 0 : 1); })
        .set_priority(p_gemm)
        .set_mapping([&](int3 kij) { return block_2_thread(kij[1], kij[2]); })
        .set_binding([](int3) { return false; })
        .set_name([](int3 kij) {
            return "GEMM_" + std::to_string(kij[0]) + "_" + std::to_string(kij[1]) + "_" + std::to_string(kij[2]);
        });

    potrf.fulfill_promise(0);
    if (!ctx) {
        const int rc = tp.compile();
        if (rc == TTAD_OK && print_edges)
            for (int t = 0; t < tp.n_tasks(); t++)
                for (int s : tp.successors(t)) printf("%s,%s\n", tp.task_name(t).c_str(), tp.task_name(s).c_str());
        if (rc != TTAD_OK) fprintf(stderr, "compile: %s\n", tp.error().c_str());
        return rc;
    }
    const int rc = tp.join(120.0);
    if (rc != TTAD_OK) fprintf(stderr, "join: %d %s %s\n", rc, tp.error().c_str(), ttad_last_error());
    return rc;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    if (mode == "dry") return run(nullptr, nullptr, atoi(argv[2]), argc > 3 ? atoi(argv[3]) : 2, false, true);
    if (mode == "binding") {
        printf("%d\n", run(nullptr, nullptr, atoi(argv[2]), 2, true, false));
        return 0;
    }
    if (mode == "gpu" && argc == 6) {
        const int n = atoi(argv[2]), b = atoi(argv[3]);
        std::vector<double> A((size_t)n * n);
        FILE* f = fopen(argv[4], "rb");
        if (!f || fread(A.data(), 8, A.size(), f) != A.size()) return 3;
        fclose(f);
        std::vector<int> sizes(n / b, b);
        if (n % b) sizes.push_back(n % b);
        const int nb = (int)sizes.size();
        ttad_context* ctx = nullptr;
        ttad_matrix* M = nullptr;
        if (ttad_init(0, &ctx) != TTAD_OK || ttad_matrix_create(ctx, nb, sizes.data(), TTAD_F64, &M) != TTAD_OK ||
            ttad_matrix_set(M, A.data(), n, 0, nullptr) != TTAD_OK) {
            fprintf(stderr, "setup: %s\n", ttad_last_error());
            return 4;
        }
        const int rc = run(ctx, M, nb, 2, false, false);
        if (rc == TTAD_OK && ttad_matrix_get(M, A.data(), n, 0, nullptr) != TTAD_OK) return 5;
        ttad_matrix_destroy(M);
        ttad_finalize(ctx);
        if (rc != TTAD_OK) return 6;
        f = fopen(argv[5], "wb");
        if (!f || fwrite(A.data(), 8, A.size(), f) != A.size()) return 7;
        fclose(f);
        return 0;
    }
    return 2;
}
