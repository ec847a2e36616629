/*
 * tasktorrent_amd: C-ABI of an MI355X-native (gfx950) tiled dense Cholesky.
 *
 * This is the drop-in boundary for the hot path of the reference's
 * miniapp/dense_cholesky/2d_cholesky.cpp (tile tasks POTRF/TRSM/SYRK/GEMM
 * seeded with indegree/fulfill_promise and run on a Threadpool). Every entry
 * point below names the reference interface it replaces (file:line, paths
 * relative to the reference repository root).
 *
 * Conventions: plain pointers and sizes; matrices column-major; int status
 * codes (TTAD_OK = 0); ttad_last_error() gives a message for the calling
 * thread. `stream` is a hipStream_t passed as void* (NULL = the context's own
 * stream). Device pointers are HBM addresses on the context's device.
 * One host thread per context at a time.
 */
#ifndef TASKTORRENT_AMD_H
#define TASKTORRENT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define TTAD_OK 0
#define TTAD_ERR_INVALID 1     /* bad argument */
#define TTAD_ERR_OOM 2         /* device allocation failed or over the HBM budget */
#define TTAD_ERR_NOT_SPD 3     /* non-positive pivot; *info = 1-based global column */
#define TTAD_ERR_TIMEOUT 4     /* device watchdog or host wait timed out */
#define TTAD_ERR_DEVICE 5      /* HIP runtime error */
#define TTAD_ERR_IR_STALLED 6  /* iterative refinement did not reach the target */
#define TTAD_ERR_GRAPH 7       /* inconsistent task graph (indegree vs edges) */
#define TTAD_ERR_BUSY 8        /* resource in use (graph running, serving queue full) */
#define TTAD_ERR_NOT_FOUND 9   /* unknown handle/id */

/* ---- arithmetic / storage ----------------------------------------------- */
#define TTAD_F64 0   /* FP64 tiles, FP64 MFMA */
#define TTAD_F32 1   /* FP32 tiles, FP32 MFMA */
#define TTAD_BF16 2  /* FP32 tiles, BF16 MFMA inputs with FP32 accumulation */
#define TTAD_BF16X3 3 /* FP32 tiles, operands split hi+lo into two BF16s, 3 MFMAs, FP32 accumulation */

/* ---- task kinds of the Cholesky PTG (2d_cholesky.cpp:316-608) ------------ */
#define TTAD_POTRF 0
#define TTAD_TRSM 1
#define TTAD_SYRK 2
#define TTAD_GEMM 3
/* No tile access: the workgroup spins for k microseconds (scheduler-overhead benchmark,
 * replaces the spin_for bodies of miniapp/mini_benchmarks/shared/ttor_cholesky.cpp:22-28);
 * mat is ignored and no matrix needs to be bound; i only places the task's queue (as for GEMM). */
#define TTAD_SPIN 5
/* GEMM / SYRK in accumulate mode (2d_cholesky.cpp --accumulate, :149-155, 479-608): the product
 * is formed from zero and added into the C tile under a lock per 128x128 sub-tile, in the
 * order the tasks finish (the reference's ACCU task, fused into the product). */
#define TTAD_GEMM_ACC 6
#define TTAD_SYRK_ACC 7
/* OR into prio_kind of ttad_potrf_tiled / ttad_cholesky_dag: the accumulate-mode DAG (GEMM and
 * SYRK tasks of one tile run in any order; POTRF(j) waits for j of them, TRSM(i,j) for 1 + j). */
#define TTAD_ACCUMULATE 0x100

/* ---- schedules of ttad_potrf_tiled ------------------------------------------ */
#define TTAD_SCHED_STREAM 0      /* one launch per panel step (bulk-synchronous) */
#define TTAD_SCHED_PERSISTENT 1  /* device-resident DAG with dependency counters */

typedef struct ttad_context ttad_context;
typedef struct ttad_matrix ttad_matrix;
typedef struct ttad_graph ttad_graph;

/* ---- library / context ---------------------------------------------------- */
const char* ttad_last_error(void);
int ttad_version(void);
/* Replaces the Threadpool construction (threadpool_shared.cpp:17-33): binds a
 * device, creates its stream and scratch. */
int ttad_init(int device, ttad_context** out);
/* Matrices, graphs and servers created on ctx must be destroyed before it: their
 * handles refer to it. */
int ttad_finalize(ttad_context* ctx);
int ttad_device_info(ttad_context* ctx, int* n_cu, long long* hbm_bytes, char* name, int name_len);
int ttad_synchronize(ttad_context* ctx, void* stream);
/* Context options (no reference counterpart).
 * TTAD_OPT_FACTOR_RESIDUAL: value != 0 makes ttad_solve_ir and ttad_posv_mixed also compute
 *   ||L L^T - A||_F / ||A||_F of the factor they used (ttad_factor_residual, O(n^3) work done after
 *   the solution is complete and not counted in its times) into ttad_solve_stats.factor_residual;
 *   0 (default) leaves that field at -1.
 * TTAD_OPT_RESIDUAL_ROWS: value != 0 makes every FP64 residual (ttad_residual and the refinement
 *   steps) use the row-block kernels, which read each stored block of A twice; 0 (default) uses the
 *   tile-pair kernels, which read A once, wherever the tiles have at most 512 rows and their
 *   partial sums are smaller than half of A. Both give r = b - A x to rounding. */
#define TTAD_OPT_FACTOR_RESIDUAL 1
#define TTAD_OPT_RESIDUAL_ROWS 2
int ttad_set_option(ttad_context* ctx, int option, long long value);

/* ---- tile kernels: BLAS/LAPACK-shaped, device pointers, column-major ----------- */
/* LAPACKE_dpotrf(LAPACK_COL_MAJOR, 'L', n, A, lda)  (2d_cholesky.cpp:319).
 * *info as LAPACK (0 ok, k > 0 non-positive pivot at column k). Synchronous. */
int ttad_dpotrf(ttad_context* ctx, int n, double* A, int lda, int* info, void* stream);
/* cblas_dtrsm(ColMajor, Right, Lower, Trans, NonUnit, m, n, 1.0, L, ldl, B, ldb)
 * (2d_cholesky.cpp:404-405): B := B L^{-T}. */
int ttad_dtrsm(ttad_context* ctx, int m, int n, const double* L, int ldl, double* B, int ldb, void* stream);
/* ttad_dtrsm without any allocation: no counterpart in the reference. For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops). The inverses of L's 64 x 64 diagonal blocks go to a
 * static per-GPU scratch of 2 MiB, so n <= 4096 (TTAD_ERR_INVALID above that); m is unbounded. Asynchronous on
 * `stream`; calls that overlap on one GPU share that scratch and race. */
#define TTAD_DTRSM_NOALLOC_MAX_N 4096
int ttad_dtrsm_noalloc(ttad_context* ctx, int m, int n, const double* L, int ldl, double* B, int ldb,
                       void* stream);
/* cblas_dsyrk(ColMajor, Lower, NoTrans, n, k, alpha, A, lda, beta, C, ldc)
 * (2d_cholesky.cpp:490-491): lower triangle of C := beta C + alpha A A^T. */
int ttad_dsyrk(ttad_context* ctx, int n, int k, double alpha, const double* A, int lda, double beta, double* C,
               int ldc, void* stream);
/* cblas_dgemm(ColMajor, NoTrans, Trans, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
 * (2d_cholesky.cpp:493-494): C := beta C + alpha A B^T. */
int ttad_dgemm(ttad_context* ctx, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
               int ldb, double beta, double* C, int ldc, void* stream);
/* FP32 versions of the same four calls (mixed-precision variants; no reference
 * counterpart, the reference is FP64 only). ttad_spotrf factors in FP64. */
int ttad_spotrf(ttad_context* ctx, int n, float* A, int lda, int* info, void* stream);
int ttad_strsm(ttad_context* ctx, int m, int n, const float* L, int ldl, float* B, int ldb, void* stream);
int ttad_ssyrk(ttad_context* ctx, int n, int k, float alpha, const float* A, int lda, float beta, float* C, int ldc,
               void* stream);
int ttad_sgemm(ttad_context* ctx, int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
               float beta, float* C, int ldc, void* stream);
/* C := beta C + alpha bf16(A) bf16(B)^T with FP32 storage and accumulation
 * (BF16 trailing-update variant; syrk != 0 updates the lower triangle only). */
int ttad_gemm_bf16(ttad_context* ctx, int m, int n, int k, float alpha, const float* A, int lda, const float* B,
                   int ldb, float beta, float* C, int ldc, int syrk, void* stream);
/* Same with each operand split hi+lo into two BF16s and three MFMAs per product
 * (about 16 significant bits per operand; BF16X3 trailing-update variant). */
int ttad_gemm_bf16x3(ttad_context* ctx, int m, int n, int k, float alpha, const float* A, int lda, const float* B,
                     int ldb, float beta, float* C, int ldc, int syrk, void* stream);

/* ---- tiled symmetric matrices in HBM ------------------------------------------ */
/* Replaces the per-tile storage `blocks[ii + jj*num_blocks]` of 2d_cholesky.cpp:134-147
 * with lda = block size (block_sizes_lda, :86-89). Only tiles i >= j exist.
 * storage: TTAD_F64 or TTAD_F32. */
int ttad_matrix_create(ttad_context* ctx, int num_blocks, const int* block_sizes, int storage, ttad_matrix** out);
int ttad_matrix_destroy(ttad_matrix* M);
/* Copy the lower triangle of a full n x n column-major FP64 array in (set) or out
 * (get). *_device != 0: the array is in HBM, else in host memory. get writes the
 * strict upper triangle as zeros. */
int ttad_matrix_set(ttad_matrix* M, const double* A, long long lda, int src_device, void* stream);
int ttad_matrix_get(ttad_matrix* M, double* A, long long lda, int dst_device, void* stream);
/* Tile-wise copy between matrices of the same partition, converting storage. */
int ttad_matrix_copy(ttad_matrix* dst, const ttad_matrix* src, void* stream);
int ttad_matrix_info(const ttad_matrix* M, int* n, int* num_blocks, int* storage, long long* bytes);
int ttad_matrix_tile(const ttad_matrix* M, int i, int j, void** ptr, int* ld);

/* ---- the tiled Cholesky --------------------------------------------------- */
/* Replaces cholesky() of 2d_cholesky.cpp:41-690 on one rank, non-accumulate mode:
 * the whole factorization of M in place (lower tiles become L).
 * prec: TTAD_F64 (M FP64) or TTAD_F32 / TTAD_BF16 / TTAD_BF16X3 for the trailing update (M FP32;
 * diagonal POTRF in FP64). prio_kind 0..3 as PrioKind (:39, :157-204); 4 (no counterpart in
 * the reference): longest remaining path of each task in the DAG ("bottom level").
 * Synchronous; *info as ttad_dpotrf. */
int ttad_potrf_tiled(ttad_context* ctx, ttad_matrix* M, int schedule, int prio_kind, int prec, void* stream,
                     int* info);

/* Device time (HIP events on the factorization's stream) of the last
 * ttad_potrf_tiled on this context: the persistent executor's launch for
 * TTAD_SCHED_PERSISTENT, all launches for TTAD_SCHED_STREAM; replaces the t0/t1
 * wall-clock window of 2d_cholesky.cpp:611-619. */
int ttad_last_elapsed(ttad_context* ctx, double* seconds);

/* ---- device task graphs: the Taskflow<K>/Threadpool drop-in ------------------- */
typedef struct {
    int32_t kind;  /* TTAD_POTRF .. TTAD_GEMM, or TTAD_SPIN */
    int32_t prec;  /* TTAD_F64 / TTAD_F32 / TTAD_BF16 / TTAD_BF16X3 */
    int32_t mat;   /* matrix slot bound with ttad_graph_bind */
    int32_t i, j, k;
} ttad_task_desc;

typedef struct {
    double elapsed_s;       /* launch -> completion, device events */
    long long items_done;
    long long items_total;
    int workers;            /* persistent workgroups */
    int abort_code;         /* 0 ok, 1 watchdog, 2 queue overflow, 3 item variant missing from the executor */
} ttad_graph_stats;

/* A graph is the compiled form of the Taskflow callbacks (taskflow.hpp:40-297):
 * for every task its kind/coordinates (set_task), indegree (set_indegree),
 * out-edges in CSR form (the fulfill_promise calls of set_fulfill) and priority
 * (set_priority, higher first). Validated: every edge target must receive exactly
 * indegree fulfills, counting host seeds. */
int ttad_graph_create(ttad_context* ctx, int n_tasks, const ttad_task_desc* tasks, const int32_t* indegree,
                      const int32_t* out_ptr, const int32_t* out_idx, const double* priority, ttad_graph** out);
int ttad_graph_bind(ttad_graph* g, int slot, ttad_matrix* M);
/* Taskflow::fulfill_promise(k) called from the host (2d_cholesky.cpp:614). */
int ttad_graph_seed(ttad_graph* g, int task);
int ttad_graph_clear_seeds(ttad_graph* g);
/* Starts the persistent executor (threadpool_shared.cpp:81-198). n_workers <= 0:
 * all resident slots. trace != 0 records per-item device timestamps. */
int ttad_graph_launch(ttad_graph* g, int n_workers, int trace, void* stream);
/* Threadpool::join (threadpool_shared.cpp:234-241): waits for completion,
 * timeout_s <= 0 waits forever. Returns NOT_SPD if a bound matrix hit a bad pivot. */
int ttad_graph_wait(ttad_graph* g, double timeout_s, ttad_graph_stats* stats);
/* Writes the last run's trace in the reference Logger CSV format (util.cpp:117-127, events of
 * threadpool_shared.cpp:36-57): per task one "<prefix><worker>>run>NAME,start,end" row (first work
 * item's start to last item's end, worker = the workgroup of its first item) and one
 * "<prefix><worker>>deps>NAME,start,end" row (the release of its successors), times in seconds
 * since the epoch; NAME as in ttad_graph_write_deps (POTRF_j, TRSM_i_j, GEMM_k_i_j). */
int ttad_graph_write_trace(ttad_graph* g, const char* path, const char* prefix);
/* Writes the graph's edges in the reference DepsLogger CSV format (util.cpp:103-112):
 * one "SRC_NAME,DST_NAME" line per fulfill_promise edge. No run needed. */
int ttad_graph_write_deps(ttad_graph* g, const char* path);
int ttad_graph_destroy(ttad_graph* g);

/* ---- solve phase ------------------------------------------------------------- */
typedef struct {
    int iterations;          /* refinement steps taken (all attempts) */
    int fell_back;           /* 1 if the FP64 factorization was used after the low-precision one failed */
    double backward_error;   /* max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) */
    double factor_s;         /* host wall time of the (first) factorization, seconds */
    double solve_s;          /* host wall time of everything after it (refinement, fallback) */
    double total_s;          /* time to solution */
    int info;                /* factorization info of the factor used */
    int prec_used;           /* TTAD_F64 / TTAD_F32 / TTAD_BF16 of the factor used */
    double factor_residual;  /* ||L L^T - A||_F / ||A||_F of the factor used, with TTAD_OPT_FACTOR_RESIDUAL;
                                else -1 */
} ttad_solve_stats;

/* X = (L L^T)^{-1} B with a factor left by ttad_potrf_tiled (FP64 or FP32 tiles; an
 * FP32 factor solves in FP32). B, X: device FP64 n x nrhs, column-major. Tile sizes
 * must be multiples of 64 except the last. Replaces the check solve
 * L.solveInPlace(b); L.transpose().solveInPlace(b) of 2d_cholesky.cpp:676-681. */
int ttad_solve(ttad_context* ctx, ttad_matrix* L, int nrhs, const double* B, long long ldb, double* X, long long ldx,
               void* stream);
/* R = B - A X in FP64 (A: FP64 tiles, symmetric from its lower triangle) and the
 * normwise backward error max_j ||r_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
 * (the b = A*x of 2d_cholesky.cpp:677 and its error check :682-687). */
int ttad_residual(ttad_context* ctx, ttad_matrix* A, int nrhs, const double* B, long long ldb, const double* X,
                  long long ldx, double* R, long long ldr, double* eta, void* stream);
/* ||L L^T - A||_F / ||A||_F over the full symmetric matrices, from the lower tiles of A (FP64) and
 * of its factor L (FP64 or FP32 tiles, same partition; the strict upper triangle of L's diagonal
 * tiles is ignored). FP64 products on the device, partial sums added in a fixed order.
 * Replaces the factor check of the reference's tests (tests/shared/tests.cpp:568-583: tiled
 * L against the monolithic factor) and complements the solve check of 2d_cholesky.cpp:646-689.
 * Synchronous. */
int ttad_factor_residual(ttad_context* ctx, ttad_matrix* A, ttad_matrix* L, double* rel, void* stream);
/* Iterative refinement in FP64 with the factor L of A (any precision): X from 0,
 * r = B - A X, X += (L L^T)^{-1} r until eta <= tol or max_iter steps or a stall
 * (eta not halving twice). TTAD_ERR_IR_STALLED if tol was not reached.
 * No reference counterpart (the reference is FP64 only). */
int ttad_solve_ir(ttad_context* ctx, ttad_matrix* A, ttad_matrix* L, int nrhs, const double* B, long long ldb,
                  double* X, long long ldx, double tol, int max_iter, ttad_solve_stats* stats, void* stream);
/* Mixed-precision solver: factors a working copy of A with the trailing update in
 * prec (TTAD_F32 / TTAD_BF16 / TTAD_BF16X3; FP64 diagonal POTRF), refines in FP64 to eta <= tol,
 * and falls back to the FP64 factorization when the low-precision factor is not
 * positive definite or refinement stalls. The factor stays cached on the device
 * for further right-hand sides of the same A. No reference counterpart. */
int ttad_posv_mixed(ttad_context* ctx, ttad_matrix* A, int prec, int nrhs, const double* B, long long ldb, double* X,
                    long long ldx, double tol, int max_iter, ttad_solve_stats* stats, void* stream);

/* ---- serving loop: device-resident continuous batching of SPD solves ---------------- */
/* No reference counterpart (the reference runs one factorization per process). A
 * server owns a persistent kernel that runs until ttad_server_destroy; requests posted
 * with ttad_server_submit are admitted while it runs, as their HBM footprint (factor +
 * vectors, from an arena of hbm_budget bytes), task-pool range and request slot become
 * free, and their task graphs share the same workgroups. */
typedef struct ttad_server ttad_server;
typedef struct {
    long long hbm_budget;  /* arena bytes for factors and vectors (admission limit) */
    int max_tasks;         /* task pool size */
    int max_requests;      /* concurrent request slots */
    int workers;           /* persistent workgroups; <= 0: all resident slots */
    int idle_timeout_s;    /* the device loop exits after this long without work (default 600) */
} ttad_server_config;
typedef struct {
    long long id;
    int status;             /* TTAD_OK, TTAD_ERR_IR_STALLED, TTAD_ERR_NOT_SPD, or an admission error. A request
                               whose low-precision factor is not positive definite or whose refinement stalls
                               (or reaches max_iter) is re-run with an FP64 factor before it is reported; the
                               status is that of the FP64 attempt */
    int n, nrhs;
    int iterations;         /* refinement steps, all attempts */
    int reused;             /* 1 if the request used a cached factor */
    long long factor_id;    /* id of the factor used (pass as reuse_factor to solve again) */
    double backward_error;  /* normwise backward error of the returned solution */
    double latency_s;       /* submit -> solution copied back to the caller's buffer */
    double queue_s;         /* submit -> first admission */
    int fell_back;          /* 1 if the FP64 factorization was used after the low-precision one failed */
    int info;               /* factorization info of the factor used (1-based column of a non-positive pivot) */
} ttad_server_result;
typedef struct {
    int pending, active, cached_factors, workers;
    long long arena_used, arena_size, items_run;
} ttad_server_stats;

int ttad_server_create(ttad_context* ctx, const ttad_server_config* cfg, ttad_server** out);
/* A: FP64 tiled matrix (tile sizes multiples of 64), resident and unchanged until the
 * request completes. B: host n x nrhs column-major, read at admission (again if the request
 * falls back to FP64), so it stays valid until ttad_server_poll reports the request; X: host
 * buffer that receives the solution when ttad_server_poll reports the request.
 * prec: TTAD_F64 / TTAD_F32 / TTAD_BF16 / TTAD_BF16X3 factorization; reuse_factor >= 0
 * solves with a factor kept by an earlier request (keep_factor != 0) instead of
 * factoring. Never blocks: requests that do not fit yet wait in a host queue. */
int ttad_server_submit(ttad_server* S, ttad_matrix* A, int prec, int nrhs, const double* B, double* X, double tol,
                       int max_iter, long long reuse_factor, int keep_factor, long long* id);
/* Collects completed requests (up to max_out) and admits waiting ones. */
int ttad_server_poll(ttad_server* S, ttad_server_result* out, int max_out, int* n_out);
int ttad_server_query(ttad_server* S, ttad_server_stats* stats);
/* Stops the device loop after the requests in flight complete, frees everything. */
int ttad_server_destroy(ttad_server* S);

/* The Cholesky PTG of 2d_cholesky.cpp:281-614 (one rank, non-accumulate mode),
 * declared through the Taskflow drop-in and compiled on the host without a
 * device: tasks (POTRF/TRSM/SYRK/GEMM with tile coordinates, k = step), their
 * indegrees, priorities (PrioKind prio_kind) and out-edges in CSR form.
 * Call once with NULL arrays to get *n_tasks / *n_edges, then with arrays of
 * at least that size. */
int ttad_cholesky_dag(int num_blocks, int prio_kind, int* n_tasks, int* n_edges, ttad_task_desc* tasks,
                      int32_t* indegree, double* priority, int32_t* out_ptr, int32_t* out_idx);

#ifdef __cplusplus
}
#endif
#endif /* TASKTORRENT_AMD_H */
