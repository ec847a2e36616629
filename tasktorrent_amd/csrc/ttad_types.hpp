// Shared host/device definitions of the tile DAG executor.
#pragma once
#include <stdint.h>

namespace ttad {

// Task kinds of the tiled Cholesky PTG (miniapp/dense_cholesky/2d_cholesky.cpp:316-608).
enum Kind : int32_t { K_POTRF = 0, K_TRSM = 1, K_SYRK = 2, K_GEMM = 3, K_ACCU = 4, K_NOP = 5 };
// Arithmetic of a task. Tiles are stored in FP64 (P_F64) or FP32 (the mixed variants).
enum Prec : int32_t { P_F64 = 0, P_F32 = 1, P_BF16 = 2, P_BF16X3 = 3 };

constexpr int NTHREADS = 256;  // one workgroup = 4 wave64
constexpr int GEMM_BM = 128;   // GEMM/SYRK work item = 128x128 output sub-tile
constexpr int TRSM_BM = 32;      // TRSM work item of an FP64 tile: 32-row block (its latency is on the critical path)
constexpr int TRSM_BM_F32 = 64;  // of an FP32 tile: 64-row block (half the items: fewer re-reads of the diagonal tile)
__host__ __device__ constexpr int trsm_rows(int elem_bytes) { return elem_bytes == 8 ? TRSM_BM : TRSM_BM_F32; }
constexpr int DB = 64;         // diagonal sub-block size whose inverse POTRF leaves for TRSM

// One tiled symmetric matrix in HBM. Tile (i,j), i >= j, is a column-major
// bsz[i] x bsz[j] array with leading dimension bsz[i] (block_sizes_lda,
// 2d_cholesky.cpp:86-89), reached through the pointer table.
struct DevMat {
    int32_t nt;
    int32_t storage;  // P_F64 or P_F32
    const int32_t* bsz;
    const int32_t* boff;
    void* const* tiles;    // [nt*nt], entry i + j*nt
    double* const* dinv;   // [nt]: inverses of the 64x64 diagonal blocks of L_jj (FP64, 64x64 each, ld 64)
    int32_t* info;         // first non-positive pivot, 1-based global column; 0 = none
    int32_t* pflags;       // [nt * 128]: progress flags of the row-block POTRF items of each diagonal tile (zero between runs)
    int32_t* locks;        // [nt * nt * 16]: accumulate-mode sub-tile locks of tile (i, j) (zero when free)
    // [nt*nt] or null: bf16 hi/lo planes of the off-diagonal FP32 tiles (i > j, bsz[j] % 32 == 0; other
    // entries null), written by the TRSM items that produce the tile and read by the bf16x3 GEMM items.
    // Tile (i, j) with m = bsz[i] rows: 16-byte granule ((c*2 + p)*4 + g)*m + r holds bf16 plane p
    // (0: hi = bf16(x), 1: lo = bf16(x - hi), both rounded to nearest even) of columns 32c + 8g ... +7 of row r.
    void* const* planes;
};
// bf16 split planes: granule of (chunk of 32 columns, plane, 8-column group, row) in a tile of m rows.
__host__ __device__ inline long plane_granule(int c, int p, int g, int r, int m) {
    return ((long)(c * 2 + p) * 4 + g) * m + r;
}

// Device task record.
struct DevTask {
    int32_t kind;
    int32_t prec;
    int32_t mat;
    int32_t n_items;
    int32_t i, j, k;
    int32_t bucket;
    int32_t crit;  // on a longest path of the DAG: runs with no other item started on its CU
    int32_t accum; // GEMM/SYRK: add the product into C under the sub-tile lock (accumulate mode)
};

// Work items per task of tiles with elem_bytes-byte elements; host and device must agree.
__host__ __device__ inline int items_of(int kind, int bi, int bj, int elem_bytes) {
    if (kind == K_POTRF) return 1;
    if (kind == K_TRSM) return (bi + trsm_rows(elem_bytes) - 1) / trsm_rows(elem_bytes);
    const int tr = (bi + GEMM_BM - 1) / GEMM_BM;
    if (kind == K_SYRK) return tr * (tr + 1) / 2;
    if (kind == K_GEMM) return tr * ((bj + GEMM_BM - 1) / GEMM_BM);
    return 1;
}

}  // namespace ttad
