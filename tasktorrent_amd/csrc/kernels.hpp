// Host-side launch wrappers of kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "ttad_types.hpp"

namespace ttad {

struct DevGraph;

hipError_t kernels_init();
// Resident persistent workers per CU of an executor instantiation (index into the variant sets).
int persistent_blocks_per_cu(int vset);
// vset: index into the compiled item-variant sets (vset_index of the variants a graph needs).
void launch_items(int vset, const DevMat* mats, const DevTask* tasks, const int32_t* prefix, int n_tasks,
                  int n_items, int32_t* err, hipStream_t s);
void launch_graph_reset(const DevGraph& g, int n_slots, hipStream_t s);
// Writes the device wall clock (wall_clock64) to *out.
void launch_read_clock(uint64_t* out, hipStream_t s);
void launch_persistent(int vset, const DevGraph& g, int n_workers, hipStream_t s);
void launch_raw_gemm(int prec, const void* A, long lda, const void* B, long ldb, void* C, long ldc, int m, int n,
                     int K, int syrk, double alpha, double beta, hipStream_t s);
void launch_raw_potrf(int prec, void* A, long lda, int n, double* dinv, int* info, hipStream_t s);
void launch_trtri_diag(int prec, const void* L, long ldl, int n, double* dinv, hipStream_t s);
void launch_raw_trsm(int prec, const void* L, long ldl, const double* dinv, void* X, long ldx, int m, int n,
                     hipStream_t s);
void launch_pack(int storage, const double* src, long lda, const DevMat* M, int nt, hipStream_t s);
void launch_unpack(int storage, double* dst, long ldd, const DevMat* M, int nt, int zero_upper, hipStream_t s);
void launch_copy_tiles(int dst_storage, int src_storage, const DevMat* S, const DevMat* D, int nt, hipStream_t s);

}  // namespace ttad
