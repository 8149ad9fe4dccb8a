// irc_kernels.h -- host-callable launchers of the gfx950 kernels.
#pragma once
#include "irc_common.h"

namespace irc {

hipError_t launch_pack_fp32(hipStream_t st, int64_t n_cols, int D, int DP, const float* F, float* feat);
hipError_t launch_block_runs(hipStream_t st, int S, int C, const int64_t* site_ptr, const int64_t* site_np,
                             const int64_t* lim, const int* col_spike, int* run_beg, int* run_len);
hipError_t launch_batch_queries(hipStream_t st, int64_t N, int S, int C, const int64_t* site_ptr,
                                const int64_t* site_np, const int64_t* prim_ptr, const int64_t* lim,
                                const int* col_spike, int64_t n_cols, int* q_col, int* q_blk, int* q_out,
                                int* flags);
size_t knn_exact_lds_bytes(int k, int DP);
hipError_t launch_knn_exact(hipStream_t st, const DevPlan& p, const int* qlist, int nq, int grid,
                            int* knn_col, double* dk2, int* n_cand);
hipError_t launch_knn_exact_dev(hipStream_t st, const DevPlan& p, const int* qlist, const int* nq_dev, int nq_max,
                                int grid, int* knn_col, double* dk2);
hipError_t launch_knn_finish(hipStream_t st, const DevPlan& p, int nq, int mode, const int* knn_col,
                             const double* dk2, float* rho_slot, float* dk_slot, int* knn32,
                             int64_t* knn64 = nullptr);
// mode 0 of launch_knn_finish for the queries list[0 .. *n_list) (at most nq_max)
hipError_t launch_knn_finish_list(hipStream_t st, const DevPlan& p, const int* list, const int* n_list, int nq_max,
                                  const int* knn_col, const double* dk2, float* rho_slot, float* dk_slot,
                                  int64_t* knn64, float* rho_col = nullptr);
// out[q * k + j] = float(sqrt(d2)) of the j-th listed neighbour of query q (canonical FP64 distance)
hipError_t launch_knn_dists(hipStream_t st, const DevPlan& p, int nq, const int* knn_col, float* out);
hipError_t launch_rho_cols(hipStream_t st, int64_t n_cols, const int* col_key, const float* rho_slot,
                           float* rho_col);
// d2head (optional): exact d^2 of the first IRC_D2HEAD entries of every list ([slot][nq], NaN unknown)
hipError_t launch_delta_from_knn(hipStream_t st, const DevPlan& p, int nq, const int* knn_col,
                                 const float* rho_col, float* delta_slot,
                                 int64_t* nneigh_slot, int* unres, int* n_unres, const double* d2head = nullptr);
hipError_t launch_delta_exact(hipStream_t st, const DevPlan& p, const int* qlist, int nq, const int* nq_dev,
                              int grid, int mode, const float* rho_col, float* delta_slot, int64_t* nneigh64,
                              int* nneigh32);

// legacy dc-density variants
hipError_t launch_rho_dc(hipStream_t st, int D, int n1, int n12, const int* miSite12, const float* fet1,
                         const float* fet2, float dc2, float* rho_out, int grid);
hipError_t launch_delta_dc(hipStream_t st, int D, int n1, int n12, const int* miSite12, const float* fet1,
                           const float* fet2, const float* rho12, float dc2, float* delta_out,
                           unsigned* nneigh_out, int grid);

}  // namespace irc
