// irc_common.h -- shared device/host definitions of the gfx950 density-peak library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define IRC_WAVE 64
#define IRC_DMAX 64          // largest padded feature dimension handled by the kernels
#define IRC_KMAX 64          // largest k handled by the kernels
#define IRC_D2HEAD 4         // list entries whose exact d^2 the re-rank keeps for the delta pass

// The drop-in value reported for a spike with no denser neighbour:
// single(sqrt(3.402e38)) (matlab/irc2.m:3686, 3723-3725).
#define IRC_SINGLE_INF_D 3.402E+38

// Work description shared by the kernels.  Columns are the site-grouped feature
// columns; a block is one (site, drift chunk) pair and owns up to two contiguous
// column runs (primary-site spikes, secondary-site spikes).
struct DevPlan {
    const float* feat;      // [n_cols][DP] fp32, zero padded beyond D
    int DP;                 // padded dimension (multiple of 4)
    int D;
    int C;                  // drift chunks
    int k;
    const int* col_key;     // tie-break key per column (spike id or column index)
    const int* col_ret;     // value reported for a column (1-based spike id or column)
    int n_blocks;           // (site, chunk) blocks
    const int* run_beg;     // [n_blocks*2]
    const int* run_len;     // [n_blocks*2]
    const int* link_ptr;    // [C+1]
    const int* link_idx;    // chunks r with link(r, c), ascending
    int nq;                 // queries of the plan; the kNN lists are slot-major, [k][nq]
    const int* q_col;       // [nq] query column
    const int* q_blk;       // [nq] block of the query (site*C + chunk)
    const int* q_out;       // [nq] output slot
};

__device__ __forceinline__ bool irc_less(double t1, int k1, double t2, int k2) {
    return (t1 < t2) || (t1 == t2 && k1 < k2);
}
