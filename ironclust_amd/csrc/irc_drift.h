// irc_drift.h -- drift linkage kernels (calc_drift_).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace irc {

// counts: [C][n_amp*n_pos] int32 (bin iAmp + n_amp*iPos); d2: [C][C] int64 scratch;
// link: [C][C] uint8, link[c*C + r] = mlDrift(r, c)
hipError_t launch_drift(hipStream_t st, int64_t N, int C, const int64_t* lim, const float* amp, const float* pos,
                        const float* amp_edges, int n_amp, const float* pos_edges, int n_pos, int n_batch,
                        int n_time_max, int* counts, int64_t* d2, uint8_t* link);

}  // namespace irc
