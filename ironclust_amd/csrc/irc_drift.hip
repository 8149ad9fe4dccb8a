// irc_drift.hip -- drift linkage (calc_drift_, matlab/irc2.m:6373-6447) on gfx950.
//
// Time is cut into C equal-count chunks; every chunk gets a 2-D histogram of its spikes over
// amplitude quantile bins x position quantile bins; chunk c is linked to the n_batch chunks
// whose histograms are nearest (Euclidean), restricted to a band of chunks for long recordings
// (dist_mask_ml_, matlab/irc2.m:6452-6462).  Counts are integers, so squared distances are
// exact and the ranking is by (distance^2, chunk index) -- the order of MATLAB's stable sort.
//   k_drift_hist    one workgroup per chunk, histogram in LDS (HBM-bound: 8 bytes per spike)
//   k_drift_d2      one thread per (row, column) pair of the mask, exact int64 distance
//   k_drift_select  one workgroup per column, rank of every masked row, links the first n_batch
#include <hip/hip_runtime.h>

#include "irc_drift.h"

namespace irc {

namespace {

// bin of v among sorted edges e[0..n] with MATLAB/numpy semantics: [e_i, e_i+1), the last bin
// closed; -1 outside
__device__ __forceinline__ int edge_bin(float v, const float* e, int n) {
    if (!(v >= e[0] && v <= e[n])) return -1;
    if (v == e[n]) {
        // the closed last bin holds v == e[n]; with repeated top edges it is the last bin whose
        // left edge is below or at v
        int b = n - 1;
        return b;
    }
    int lo = 0, hi = n - 1;   // largest i with e[i] <= v
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (e[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_drift_hist(int64_t N, int C, const int64_t* __restrict__ lim, const float* __restrict__ amp,
                             const float* __restrict__ pos, const float* __restrict__ amp_edges, int n_amp,
                             const float* __restrict__ pos_edges, int n_pos, int* __restrict__ counts) {
    extern __shared__ int sh[];
    float* ea = reinterpret_cast<float*>(sh);
    float* ep = ea + (n_amp + 1);
    int* h = reinterpret_cast<int*>(ep + (n_pos + 1));
    const int nb = n_amp * n_pos;
    const int c = blockIdx.x;
    for (int i = threadIdx.x; i <= n_amp; i += blockDim.x) ea[i] = amp_edges[i];
    for (int i = threadIdx.x; i <= n_pos; i += blockDim.x) ep[i] = pos_edges[i];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int64_t b = lim[c] - 1, e = lim[c + 1] - 1;
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const int ia = edge_bin(amp[i], ea, n_amp);
        const int ip = edge_bin(pos[i], ep, n_pos);
        if (ia >= 0 && ip >= 0) atomicAdd(&h[ia + n_amp * ip], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) counts[(int64_t)c * nb + i] = h[i];
}

// mask(r, c): every row without a band; else |r - c| <= q plus the corner blocks of calc_drift_
__device__ __forceinline__ bool drift_mask(int r, int c, int C, int q, bool band) {
    if (!band) return true;
    if (abs(r - c) <= q) return true;
    if (r < 2 * q + 1 && c < q) return true;
    if (r >= C - 2 * q - 1 && c >= C - q) return true;
    return false;
}

__global__ void k_drift_d2(int C, int nb, const int* __restrict__ counts, int q, int band,
                           int64_t* __restrict__ d2) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)C * C) return;
    const int c = (int)(e / C), r = (int)(e - (int64_t)c * C);
    if (!drift_mask(r, c, C, q, band != 0)) return;
    const int* a = counts + (int64_t)r * nb;
    const int* b = counts + (int64_t)c * nb;
    int64_t s = 0;
    for (int i = 0; i < nb; ++i) {
        const int64_t d = (int64_t)a[i] - b[i];
        s += d * d;
    }
    d2[e] = s;
}

__global__ void k_drift_select(int C, int q, int band, int n_batch, const int64_t* __restrict__ d2,
                               uint8_t* __restrict__ link) {
    const int c = blockIdx.x;
    for (int r = threadIdx.x; r < C; r += blockDim.x) {
        uint8_t v = 0;
        if (drift_mask(r, c, C, q, band != 0)) {
            const int64_t dr = d2[(int64_t)c * C + r];
            int rank = 0;
            for (int r2 = 0; r2 < C && rank < n_batch; ++r2) {
                if (!drift_mask(r2, c, C, q, band != 0)) continue;
                const int64_t d = d2[(int64_t)c * C + r2];
                rank += (d < dr || (d == dr && r2 < r)) ? 1 : 0;
            }
            v = rank < n_batch ? 1 : 0;
        }
        link[(int64_t)c * C + r] = v;
    }
}

}  // namespace

hipError_t launch_drift(hipStream_t st, int64_t N, int C, const int64_t* lim, const float* amp, const float* pos,
                        const float* amp_edges, int n_amp, const float* pos_edges, int n_pos, int n_batch,
                        int n_time_max, int* counts, int64_t* d2, uint8_t* link) {
    const int nb = n_amp * n_pos;
    const size_t lds = sizeof(float) * (n_amp + 1 + n_pos + 1) + sizeof(int) * (size_t)nb;
    k_drift_hist<<<C, 256, lds, st>>>(N, C, lim, amp, pos, amp_edges, n_amp, pos_edges, n_pos, counts);
    const int band = C > n_time_max ? 1 : 0;
    const int q = n_time_max / 4;
    const int64_t pairs = (int64_t)C * C;
    k_drift_d2<<<(unsigned)((pairs + 255) / 256), 256, 0, st>>>(C, nb, counts, q, band, d2);
    k_drift_select<<<C, 256, 0, st>>>(C, q, band, n_batch, d2, link);
    return hipGetLastError();
}

}  // namespace irc
