// irc_legacy.hip -- gfx950 kernels of the legacy all-sites drift mode (irc v4
// vcCluster='drift'): dc-count density and delta over one drift-chunk group with
// the primary/secondary site-match rule.
//   jrc3_cuda_rho_drift    matlab/irc_cuda_rho_drift.cu:28-108   (caller irc.m:24612-24637)
//   jrc3_cuda_delta_drift  matlab/irc_cuda_delta_drift.cu:28-132 (caller irc.m:24641-24666)
// Distances are the canonical FP64 fma chain; one wavefront per query, lanes over
// candidates, wave reductions instead of per-thread shared-memory slots.
#include "irc_common.h"
#include "irc_kernels.h"

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ double d2_rows(const float* __restrict__ a, const float* __restrict__ b, int D) {
    double acc = 0.0;
    for (int d = 0; d < D; ++d) {
        double t = (double)a[d] - (double)b[d];
        acc = fma(t, t, acc);
    }
    return acc;
}

__global__ __launch_bounds__(64) void k_rho_dc(int D, int n1, int n12, const int* __restrict__ miSite12,
                                               const float* __restrict__ fet1, const float* __restrict__ fet2,
                                               float dc2, float* __restrict__ rho_out) {
    const int lane = threadIdx.x;
    for (int i = blockIdx.x; i < n1; i += gridDim.x) {
        const int s = miSite12[i];
        const float* q = fet1 + (int64_t)i * D;
        long long ncmp = 0, ncnt = 0;
        for (int j = lane; j < n12; j += 64) {
            const float* x;
            if (miSite12[j] == s) x = fet1 + (int64_t)j * D;
            else if (miSite12[j + n12] == s) x = fet2 + (int64_t)j * D;
            else continue;
            const double t = d2_rows(q, x, D);
            ++ncmp;
            if (t <= (double)dc2) ++ncnt;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            ncmp += __shfl_xor(ncmp, off, 64);
            ncnt += __shfl_xor(ncnt, off, 64);
        }
        if (lane == 0) rho_out[i] = ncmp > 0 ? (float)((double)ncnt / (double)ncmp) : 0.0f;
    }
}

__global__ __launch_bounds__(64) void k_delta_dc(int D, int n1, int n12, const int* __restrict__ miSite12,
                                                 const float* __restrict__ fet1, const float* __restrict__ fet2,
                                                 const float* __restrict__ rho12, float dc2,
                                                 float* __restrict__ delta_out, unsigned* __restrict__ nneigh_out) {
    const int lane = threadIdx.x;
    const double init = (double)(float)IRC_SINGLE_INF_D;
    for (int i = blockIdx.x; i < n1; i += gridDim.x) {
        const int s = miSite12[i];
        const float ri = rho12[i];
        const float* q = fet1 + (int64_t)i * D;
        double bt = init;
        int bj = 0x7fffffff;
        for (int j = lane; j < n12; j += 64) {
            if (!(rho12[j] > ri)) continue;
            const float* x;
            if (miSite12[j] == s) x = fet1 + (int64_t)j * D;
            else if (miSite12[j + n12] == s) x = fet2 + (int64_t)j * D;
            else continue;
            const double t = d2_rows(q, x, D);
            if (t < bt || (t == bt && j < bj && t < init)) { bt = t; bj = j; }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            double ot = __shfl_xor(bt, off, 64);
            int oj = __shfl_xor(bj, off, 64);
            if (ot < bt || (ot == bt && oj < bj)) { bt = ot; bj = oj; }
        }
        if (lane == 0) {
            const bool none = !(bt < init);
            delta_out[i] = (float)sqrt(bt / (double)dc2);
            nneigh_out[i] = none ? (unsigned)(i + 1) : (unsigned)(bj + 1);
        }
    }
}

}  // namespace

namespace irc {

hipError_t launch_rho_dc(hipStream_t st, int D, int n1, int n12, const int* miSite12, const float* fet1,
                         const float* fet2, float dc2, float* rho_out, int grid) {
    if (n1 <= 0) return hipSuccess;
    k_rho_dc<<<grid < n1 ? grid : n1, 64, 0, st>>>(D, n1, n12, miSite12, fet1, fet2, dc2, rho_out);
    return hipGetLastError();
}

hipError_t launch_delta_dc(hipStream_t st, int D, int n1, int n12, const int* miSite12, const float* fet1,
                           const float* fet2, const float* rho12, float dc2, float* delta_out,
                           unsigned* nneigh_out, int grid) {
    if (n1 <= 0) return hipSuccess;
    k_delta_dc<<<grid < n1 ? grid : n1, 64, 0, st>>>(D, n1, n12, miSite12, fet1, fet2, rho12, dc2, delta_out,
                                                     nneigh_out);
    return hipGetLastError();
}

}  // namespace irc
