// irc_plan.hip -- device-side work planning of the batched kNN pass (gfx950).
//
// The batch form replaces sort_page_'s loops over pages, sites and query chunks
// (matlab/irc2.m:3456-3503, 3539-3561; search_knn_drift_'s per-chunk launches 3768-3787) by
// one work list built on the device from the (site, chunk) column runs:
//   block b = s*C + c owns the queries of its primary run and streams the runs of the
//   chunks linked to c (mlDrift(:, c), matlab/irc2.m:3664, 3770);
//   it becomes ceil(nq/64) screen items of up to 64 queries, or goes whole to the exact
//   kernel when its stream is longer than the screen's keys can address;
//   the items are split into 8 contiguous segments of about equal cost, one per XCD.
// Everything stays on the device, so a batch call needs no host round trip between its
// kernels (and can be captured in a HIP graph).
#include <hip/hip_runtime.h>

#include "irc_common.h"
#include "irc_plan.h"

namespace irc {

namespace {

// one thread per block: query range, candidate count, items and exact queries of the block
__global__ void k_block_plan(int S, int C, const int* __restrict__ link_ptr, const int* __restrict__ link_idx,
                             const int* __restrict__ run_beg, const int* __restrict__ run_len,
                             const int64_t* __restrict__ site_ptr, const int64_t* __restrict__ prim,
                             int min_item, int max_cand, int max_runs, PlanBuf pb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= S * C) return;
    const int s = b / C, c = b - s * C;
    const int nq = run_len[2 * b];
    int64_t ncand = 0;
    int nruns = 0;
    for (int li = link_ptr[c]; li < link_ptr[c + 1]; ++li) {
        const int b2 = s * C + link_idx[li];
        for (int part = 0; part < 2; ++part) {
            const int len = run_len[2 * b2 + part];
            ncand += len;
            nruns += len > 0;
        }
    }
    const int q0 = (int)(prim[s] + (run_beg[2 * b] - site_ptr[s]));
    int nitems = 0, nexact = 0;
    if (nq > 0) {
        if (ncand > max_cand || nruns > max_runs) {
            nexact = nq;
        } else {
            nitems = nq / 64;
            const int tail = nq - 64 * nitems;
            if (tail > 0) {
                if (tail >= min_item) ++nitems; else nexact = tail;
            }
        }
    }
    pb.blk_q0[b] = q0;
    pb.blk_ncand[b] = (int)(ncand < 0x7fffffff ? ncand : 0x7fffffff);
    pb.blk_nitems[b] = nitems;
    pb.blk_nexact[b] = nexact;
    pb.blk_cost[b] = (int64_t)nitems * (ncand + 64);
    if (nq > 0) atomicAdd(reinterpret_cast<unsigned long long*>(pb.n_pairs), (unsigned long long)nq * ncand);
}

// Exclusive scans of the per-block item, exact-query and cost counts in three coalesced
// steps: per-tile sums, a scan of the tile sums, per-tile scans plus the tile offset.
constexpr int SCAN_T = 1024;

// inclusive scan of v over the workgroup (Hillis-Steele in LDS); returns the inclusive value
__device__ __forceinline__ int64_t block_scan_incl(int64_t v, int64_t* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        const int64_t add = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const int64_t r = sh[t];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(int n, PlanBuf pb, int64_t* __restrict__ tile_sum) {
    __shared__ int64_t sh[SCAN_T];
    const int i = blockIdx.x * SCAN_T + threadIdx.x;
    const bool in = i < n;
    const int64_t a = block_scan_incl(in ? pb.blk_nitems[i] : 0, sh);
    const int64_t e = block_scan_incl(in ? pb.blk_nexact[i] : 0, sh);
    const int64_t w = block_scan_incl(in ? pb.blk_cost[i] : 0, sh);
    if (threadIdx.x == SCAN_T - 1) {
        tile_sum[3 * blockIdx.x + 0] = a;
        tile_sum[3 * blockIdx.x + 1] = e;
        tile_sum[3 * blockIdx.x + 2] = w;
    }
}

__global__ void k_scan_top(int ntiles, int64_t* __restrict__ tile_sum, PlanBuf pb) {
    const int t = threadIdx.x;
    if (t >= 3) return;
    int64_t run = 0;
    for (int i = 0; i < ntiles; ++i) {
        const int64_t v = tile_sum[3 * i + t];
        tile_sum[3 * i + t] = run;
        run += v;
    }
    pb.totals[t] = run;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_apply(int n, PlanBuf pb, const int64_t* __restrict__ tile_sum) {
    __shared__ int64_t sh[SCAN_T];
    const int i = blockIdx.x * SCAN_T + threadIdx.x;
    const bool in = i < n;
    const int64_t va = in ? pb.blk_nitems[i] : 0, ve = in ? pb.blk_nexact[i] : 0, vw = in ? pb.blk_cost[i] : 0;
    const int64_t a = block_scan_incl(va, sh) - va + tile_sum[3 * blockIdx.x + 0];
    const int64_t e = block_scan_incl(ve, sh) - ve + tile_sum[3 * blockIdx.x + 1];
    const int64_t w = block_scan_incl(vw, sh) - vw + tile_sum[3 * blockIdx.x + 2];
    if (in) {
        pb.item_off[i] = (int)a;
        pb.exact_off[i] = (int)e;
        pb.cost_off[i] = w;
    }
}

// one thread per block: write its items (groups of 64 queries, the last one the remainder)
// and its exact queries (a short remainder, or the whole block when it is not screened)
__global__ void k_block_items(int nb, const int* __restrict__ run_len, PlanBuf pb, ScreenItem* __restrict__ items,
                              int* __restrict__ exact_q) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int nq = run_len[2 * b];
    const int ni = pb.blk_nitems[b], ne = pb.blk_nexact[b];
    const int q0 = pb.blk_q0[b], ncand = pb.blk_ncand[b];
    const int io = pb.item_off[b], eo = pb.exact_off[b];
    const int nscreen = nq - ne;
    for (int i = 0; i < ni; ++i) {
        ScreenItem it;
        it.blk = b;
        it.q0 = q0 + 64 * i;
        it.nq = nscreen - 64 * i < 64 ? nscreen - 64 * i : 64;
        it.ncand = ncand;
        items[io + i] = it;
    }
    for (int j = 0; j < ne; ++j) exact_q[eo + j] = q0 + nscreen + j;
}

__global__ void k_zero_i64(int64_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// per-XCD segment bounds (thread j finds bound j by binary search over the cost prefix), the
// fallback counter's start and the item total
__global__ void k_plan_segments(int nb, PlanBuf pb, int* __restrict__ seg_beg, int* __restrict__ n_fallback) {
    const int j = threadIdx.x;
    const int64_t n_items = pb.totals[0], total = pb.totals[2];
    if (j == 0) {
        seg_beg[0] = 0;
        seg_beg[8] = (int)n_items;
        *n_fallback = (int)pb.totals[1];
    }
    if (j < 1 || j > 7) return;
    const int64_t target = total * j / 8;
    // last block whose cost range starts at or before the target
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pb.cost_off[mid] <= target) lo = mid; else hi = mid - 1;
    }
    const int64_t per = (int64_t)pb.blk_ncand[lo] + 64;
    int64_t i = pb.blk_nitems[lo] > 0 ? (target - pb.cost_off[lo] + per - 1) / per : 0;
    if (i > pb.blk_nitems[lo]) i = pb.blk_nitems[lo];
    int64_t v = pb.item_off[lo] + i;
    if (v > n_items) v = n_items;
    seg_beg[j] = (int)v;
}

}  // namespace

// largest candidate count of any block (thread per block), into *out with atomicMax
__global__ void k_max_ncand(int S, int C, const int* __restrict__ link_ptr, const int* __restrict__ link_idx,
                            const int* __restrict__ run_len, int* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= S * C || run_len[2 * b] == 0) return;
    const int s = b / C, c = b - s * C;
    int64_t n = 0;
    for (int li = link_ptr[c]; li < link_ptr[c + 1]; ++li) {
        const int b2 = s * C + link_idx[li];
        n += run_len[2 * b2] + run_len[2 * b2 + 1];
    }
    atomicMax(out, (int)(n < 0x7fffffff ? n : 0x7fffffff));
}

hipError_t launch_max_ncand(hipStream_t st, int S, int C, const int* link_ptr, const int* link_idx, const int* run_len,
                            int* out) {
    const int nb = S * C;
    k_max_ncand<<<(nb + 255) / 256, 256, 0, st>>>(S, C, link_ptr, link_idx, run_len, out);
    return hipGetLastError();
}

__global__ void k_stat_copy(const int* __restrict__ s32, const int64_t* __restrict__ s64, int64_t* __restrict__ dst) {
    if (threadIdx.x == 0) *dst = s32 ? (int64_t)*s32 : *s64;
}

hipError_t launch_stat_copy(hipStream_t st, const int* s32, const int64_t* s64, int64_t* dst) {
    k_stat_copy<<<1, 64, 0, st>>>(s32, s64, dst);
    return hipGetLastError();
}

namespace {
__global__ void k_count_diff(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, int64_t n,
                             unsigned long long* __restrict__ count) {
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += a[i] != b[i];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}
}  // namespace

hipError_t launch_count_diff(hipStream_t st, const uint32_t* a, const uint32_t* b, int64_t n_words,
                             unsigned long long* count) {
    if (n_words <= 0) return hipSuccess;
    const int64_t blocks = (n_words + 255) / 256;
    k_count_diff<<<(unsigned)(blocks < 4096 ? blocks : 4096), 256, 0, st>>>(a, b, n_words, count);
    return hipGetLastError();
}

hipError_t launch_zero_i64(hipStream_t st, int64_t* p, int n) {
    k_zero_i64<<<(n + 63) / 64, 64, 0, st>>>(p, n);
    return hipGetLastError();
}

size_t plan_buf_bytes(int64_t n_blocks) {
    const int64_t ntiles = (n_blocks + SCAN_T - 1) / SCAN_T;
    return (size_t)n_blocks * (6 * sizeof(int) + 2 * sizeof(int64_t)) + 8 * sizeof(int64_t) +
           (size_t)ntiles * 3 * sizeof(int64_t) + 256;
}

PlanBuf plan_buf_carve(void* base, int64_t n_blocks) {
    char* p = reinterpret_cast<char*>(base);
    PlanBuf pb;
    pb.totals = reinterpret_cast<int64_t*>(p);
    pb.n_pairs = pb.totals + 4;
    p += 8 * sizeof(int64_t);
    pb.cost_off = reinterpret_cast<int64_t*>(p);
    p += n_blocks * sizeof(int64_t);
    pb.blk_cost = reinterpret_cast<int64_t*>(p);
    p += n_blocks * sizeof(int64_t);
    pb.blk_q0 = reinterpret_cast<int*>(p);
    p += n_blocks * sizeof(int);
    pb.blk_ncand = reinterpret_cast<int*>(p);
    p += n_blocks * sizeof(int);
    pb.blk_nitems = reinterpret_cast<int*>(p);
    p += n_blocks * sizeof(int);
    pb.blk_nexact = reinterpret_cast<int*>(p);
    p += n_blocks * sizeof(int);
    pb.item_off = reinterpret_cast<int*>(p);
    p += n_blocks * sizeof(int);
    pb.exact_off = reinterpret_cast<int*>(p);
    p += n_blocks * sizeof(int);
    p = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(p) + 7) & ~(uintptr_t)7);
    pb.tile_sum = reinterpret_cast<int64_t*>(p);
    return pb;
}

hipError_t launch_plan(hipStream_t st, int S, int C, const DevPlan& p, const int64_t* site_ptr, const int64_t* prim,
                       int min_item, int max_cand, int max_runs, const PlanBuf& pb, ScreenItem* items, int* exact_q,
                       int* seg_beg, int* n_fallback) {
    const int nb = S * C;
    k_zero_i64<<<1, 64, 0, st>>>(pb.totals, 8);
    k_block_plan<<<(nb + 255) / 256, 256, 0, st>>>(S, C, p.link_ptr, p.link_idx, p.run_beg, p.run_len, site_ptr,
                                                   prim, min_item, max_cand, max_runs, pb);
    const int ntiles = (nb + SCAN_T - 1) / SCAN_T;
    k_scan_tiles<<<ntiles, SCAN_T, 0, st>>>(nb, pb, pb.tile_sum);
    k_scan_top<<<1, 64, 0, st>>>(ntiles, pb.tile_sum, pb);
    k_scan_apply<<<ntiles, SCAN_T, 0, st>>>(nb, pb, pb.tile_sum);
    k_block_items<<<(nb + 255) / 256, 256, 0, st>>>(nb, p.run_len, pb, items, exact_q);
    k_plan_segments<<<1, 64, 0, st>>>(nb, pb, seg_beg, n_fallback);
    return hipGetLastError();
}

}  // namespace irc
