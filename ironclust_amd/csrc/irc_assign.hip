// irc_assign.hip -- cluster labels along the nearest-denser chains (assignCluster_knn_,
// matlab/irc2.m:6640-6694, its final loop).
//
// The reference visits spikes in descending rho and copies the label of nneigh, which is always
// denser and therefore already visited: every spike ends with the label of the first labelled
// spike on its chain.  Here that is pointer jumping: each round, an unlabelled spike either takes
// the label at its current jump target or doubles its jump, so ceil(log2 n) + 1 rounds resolve
// every chain (chains are acyclic: rho increases strictly along nneigh, and the only fixed points
// are peaks).  Reads of other spikes' entries during a round see either the old or the new value,
// both of which lie on the same chain, so the result does not depend on the interleaving.
#include <hip/hip_runtime.h>

#include "irc_assign.h"

namespace irc {

namespace {

__global__ void k_chain_init(int64_t n, const int64_t* __restrict__ nneigh, const int32_t* __restrict__ lab,
                             int32_t* __restrict__ nxt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j = nneigh[i] - 1;
    nxt[i] = (lab[i] > 0 || j < 0 || j >= n) ? (int32_t)i : (int32_t)j;
}

__global__ void k_chain_jump(int64_t n, int32_t* lab, int32_t* nxt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (lab[i] > 0) return;
    const int32_t j = nxt[i];
    const int32_t l = lab[j];
    if (l > 0) {
        lab[i] = l;
    } else {
        nxt[i] = nxt[j];
    }
}

// one thread per spike: its own-label neighbours are added with one atomic, the others one each
// (well-separated clusters have few)
__global__ void k_isolation_counts(int64_t n, int k, const int64_t* __restrict__ knn, const int32_t* __restrict__ clu,
                                   int n_clu, unsigned long long* __restrict__ counts) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int c = clu[s];
    if (c < 1 || c > n_clu) return;
    unsigned long long* row = counts + (int64_t)(c - 1) * (n_clu + 1);
    unsigned long long own = 0;
    for (int j = 0; j < k; ++j) {
        const int64_t t = knn[s * k + j] - 1;
        int d = (t >= 0 && t < n) ? clu[t] : 0;
        if (d < 1 || d > n_clu) d = 0;
        if (d == c) ++own;
        else atomicAdd(&row[d], 1ull);
    }
    if (own) atomicAdd(&row[c], own);
}

}  // namespace

hipError_t launch_isolation_counts(hipStream_t st, int64_t n, int k, const int64_t* knn, const int32_t* clu,
                                   int n_clu, unsigned long long* counts) {
    if (n <= 0) return hipSuccess;
    k_isolation_counts<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, k, knn, clu, n_clu, counts);
    return hipGetLastError();
}

hipError_t launch_assign_chain(hipStream_t st, int64_t n, const int64_t* nneigh, int32_t* lab, int32_t* nxt) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + 255) / 256);
    k_chain_init<<<grid, 256, 0, st>>>(n, nneigh, lab, nxt);
    int rounds = 1;
    while (((int64_t)1 << (rounds - 1)) < n) ++rounds;
    for (int r = 0; r <= rounds; ++r) k_chain_jump<<<grid, 256, 0, st>>>(n, lab, nxt);
    return hipGetLastError();
}

}  // namespace irc
