// irc_assign.h -- density-peak label propagation along nneigh chains (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace irc {

// lab[a] (in/out): cluster of a (> 0) or 0 (unassigned); nxt: scratch of n ints.  On return every
// spike whose nneigh chain reaches a labelled spike carries that spike's label (the first labelled
// spike on the chain a, nneigh(a), nneigh(nneigh(a)), ...).  nneigh is 1-based.
hipError_t launch_assign_chain(hipStream_t st, int64_t n, const int64_t* nneigh, int32_t* lab, int32_t* nxt);

// counts[(c - 1) * (n_clu + 1) + d] += #{(s, j): clu[s] = c, clu[knn(j, s)] = d} for clusters c >= 1,
// with d = 0 for neighbours outside every cluster; knn is k x n column-major, 1-based.  counts must
// be zeroed by the caller.
hipError_t launch_isolation_counts(hipStream_t st, int64_t n, int k, const int64_t* knn, const int32_t* clu,
                                   int n_clu, unsigned long long* counts);

}  // namespace irc
