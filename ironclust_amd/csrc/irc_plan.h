// irc_plan.h -- device-side work planning of the batched kNN pass.
#pragma once
#include "irc_common.h"
#include "irc_screen.h"

namespace irc {

// Per-block planning arrays (carved from one device buffer).
struct PlanBuf {
    int64_t* totals;      // [0] items, [1] exact queries, [2] cost
    int64_t* n_pairs;     // sum over queries of |C(a)|
    int64_t* cost_off;    // [n_blocks] exclusive prefix of the block costs
    int64_t* blk_cost;    // [n_blocks] items * (candidates + 64)
    int* blk_q0;          // [n_blocks] first query of the block
    int* blk_ncand;       // [n_blocks] candidates of the block
    int* blk_nitems;      // [n_blocks] screen items
    int* blk_nexact;      // [n_blocks] queries sent to the exact kernel
    int* item_off;        // [n_blocks] exclusive prefix of blk_nitems
    int* exact_off;       // [n_blocks] exclusive prefix of blk_nexact
    int64_t* tile_sum;    // [3 * tiles] scan scratch
};

size_t plan_buf_bytes(int64_t n_blocks);
PlanBuf plan_buf_carve(void* base, int64_t n_blocks);
// Build the screen items, the exact-query list (its length goes to *n_fallback, where the
// screen appends its uncertified queries) and the 9 per-XCD segment offsets.  items needs room
// for n_queries / 64 + n_blocks entries, exact_q for n_queries.
hipError_t launch_plan(hipStream_t st, int S, int C, const DevPlan& p, const int64_t* site_ptr, const int64_t* prim,
                       int min_item, int max_cand, int max_runs, const PlanBuf& pb, ScreenItem* items, int* exact_q,
                       int* seg_beg, int* n_fallback);

// device counters of a call (read once after the call's last kernel): *dst = s32 ? *s32 : *s64
hipError_t launch_stat_copy(hipStream_t st, const int* s32, const int64_t* s64, int64_t* dst);
hipError_t launch_zero_i64(hipStream_t st, int64_t* p, int n);
// *count += number of 32-bit words that differ between a and b (n_words each); *count starts at 0
hipError_t launch_count_diff(hipStream_t st, const uint32_t* a, const uint32_t* b, int64_t n_words,
                             unsigned long long* count);
// largest candidate count of a block holding queries (atomicMax into *out, which starts at 0)
hipError_t launch_max_ncand(hipStream_t st, int S, int C, const int* link_ptr, const int* link_idx, const int* run_len,
                            int* out);

}  // namespace irc
