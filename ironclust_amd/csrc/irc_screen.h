// irc_screen.h -- MFMA distance screen for the kNN pass (gfx950).
#pragma once
#include "irc_common.h"

namespace irc {

// One screen work item: up to 64 queries of one (site, chunk) block against the
// block's whole candidate stream (the runs of the linked chunks).
struct ScreenItem {
    int blk;     // site * C + chunk
    int q0;      // first query index (queries of a block are contiguous)
    int nq;      // queries in this item (1..64)
    int ncand;   // candidates of the block (<= IRC_SCREEN_MAXCAND)
};

#define IRC_SCREEN_MAXCAND 65536  // candidate positions fit 16 bits of the packed key
#define IRC_SCREEN_MAXRUNS 128    // runs per block: 2 per linked chunk

// Packed operand rows are stored in 16-byte pieces (8 bf16 or 4 fp32 values), tiled by groups
// of 32 columns: [col / 32][piece][col % 32], so one wavefront load of piece p over 32
// consecutive columns reads 512 contiguous bytes.  np = pieces per row; the result is the
// piece's index in 16-byte units.  Buffers hold screen_cols_padded(n_cols) rows.
__host__ __device__ inline int64_t screen_piece(int64_t col, int piece, int np) {
    return (((col >> 5) * np + piece) << 5) | (col & 31);
}
inline int64_t screen_cols_padded(int64_t n_cols) { return (n_cols + 31) & ~(int64_t)31; }

struct ScreenArgs {
    DevPlan p;
    const void* hi;          // tiled rows: bf16 bits (bf16 form) or fp32 values (fp32 form)
    const void* lo;          // tiled rows: bf16 low parts (bf16 form)
    const float* norm;       // [n_cols] fp32 |x|^2
    const float* run_nmax;   // [n_blocks * 2] largest |x|^2 of each (block, part) run
    const ScreenItem* items;
    int n_items;
    int kprime;              // over-fetched candidates per query (32, 40, 48 or 64)
    int large_ncand;         // blocks with more candidates keep 64 (when kprime > 32)
    int hist;                // run the histogram pre-pass that sets each query's starting threshold
    double eta;              // |screened - exact| <= eta * (|q|^2 + |c|^2)
    // certified candidate lists, slot-major ([slot][n_q]) so a wavefront's 64 queries write
    // and the re-rank's lanes read contiguous words
    int64_t n_q;             // queries of the plan (row stride of the lists)
    int* cand_n;             // per query: r | (min(k, |C|) << 16), or -1 (not screened / uncertified)
    int* cand_col;           // [slots][n_q] candidate columns in ascending screened order
    float* cand_d;           // [slots][n_q] their truncated screened distances
    double* cand_eps;        // per query: the screen's error bound
    int* fallback;           // queries to redo with the exact kernel
    int* n_fallback;
};

enum { SCREEN_FP32 = 0, SCREEN_BF16 = 1 };

int screen_kp(int D);                                  // padded operand width (multiple of 16)
double screen_eta(int prec, int D, int KP);            // rigorous error factor
// feat (optional): also write the [n_cols][DP] zero-padded FP32 rows (F must have stride D)
hipError_t launch_screen_pack(hipStream_t st, int prec, int64_t n_cols, int D, int FS, int KP, const float* F,
                              void* hi, void* lo, float* norm, float* feat = nullptr,
                              int DP = 0);
// seg_ctr: 8 device counters (zeroed by the launch); seg_beg: 9 device offsets of the
// per-XCD item segments (items of segment x are dequeued first by workgroups on XCD x).
hipError_t launch_screen_knn_q(hipStream_t st, int prec, int KP, const ScreenArgs& a, int grid, int* seg_ctr,
                               const int* seg_beg);
// FP64 re-rank of the certified lists into the k-lists (knn_col, dk2) in two passes: a short
// window over every query, whose overflows (spill[0 .. *n_spill), n_q ints of room; the counter
// is zeroed by the launch) take a second pass with the long window; queries whose ambiguous
// prefix does not fit that either join the fallback list.
hipError_t launch_screen_rerank(hipStream_t st, const ScreenArgs& a, int* knn_col, double* dk2, int* spill,
                                int* n_spill);
hipError_t launch_fill_int(hipStream_t st, int* p, int64_t n, int v);
// largest |x|^2 of every (block, part) column run
hipError_t launch_run_nmax(hipStream_t st, int64_t n_runs, const int* run_beg, const int* run_len, const float* norm,
                           float* out);
hipError_t launch_mfma_layout_probe(hipStream_t st, float* out);

}  // namespace irc
