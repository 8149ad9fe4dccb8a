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

enum { SCREEN_FP32 = 0, SCREEN_BF16 = 1, SCREEN_FP8 = 2 };

// Packed operand rows: every column's candidate-form row [x/s, w, |x|^2/(s^2 w), 0...] is stored in
// 16-byte pieces (4 fp32, 8 bf16 or 16 fp8 values), NP pieces per column:
//   FP32  KP = 16 KB fp32 values                      -> NP = 4 KB
//   BF16  KP bf16 high parts and KP bf16 low parts    -> NP = 4 KB
//   FP8   64 e4m3 values in three terms (x = h + m + l, each rounded to e4m3) -> NP = 12
// (s = 1, w = 1 for FP32/BF16; for FP8 s is the site's power-of-two scale and w = 256).
// A 32x32 MFMA takes from lane (m, h) the pieces of lane half h only, so the pieces are stored
// by lane half: for a group of 32 columns, [h][j][col % 32], j < NP / 2 in the order the MFMAs use
// them; one load instruction over 32 consecutive columns reads 512 contiguous bytes per half, and a
// lane's NP / 2 pieces of one column sit 512 bytes apart (one address, immediate offsets).
//   FP32  piece j of half h: dimensions 8 j + 4 h .. + 3
//   BF16  piece j = 2 kb + t of half h: dimensions 16 kb + 8 h .. + 7, t = 0 high / 1 low parts
//   FP8   piece j = 2 t + p of half h: dimensions 32 h + 16 p .. + 15 of term t (h, m, l)
__host__ __device__ inline int screen_np(int prec, int KB) { return prec == SCREEN_FP8 ? 12 : 4 * KB; }
// index (in 16-byte units) of piece 0 of lane half h of column col; piece j is 32 further on per j
__host__ __device__ inline int64_t screen_off(int64_t col, int h, int np) {
    return (((col >> 5) * np + h * (np >> 1)) << 5) | (col & 31);
}
inline int64_t screen_cols_padded(int64_t n_cols) { return (n_cols + 31) & ~(int64_t)31; }

struct ScreenArgs {
    DevPlan p;
    const void* rows;        // packed operand rows (layout above)
    const float* norm;       // [n_cols] fp32 |x|^2
    const float* run_nmax;   // [n_blocks * 2] largest |x|^2 of each (block, part) run
    const float* site_scale; // FP8: power-of-two scale s of each site (block / C), else unused
    const ScreenItem* items;
    ScreenItem* items_kind;  // [n_items] the items reordered by the launch: per XCD segment, those of
                             // more than 32 queries, then those of at most 32
    int n_items;
    int kprime;              // over-fetched candidates per query (32, 40, 48 or 64)
    int large_ncand;         // blocks with more candidates keep 64 (when kprime > 32)
    // |screened - exact| <= eta (|q|^2 + |c|^2) + eta2 s (|q| + |c|) + eta3 s^2
    double eta, eta2, eta3;
    // certified candidate lists, slot-major ([slot][n_q]) so a wavefront's 64 queries write
    // and the re-rank's lanes read contiguous pairs
    int64_t n_q;             // queries of the plan (row stride of the lists)
    int* cand_n;             // per query: r | (min(k, |C|) << 16), or -1 (not screened / uncertified)
    int2* cand;              // [slots][n_q] (candidate column, bits of its truncated screened
                             // distance) in ascending screened order
    double* cand_eps;        // per query: the screen's error bound
    int* fallback;           // queries to redo with the exact kernel
    int* n_fallback;
    // batch form: the re-rank also finishes the queries it completes (rho = float(1/sqrt(d2_k)),
    // d_k, and the int64 spike ids of the list in slot order, row q_out * k); null: not fused
    float* fin_rho;
    float* fin_dk;
    int64_t* fin_knn;
    // and the exact d^2 of the first IRC_D2HEAD list entries ([slot][n_q]; NaN for queries it does
    // not finish), which the delta pass uses instead of recomputing them
    double* fin_d2head;
    float* fin_rho_col;      // rho of every query's own column (the delta pass's per-column rho)
};

int screen_kp(int D);                                  // padded operand width of the FP32/BF16 forms
// FP8 needs one 64-wide K block: D + 2 <= 64
inline bool screen_fits(int prec, int D) { return prec == SCREEN_FP8 ? D + 2 <= 64 : screen_kp(D) <= 64; }
// rigorous error factors of each form (eta2, eta3 are 0 except for FP8)
void screen_eta(int prec, int D, int KP, double* eta, double* eta2, double* eta3);
// The pack pass over F (stride FS): operand rows, |x|^2 per column, and optionally the [n_cols][DP]
// zero-padded FP32 rows (feat; F must have stride D).  FP8: column c of segment s (seg_ptr[s] <= c <
// seg_ptr[s + 1]) is scaled by seg_scale[s].
hipError_t launch_screen_pack(hipStream_t st, int prec, int64_t n_cols, int D, int FS, int KB, const float* F,
                              void* rows, float* norm, float* feat, int DP, const int64_t* seg_ptr, int n_seg,
                              const float* seg_scale);
// FP8: seg_scale[s] = 2^e, the smallest power of two with max |F| over segment s's columns <= 32 * 2^e
hipError_t launch_site_scale(hipStream_t st, int n_seg, const int64_t* seg_ptr, const float* F, int FS, int D,
                             float* seg_scale);
// seg_ctr: 25 ints of device scratch (counters and the split offsets; set by the launch);
// seg_beg: 9 device offsets of the per-XCD item segments (items of segment x are dequeued first
// by workgroups on XCD x).  Items of at most 32 queries run in a kernel of their own.
hipError_t launch_screen_knn_q(hipStream_t st, int prec, int KB, const ScreenArgs& a, int grid, int* seg_ctr,
                               const int* seg_beg);
// FP64 re-rank of the certified lists into the k-lists (knn_col, dk2) in two passes: a short
// window over every query, whose overflows (spill[0 .. *n_spill), n_q ints of room; the counter
// is zeroed by the launch) take a second pass with the long window; queries whose ambiguous
// prefix does not fit that either join the fallback list.
hipError_t launch_screen_rerank(hipStream_t st, const ScreenArgs& a, int* knn_col, double* dk2, int* spill,
                                int* n_spill);
// the re-rank finishes its queries itself (fin_* set) only for lists of at most 32
inline bool screen_rerank_finishes(const ScreenArgs& a) {
    return a.p.k <= 32 && (a.fin_rho || a.fin_knn || a.fin_d2head || a.fin_rho_col);
}
hipError_t launch_fill_int(hipStream_t st, int* p, int64_t n, int v);
// largest |x|^2 of every (block, part) column run
hipError_t launch_run_nmax(hipStream_t st, int64_t n_runs, const int* run_beg, const int* run_len, const float* norm,
                           float* out);
hipError_t launch_mfma_layout_probe(hipStream_t st, float* out);
hipError_t launch_mfma_accum_probe(hipStream_t st, int prec, int K, int n_tiles, const float* A, const float* B,
                                   float* out);

}  // namespace irc
