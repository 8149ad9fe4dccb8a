/*
 * irc_hip.h -- C ABI of the MI355X (gfx950) density-peak hot path of IronClust:
 * drift-constrained per-site kNN, rho = 1/d_k, and delta/nneigh (nearest denser
 * spike within the linked drift chunks).
 *
 * The entry points replace the parallel.gpu.CUDAKernel/PTX launches of the
 * reference (paths relative to the ironclust repository):
 *
 *   irc_knn_drift       search_knn_drift_   matlab/irc2.m:3731-3789  (kernel search_min_drift.cu:24-94)
 *   irc_delta_drift     search_delta_drift_ matlab/irc2.m:3794-3865  (kernel search_delta_drift.cu:24-106)
 *   irc_rho_delta_batch sort_page_          matlab/irc2.m:3456-3503  (rho_page_ 3524, delta_page_ 3566,
 *                                            rho_paged_site_ 3635, delta_paged_site_ 3684), all sites and
 *                                            drift chunks in one call
 *   irc_knn_ab          cuda_knn_           matlab/irc2.m:4202-4238, irc.m:24365-24401
 *                                            (kernels cuda_knn_index.cu:27-111, cuda_knn.cu)
 *   irc_knn_ab_full     cuda_knn_full       matlab/cuda_knn_full.cu:27-100 (all k distances)
 *   irc_knn_ab_ib       cuda_knn_ib         matlab/cuda_knn_ib.cu:27-97 (indirect candidate list)
 *   irc_delta_knn_ab    cuda_delta_knn_     irc.m:24519-24553        (kernel cuda_delta_knn.cu:19-88)
 *   irc_rho_drift_dc    cuda_rho_drift_     irc.m:24612-24637        (kernel irc_cuda_rho_drift.cu:28-108)
 *   irc_delta_drift_dc  cuda_delta_drift_   irc.m:24641-24666        (kernel irc_cuda_delta_drift.cu:28-132)
 *   irc_calc_drift      calc_drift_         matlab/irc2.m:6373-6447  (drift linkage, the path's input)
 *   irc_assign_chain    assignCluster_knn_  matlab/irc2.m:6681-6692  (cluster labels along nneigh, the
 *                                            path's consumer)
 *   irc_isolation_counts isolation_score_   matlab/irc2.m:6699-6731  (kNN-graph merges, knn_overlap_merge_)
 *
 * Results are exact to the canonical definition (FP64 distances accumulated with
 * fused multiply-add in dimension order; ties broken by lower spike id), whatever
 * precision option is chosen: a reduced-precision screen only selects candidates,
 * and every reported distance, index, rho and delta comes from the FP64 re-rank.
 *
 * Conventions: features are D x n single precision, column-major (MATLAB layout),
 * i.e. n consecutive groups of D floats.  Indices returned to the caller are 1-based.
 * All calls are synchronous.  No C++ exception crosses this boundary.
 * One context per host thread; the library is re-entrant across contexts.
 */
#ifndef IRC_HIP_H
#define IRC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct irc_ctx irc_ctx;

typedef enum {
    IRC_OK = 0,
    IRC_E_ARG = 1,          /* invalid argument (sizes, null pointers, out-of-range indices) */
    IRC_E_NOMEM = 2,        /* device or host allocation failed */
    IRC_E_HIP = 3,          /* HIP runtime error */
    IRC_E_UNSUPPORTED = 4,  /* valid but unsupported (e.g. D too large) */
    IRC_E_FORCED = 5        /* failure forced through irc_options.force_fail (test hook) */
} irc_status;

typedef enum {
    IRC_PREC_EXACT = 0,        /* FP64 over every candidate pair */
    IRC_PREC_FP32 = 1,         /* FP32 MFMA distance screen + FP64 re-rank */
    IRC_PREC_BF16_SCREEN = 2,  /* BF16 (hi/lo split) MFMA screen, FP32 accumulate, + FP64 re-rank */
    IRC_PREC_AUTO = 3,         /* BF16 screen where supported, else FP32 */
    IRC_PREC_FP8_SCREEN = 4    /* FP8 e4m3 MFMA screen (three-term split, per-site power-of-two scale),
                                  FP32 accumulate, + FP64 re-rank */
} irc_precision;

typedef enum {
    IRC_MEM_HOST = 0,    /* pointers in irc_batch / irc_outputs are host memory */
    IRC_MEM_DEVICE = 1   /* pointers are device memory of the context's device */
} irc_mem;

/* irc_options.flags */
#define IRC_FLAG_GRAPH 4  /* irc_rho_delta_batch: capture the device phase (kNN, rho, delta) in a HIP
                             graph and replay it on later calls with the same descriptor: same
                             pointers, sizes, options and host descriptor contents, and unchanged
                             col_spike contents (only F may change between calls).  A call that grows
                             a context buffer invalidates the graph; the next call captures again. */
#define IRC_FLAG_AUTOTUNE 8 /* irc_rho_delta_batch: use the context's tuned options for this workload
                             shape (irc_autotune runs first when there are none yet) */
#define IRC_FLAG_SITE_LAUNCHES 16 /* irc_rho_delta_batch: launch the screen once per site over that
                             site's work items instead of once persistently over all of them (the
                             reference's per-site launch pattern, matlab/irc2.m:3537-3561; for
                             comparison; not with IRC_FLAG_GRAPH) */

typedef struct {
    int32_t precision;  /* irc_precision; default IRC_PREC_AUTO */
    int32_t k_over;     /* over-fetched candidates per query for the screen: 32, 40, 48 or 64
                           (0 = 48 for the BF16 form, 40 for the FP32 form, 64 for the FP8 form) */
    int32_t tile;       /* smallest query group screened per wavefront; smaller groups (the
                           leftovers of a block) use the exact kernel (0 = never) */
    int32_t force_fail; /* test hook: non-zero -> return IRC_E_FORCED, outputs untouched */
    int32_t mem;        /* irc_mem for irc_rho_delta_batch */
    int32_t flags;      /* IRC_FLAG_* bits */
    int32_t waves_per_cu; /* persistent grid of the screen, wavefronts per CU (0 = 8) */
    int32_t large_ncand;  /* blocks with more candidates keep 64 list slots (0 = 8192) */
} irc_options;

/* Fill *opt with defaults. */
void irc_options_default(irc_options* opt);

/* Create / destroy a context on HIP device `device` (ordinal).  The context owns
 * device buffers, the stream and tuning state, and keeps them across calls
 * (the role of the `persistent CK` kernel objects, matlab/irc2.m:3732). */
irc_status irc_ctx_create(irc_ctx** ctx, int device);
void irc_ctx_destroy(irc_ctx* ctx);
const char* irc_last_error(const irc_ctx* ctx);

/* Library / build information string (static storage). */
const char* irc_version(void);

/* ---------------------------------------------------------------------------
 * Per-site problem: one site, one page (prepare_page_site_, matlab/irc2.m:3869-3910).
 *   F            mrFet_out, D x n_cols single (columns: primary spikes by ascending id, then
 *                secondary spikes by ascending id)
 *   drift_of_col viDrift_out, 1-based drift chunk of each column (1..n_drift)
 *   id_of_col    optional global spike id per column, used only to break distance ties
 *                (NULL: column index)
 *   link         mlDrift1, n_drift x n_drift column-major: link[r + c*n_drift] != 0 when
 *                chunk r is a candidate chunk for queries in chunk c
 *   in_begin     0-based first query column (find(vl_in,1)-1); n_in query columns follow
 */
typedef struct {
    int32_t D;
    int32_t k;
    int64_t n_cols;
    const float* F;
    const int32_t* drift_of_col;
    const int64_t* id_of_col;
    int32_t n_drift;
    const uint8_t* link;
    int64_t in_begin;
    int64_t n_in;
} irc_site_problem;

/* search_knn_drift_: dk_out[n_in] = sqrt(d2 of the k-th neighbour) (vrKnn_in),
 * knn_out[k x n_in] = 1-based column indices of the k nearest (miKnn_in), nearest first.
 * Host memory. */
irc_status irc_knn_drift(irc_ctx* ctx, const irc_site_problem* prob, const irc_options* opt,
                         float* dk_out, int32_t* knn_out);

/* search_delta_drift_: rho_cols[n_cols] = rho per column (vrRho_out);
 * dmin_out[n_in] = sqrt(d2) to the nearest column with larger rho (vrDelta_in, not yet
 * multiplied by rho), +Inf when there is none; nneigh_out[n_in] = its 1-based column
 * (the query's own column when there is none).  Host memory. */
irc_status irc_delta_drift(irc_ctx* ctx, const irc_site_problem* prob, const irc_options* opt,
                           const float* rho_cols, float* dmin_out, int32_t* nneigh_out);

/* ---------------------------------------------------------------------------
 * Batch problem: every site and drift chunk of a recording (sort_page_).
 *   F            D x n_cols single, grouped by site: site s owns columns
 *                [site_ptr[s], site_ptr[s+1]); its first site_n_primary[s] columns are the
 *                spikes whose primary site is s (features at s, ascending id), the rest are
 *                spikes whose secondary site is s (features at s, ascending id)
 *                (the layout of _fet_/_fet2_ files, matlab/irc2.m:4583-4618, 4039-4135)
 *   col_spike    1-based spike id of each column (int32; n_spikes < 2^31)
 *   lim_drift    viLim_drift, n_drift+1 entries, 1-based: chunk c holds ids
 *                lim_drift[c] .. lim_drift[c+1]-1 (matlab/irc2.m:6395)
 *   link         mlDrift, n_drift x n_drift column-major (matlab/irc2.m:6433-6437)
 * site_ptr / site_n_primary / lim_drift / link are always host memory; F and col_spike
 * follow irc_options.mem.
 */
typedef struct {
    int64_t n_spikes;
    int32_t D;
    int32_t k;
    int32_t n_sites;
    int32_t n_drift;
    int64_t n_cols;
    const float* F;
    const int32_t* col_spike;
    const int64_t* site_ptr;
    const int64_t* site_n_primary;
    const int64_t* lim_drift;
    const uint8_t* link;
} irc_batch;

/* Outputs indexed by spike id - 1 (irc_options.mem selects host or device memory).
 *   rho     float32(1/sqrt(d2_k))                    (vrRho before rho/max(rho), irc2.m:3669)
 *   delta   float32(sqrt(d2_min) * rho)              (vrDelta, irc2.m:3720-3725)
 *   nneigh  1-based spike id of the nearest denser spike, or self
 *   knn     k x n_spikes int64, 1-based spike ids, nearest first (may be NULL): miKnn_in after
 *           int64(viSpk_out(miKnn_in)) (irc2.m:3670), the type _knn_<site>.irc holds (irc2.m:3395)
 *   dk      sqrt(d2_k) (may be NULL)
 */
typedef struct {
    float* rho;
    float* delta;
    int64_t* nneigh;
    int64_t* knn;
    float* dk;
} irc_outputs;

irc_status irc_rho_delta_batch(irc_ctx* ctx, const irc_batch* batch, const irc_options* opt,
                               irc_outputs* out);

/* irc_rho_delta_batch_noalloc: irc_rho_delta_batch with no device memory allocation (no counterpart in
 * the reference).  For callers that cannot call hipMalloc/hipFree: code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation.  Every device buffer the call needs (plan, packed
 * operands, candidate lists, kNN lists, and in host mode the staging copies) is carved from a 64 MiB
 * array held by each GPU the library runs on, for the life of the process; a problem that needs more
 * (about 55 000 spikes at D = 42, k = 30 with 384 sites x 200 chunks) fails with IRC_E_UNSUPPORTED before
 * any output is written.  Inputs, outputs and options are those of irc_rho_delta_batch (host or device
 * memory); IRC_FLAG_GRAPH, IRC_FLAG_AUTOTUNE and IRC_FLAG_SITE_LAUNCHES are rejected.  All calls on one
 * GPU share its array, which keeps the last call's contents: calls that overlap on one GPU (from
 * several contexts or threads) race.  Host-side descriptor vectors (O(sites + links)) are the only
 * allocations it makes. */
irc_status irc_rho_delta_batch_noalloc(irc_ctx* ctx, const irc_batch* batch, const irc_options* opt,
                                       irc_outputs* out);

/* Autotuner of irc_rho_delta_batch (no reference counterpart: the reference fixes its launch shapes,
 * matlab/irc2.m:3755-3757, 3775).  Times the call on the caller's batch and outputs with each
 * candidate setting of the over-fetch depth (k_over), the screen's persistent grid (waves_per_cu), the
 * smallest query group screened per wavefront (tile, the queue granularity) and the deep-list
 * threshold (large_ncand), starting from *base (its precision and mem are kept), and writes the
 * fastest whose outputs equal base's bit for bit to *best.  The choice is also kept in the context for
 * IRC_FLAG_AUTOTUNE calls of the same workload shape.  The outputs hold base's results on return.
 * rep (may be NULL) receives the trials. */
typedef struct {
    int32_t n_trials;
    int32_t best_trial;        /* index into the trial arrays */
    double ms_base;            /* device time of base's setting */
    double ms_best;
    int32_t k_over[16], waves_per_cu[16], tile[16], large_ncand[16];
    double ms[16];
    int32_t identical[16];     /* the trial reproduced base's outputs */
} irc_tune_report;

irc_status irc_autotune(irc_ctx* ctx, const irc_batch* batch, const irc_options* base, irc_outputs* out,
                        irc_options* best, irc_tune_report* rep);

/* Statistics of the last irc_rho_delta_batch / irc_knn_drift call (for benchmarks and tests). */
typedef struct {
    int64_t n_queries;
    int64_t n_pairs;          /* sum over queries of |C(a)| */
    int64_t n_screen_fallback;/* queries whose screened candidate set could not be certified */
    int64_t n_delta_full;     /* queries whose nearest denser spike was not in their kNN list */
    double ms_total;          /* device time of the whole call, HIP events */
    double ms_knn;            /* kNN + rho kernels */
    double ms_delta;          /* delta kernels */
    double ms_screen;         /* the screen kernel alone (0 on the exact path) */
    int32_t precision_used;
    double ms_rerank;         /* the FP64 re-rank of the certified lists (0 on the exact path) */
    int64_t n_rerank_overflow;/* of n_screen_fallback: queries whose ambiguous run overflowed the re-rank window */
    double ms_graph;          /* device time of the replayed graph (IRC_FLAG_GRAPH replays; else 0) */
    int64_t n_rerank_second;  /* queries the re-rank's long-window second pass redid (their ambiguous
                                 run overflowed the first pass's short window) */
} irc_stats;

irc_status irc_get_stats(const irc_ctx* ctx, irc_stats* st);

/* Self-test of the MFMA lane layouts, the e4m3 decoding of the FP8 MFMA and the cross-half lane swap
 * the screen kernel relies on (no reference counterpart: a diagnostic).  out: 6144 floats, host memory. */
irc_status irc_selftest_mfma(irc_ctx* ctx, float* out);

/* Accumulation self-test of the screen's MFMA chains (no reference counterpart: a diagnostic of the
 * rounding model the certified screen bound assumes).  For t < n_tiles: out[t] (32 x 32) = A[t]
 * (32 x K) times B[t] (K x 32), all row-major float, host memory, computed by one chain of
 * v_mfma_f32_32x32x2_f32 (prec 1, K even) or of v_mfma_f32_32x32x16_bf16 (prec 2, K a multiple of
 * 16; the entries of A and B must be representable in bf16).  2 <= K <= 64, 1 <= n_tiles <= 65536. */
irc_status irc_selftest_mfma_accum(irc_ctx* ctx, int prec, int K, int n_tiles, const float* A, const float* B,
                                   float* out);

/* ---------------------------------------------------------------------------
 * Legacy A/B forms (irc v4 path).
 * irc_knn_ab: B = candidates (D x nB), A = queries (D x nA).  K_out[nA] = sqrt(d2_k);
 * I_out[k x nA] = 1-based indices into B (NULL: not returned; the cuda_knn variant).
 * Exact, ties by lower B index; d = 0 included (cuda_knn_index.cu:27-111 semantics made exact).
 */
irc_status irc_knn_ab(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, int64_t nA, const float* A,
                      int32_t k, float* K_out, int32_t* I_out);

/* irc_knn_ab_full: the cuda_knn_full form (cuda_knn_full.cu:27-100; the kernel has no caller in the
 * reference's MATLAB): K_out[k x nA] = sqrt(d2) of every one of the k nearest (nearest first, not only
 * the k-th), I_out[k x nA] = 1-based indices into B.  Exact, ties by lower B index. */
irc_status irc_knn_ab_full(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, int64_t nA, const float* A,
                           int32_t k, float* K_out, int32_t* I_out);

/* irc_knn_ab_ib: the cuda_knn_ib form (cuda_knn_ib.cu:27-97): the candidates are the columns
 * IB[0..nIB) (1-based, cuda_knn_ib.cu:59) of B.  K_out[nA] = sqrt(d2_k); I_out[k x nA] (may be
 * NULL) = 1-based columns of B (not positions in IB).  Exact, ties by lower B column. */
irc_status irc_knn_ab_ib(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, int64_t nIB, const int32_t* IB,
                         int64_t nA, const float* A, int32_t k, float* K_out, int32_t* I_out);

/* irc_delta_knn_ab: min distance from each A to B with R_B > R_A (cuda_delta_knn.cu:19-88);
 * D_out = sqrt(d2_min) (+Inf when none), N_out 1-based index into B (0 when none). */
irc_status irc_delta_knn_ab(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, const float* RB,
                            int64_t nA, const float* A, const float* RA, float* D_out, int32_t* N_out);

/* irc_rho_drift_dc / irc_delta_drift_dc: dc-count density and delta over one drift-chunk
 * group with the primary/secondary site match rule (irc_cuda_rho_drift.cu:81-88).
 *   miSite12  n12 x 2 column-major (primary site column, secondary site column)
 *   fet1/fet2 D x n12 single; queries are the first n1 spikes
 * rho_out[n1] = #(d2 <= dc2) / #compared; delta_out[n1] = sqrt(d2_min/dc2),
 * nneigh_out[n1] 1-based (self when none). */
irc_status irc_rho_drift_dc(irc_ctx* ctx, int32_t D, int64_t n1, int64_t n12, const int32_t* miSite12,
                            const float* fet1, const float* fet2, float dc2, float* rho_out);
irc_status irc_delta_drift_dc(irc_ctx* ctx, int32_t D, int64_t n1, int64_t n12, const int32_t* miSite12,
                              const float* fet1, const float* fet2, const float* rho12, float dc2,
                              float* delta_out, uint32_t* nneigh_out);

/*
 * Drift linkage: calc_drift_ (matlab/irc2.m:6373-6447) with dist_mask_ml_ (6452-6462),
 * quantile_vr_ (6478-6486), subsample_vr_ (1557-1572) and mi2ml_drift_ (6508-6515).
 * Produces the viLim_drift / mlDrift pair that the kNN and delta entry points consume.
 *   amp, pos    n_spikes single (vrAmp_spk, mrPos_spk(:,2)), host memory
 *   n_amp       nQuantile_drift (amplitude quantile bins), n_pos = numel(viSite2Chan)
 *   n_drift     nTime_drift (chunks), n_batch = nTime_batch (links per chunk),
 *   n_time_max  nTime_max_drift: above it, chunk c may link only to chunks within
 *               n_time_max/4 (plus the corner blocks of calc_drift_)
 * Outputs: *n_drift_out = C (1 when n_batch >= n_drift: one chunk, linked to itself);
 * lim_out[C+1] (1-based, viLim_drift); link_out[C*C] column-major, link_out[r + c*C] =
 * mlDrift(r, c); counts_out (optional) [(n_amp*n_pos) x C] int32 histograms (mrCount_drift).
 * Ranking is by exact squared distance between count vectors, ties to the lower chunk.
 */
typedef struct {
    int64_t n_spikes;
    const float* amp;
    const float* pos;
    int32_t n_amp;
    int32_t n_pos;
    int32_t n_drift;
    int32_t n_batch;
    int32_t n_time_max;
} irc_drift_problem;

irc_status irc_calc_drift(irc_ctx* ctx, const irc_drift_problem* pr, int32_t* n_drift_out, int64_t* lim_out,
                          uint8_t* link_out, int32_t* counts_out);

/* Cluster labels along the nearest-denser chains: the final loop of assignCluster_knn_
 * (matlab/irc2.m:6681-6692), which visits spikes by descending rho and copies the cluster of
 * nneigh.  labels[n] (in/out): the cluster of each peak (> 0) and 0 elsewhere on entry; on return
 * every spike carries the cluster of the first labelled spike on its chain a, nneigh(a), ...
 * nneigh[n]: 1-based (the batch call's nneigh output).  opt->mem selects host or device
 * pointers (opt may be NULL: host). */
irc_status irc_assign_chain(irc_ctx* ctx, int64_t n, const int64_t* nneigh, int32_t* labels,
                            const irc_options* opt);

/* Neighbour-cluster counts behind isolation_score_ (matlab/irc2.m:6699-6731), used by
 * knn_overlap_merge_ (matlab/irc2.m:2156-2174): counts[(c-1)*(n_clu+1) + d] = how many kNN entries
 * of the spikes of cluster c (1..n_clu) lie in cluster d (0: in no cluster).  knn: k x n int64
 * 1-based spike ids (miKnn, the batch call's output); clu[n]: cluster per spike (<= 0: none).
 * counts (n_clu x (n_clu+1), uint64) is overwritten.  n_clu <= 4096.  opt->mem: host or device. */
irc_status irc_isolation_counts(irc_ctx* ctx, int64_t n, int32_t k, const int64_t* knn, const int32_t* clu,
                                int32_t n_clu, uint64_t* counts, const irc_options* opt);

#ifdef __cplusplus
}
#endif

#endif /* IRC_HIP_H */
