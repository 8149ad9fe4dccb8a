/*
 * oracle.c -- CPU restatement (FP64, exact) of IronClust's density-peak hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product path (ironclust_amd, libirc_hip.so) never links or calls it.
 *
 * Canonical semantics restated here (reference paths relative to the
 * ironclust repository):
 *
 *   candidate set  C(a) for a query a with primary site s and drift chunk c:
 *                  {(b, F1_b) : site1[b]==s, chunk(b) in L(c)} U
 *                  {(b, F2_b) : site2[b]==s, chunk(b) in L(c)}
 *                  with L(c) = {r : mlDrift(r,c)}             matlab/irc2.m:3885-3900, 3770
 *                  self included (mlDrift(c,c) is true)         matlab/irc2.m:3664-3665
 *   distance       d2(a,b) = fma-chain over d=0..D-1 of (double(x_d)-double(y_d))^2,
 *                  accumulated in FP64 in dimension order.       (knn_cpu_ computes the same
 *                  quantity in single precision via pdist2,      matlab/irc2.m:4253)
 *   kNN(a)         first k of C(a) ordered by (d2, spike id); if |C(a)|<k the list
 *                  is padded with its last entry                 matlab/irc2.m:4244-4263, 3673-3676
 *   rho(a)         float32(1/sqrt(d2 of k-th entry)), evaluated in FP64
 *                                                                matlab/irc2.m:3669
 *   nneigh(a)      argmin over {b in C(a): rho_b > rho_a} by (d2, id)
 *                                                                matlab/irc2.m:3981-3994 (delta_cpu_)
 *   delta(a)       float32(sqrt(d2_min) * double(rho_a)); if the denser set is empty or
 *                  the product is not finite: nneigh = a, delta = float32(sqrt(3.402e38))
 *                                                                matlab/irc2.m:3720-3725
 *
 * Per-site entry points restate search_knn_drift_ / search_delta_drift_
 * (matlab/irc2.m:3731, 3794) on the site-local column layout built by
 * prepare_page_site_ (matlab/irc2.m:3869-3910).
 *
 * Legacy variants restated for parity of the variant kernels:
 *   orc_rho_drift_dc    jrc3_cuda_rho_drift   matlab/irc_cuda_rho_drift.cu:28-108
 *   orc_delta_drift_dc  jrc3_cuda_delta_drift matlab/irc_cuda_delta_drift.cu:28-132
 *   orc_knn_ab          exact form of cuda_knn_index / knn_cpu_  matlab/cuda_knn_index.cu:27-111
 *   orc_knn_ab_list     exact forms of cuda_knn_full / cuda_knn_ib matlab/cuda_knn_full.cu:27-100,
 *                                                                 matlab/cuda_knn_ib.cu:27-97
 *   orc_delta_knn_ab    cuda_delta_knn / delta_knn_cpu_         matlab/cuda_delta_knn.cu:19-88
 *
 * Build with -ffp-contract=off: every fused multiply-add here is explicit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_SINGLE_INF 3.402E+38

static inline double orc_d2(const float* q, const float* x, int D) {
    double acc = 0.0;
    for (int d = 0; d < D; ++d) {
        double t = (double)q[d] - (double)x[d];
        acc = fma(t, t, acc);
    }
    return acc;
}

/* exported so tests can pin the distance itself */
double orc_dist2(const float* q, const float* x, int D) { return orc_d2(q, x, D); }

float orc_rho_from_d2(double d2) { return (float)(1.0 / sqrt(d2)); }

float orc_sentinel_delta(void) { return (float)sqrt(ORC_SINGLE_INF); }

typedef struct {
    double t;
    int64_t key;   /* tie-break key (spike id) */
    int64_t col;   /* payload (column / spike id) */
} orc_cand;

static int cmp_cand(const void* a, const void* b) {
    const orc_cand* x = (const orc_cand*)a;
    const orc_cand* y = (const orc_cand*)b;
    if (x->t < y->t) return -1;
    if (x->t > y->t) return 1;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return 0;
}

static inline int cand_less(double t1, int64_t k1, double t2, int64_t k2) {
    return (t1 < t2) || (t1 == t2 && k1 < k2);
}

/* ------------------------------------------------------------------------ */
/* Batch form: all sites, all chunks.  Mirrors sort_page_ (irc2.m:3456-3503)
 * without paging (paging does not change results under the link mask; the
 * host-side planner plan_sort_page_ is tested for that property).
 *
 * Inputs use the reference's spike-level arrays:
 *   F1, F2 : D x N single, column-major (i.e. N rows of D floats); F2 may be NULL
 *   site1  : 1..S primary site (viSite_spk); site2: 0 (none) or 1..S (viSite2_spk), may be NULL
 *   lim    : C+1 entries of viLim_drift (1-based; chunk c holds ids lim[c]..lim[c+1]-1)
 *   link   : C x C uint8, column-major, link[r + c*C] = mlDrift(r,c)
 * Outputs (N each, knn is k x N column-major, 1-based spike ids, may be NULL):
 */
typedef struct {
    int64_t* ids;   /* sorted ascending spike ids (0-based) */
    int64_t n;
    int64_t* chunk_ptr; /* C+1 offsets into ids by chunk */
} orc_group;

static void build_group(orc_group* g, int64_t N, const int32_t* site, int s, const int32_t* chunk_of, int C) {
    int64_t n = 0;
    for (int64_t i = 0; i < N; ++i) if (site[i] == s) ++n;
    g->n = n;
    g->ids = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    g->chunk_ptr = (int64_t*)calloc((size_t)C + 1, sizeof(int64_t));
    int64_t j = 0;
    for (int64_t i = 0; i < N; ++i) if (site[i] == s) { g->ids[j++] = i; g->chunk_ptr[chunk_of[i] + 1]++; }
    for (int c = 0; c < C; ++c) g->chunk_ptr[c + 1] += g->chunk_ptr[c];
}

static void free_group(orc_group* g) { free(g->ids); free(g->chunk_ptr); }

int orc_sort_batch(int64_t N, int D, int S, int C, int k,
                   const float* F1, const float* F2,
                   const int32_t* site1, const int32_t* site2,
                   const int64_t* lim, const uint8_t* link,
                   float* rho, float* delta, int64_t* nneigh, int64_t* knn, double* dk2_out,
                   int nthreads) {
    if (N <= 0 || D <= 0 || S <= 0 || C <= 0 || k <= 0) return 1;
    int32_t* chunk_of = (int32_t*)malloc(sizeof(int32_t) * N);
    for (int c = 0; c < C; ++c)
        for (int64_t id = lim[c]; id < lim[c + 1]; ++id)
            if (id >= 1 && id <= N) chunk_of[id - 1] = c;
    /* link lists L(c) in ascending r */
    int* lptr = (int*)calloc((size_t)C + 1, sizeof(int));
    for (int c = 0; c < C; ++c) { lptr[c + 1] = lptr[c]; for (int r = 0; r < C; ++r) if (link[r + (int64_t)c * C]) lptr[c + 1]++; }
    int* lidx = (int*)malloc(sizeof(int) * (lptr[C] > 0 ? lptr[C] : 1));
    for (int c = 0; c < C; ++c) { int j = lptr[c]; for (int r = 0; r < C; ++r) if (link[r + (int64_t)c * C]) lidx[j++] = r; }

    orc_group* g1 = (orc_group*)calloc((size_t)S, sizeof(orc_group));
    orc_group* g2 = (orc_group*)calloc((size_t)S, sizeof(orc_group));
    for (int s = 0; s < S; ++s) {
        build_group(&g1[s], N, site1, s + 1, chunk_of, C);
        if (site2) build_group(&g2[s], N, site2, s + 1, chunk_of, C);
    }
    double* dk2 = (double*)malloc(sizeof(double) * N);
    int nt = nthreads > 0 ? nthreads : 1;
#ifdef _OPENMP
    omp_set_num_threads(nt);
#endif
    /* pass 1: kNN, rho */
#pragma omp parallel
    {
        orc_cand* buf = NULL; int64_t cap = 0;
        /* every spike is a query of its primary site; queries are independent, so the loop runs over
           spikes (a few large sites still use every thread) */
#pragma omp for schedule(dynamic, 64)
        for (int64_t a = 0; a < N; ++a) {
            {
                const int s = site1[a] - 1;
                if (s < 0 || s >= S) continue;
                int c = chunk_of[a];
                int64_t n = 0;
                for (int li = lptr[c]; li < lptr[c + 1]; ++li) {
                    int r = lidx[li];
                    n += g1[s].chunk_ptr[r + 1] - g1[s].chunk_ptr[r];
                    if (site2) n += g2[s].chunk_ptr[r + 1] - g2[s].chunk_ptr[r];
                }
                if (n > cap) { cap = n * 2; buf = (orc_cand*)realloc(buf, sizeof(orc_cand) * cap); }
                int64_t m = 0;
                const float* q = F1 + a * D;
                for (int li = lptr[c]; li < lptr[c + 1]; ++li) {
                    int r = lidx[li];
                    for (int64_t j = g1[s].chunk_ptr[r]; j < g1[s].chunk_ptr[r + 1]; ++j) {
                        int64_t b = g1[s].ids[j];
                        buf[m].t = orc_d2(q, F1 + b * D, D); buf[m].key = b; buf[m].col = b; ++m;
                    }
                    if (site2) for (int64_t j = g2[s].chunk_ptr[r]; j < g2[s].chunk_ptr[r + 1]; ++j) {
                        int64_t b = g2[s].ids[j];
                        buf[m].t = orc_d2(q, F2 + b * D, D); buf[m].key = b; buf[m].col = b; ++m;
                    }
                }
                qsort(buf, (size_t)m, sizeof(orc_cand), cmp_cand);
                int64_t kk = m < k ? m : k;
                if (knn) {
                    for (int j = 0; j < k; ++j) knn[a * k + j] = buf[j < kk ? j : kk - 1].col + 1;
                }
                dk2[a] = buf[kk - 1].t;
                rho[a] = (float)(1.0 / sqrt(dk2[a]));
            }
        }
        free(buf);
    }
    /* pass 2: delta, nneigh */
    const float sent = (float)sqrt(ORC_SINGLE_INF);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t a = 0; a < N; ++a) {
        {
            const int s = site1[a] - 1;
            if (s < 0 || s >= S) continue;
            int c = chunk_of[a];
            const float* q = F1 + a * D;
            float ra = rho[a];
            double best_t = INFINITY; int64_t best = -1;
            for (int li = lptr[c]; li < lptr[c + 1]; ++li) {
                int r = lidx[li];
                for (int part = 0; part < (site2 ? 2 : 1); ++part) {
                    orc_group* g = part == 0 ? &g1[s] : &g2[s];
                    const float* F = part == 0 ? F1 : F2;
                    for (int64_t j = g->chunk_ptr[r]; j < g->chunk_ptr[r + 1]; ++j) {
                        int64_t b = g->ids[j];
                        if (!(rho[b] > ra)) continue;
                        double t = orc_d2(q, F + b * D, D);
                        if (best < 0 || cand_less(t, b, best_t, best)) { best_t = t; best = b; }
                    }
                }
            }
            if (best < 0) { delta[a] = sent; nneigh[a] = a + 1; }
            else {
                float dl = (float)(sqrt(best_t) * (double)ra);
                if (isnan(dl) || isinf(dl)) { delta[a] = sent; nneigh[a] = a + 1; }
                else { delta[a] = dl; nneigh[a] = best + 1; }
            }
        }
    }
    if (dk2_out) memcpy(dk2_out, dk2, sizeof(double) * N);
    for (int s = 0; s < S; ++s) { free_group(&g1[s]); if (site2) free_group(&g2[s]); }
    free(g1); free(g2); free(dk2); free(chunk_of); free(lptr); free(lidx);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-site forms.  F is D x n_cols (column-major) as mrFet_out; drift_of_col is
 * viDrift_out (1-based); link is mlDrift1 (n_drift x n_drift, column-major);
 * queries are columns in_begin .. in_begin+n_in-1 (0-based, vl_in).
 * id_of_col (may be NULL) gives the tie-break key; NULL uses the column index.
 */
static int64_t site_cands(int D, int64_t n_cols, const float* F, const int32_t* drift_of_col,
                          const int64_t* id_of_col, int n_drift, const uint8_t* link,
                          int64_t qcol, orc_cand* buf) {
    int c = drift_of_col[qcol] - 1;
    const float* q = F + qcol * D;
    int64_t m = 0;
    for (int64_t j = 0; j < n_cols; ++j) {
        int r = drift_of_col[j] - 1;
        if (r < 0 || r >= n_drift || !link[r + (int64_t)c * n_drift]) continue;
        buf[m].t = orc_d2(q, F + j * D, D);
        buf[m].key = id_of_col ? id_of_col[j] : j;
        buf[m].col = j;
        ++m;
    }
    return m;
}

/* search_knn_drift_ (irc2.m:3731-3789), exact: dk_out = float32(sqrt(d2_k)),
 * knn_out = k x n_in, 1-based column indices. */
int orc_knn_site(int D, int64_t n_cols, const float* F, const int32_t* drift_of_col,
                 const int64_t* id_of_col, int n_drift, const uint8_t* link,
                 int64_t in_begin, int64_t n_in, int k, float* dk_out, int32_t* knn_out) {
    orc_cand* buf = (orc_cand*)malloc(sizeof(orc_cand) * (n_cols > 0 ? n_cols : 1));
    for (int64_t i = 0; i < n_in; ++i) {
        int64_t qcol = in_begin + i;
        int64_t m = site_cands(D, n_cols, F, drift_of_col, id_of_col, n_drift, link, qcol, buf);
        if (m <= 0) { free(buf); return 1; }
        qsort(buf, (size_t)m, sizeof(orc_cand), cmp_cand);
        int64_t kk = m < k ? m : k;
        for (int j = 0; j < k; ++j) knn_out[i * k + j] = (int32_t)(buf[j < kk ? j : kk - 1].col + 1);
        dk_out[i] = (float)sqrt(buf[kk - 1].t);
    }
    free(buf);
    return 0;
}

/* search_delta_drift_ (irc2.m:3794-3865), exact: dmin_out = float32(sqrt(d2_min)) over
 * candidates with rho_cols[b] > rho_cols[q]; +Inf and own column when none. */
int orc_delta_site(int D, int64_t n_cols, const float* F, const int32_t* drift_of_col,
                   const int64_t* id_of_col, int n_drift, const uint8_t* link,
                   int64_t in_begin, int64_t n_in, const float* rho_cols,
                   float* dmin_out, int32_t* nneigh_out) {
    for (int64_t i = 0; i < n_in; ++i) {
        int64_t qcol = in_begin + i;
        int c = drift_of_col[qcol] - 1;
        const float* q = F + qcol * D;
        float ra = rho_cols[qcol];
        double bt = INFINITY; int64_t bk = 0, bc = -1;
        for (int64_t j = 0; j < n_cols; ++j) {
            int r = drift_of_col[j] - 1;
            if (r < 0 || r >= n_drift || !link[r + (int64_t)c * n_drift]) continue;
            if (!(rho_cols[j] > ra)) continue;
            double t = orc_d2(q, F + j * D, D);
            int64_t key = id_of_col ? id_of_col[j] : j;
            if (bc < 0 || cand_less(t, key, bt, bk)) { bt = t; bk = key; bc = j; }
        }
        if (bc < 0) { dmin_out[i] = INFINITY; nneigh_out[i] = (int32_t)(qcol + 1); }
        else { dmin_out[i] = (float)sqrt(bt); nneigh_out[i] = (int32_t)(bc + 1); }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* A/B forms (legacy irc v4 path): B = candidates (nC x nB), A = queries.
 * Exact kNN (knn_cpu_, irc2.m:4244-4263; cuda_knn_index.cu:27-111 approximates it):
 * K = float32(sqrt(d2_k)), I = k x nA, 1-based index into B, ties by B index. */
int orc_knn_ab(int D, int64_t nB, const float* B, int64_t nA, const float* A, int k,
               float* K_out, int32_t* I_out) {
    orc_cand* buf = (orc_cand*)malloc(sizeof(orc_cand) * (nB > 0 ? nB : 1));
    if (nB <= 0) { free(buf); return 1; }
    for (int64_t i = 0; i < nA; ++i) {
        for (int64_t j = 0; j < nB; ++j) { buf[j].t = orc_d2(A + i * D, B + j * D, D); buf[j].key = j; buf[j].col = j; }
        qsort(buf, (size_t)nB, sizeof(orc_cand), cmp_cand);
        int64_t kk = nB < k ? nB : k;
        for (int j = 0; j < k; ++j) I_out[i * k + j] = (int32_t)(buf[j < kk ? j : kk - 1].col + 1);
        K_out[i] = (float)sqrt(buf[kk - 1].t);
    }
    free(buf);
    return 0;
}

/* cuda_knn_full (cuda_knn_full.cu:27-100) and cuda_knn_ib (cuda_knn_ib.cu:27-97), exact:
 * the candidates are the columns IB[0..nIB) of B (1-based, cuda_knn_ib.cu:59; IB NULL: every
 * column of B), ordered by (d2, B column); fewer than k are padded with the last one.
 * K_all (k x nA, may be NULL) = float32(sqrt(d2)) of every listed neighbour (cuda_knn_full.cu:97-100),
 * K_out (nA, may be NULL) = that of the k-th (cuda_knn_ib.cu:90), I_out (k x nA) = 1-based B columns. */
int orc_knn_ab_list(int D, int64_t nB, const float* B, int64_t nIB, const int32_t* IB, int64_t nA,
                    const float* A, int k, float* K_all, float* K_out, int32_t* I_out) {
    const int64_t n = IB ? nIB : nB;
    if (n <= 0 || k <= 0) return 1;
    for (int64_t j = 0; IB && j < n; ++j)
        if (IB[j] < 1 || IB[j] > nB) return 2;
    orc_cand* buf = (orc_cand*)malloc(sizeof(orc_cand) * (size_t)n);
    for (int64_t i = 0; i < nA; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            const int64_t col = IB ? (int64_t)IB[j] - 1 : j;
            buf[j].t = orc_d2(A + i * D, B + col * D, D);
            buf[j].key = col;
            buf[j].col = col;
        }
        qsort(buf, (size_t)n, sizeof(orc_cand), cmp_cand);
        const int64_t kk = n < k ? n : k;
        for (int j = 0; j < k; ++j) {
            const orc_cand* c = &buf[j < kk ? j : kk - 1];
            I_out[i * k + j] = (int32_t)(c->col + 1);
            if (K_all) K_all[i * k + j] = (float)sqrt(c->t);
        }
        if (K_out) K_out[i] = (float)sqrt(buf[kk - 1].t);
    }
    free(buf);
    return 0;
}

/* cuda_delta_knn (cuda_delta_knn.cu:19-88) / delta_knn_cpu_ (irc.m:24558):
 * min over b with R_B[b] > R_A[a] (d = 0 allowed); ties by B index.
 * D_out = float32(sqrt(d2_min)), N_out 1-based; none -> +Inf and 0. */
int orc_delta_knn_ab(int D, int64_t nB, const float* B, const float* RB,
                     int64_t nA, const float* A, const float* RA,
                     float* D_out, int32_t* N_out) {
    for (int64_t i = 0; i < nA; ++i) {
        double bt = INFINITY; int64_t bj = -1;
        for (int64_t j = 0; j < nB; ++j) {
            if (!(RB[j] > RA[i])) continue;
            double t = orc_d2(A + i * D, B + j * D, D);
            if (bj < 0 || t < bt) { bt = t; bj = j; }
        }
        if (bj < 0) { D_out[i] = INFINITY; N_out[i] = 0; }
        else { D_out[i] = (float)sqrt(bt); N_out[i] = (int32_t)(bj + 1); }
    }
    return 0;
}

/* jrc3_cuda_rho_drift (irc_cuda_rho_drift.cu:28-108), FP64 distance:
 * for query i (0..n1-1; queries are the first n1 of the n12 spikes of a drift
 * chunk group), over every candidate j (0..n12-1, self included):
 *   site match: site1[i] == site1[j] -> compare with fet1[j]; else if
 *   site1[i] == site2[j] -> compare with fet2[j]; else skip   (:81-88)
 * rho = float32(#(d2 <= dc2) / #compared), the division in FP64 (:104).
 * miSite12 is n12 x 2 column-major (site1 column then site2 column). */
int orc_rho_drift_dc(int D, int64_t n1, int64_t n12, const int32_t* miSite12,
                     const float* fet1, const float* fet2, float dc2, float* rho_out) {
    for (int64_t i = 0; i < n1; ++i) {
        int s = miSite12[i];
        int64_t ncmp = 0, ncnt = 0;
        for (int64_t j = 0; j < n12; ++j) {
            const float* x;
            if (miSite12[j] == s) x = fet1 + j * D;
            else if (miSite12[j + n12] == s) x = fet2 + j * D;
            else continue;
            double t = orc_d2(fet1 + i * D, x, D);
            ++ncmp;
            if (t <= (double)dc2) ++ncnt;
        }
        rho_out[i] = ncmp > 0 ? (float)((double)ncnt / (double)ncmp) : 0.0f;
    }
    return 0;
}

/* jrc3_cuda_delta_drift (irc_cuda_delta_drift.cu:28-132), FP64 distance:
 * min d2 over site-matched candidates with rho[j] > rho[i] (:90), ties by lower j;
 * the running minimum starts at single(3.402e38) with nneigh = self (:64-67).
 * delta = float32(sqrt(d2_min / dc2)), nneigh 1-based (:118-129). */
int orc_delta_drift_dc(int D, int64_t n1, int64_t n12, const int32_t* miSite12,
                       const float* fet1, const float* fet2, const float* rho12, float dc2,
                       float* delta_out, uint32_t* nneigh_out) {
    for (int64_t i = 0; i < n1; ++i) {
        int s = miSite12[i];
        double bt = (double)(float)ORC_SINGLE_INF; int64_t bj = -1;
        for (int64_t j = 0; j < n12; ++j) {
            const float* x;
            if (miSite12[j] == s) x = fet1 + j * D;
            else if (miSite12[j + n12] == s) x = fet2 + j * D;
            else continue;
            if (!(rho12[j] > rho12[i])) continue;
            double t = orc_d2(fet1 + i * D, x, D);
            if (t < bt) { bt = t; bj = j; }
        }
        delta_out[i] = (float)sqrt(bt / (double)dc2);
        nneigh_out[i] = (uint32_t)(bj < 0 ? i + 1 : bj + 1);
    }
    return 0;
}
