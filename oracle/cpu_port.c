/*
 * cpu_port.c -- FP32 CPU restatement of the reference's CPU path, used ONLY as
 * the timed CPU baseline (bench.py cpu_baseline leg) and checked by tests.
 * TEST / MEASUREMENT INFRASTRUCTURE: the product path never links or calls it.
 *
 * It follows the reference's CPU fallback literally:
 *   rho_paged_site_  (matlab/irc2.m:3653-3668): per site, per drift chunk, the
 *                    candidates are the site's spikes in the linked chunks;
 *   knn_cpu_         (matlab/irc2.m:4244-4263): queries in blocks of 1000,
 *                    Euclidean distances in single precision, k smallest with
 *                    their ids (pdist2 'Smallest'), d_k = k-th distance;
 *                    rho = 1./d_k (:3669); the ids are mapped to global 1-based
 *                    spike ids and padded with the last one (:3666-3676)
 *   delta_paged_site_(matlab/irc2.m:3703-3725) + delta_cpu_ (matlab/irc2.m:3981-3994):
 *                    d2 = |y|^2 + |x|^2 - 2 x'y in single precision (GEMM form),
 *                    rho2 <= rho1 masked to inf, column min (first occurrence),
 *                    delta = sqrt(|d2|) * rho, NaN/Inf -> sqrt(3.402e38), nneigh = self.
 * Sites run in parallel with OpenMP, as parfor iSite does (matlab/irc2.m:3539,3580).
 * Spike-level inputs are the same as orc_sort_batch in oracle.c.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define STEP_KNN 1000

typedef struct {
    int64_t* ids;
    int64_t n;
    int64_t* chunk_ptr;
} grp;

static void grp_build(grp* g, int64_t N, const int32_t* site, int s, const int32_t* chunk_of, int C) {
    int64_t n = 0;
    for (int64_t i = 0; i < N; ++i) if (site[i] == s) ++n;
    g->n = n;
    g->ids = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    g->chunk_ptr = (int64_t*)calloc((size_t)C + 1, sizeof(int64_t));
    int64_t j = 0;
    for (int64_t i = 0; i < N; ++i) if (site[i] == s) { g->ids[j++] = i; g->chunk_ptr[chunk_of[i] + 1]++; }
    for (int c = 0; c < C; ++c) g->chunk_ptr[c + 1] += g->chunk_ptr[c];
}

/* Build the candidate matrix of one (site, chunk): transposed (D x m, row d contiguous)
 * for vectorised distance loops, plus spike ids, in the reference's order:
 * per linked chunk (ascending), primary columns then secondary columns. */
static int64_t gather_cands(int D, const float* F1, const float* F2, const grp* g1, const grp* g2,
                            const int* lidx, int l0, int l1, float* XT, int64_t ldx, int64_t* ids) {
    int64_t m = 0;
    for (int li = l0; li < l1; ++li) {
        int r = lidx[li];
        for (int part = 0; part < (g2 ? 2 : 1); ++part) {
            const grp* g = part == 0 ? g1 : g2;
            const float* F = part == 0 ? F1 : F2;
            for (int64_t j = g->chunk_ptr[r]; j < g->chunk_ptr[r + 1]; ++j) {
                int64_t b = g->ids[j];
                for (int d = 0; d < D; ++d) XT[d * ldx + m] = F[b * D + d];
                ids[m++] = b;
            }
        }
    }
    return m;
}

int cpu_sort_batch(int64_t N, int D, int S, int C, int k,
                   const float* F1, const float* F2,
                   const int32_t* site1, const int32_t* site2,
                   const int64_t* lim, const uint8_t* link,
                   const int32_t* site_list, int n_site_list,
                   float* rho, float* delta, int64_t* nneigh, int64_t* knn, int nthreads) {
    if (N <= 0 || D <= 0 || S <= 0 || C <= 0 || k <= 0) return 1;
    int32_t* chunk_of = (int32_t*)malloc(sizeof(int32_t) * N);
    for (int c = 0; c < C; ++c)
        for (int64_t id = lim[c]; id < lim[c + 1]; ++id)
            if (id >= 1 && id <= N) chunk_of[id - 1] = c;
    int* lptr = (int*)calloc((size_t)C + 1, sizeof(int));
    for (int c = 0; c < C; ++c) { lptr[c + 1] = lptr[c]; for (int r = 0; r < C; ++r) if (link[r + (int64_t)c * C]) lptr[c + 1]++; }
    int* lidx = (int*)malloc(sizeof(int) * (lptr[C] > 0 ? lptr[C] : 1));
    for (int c = 0; c < C; ++c) { int j = lptr[c]; for (int r = 0; r < C; ++r) if (link[r + (int64_t)c * C]) lidx[j++] = r; }

    int ns = site_list ? n_site_list : S;
    grp* g1 = (grp*)calloc((size_t)S, sizeof(grp));
    grp* g2 = (grp*)calloc((size_t)S, sizeof(grp));
    for (int si = 0; si < ns; ++si) {
        int s = site_list ? site_list[si] - 1 : si;
        grp_build(&g1[s], N, site1, s + 1, chunk_of, C);
        if (site2) grp_build(&g2[s], N, site2, s + 1, chunk_of, C);
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const float sent = (float)sqrt(3.402E+38);
    /* pass 1: knn_cpu_ per (site, chunk) -> rho */
#pragma omp parallel
    {
        int64_t cap = 0;
        float* XT = NULL; int64_t* ids = NULL; float* dist = NULL;
        float* topd = (float*)malloc(sizeof(float) * (k + 1));
        int64_t* topi = (int64_t*)malloc(sizeof(int64_t) * (k + 1));
#pragma omp for schedule(dynamic, 1)
        for (int si = 0; si < ns; ++si) {
            int s = site_list ? site_list[si] - 1 : si;
            for (int c = 0; c < C; ++c) {
                int64_t q0 = g1[s].chunk_ptr[c], q1 = g1[s].chunk_ptr[c + 1];
                if (q1 <= q0) continue;
                int64_t m = 0;
                for (int li = lptr[c]; li < lptr[c + 1]; ++li) {
                    int r = lidx[li];
                    m += g1[s].chunk_ptr[r + 1] - g1[s].chunk_ptr[r];
                    if (site2) m += g2[s].chunk_ptr[r + 1] - g2[s].chunk_ptr[r];
                }
                if (m > cap) {
                    cap = m + 64;
                    XT = (float*)realloc(XT, sizeof(float) * D * cap);
                    ids = (int64_t*)realloc(ids, sizeof(int64_t) * cap);
                    dist = (float*)realloc(dist, sizeof(float) * cap);
                }
                m = gather_cands(D, F1, F2, &g1[s], site2 ? &g2[s] : NULL, lidx, lptr[c], lptr[c + 1], XT, cap, ids);
                int kk = m < k ? (int)m : k;
                for (int64_t qb = q0; qb < q1; qb += STEP_KNN) {
                    int64_t qe = qb + STEP_KNN < q1 ? qb + STEP_KNN : q1;
                    for (int64_t qi = qb; qi < qe; ++qi) {
                        int64_t a = g1[s].ids[qi];
                        const float* q = F1 + a * D;
                        for (int64_t j = 0; j < m; ++j) dist[j] = 0.0f;
                        for (int d = 0; d < D; ++d) {
                            const float qd = q[d];
                            const float* xr = XT + d * cap;
                            for (int64_t j = 0; j < m; ++j) { float t = qd - xr[j]; dist[j] += t * t; }
                        }
                        /* k smallest with their indices (pdist2 'Smallest'): sorted insertion with threshold */
                        int nt = 0;
                        for (int64_t j = 0; j < m; ++j) {
                            float v = dist[j];
                            if (nt == kk && !(v < topd[kk - 1])) continue;
                            int p = nt < kk ? nt++ : kk - 1;
                            while (p > 0 && topd[p - 1] > v) { topd[p] = topd[p - 1]; topi[p] = topi[p - 1]; --p; }
                            topd[p] = v;
                            topi[p] = j;
                        }
                        float dk = sqrtf(topd[kk - 1]);
                        rho[a] = 1.0f / dk;
                        /* global ids, padded with the last entry (rho_paged_site_, matlab/irc2.m:3666-3676) */
                        if (knn)
                            for (int i = 0; i < k; ++i) knn[a * k + i] = ids[topi[i < kk ? i : kk - 1]] + 1;
                    }
                }
            }
        }
        free(XT); free(ids); free(dist); free(topd); free(topi);
    }
    /* pass 2: delta_cpu_ per (site, chunk) */
#pragma omp parallel
    {
        int64_t cap = 0;
        float* XT = NULL; int64_t* ids = NULL; float* n2 = NULL; float* r2 = NULL;
#pragma omp for schedule(dynamic, 1)
        for (int si = 0; si < ns; ++si) {
            int s = site_list ? site_list[si] - 1 : si;
            for (int c = 0; c < C; ++c) {
                int64_t q0 = g1[s].chunk_ptr[c], q1 = g1[s].chunk_ptr[c + 1];
                if (q1 <= q0) continue;
                int64_t m = 0;
                for (int li = lptr[c]; li < lptr[c + 1]; ++li) {
                    int r = lidx[li];
                    m += g1[s].chunk_ptr[r + 1] - g1[s].chunk_ptr[r];
                    if (site2) m += g2[s].chunk_ptr[r + 1] - g2[s].chunk_ptr[r];
                }
                if (m > cap) {
                    cap = m + 64;
                    XT = (float*)realloc(XT, sizeof(float) * D * cap);
                    ids = (int64_t*)realloc(ids, sizeof(int64_t) * cap);
                    n2 = (float*)realloc(n2, sizeof(float) * cap);
                    r2 = (float*)realloc(r2, sizeof(float) * cap);
                }
                m = gather_cands(D, F1, F2, &g1[s], site2 ? &g2[s] : NULL, lidx, lptr[c], lptr[c + 1], XT, cap, ids);
                for (int64_t j = 0; j < m; ++j) { n2[j] = 0.0f; r2[j] = rho[ids[j]]; }
                for (int d = 0; d < D; ++d) { const float* xr = XT + d * cap; for (int64_t j = 0; j < m; ++j) n2[j] += xr[j] * xr[j]; }
                for (int64_t qi = q0; qi < q1; ++qi) {
                    int64_t a = g1[s].ids[qi];
                    const float* q = F1 + a * D;
                    float nq = 0.0f;
                    for (int d = 0; d < D; ++d) nq += q[d] * q[d];
                    float r1 = rho[a];
                    float best = INFINITY; int64_t bj = 0;
                    /* dot products for this query; the GEMM form of delta_cpu_ */
                    for (int64_t j0 = 0; j0 < m; j0 += 256) {
                        int64_t j1 = j0 + 256 < m ? j0 + 256 : m;
                        float dot[256];
                        for (int64_t j = j0; j < j1; ++j) dot[j - j0] = 0.0f;
                        for (int d = 0; d < D; ++d) {
                            const float qd = q[d];
                            const float* xr = XT + d * cap;
                            for (int64_t j = j0; j < j1; ++j) dot[j - j0] += xr[j] * qd;
                        }
                        for (int64_t j = j0; j < j1; ++j) {
                            float v = nq + (n2[j] - 2.0f * dot[j - j0]);
                            if (r2[j] <= r1) v = INFINITY;
                            if (v < best) { best = v; bj = j; }
                        }
                    }
                    float dl = sqrtf(fabsf(best)) * r1;
                    if (isnan(dl) || isinf(dl)) { delta[a] = sent; nneigh[a] = a + 1; }
                    else { delta[a] = dl; nneigh[a] = ids[bj] + 1; }
                }
            }
        }
        free(XT); free(ids); free(n2); free(r2);
    }
    for (int si = 0; si < ns; ++si) {
        int s = site_list ? site_list[si] - 1 : si;
        free(g1[s].ids); free(g1[s].chunk_ptr);
        if (site2) { free(g2[s].ids); free(g2[s].chunk_ptr); }
    }
    free(g1); free(g2); free(chunk_of); free(lptr); free(lidx);
    return 0;
}

int cpu_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
