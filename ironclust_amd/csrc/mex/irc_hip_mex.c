/*
 * irc_hip_mex.c -- MATLAB MEX gateway over the C ABI of libirc_hip.so (include/irc_hip.h).
 * It holds no logic: it checks the mxArray sizes, maps them to the ABI structs and maps errors to
 * mexErrMsgIdAndTxt('irc:hip', ...), which the caller's try/catch turns into its CPU fallback
 * (matlab/irc2.m:3643-3652, 3694-3702).  Build where MATLAB's mex.h exists:
 *   mex -R2018a irc_hip_mex.c -I<repo>/include -L<repo>/ironclust_amd/_lib -lirc_hip
 *
 * Commands (1-based indices in and out, MATLAB orientation):
 *   [vrKnn_in, miKnn_in]    = irc_hip_mex('knn_drift', vl_in, mrFet_out, viDrift_out, mlDrift1, knn)
 *                              replaces search_knn_drift_ (matlab/irc2.m:3731-3789); miKnn_in int32
 *                              column indices, as the reference
 *   [vrDelta_in, viNneigh_in] = irc_hip_mex('delta_drift', vl_in, mrFet_out, vrRho_out, viDrift_out, mlDrift1)
 *                              replaces search_delta_drift_ (matlab/irc2.m:3794-3865)
 *   [vrRho, vrDelta, viNneigh, miKnn] = irc_hip_mex('rho_delta_page', mrFet, viSpk_col, viSite_ptr,
 *                              vnPrimary_site, viLim_drift, mlDrift, knn)
 *                              replaces rho_page_ + delta_page_ (matlab/irc2.m:3524-3607); viNneigh and
 *                              miKnn are int64 spike ids (irc2.m:3670, 3723), the type _knn_<site>.irc holds
 *   [vrKnn, miKnn]          = irc_hip_mex('knn_ab', mrFet2, mrFet1, knn)   replaces cuda_knn_ (irc2.m:4202)
 *   [mrKnn, miKnn]          = irc_hip_mex('knn_ab_full', mrFet2, mrFet1, knn)   the cuda_knn_full form
 *   [vrKnn, miKnn]          = irc_hip_mex('knn_ab_ib', mrFet2, viB, mrFet1, knn) the cuda_knn_ib form
 *   [vrOpt, mrTrial]        = irc_hip_mex('autotune', mrFet, viSpk_col, viSite_ptr, vnPrimary_site,
 *                              viLim_drift, mlDrift, knn [, vrOpt0])
 *                              times rho_delta_page's call with each candidate setting (irc_autotune)
 *                              and returns the fastest that reproduces vrOpt0's outputs bit for bit;
 *                              the context keeps it for later calls with the autotune flag (8).
 *                              mrTrial has one row per trial: [k_over, waves_per_cu, tile,
 *                              large_ncand, ms, identical, chosen]
 *   irc_hip_mex('reset')   destroys the persistent context (and its tuned settings)
 * Optional trailing argument vrOpt = [precision, flags, k_over, tile, waves_per_cu, large_ncand], any
 * leading part of it (a scalar is the precision): precision 0 exact, 1 FP32 screen, 2 BF16 screen,
 * 3 auto, 4 FP8 screen; flags 4 graph replay, 8 autotuned settings, 16 per-site launches; the other
 * entries as irc_options (0 = the library's default).
 */
#include <stdint.h>
#include <string.h>

#include "mex.h"
#include "irc_hip.h"

static irc_ctx* g_ctx = NULL;

static void cleanup(void) {
    if (g_ctx) irc_ctx_destroy(g_ctx);
    g_ctx = NULL;
}

static irc_ctx* ctx(void) {
    if (!g_ctx) {
        if (irc_ctx_create(&g_ctx, 0) != IRC_OK) mexErrMsgIdAndTxt("irc:hip", "irc_ctx_create failed");
        mexAtExit(cleanup);
    }
    return g_ctx;
}

static void check(irc_status st) {
    if (st != IRC_OK) mexErrMsgIdAndTxt("irc:hip", "%s", irc_last_error(g_ctx));
}

static void need(int ok, const char* what) {
    if (!ok) mexErrMsgIdAndTxt("irc:hip:arg", "%s", what);
}

static mwSize numel(const mxArray* a) { return mxGetNumberOfElements(a); }

static int32_t* to_i32(const mxArray* a) {
    const mwSize n = numel(a);
    int32_t* v = (int32_t*)mxMalloc(sizeof(int32_t) * (n ? n : 1));
    const double* d = mxIsDouble(a) ? mxGetPr(a) : NULL;
    for (mwSize i = 0; i < n; ++i) {
        if (d) v[i] = (int32_t)d[i];
        else if (mxIsInt32(a)) v[i] = ((const int32_t*)mxGetData(a))[i];
        else if (mxIsInt64(a)) v[i] = (int32_t)((const int64_t*)mxGetData(a))[i];
        else if (mxIsSingle(a)) v[i] = (int32_t)((const float*)mxGetData(a))[i];
        else mexErrMsgIdAndTxt("irc:hip:arg", "unsupported index class");
    }
    return v;
}

static int64_t* to_i64(const mxArray* a) {
    const mwSize n = numel(a);
    int64_t* v = (int64_t*)mxMalloc(sizeof(int64_t) * (n ? n : 1));
    if (mxIsInt64(a)) {
        for (mwSize i = 0; i < n; ++i) v[i] = ((const int64_t*)mxGetData(a))[i];
    } else if (mxIsDouble(a)) {
        for (mwSize i = 0; i < n; ++i) v[i] = (int64_t)mxGetPr(a)[i];
    } else {
        int32_t* t = to_i32(a);
        for (mwSize i = 0; i < n; ++i) v[i] = t[i];
        mxFree(t);
    }
    return v;
}

static uint8_t* to_u8(const mxArray* a) {
    const mwSize n = numel(a);
    uint8_t* v = (uint8_t*)mxMalloc(n ? n : 1);
    if (mxIsLogical(a)) { const mxLogical* l = mxGetLogicals(a); for (mwSize i = 0; i < n; ++i) v[i] = l[i] ? 1 : 0; }
    else { int32_t* t = to_i32(a); for (mwSize i = 0; i < n; ++i) v[i] = t[i] != 0; mxFree(t); }
    return v;
}

static void opts_from(irc_options* o, int nrhs, const mxArray* prhs[], int pos) {
    irc_options_default(o);
    if (nrhs <= pos) return;
    const mwSize n = numel(prhs[pos]);
    need(n >= 1 && n <= 6, "options must be [precision, flags, k_over, tile, waves_per_cu, large_ncand] or a leading part");
    int32_t* v = to_i32(prhs[pos]);
    int32_t* f[6] = {&o->precision, &o->flags, &o->k_over, &o->tile, &o->waves_per_cu, &o->large_ncand};
    for (mwSize i = 0; i < n; ++i) *f[i] = v[i];
    mxFree(v);
}

/* rho_delta_page / autotune arguments prhs[1..7] -> irc_batch; the index arrays it converts are
 * returned in tmp[0..4] for mxFree */
static void page_batch(int nrhs, const mxArray* prhs[], irc_batch* b, void* tmp[5]) {
    need(nrhs >= 8 && mxIsSingle(prhs[1]), "bad arguments");
    const mwSize n_cols = mxGetN(prhs[1]);
    const mwSize S = numel(prhs[4]);
    need(S >= 1, "vnPrimary_site is empty");
    need(numel(prhs[2]) == n_cols, "numel(viSpk_col) must equal size(mrFet, 2)");
    need(numel(prhs[3]) == S + 1, "numel(viSite_ptr) must equal numel(vnPrimary_site) + 1");
    need(numel(prhs[5]) >= 2, "viLim_drift needs at least two entries");
    const mwSize C = numel(prhs[5]) - 1;
    need(mxGetM(prhs[6]) == C && mxGetN(prhs[6]) == C, "mlDrift must be n_drift x n_drift");
    memset(b, 0, sizeof(*b));
    b->D = (int32_t)mxGetM(prhs[1]);
    b->n_cols = (int64_t)n_cols;
    b->F = (const float*)mxGetData(prhs[1]);
    int64_t* sp = to_i64(prhs[3]);
    tmp[0] = sp;
    need(sp[S] == (int64_t)n_cols, "viSite_ptr(end) must equal size(mrFet, 2)");
    int32_t* spk = to_i32(prhs[2]);
    int64_t* np = to_i64(prhs[4]);
    int64_t* lim = to_i64(prhs[5]);
    uint8_t* lk = to_u8(prhs[6]);
    tmp[1] = spk; tmp[2] = np; tmp[3] = lim; tmp[4] = lk;
    b->col_spike = spk;
    b->site_ptr = sp;
    b->site_n_primary = np;
    b->n_sites = (int32_t)S;
    b->lim_drift = lim;
    b->n_drift = (int32_t)C;
    b->link = lk;
    b->k = (int32_t)mxGetScalar(prhs[7]);
    need(b->k >= 1, "knn must be positive");
    b->n_spikes = 0;
    for (mwSize s = 0; s < S; ++s) b->n_spikes += np[s];
}

/* contiguous run marked by vl_in -> (in_begin, n_in) */
static void in_range(const mxArray* vl, int64_t* beg, int64_t* n) {
    const mwSize m = numel(vl);
    const mxLogical* l = mxGetLogicals(vl);
    int64_t b = -1, e = -1;
    for (mwSize i = 0; i < m; ++i) if (l[i]) { if (b < 0) b = (int64_t)i; e = (int64_t)i; }
    for (int64_t i = b; b >= 0 && i <= e; ++i) need(l[i] != 0, "vl_in must mark a contiguous run of columns");
    *beg = b < 0 ? 0 : b;
    *n = b < 0 ? 0 : e - b + 1;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    char cmd[64];
    need(nrhs >= 1 && mxIsChar(prhs[0]), "first argument must be a command string");
    mxGetString(prhs[0], cmd, sizeof(cmd));
    if (!strcmp(cmd, "reset")) { cleanup(); return; }
    irc_options o;
    if (!strcmp(cmd, "knn_drift") || !strcmp(cmd, "delta_drift")) {
        const int is_knn = !strcmp(cmd, "knn_drift");
        need(nrhs >= 6 && mxIsLogical(prhs[1]) && mxIsSingle(prhs[2]), "bad arguments");
        const mxArray* F = prhs[2];
        const mxArray* drift = prhs[is_knn ? 3 : 4];
        const mxArray* link = prhs[is_knn ? 4 : 5];
        const mwSize n_cols = mxGetN(F);
        need(numel(prhs[1]) == n_cols, "numel(vl_in) must equal size(mrFet_out, 2)");
        need(numel(drift) == n_cols, "numel(viDrift_out) must equal size(mrFet_out, 2)");
        need(mxGetM(link) == mxGetN(link) && mxGetM(link) > 0, "mlDrift1 must be square");
        irc_site_problem pr;
        memset(&pr, 0, sizeof(pr));
        pr.D = (int32_t)mxGetM(F);
        pr.n_cols = (int64_t)n_cols;
        pr.F = (const float*)mxGetData(F);
        in_range(prhs[1], &pr.in_begin, &pr.n_in);
        int32_t* dr = to_i32(drift);
        uint8_t* lk = to_u8(link);
        pr.drift_of_col = dr;
        pr.n_drift = (int32_t)mxGetM(link);
        pr.link = lk;
        if (is_knn) {
            pr.k = (int32_t)mxGetScalar(prhs[5]);
            need(pr.k >= 1, "knn must be positive");
            opts_from(&o, nrhs, prhs, 6);
            plhs[0] = mxCreateNumericMatrix((mwSize)pr.n_in, 1, mxSINGLE_CLASS, mxREAL);
            plhs[1] = mxCreateNumericMatrix((mwSize)pr.k, (mwSize)pr.n_in, mxINT32_CLASS, mxREAL);
            check(irc_knn_drift(ctx(), &pr, &o, (float*)mxGetData(plhs[0]), (int32_t*)mxGetData(plhs[1])));
        } else {
            need(mxIsSingle(prhs[3]), "vrRho_out must be single");
            need(numel(prhs[3]) == n_cols, "numel(vrRho_out) must equal size(mrFet_out, 2)");
            pr.k = 1;
            opts_from(&o, nrhs, prhs, 6);
            plhs[0] = mxCreateNumericMatrix((mwSize)pr.n_in, 1, mxSINGLE_CLASS, mxREAL);
            plhs[1] = mxCreateNumericMatrix((mwSize)pr.n_in, 1, mxINT32_CLASS, mxREAL);
            check(irc_delta_drift(ctx(), &pr, &o, (const float*)mxGetData(prhs[3]), (float*)mxGetData(plhs[0]),
                                  (int32_t*)mxGetData(plhs[1])));
        }
        mxFree(dr);
        mxFree(lk);
        return;
    }
    if (!strcmp(cmd, "rho_delta_page")) {
        irc_batch b;
        void* tmp[5] = {0};
        page_batch(nrhs, prhs, &b, tmp);
        opts_from(&o, nrhs, prhs, 8);
        plhs[0] = mxCreateNumericMatrix((mwSize)b.n_spikes, 1, mxSINGLE_CLASS, mxREAL);
        plhs[1] = mxCreateNumericMatrix((mwSize)b.n_spikes, 1, mxSINGLE_CLASS, mxREAL);
        plhs[2] = mxCreateNumericMatrix((mwSize)b.n_spikes, 1, mxINT64_CLASS, mxREAL);
        irc_outputs out;
        memset(&out, 0, sizeof(out));
        out.rho = (float*)mxGetData(plhs[0]);
        out.delta = (float*)mxGetData(plhs[1]);
        out.nneigh = (int64_t*)mxGetData(plhs[2]);
        if (nlhs > 3) {
            plhs[3] = mxCreateNumericMatrix((mwSize)b.k, (mwSize)b.n_spikes, mxINT64_CLASS, mxREAL);
            out.knn = (int64_t*)mxGetData(plhs[3]);
        }
        const irc_status st = irc_rho_delta_batch(ctx(), &b, &o, &out);
        for (int i = 0; i < 5; ++i) mxFree(tmp[i]);
        check(st);
        return;
    }
    if (!strcmp(cmd, "autotune")) {
        irc_batch b;
        void* tmp[5] = {0};
        page_batch(nrhs, prhs, &b, tmp);
        opts_from(&o, nrhs, prhs, 8);
        irc_outputs out;
        memset(&out, 0, sizeof(out));
        out.rho = (float*)mxMalloc(sizeof(float) * (size_t)b.n_spikes);
        out.delta = (float*)mxMalloc(sizeof(float) * (size_t)b.n_spikes);
        out.nneigh = (int64_t*)mxMalloc(sizeof(int64_t) * (size_t)b.n_spikes);
        irc_options best;
        irc_tune_report rep;
        memset(&rep, 0, sizeof(rep));
        const irc_status st = irc_autotune(ctx(), &b, &o, &out, &best, &rep);
        for (int i = 0; i < 5; ++i) mxFree(tmp[i]);
        mxFree(out.rho); mxFree(out.delta); mxFree(out.nneigh);
        check(st);
        plhs[0] = mxCreateNumericMatrix(1, 6, mxDOUBLE_CLASS, mxREAL);
        double* v = mxGetPr(plhs[0]);
        v[0] = best.precision; v[1] = best.flags; v[2] = best.k_over;
        v[3] = best.tile; v[4] = best.waves_per_cu; v[5] = best.large_ncand;
        if (nlhs > 1) {
            const mwSize n = (mwSize)rep.n_trials;
            plhs[1] = mxCreateNumericMatrix(n, 7, mxDOUBLE_CLASS, mxREAL);
            double* t = mxGetPr(plhs[1]);
            for (mwSize i = 0; i < n; ++i) {
                t[i] = rep.k_over[i]; t[i + n] = rep.waves_per_cu[i]; t[i + 2 * n] = rep.tile[i];
                t[i + 3 * n] = rep.large_ncand[i]; t[i + 4 * n] = rep.ms[i]; t[i + 5 * n] = rep.identical[i];
                t[i + 6 * n] = (int32_t)i == rep.best_trial;
            }
        }
        return;
    }
    if (!strcmp(cmd, "knn_ab") || !strcmp(cmd, "knn_ab_full") || !strcmp(cmd, "knn_ab_ib")) {
        const int ib = !strcmp(cmd, "knn_ab_ib"), full = !strcmp(cmd, "knn_ab_full");
        const int ia = ib ? 3 : 2;   /* position of mrFet1 */
        need(nrhs >= ia + 2 && mxIsSingle(prhs[1]) && mxIsSingle(prhs[ia]), "bad arguments");
        need(mxGetM(prhs[1]) == mxGetM(prhs[ia]), "mrFet2 and mrFet1 must have the same number of rows");
        const int32_t D = (int32_t)mxGetM(prhs[1]);
        const int64_t nB = (int64_t)mxGetN(prhs[1]), nA = (int64_t)mxGetN(prhs[ia]);
        const int64_t nC = ib ? (int64_t)numel(prhs[2]) : nB;   /* candidates */
        need(nB >= 1 && nC >= 1, "no candidates");
        int32_t k = (int32_t)mxGetScalar(prhs[ia + 1]);
        need(k >= 1, "knn must be positive");
        if (k > nC) k = (int32_t)nC;
        plhs[0] = mxCreateNumericMatrix(full ? (mwSize)k : (mwSize)nA, full ? (mwSize)nA : 1, mxSINGLE_CLASS, mxREAL);
        plhs[1] = mxCreateNumericMatrix((mwSize)k, (mwSize)nA, mxINT32_CLASS, mxREAL);
        const float* B = (const float*)mxGetData(prhs[1]);
        const float* A = (const float*)mxGetData(prhs[ia]);
        float* K = (float*)mxGetData(plhs[0]);
        int32_t* I = (int32_t*)mxGetData(plhs[1]);
        if (ib) {
            int32_t* vb = to_i32(prhs[2]);
            const irc_status st = irc_knn_ab_ib(ctx(), D, nB, B, nC, vb, nA, A, k, K, I);
            mxFree(vb);
            check(st);
        } else if (full) {
            check(irc_knn_ab_full(ctx(), D, nB, B, nA, A, k, K, I));
        } else {
            check(irc_knn_ab(ctx(), D, nB, B, nA, A, k, K, I));
        }
        return;
    }
    mexErrMsgIdAndTxt("irc:hip:arg", "unknown command '%s'", cmd);
}
