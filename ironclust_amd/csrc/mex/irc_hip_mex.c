/*
 * irc_hip_mex.c -- MATLAB MEX gateway over the C ABI of libirc_hip.so (include/irc_hip.h).
 * It holds no logic: it maps mxArrays to the ABI structs and errors to
 * mexErrMsgIdAndTxt('irc:hip', ...), which the caller's try/catch turns into its CPU
 * fallback (matlab/irc2.m:3643-3652, 3694-3702).  Build where MATLAB's mex.h exists:
 *   mex -R2018a irc_hip_mex.c -I<repo>/include -L<repo>/ironclust_amd/_lib -lirc_hip
 *
 * Commands (1-based indices in and out, MATLAB orientation):
 *   [vrKnn_in, miKnn_in]    = irc_hip_mex('knn_drift', vl_in, mrFet_out, viDrift_out, mlDrift1, knn)
 *                              replaces search_knn_drift_ (matlab/irc2.m:3731-3789)
 *   [vrDelta_in, viNneigh_in] = irc_hip_mex('delta_drift', vl_in, mrFet_out, vrRho_out, viDrift_out, mlDrift1)
 *                              replaces search_delta_drift_ (matlab/irc2.m:3794-3865)
 *   [vrRho, vrDelta, viNneigh, miKnn] = irc_hip_mex('rho_delta_page', mrFet, viSpk_col, viSite_ptr,
 *                              vnPrimary_site, viLim_drift, mlDrift, knn)
 *                              replaces rho_page_ + delta_page_ (matlab/irc2.m:3524-3607)
 *   [vrKnn, miKnn]          = irc_hip_mex('knn_ab', mrFet2, mrFet1, knn)   replaces cuda_knn_ (irc2.m:4202)
 *   irc_hip_mex('reset')   destroys the persistent context
 * Optional trailing argument: precision (0 exact, 1 FP32 screen, 2 BF16 screen, 3 auto).
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

static int32_t* to_i32(const mxArray* a) {
    const mwSize n = mxGetNumberOfElements(a);
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
    const mwSize n = mxGetNumberOfElements(a);
    int64_t* v = (int64_t*)mxMalloc(sizeof(int64_t) * (n ? n : 1));
    int32_t* t = to_i32(a);
    for (mwSize i = 0; i < n; ++i) v[i] = t[i];
    if (mxIsInt64(a)) for (mwSize i = 0; i < n; ++i) v[i] = ((const int64_t*)mxGetData(a))[i];
    mxFree(t);
    return v;
}

static uint8_t* to_u8(const mxArray* a) {
    const mwSize n = mxGetNumberOfElements(a);
    uint8_t* v = (uint8_t*)mxMalloc(n ? n : 1);
    if (mxIsLogical(a)) { const mxLogical* l = mxGetLogicals(a); for (mwSize i = 0; i < n; ++i) v[i] = l[i] ? 1 : 0; }
    else { int32_t* t = to_i32(a); for (mwSize i = 0; i < n; ++i) v[i] = t[i] != 0; mxFree(t); }
    return v;
}

static void opts_from(irc_options* o, int nrhs, const mxArray* prhs[], int pos) {
    irc_options_default(o);
    if (nrhs > pos) o->precision = (int32_t)mxGetScalar(prhs[pos]);
}

/* contiguous run marked by vl_in -> (in_begin, n_in) */
static void in_range(const mxArray* vl, int64_t* beg, int64_t* n) {
    const mwSize m = mxGetNumberOfElements(vl);
    const mxLogical* l = mxGetLogicals(vl);
    int64_t b = -1, e = -1;
    for (mwSize i = 0; i < m; ++i) if (l[i]) { if (b < 0) b = (int64_t)i; e = (int64_t)i; }
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
        irc_site_problem pr;
        memset(&pr, 0, sizeof(pr));
        pr.D = (int32_t)mxGetM(F);
        pr.n_cols = (int64_t)mxGetN(F);
        pr.F = (const float*)mxGetData(F);
        int32_t* dr = to_i32(drift);
        uint8_t* lk = to_u8(link);
        pr.drift_of_col = dr;
        pr.n_drift = (int32_t)mxGetM(link);
        pr.link = lk;
        in_range(prhs[1], &pr.in_begin, &pr.n_in);
        if (is_knn) {
            pr.k = (int32_t)mxGetScalar(prhs[5]);
            opts_from(&o, nrhs, prhs, 6);
            plhs[0] = mxCreateNumericMatrix((mwSize)pr.n_in, 1, mxSINGLE_CLASS, mxREAL);
            plhs[1] = mxCreateNumericMatrix((mwSize)pr.k, (mwSize)pr.n_in, mxINT32_CLASS, mxREAL);
            check(irc_knn_drift(ctx(), &pr, &o, (float*)mxGetData(plhs[0]), (int32_t*)mxGetData(plhs[1])));
        } else {
            need(mxIsSingle(prhs[3]), "vrRho_out must be single");
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
        need(nrhs >= 8 && mxIsSingle(prhs[1]), "bad arguments");
        irc_batch b;
        memset(&b, 0, sizeof(b));
        b.D = (int32_t)mxGetM(prhs[1]);
        b.n_cols = (int64_t)mxGetN(prhs[1]);
        b.F = (const float*)mxGetData(prhs[1]);
        int32_t* spk = to_i32(prhs[2]);
        int64_t* sp = to_i64(prhs[3]);
        int64_t* np = to_i64(prhs[4]);
        int64_t* lim = to_i64(prhs[5]);
        uint8_t* lk = to_u8(prhs[6]);
        b.col_spike = spk;
        b.site_ptr = sp;
        b.site_n_primary = np;
        b.n_sites = (int32_t)mxGetNumberOfElements(prhs[4]);
        b.lim_drift = lim;
        b.n_drift = (int32_t)mxGetNumberOfElements(prhs[5]) - 1;
        b.link = lk;
        b.k = (int32_t)mxGetScalar(prhs[7]);
        b.n_spikes = 0;
        for (int32_t s = 0; s < b.n_sites; ++s) b.n_spikes += np[s];
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
            plhs[3] = mxCreateNumericMatrix((mwSize)b.k, (mwSize)b.n_spikes, mxINT32_CLASS, mxREAL);
            out.knn = (int32_t*)mxGetData(plhs[3]);
        }
        check(irc_rho_delta_batch(ctx(), &b, &o, &out));
        mxFree(spk); mxFree(sp); mxFree(np); mxFree(lim); mxFree(lk);
        return;
    }
    if (!strcmp(cmd, "knn_ab")) {
        need(nrhs >= 4 && mxIsSingle(prhs[1]) && mxIsSingle(prhs[2]), "bad arguments");
        const int32_t D = (int32_t)mxGetM(prhs[1]);
        const int64_t nB = (int64_t)mxGetN(prhs[1]), nA = (int64_t)mxGetN(prhs[2]);
        int32_t k = (int32_t)mxGetScalar(prhs[3]);
        if (k > nB) k = (int32_t)nB;
        plhs[0] = mxCreateNumericMatrix((mwSize)nA, 1, mxSINGLE_CLASS, mxREAL);
        plhs[1] = mxCreateNumericMatrix((mwSize)k, (mwSize)nA, mxINT32_CLASS, mxREAL);
        check(irc_knn_ab(ctx(), D, nB, (const float*)mxGetData(prhs[1]), nA, (const float*)mxGetData(prhs[2]), k,
                         (float*)mxGetData(plhs[0]), (int32_t*)mxGetData(plhs[1])));
        return;
    }
    mexErrMsgIdAndTxt("irc:hip:arg", "unknown command '%s'", cmd);
}
