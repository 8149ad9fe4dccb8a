/*
 * mex_mock.c -- the MEX API stand-in of mex.h plus a C entry point for ctypes: mock_call() builds
 * mxArrays from plain buffers, runs mexFunction and returns 0 or the identifier of the error it
 * raised.  Test infrastructure only.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

static jmp_buf g_jmp;
static char g_err[512];

static size_t elsize(mxClassID c) {
    switch (c) {
        case mxLOGICAL_CLASS: case mxCHAR_CLASS: return 1;
        case mxSINGLE_CLASS: case mxINT32_CLASS: return 4;
        default: return 8;
    }
}
mwSize mxGetNumberOfElements(const mxArray* a) { return a->m * a->n; }
mwSize mxGetM(const mxArray* a) { return a->m; }
mwSize mxGetN(const mxArray* a) { return a->n; }
int mxIsChar(const mxArray* a) { return a->cls == mxCHAR_CLASS; }
int mxIsLogical(const mxArray* a) { return a->cls == mxLOGICAL_CLASS; }
int mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
int mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsInt32(const mxArray* a) { return a->cls == mxINT32_CLASS; }
int mxIsInt64(const mxArray* a) { return a->cls == mxINT64_CLASS; }
void* mxGetData(const mxArray* a) { return a->data; }
double* mxGetPr(const mxArray* a) { return (double*)a->data; }
mxLogical* mxGetLogicals(const mxArray* a) { return (mxLogical*)a->data; }
double mxGetScalar(const mxArray* a) {
    switch (a->cls) {
        case mxDOUBLE_CLASS: return ((double*)a->data)[0];
        case mxSINGLE_CLASS: return ((float*)a->data)[0];
        case mxINT32_CLASS: return ((int32_t*)a->data)[0];
        case mxINT64_CLASS: return (double)((int64_t*)a->data)[0];
        default: return ((unsigned char*)a->data)[0];
    }
}
int mxGetString(const mxArray* a, char* buf, mwSize len) {
    snprintf(buf, len, "%s", a->str ? a->str : "");
    return 0;
}
void* mxMalloc(size_t n) { return malloc(n); }
void mxFree(void* p) { free(p); }
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c) {
    (void)c;
    mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
    a->cls = cls;
    a->m = m;
    a->n = n;
    a->data = calloc(m * n ? m * n : 1, elsize(cls));
    return a;
}
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    char msg[400];
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    snprintf(g_err, sizeof(g_err), "%s|%s", id, msg);
    longjmp(g_jmp, 1);
}
int mexAtExit(void (*f)(void)) { (void)f; return 0; }

/* Argument i: class cls[i], size m[i] x n[i], data[i] (copied; char arrays take str[i]).
 * Outputs: out_data[j] receives a pointer to the data of plhs[j] (valid until the next call),
 * out_m/out_n its size and out_cls its class.  Returns 0, or 1 with the error text in err. */
int mock_call(int nrhs, const int* cls, const int64_t* m, const int64_t* n, const void* const* data,
              const char* const* str, int nlhs, void** out_data, int64_t* out_m, int64_t* out_n, int* out_cls,
              char* err, int err_len) {
    static mxArray* keep[8];
    for (int j = 0; j < 8; ++j) {
        if (keep[j]) { free(keep[j]->data); free(keep[j]); keep[j] = NULL; }
    }
    mxArray* in[16];
    for (int i = 0; i < nrhs && i < 16; ++i) {
        in[i] = mxCreateNumericMatrix((mwSize)m[i], (mwSize)n[i], (mxClassID)cls[i], mxREAL);
        if (cls[i] == mxCHAR_CLASS) in[i]->str = (char*)str[i];
        else if (data[i]) memcpy(in[i]->data, data[i], (size_t)(m[i] * n[i]) * elsize((mxClassID)cls[i]));
    }
    mxArray* out[8] = {0};
    int rc = 0;
    g_err[0] = 0;
    if (setjmp(g_jmp) == 0) {
        mexFunction(nlhs, out, nrhs, (const mxArray**)in);
    } else {
        rc = 1;
    }
    for (int i = 0; i < nrhs && i < 16; ++i) { free(in[i]->data); free(in[i]); }
    for (int j = 0; j < nlhs && j < 8; ++j) {
        keep[j] = out[j];
        out_data[j] = out[j] ? out[j]->data : NULL;
        out_m[j] = out[j] ? (int64_t)out[j]->m : 0;
        out_n[j] = out[j] ? (int64_t)out[j]->n : 0;
        out_cls[j] = out[j] ? (int)out[j]->cls : -1;
    }
    snprintf(err, (size_t)err_len, "%s", g_err);
    return rc;
}
