/*
 * mex.h -- a minimal stand-in for the subset of MATLAB's MEX C API that
 * ironclust_amd/csrc/mex/irc_hip_mex.c uses, so the gateway can be compiled and driven by the
 * tests without MATLAB.  Test infrastructure only: real builds use MATLAB's own mex.h.
 * mexErrMsgIdAndTxt longjmps back to mock_call(), which reports the error identifier.
 */
#ifndef IRC_MEX_MOCK_H
#define IRC_MEX_MOCK_H
#include <stddef.h>
#include <stdint.h>

typedef size_t mwSize;
typedef unsigned char mxLogical;
typedef enum { mxLOGICAL_CLASS, mxCHAR_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT32_CLASS, mxINT64_CLASS } mxClassID;
typedef enum { mxREAL } mxComplexity;
typedef struct mxArray_tag {
    mxClassID cls;
    mwSize m, n;
    void* data;
    char* str;
} mxArray;

mwSize mxGetNumberOfElements(const mxArray* a);
mwSize mxGetM(const mxArray* a);
mwSize mxGetN(const mxArray* a);
int mxIsChar(const mxArray* a);
int mxIsLogical(const mxArray* a);
int mxIsSingle(const mxArray* a);
int mxIsDouble(const mxArray* a);
int mxIsInt32(const mxArray* a);
int mxIsInt64(const mxArray* a);
void* mxGetData(const mxArray* a);
double* mxGetPr(const mxArray* a);
mxLogical* mxGetLogicals(const mxArray* a);
double mxGetScalar(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize len);
void* mxMalloc(size_t n);
void mxFree(void* p);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexAtExit(void (*f)(void));
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
#endif
