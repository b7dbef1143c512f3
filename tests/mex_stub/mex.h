/* Minimal stand-in for MATLAB's mex.h: the gateway entry point and the error call.  In the stub,
 * mexErrMsgIdAndTxt throws (the test harness in mex_stub.cpp catches it and reports id and message);
 * MATLAB unwinds the MEX call instead. */
#ifndef WUST_STUB_MEX_H
#define WUST_STUB_MEX_H
#include "matrix.h"

#ifdef __cplusplus
extern "C" {
#endif
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
[[noreturn]] void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
#ifdef __cplusplus
}
#endif
#endif
