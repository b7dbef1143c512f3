/* Minimal stand-in for MATLAB's matrix.h (R2018a interleaved-complex API), for compiling and unit-testing
 * the MEX gateway without MATLAB.  Only what waveforminversionust_amd/matlab/wust_hip_mex.cpp uses is
 * declared; the implementation (mex_stub.cpp) keeps every array on the host heap. */
#ifndef WUST_STUB_MATRIX_H
#define WUST_STUB_MATRIX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef bool mxLogical;
typedef struct {
    double real, imag;
} mxComplexDouble;
typedef enum { mxUNKNOWN_CLASS = 0, mxLOGICAL_CLASS = 3, mxCHAR_CLASS = 4, mxDOUBLE_CLASS = 6, mxUINT64_CLASS = 15 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
bool mxIsChar(const mxArray *a);
bool mxIsDouble(const mxArray *a);
bool mxIsLogical(const mxArray *a);
bool mxIsUint64(const mxArray *a);
bool mxIsComplex(const mxArray *a);
double mxGetScalar(const mxArray *a);
double *mxGetDoubles(const mxArray *a);
mxComplexDouble *mxGetComplexDoubles(const mxArray *a);
mxLogical *mxGetLogicals(const mxArray *a);
uint64_t *mxGetUint64s(const mxArray *a);
char *mxArrayToString(const mxArray *a);
void mxFree(void *p);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateDoubleScalar(double v);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c);
void mxDestroyArray(mxArray *a);

#ifdef __cplusplus
}
#endif
#endif
