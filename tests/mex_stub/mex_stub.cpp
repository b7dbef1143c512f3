// Host-heap implementation of the stub matrix.h / mex.h, plus a C harness (stub_*) through which the
// tests build MATLAB-layout arrays, call mexFunction and read its outputs with ctypes.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls = mxDOUBLE_CLASS;
    bool complex = false;
    std::vector<mwSize> dims;
    std::vector<unsigned char> data;   // column-major; complex interleaved (re, im)
    size_t numel() const {
        size_t n = 1;
        for (mwSize d : dims) n *= d;
        return n;
    }
};

namespace {
struct StubMexError {
    std::string id, msg;
};
size_t elem_bytes(mxClassID c, bool cx) {
    switch (c) {
        case mxDOUBLE_CLASS: return cx ? 16 : 8;
        case mxUINT64_CLASS: return 8;
        case mxLOGICAL_CLASS: return 1;
        case mxCHAR_CLASS: return 2;
        default: return 8;
    }
}
mxArray *make(mxClassID cls, bool cx, std::vector<mwSize> dims) {
    if (dims.size() < 2) dims.resize(2, 1);
    mxArray *a = new mxArray_tag();
    a->cls = cls;
    a->complex = cx;
    a->dims = dims;
    a->data.assign(a->numel() * elem_bytes(cls, cx), 0);
    return a;
}
}  // namespace

extern "C" {

mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->dims.size(); }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims.data(); }
size_t mxGetNumberOfElements(const mxArray *a) { return a->numel(); }
bool mxIsChar(const mxArray *a) { return a->cls == mxCHAR_CLASS; }
bool mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsLogical(const mxArray *a) { return a->cls == mxLOGICAL_CLASS; }
bool mxIsUint64(const mxArray *a) { return a->cls == mxUINT64_CLASS; }
bool mxIsComplex(const mxArray *a) { return a->complex; }
double mxGetScalar(const mxArray *a) {
    if (a->numel() == 0) return 0.0;
    switch (a->cls) {
        case mxDOUBLE_CLASS: return reinterpret_cast<const double *>(a->data.data())[0];
        case mxLOGICAL_CLASS: return a->data[0] ? 1.0 : 0.0;
        case mxUINT64_CLASS: return (double)reinterpret_cast<const uint64_t *>(a->data.data())[0];
        default: return 0.0;
    }
}
double *mxGetDoubles(const mxArray *a) {
    return (a->cls == mxDOUBLE_CLASS && !a->complex) ? reinterpret_cast<double *>(const_cast<unsigned char *>(a->data.data())) : nullptr;
}
mxComplexDouble *mxGetComplexDoubles(const mxArray *a) {
    return (a->cls == mxDOUBLE_CLASS && a->complex)
               ? reinterpret_cast<mxComplexDouble *>(const_cast<unsigned char *>(a->data.data()))
               : nullptr;
}
mxLogical *mxGetLogicals(const mxArray *a) {
    return a->cls == mxLOGICAL_CLASS ? reinterpret_cast<mxLogical *>(const_cast<unsigned char *>(a->data.data())) : nullptr;
}
uint64_t *mxGetUint64s(const mxArray *a) {
    return a->cls == mxUINT64_CLASS ? reinterpret_cast<uint64_t *>(const_cast<unsigned char *>(a->data.data())) : nullptr;
}
char *mxArrayToString(const mxArray *a) {
    if (a->cls != mxCHAR_CLASS) return nullptr;
    const size_t n = a->numel();
    char *s = static_cast<char *>(std::malloc(n + 1));
    const char16_t *c = reinterpret_cast<const char16_t *>(a->data.data());
    for (size_t i = 0; i < n; ++i) s[i] = (char)c[i];
    s[n] = 0;
    return s;
}
void mxFree(void *p) { std::free(p); }
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) { return make(mxDOUBLE_CLASS, c == mxCOMPLEX, {m, n}); }
mxArray *mxCreateDoubleScalar(double v) {
    mxArray *a = make(mxDOUBLE_CLASS, false, {1, 1});
    std::memcpy(a->data.data(), &v, 8);
    return a;
}
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c) {
    return make(cls, c == mxCOMPLEX, {m, n});
}
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c) {
    return make(cls, c == mxCOMPLEX, std::vector<mwSize>(dims, dims + ndim));
}
void mxDestroyArray(mxArray *a) { delete a; }

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    char buf[4096];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw StubMexError{id, buf};
}

// ---- harness ---------------------------------------------------------------------------------
// class codes: 0 double, 1 logical, 2 char, 3 uint64
void *stub_array(int cls, int complex, int ndim, const size_t *dims, const void *data) {
    const mxClassID c = cls == 1 ? mxLOGICAL_CLASS : cls == 2 ? mxCHAR_CLASS : cls == 3 ? mxUINT64_CLASS : mxDOUBLE_CLASS;
    mxArray *a = make(c, complex != 0, std::vector<mwSize>(dims, dims + ndim));
    if (data && !a->data.empty()) std::memcpy(a->data.data(), data, a->data.size());
    return a;
}
void *stub_string(const char *s) {
    const size_t n = std::strlen(s);
    mxArray *a = make(mxCHAR_CLASS, false, {1, n});
    char16_t *c = reinterpret_cast<char16_t *>(a->data.data());
    for (size_t i = 0; i < n; ++i) c[i] = (char16_t)s[i];
    return a;
}
int stub_describe(const void *p, int *cls, int *complex, int *ndim, size_t *dims, int max_dims) {
    const mxArray *a = static_cast<const mxArray *>(p);
    *cls = a->cls == mxLOGICAL_CLASS ? 1 : a->cls == mxCHAR_CLASS ? 2 : a->cls == mxUINT64_CLASS ? 3 : 0;
    *complex = a->complex ? 1 : 0;
    *ndim = (int)a->dims.size();
    for (int i = 0; i < *ndim && i < max_dims; ++i) dims[i] = a->dims[i];
    return 0;
}
const void *stub_data(const void *p) { return static_cast<const mxArray *>(p)->data.data(); }
size_t stub_bytes(const void *p) { return static_cast<const mxArray *>(p)->data.size(); }
void stub_destroy(void *p) { delete static_cast<mxArray *>(p); }
// 0: returned normally; 1: the gateway raised (id / msg filled); 2: another C++ exception escaped
int stub_call(int nlhs, void **plhs, int nrhs, void **prhs, char *id, char *msg, int cap) {
    try {
        mexFunction(nlhs, reinterpret_cast<mxArray **>(plhs), nrhs, const_cast<const mxArray **>(reinterpret_cast<mxArray **>(prhs)));
        return 0;
    } catch (const StubMexError &e) {
        std::snprintf(id, cap, "%s", e.id.c_str());
        std::snprintf(msg, cap, "%s", e.msg.c_str());
        return 1;
    } catch (const std::exception &e) {
        std::snprintf(id, cap, "%s", "stub:UncaughtException");
        std::snprintf(msg, cap, "%s", e.what());
        return 2;
    }
}

}  // extern "C"
