// Shared device/host helpers for the gfx950 Helmholtz engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/wust_hip.h"

namespace wust {

// ---------------------------------------------------------------------------
// complex arithmetic (interleaved re/im; float2 / double2 storage)
// ---------------------------------------------------------------------------
struct __attribute__((aligned(8))) cf {
    float re, im;
};
struct __attribute__((aligned(16))) cd {
    double re, im;
};

__host__ __device__ __forceinline__ cf mk(float r, float i) { return cf{r, i}; }
__host__ __device__ __forceinline__ cd mkd(double r, double i) { return cd{r, i}; }

__host__ __device__ __forceinline__ cf operator+(cf a, cf b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ __forceinline__ cf operator-(cf a, cf b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ __forceinline__ cf operator-(cf a) { return {-a.re, -a.im}; }
__host__ __device__ __forceinline__ cf operator*(cf a, cf b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ __forceinline__ cf conjf_(cf a) { return {a.re, -a.im}; }
// a*b with its rounding fixed by the source: each part rounds its first product, then adds the
// second by one fma (what cfma gives from zero).  Kernels whose results must agree bit for bit use
// it where operator* would leave the choice of which product to fuse to the compiler.
__host__ __device__ __forceinline__ cf cmul_fixed(cf a, cf b) {
    return {fmaf(-a.im, b.im, a.re * b.re), fmaf(a.im, b.re, a.re * b.im)};
}
__host__ __device__ __forceinline__ float abs2(cf a) { return a.re * a.re + a.im * a.im; }
// c += a*b
__host__ __device__ __forceinline__ void cfma(cf &c, cf a, cf b) {
    c.re = fmaf(a.re, b.re, c.re);
    c.re = fmaf(-a.im, b.im, c.re);
    c.im = fmaf(a.re, b.im, c.im);
    c.im = fmaf(a.im, b.re, c.im);
}
__host__ __device__ __forceinline__ cf crecip(cf a) {
    float s = 1.0f / (a.re * a.re + a.im * a.im);
    return {a.re * s, -a.im * s};
}

// Complex reciprocal with the hardware v_rcp_f32 (1 ulp); used on latency-critical
// pivot chains whose result is corrected by FP64 refinement.
__device__ __forceinline__ cf crecip_fast(cf a) {
#if defined(__HIP_DEVICE_COMPILE__)
    float s = __builtin_amdgcn_rcpf(a.re * a.re + a.im * a.im);
#else
    float s = 1.0f / (a.re * a.re + a.im * a.im);
#endif
    return {a.re * s, -a.im * s};
}

// Workgroup barrier that orders LDS only: it waits for this wave's LDS
// operations (lgkmcnt) but not for global loads still in flight, which the
// HIP __syncthreads() fence would drain.  The "memory" clobber keeps the
// compiler from moving shared-memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
#endif
}

__host__ __device__ __forceinline__ cd operator+(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ __forceinline__ cd operator-(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ __forceinline__ cd operator-(cd a) { return {-a.re, -a.im}; }
__host__ __device__ __forceinline__ cd operator*(cd a, cd b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ __forceinline__ cd operator*(double s, cd a) { return {s * a.re, s * a.im}; }
__host__ __device__ __forceinline__ cd conjd(cd a) { return {a.re, -a.im}; }
__host__ __device__ __forceinline__ double abs2(cd a) { return a.re * a.re + a.im * a.im; }
__host__ __device__ __forceinline__ cd cdiv(cd a, cd b) {
    // Smith's algorithm (robust complex division).
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, den = b.re + b.im * r;
        return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
    }
    double r = b.re / b.im, den = b.re * r + b.im;
    return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}
__host__ __device__ __forceinline__ cf tof(cd a) { return {(float)a.re, (float)a.im}; }
__host__ __device__ __forceinline__ cd tod(cf a) { return {(double)a.re, (double)a.im}; }

// ---------------------------------------------------------------------------
// stencil direction indices of the per-node coefficient arrays
// row (y, x) couples to (y+dy, x+dx):
//   C(0,0) N(-1,0) S(+1,0) W(0,-1) NW(-1,-1) SW(+1,-1) E(0,+1) NE(-1,+1) SE(+1,+1)
// In block-tridiagonal terms over grid columns: D_j = tridiag(N, C, S),
// L_j (to column j-1) = tridiag(NW, W, SW), U_j (to column j+1) = tridiag(NE, E, SE).
// ---------------------------------------------------------------------------
enum Dir { DC = 0, DN = 1, DS = 2, DW = 3, DNW = 4, DSW = 5, DE = 6, DNE = 7, DSE = 8 };

// device status word bits (set by kernels, read by the host after a sync)
enum StatusBits : uint32_t {
    ST_ZERO_PIVOT = 1u,
    ST_NONFINITE_PIVOT = 2u,
    ST_SCALE_FAILURE = 4u,
    ST_NONFINITE = 8u,
    ST_INVERSE_UNCONVERGED = 16u,   // a Newton-Schulz block inverse missed its residual bound
};

// ---------------------------------------------------------------------------
// host-side error plumbing
// ---------------------------------------------------------------------------
struct Error : public std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

#define WUST_HIP_CHECK(expr)                                                                      \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            throw ::wust::Error(WUST_HIP_ERROR, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

#ifdef WUST_DEBUG_SYNC
// Debug build: every launch is followed by a device synchronization, so an asynchronous failure
// (a fault, an abort) is reported at the launch that caused it, with its file and line.
#define WUST_LAUNCH_CHECK()                                                                            \
    do {                                                                                               \
        WUST_HIP_CHECK(hipGetLastError());                                                             \
        const hipError_t _s = hipDeviceSynchronize();                                                  \
        if (_s != hipSuccess)                                                                          \
            throw ::wust::Error(WUST_HIP_ERROR, std::string(__FILE__ ":") + std::to_string(__LINE__) + \
                                                    ": kernel failed: " + hipGetErrorString(_s));      \
    } while (0)
#else
#define WUST_LAUNCH_CHECK() WUST_HIP_CHECK(hipGetLastError())
#endif

inline void require(bool ok, const std::string &msg) {
    if (!ok) throw Error(WUST_INVALID_INPUT, msg);
}

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

}  // namespace wust
