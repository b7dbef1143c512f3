// NCG step around the oracle, FP64 on the device: ROI masking and the radial
// Butterworth filter (Runtime/matlab/wust_roi_gradient.m:1-10,
// Runtime/solver/ringingRemovalFilt.m:1-45), the clipped PR/FR direction
// (Runtime/matlab/wust_ncg_direction.m:1-17), the linearized step length
// (Runtime/matlab/linearized_step_length.m:1-11) and the bounded slowness
// update (Runtime/matlab/wust_project_slowness_update.m:1-10).
//
// Grid arrays are MATLAB [Ny, Nx] column-major (y fastest).  The filter runs
// in the natural FFT order: bin j of an axis of n points spaced h carries the
// frequency (j < (n+1)/2 ? j : j - n) / (n h), which is where the reference's
// fftshift / wrapped k grid / ifftshift sequence puts each frequency; only
// its square enters the filter.  Reductions are two-stage (per-block partial
// sums, then one block), so results do not depend on atomic ordering.
#include <hipfft/hipfft.h>

#include <map>
#include <mutex>
#include <utility>

#include "ncg.h"

namespace wust {

namespace {

constexpr int RB = 256;   // threads per block of the elementwise / reduction kernels

__global__ void roi_kernel(const double *g, const uint8_t *mask, int64_t n, double *roi, double2 *zc) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = mask[i] ? g[i] : 0.0;
    roi[i] = v;
    if (zc) zc[i] = make_double2(v, 0.0);
}

// zc holds [Nx][Ny] (y fastest) = MATLAB [Ny, Nx]; scale folds in the 1/(Nx Ny) of the inverse
__global__ void butterworth_kernel(double2 *zc, int ny, int nx, double dx, double dy, double kcut2, double order,
                                   double scale) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)ny * nx) return;
    const int iy = (int)(i % ny), ix = (int)(i / ny);
    const double fy = (iy < (ny + 1) / 2 ? iy : iy - ny) / (ny * dy);
    const double fx = (ix < (nx + 1) / 2 ? ix : ix - nx) / (nx * dx);
    const double h = scale / (1.0 + pow((fx * fx + fy * fy) / kcut2, order));
    zc[i] = make_double2(zc[i].x * h, zc[i].y * h);
}

__global__ void unmask_real_kernel(const double2 *zc, const uint8_t *mask, int64_t n, double *z) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    z[i] = mask[i] ? zc[i].x : 0.0;
}

// block sums of K products: out[block][k] = sum_i a_k[i] b_k[i]
template <int K>
__global__ void __launch_bounds__(RB) dots_kernel(const double *const *a, const double *const *b, int64_t n,
                                                  double *partial) {
    __shared__ double sh[K][RB];
    const int64_t i0 = blockIdx.x * (int64_t)RB * 4 + threadIdx.x;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t i = i0 + (int64_t)u * RB;
        if (i < n) {
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] += a[k][i] * b[k][i];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < K; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < K) partial[blockIdx.x * K + threadIdx.x] = sh[threadIdx.x][0];
}

// one block: out[k] = sum_b partial[b][k]
template <int K>
__global__ void __launch_bounds__(RB) finish_kernel(const double *partial, int nblocks, double *out) {
    __shared__ double sh[K][RB];
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += RB) {
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += partial[b * K + k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < K; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < K) out[threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void direction_kernel(const double *z, const double *prev_dir, double beta, int64_t n, double *d,
                                 uint32_t *bad) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = (beta != 0.0) ? -z[i] + beta * prev_dir[i] : -z[i];
    d[i] = v;
    if (!isfinite(v)) atomicOr(bad, 1u);
}

// c_new and, per block, [sum (s_new - s_old)^2, sum s_old^2] over the update mask
__global__ void __launch_bounds__(RB) project_kernel(const double *c, const double *d, double step, double max_up,
                                                     double lo, double hi, const uint8_t *mask, int64_t n,
                                                     double *c_new, double *partial, uint32_t *bad) {
    __shared__ double sh[2][RB];
    const int64_t i = blockIdx.x * (int64_t)RB + threadIdx.x;
    double num = 0.0, den = 0.0;
    if (i < n) {
        const double ci = c[i];
        const double c_lo = fmax(ci - max_up, lo), c_hi = fmin(ci + max_up, hi);
        const double trial = 1.0 / ci + step * d[i];
        if (!isfinite(trial)) atomicOr(bad, 1u);
        const double s_new = fmax(1.0 / c_hi, fmin(1.0 / c_lo, trial));
        if (mask[i]) {
            const double cn = 1.0 / s_new;
            c_new[i] = cn;
            const double s_old = 1.0 / ci, ds = 1.0 / cn - s_old;
            num = ds * ds;
            den = s_old * s_old;
        } else {
            c_new[i] = ci;
        }
    }
    sh[0][threadIdx.x] = num;
    sh[1][threadIdx.x] = den;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2) partial[blockIdx.x * 2 + threadIdx.x] = sh[threadIdx.x][0];
}

// 2-D complex-double FFT plans by (ny, nx), created once per process and shape
hipfftHandle fft_plan(int ny, int nx) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, hipfftHandle> plans;
    std::lock_guard<std::mutex> lk(mu);
    auto it = plans.find({ny, nx});
    if (it != plans.end()) return it->second;
    hipfftHandle p;
    // row-major [Nx][Ny]: Ny is the contiguous (fastest) dimension
    if (hipfftPlan2d(&p, nx, ny, HIPFFT_Z2Z) != HIPFFT_SUCCESS)
        throw Error(WUST_LIBRARY_ERROR, "hipfftPlan2d failed");
    plans[{ny, nx}] = p;
    return p;
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    explicit DevBuf(int64_t n) { WUST_HIP_CHECK(hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1))); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

unsigned blocks_of(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void ncg_roi_gradient(const double *g, const uint8_t *mask, int ny, int nx, double dx, double dy, double c0,
                      double f, double cutoff, double order, double *roi, double *z, hipStream_t s) {
    const int64_t n = (int64_t)ny * nx;
    const unsigned nb = blocks_of(n, RB);
    if (cutoff == 0.0) {
        hipLaunchKernelGGL(roi_kernel, dim3(nb), dim3(RB), 0, s, g, mask, n, roi, (double2 *)nullptr);
        WUST_LAUNCH_CHECK();
        WUST_HIP_CHECK(hipMemcpyAsync(z, roi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        return;
    }
    DevBuf<double2> zc(n);
    hipLaunchKernelGGL(roi_kernel, dim3(nb), dim3(RB), 0, s, g, mask, n, roi, zc.p);
    WUST_LAUNCH_CHECK();
    hipfftHandle plan = fft_plan(ny, nx);
    if (hipfftSetStream(plan, s) != HIPFFT_SUCCESS) throw Error(WUST_LIBRARY_ERROR, "hipfftSetStream failed");
    auto *zz = reinterpret_cast<hipfftDoubleComplex *>(zc.p);
    if (hipfftExecZ2Z(plan, zz, zz, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        throw Error(WUST_LIBRARY_ERROR, "hipfftExecZ2Z (forward) failed");
    const double kcut = 2.0 * f / c0 * cutoff;
    hipLaunchKernelGGL(butterworth_kernel, dim3(nb), dim3(RB), 0, s, zc.p, ny, nx, dx, dy, kcut * kcut, order,
                       1.0 / (double)n);
    WUST_LAUNCH_CHECK();
    if (hipfftExecZ2Z(plan, zz, zz, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
        throw Error(WUST_LIBRARY_ERROR, "hipfftExecZ2Z (inverse) failed");
    hipLaunchKernelGGL(unmask_real_kernel, dim3(nb), dim3(RB), 0, s, zc.p, mask, n, z);
    WUST_LAUNCH_CHECK();
    WUST_HIP_CHECK(hipStreamSynchronize(s));   // zc is freed on return
}

void ncg_direction(const double *g, const double *z, const double *prev_z, double prev_gz, const double *prev_dir,
                   bool same_frequency, int64_t n, double *direction, double *beta_out, double *gz_out,
                   double *slope_out, hipStream_t s) {
    // gz = <g, z>, gp = <g, prev_z>, gd = <g, prev_dir> in one pass
    const int64_t per = (int64_t)RB * 4;
    const unsigned nb = blocks_of(n, per);
    DevBuf<double> partial(3 * (int64_t)nb + 3);
    DevBuf<const double *> ptrs(6);
    DevBuf<uint32_t> bad(1);
    const double *h_ptrs[6] = {g, g, g, z, prev_z ? prev_z : z, prev_dir ? prev_dir : z};
    WUST_HIP_CHECK(hipMemcpyAsync(ptrs.p, h_ptrs, sizeof(h_ptrs), hipMemcpyHostToDevice, s));
    WUST_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(dots_kernel<3>, dim3(nb), dim3(RB), 0, s, ptrs.p, ptrs.p + 3, n, partial.p);
    WUST_LAUNCH_CHECK();
    hipLaunchKernelGGL(finish_kernel<3>, dim3(1), dim3(RB), 0, s, partial.p, (int)nb, partial.p + 3 * nb);
    WUST_LAUNCH_CHECK();
    double sums[3];
    WUST_HIP_CHECK(hipMemcpyAsync(sums, partial.p + 3 * nb, sizeof(sums), hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    const double gz = sums[0];
    double beta = 0.0;
    if (same_frequency && prev_gz > 0) {
        const double beta_fr = gz / prev_gz;
        const double beta_pr = (gz - sums[1]) / prev_gz;
        beta = std::min(std::max(beta_pr, 0.0), beta_fr);
    }
    // <g, -z + beta d> = -gz + beta gd; restart with -z when it does not descend
    double slope = -gz + beta * sums[2];
    if (slope >= 0) {
        beta = 0.0;
        slope = -gz;
    }
    hipLaunchKernelGGL(direction_kernel, dim3(blocks_of(n, RB)), dim3(RB), 0, s, z, prev_dir, beta, n, direction,
                       bad.p);
    WUST_LAUNCH_CHECK();
    uint32_t hbad = 0;
    WUST_HIP_CHECK(hipMemcpyAsync(&hbad, bad.p, sizeof(hbad), hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    if (hbad || !std::isfinite(gz) || !(gz >= 0))
        throw Error(WUST_NUMERICAL_FAILURE, "Invalid preconditioned gradient");
    *beta_out = beta;
    *gz_out = gz;
    if (slope_out) *slope_out = slope;
}

void ncg_project(const double *c, const double *direction, double alpha, double step_damping, double max_update,
                 double lo, double hi, const uint8_t *mask, int64_t n, double *c_new, double *rel_update,
                 hipStream_t s) {
    const unsigned nb = blocks_of(n, RB);
    DevBuf<double> partial(2 * (int64_t)nb + 2);
    DevBuf<uint32_t> bad(1);
    WUST_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(project_kernel, dim3(nb), dim3(RB), 0, s, c, direction, step_damping * alpha, max_update, lo,
                       hi, mask, n, c_new, partial.p, bad.p);
    WUST_LAUNCH_CHECK();
    hipLaunchKernelGGL(finish_kernel<2>, dim3(1), dim3(RB), 0, s, partial.p, (int)nb, partial.p + 2 * nb);
    WUST_LAUNCH_CHECK();
    double sums[2];
    uint32_t hbad = 0;
    WUST_HIP_CHECK(hipMemcpyAsync(sums, partial.p + 2 * nb, sizeof(sums), hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipMemcpyAsync(&hbad, bad.p, sizeof(hbad), hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    if (hbad) throw Error(WUST_NUMERICAL_FAILURE, "Nonfinite slowness trial");
    if (rel_update) *rel_update = std::sqrt(sums[0]) / std::max(std::sqrt(sums[1]), 2.2250738585072014e-308);
}

}  // namespace wust
