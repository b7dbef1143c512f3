// "Plain complex-FP32 block LU": the reference's GPU algorithm, unrefined, on
// AMD's vendor libraries.  It is the comparator the refined engine is measured
// against and it is frozen.  Structure (Runtime/solver/decompBlockLU.cu:337-368,
// Runtime/solver/applyBlockLU.cu:433-467):
//   factor: one serial Schur recurrence over the interior columns; per block a
//           rocSOLVER cgetrf (partial pivoting) and cgetrs against the identity
//           (explicit inverse), then the Schur update kernel;
//   solve:  per column a tridiagonal coupling kernel and one rocBLAS cgemm,
//           forward then backward (adjoint: conjugate-transposed operators).
// Field layout is the reference's per-RHS column-major [R][m][npad] (interior
// rows of one grid column contiguous), so every GEMM is a plain strided call.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "plain.h"
#include "profile.h"

namespace wust {

namespace {

#define ROCBLAS_CHECK(expr)                                                                            \
    do {                                                                                               \
        rocblas_status _s = (expr);                                                                    \
        if (_s != rocblas_status_success)                                                              \
            throw ::wust::Error(WUST_LIBRARY_ERROR, std::string(#expr ": ") + rocblas_status_to_string(_s)); \
    } while (0)

__global__ void identity_kernel(cf *page, int n, int ld) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (r >= n) return;
    page[r + (int64_t)ld * c] = (r == c) ? mk(1.f, 0.f) : mk(0.f, 0.f);
}

// tmp(k, r) = src(k, j, r) - (P v)(k, r), v = u(:, jv, r); P tridiagonal,
// optionally conjugate-transposed.  One thread per (k, r), k fastest.
__global__ void coupling_kernel(const cf *__restrict__ src, const cf *__restrict__ u, cf *__restrict__ tmp,
                                const cf *__restrict__ pn, const cf *__restrict__ pc, const cf *__restrict__ ps,
                                int transposed, int has_src, int has_term, int j, int jv, int n, int npad, int m,
                                int R) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (k >= n || r >= R) return;
    const int64_t col = (int64_t)m * npad;
    cf b = has_src ? src[r * col + (int64_t)j * npad + k] : mk(0.f, 0.f);
    if (has_term) {
        const cf *v = u + r * col + (int64_t)jv * npad;
        cf acc;
        if (!transposed) {
            acc = pc[k] * v[k];
            if (k >= 1) cfma(acc, pn[k], v[k - 1]);
            if (k + 1 < n) cfma(acc, ps[k], v[k + 1]);
        } else {
            acc = conjf_(pc[k]) * v[k];
            if (k + 1 < n) cfma(acc, conjf_(pn[k + 1]), v[k + 1]);
            if (k >= 1) cfma(acc, conjf_(ps[k - 1]), v[k - 1]);
        }
        b = b - acc;
    }
    tmp[(int64_t)r * npad + k] = b;
}

__global__ void ext_to_plain_kernel(const cd *__restrict__ ext, cf *__restrict__ out, int ny, int nx, int npad,
                                    int R) {
    int yi = blockIdx.x * blockDim.x + threadIdx.x;
    int xi = blockIdx.y, r = blockIdx.z;
    int n = ny - 2, m = nx - 2;
    if (yi >= npad) return;
    cf v = mk(0.f, 0.f);
    if (yi < n) v = tof(ext[(int64_t)r * nx * ny + (int64_t)(xi + 1) * ny + yi + 1]);
    out[((int64_t)r * m + xi) * npad + yi] = v;
}

__global__ void plain_to_ext_kernel(const cf *__restrict__ in, cd *__restrict__ ext, int ny, int nx, int npad,
                                    int R) {
    int y = blockIdx.x * blockDim.x + threadIdx.x;
    int x = blockIdx.y, r = blockIdx.z;
    int m = nx - 2;
    if (y >= ny) return;
    cd v = mkd(0, 0);
    if (y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2) v = tod(in[((int64_t)r * m + (x - 1)) * npad + y - 1]);
    ext[(int64_t)r * nx * ny + (int64_t)x * ny + y] = v;
}

}  // namespace

PlainLib::~PlainLib() {
    if (handle) rocblas_destroy_handle((rocblas_handle)handle);
    if (ipiv) (void)hipFree(ipiv);
    if (info) (void)hipFree(info);
}

void PlainLib::ensure(hipStream_t s, int n) {
    if (!handle) {
        rocblas_handle h;
        ROCBLAS_CHECK(rocblas_create_handle(&h));
        handle = h;
        WUST_HIP_CHECK(hipMalloc((void **)&ipiv, sizeof(int) * (size_t)n));
        WUST_HIP_CHECK(hipMalloc((void **)&info, sizeof(int) * 2 * 65536));
    }
    ROCBLAS_CHECK(rocblas_set_stream((rocblas_handle)handle, s));
}

void factor_plain(Operator &op, PlainLib &lib, hipStream_t s) {
    const int m = op.m, n = op.n, ld = op.npad;
    lib.ensure(s, n);
    rocblas_handle h = (rocblas_handle)lib.handle;
    cf *T = op.work;
    WUST_HIP_CHECK(hipMemsetAsync(lib.info, 0, sizeof(int) * 2 * (size_t)(m < 65536 ? m : 65536), s));
    for (int j = 0; j < m; ++j) {
        ProfScope ps(PC_PLAIN_FACTOR_STEP, s);
        launch_schur_left(op, j, T, s);
        ROCBLAS_CHECK(rocsolver_cgetrf(h, n, n, (rocblas_float_complex *)T, ld, lib.ipiv, lib.info + 2 * (j % 65536)));
        cf *X = op.page(j);
        hipLaunchKernelGGL(identity_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, X, n, ld);
        WUST_LAUNCH_CHECK();
        ROCBLAS_CHECK(rocsolver_cgetrs(h, rocblas_operation_none, n, n, (rocblas_float_complex *)T, ld, lib.ipiv,
                                       (rocblas_float_complex *)X, ld));
    }
    std::vector<int> info((size_t)2 * (m < 65536 ? m : 65536));
    WUST_HIP_CHECK(hipMemcpyAsync(info.data(), lib.info, sizeof(int) * info.size(), hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    for (size_t i = 0; i < info.size(); i += 2)
        if (info[i] > 0) throw Error(WUST_SINGULAR_SCHUR_BLOCK, "getrf reported a singular Schur block");
}

void solve_plain(const Operator &op, PlainLib &lib, const cf *src, cf *u, cf *tmp, int R, bool adjoint,
                 hipStream_t s) {
    const int m = op.m, n = op.n, ld = op.npad;
    lib.ensure(s, n);
    rocblas_handle h = (rocblas_handle)lib.handle;
    const int ldu = m * ld;
    const rocblas_float_complex one = {1.0f, 0.0f}, zero = {0.0f, 0.0f}, mone = {-1.0f, 0.0f};
    const rocblas_operation opX = adjoint ? rocblas_operation_conjugate_transpose : rocblas_operation_none;
    auto coupling = [&](int j, bool has_src, int dirs[3], int cblock, int jv, bool has_term, bool transposed) {
        const cf *pn = has_term ? op.coef(dirs[0], cblock) : nullptr;
        const cf *pc = has_term ? op.coef(dirs[1], cblock) : nullptr;
        const cf *ps = has_term ? op.coef(dirs[2], cblock) : nullptr;
        hipLaunchKernelGGL(coupling_kernel, dim3((n + 255) / 256, R), dim3(256), 0, s, src, u, tmp, pn, pc, ps,
                           transposed ? 1 : 0, has_src ? 1 : 0, has_term ? 1 : 0, j, jv, n, ld, m, R);
        WUST_LAUNCH_CHECK();
    };
    auto gemm = [&](int j, const rocblas_float_complex *alpha, const rocblas_float_complex *beta) {
        ROCBLAS_CHECK(rocblas_cgemm(h, opX, rocblas_operation_none, n, R, n, alpha,
                                    (const rocblas_float_complex *)op.page(j), ld,
                                    (const rocblas_float_complex *)tmp, ld, beta,
                                    (rocblas_float_complex *)(u + (int64_t)j * ld), ldu));
    };
    int Ld[3] = {DNW, DW, DSW};
    int Ud[3] = {DNE, DE, DSE};
    // forward pass: u_j = X_j (b_j - C_j u_{j-1})  (C = L_j, or U_{j-1}^H for the adjoint)
    for (int j = 0; j < m; ++j) {
        ProfScope ps(PC_PLAIN_SOLVE_STEP, s);
        if (!adjoint) coupling(j, true, Ld, j, j - 1, j >= 1, false);
        else coupling(j, true, Ud, j - 1, j - 1, j >= 1, true);
        gemm(j, &one, &zero);
    }
    // backward pass: u_j -= X_j (C'_j u_{j+1})  (C' = U_j, or L_{j+1}^H)
    for (int j = m - 2; j >= 0; --j) {
        ProfScope ps(PC_PLAIN_SOLVE_STEP, s);
        if (!adjoint) coupling(j, false, Ud, j, j + 1, true, false);
        else coupling(j, false, Ld, j + 1, j + 1, true, true);
        // tmp holds -(C' u): coupling() computes 0 - C'u
        gemm(j, &one, &one);
    }
    (void)mone;
}

void ext_to_plain(const cd *ext, cf *out, int ny, int nx, int npad, int R, hipStream_t s) {
    hipLaunchKernelGGL(ext_to_plain_kernel, dim3((npad + 255) / 256, nx - 2, R), dim3(256), 0, s, ext, out, ny, nx,
                       npad, R);
    WUST_LAUNCH_CHECK();
}
void plain_to_ext(const cf *in, cd *ext, int ny, int nx, int npad, int R, hipStream_t s) {
    hipLaunchKernelGGL(plain_to_ext_kernel, dim3((ny + 255) / 256, nx, R), dim3(256), 0, s, in, ext, ny, nx, npad,
                       R);
    WUST_LAUNCH_CHECK();
}

}  // namespace wust

namespace wust {
namespace {
__global__ void solve64_to_plain32_kernel(const cd *__restrict__ in, cf *__restrict__ out, int m, int npad, int R,
                                          int Rpad) {
    int yi = blockIdx.x * blockDim.x + threadIdx.x;
    int xi = blockIdx.y, r = blockIdx.z;
    if (yi >= npad) return;
    out[((int64_t)r * m + xi) * npad + yi] = tof(in[((int64_t)xi * npad + yi) * Rpad + r]);
}
__global__ void plain32_to_solve64_kernel(const cf *__restrict__ in, cd *__restrict__ out, int m, int npad, int R,
                                          int Rpad) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int yi = blockIdx.y, xi = blockIdx.z;
    if (r >= Rpad) return;
    out[((int64_t)xi * npad + yi) * Rpad + r] = (r < R) ? tod(in[((int64_t)r * m + xi) * npad + yi]) : mkd(0, 0);
}
}  // namespace

void solve64_to_plain32(const cd *in, cf *out, int m, int npad, int R, int Rpad, hipStream_t s) {
    hipLaunchKernelGGL(solve64_to_plain32_kernel, dim3((npad + 255) / 256, m, R), dim3(256), 0, s, in, out, m, npad, R,
                       Rpad);
    WUST_LAUNCH_CHECK();
}
void plain32_to_solve64(const cf *in, cd *out, int m, int npad, int R, int Rpad, hipStream_t s) {
    hipLaunchKernelGGL(plain32_to_solve64_kernel, dim3((Rpad + 63) / 64, npad, m), dim3(64), 0, s, in, out, m, npad,
                       R, Rpad);
    WUST_LAUNCH_CHECK();
}
}  // namespace wust
