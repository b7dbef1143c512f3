// MATLAB gateway of the HIP engine: one MEX function over the C ABI of include/wust_hip.h.
//
// The reference reaches its GPU code through two MEX gateways that take complex-single gpuArrays,
// decompBlockLU (Runtime/solver/decompBlockLU.cu:398) and applyBlockLU
// (Runtime/solver/applyBlockLU.cu:484), called from HelmholtzSolver.m:167-168 and :183-184.  MATLAB
// gpuArrays cannot hold HIP memory, so this gateway takes host arrays (MATLAB's interleaved complex
// storage, -R2018a), keeps the device operator behind a uint64 handle and copies fields in and out:
//
//   [h, info]        = wust_hip_mex('create', x, y, vel, atten, f, signConvention, a0, L_PML,
//                                   stencil_bounds, mode, dispersion)
//   [u, v, stats]    = wust_hip_mex('solve', h, src, adjoint, tol, max_iter)
//   out              = wust_hip_mex('mass', h, v, adjoint)
//   [loss, g, scale] = wust_hip_mex('oracle', h, tx_index, rx_index, Y, mask, batch, need_gradient,
//                                   retain, tol, max_iter)
//   kappa            = wust_hip_mex('curvature', h, direction, mask, tol, max_iter)
//   info             = wust_hip_mex('info', h)
//                      wust_hip_mex('release', h)     % drop the retained oracle state
//                      wust_hip_mex('destroy', h)
//
// vel, atten, direction: real double [Ny, Nx] (atten may be []); src, v: complex double [Ny, Nx] or
// [Ny, Nx, R]; Y: complex double [n_tx, n_rx]; mask: logical [n_tx, n_rx]; tx_index / rx_index: 1-based
// linear node indices (double); mode: 'refined' | 'c64' | 'plain' | 'refined-gauss-jordan' |
// 'c64-gauss-jordan'; dispersion: [] or [time_step_s, reference_speed_mps] (kwave-ldr9).
// info = [assembly_ms, factor_ms, b, d, e, npad, factor_fallback, inverse_residual, gj_residual];
// stats = [iterations, initial_residual, final_residual, solve_ms].
//
// Errors carry the reference's identifiers (wust_status_id): every device buffer of a call is
// released by its destructor before the error reaches MATLAB, as the reference's WustMexResources
// does (Runtime/solver/block_lu_mex_utils.h:57-110).  Handles are checked against the set this
// gateway created, so a destroyed or foreign handle is WUST:InvalidInput instead of a crash.
// Calls run on the null stream and return when their results are on the host.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "mex.h"
#include "wust_hip.h"

namespace {

struct MexError {
    std::string id, msg;
};

[[noreturn]] void fail(const char *id, const std::string &msg) { throw MexError{id, msg}; }
void invalid(const std::string &msg) { fail("WUST:InvalidInput", msg); }
void check(int st) {
    if (st != WUST_OK) throw MexError{wust_status_id(st), wust_last_error()};
}
void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail("WUST:HIPError", std::string(what) + ": " + hipGetErrorString(e));
}

// Device memory owned by one call.
struct DeviceBuffer {
    void *p = nullptr;
    size_t bytes = 0;
    explicit DeviceBuffer(size_t n) : bytes(n) {
        if (n) hip_check(hipMalloc(&p, n), "hipMalloc");
    }
    DeviceBuffer(const void *host, size_t n) : DeviceBuffer(n) {
        if (n) hip_check(hipMemcpy(p, host, n, hipMemcpyHostToDevice), "hipMemcpy host to device");
    }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void download(void *host) const {
        if (bytes) hip_check(hipMemcpy(host, p, bytes, hipMemcpyDeviceToHost), "hipMemcpy device to host");
    }
};

std::set<uint64_t> &live_handles() {
    static std::set<uint64_t> s;
    return s;
}

// ---- argument readers ------------------------------------------------------------------------
std::string string_arg(const mxArray *a, const char *name) {
    if (!a || !mxIsChar(a)) invalid(std::string(name) + " must be a character vector");
    char *s = mxArrayToString(a);
    if (!s) invalid(std::string(name) + " must be a character vector");
    std::string out(s);
    mxFree(s);
    return out;
}

double scalar_arg(const mxArray *a, const char *name) {
    if (!a || mxGetNumberOfElements(a) != 1 || mxIsComplex(a) || !(mxIsDouble(a) || mxIsLogical(a)))
        invalid(std::string(name) + " must be a real scalar");
    return mxGetScalar(a);
}

wust_helmholtz *handle_arg(const mxArray *a) {
    if (!a || !mxIsUint64(a) || mxGetNumberOfElements(a) != 1 || mxIsComplex(a))
        invalid("solver handle must be a uint64 scalar");
    const uint64_t v = mxGetUint64s(a)[0];
    if (!live_handles().count(v)) invalid("invalid or destroyed solver handle");
    return reinterpret_cast<wust_helmholtz *>(static_cast<uintptr_t>(v));
}

// Dimensions of a at most `max_rank`-dimensional array (trailing singleton dimensions are MATLAB's).
std::vector<size_t> dims_of(const mxArray *a, size_t max_rank, const char *name) {
    const size_t nd = mxGetNumberOfDimensions(a);
    const mwSize *d = mxGetDimensions(a);
    std::vector<size_t> out(d, d + nd);
    while (out.size() > 2 && out.back() == 1) out.pop_back();
    if (out.size() > max_rank)
        invalid(std::string(name) + " must have at most " + std::to_string(max_rank) + " dimensions");
    while (out.size() < max_rank) out.push_back(1);
    return out;
}

const double *real_grid(const mxArray *a, int ny, int nx, const char *name) {
    if (!a || !mxIsDouble(a) || mxIsComplex(a)) invalid(std::string(name) + " must be a real double array");
    const std::vector<size_t> d = dims_of(a, 2, name);
    if (d[0] != (size_t)ny || d[1] != (size_t)nx) invalid(std::string(name) + " must be [ny, nx]");
    return mxGetDoubles(a);
}

// src / v: complex double [Ny, Nx, R] (the reference's MEX boundary also rejects real storage and
// rank-4 fields: Runtime/tests/test_runtime.m:220-231)
const mxComplexDouble *complex_field(const mxArray *a, int ny, int nx, int64_t &R, const char *name) {
    if (!a || !mxIsDouble(a) || !mxIsComplex(a)) invalid(std::string(name) + " must be complex double storage");
    const std::vector<size_t> d = dims_of(a, 3, name);
    if (d[0] != (size_t)ny || d[1] != (size_t)nx) invalid(std::string(name) + " must be [ny, nx, n_source]");
    R = (int64_t)d[2];
    if (R < 1) invalid(std::string(name) + " has no right-hand sides");
    return mxGetComplexDoubles(a);
}

std::vector<int64_t> index_arg(const mxArray *a, int64_t nodes, const char *name) {
    if (!a || !mxIsDouble(a) || mxIsComplex(a) || mxGetNumberOfElements(a) < 1)
        invalid(std::string(name) + " must be a nonempty real index vector");
    const double *p = mxGetDoubles(a);
    std::vector<int64_t> out(mxGetNumberOfElements(a));
    for (size_t i = 0; i < out.size(); ++i) {
        if (!(p[i] >= 1 && p[i] <= (double)nodes && p[i] == std::floor(p[i])))
            invalid(std::string(name) + " must hold 1-based node indices inside the grid");
        out[i] = (int64_t)p[i];
    }
    return out;
}

std::vector<uint8_t> mask_arg(const mxArray *a, int64_t rows, int64_t cols, const char *name) {
    if (!a || !(mxIsLogical(a) || mxIsDouble(a)) || mxIsComplex(a))
        invalid(std::string(name) + " must be a logical [n_tx, n_rx] mask");
    const std::vector<size_t> d = dims_of(a, 2, name);
    if ((int64_t)d[0] != rows || (int64_t)d[1] != cols) invalid(std::string(name) + " must be [n_tx, n_rx]");
    std::vector<uint8_t> out((size_t)(rows * cols));
    if (mxIsLogical(a)) {
        const mxLogical *p = mxGetLogicals(a);
        for (size_t i = 0; i < out.size(); ++i) out[i] = p[i] ? 1 : 0;
    } else {
        const double *p = mxGetDoubles(a);
        for (size_t i = 0; i < out.size(); ++i) out[i] = p[i] != 0 ? 1 : 0;
    }
    return out;
}

// x0 and spacing of a uniform increasing coordinate vector (HelmholtzSolver.m:47-51)
void axis_arg(const mxArray *a, const char *name, int &n, double &x0, double &h) {
    if (!a || !mxIsDouble(a) || mxIsComplex(a) || mxGetNumberOfElements(a) < 3)
        invalid(std::string(name) + " must be a real vector of at least 3 coordinates");
    const double *p = mxGetDoubles(a);
    n = (int)mxGetNumberOfElements(a);
    double sum = 0;
    for (int i = 1; i < n; ++i) sum += p[i] - p[i - 1];
    h = sum / (n - 1);
    for (int i = 1; i < n; ++i)
        if (!(p[i] - p[i - 1] > 0) || !(std::fabs(p[i] - p[i - 1] - h) < h * 1e-7))
            invalid("Uniform increasing grids required");
    x0 = p[0];
}

mxArray *row(const std::vector<double> &v) {
    mxArray *a = mxCreateDoubleMatrix(1, v.size(), mxREAL);
    std::memcpy(mxGetDoubles(a), v.data(), sizeof(double) * v.size());
    return a;
}

std::vector<double> info_row(const wust_helmholtz *h) {
    wust_info in{};
    check(wust_helmholtz_info(h, &in));
    return {in.assembly_ms, in.factor_ms, in.b, in.d, in.e, (double)in.npad, (double)in.factor_fallback,
            in.inverse_residual, in.gj_residual};
}

void need(int nrhs, int n, const char *cmd) {
    if (nrhs != n) invalid(std::string(cmd) + " takes " + std::to_string(n - 1) + " arguments");
}

// ---- commands --------------------------------------------------------------------------------
void cmd_create(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    need(nrhs, 12, "create");
    wust_grid_desc g{};
    axis_arg(prhs[1], "x", g.nx, g.x0, g.dx);
    axis_arg(prhs[2], "y", g.ny, g.y0, g.dy);
    const double *vel = real_grid(prhs[3], g.ny, g.nx, "vel");
    const double *atten = nullptr;
    if (mxGetNumberOfElements(prhs[4]) > 0) {
        atten = real_grid(prhs[4], g.ny, g.nx, "atten");
        bool any = false;
        for (int64_t i = 0; i < (int64_t)g.ny * g.nx; ++i) any |= atten[i] != 0;
        if (!any) atten = nullptr;
    }
    g.frequency_hz = scalar_arg(prhs[5], "f");
    g.sign_convention = scalar_arg(prhs[6], "signConvention");
    g.pml_strength = scalar_arg(prhs[7], "a0");
    g.pml_m = scalar_arg(prhs[8], "L_PML");
    if (!mxIsDouble(prhs[9]) || mxIsComplex(prhs[9]) || mxGetNumberOfElements(prhs[9]) != 2)
        invalid("stencil_bounds must be a positive ordered [min,max] pair");
    g.stencil_vmin = mxGetDoubles(prhs[9])[0];
    g.stencil_vmax = mxGetDoubles(prhs[9])[1];
    const std::string mode = string_arg(prhs[10], "mode");
    int m = 0;
    if (mode == "refined") m = WUST_MODE_REFINED;
    else if (mode == "c64") m = WUST_MODE_C64;
    else if (mode == "plain") m = WUST_MODE_PLAIN;
    else if (mode == "refined-gauss-jordan") m = WUST_MODE_REFINED | WUST_FACTOR_GAUSS_JORDAN;
    else if (mode == "c64-gauss-jordan") m = WUST_MODE_C64 | WUST_FACTOR_GAUSS_JORDAN;
    else invalid("unknown mode " + mode);
    if (mxGetNumberOfElements(prhs[11]) > 0) {
        if (!mxIsDouble(prhs[11]) || mxIsComplex(prhs[11]) || mxGetNumberOfElements(prhs[11]) != 2)
            invalid("dispersion must be [] or [time_step_s, reference_speed_mps]");
        g.dispersion = 1;
        g.ldr9_time_step_s = mxGetDoubles(prhs[11])[0];
        g.ldr9_reference_speed_mps = mxGetDoubles(prhs[11])[1];
    }
    const size_t grid_bytes = sizeof(double) * (size_t)g.ny * g.nx;
    DeviceBuffer dvel(vel, grid_bytes);
    DeviceBuffer datt(atten, atten ? grid_bytes : 0);
    wust_helmholtz *h = nullptr;
    check(wust_helmholtz_create(&g, (const double *)dvel.p, (const double *)datt.p, m, nullptr, &h));
    live_handles().insert((uint64_t)(uintptr_t)h);
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    mxGetUint64s(plhs[0])[0] = (uint64_t)(uintptr_t)h;
    if (nlhs > 1) plhs[1] = row(info_row(h));
}

void grid_of(const wust_helmholtz *h, int &ny, int &nx) {
    wust_info in{};
    check(wust_helmholtz_info(h, &in));
    ny = in.ny;
    nx = in.nx;
}

mxArray *complex_like(const mxArray *a) {
    return mxCreateNumericArray(mxGetNumberOfDimensions(a), mxGetDimensions(a), mxDOUBLE_CLASS, mxCOMPLEX);
}

void cmd_solve(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    need(nrhs, 6, "solve");
    wust_helmholtz *h = handle_arg(prhs[1]);
    int ny, nx;
    grid_of(h, ny, nx);
    int64_t R = 0;
    const mxComplexDouble *src = complex_field(prhs[2], ny, nx, R, "src");
    const int adjoint = scalar_arg(prhs[3], "adjoint") != 0;
    const double tol = scalar_arg(prhs[4], "tol");
    const double iters = scalar_arg(prhs[5], "max_iter");
    if (!(iters >= 0 && iters == std::floor(iters))) invalid("max_iter must be a nonnegative integer");
    const size_t bytes = sizeof(mxComplexDouble) * (size_t)ny * nx * (size_t)R;
    DeviceBuffer dsrc(src, bytes), dout(bytes);
    wust_solve_stats st{};
    check(wust_helmholtz_solve(h, dsrc.p, dout.p, R, adjoint, tol, (int)iters, nullptr, &st));
    plhs[0] = complex_like(prhs[2]);
    dout.download(mxGetComplexDoubles(plhs[0]));
    if (nlhs > 1) {   // v = (C .* q') .* u, HelmholtzSolver.m:197-199, formed on the device
        check(wust_helmholtz_virtual_sources(h, dout.p, dsrc.p, R, nullptr));
        plhs[1] = complex_like(prhs[2]);
        dsrc.download(mxGetComplexDoubles(plhs[1]));
    }
    if (nlhs > 2) plhs[2] = row({(double)st.iterations, st.initial_residual, st.final_residual, st.solve_ms});
}

void cmd_mass(int, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    need(nrhs, 4, "mass");
    wust_helmholtz *h = handle_arg(prhs[1]);
    int ny, nx;
    grid_of(h, ny, nx);
    int64_t R = 0;
    const mxComplexDouble *v = complex_field(prhs[2], ny, nx, R, "v");
    const int adjoint = scalar_arg(prhs[3], "adjoint") != 0;
    const size_t bytes = sizeof(mxComplexDouble) * (size_t)ny * nx * (size_t)R;
    DeviceBuffer din(v, bytes), dout(bytes);
    check(wust_helmholtz_mass_apply(h, din.p, dout.p, R, adjoint, nullptr));
    plhs[0] = complex_like(prhs[2]);
    dout.download(mxGetComplexDoubles(plhs[0]));
}

void cmd_oracle(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    need(nrhs, 11, "oracle");
    wust_helmholtz *h = handle_arg(prhs[1]);
    int ny, nx;
    grid_of(h, ny, nx);
    const int64_t nodes = (int64_t)ny * nx;
    const std::vector<int64_t> tx = index_arg(prhs[2], nodes, "tx_index");
    const std::vector<int64_t> rx = index_arg(prhs[3], nodes, "rx_index");
    const int64_t nt = (int64_t)tx.size(), nr = (int64_t)rx.size();
    if (!mxIsDouble(prhs[4]) || !mxIsComplex(prhs[4])) invalid("Y must be complex double [n_tx, n_rx]");
    const std::vector<size_t> yd = dims_of(prhs[4], 2, "Y");
    if ((int64_t)yd[0] != nt || (int64_t)yd[1] != nr) invalid("Y must be [n_tx, n_rx]");
    const std::vector<uint8_t> mask = mask_arg(prhs[5], nt, nr, "mask");
    const double batch = scalar_arg(prhs[6], "source_batch_size");
    if (!(batch >= 1 && batch == std::floor(batch))) invalid("source_batch_size must be a positive integer");
    const int need_gradient = scalar_arg(prhs[7], "need_gradient") != 0;
    const int retain = scalar_arg(prhs[8], "retain") != 0;
    const double tol = scalar_arg(prhs[9], "tol");
    const double iters = scalar_arg(prhs[10], "max_iter");
    if (!(iters >= 0 && iters == std::floor(iters))) invalid("max_iter must be a nonnegative integer");
    DeviceBuffer dY(mxGetComplexDoubles(prhs[4]), sizeof(mxComplexDouble) * (size_t)(nt * nr));
    DeviceBuffer dmask(mask.data(), mask.size());
    DeviceBuffer dgrad(need_gradient ? sizeof(double) * (size_t)nodes : 0);
    double loss = 0;
    check(wust_oracle_eval(h, tx.data(), nt, rx.data(), nr, dY.p, (const uint8_t *)dmask.p, (int64_t)batch,
                           need_gradient, retain, tol, (int)iters, &loss, dgrad.p, nullptr));
    plhs[0] = mxCreateDoubleScalar(loss);
    if (nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix(ny, nx, mxREAL);
        if (need_gradient) dgrad.download(mxGetDoubles(plhs[1]));
    }
    if (nlhs > 2) {
        plhs[2] = mxCreateNumericMatrix(1, nt, mxDOUBLE_CLASS, mxCOMPLEX);
        if (retain) check(wust_oracle_scales(h, mxGetComplexDoubles(plhs[2])));
    }
}

void cmd_curvature(int, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    need(nrhs, 6, "curvature");
    wust_helmholtz *h = handle_arg(prhs[1]);
    int ny, nx;
    grid_of(h, ny, nx);
    const double *dir = real_grid(prhs[2], ny, nx, "direction");
    if (!(mxIsLogical(prhs[3]) || mxIsDouble(prhs[3])) || mxIsComplex(prhs[3]))
        invalid("receiver_mask must be a logical [n_tx, n_rx] mask");
    const std::vector<size_t> md = dims_of(prhs[3], 2, "receiver_mask");
    const std::vector<uint8_t> mask = mask_arg(prhs[3], (int64_t)md[0], (int64_t)md[1], "receiver_mask");
    const double tol = scalar_arg(prhs[4], "tol");
    const double iters = scalar_arg(prhs[5], "max_iter");
    if (!(iters >= 0 && iters == std::floor(iters))) invalid("max_iter must be a nonnegative integer");
    DeviceBuffer ddir(dir, sizeof(double) * (size_t)ny * nx);
    DeviceBuffer dmask(mask.data(), mask.size());
    double kappa = 0;
    check(wust_directional_curvature(h, (const double *)ddir.p, (const uint8_t *)dmask.p, tol, (int)iters, &kappa,
                                     nullptr));
    plhs[0] = mxCreateDoubleScalar(kappa);
}

void run(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 1) invalid("usage: wust_hip_mex(command, ...)");
    const std::string cmd = string_arg(prhs[0], "command");
    if (cmd == "create") return cmd_create(nlhs, plhs, nrhs, prhs);
    if (cmd == "solve") return cmd_solve(nlhs, plhs, nrhs, prhs);
    if (cmd == "mass") return cmd_mass(nlhs, plhs, nrhs, prhs);
    if (cmd == "oracle") return cmd_oracle(nlhs, plhs, nrhs, prhs);
    if (cmd == "curvature") return cmd_curvature(nlhs, plhs, nrhs, prhs);
    if (cmd == "info") {
        need(nrhs, 2, "info");
        plhs[0] = row(info_row(handle_arg(prhs[1])));
        return;
    }
    if (cmd == "release") {
        need(nrhs, 2, "release");
        check(wust_oracle_release(handle_arg(prhs[1])));
        return;
    }
    if (cmd == "destroy") {
        need(nrhs, 2, "destroy");
        wust_helmholtz *h = handle_arg(prhs[1]);
        live_handles().erase((uint64_t)(uintptr_t)h);
        check(wust_helmholtz_destroy(h));
        return;
    }
    invalid("unknown command " + cmd);
}

}  // namespace

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    std::string id, msg;
    try {
        run(nlhs, plhs, nrhs, prhs);
        return;
    } catch (const MexError &e) {
        id = e.id;
        msg = e.msg;
    } catch (const std::exception &e) {
        id = "WUST:RuntimeError";
        msg = e.what();
    }
    // every DeviceBuffer of the call has been released by now
    mexErrMsgIdAndTxt(id.c_str(), "%s", msg.c_str());
}
