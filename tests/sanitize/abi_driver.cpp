// Driver of the C ABI (include/wust_hip.h) for the sanitized debug build of the library
// (libwust_hip_debug.so: host code under AddressSanitizer and UndefinedBehaviorSanitizer, a device
// synchronization and error check after every launch).  Linked with the sanitizer runtime, so that
// the library's host code runs instrumented.
//
//   abi_driver validate   argument checks that fail before any device work (no GPU needed)
//   abi_driver full       every entry point on a small operator, on the GPU
//
// Prints "abi_driver <mode>: ok" and exits 0 when every expectation holds; a sanitizer report
// aborts the process (halt_on_error).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "wust_hip.h"

static int failures = 0;
#define EXPECT(cond, what)                                                                                   \
    do {                                                                                                     \
        if (!(cond)) {                                                                                       \
            std::fprintf(stderr, "FAIL %s:%d: %s (last error: %s)\n", __FILE__, __LINE__, what, wust_last_error()); \
            ++failures;                                                                                      \
        }                                                                                                    \
    } while (0)
#define HIP_OK(expr)                                                                        \
    do {                                                                                    \
        if ((expr) != hipSuccess) {                                                         \
            std::fprintf(stderr, "HIP call failed at %s:%d: %s\n", __FILE__, __LINE__, #expr); \
            std::exit(2);                                                                   \
        }                                                                                   \
    } while (0)

static wust_grid_desc grid(int ny, int nx) {
    wust_grid_desc g{};
    g.ny = ny;
    g.nx = nx;
    g.dx = g.dy = 2.5e-4;
    g.x0 = -0.5 * (nx - 1) * g.dx;
    g.y0 = -0.5 * (ny - 1) * g.dy;
    g.frequency_hz = 3e5;
    g.sign_convention = -1;
    g.pml_strength = 10.0;
    g.pml_m = 2e-3;
    g.stencil_vmin = 1400;
    g.stencil_vmax = 1700;
    return g;
}

template <typename T>
static T *dev(size_t n) {
    void *p = nullptr;
    HIP_OK(hipMalloc(&p, n * sizeof(T) + 16));
    HIP_OK(hipMemset(p, 0, n * sizeof(T) + 16));
    return static_cast<T *>(p);
}
template <typename T>
static std::vector<T> host(const T *d, size_t n) {
    std::vector<T> h(n);
    HIP_OK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}
template <typename T>
static void upload(T *d, const std::vector<T> &h) {
    HIP_OK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}
static bool finite(const std::vector<std::complex<double>> &v) {
    for (auto &z : v)
        if (!std::isfinite(z.real()) || !std::isfinite(z.imag())) return false;
    return true;
}

static void validate() {
    EXPECT(wust_abi_version() == WUST_ABI_VERSION, "ABI version");
    for (int s = 0; s <= WUST_NOT_CONVERGED; ++s) EXPECT(wust_status_id(s) && std::strlen(wust_status_id(s)) > 5, "status id");
    double b = 0, d = 0, e = 0, alpha = -1;
    EXPECT(wust_stencil_opt_params(1400, 1700, 5e5, 2e-4, 1.0, &b, &d, &e) == WUST_OK, "stencil parameters");
    EXPECT(std::fabs(b - 5.0 / 6.0) < 1e-15 && std::isfinite(d) && std::isfinite(e), "b = 5/6");
    EXPECT(wust_linearized_step_length(-2.0, 4.0, &alpha) == WUST_OK && alpha == 0.5, "step length");
    EXPECT(wust_linearized_step_length(1.0, -1.0, &alpha) == WUST_NUMERICAL_FAILURE, "negative curvature");
    EXPECT(wust_linearized_step_length(NAN, 1.0, &alpha) == WUST_NUMERICAL_FAILURE, "non-finite slope");

    // create: every rejected description leaves *out null and touches no device memory
    double fake_vel = 1500;
    auto reject = [&](wust_grid_desc g, int mode, const char *what) {
        wust_helmholtz *h = reinterpret_cast<wust_helmholtz *>(&fake_vel);
        EXPECT(wust_helmholtz_create(&g, &fake_vel, nullptr, mode, nullptr, &h) == WUST_INVALID_INPUT, what);
        EXPECT(h == nullptr, "out cleared on failure");
        EXPECT(std::strlen(wust_last_error()) > 0, "message recorded");
    };
    wust_grid_desc g = grid(20, 18);
    wust_grid_desc bad = g;
    bad.ny = 2;
    reject(bad, WUST_MODE_REFINED, "ny < 3");
    bad = g;
    bad.dx = -1;
    reject(bad, WUST_MODE_REFINED, "negative spacing");
    bad = g;
    bad.frequency_hz = 0;
    reject(bad, WUST_MODE_REFINED, "zero frequency");
    bad = g;
    bad.stencil_vmax = bad.stencil_vmin;
    reject(bad, WUST_MODE_REFINED, "unordered stencil bounds");
    bad = g;
    bad.dispersion = 1;
    bad.dy = 2 * bad.dx;
    reject(bad, WUST_MODE_REFINED, "LDR9 on rectangular cells");
    reject(g, 7, "unknown mode");
    reject(g, WUST_MODE_REFINED | 0x400, "unknown mode flag");
    reject(g, WUST_MODE_REFINED | WUST_FACTOR_GAUSS_JORDAN | WUST_FACTOR_FIXED_BUDGET, "Gauss-Jordan with a budget");
    EXPECT(wust_helmholtz_create(&g, nullptr, nullptr, 0, nullptr, nullptr) == WUST_INVALID_INPUT, "null arguments");

    // entry points on a null handle
    int dummy = 0;
    EXPECT(wust_helmholtz_destroy(nullptr) == WUST_OK, "destroy(NULL) is a no-op");
    EXPECT(wust_helmholtz_solve(nullptr, &dummy, &dummy, 1, 0, 1e-10, 5, nullptr, nullptr) == WUST_INVALID_INPUT, "solve(NULL)");
    wust_info info;
    EXPECT(wust_helmholtz_info(nullptr, &info) == WUST_INVALID_INPUT, "info(NULL)");
    double kappa = 0;
    EXPECT(wust_directional_curvature(nullptr, nullptr, nullptr, 1e-10, 5, &kappa, nullptr) == WUST_INVALID_INPUT, "curvature(NULL)");
    double f = 1e6;
    EXPECT(wust_rf_dtft_noalloc(&dummy, 2, 100, 1, 1, &fake_vel, 1e-8, &f, 1, nullptr, &dummy, nullptr) == WUST_INVALID_INPUT, "dtft null active");
}

static void full() {
    const int ny = 40, nx = 36, R = 3;
    const size_t nodes = (size_t)ny * nx;
    wust_grid_desc g = grid(ny, nx);
    std::vector<double> vel(nodes);
    for (int j = 0; j < nx; ++j)
        for (int i = 0; i < ny; ++i) {
            const double x = g.x0 + j * g.dx, y = g.y0 + i * g.dy;
            vel[i + (size_t)ny * j] = 1500 + 40 * std::exp(-(x * x + y * y) / 4e-6);
        }
    double *d_vel = dev<double>(nodes);
    upload(d_vel, vel);
    wust_helmholtz *h = nullptr;
    EXPECT(wust_helmholtz_create(&g, d_vel, nullptr, WUST_MODE_REFINED, nullptr, &h) == WUST_OK && h, "create");
    if (!h) return;
    wust_info info{};
    EXPECT(wust_helmholtz_info(h, &info) == WUST_OK && info.n == ny - 2 && info.m == nx - 2, "info");

    // solve / adjoint solve / virtual sources / mass apply
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd;
    std::vector<std::complex<double>> src(nodes * R);
    for (int r = 0; r < R; ++r)
        for (int j = 1; j < nx - 1; ++j)
            for (int i = 1; i < ny - 1; ++i) src[i + (size_t)ny * j + nodes * r] = {nd(rng), nd(rng)};
    auto *d_src = dev<std::complex<double>>(nodes * R);
    auto *d_u = dev<std::complex<double>>(nodes * R);
    auto *d_v = dev<std::complex<double>>(nodes * R);
    upload(d_src, src);
    for (int adjoint = 0; adjoint < 2; ++adjoint) {
        wust_solve_stats st{};
        EXPECT(wust_helmholtz_solve(h, d_src, d_u, R, adjoint, 1e-10, 20, nullptr, &st) == WUST_OK, "solve");
        EXPECT(st.final_residual <= 1e-10 && st.iterations >= 1, "refined to tolerance");
        EXPECT(finite(host(d_u, nodes * R)), "finite solution");
        EXPECT(wust_helmholtz_mass_apply(h, d_u, d_v, R, adjoint, nullptr) == WUST_OK, "mass apply");
    }
    EXPECT(wust_helmholtz_virtual_sources(h, d_u, d_v, R, nullptr) == WUST_OK, "virtual sources");
    EXPECT(wust_helmholtz_solve(h, d_src, d_u, R, 0, 1e-20, 1, nullptr, nullptr) == WUST_NOT_CONVERGED, "not converged");
    EXPECT(wust_helmholtz_solve(h, d_src, d_src, R, 0, 1e-10, 20, nullptr, nullptr) == WUST_INVALID_INPUT, "aliased output");
    auto *d_coef = dev<std::complex<double>>(nodes);
    EXPECT(wust_helmholtz_coefficients(h, d_coef, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                       nullptr, nullptr) == WUST_OK, "coefficients");

    // oracle: loss, gradient, retained state, curvature, release
    const int n_tx = 4, n_rx = 12;
    std::vector<int64_t> tx(n_tx), rx(n_rx);
    for (int t = 0; t < n_tx; ++t) tx[t] = 1 + (3 + 9 * t) + (int64_t)ny * 4;            // column 5, 1-based
    for (int r = 0; r < n_rx; ++r) rx[r] = 1 + (2 + 3 * r) + (int64_t)ny * (nx - 5);       // column nx - 4
    std::vector<std::complex<double>> Y((size_t)n_tx * n_rx);
    for (auto &y : Y) y = {nd(rng) * 1e-3, nd(rng) * 1e-3};
    auto *d_Y = dev<std::complex<double>>(Y.size());
    upload(d_Y, Y);
    auto *d_mask = dev<uint8_t>(Y.size());
    upload(d_mask, std::vector<uint8_t>(Y.size(), 1));
    double *d_grad = dev<double>(nodes);
    double loss = -1;
    for (int64_t batch : {int64_t(0), int64_t(3)})
        EXPECT(wust_oracle_eval(h, tx.data(), n_tx, rx.data(), n_rx, d_Y, d_mask, batch, 1, 1, 1e-10, 20, &loss,
                                d_grad, nullptr) == WUST_OK && std::isfinite(loss) && loss > 0, "oracle eval");
    std::vector<std::complex<double>> scales(n_tx);
    EXPECT(wust_oracle_scales(h, scales.data()) == WUST_OK && finite(scales), "scales");
    std::vector<double> dir(nodes, 1e-6);
    double *d_dir = dev<double>(nodes);
    upload(d_dir, dir);
    double kappa = -1;
    EXPECT(wust_directional_curvature(h, d_dir, d_mask, 1e-10, 20, &kappa, nullptr) == WUST_OK && kappa >= 0, "curvature");
    EXPECT(wust_oracle_release(h) == WUST_OK, "release");
    EXPECT(wust_directional_curvature(h, d_dir, d_mask, 1e-10, 20, &kappa, nullptr) == WUST_INVALID_INPUT, "curvature after release");
    std::vector<int64_t> bad_tx = tx;
    bad_tx[0] = 0;   // 1-based: 0 is out of range
    EXPECT(wust_oracle_eval(h, bad_tx.data(), n_tx, rx.data(), n_rx, d_Y, d_mask, 0, 1, 0, 1e-10, 20, &loss, d_grad,
                            nullptr) == WUST_INVALID_INPUT, "transmitter index out of range");

    // NCG helpers on the gradient
    auto *d_upd = dev<uint8_t>(nodes);
    upload(d_upd, std::vector<uint8_t>(nodes, 1));
    double *d_roi = dev<double>(nodes), *d_z = dev<double>(nodes), *d_p = dev<double>(nodes), *d_c = dev<double>(nodes);
    EXPECT(wust_roi_gradient(d_grad, d_upd, ny, nx, g.dx, g.dy, 1500, g.frequency_hz, 0.5, 8, d_roi, d_z, nullptr) == WUST_OK, "ROI gradient");
    double beta = -1, gz = -1, slope = 0, rel = -1;
    EXPECT(wust_ncg_direction(d_roi, d_z, nullptr, 0, nullptr, 0, (int64_t)nodes, d_p, &beta, &gz, &slope, nullptr) == WUST_OK, "NCG direction");
    EXPECT(wust_project_slowness_update(d_vel, d_p, 1e-3, 1.0, 50, 1300, 1800, d_upd, (int64_t)nodes, d_c, &rel, nullptr) == WUST_OK && std::isfinite(rel), "projection");

    EXPECT(wust_helmholtz_destroy(h) == WUST_OK, "destroy");
    EXPECT(wust_pool_trim() == WUST_OK, "pool trim");

    // DTFT: the allocating entry and the allocation-free one agree
    const int nt = 300, nrx = 5, ntx = 3, nf = 2;
    std::vector<float> rf((size_t)nt * nrx * ntx);
    for (auto &x : rf) x = (float)nd(rng);
    std::vector<double> t(nt);
    const double dt = 2.5e-8;
    for (int i = 0; i < nt; ++i) t[i] = 1e-6 + i * dt;
    const double freqs[nf] = {1.0e6, 2.5e6};
    float *d_rf = dev<float>(rf.size());
    upload(d_rf, rf);
    double *d_t = dev<double>(nt);
    upload(d_t, t);
    auto *d_act = dev<uint8_t>((size_t)ntx * nrx);
    upload(d_act, std::vector<uint8_t>((size_t)ntx * nrx, 1));
    const size_t np = (size_t)ntx * nrx * nf;
    auto *d_p1 = dev<std::complex<float>>(np), *d_p2 = dev<std::complex<float>>(np);
    EXPECT(wust_rf_dtft(d_rf, 0, nt, nrx, ntx, d_t, dt, freqs, nf, d_act, d_p1, nullptr, nullptr) == WUST_OK, "DTFT");
    EXPECT(wust_rf_dtft_noalloc(d_rf, 0, nt, nrx, ntx, d_t, dt, freqs, nf, d_act, d_p2, nullptr) == WUST_OK, "DTFT without allocation");
    const auto p1 = host(d_p1, np), p2 = host(d_p2, np);
    EXPECT(std::memcmp(p1.data(), p2.data(), np * sizeof(p1[0])) == 0, "DTFT entries agree");
    EXPECT(wust_rf_dtft_noalloc(d_rf, 0, WUST_DTFT_NOALLOC_MAX_TIME + 1, nrx, ntx, d_t, dt, freqs, nf, d_act, d_p2,
                                nullptr) == WUST_INVALID_INPUT, "DTFT without allocation beyond its limit");

    for (void *p : {(void *)d_vel, (void *)d_src, (void *)d_u, (void *)d_v, (void *)d_coef, (void *)d_Y, (void *)d_mask,
                    (void *)d_grad, (void *)d_dir, (void *)d_upd, (void *)d_roi, (void *)d_z, (void *)d_p, (void *)d_c,
                    (void *)d_rf, (void *)d_t, (void *)d_act, (void *)d_p1, (void *)d_p2})
        HIP_OK(hipFree(p));
}

int main(int argc, char **argv) {
    const char *mode = argc > 1 ? argv[1] : "validate";
    if (std::strcmp(mode, "validate") == 0) validate();
    else if (std::strcmp(mode, "full") == 0) {
        validate();
        full();
    } else {
        std::fprintf(stderr, "usage: abi_driver validate|full\n");
        return 2;
    }
    if (failures) {
        std::fprintf(stderr, "abi_driver %s: %d expectation(s) failed\n", mode, failures);
        return 1;
    }
    std::printf("abi_driver %s: ok\n", mode);
    return 0;
}
