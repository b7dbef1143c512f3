// C ABI of the engine (declared in include/wust_hip.h): handle lifetime,
// validation, refinement driver, oracle (loss/gradient) and curvature drivers.
#include <cstdlib>
#include <mutex>
#include <map>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "engine.h"
#include "fields.h"
#include "dtft.h"
#include "ncg.h"
#include "plain.h"
#include "profile.h"

using namespace wust;

namespace {

thread_local std::string g_last_error;

int fail(int status, const std::string &msg) {
    g_last_error = msg;
    return status;
}

template <typename F>
int guarded(F &&fn) {
    try {
        fn();
        g_last_error.clear();
        return WUST_OK;
    } catch (const Error &e) {
        return fail(e.status, e.what());
    } catch (const std::bad_alloc &) {
        return fail(WUST_RUNTIME_ERROR, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(WUST_RUNTIME_ERROR, e.what());
    }
}

// Device buffers of destroyed operators are kept for the next operator with the same sizes
// (one operator is typically created per model update, with identical shapes): the
// destructor synchronizes the device once and hands its buffers to a free list keyed by
// (device, bytes); an allocation of that size takes one back instead of calling hipMalloc.
// Only the destructor pools buffers, so none can still be in use by queued kernels.  The
// pool holds at most WUST_POOL_MB megabytes (default 98304); wust_pool_trim() frees it.
struct DevicePool {
    std::mutex mu;
    std::map<std::pair<int, size_t>, std::vector<void *>> free_list;
    size_t held = 0;
    size_t cap = [] {
        const char *e = std::getenv("WUST_POOL_MB");
        const long long mb = e ? std::atoll(e) : 98304;
        return (size_t)(mb > 0 ? mb : 0) << 20;
    }();
};
DevicePool &pool() {
    static DevicePool *p = new DevicePool();   // never destroyed: buffers outlive static teardown
    return *p;
}
int current_device() {
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

template <typename T>
T *dalloc(int64_t count) {
    void *p = nullptr;
    if (count <= 0) return nullptr;
    const size_t bytes = sizeof(T) * (size_t)count;
    {
        DevicePool &pl = pool();
        std::lock_guard<std::mutex> lk(pl.mu);
        auto it = pl.free_list.find({current_device(), bytes});
        if (it != pl.free_list.end() && !it->second.empty()) {
            p = it->second.back();
            it->second.pop_back();
            pl.held -= bytes;
            return static_cast<T *>(p);
        }
    }
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
        throw Error(WUST_HIP_ERROR, std::string("hipMalloc(") + std::to_string(bytes) + " bytes): " +
                                         hipGetErrorString(e));
    return static_cast<T *>(p);
}
// immediate release (hipFree synchronizes with all queued work)
template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}
// release into the pool; the caller has synchronized the device
template <typename T>
void dpool(T *&p, int64_t count) {
    if (!p) return;
    const size_t bytes = sizeof(T) * (size_t)count;
    DevicePool &pl = pool();
    std::lock_guard<std::mutex> lk(pl.mu);
    if (count > 0 && pl.held + bytes <= pl.cap) {
        pl.free_list[{current_device(), bytes}].push_back((void *)p);
        pl.held += bytes;
    } else {
        (void)hipFree((void *)p);
    }
    p = nullptr;
}

// Householder least squares for the 2-column stencil fit
// (Runtime/solver/stencilOptParams.m:33-37 solves [S2 S3] [d; e] = S4 - b S1 with MATLAB's QR "\").
void stencil_params(double vmin, double vmax, double f, double h, double g, double &b, double &d, double &e) {
    const int na = 100, nw = 10;
    const double pi = M_PI;
    double Gmin = vmin / (f * h), Gmax = vmax / (f * h);
    std::vector<double> c0, c1, rhs;
    b = 5.0 / 6.0;
    for (int m = 1; m <= na; ++m)       // column-major flatten of [nw, na] meshgrid
        for (int n = 1; n <= nw; ++n) {
            double th = (m - 1) * pi / (4.0 * (na - 1));
            double G = 1.0 / (1.0 / Gmax + ((double)(n - 1) / (nw - 1)) * (1.0 / Gmin - 1.0 / Gmax));
            double P = std::cos(g * 2 * pi * std::cos(th) / G);
            double Q = std::cos(2 * pi * std::sin(th) / G);
            double S1 = (1 + 1 / (g * g)) * (G * G) * (1 - P - Q + P * Q);
            double S2 = (pi * pi) * (2 - P - Q);
            double S3 = (2 * pi * pi) * (1 - P * Q);
            double S4 = 2 * pi * pi + (G * G) * ((1 + 1 / (g * g)) * P * Q - P - Q / (g * g));
            c0.push_back(S2);
            c1.push_back(S3);
            rhs.push_back(S4 - b * S1);
        }
    const size_t M = rhs.size();
    std::vector<double> *cols[2] = {&c0, &c1};
    double R[2][2] = {{0, 0}, {0, 0}};
    for (int k = 0; k < 2; ++k) {
        std::vector<double> &a = *cols[k];
        double norm = 0;
        for (size_t i = k; i < M; ++i) norm += a[i] * a[i];
        norm = std::sqrt(norm);
        double alpha = a[k] > 0 ? -norm : norm;
        std::vector<double> v(M, 0.0);
        for (size_t i = k; i < M; ++i) v[i] = a[i];
        v[k] -= alpha;
        double vv = 0;
        for (size_t i = k; i < M; ++i) vv += v[i] * v[i];
        auto reflect = [&](std::vector<double> &x) {
            double s = 0;
            for (size_t i = k; i < M; ++i) s += v[i] * x[i];
            s = 2 * s / vv;
            for (size_t i = k; i < M; ++i) x[i] -= s * v[i];
        };
        for (int j = k; j < 2; ++j) reflect(*cols[j]);
        reflect(rhs);
        for (int j = k; j < 2; ++j) R[k][j] = (*cols[j])[k];
    }
    double x1 = rhs[1] / R[1][1];
    double x0 = (rhs[0] - R[0][1] * x1) / R[0][0];
    d = x0;
    e = x1;
}

// Whether the last refined solve of an operator shape, frequency and direction needed more than one
// correction (1088^2 x 512 at 0.3-1.4 MHz: every solve; 560^2 x 256: none).  It decides whether the
// pass after the first correction stores U and R at once or only measures the residual; both give
// the same values, so the choice changes the work only.
struct CorrectionKey {
    int ny, nx, npad, Rpad, adjoint;
    double f, dx, dy;
    bool operator<(const CorrectionKey &o) const {
        return std::tie(ny, nx, npad, Rpad, adjoint, f, dx, dy) <
               std::tie(o.ny, o.nx, o.npad, o.Rpad, o.adjoint, o.f, o.dx, o.dy);
    }
};
std::mutex g_corrections_mu;
std::map<CorrectionKey, bool> g_more_corrections;

struct Retained {
    cd *buf = nullptr;     // storage of the forward field (c128-sized)
    FieldView u;           // the forward field in buf
    cd *scale = nullptr;   // device [Rpad]
    int first = 0, nb = 0, Rpad = 0;
};

}  // namespace

struct wust_helmholtz {
    Operator op;
    PlainLib plain;
    int ws_Rpad = 0;
    cd *b64 = nullptr, *t64 = nullptr, *u64 = nullptr, *lam64 = nullptr;
    // leading elements of b64 known to hold zeros (the oracle evaluation undoes what it scatters); the
    // solve layout of a batch of Rpad right-hand sides is the first field_elems(Rpad) elements
    int64_t b64_zero_elems = 0;
    cf *s32 = nullptr, *d32 = nullptr, *b32 = nullptr;   // d32: the corrections after the first (refined)
    cf *r32 = nullptr;          // refinement residuals, the correction solves' right-hand sides
    int64_t s32_zero_elems = 0; // leading elements of s32 known to be zero (point right-hand sides
                                // write their nodes into it and clear them after their solve)
    int64_t s32_points_elems = 0; // nonzero while s32 holds a point right-hand side's nodes: the
                                  // zero extent to restore once they are cleared
    uint8_t *bmask = nullptr;   // per 64 values of the current right-hand side: any nonzero
    bool bmask_valid = false;
    cf *p_src = nullptr, *p_u = nullptr, *p_tmp = nullptr;
    double *norms = nullptr;    // [2][Rpad]
    double *loss_r = nullptr;   // [Rpad]
    cd *scales_dev = nullptr;   // source scales of every transmitter of the evaluation: [scales_cap]
    int64_t scales_cap = 0;
    cd *scale = nullptr;        // [Rpad]
    int *idx = nullptr;         // scratch for node coordinates: 4 * max(ntx, nrx)
    int idx_cap = 0;
    int idx_stride = 0;         // stride between the four coordinate arrays of the last evaluation
    std::vector<Retained> retained;
    std::vector<int> rx_x, rx_y;
    int ret_ntx = 0, ret_nrx = 0;
    std::vector<cd> scales_host;

    ~wust_helmholtz() {
        release_retained();
        if (hipDeviceSynchronize() == hipSuccess) {
            // big buffers go back to the pool for the next operator of this shape
            const int64_t nn = (int64_t)op.ny * op.nx, pg = (int64_t)op.npad * op.npad;
            pool_ws();
            dpool(op.ck2, nn);
            dpool(op.vsm, nn);
            dpool(op.w64, 9 * (int64_t)op.m * op.n);
            dpool(op.w32, 9 * (int64_t)op.m * op.npad);
            dpool(op.pages, pg * op.m);
            dpool(op.work, 6 * pg);
            if (op.nswork) dpool(op.nswork, 6 * pg);
        }
        release_ws();
        dfree(op.ck2);
        dfree(op.vsm);
        dfree(op.w64);
        dfree(op.w32);
        dfree(op.pages);
        dfree(op.work);
        dfree(op.pbuf);
        dfree(op.nswork);
        dfree(op.ns_norm);
        dfree(op.ns_rounds);
        dfree(op.colflags);
        dfree(op.gj_check);
        dfree(op.status);
        dfree(idx);
        dfree(scales_dev);
    }
    // the solve workspace of Rpad RHS into the pool (same sizes as ensure_ws allocates)
    void pool_ws() {
        if (!ws_Rpad) return;
        const int64_t cnt = field_elems(ws_Rpad);
        dpool(b64, cnt);
        dpool(t64, cnt);
        dpool(u64, cnt);
        dpool(lam64, cnt);
        if (s32) dpool(s32, cnt);
        if (d32) dpool(d32, cnt);
        if (r32) dpool(r32, cnt);
        if (bmask) dpool(bmask, cnt / 64);
        if (b32) dpool(b32, solve_work_elems(op, ws_Rpad));
    }
    void release_retained() {
        for (auto &r : retained) {
            dfree(r.buf);
            dfree(r.scale);
        }
        retained.clear();
    }
    void release_ws() {
        dfree(b64);
        dfree(t64);
        dfree(u64);
        dfree(lam64);
        dfree(s32);
        dfree(d32);
        dfree(r32);
        dfree(bmask);
        dfree(b32);
        dfree(p_src);
        dfree(p_u);
        dfree(p_tmp);
        dfree(norms);
        dfree(loss_r);
        dfree(scale);
        ws_Rpad = 0;
    }
    int64_t field_elems(int Rpad) const { return (int64_t)op.m * op.npad * Rpad; }
    void ensure_ws(int Rpad) {
        if (Rpad <= ws_Rpad) return;
        release_ws();
        int64_t cnt = field_elems(Rpad);
        b64 = dalloc<cd>(cnt);
        b64_zero_elems = 0;
        t64 = dalloc<cd>(cnt);
        u64 = dalloc<cd>(cnt);
        lam64 = dalloc<cd>(cnt);
        if (op.mode == WUST_MODE_PLAIN) {
            p_src = dalloc<cf>(cnt);
            p_u = dalloc<cf>(cnt);
            p_tmp = dalloc<cf>((int64_t)op.npad * Rpad);
        } else {
            s32 = dalloc<cf>(cnt);
            s32_zero_elems = s32_points_elems = 0;
            if (op.mode == WUST_MODE_REFINED) {
                d32 = dalloc<cf>(cnt);
                r32 = dalloc<cf>(cnt);
            }
            bmask = dalloc<uint8_t>(cnt / 64);
            b32 = dalloc<cf>(solve_work_elems(op, Rpad));
        }
        norms = dalloc<double>(2 * (int64_t)Rpad);
        loss_r = dalloc<double>(Rpad);
        scale = dalloc<cd>(Rpad);
        ws_Rpad = Rpad;
    }
    void ensure_scales(int64_t count) {
        if (count <= scales_cap) return;
        dfree(scales_dev);
        scales_dev = dalloc<cd>(count);
        scales_cap = count;
    }
    void ensure_idx(int count) {
        if (count <= idx_cap) return;
        dfree(idx);
        idx = dalloc<int>(4 * (int64_t)count);
        idx_cap = count;
    }
    uint32_t read_status(hipStream_t s) {
        uint32_t st = 0;
        WUST_HIP_CHECK(hipMemcpyAsync(&st, op.status, sizeof(st), hipMemcpyDeviceToHost, s));
        WUST_HIP_CHECK(hipStreamSynchronize(s));
        return st;
    }

    // b -> s32, and the numbers of leading and trailing block columns of b that are zero (point
    // sources and receivers leave the columns between the array and the domain edges empty; the
    // sweeps skip their forward elimination).  px / py (npts nodes, or null): every nonzero of b
    // lies at one of these nodes, and only they are read.  hpx (host copy of px, or null): the
    // zero columns are those without a node (a node whose values are all zero keeps its column, a
    // conservative count), found on the host without reading the device's column flags back.
    void demote_rhs(const cd *b, int64_t cnt, int Rpad, hipStream_t s, int &lead, int &trail, const int *px = nullptr,
                    const int *py = nullptr, int npts = 0, const int *hpx = nullptr) {
        lead = trail = 0;
        bmask_valid = false;
        const char *e = std::getenv("WUST_SKIP_ZERO_COLUMNS");
        if (!op.colflags || !bmask || (e && e[0] == '0')) {
            demote(b, s32, cnt, s);
            s32_zero_elems = s32_points_elems = 0;
            return;
        }
        WUST_HIP_CHECK(hipMemsetAsync(op.colflags, 0, sizeof(int) * op.m, s));
        if (px) {
            // s32 stays zero between point right-hand sides (their nodes are cleared after the solve),
            // so only its first use, or a dense use before, needs the full clear; until the nodes are
            // cleared it counts as not zero (an error in between leaves the next use a full clear)
            sparse_demote_flag(b, s32, cnt, px, py, npts, op.npad, Rpad, op.colflags, bmask, s, s32_zero_elems >= cnt);
            s32_zero_elems = 0;
            s32_points_elems = cnt;
        } else {
            demote_flag(b, s32, cnt, (int64_t)op.npad * Rpad, op.colflags, bmask, s);
            s32_zero_elems = s32_points_elems = 0;
        }
        bmask_valid = true;
        if (px && hpx) {
            int lo = op.m, hi = -1;
            for (int q = 0; q < npts; ++q)
                if (hpx[q] >= 0) {
                    lo = std::min(lo, hpx[q]);
                    hi = std::max(hi, hpx[q]);
                }
            lead = hi < 0 ? op.m : lo;
            trail = hi < 0 ? op.m : op.m - 1 - hi;
            return;
        }
        std::vector<int> f((size_t)op.m);
        WUST_HIP_CHECK(hipMemcpyAsync(f.data(), op.colflags, sizeof(int) * op.m, hipMemcpyDeviceToHost, s));
        WUST_HIP_CHECK(hipStreamSynchronize(s));
        while (lead < op.m && !f[lead]) ++lead;
        while (trail < op.m && !f[op.m - 1 - trail]) ++trail;
    }

    // Solve in the solve layout: b (c128, Rpad) -> out (c128).  b is not modified.  px / py (npts
    // nodes, or null): every nonzero of b lies at one of these nodes.
    // The solution lands in `out` (c128-sized) and `res` says how: the FP64 field, the c64 iterate in
    // out's first half (c64 mode, or refined with no correction needed), or the pair (iterate,
    // correction) in its two halves (refined with one correction, the usual case).
    wust_solve_stats solve_internal(const cd *b, cd *out, FieldView &res, int R, int Rpad, bool adjoint, double tol,
                                    int max_iter, bool want_residual, hipStream_t s, const int *px = nullptr,
                                    const int *py = nullptr, int npts = 0, const int *hpx = nullptr) {
        wust_solve_stats st{};
        res = FieldView{};
        const int64_t cnt = field_elems(Rpad);
        double *nb2 = norms, *nr2 = norms + Rpad;
        bmask_valid = false;   // set by demote_rhs for this b
        // one pass: U = U_old + D (FP64 U_old) or D0 + D (two c64 corrections), stored in U_new when
        // given; R = b - H U, per-RHS |R|^2 (and |b|^2 on the first pass); returns the max over RHS
        // of |R| / |b|
        bool first_pass = true;
        auto residual_pass = [&](const cd *u_old, const cf *d0, const cf *d, cd *u_new, cf *res32) {
            WUST_HIP_CHECK(hipMemsetAsync(first_pass ? nb2 : nr2, 0, sizeof(double) * Rpad * (first_pass ? 2 : 1), s));
            ResidualArgs a{};
            a.w64 = op.w64;
            a.u_old = u_old;
            a.d0 = d0;
            a.d32 = d;
            a.b64 = b;
            a.bmask = bmask_valid ? bmask : nullptr;
            a.u_new = u_new;
            a.res32 = res32;
            a.norm2 = nr2;
            a.bnorm2 = first_pass ? nb2 : nullptr;
            a.n = op.n;
            a.m = op.m;
            a.npad = op.npad;
            a.Rpad = Rpad;
            a.adjoint = adjoint ? 1 : 0;
            launch_residual(a, s);
            first_pass = false;
            std::vector<double> h(2 * (size_t)Rpad);
            WUST_HIP_CHECK(hipMemcpyAsync(h.data(), norms, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
            WUST_HIP_CHECK(hipStreamSynchronize(s));
            double worst = 0;
            for (int r = 0; r < R; ++r)
                if (h[r] > 0) worst = std::max(worst, std::sqrt(h[Rpad + r] / h[r]));
                else if (h[Rpad + r] > 0) worst = INFINITY;
            if (const char *e = std::getenv("WUST_REFINE_LOG"); e && e[0] == '1') {
                std::vector<double> rel;
                for (int r = 0; r < R; ++r) rel.push_back(h[r] > 0 ? std::sqrt(h[Rpad + r] / h[r]) : 0.0);
                std::sort(rel.begin(), rel.end());
                int over = 0;
                for (double v : rel) over += v > tol;
                std::fprintf(stderr, "refine %s R=%d: min %.3g p50 %.3g p90 %.3g max %.3g, %d above tol %.3g\n",
                             adjoint ? "adjoint" : "forward", R, rel.front(), rel[rel.size() / 2],
                             rel[(rel.size() * 9) / 10], rel.back(), over, tol);
            }
            return worst;
        };
        if (op.mode == WUST_MODE_PLAIN) {
            solve64_to_plain32(b, p_src, op.m, op.npad, R, Rpad, s);
            solve_plain(op, plain, p_src, p_u, p_tmp, R, adjoint, s);
            plain32_to_solve64(p_u, out, op.m, op.npad, R, Rpad, s);
            res.u64 = out;
            if (want_residual) {
                st.initial_residual = st.final_residual = residual_pass(out, nullptr, nullptr, nullptr, nullptr);
            }
            return st;
        }
        int lead = 0, trail = 0;
        cf *d0 = reinterpret_cast<cf *>(out), *d1 = d0 + cnt;   // out's two c64 halves
        demote_rhs(b, cnt, Rpad, s, lead, trail, px, py, npts, hpx);
        solve_engine(op, s32, d0, b32, Rpad, adjoint, s, lead, trail);
        if (px && s32_points_elems > 0) {
            launch_clear_points32(s32, px, py, npts, op.npad, Rpad, s);
            s32_zero_elems = s32_points_elems;
            s32_points_elems = 0;
        }
        res.a = d0;
        if (op.mode == WUST_MODE_C64) {
            if (want_residual) st.initial_residual = st.final_residual = residual_pass(nullptr, nullptr, d0, nullptr, nullptr);
            return st;
        }
        // refinement: U = D_1 + D_2 + ...; R = b - H U; D_{k+1} = solve(R).  D_1 and D_2 stay in single
        // precision in out's halves, and the pass after the first correction only measures the residual
        // of D_1 + D_2: with it below tol (the usual case) the result is that pair, and neither the FP64
        // sum nor the residual is stored.  Otherwise the pass runs again to store both (the same
        // values; it stores them at once where the last solve of this key needed another correction
        // too), and later passes accumulate in FP64, alternating between two buffers ordered so that
        // the result ends in `out`.
        double rel = residual_pass(nullptr, nullptr, d0, nullptr, r32);
        st.initial_residual = st.final_residual = rel;
        int it = 0;
        const CorrectionKey ckey{op.ny, op.nx, op.npad, Rpad, adjoint ? 1 : 0, op.ap.f, op.ap.dx, op.ap.dy};
        if (rel > tol && max_iter > 0) {
            solve_engine(op, r32, d1, b32, Rpad, adjoint, s);
            it = 1;
            res.b = d1;
            bool store = false;
            {
                std::lock_guard<std::mutex> lk(g_corrections_mu);
                auto f = g_more_corrections.find(ckey);
                store = f != g_more_corrections.end() && f->second;
            }
            cd *bufs[2] = {t64 == out ? u64 : t64, out};
            rel = st.final_residual = residual_pass(nullptr, d0, d1, store ? bufs[0] : nullptr, store ? r32 : nullptr);
            if (rel > tol && it < max_iter) {
                int cur = 0;
                if (!store) rel = residual_pass(nullptr, d0, d1, bufs[cur], r32);
                while (rel > tol && it < max_iter) {
                    solve_engine(op, r32, d32, b32, Rpad, adjoint, s);
                    ++it;
                    const cd *u_old = bufs[cur];
                    cur ^= 1;
                    rel = st.final_residual = residual_pass(u_old, nullptr, d32, bufs[cur], r32);
                }
                if (bufs[cur] != out)
                    WUST_HIP_CHECK(hipMemcpyAsync(out, bufs[cur], sizeof(cd) * cnt, hipMemcpyDeviceToDevice, s));
                res = FieldView{};
                res.u64 = out;
            }
        }
        st.iterations = it;
        if (it >= 1 && std::isfinite(st.final_residual)) {
            std::lock_guard<std::mutex> lk(g_corrections_mu);
            g_more_corrections[ckey] = it >= 2;
        }
        if (!std::isfinite(st.final_residual))
            throw Error(WUST_NUMERICAL_FAILURE, "non-finite residual in the refined solve");
        if (st.final_residual > tol && max_iter > 0)
            throw Error(WUST_NOT_CONVERGED, "refinement stopped at relative residual " +
                                                std::to_string(st.final_residual) + " > tol " + std::to_string(tol));
        return st;
    }

    void node_coords(const int64_t *index, int64_t count, std::vector<int> &xs, std::vector<int> &ys, bool allow_ring,
                     const char *what) {
        xs.resize(count);
        ys.resize(count);
        const int64_t N = (int64_t)op.ny * op.nx;
        for (int64_t i = 0; i < count; ++i) {
            int64_t node = index[i] - 1;
            require(node >= 0 && node < N, std::string(what) + " index outside the grid");
            int y = (int)(node % op.ny), x = (int)(node / op.ny);
            bool ring = (y == 0 || y == op.ny - 1 || x == 0 || x == op.nx - 1);
            if (ring) {
                require(allow_ring, std::string(what) + " must sample interior grid points");
                xs[i] = -1;
                ys[i] = -1;
            } else {
                xs[i] = x - 1;
                ys[i] = y - 1;
            }
        }
    }
};

extern "C" {

int wust_abi_version(void) { return WUST_ABI_VERSION; }

const char *wust_status_id(int status) {
    switch (status) {
        case WUST_OK: return "WUST:OK";
        case WUST_INVALID_INPUT: return "WUST:InvalidInput";
        case WUST_INVALID_OUTPUT: return "WUST:InvalidOutput";
        case WUST_HIP_ERROR: return "WUST:HIPError";
        case WUST_LIBRARY_ERROR: return "WUST:LibraryError";
        case WUST_SINGULAR_SCHUR_BLOCK: return "WUST:SingularSchurBlock";
        case WUST_NUMERICAL_FAILURE: return "WUST:NumericalFailure";
        case WUST_RESOURCE_CLEANUP: return "WUST:ResourceCleanup";
        case WUST_RUNTIME_ERROR: return "WUST:RuntimeError";
        case WUST_NOT_CONVERGED: return "WUST:NotConverged";
        default: return "WUST:Unknown";
    }
}

const char *wust_last_error(void) { return g_last_error.c_str(); }

int wust_stencil_opt_params(double vmin, double vmax, double f, double h, double g, double *b, double *d, double *e) {
    return guarded([&] {
        require(b && d && e, "null output pointer");
        require(std::isfinite(vmin) && std::isfinite(vmax) && vmin > 0 && vmax > vmin,
                "stencil_bounds must be a positive ordered [min,max] pair");
        require(f > 0 && h > 0 && g > 0, "frequency and spacings must be positive");
        stencil_params(vmin, vmax, f, h, g, *b, *d, *e);
    });
}

int wust_helmholtz_create(const wust_grid_desc *desc, const double *d_vel, const double *d_atten, int mode,
                          void *stream, wust_helmholtz **out) {
    if (out) *out = nullptr;
    return guarded([&] {
        require(desc && d_vel && out, "null argument");
        const int flags = mode & ~0xff;
        mode &= 0xff;
        require(mode == WUST_MODE_REFINED || mode == WUST_MODE_C64 || mode == WUST_MODE_PLAIN, "unknown mode");
        require((flags & ~(WUST_FACTOR_GAUSS_JORDAN | WUST_FACTOR_FIXED_BUDGET)) == 0, "unknown mode flag");
        require(!((flags & WUST_FACTOR_GAUSS_JORDAN) && (flags & WUST_FACTOR_FIXED_BUDGET)),
                "the Gauss-Jordan factor has no Newton-Schulz budget");
        const wust_grid_desc &g = *desc;
        require(g.ny >= 3 && g.nx >= 3, "Expected a positive finite [ny,nx] medium (min size 3)");
        require(std::isfinite(g.dx) && std::isfinite(g.dy) && g.dx > 0 && g.dy > 0, "Uniform increasing grids required");
        require(std::isfinite(g.frequency_hz) && g.frequency_hz > 0, "frequency must be positive");
        require(std::isfinite(g.pml_m) && g.pml_m > 0, "pml_m must be positive");
        require(std::isfinite(g.pml_strength), "pml_strength must be finite");
        require(std::isfinite(g.stencil_vmin) && std::isfinite(g.stencil_vmax) && g.stencil_vmin > 0 &&
                    g.stencil_vmax > g.stencil_vmin,
                "stencil_bounds must be a positive ordered [min,max] pair");
        if (g.dispersion) {
            require(d_atten == nullptr, "LDR9 requires zero attenuation");
            require(std::fabs(g.dy / g.dx - 1) < 1e-10, "LDR9 requires square grid cells");
            require(g.ldr9_time_step_s > 0 && g.ldr9_reference_speed_mps > 0, "LDR9 needs time step and reference speed");
        }
        hipStream_t s = (hipStream_t)stream;
        std::unique_ptr<wust_helmholtz> h(new wust_helmholtz());
        Operator &op = h->op;
        op.ny = g.ny;
        op.nx = g.nx;
        op.n = g.ny - 2;
        op.m = g.nx - 2;
        op.npad = (int)round_up(op.n, 32);
        // the factor's GEMMs address an inverse page with 32-bit byte offsets (npad^2 x 8 bytes)
        require((int64_t)op.npad * op.npad * 8 <= INT32_MAX, "grid too large: ny - 2 must not exceed 16352");
        op.p = op.m / 2;
        op.mode = mode;
        op.factor_method = (flags & WUST_FACTOR_GAUSS_JORDAN) ? FACTOR_GJ : FACTOR_AUTO;
        op.fixed_budget = (flags & WUST_FACTOR_FIXED_BUDGET) != 0;
        stencil_params(g.stencil_vmin, g.stencil_vmax, g.frequency_hz, g.dx, g.dy / g.dx, op.b, op.d, op.e);
        AssemblyParams &ap = op.ap;
        ap.ny = g.ny;
        ap.nx = g.nx;
        ap.npad = op.npad;
        ap.x0 = g.x0;
        ap.dx = g.dx;
        ap.y0 = g.y0;
        ap.dy = g.dy;
        ap.f = g.frequency_hz;
        ap.sgn = (g.sign_convention > 0) ? 1.0 : (g.sign_convention < 0 ? -1.0 : 0.0);
        ap.a0 = g.pml_strength;
        ap.L = g.pml_m;
        ap.b = op.b;
        ap.d = op.d;
        ap.e = op.e;
        ap.dispersion = g.dispersion;
        ap.ldr9_dt = g.ldr9_time_step_s;
        ap.ldr9_cr = g.ldr9_reference_speed_mps;
        const int64_t nodes = (int64_t)g.ny * g.nx;
        op.ck2 = dalloc<cd>(nodes);
        op.vsm = dalloc<cd>(nodes);
        op.w64 = dalloc<cd>(9 * (int64_t)op.m * op.n);
        op.w32 = dalloc<cf>(9 * (int64_t)op.m * op.npad);
        op.status = dalloc<uint32_t>(16);
        const int64_t pg = (int64_t)op.npad * op.npad;
        op.pages = dalloc<cf>(pg * op.m);
        op.work = dalloc<cf>(6 * pg);
        op.pbuf = dalloc<cf>(4 * 32 * 32);
        if (mode != WUST_MODE_PLAIN && op.factor_method == FACTOR_AUTO) {
            op.nswork = dalloc<cf>(6 * pg);
            op.ns_norm = dalloc<float>(2 * 8);
            op.ns_rounds = dalloc<int>(op.m);
        }
        if (mode != WUST_MODE_PLAIN) {
            op.colflags = dalloc<int>(op.m);
            op.gj_check = dalloc<float>(op.m);
        }
        op.factor_bytes = (int64_t)sizeof(cf) * pg * op.m;
        WUST_HIP_CHECK(hipMemsetAsync(op.status, 0, 8 * sizeof(uint32_t), s));

        hipEvent_t e0, e1, e2;
        WUST_HIP_CHECK(hipEventCreate(&e0));
        WUST_HIP_CHECK(hipEventCreate(&e1));
        WUST_HIP_CHECK(hipEventCreate(&e2));
        WUST_HIP_CHECK(hipEventRecord(e0, s));
        {
            ProfScope ps(PC_ASSEMBLY, s);
            launch_assembly(ap, d_vel, d_atten, op.ck2, op.vsm, op.w64, op.w32, op.status, s);
        }
        WUST_HIP_CHECK(hipEventRecord(e1, s));
        uint32_t st = h->read_status(s);
        if (st & ST_NONFINITE) throw Error(WUST_NUMERICAL_FAILURE, "LDR-9 produced invalid wavenumber or derivative");
        if (mode == WUST_MODE_PLAIN) factor_plain(op, h->plain, s);
        else factor_engine(op, h->plain, s);
        WUST_HIP_CHECK(hipEventRecord(e2, s));
        WUST_HIP_CHECK(hipEventSynchronize(e2));
        float ms1 = 0, ms2 = 0;
        WUST_HIP_CHECK(hipEventElapsedTime(&ms1, e0, e1));
        WUST_HIP_CHECK(hipEventElapsedTime(&ms2, e1, e2));
        op.assembly_ms = ms1;
        op.factor_ms = ms2;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        *out = h.release();
    });
}

int wust_helmholtz_destroy(wust_helmholtz *h) {
    return guarded([&] { delete h; });
}

int wust_pool_trim(void) {
    return guarded([&] {
        DevicePool &pl = pool();
        std::lock_guard<std::mutex> lk(pl.mu);
        int dev0 = current_device();
        for (auto &kv : pl.free_list) {
            WUST_HIP_CHECK(hipSetDevice(kv.first.first));
            for (void *p : kv.second) (void)hipFree(p);
        }
        WUST_HIP_CHECK(hipSetDevice(dev0));
        pl.free_list.clear();
        pl.held = 0;
        seed_cache_trim();
    });
}

int wust_helmholtz_info(const wust_helmholtz *h, wust_info *info) {
    return guarded([&] {
        require(h && info, "null argument");
        const Operator &op = h->op;
        info->ny = op.ny;
        info->nx = op.nx;
        info->n = op.n;
        info->m = op.m;
        info->npad = op.npad;
        info->mode = op.mode;
        info->b = op.b;
        info->d = op.d;
        info->e = op.e;
        info->assembly_ms = op.assembly_ms;
        info->factor_ms = op.factor_ms;
        info->min_pivot_ratio = op.min_pivot_ratio;
        info->factor_bytes = op.factor_bytes;
        info->inverse_residual = op.inverse_residual;
        info->iterative_blocks = op.ns_blocks;
        info->factor_fallback = op.factor_fallback;
        info->gj_residual = op.gj_residual;
    });
}

int wust_helmholtz_coefficients(const wust_helmholtz *h, void *Dd, void *Dl, void *Du, void *Ld, void *Ll, void *Lu,
                                void *Ud, void *Ul, void *Uu, void *stream) {
    return guarded([&] {
        require(h != nullptr, "null handle");
        launch_export(h->op.ny, h->op.nx, h->op.w64, (cd *)Dd, (cd *)Dl, (cd *)Du, (cd *)Ld, (cd *)Ll, (cd *)Lu,
                      (cd *)Ud, (cd *)Ul, (cd *)Uu, (hipStream_t)stream);
    });
}

int wust_helmholtz_solve(wust_helmholtz *h, const void *d_src, void *d_out, int64_t R, int adjoint, double tol,
                         int max_iter, void *stream, wust_solve_stats *stats) {
    return guarded([&] {
        require(h && d_src && d_out, "null argument");
        require(R >= 1 && R <= (1 << 20), "number of right-hand sides out of range");
        require(d_src != d_out, "solve output must not alias its source");
        require(max_iter >= 0, "max_iter must be non-negative");
        hipStream_t s = (hipStream_t)stream;
        hipEvent_t e0, e1;
        WUST_HIP_CHECK(hipEventCreate(&e0));
        WUST_HIP_CHECK(hipEventCreate(&e1));
        WUST_HIP_CHECK(hipEventRecord(e0, s));
        const int Rpad = (int)round_up(R, 16);
        h->ensure_ws(Rpad);
        h->b64_zero_elems = 0;
        ext_to_solve64((const cd *)d_src, h->b64, h->op.ny, h->op.nx, h->op.npad, (int)R, Rpad, nullptr, s);
        FieldView u;
        wust_solve_stats st =
            h->solve_internal(h->b64, h->u64, u, (int)R, Rpad, adjoint != 0, tol, max_iter, stats != nullptr, s);
        solve_to_ext(u, (cd *)d_out, h->op.ny, h->op.nx, h->op.npad, (int)R, Rpad, s);
        WUST_HIP_CHECK(hipEventRecord(e1, s));
        WUST_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        WUST_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        st.solve_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (stats) *stats = st;
    });
}

int wust_helmholtz_virtual_sources(const wust_helmholtz *h, const void *d_u, void *d_v, int64_t R, void *stream) {
    return guarded([&] {
        require(h && d_u && d_v, "null argument");
        require(R >= 1, "R must be positive");
        launch_virtual_sources((const cd *)d_u, h->op.vsm, (cd *)d_v, (int64_t)h->op.ny * h->op.nx, R,
                               (hipStream_t)stream);
    });
}

int wust_helmholtz_mass_apply(const wust_helmholtz *h, const void *d_in, void *d_out, int64_t R, int adjoint,
                              void *stream) {
    return guarded([&] {
        require(h && d_in && d_out, "null argument");
        require(d_in != d_out, "mass_apply output must not alias its input");
        require(R >= 1, "R must be positive");
        launch_mass_apply((const cd *)d_in, (cd *)d_out, h->op.ny, h->op.nx, R, h->op.d, h->op.e, adjoint ? 1 : 0,
                          (hipStream_t)stream);
    });
}

int wust_oracle_eval(wust_helmholtz *h, const int64_t *tx_index, int64_t n_tx, const int64_t *rx_index, int64_t n_rx,
                     const void *d_Y, const uint8_t *d_mask, int64_t batch, int need_gradient, int retain, double tol,
                     int max_iter, double *loss, void *d_grad, void *stream) {
    return guarded([&] {
        require(h && tx_index && rx_index && d_Y && d_mask && loss, "null argument");
        require(n_tx >= 1 && n_rx >= 1, "empty transmitter or receiver set");
        require(!need_gradient || d_grad, "gradient output required");
        hipStream_t s = (hipStream_t)stream;
        Operator &op = h->op;
        std::vector<int> tx_x, tx_y;
        h->node_coords(tx_index, n_tx, tx_x, tx_y, false, "tx_index");
        h->node_coords(rx_index, n_rx, h->rx_x, h->rx_y, true, "rx_index");
        if (batch <= 0 || batch > n_tx) batch = n_tx;
        h->release_retained();
        h->scales_host.assign((size_t)n_tx, mkd(0, 0));
        h->ret_ntx = (int)n_tx;
        h->ret_nrx = (int)n_rx;
        const int cap = (int)std::max(n_tx, n_rx);
        h->ensure_idx(cap);
        h->idx_stride = cap;
        int *rxx = h->idx, *rxy = h->idx + cap, *txx = h->idx + 2 * cap, *txy = h->idx + 3 * cap;
        WUST_HIP_CHECK(hipMemcpyAsync(rxx, h->rx_x.data(), sizeof(int) * n_rx, hipMemcpyHostToDevice, s));
        WUST_HIP_CHECK(hipMemcpyAsync(rxy, h->rx_y.data(), sizeof(int) * n_rx, hipMemcpyHostToDevice, s));
        WUST_HIP_CHECK(hipMemcpyAsync(txx, tx_x.data(), sizeof(int) * n_tx, hipMemcpyHostToDevice, s));
        WUST_HIP_CHECK(hipMemcpyAsync(txy, tx_y.data(), sizeof(int) * n_tx, hipMemcpyHostToDevice, s));
        // status words 6-9 (the factor uses 0-4): the gradient's non-finite flag, the source fits'
        // failure flags and the loss summed over the batches (a double), read back once at the end
        int *nonfinite = reinterpret_cast<int *>(op.status + 6);
        uint32_t *fit_status = op.status + 7;
        double *loss_acc = reinterpret_cast<double *>(op.status + 8);
        WUST_HIP_CHECK(hipMemsetAsync(op.status + 6, 0, 4 * sizeof(uint32_t), s));
        if (need_gradient)
            WUST_HIP_CHECK(hipMemsetAsync(d_grad, 0, sizeof(double) * (size_t)op.ny * op.nx, s));
        h->ensure_scales(n_tx);
        for (int64_t first = 0; first < n_tx; first += batch) {
            const int nb = (int)std::min<int64_t>(batch, n_tx - first);
            const int Rpad = (int)round_up(nb, 16);
            h->ensure_ws(Rpad);
            const int64_t cnt = h->field_elems(Rpad);
            // unit impulses at the batch's transmitters (Runtime/matlab/wust_oracle.m:31-34), into b64
            // kept zero between evaluations: each scatter below is undone after its solve
            if (h->b64_zero_elems < cnt) {
                WUST_HIP_CHECK(hipMemsetAsync(h->b64, 0, sizeof(cd) * cnt, s));
                h->b64_zero_elems = cnt;
            }
            const int64_t zero_after = h->b64_zero_elems;   // restored once this batch undoes its scatters
            h->b64_zero_elems = 0;
            launch_impulses(h->b64, txx + first, txy + first, nb, op.npad, Rpad, s);
            cd *ufwd = h->u64;
            if (retain) ufwd = dalloc<cd>(cnt);
            try {
                FieldView fwd;
                h->solve_internal(h->b64, ufwd, fwd, nb, Rpad, false, tol, max_iter, false, s, txx + first,
                                  txy + first, nb, tx_x.data() + first);
                // source fit, residual, loss and the adjoint source
                cd *adj = h->b64;  // forward RHS no longer needed: undo its impulses
                launch_impulses(adj, txx + first, txy + first, nb, op.npad, Rpad, s, 0.0);
                SourceFitArgs fa{};
                fa.u = fwd;
                fa.Y = (const cd *)d_Y;
                fa.mask = d_mask;
                fa.rx_x = rxx;
                fa.rx_y = rxy;
                fa.scale = h->scale;
                fa.loss = h->loss_r;
                fa.adj_src = need_gradient ? adj : nullptr;
                fa.status = fit_status;
                fa.ntx = (int)n_tx;
                fa.nrx = (int)n_rx;
                fa.tx_offset = (int)first;
                fa.npad = op.npad;
                fa.Rpad = Rpad;
                launch_source_fit(fa, nb, s);
                // the batch's losses into the total in RHS order, its scales beside the others'; a
                // failed fit is reported at the end (Runtime/matlab/wust_oracle.m:59 checks each
                // batch: the result is the same error, after the remaining batches' work)
                launch_accumulate(h->loss_r, nb, loss_acc, s);
                WUST_HIP_CHECK(hipMemcpyAsync(h->scales_dev + first, h->scale, sizeof(cd) * nb,
                                              hipMemcpyDeviceToDevice, s));
                if (need_gradient) {
                    FieldView lam;
                    h->solve_internal(adj, h->lam64, lam, nb, Rpad, true, tol, max_iter, false, s, rxx, rxy,
                                      (int)n_rx, h->rx_x.data());
                    GradientArgs ga{};
                    ga.u = fwd;
                    ga.lam = lam;
                    ga.vsm = op.vsm;
                    ga.scale = h->scale;
                    ga.grad = (double *)d_grad;
                    ga.d = op.d;
                    ga.e = op.e;
                    ga.ny = op.ny;
                    ga.n = op.n;
                    ga.m = op.m;
                    ga.npad = op.npad;
                    ga.Rpad = Rpad;
                    ga.R = nb;
                    launch_gradient(ga, s);
                }
                // undo the adjoint source (written at the receivers' nodes for RHS r < nb)
                launch_clear_points(adj, rxx, rxy, (int)n_rx, nb, op.npad, Rpad, s);
                h->b64_zero_elems = zero_after;
                if (retain) {
                    Retained rt;
                    rt.buf = ufwd;
                    rt.u = fwd;
                    rt.scale = dalloc<cd>(Rpad);
                    WUST_HIP_CHECK(hipMemsetAsync(rt.scale, 0, sizeof(cd) * Rpad, s));
                    WUST_HIP_CHECK(hipMemcpyAsync(rt.scale, h->scale, sizeof(cd) * nb, hipMemcpyDeviceToDevice, s));
                    rt.first = (int)first;
                    rt.nb = nb;
                    rt.Rpad = Rpad;
                    h->retained.push_back(rt);
                    ufwd = nullptr;
                }
            } catch (...) {
                if (retain && ufwd) dfree(ufwd);
                throw;
            }
        }
        // the reference's closing check (Runtime/matlab/wust_oracle.m:79-80): a finite loss and,
        // when computed, a finite gradient
        if (need_gradient) launch_nonfinite_flag((const double *)d_grad, (int64_t)op.ny * op.nx, nonfinite, s);
        uint32_t words[4];   // status words 6-9
        WUST_HIP_CHECK(hipMemcpyAsync(words, op.status + 6, sizeof(words), hipMemcpyDeviceToHost, s));
        WUST_HIP_CHECK(hipMemcpyAsync(h->scales_host.data(), h->scales_dev, sizeof(cd) * n_tx, hipMemcpyDeviceToHost,
                                      s));
        WUST_HIP_CHECK(hipStreamSynchronize(s));
        if (words[1] & ST_SCALE_FAILURE)
            throw Error(WUST_NUMERICAL_FAILURE, "Unusable source-scale denominator/numerator");
        double total;
        std::memcpy(&total, words + 2, sizeof(total));
        if (!std::isfinite(total) || words[0] != 0) throw Error(WUST_NUMERICAL_FAILURE, "Nonfinite objective/gradient");
        *loss = total;
    });
}

int wust_oracle_scales(const wust_helmholtz *h, void *scales_host) {
    return guarded([&] {
        require(h && scales_host, "null argument");
        std::memcpy(scales_host, h->scales_host.data(), sizeof(cd) * h->scales_host.size());
    });
}

int wust_directional_curvature(wust_helmholtz *h, const double *d_direction, const uint8_t *d_mask, double tol,
                               int max_iter, double *kappa, void *stream) {
    return guarded([&] {
        require(h && d_direction && d_mask && kappa, "null argument");
        require(!h->retained.empty(), "no retained oracle state (call wust_oracle_eval with retain)");
        hipStream_t s = (hipStream_t)stream;
        Operator &op = h->op;
        const int cap = h->idx_stride;
        int *rxx = h->idx, *rxy = h->idx + cap;
        double total = 0;
        for (auto &rt : h->retained) {
            h->ensure_ws(rt.Rpad);
            CurvSrcArgs ca{};
            ca.u = rt.u;
            ca.vsm = op.vsm;
            ca.direction = d_direction;
            h->b64_zero_elems = 0;
            ca.src = h->b64;
            ca.d = op.d;
            ca.e = op.e;
            ca.ny = op.ny;
            ca.n = op.n;
            ca.m = op.m;
            ca.npad = op.npad;
            ca.Rpad = rt.Rpad;
            launch_curvature_source(ca, s);
            FieldView w;
            h->solve_internal(h->b64, h->lam64, w, rt.nb, rt.Rpad, false, tol, max_iter, false, s);
            CurvGatherArgs ga{};
            ga.u = w;
            ga.scale = rt.scale;
            ga.mask = d_mask;
            ga.rx_x = rxx;
            ga.rx_y = rxy;
            ga.out = h->loss_r;
            ga.ntx = h->ret_ntx;
            ga.nrx = h->ret_nrx;
            ga.tx_offset = rt.first;
            ga.npad = op.npad;
            ga.Rpad = rt.Rpad;
            launch_curvature_gather(ga, rt.nb, s);
            std::vector<double> part(rt.nb);
            WUST_HIP_CHECK(hipMemcpyAsync(part.data(), h->loss_r, sizeof(double) * rt.nb, hipMemcpyDeviceToHost, s));
            WUST_HIP_CHECK(hipStreamSynchronize(s));
            for (double v : part) total += v;
        }
        *kappa = total;
    });
}

int wust_oracle_release(wust_helmholtz *h) {
    return guarded([&] {
        require(h != nullptr, "null handle");
        h->release_retained();
    });
}

}  // extern "C"

// ---- NCG step ----------------------------------------------------------------------

extern "C" {

int wust_roi_gradient(const double *d_g, const uint8_t *d_mask, int32_t ny, int32_t nx, double dx, double dy,
                      double c0, double frequency_hz, double cutoff, double order, double *d_roi, double *d_z,
                      void *stream) {
    return guarded([&] {
        require(d_g && d_mask && d_roi && d_z, "null argument");
        require(ny >= 1 && nx >= 1, "Expected a nonempty [ny,nx] grid");
        require(std::isfinite(cutoff) && cutoff >= 0, "cutoff must be a finite nonnegative scalar");
        if (cutoff > 0) {
            require(std::isfinite(dx) && std::isfinite(dy) && dx > 0 && dy > 0, "Uniform increasing grids required");
            require(std::isfinite(c0) && c0 > 0 && std::isfinite(frequency_hz) && frequency_hz > 0,
                    "c0 and frequency must be positive");
            require(std::isfinite(order) && order > 0, "filter order must be positive");
        }
        ncg_roi_gradient(d_g, d_mask, ny, nx, dx, dy, c0, frequency_hz, cutoff, order, d_roi, d_z,
                         (hipStream_t)stream);
    });
}

int wust_ncg_direction(const double *d_g, const double *d_z, const double *d_prev_z, double prev_gz,
                       const double *d_prev_direction, int same_frequency, int64_t count, double *d_direction,
                       double *beta, double *gz, double *slope, void *stream) {
    return guarded([&] {
        require(d_g && d_z && d_direction && beta && gz && count > 0, "null or empty argument");
        if (same_frequency && prev_gz > 0)
            require(d_prev_z && d_prev_direction, "previous z and direction are required on the same frequency");
        ncg_direction(d_g, d_z, d_prev_z, prev_gz, d_prev_direction, same_frequency != 0, count, d_direction, beta,
                      gz, slope, (hipStream_t)stream);
    });
}

int wust_linearized_step_length(double slope, double curvature, double *alpha) {
    return guarded([&] {
        require(alpha != nullptr, "null argument");
        if (!(std::isfinite(slope) && std::isfinite(curvature) && curvature >= 0))
            throw Error(WUST_NUMERICAL_FAILURE, "Nonfinite slope/curvature");
        if (!(curvature > 0 || slope == 0))
            throw Error(WUST_NUMERICAL_FAILURE, "Zero curvature with nonzero directional slope");
        *alpha = (curvature > 0) ? std::max(0.0, -slope / curvature) : 0.0;
    });
}

int wust_project_slowness_update(const double *d_c, const double *d_direction, double alpha, double step_damping,
                                 double max_update_mps, double bound_lo_mps, double bound_hi_mps,
                                 const uint8_t *d_update_mask, int64_t count, double *d_c_new,
                                 double *relative_update, void *stream) {
    return guarded([&] {
        require(d_c && d_direction && d_update_mask && d_c_new && count > 0, "null or empty argument");
        require(std::isfinite(bound_lo_mps) && std::isfinite(bound_hi_mps) && bound_lo_mps > 0 &&
                    bound_hi_mps > bound_lo_mps,
                "Invalid speed bounds");
        require(std::isfinite(max_update_mps) && max_update_mps > 0 && std::isfinite(step_damping) &&
                    step_damping > 0,
                "Expected positive finite max_update_mps and step_damping");
        require(std::isfinite(alpha) && alpha >= 0, "alpha must be finite and nonnegative");
        ncg_project(d_c, d_direction, alpha, step_damping, max_update_mps, bound_lo_mps, bound_hi_mps,
                    d_update_mask, count, d_c_new, relative_update, (hipStream_t)stream);
    });
}

// ---- RF -> frequency DTFT ---------------------------------------------------------------

int wust_rf_dtft(const void *d_rf, int rf_dtype, int64_t n_time, int32_t n_rx, int32_t n_tx, const double *d_time_s,
                 double dt, const double *frequencies_hz, int32_t n_freq, const uint8_t *d_active, void *d_pressure,
                 float *kernel_ms, void *stream) {
    return guarded([&] {
        require(d_rf && d_time_s && frequencies_hz && d_active && d_pressure, "null argument");
        require(rf_dtype == 0 || rf_dtype == 1, "rf_dtype must be 0 (float32) or 1 (float64)");
        require(n_time >= 2 && n_rx >= 1 && n_tx >= 1 && n_freq >= 1, "RF must be real [time,rx,tx] pressure");
        require(std::isfinite(dt) && dt > 0, "Invalid sample times");
        for (int i = 0; i < n_freq; ++i)
            require(std::isfinite(frequencies_hz[i]) && frequencies_hz[i] > 0 && frequencies_hz[i] < 0.5 / dt,
                    "Frequency outside saved Nyquist");
        rf_dtft(d_rf, rf_dtype == 1, n_time, n_rx, n_tx, d_time_s, dt, frequencies_hz, n_freq, d_active, d_pressure,
                (hipStream_t)stream, kernel_ms);
    });
}

int wust_rf_dtft_noalloc(const void *d_rf, int rf_dtype, int64_t n_time, int32_t n_rx, int32_t n_tx,
                         const double *d_time_s, double dt, const double *frequencies_hz, int32_t n_freq,
                         const uint8_t *d_active, void *d_pressure, void *stream) {
    return guarded([&] {
        require(d_rf && d_time_s && frequencies_hz && d_active && d_pressure, "null argument");
        require(rf_dtype == 0 || rf_dtype == 1, "rf_dtype must be 0 (float32) or 1 (float64)");
        require(n_time >= 2 && n_rx >= 1 && n_tx >= 1 && n_freq >= 1, "RF must be real [time,rx,tx] pressure");
        require(std::isfinite(dt) && dt > 0, "Invalid sample times");
        for (int i = 0; i < n_freq; ++i)
            require(std::isfinite(frequencies_hz[i]) && frequencies_hz[i] > 0 && frequencies_hz[i] < 0.5 / dt,
                    "Frequency outside saved Nyquist");
        rf_dtft_noalloc(d_rf, rf_dtype == 1, n_time, n_rx, n_tx, d_time_s, dt, frequencies_hz, n_freq, d_active,
                        d_pressure, (hipStream_t)stream);
    });
}

}  // extern "C"
