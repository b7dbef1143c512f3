// Nine-point PML Helmholtz assembly on the device (FP64).
//
// Restates Runtime/solver/HelmholtzSolver.m:55-113 (complex speed, k, fine-mesh
// PML stretch, staggered A/B/C, q') and Runtime/solver/ninePointCoefficients.m
// (the nine coefficient formulas), one thread per grid node, without building
// any of the intermediate [2Ny-1, 2Nx-1] fine-mesh arrays: the PML factors are
// separable (ex depends on the fine column only, ey on the fine row only), so
// A, B and C are evaluated on the fly from two 1-D profiles.
#include "assembly.h"

namespace wust {

namespace {

struct Pml1D {
    double lo, hi;   // first / last node coordinate
    int nfine;       // 2*(N-1)+1 samples of the fine mesh
    double ctr, span, L, a0f2pi, inv2pif, sgn;

    __device__ __forceinline__ cd stretch(int fine) const {
        // linspace(lo, hi, nfine)[fine]; the last sample is exactly hi.
        double t = (fine == nfine - 1) ? hi : lo + fine * ((hi - lo) / (nfine - 1));
        double s = fmax(fabs(t - ctr) - span + L, 0.0) / L;
        double sig = a0f2pi * (s * s);
        return mkd(1.0, sig * sgn * inv2pif);
    }
};

__device__ __forceinline__ Pml1D make_pml(double lo, double hi, int n, double a0, double f, double L,
                                          double sgn) {
    Pml1D p;
    p.lo = lo;
    p.hi = hi;
    p.nfine = 2 * (n - 1) + 1;
    p.ctr = (lo + hi) / 2;
    p.span = (hi - lo) / 2;
    p.L = L;
    p.a0f2pi = 2 * M_PI * a0 * f;
    p.inv2pif = 1.0 / (2 * M_PI * f);
    p.sgn = sgn;
    return p;
}

// k-Wave matched LDR9 q and dq/ds (Runtime/solver/kwaveLdr9WavenumberSquared.m:13-61);
// standard 8-point Gauss-Legendre rule on [0, pi/4].
__device__ void ldr9(double c, double f, double h, double b, double d, double e, double dt, double cr,
                     double *q_out, double *dq_out) {
    const double nodes[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                             -0.1834346424956498, 0.1834346424956498,  0.5255324099163290,
                             0.7966664774136267,  0.9602898564975363};
    const double wts[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                           0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                           0.2223810344533745, 0.1012285362903763};
    double omega = 2 * M_PI * f;
    double ts = sin(omega * dt / 2);
    double z = (cr / c) * ts;
    double K = (2 * h / (cr * dt)) * asin(z);
    double N = 0, D = 0, Np = 0, Dp = 0;
    for (int j = 0; j < 8; ++j) {
        double th = (nodes[j] + 1) * (M_PI / 8);
        double ct = cos(th), st = sin(th);
        double X = K * ct, Y = K * st;
        double cx = cos(X), cy = cos(Y), sx = sin(X), sy = sin(Y);
        double lam = -4 * b + 2 * (2 * b - 1) * (cx + cy) + 4 * (1 - b) * cx * cy;
        double mass = 1 - d - e + (d / 2) * (cx + cy) + e * cx * cy;
        double lamp = -2 * (2 * b - 1) * (sx * ct + sy * st) - 4 * (1 - b) * (sx * ct * cy + cx * sy * st);
        double massp = -(d / 2) * (sx * ct + sy * st) - e * (sx * ct * cy + cx * sy * st);
        double w = 0.5 * wts[j];
        N += w * lam * mass;
        D += w * mass * mass;
        Np += w * (lamp * mass + lam * massp);
        Dp += w * (2 * mass * massp);
    }
    double Q = -N / D;
    double Qp = -(Np * D - N * Dp) / (D * D);
    double dKds = (2 * h * ts) / (dt * sqrt(1 - z * z));
    *q_out = Q / (h * h);
    *dq_out = (Qp / (h * h)) * dKds;
}

// Per-node: CK2 = C .* k.^2 and the virtual-source multiplier vsm = C .* q'.
__global__ void node_terms_kernel(AssemblyParams p, const double *__restrict__ vel,
                                  const double *__restrict__ atten, cd *__restrict__ ck2,
                                  cd *__restrict__ vsm, uint32_t *__restrict__ status) {
    int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t total = (int64_t)p.ny * p.nx;
    if (idx >= total) return;
    int y = (int)(idx % p.ny);
    int x = (int)(idx / p.ny);
    Pml1D px = make_pml(p.x0, p.x0 + (p.nx - 1) * p.dx, p.nx, p.a0, p.f, p.L, p.sgn);
    Pml1D py = make_pml(p.y0, p.y0 + (p.ny - 1) * p.dy, p.ny, p.a0, p.f, p.L, p.sgn);
    cd C = px.stretch(2 * x) * py.stretch(2 * y);
    double c = vel[idx];
    double at = atten ? atten[idx] : 0.0;
    // velocity = 1 ./ (1 ./ vel + 1i * SI * sign), SI = atten / (2 pi)
    cd inv_v = mkd(1.0 / c, at / (2 * M_PI) * p.sgn);
    cd velocity = cdiv(mkd(1.0, 0.0), inv_v);
    cd k2, dq;
    if (p.dispersion) {
        double q, dqs;
        ldr9(c, p.f, p.dx, p.b, p.d, p.e, p.ldr9_dt, p.ldr9_cr, &q, &dqs);
        // k = sqrt(q); k.^2 (q > 0 real)
        double k = sqrt(q);
        k2 = mkd(k * k, 0.0);
        dq = mkd(dqs, 0.0);
        if (!(isfinite(q) && q > 0 && isfinite(dqs))) atomicOr(status, (uint32_t)ST_NONFINITE);
    } else {
        cd k = cdiv(mkd(2 * M_PI * p.f, 0.0), velocity);
        k2 = k * k;
        dq = cdiv(mkd(8 * M_PI * M_PI * p.f * p.f, 0.0), velocity);
    }
    ck2[idx] = C * k2;
    vsm[idx] = C * dq;
}

// Per interior node: the nine coefficients of ninePointCoefficients.m:38-72,
// written as FP64 per-node rows w64[dir][xi][yi] (interior indices) and as the
// complex-FP32 block coefficients w32[dir][xi][npad] with the couplings to
// the Dirichlet ring (and the padding) zeroed.
__global__ void coefficients_kernel(AssemblyParams p, const cd *__restrict__ ck2, cd *__restrict__ w64,
                                    cf *__restrict__ w32) {
    int n = p.ny - 2, m = p.nx - 2;
    int yi = blockIdx.x * blockDim.x + threadIdx.x;  // 0..npad-1
    int xi = blockIdx.y;                             // 0..m-1
    if (yi >= p.npad || xi >= m) return;
    int64_t plane32 = (int64_t)m * p.npad;
    int64_t o32 = (int64_t)xi * p.npad + yi;
    if (yi >= n) {
        for (int dd = 0; dd < 9; ++dd) w32[dd * plane32 + o32] = mk(0.f, 0.f);
        return;
    }
    int y = yi + 1, x = xi + 1, ny = p.ny;
    Pml1D px = make_pml(p.x0, p.x0 + (p.nx - 1) * p.dx, p.nx, p.a0, p.f, p.L, p.sgn);
    Pml1D py = make_pml(p.y0, p.y0 + (p.ny - 1) * p.dy, p.ny, p.a0, p.f, p.L, p.sgn);
    // staggered terms: A(y,x) = ey(2y)/ex(2x+1) [x < Nx-1]; B(y,x) = ex(2x)/ey(2y+1) [y < Ny-1]
    auto A = [&](int yy, int xx) { return cdiv(py.stretch(2 * yy), px.stretch(2 * xx + 1)); };
    auto B = [&](int yy, int xx) { return cdiv(px.stretch(2 * xx), py.stretch(2 * yy + 1)); };
    auto CK = [&](int yy, int xx) { return ck2[yy + (int64_t)ny * xx]; };
    const double b = p.b, d = p.d, e = p.e;
    const double h2 = p.dx * p.dx;
    const double g = p.dy / p.dx, g2 = g * g;
    const double hb = (1 - b) / 2;
    cd w[9];
    // Dd(y,x)
    w[DC] = (1 - d - e) * CK(y, x) -
            b * ((1.0 / h2) * (A(y, x) + A(y, x - 1) + (1.0 / g2) * B(y, x) + (1.0 / g2) * B(y - 1, x)));
    // Dl(y-1,x): coefficient toward (y-1, x)
    w[DN] = (1.0 / h2) * ((b / g2) * B(y - 1, x) - hb * (A(y - 1, x) + A(y - 1, x - 1))) +
            (d / 4) * CK(y - 1, x);
    // Du(y,x): toward (y+1, x)
    w[DS] = (1.0 / h2) * ((b / g2) * B(y, x) - hb * (A(y + 1, x) + A(y + 1, x - 1))) + (d / 4) * CK(y + 1, x);
    // Ld(y,x-1): toward (y, x-1)
    w[DW] = (1.0 / h2) * (b * A(y, x - 1) - hb * ((1.0 / g2) * B(y, x - 1) + (1.0 / g2) * B(y - 1, x - 1))) +
            (d / 4) * CK(y, x - 1);
    // Ll(y-1,x-1): toward (y-1, x-1)
    w[DNW] = (1.0 / h2) * (hb * (A(y - 1, x - 1) + (1.0 / g2) * B(y - 1, x - 1))) + (e / 4) * CK(y - 1, x - 1);
    // Lu(y,x-1): toward (y+1, x-1)
    w[DSW] = (1.0 / h2) * (hb * (A(y + 1, x - 1) + (1.0 / g2) * B(y, x - 1))) + (e / 4) * CK(y + 1, x - 1);
    // Ud(y,x): toward (y, x+1)
    w[DE] = (1.0 / h2) * (b * A(y, x) - hb * ((1.0 / g2) * B(y, x + 1) + (1.0 / g2) * B(y - 1, x + 1))) +
            (d / 4) * CK(y, x + 1);
    // Ul(y-1,x): toward (y-1, x+1)
    w[DNE] = (1.0 / h2) * (hb * (A(y - 1, x) + (1.0 / g2) * B(y - 1, x + 1))) + (e / 4) * CK(y - 1, x + 1);
    // Uu(y,x): toward (y+1, x+1)
    w[DSE] = (1.0 / h2) * (hb * (A(y + 1, x) + (1.0 / g2) * B(y, x + 1))) + (e / 4) * CK(y + 1, x + 1);

    int64_t plane64 = (int64_t)m * n;
    int64_t o64 = (int64_t)xi * n + yi;
    for (int dd = 0; dd < 9; ++dd) w64[dd * plane64 + o64] = w[dd];
    // block coefficients: drop couplings that leave the interior
    bool top = (yi == 0), bot = (yi == n - 1), left = (xi == 0), right = (xi == m - 1);
    for (int dd = 0; dd < 9; ++dd) {
        bool drop = false;
        if ((dd == DN || dd == DNW || dd == DNE) && top) drop = true;
        if ((dd == DS || dd == DSW || dd == DSE) && bot) drop = true;
        if ((dd == DW || dd == DNW || dd == DSW) && left) drop = true;
        if ((dd == DE || dd == DNE || dd == DSE) && right) drop = true;
        w32[dd * plane32 + o32] = drop ? mk(0.f, 0.f) : tof(w[dd]);
    }
}

// Scatter the per-node rows into the reference's nine arrays (zeros elsewhere,
// Dd = 1 on the Dirichlet ring).
__global__ void export_kernel(int ny, int nx, const cd *__restrict__ w64, cd *Dd, cd *Dl, cd *Du, cd *Ld,
                              cd *Ll, cd *Lu, cd *Ud, cd *Ul, cd *Uu) {
    int y = blockIdx.x * blockDim.x + threadIdx.x;
    int x = blockIdx.y;
    if (y >= ny || x >= nx) return;
    int n = ny - 2, m = nx - 2;
    bool interior = (y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2);
    int64_t plane = (int64_t)n * m;
    auto W = [&](int dd, int yy, int xx) { return w64[dd * plane + (int64_t)(xx - 1) * n + (yy - 1)]; };
    const cd z = mkd(0, 0);
    if (Dd) Dd[y + (int64_t)ny * x] = interior ? W(DC, y, x) : mkd(1, 0);
    // arrays with Ny-1 rows: entry (yy, x) belongs to row node yy+1 (N-type) or yy (S-type)
    if (y < ny - 1) {
        int yn = y + 1;  // row node whose N-coefficient is stored at yy = y
        bool in_n = (yn >= 1 && yn <= ny - 2);
        bool in_s = (y >= 1 && y <= ny - 2);
        bool xin = (x >= 1 && x <= nx - 2);
        if (Dl) Dl[y + (int64_t)(ny - 1) * x] = (in_n && xin) ? W(DN, yn, x) : z;
        if (Du) Du[y + (int64_t)(ny - 1) * x] = (in_s && xin) ? W(DS, y, x) : z;
        if (x < nx - 1) {
            int64_t o = y + (int64_t)(ny - 1) * x;
            bool xl = (x + 1 >= 1 && x + 1 <= nx - 2);  // row column x+1 stores its W-type at x
            bool xu = (x >= 1 && x <= nx - 2);          // row column x stores its E-type at x
            if (Ll) Ll[o] = (in_n && xl) ? W(DNW, yn, x + 1) : z;
            if (Lu) Lu[o] = (in_s && xl) ? W(DSW, y, x + 1) : z;
            if (Ul) Ul[o] = (in_n && xu) ? W(DNE, yn, x) : z;
            if (Uu) Uu[o] = (in_s && xu) ? W(DSE, y, x) : z;
        }
    }
    if (x < nx - 1) {
        int64_t o = y + (int64_t)ny * x;
        bool yin = (y >= 1 && y <= ny - 2);
        if (Ld) Ld[o] = (yin && x + 1 <= nx - 2) ? W(DW, y, x + 1) : z;
        if (Ud) Ud[o] = (yin && x >= 1) ? W(DE, y, x) : z;
    }
}

}  // namespace

void launch_assembly(const AssemblyParams &p, const double *vel, const double *atten, cd *ck2, cd *vsm, cd *w64,
                     cf *w32, uint32_t *status, hipStream_t s) {
    int64_t total = (int64_t)p.ny * p.nx;
    hipLaunchKernelGGL(node_terms_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, p, vel, atten,
                       ck2, vsm, status);
    WUST_LAUNCH_CHECK();
    dim3 grid((p.npad + 63) / 64, p.nx - 2);
    hipLaunchKernelGGL(coefficients_kernel, grid, dim3(64), 0, s, p, ck2, w64, w32);
    WUST_LAUNCH_CHECK();
}

void launch_export(int ny, int nx, const cd *w64, cd *Dd, cd *Dl, cd *Du, cd *Ld, cd *Ll, cd *Lu, cd *Ud, cd *Ul,
                   cd *Uu, hipStream_t s) {
    dim3 grid((ny + 63) / 64, nx);
    hipLaunchKernelGGL(export_kernel, grid, dim3(64), 0, s, ny, nx, w64, Dd, Dl, Du, Ld, Ll, Lu, Ud, Ul, Uu);
    WUST_LAUNCH_CHECK();
}

}  // namespace wust
