// FP64 field kernels: layout conversion, matrix-free nine-point residual with
// the refinement update fused in, source fit and loss, slowness gradient, mass
// stencil, virtual sources and the curvature source.  All of these are
// HBM-bound; each thread owns one (node, RHS) value and RHS is the fastest
// index of every solve-layout array so that the stencil neighbours, the
// coefficient rows and the stores are coalesced.
#include "fields.h"
#include "profile.h"

namespace wust {

namespace {

// ---------------------------------------------------------------------------
// external [R][Nx][Ny] complex128  <->  solve layout [m][npad][Rpad]
// ---------------------------------------------------------------------------
template <typename TO>
__device__ __forceinline__ TO conv_out(cd v);
template <>
__device__ __forceinline__ cd conv_out<cd>(cd v) { return v; }
template <>
__device__ __forceinline__ cf conv_out<cf>(cd v) { return tof(v); }

template <typename TO>
__global__ void __launch_bounds__(256) ext_to_solve_kernel(const cd *__restrict__ ext, TO *__restrict__ out, int ny,
                                                           int nx, int n, int npad, int R, int Rpad,
                                                           double *__restrict__ norm2) {
    __shared__ cd tile[32][33];
    const int xi = blockIdx.z;
    const int y0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    // load: lanes along y (contiguous in ext)
    for (int k = ty; k < 32; k += 8) {
        int yi = y0 + tx, r = r0 + k;
        cd v = mkd(0, 0);
        if (yi < n && r < R) v = ext[(int64_t)r * nx * ny + (int64_t)(xi + 1) * ny + (yi + 1)];
        tile[k][tx] = v;
    }
    __syncthreads();
    // store: lanes along r (contiguous in the solve layout)
    double acc = 0;
    for (int k = ty; k < 32; k += 8) {
        int yi = y0 + k, r = r0 + tx;
        if (yi < npad && r < Rpad) {
            cd v = tile[tx][k];
            out[((int64_t)xi * npad + yi) * Rpad + r] = conv_out<TO>(v);
            acc += abs2(v);
        }
    }
    if (norm2) {
        __shared__ double part[8][33];
        part[ty][tx] = acc;
        __syncthreads();
        if (ty == 0) {
            double s = 0;
            for (int k = 0; k < 8; ++k) s += part[k][tx];
            int r = r0 + tx;
            if (r < Rpad && s != 0) atomicAdd(norm2 + r, s);
        }
    }
}

__global__ void __launch_bounds__(256) solve_to_ext_kernel(FieldView in, cd *__restrict__ ext, int ny, int nx, int n,
                                                           int npad, int R, int Rpad) {
    __shared__ cd tile[32][33];
    const int x = blockIdx.z;  // full grid column 0..nx-1
    const int y0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const bool xin = (x >= 1 && x <= nx - 2);
    for (int k = ty; k < 32; k += 8) {
        int y = y0 + k, r = r0 + tx;
        cd v = mkd(0, 0);
        if (xin && y >= 1 && y <= ny - 2 && r < R) v = in.at(((int64_t)(x - 1) * npad + (y - 1)) * Rpad + r);
        tile[k][tx] = v;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        int y = y0 + tx, r = r0 + k;
        if (y < ny && r < R) ext[(int64_t)r * nx * ny + (int64_t)x * ny + y] = tile[tx][k];
    }
}

// ---------------------------------------------------------------------------
// fused refinement update + FP64 residual:  U_new = U_old + D;  Rs = B - H U_new
// (or H^H).  Couplings leaving the interior see zero (Dirichlet ring).
// ---------------------------------------------------------------------------
__constant__ int c_dy[9] = {0, -1, 1, 0, -1, 1, 0, -1, 1};
__constant__ int c_dx[9] = {0, 0, 0, -1, -1, -1, 1, 1, 1};
__constant__ int c_opp[9] = {DC, DS, DN, DE, DSE, DNE, DW, DSW, DNW};

// One thread per (RHS r, row y), a workgroup per 16 RHS x 16 rows that marches along a run of
// a.cols <= RES_COLS columns.  The iterate values of columns x-1, x, x+1 over the tile's rows and one halo
// row on each side sit in an LDS ring of four column slots: a column step writes the column that
// the previous step loaded (each thread its own row, 32 threads the halo rows), issues the loads
// of the column after it, and after one barrier reads its nine neighbours from the ring, so every
// iterate value is loaded from memory once per workgroup and the loads stay in flight for a whole
// step.  Neighbour addresses are clamped in range and the coefficient of a coupling that leaves
// the interior is zero, so a clamped node (one inside the thread's own stencil) adds an exact zero.
// The 9 x 16 coefficients of a column (the same for all RHS) are staged the same way,
// double-buffered; the right-hand-side mask bytes of the whole run are staged once.  Fused
// u += d; per-RHS |r|^2 and |b|^2 are summed over the run in registers and reduced per workgroup
// in LDS before the atomics.  The iterate is U_old (c128), D0 (c64: the first correction, kept in
// single precision so that the first pass stores no FP64 iterate) or zero.
constexpr int RES_COLS = 32;   // longest run; shorter runs keep small grids' workgroup counts up

template <bool HAS_U, bool HAS_D0, bool HAS_D>
__global__ void __launch_bounds__(256) residual_kernel(ResidualArgs a) {
    const int lr = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int r = blockIdx.x * 16 + lr;
    const int y0 = blockIdx.y * 16, yi = y0 + ly;
    const int n = a.n, m = a.m, npad = a.npad, Rpad = a.Rpad;
    const int xa = blockIdx.z * a.cols, xb = min(xa + a.cols, m);
    const int64_t colstride = (int64_t)npad * Rpad;
    const bool live = yi < n;
    // ring rows j = 0 .. 17 hold grid rows y0 - 1 + j, clamped to the interior; thread t holds its
    // own row (j = t / 16 + 1) and threads t < 32 also the halo rows j = 0 (t < 16) and j = 17
    const bool halo = threadIdx.x < 32;
    const int hj = threadIdx.x < 16 ? 0 : 17;
    const int64_t own_l = (int64_t)min(yi, n - 1) * Rpad + r;
    const int64_t halo_l = (int64_t)(threadIdx.x < 16 ? max(y0 - 1, 0) : min(y0 + 16, n - 1)) * Rpad + r;

    // the coefficient this thread stages (threads < 144): direction cdir = 3 ix + iy of tile row t % 16
    const int t = threadIdx.x < 144 ? threadIdx.x : 0;
    const int cdir = t >> 4, ty = y0 + (t & 15);
    const int iy = cdir % 3, ix = cdir / 3;
    const int cy = iy == 1 ? -1 : (iy == 2 ? 1 : 0), cx = ix == 1 ? -1 : (ix == 2 ? 1 : 0);
    const int yy = ty + cy;
    const bool yvalid = ty < n && yy >= 0 && yy < n;
    const int64_t plane = (int64_t)m * n;
    const int opp = 3 * (ix == 1 ? 2 : (ix == 2 ? 1 : 0)) + (iy == 1 ? 2 : (iy == 2 ? 1 : 0));
    const cd *wsrc = !a.adjoint ? a.w64 + cdir * plane + min(ty, n - 1)
                                : a.w64 + opp * plane + min(max(yy, 0), n - 1);
    auto coef = [&](int x) -> cd {        // column x's coefficient (x in range; neighbour clamped)
        const int xc = a.adjoint ? min(max(x + cx, 0), m - 1) : x;
        return wsrc[(int64_t)xc * n];
    };

    __shared__ cd ru[HAS_U ? 4 : 1][18][16];
    __shared__ cf rp[HAS_D0 ? 4 : 1][18][16];
    __shared__ cf rq[HAS_D ? 4 : 1][18][16];
    __shared__ uint8_t bms[RES_COLS][16];
    __shared__ cd wt[2][9][16];
    cd pu = mkd(0, 0), hu = mkd(0, 0);   // loaded column: own row, halo row
    cf pp = mk(0.f, 0.f), hp = mk(0.f, 0.f), pq = mk(0.f, 0.f), hq = mk(0.f, 0.f);
    auto load_col = [&](int x) {
        const int64_t base = (int64_t)min(max(x, 0), m - 1) * colstride;
        if (HAS_U) pu = a.u_old[base + own_l];
        if (HAS_D0) pp = a.d0[base + own_l];
        if (HAS_D) pq = a.d32[base + own_l];
        if (halo) {
            if (HAS_U) hu = a.u_old[base + halo_l];
            if (HAS_D0) hp = a.d0[base + halo_l];
            if (HAS_D) hq = a.d32[base + halo_l];
        }
    };
    auto store_col = [&](int x) {         // the loaded column into its ring slot (x may be -1)
        const int s = x & 3;
        if (HAS_U) ru[s][ly + 1][lr] = pu;
        if (HAS_D0) rp[s][ly + 1][lr] = pp;
        if (HAS_D) rq[s][ly + 1][lr] = pq;
        if (halo) {
            if (HAS_U) ru[s][hj][lr] = hu;
            if (HAS_D0) rp[s][hj][lr] = hp;
            if (HAS_D) rq[s][hj][lr] = hq;
        }
    };

    for (int e = threadIdx.x; e < a.cols * 16; e += 256) {
        // mask byte of column xa + e / 16, tile row e % 16 (the 16 RHS share one byte)
        const int bx = xa + (e >> 4), brow = min(y0 + (e & 15), n - 1);
        uint8_t v = 1;
        if (a.bmask && bx < xb) v = a.bmask[(((int64_t)bx * npad + brow) * Rpad + blockIdx.x * 16) >> 6];
        bms[e >> 4][e & 15] = v;
    }
    load_col(xa - 1);
    store_col(xa - 1);
    load_col(xa);
    store_col(xa);
    lds_barrier();
    const int64_t own = (int64_t)yi * Rpad + r;   // this thread's node in a column (unclamped row)
    load_col(xa + 1);
    cd wl = coef(xa);
    cd bn = (live && a.b64 && bms[0][ly]) ? a.b64[(int64_t)xa * colstride + own] : mkd(0, 0);

    double contrib = 0, bcontrib = 0;
#pragma unroll 1
    for (int x = xa; x < xb; ++x) {
        const int64_t o = (int64_t)x * colstride + own;
        // stage what the previous step loaded (column x + 1, column x's coefficient and b) ...
        store_col(x + 1);
        if (threadIdx.x < 144) {
            const bool valid = yvalid && x + cx >= 0 && x + cx < m;
            wt[x & 1][cdir][threadIdx.x & 15] = valid ? (a.adjoint ? conjd(wl) : wl) : mkd(0, 0);
        }
        const cd b = bn;
        // ... then issue the next step's loads, which stay in flight through this step
        load_col(x + 2);
        wl = coef(min(x + 1, m - 1));
        bn = mkd(0, 0);
        if (x + 1 < xb && live && a.b64 && bms[x + 1 - xa][ly]) bn = a.b64[o + colstride];
        lds_barrier();
        if (!live) {
            if (a.u_new) a.u_new[o] = mkd(0, 0);
            if (a.res32) a.res32[o] = mk(0.f, 0.f);
            if (a.res64) a.res64[o] = mkd(0, 0);
            continue;
        }
        // direction dd = 3 ix + iy: column x + DX[dd], row y + DY[dd]
        constexpr int DY[9] = {0, -1, 1, 0, -1, 1, 0, -1, 1};
        constexpr int DX[9] = {0, 0, 0, -1, -1, -1, 1, 1, 1};
        cd u[9];
#pragma unroll
        for (int dd = 0; dd < 9; ++dd) {
            const int s = (x + DX[dd]) & 3, j = ly + 1 + DY[dd];
            if (HAS_U || HAS_D0) {
                u[dd] = HAS_U ? ru[s][j][lr] : tod(rp[s][j][lr]);
                if (HAS_D) u[dd] = u[dd] + tod(rq[s][j][lr]);
            } else {
                u[dd] = HAS_D ? tod(rq[s][j][lr]) : mkd(0, 0);
            }
        }
        cd huacc = mkd(0, 0);
#pragma unroll
        for (int dd = 0; dd < 9; ++dd) huacc = huacc + wt[x & 1][dd][ly] * u[dd];
        const cd res = b - huacc;
        if (a.u_new) a.u_new[o] = u[DC];
        if (a.res32) a.res32[o] = tof(res);
        if (a.res64) a.res64[o] = res;
        contrib += abs2(res);
        bcontrib += abs2(b);
    }
    if (a.norm2 || a.bnorm2) {
        __shared__ double part[2][16][17];
        part[0][threadIdx.x >> 4][threadIdx.x & 15] = contrib;
        part[1][threadIdx.x >> 4][threadIdx.x & 15] = bcontrib;
        __syncthreads();
        if (threadIdx.x < 32) {
            const int which = threadIdx.x >> 4, lane = threadIdx.x & 15;
            double *dst = which ? a.bnorm2 : a.norm2;
            double s = 0;
            for (int k = 0; k < 16; ++k) s += part[which][k][lane];
            const int rr = blockIdx.x * 16 + lane;
            if (dst && s != 0) atomicAdd(dst + rr, s);
        }
    }
}

__global__ void demote_kernel(const cd *__restrict__ in, cf *__restrict__ out, int64_t count) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < count) out[i] = tof(in[i]);
}
// A wave's 64 values lie in one block of `per` (a multiple of 64), so one lane stores the flag.
__global__ void demote_flag_kernel(const cd *__restrict__ in, cf *__restrict__ out, int64_t count, int64_t per,
                                   int *__restrict__ flags, uint8_t *__restrict__ mask) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool nz = false;
    if (i < count) {
        const cd v = in[i];
        out[i] = tof(v);
        nz = (v.re != 0.0) || (v.im != 0.0);
    }
    const bool any = __any(nz);
    if ((threadIdx.x & 63) == 0 && i < count) {
        if (any) flags[i / per] = 1;
        mask[i >> 6] = any ? 1 : 0;
    }
}

// demote_flag for a right-hand side whose nonzeros all lie at the nodes (px[q], py[q]): converts
// those nodes' values for every RHS (out and mask pre-zeroed) and flags the block columns and
// 64-value chunks that hold a nonzero, as demote_flag does over the whole field
__global__ void sparse_demote_flag_kernel(const cd *__restrict__ in, cf *__restrict__ out, const int *__restrict__ px,
                                          const int *__restrict__ py, int npts, int npad, int Rpad,
                                          int *__restrict__ flags, uint8_t *__restrict__ mask) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (r >= Rpad || q >= npts || px[q] < 0) return;
    const int64_t i = ((int64_t)px[q] * npad + py[q]) * Rpad + r;
    const cd v = in[i];
    out[i] = tof(v);
    if (v.re != 0.0 || v.im != 0.0) {
        flags[px[q]] = 1;
        mask[i >> 6] = 1;
    }
}

// unit impulses: B[xi][yi][r] = 1 at the source node of RHS r (B pre-zeroed)
__global__ void impulse_kernel(cd *__restrict__ b, const int *__restrict__ sx, const int *__restrict__ sy, int R,
                               int npad, int Rpad, double value) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    b[((int64_t)sx[r] * npad + sy[r]) * Rpad + r] = mkd(value, 0);
}

// zeroes b at nodes (px[q], py[q]) for every RHS r < R (px[q] < 0: no node)
__global__ void clear_points_kernel(cd *__restrict__ b, const int *__restrict__ px, const int *__restrict__ py,
                                    int npts, int R, int npad, int Rpad) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (r >= R || q >= npts || px[q] < 0) return;
    b[((int64_t)px[q] * npad + py[q]) * Rpad + r] = mkd(0, 0);
}

// zeroes b at nodes (px[q], py[q]) for every RHS r < Rpad (single precision; px[q] < 0: no node)
__global__ void clear_points32_kernel(cf *__restrict__ b, const int *__restrict__ px, const int *__restrict__ py,
                                      int npts, int npad, int Rpad) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (r >= Rpad || q >= npts || px[q] < 0) return;
    b[((int64_t)px[q] * npad + py[q]) * Rpad + r] = mk(0.f, 0.f);
}

// *acc += v[0] + v[1] + ... + v[n-1], one addition at a time in index order (one thread: the
// same sum, rounding for rounding, as a host loop over the values)
__global__ void accumulate_kernel(const double *__restrict__ v, int n, double *__restrict__ acc) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double t = *acc;
    for (int i = 0; i < n; ++i) t += v[i];
    *acc = t;
}

// ---------------------------------------------------------------------------
// source fit (Runtime/matlab/wust_oracle.m:43-64): one block per RHS
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
    return s;
}

__global__ void __launch_bounds__(256) source_fit_kernel(SourceFitArgs a) {
    const int r = blockIdx.x;
    const int t = a.tx_offset + r;
    __shared__ double sh[8];
    double nr = 0, ni = 0, den = 0;
    for (int q = threadIdx.x; q < a.nrx; q += blockDim.x) {
        double mk_ = a.mask[t + (int64_t)a.ntx * q] ? 1.0 : 0.0;
        cd p = mkd(0, 0);
        if (a.rx_x[q] >= 0) p = a.u.at(((int64_t)a.rx_x[q] * a.npad + a.rx_y[q]) * a.Rpad + r);
        p = mk_ * p;
        cd d = mk_ * a.Y[t + (int64_t)a.ntx * q];
        cd cp = conjd(p) * d;
        nr += cp.re;
        ni += cp.im;
        den += abs2(p);
    }
    nr = block_sum(nr, sh);
    ni = block_sum(ni, sh);
    den = block_sum(den, sh);
    cd num = mkd(nr, ni);
    bool ok = isfinite(den) && den > 2.2250738585072014e-308 && isfinite(nr) && isfinite(ni);
    cd scale = ok ? cdiv(num, mkd(den, 0)) : mkd(0, 0);
    if (!(isfinite(scale.re) && isfinite(scale.im))) ok = false;
    if (!ok) {
        if (threadIdx.x == 0) atomicOr(a.status, (uint32_t)ST_SCALE_FAILURE);
        scale = mkd(0, 0);
    }
    double loss = 0;
    for (int q = threadIdx.x; q < a.nrx; q += blockDim.x) {
        double mk_ = a.mask[t + (int64_t)a.ntx * q] ? 1.0 : 0.0;
        cd p = mkd(0, 0);
        if (a.rx_x[q] >= 0) p = a.u.at(((int64_t)a.rx_x[q] * a.npad + a.rx_y[q]) * a.Rpad + r);
        p = mk_ * p;
        cd d = mk_ * a.Y[t + (int64_t)a.ntx * q];
        cd res = mk_ * (p * scale - d);
        loss += abs2(res);
        if (a.adj_src && a.rx_x[q] >= 0) a.adj_src[((int64_t)a.rx_x[q] * a.npad + a.rx_y[q]) * a.Rpad + r] = res;
    }
    loss = block_sum(loss, sh);
    if (threadIdx.x == 0) {
        a.scale[r] = scale;
        a.loss[r] = 0.5 * loss;
    }
}

// ---------------------------------------------------------------------------
// gradient (Runtime/matlab/wust_oracle.m:66-68):
// g(y,x) += sum_r -Re( conj(a_r vsm u) .* (K lambda) ),  K = mass stencil,
// lambda = 0 on the Dirichlet ring (massApply(., true) zeroes the input border)
// ---------------------------------------------------------------------------
// One workgroup per 8 rows x a run of a.cols <= GRAD_COLS columns, 32 RHS lanes per row; the
// lambda values of columns x-1, x, x+1 over the rows and one halo row on each side sit in an LDS
// ring of four column slots (zero outside the interior), filled one column per step as in the
// residual, so each lambda value is loaded once per workgroup instead of nine times.  The RHS run
// in chunks of 32; a node's sum over a chunk's lanes is a shuffle tree, added in chunk order to the
// node's accumulator in LDS, so every node is summed by one workgroup in a fixed order (no atomics).
constexpr int GRAD_COLS = 32;

// a field access split into its loads and their use, for a field known to be a c64 pair (PAIR) or
// an FP64 field, so that a loaded value occupies the registers of its one form
template <bool PAIR>
__device__ __forceinline__ FieldRaw field_load(const FieldView &v, int64_t i) {
    FieldRaw r{};
    if constexpr (PAIR) {
        r.a = v.a[i];
        if (v.b) r.b = v.b[i];
    } else {
        r.u = v.u64[i];
    }
    return r;
}
template <bool PAIR>
__device__ __forceinline__ cd field_value(const FieldView &v, const FieldRaw &r) {
    if constexpr (PAIR) return v.b ? tod(r.a) + tod(r.b) : tod(r.a);
    else return r.u;
}

template <bool PAIR_U, bool PAIR_L>
__global__ void __launch_bounds__(256) gradient_kernel(GradientArgs a) {
    const int lr = threadIdx.x & 31, ly = threadIdx.x >> 5;
    const int y0 = blockIdx.x * 8, yi = y0 + ly;
    const int n = a.n, m = a.m, npad = a.npad, Rpad = a.Rpad;
    const int xa = blockIdx.y * a.cols, xb = min(xa + a.cols, m);
    const int64_t colstride = (int64_t)npad * Rpad;
    const double kc = 1 - a.d - a.e, ke = a.d / 4, kd = a.e / 4;
    // ring rows j = 0 .. 9 hold grid rows y0 - 1 + j: thread t its own row (j = t / 32 + 1), threads
    // t < 64 also the halo rows j = 0 (t < 32) and j = 9
    const bool halo = threadIdx.x < 64;
    const int hj = threadIdx.x < 32 ? 0 : 9, hy = threadIdx.x < 32 ? y0 - 1 : y0 + 8;
    const bool own_in = yi < n, halo_in = halo && hy >= 0 && hy < n;
    const int own_y = min(yi, n - 1), halo_y = min(max(hy, 0), n - 1);
    __shared__ cd ring[4][10][32];
    __shared__ double acc[GRAD_COLS][8];
    for (int e = threadIdx.x; e < GRAD_COLS * 8; e += 256) acc[e >> 3][e & 7] = 0;
    for (int r0 = 0; r0 < a.R; r0 += 32) {
        const int r = r0 + lr;
        const bool rin = r < a.R;
        const int rc = min(r, Rpad - 1);              // in-range address for lanes beyond the RHS
        const cd sc = rin ? a.scale[r] : mkd(0, 0);
        // loaded column (own row, halo row), kept as loaded until it is stored into the ring
        FieldRaw pl{}, hl{};
        bool pl_in = false, hl_in = false;
        auto load_col = [&](int x) {
            const bool xin = x >= 0 && x < m;
            const int64_t base = (int64_t)min(max(x, 0), m - 1) * colstride + rc;
            pl = field_load<PAIR_L>(a.lam, base + (int64_t)own_y * Rpad);
            pl_in = xin && own_in && rin;
            if (halo) {
                hl = field_load<PAIR_L>(a.lam, base + (int64_t)halo_y * Rpad);
                hl_in = xin && halo_in && rin;
            }
        };
        auto store_col = [&](int x) {
            const int sl = x & 3;
            ring[sl][ly + 1][lr] = pl_in ? field_value<PAIR_L>(a.lam, pl) : mkd(0, 0);
            if (halo) ring[sl][hj][lr] = hl_in ? field_value<PAIR_L>(a.lam, hl) : mkd(0, 0);
        };
        auto load_u = [&](int x) {
            return field_load<PAIR_U>(a.u, (int64_t)min(x, m - 1) * colstride + (int64_t)own_y * Rpad + rc);
        };
        auto load_v = [&](int x) { return a.vsm[(int64_t)(min(x, m - 1) + 1) * a.ny + (own_y + 1)]; };
        __syncthreads();                              // the previous chunk's last reads of the ring
        load_col(xa - 1);
        store_col(xa - 1);
        load_col(xa);
        store_col(xa);
        load_col(xa + 1);
        FieldRaw un = load_u(xa);
        cd vn = load_v(xa);
#pragma unroll 1
        for (int x = xa; x < xb; ++x) {
            store_col(x + 1);
            const cd u = field_value<PAIR_U>(a.u, un), v = vn;
            load_col(x + 2);
            un = load_u(x + 1);
            vn = load_v(x + 1);
            __syncthreads();
            auto L = [&](int dy, int dx) { return ring[(x + dx) & 3][ly + 1 + dy][lr]; };
            // the reference's order: centre, the four edge neighbours, the four corners
            cd kl = kc * L(0, 0);
            kl = kl + ke * (L(-1, 0) + L(1, 0) + L(0, -1) + L(0, 1));
            kl = kl + kd * (L(-1, -1) + L(1, -1) + L(-1, 1) + L(1, 1));
            const cd av = sc * (v * u);
            double c = rin ? -(av.re * kl.re + av.im * kl.im) : 0.0;   // -Re(conj(av) kl)
#pragma unroll
            for (int off = 16; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
            if (lr == 0) acc[x - xa][ly] += c;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < GRAD_COLS * 8; e += 256) {
        const int x = xa + (e >> 3), y = y0 + (e & 7);
        if (x < xb && y < n) a.grad[(int64_t)(x + 1) * a.ny + (y + 1)] += acc[e >> 3][e & 7];
    }
}

// flag = 1 when any of v[0 .. count) is not finite (flag zeroed beforehand; only set here)
__global__ void __launch_bounds__(256) nonfinite_flag_kernel(const double *__restrict__ v, int64_t count,
                                                             int *__restrict__ flag) {
    bool bad = false;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        bad |= !isfinite(v[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// curvature source: src(i, r) = sum_j K(j - i) vsm(j) u(j, r) delta(j), j interior
__global__ void __launch_bounds__(256) curvature_source_kernel(CurvSrcArgs a) {
    const int r = blockIdx.x * 16 + (threadIdx.x & 15);
    const int yi = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int xi = blockIdx.z;
    if (yi >= a.npad || r >= a.Rpad) return;
    const int64_t o = ((int64_t)xi * a.npad + yi) * a.Rpad + r;
    if (yi >= a.n) {
        a.src[o] = mkd(0, 0);
        return;
    }
    const double kc = 1 - a.d - a.e, ke = a.d / 4, kd = a.e / 4;
    cd s = mkd(0, 0);
    for (int dd = 0; dd < 9; ++dd) {
        int yy = yi + c_dy[dd], xx = xi + c_dx[dd];
        if (yy < 0 || yy >= a.n || xx < 0 || xx >= a.m) continue;
        double kw = (dd == DC) ? kc : ((dd == DN || dd == DS || dd == DW || dd == DE) ? ke : kd);
        int64_t node = (int64_t)(xx + 1) * a.ny + (yy + 1);
        cd v = a.vsm[node] * a.u.at(((int64_t)xx * a.npad + yy) * a.Rpad + r);
        s = s + (kw * a.direction[node]) * v;
    }
    a.src[o] = s;
}

// kappa partials: sum_q | u(rx_q, r) a_r mask |^2  (one block per RHS)
__global__ void __launch_bounds__(256) curvature_gather_kernel(CurvGatherArgs a) {
    const int r = blockIdx.x;
    const int t = a.tx_offset + r;
    __shared__ double sh[8];
    double s = 0;
    for (int q = threadIdx.x; q < a.nrx; q += blockDim.x) {
        if (!a.mask[t + (int64_t)a.ntx * q] || a.rx_x[q] < 0) continue;
        cd u = a.u.at(((int64_t)a.rx_x[q] * a.npad + a.rx_y[q]) * a.Rpad + r);
        s += abs2(u * a.scale[r]);
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) a.out[r] = s;
}

// mass stencil on external-layout fields (Runtime/solver/HelmholtzSolver.m:205-220)
__global__ void mass_apply_kernel(const cd *__restrict__ in, cd *__restrict__ out, int ny, int nx, int64_t R,
                                  double d, double e, int adjoint) {
    int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t total = (int64_t)ny * nx * R;
    if (idx >= total) return;
    int y = (int)(idx % ny);
    int64_t t = idx / ny;
    int x = (int)(t % nx);
    int64_t r = t / nx;
    bool border = (y == 0 || y == ny - 1 || x == 0 || x == nx - 1);
    if (!adjoint && border) {
        out[idx] = mkd(0, 0);
        return;
    }
    const double kc = 1 - d - e, ke = d / 4, kd = e / 4;
    cd s = mkd(0, 0);
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int yy = y + dy, xx = x + dx;
            if (yy < 0 || yy >= ny || xx < 0 || xx >= nx) continue;
            if (adjoint && (yy == 0 || yy == ny - 1 || xx == 0 || xx == nx - 1)) continue;
            double kw = (dy == 0 && dx == 0) ? kc : ((dy == 0 || dx == 0) ? ke : kd);
            s = s + kw * in[r * ny * nx + (int64_t)xx * ny + yy];
        }
    out[idx] = s;
}

__global__ void virtual_sources_kernel(const cd *__restrict__ u, const cd *__restrict__ vsm, cd *__restrict__ v,
                                       int64_t nodes, int64_t R) {
    int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= nodes * R) return;
    v[idx] = vsm[idx % nodes] * u[idx];
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void ext_to_solve64(const cd *ext, cd *out, int ny, int nx, int npad, int R, int Rpad, double *norm2, hipStream_t s) {
    int n = ny - 2, m = nx - 2;
    dim3 grid((npad + 31) / 32, (Rpad + 31) / 32, m);
    hipLaunchKernelGGL(ext_to_solve_kernel<cd>, grid, dim3(256), 0, s, ext, out, ny, nx, n, npad, R, Rpad, norm2);
    WUST_LAUNCH_CHECK();
}
void solve_to_ext(const FieldView &in, cd *ext, int ny, int nx, int npad, int R, int Rpad, hipStream_t s) {
    int n = ny - 2;
    dim3 grid((ny + 31) / 32, (R + 31) / 32, nx);
    hipLaunchKernelGGL(solve_to_ext_kernel, grid, dim3(256), 0, s, in, ext, ny, nx, n, npad, R, Rpad);
    WUST_LAUNCH_CHECK();
}
void launch_residual(const ResidualArgs &a, hipStream_t s) {
    // runs of up to RES_COLS columns, as long as at least 4096 workgroups remain
    const int64_t tiles = (int64_t)(a.Rpad / 16) * (a.npad / 16) * a.m;
    ResidualArgs k = a;
    k.cols = (int)std::max<int64_t>(1, std::min<int64_t>(RES_COLS, tiles / 4096));
    dim3 grid(a.Rpad / 16, a.npad / 16, (a.m + k.cols - 1) / k.cols);
    require(!(a.u_old && a.d0), "residual: U_old and D0 are exclusive");
    require(a.Rpad % 16 == 0 && a.npad % 16 == 0 && a.n >= 1 && a.n <= a.npad, "residual: padded sizes");
    if (a.u_old && a.d32) launch_k(PC_RESIDUAL, residual_kernel<true, false, true>, grid, dim3(256), 0, s, k);
    else if (a.u_old) launch_k(PC_RESIDUAL, residual_kernel<true, false, false>, grid, dim3(256), 0, s, k);
    else if (a.d0 && a.d32) launch_k(PC_RESIDUAL, residual_kernel<false, true, true>, grid, dim3(256), 0, s, k);
    else if (a.d0) launch_k(PC_RESIDUAL, residual_kernel<false, true, false>, grid, dim3(256), 0, s, k);
    else if (a.d32) launch_k(PC_RESIDUAL, residual_kernel<false, false, true>, grid, dim3(256), 0, s, k);
    else launch_k(PC_RESIDUAL, residual_kernel<false, false, false>, grid, dim3(256), 0, s, k);
}
void demote_flag(const cd *in, cf *out, int64_t count, int64_t per, int *flags, uint8_t *mask, hipStream_t s) {
    require(count % 64 == 0 && per % 64 == 0, "demote_flag: sizes must be multiples of 64");
    hipLaunchKernelGGL(demote_flag_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, in, out, count,
                       per, flags, mask);
    WUST_LAUNCH_CHECK();
}
void demote(const cd *in, cf *out, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL(demote_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, in, out, count);
    WUST_LAUNCH_CHECK();
}
void launch_impulses(cd *b, const int *sx, const int *sy, int R, int npad, int Rpad, hipStream_t s, double value) {
    hipLaunchKernelGGL(impulse_kernel, dim3((R + 255) / 256), dim3(256), 0, s, b, sx, sy, R, npad, Rpad, value);
    WUST_LAUNCH_CHECK();
}
void sparse_demote_flag(const cd *in, cf *out, int64_t count, const int *px, const int *py, int npts, int npad,
                        int Rpad, int *flags, uint8_t *mask, hipStream_t s, bool out_zero) {
    require(count % 64 == 0, "sparse_demote_flag: size must be a multiple of 64");
    if (!out_zero) WUST_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(cf) * count, s));
    WUST_HIP_CHECK(hipMemsetAsync(mask, 0, count / 64, s));
    if (npts <= 0) return;
    hipLaunchKernelGGL(sparse_demote_flag_kernel, dim3((Rpad + 255) / 256, npts), dim3(256), 0, s, in, out, px, py,
                       npts, npad, Rpad, flags, mask);
    WUST_LAUNCH_CHECK();
}
void launch_clear_points(cd *b, const int *px, const int *py, int npts, int R, int npad, int Rpad, hipStream_t s) {
    if (npts <= 0 || R <= 0) return;
    hipLaunchKernelGGL(clear_points_kernel, dim3((R + 255) / 256, npts), dim3(256), 0, s, b, px, py, npts, R, npad,
                       Rpad);
    WUST_LAUNCH_CHECK();
}
void launch_clear_points32(cf *b, const int *px, const int *py, int npts, int npad, int Rpad, hipStream_t s) {
    if (npts <= 0 || Rpad <= 0) return;
    hipLaunchKernelGGL(clear_points32_kernel, dim3((Rpad + 255) / 256, npts), dim3(256), 0, s, b, px, py, npts, npad,
                       Rpad);
    WUST_LAUNCH_CHECK();
}
void launch_accumulate(const double *v, int n, double *acc, hipStream_t s) {
    hipLaunchKernelGGL(accumulate_kernel, dim3(1), dim3(64), 0, s, v, n, acc);
    WUST_LAUNCH_CHECK();
}
void launch_source_fit(const SourceFitArgs &a, int R, hipStream_t s) {
    launch_k(PC_SOURCE_FIT, source_fit_kernel, dim3(R), dim3(256), 0, s, a);
}
void launch_gradient(const GradientArgs &a, hipStream_t s) {
    // runs of up to GRAD_COLS columns, as long as at least 2048 workgroups remain
    require(a.Rpad % 16 == 0 && a.npad % 8 == 0 && a.n >= 1 && a.R <= a.Rpad, "gradient: padded sizes");
    GradientArgs k = a;
    const int64_t tiles = (int64_t)(a.npad / 8) * a.m;
    k.cols = (int)std::max<int64_t>(1, std::min<int64_t>(GRAD_COLS, tiles / 2048));
    dim3 grid(a.npad / 8, (a.m + k.cols - 1) / k.cols);
    require((a.u.u64 || a.u.a) && (a.lam.u64 || a.lam.a), "gradient: missing field");
    const bool pu = !a.u.u64, pl = !a.lam.u64;
    if (pu && pl) launch_k(PC_GRADIENT, gradient_kernel<true, true>, grid, dim3(256), 0, s, k);
    else if (pu) launch_k(PC_GRADIENT, gradient_kernel<true, false>, grid, dim3(256), 0, s, k);
    else if (pl) launch_k(PC_GRADIENT, gradient_kernel<false, true>, grid, dim3(256), 0, s, k);
    else launch_k(PC_GRADIENT, gradient_kernel<false, false>, grid, dim3(256), 0, s, k);
}
void launch_nonfinite_flag(const double *v, int64_t count, int *flag, hipStream_t s) {
    if (count <= 0) return;
    const unsigned blocks = (unsigned)std::min<int64_t>((count + 255) / 256, 1024);
    hipLaunchKernelGGL(nonfinite_flag_kernel, dim3(blocks), dim3(256), 0, s, v, count, flag);
    WUST_LAUNCH_CHECK();
}
void launch_curvature_source(const CurvSrcArgs &a, hipStream_t s) {
    dim3 grid((a.Rpad + 15) / 16, (a.npad + 15) / 16, a.m);
    hipLaunchKernelGGL(curvature_source_kernel, grid, dim3(256), 0, s, a);
    WUST_LAUNCH_CHECK();
}
void launch_curvature_gather(const CurvGatherArgs &a, int R, hipStream_t s) {
    hipLaunchKernelGGL(curvature_gather_kernel, dim3(R), dim3(256), 0, s, a);
    WUST_LAUNCH_CHECK();
}
void launch_mass_apply(const cd *in, cd *out, int ny, int nx, int64_t R, double d, double e, int adjoint,
                       hipStream_t s) {
    int64_t total = (int64_t)ny * nx * R;
    hipLaunchKernelGGL(mass_apply_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, ny, nx, R,
                       d, e, adjoint);
    WUST_LAUNCH_CHECK();
}
void launch_virtual_sources(const cd *u, const cd *vsm, cd *v, int64_t nodes, int64_t R, hipStream_t s) {
    int64_t total = nodes * R;
    hipLaunchKernelGGL(virtual_sources_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, u, vsm, v,
                       nodes, R);
    WUST_LAUNCH_CHECK();
}

}  // namespace wust
