#pragma once
#include "wust_common.h"

namespace wust {

// A field in the solve layout: the FP64 field u64, or the FP64 sum of two c64 fields a + b (b null:
// a alone).  A refined solve with a single correction leaves its result as the pair (first iterate,
// correction), so that no FP64 copy of it is written; the sum is formed where it is read, by the same
// FP64 addition that storing it would have used.
// load() and value() split one access, so that a load issued ahead of its use stays in flight until
// the use (the sum waits for both loads).
struct FieldRaw {
    cd u;
    cf a, b;
};
struct FieldView {
    const cd *u64 = nullptr;
    const cf *a = nullptr, *b = nullptr;
    __host__ __device__ __forceinline__ FieldRaw load(int64_t i) const {
        FieldRaw r{};
        if (u64) {
            r.u = u64[i];
        } else {
            r.a = a[i];
            if (b) r.b = b[i];
        }
        return r;
    }
    __host__ __device__ __forceinline__ cd value(const FieldRaw &r) const {
        if (u64) return r.u;
        return b ? tod(r.a) + tod(r.b) : tod(r.a);
    }
    __host__ __device__ __forceinline__ cd at(int64_t i) const { return value(load(i)); }
};

struct ResidualArgs {
    const cd *w64;        // [9][m][n]
    const cd *u_old;      // solve layout c128 or null
    const cf *d0;         // c64 iterate (the first correction) in place of u_old, or null
    const cf *d32;        // correction c64 or null
    const cd *b64;        // right-hand side c128 or null (= 0)
    const uint8_t *bmask; // null, or per 64 consecutive values of b64: 0 when all of them are zero
    cd *u_new;            // U_old + D (may be null)
    cf *res32;            // B - H U_new as c64 (may be null)
    cd *res64;            // same in c128 (may be null)
    double *norm2;        // per-RHS sum |res|^2 accumulated with atomics (may be null)
    double *bnorm2;       // per-RHS sum |b|^2 accumulated with atomics (may be null)
    int n, m, npad, Rpad, adjoint;
    int cols;             // columns per workgroup (set by launch_residual)
};

struct SourceFitArgs {
    FieldView u;          // forward field
    const cd *Y;          // [ntx, nrx] column-major
    const uint8_t *mask;  // [ntx, nrx]
    const int *rx_x, *rx_y;  // interior coordinates, -1 for receivers on the ring
    cd *scale;            // [R]
    double *loss;         // [R]
    cd *adj_src;          // solve layout c128, pre-zeroed (or null)
    uint32_t *status;
    int ntx, nrx, tx_offset, npad, Rpad;
};

struct GradientArgs {
    FieldView u, lam;     // forward and adjoint fields
    const cd *vsm;        // [Ny*Nx]
    const cd *scale;      // [R]
    double *grad;         // [Ny*Nx], accumulated
    double d, e;
    int ny, n, m, npad, Rpad, R;
    int cols;             // columns per workgroup (set by launch_gradient)
};

struct CurvSrcArgs {
    FieldView u;          // retained forward field
    const cd *vsm;
    const double *direction;  // [Ny*Nx]
    cd *src;              // solve layout c128
    double d, e;
    int ny, n, m, npad, Rpad;
};

struct CurvGatherArgs {
    FieldView u;          // the curvature solve's field
    const cd *scale;
    const uint8_t *mask;
    const int *rx_x, *rx_y;
    double *out;          // [R]
    int ntx, nrx, tx_offset, npad, Rpad;
};

void ext_to_solve64(const cd *ext, cd *out, int ny, int nx, int npad, int R, int Rpad, double *norm2, hipStream_t s);
void solve_to_ext(const FieldView &in, cd *ext, int ny, int nx, int npad, int R, int Rpad, hipStream_t s);
void launch_residual(const ResidualArgs &a, hipStream_t s);
void demote(const cd *in, cf *out, int64_t count, hipStream_t s);
// demote, and flags[i / per] = 1 for every block of `per` consecutive values that holds a nonzero
// (flags pre-zeroed; per a multiple of 64); mask[i / 64] = whether values 64 (i / 64) .. + 63 hold
// a nonzero (count a multiple of 64)
void demote_flag(const cd *in, cf *out, int64_t count, int64_t per, int *flags, uint8_t *mask, hipStream_t s);
// the same for a right-hand side whose nonzeros all lie at the nodes (px[q], py[q]) (px < 0: none),
// reading only those nodes (flags pre-zeroed; count a multiple of 64); out is zeroed first unless
// out_zero says that it already is
void sparse_demote_flag(const cd *in, cf *out, int64_t count, const int *px, const int *py, int npts, int npad,
                        int Rpad, int *flags, uint8_t *mask, hipStream_t s, bool out_zero = false);
// b = 0 at nodes (px[q], py[q]) for every RHS r < Rpad, single precision (px[q] < 0: no node)
void launch_clear_points32(cf *b, const int *px, const int *py, int npts, int npad, int Rpad, hipStream_t s);
// *acc += v[0] + ... + v[n-1], added one at a time in index order on the device
void launch_accumulate(const double *v, int n, double *acc, hipStream_t s);
// b at node (sx[r], sy[r]) of RHS r = value, r < R (value 0 undoes the impulses)
void launch_impulses(cd *b, const int *sx, const int *sy, int R, int npad, int Rpad, hipStream_t s,
                     double value = 1.0);
// b = 0 at nodes (px[q], py[q]) for every RHS r < R (px[q] < 0: no node)
void launch_clear_points(cd *b, const int *px, const int *py, int npts, int R, int npad, int Rpad, hipStream_t s);
void launch_source_fit(const SourceFitArgs &a, int R, hipStream_t s);
void launch_gradient(const GradientArgs &a, hipStream_t s);
// *flag |= 1 when any of v[0 .. count) is not finite (the caller zeroes *flag first)
void launch_nonfinite_flag(const double *v, int64_t count, int *flag, hipStream_t s);
void launch_curvature_source(const CurvSrcArgs &a, hipStream_t s);
void launch_curvature_gather(const CurvGatherArgs &a, int R, hipStream_t s);
void launch_mass_apply(const cd *in, cd *out, int ny, int nx, int64_t R, double d, double e, int adjoint,
                       hipStream_t s);
void launch_virtual_sources(const cd *u, const cd *vsm, cd *v, int64_t nodes, int64_t R, hipStream_t s);

}  // namespace wust
