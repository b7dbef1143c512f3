/*
 * wust_hip.h -- C ABI of the MI355X (gfx950) Helmholtz solve-and-gradient engine.
 *
 * One opaque handle (wust_helmholtz) owns one device-resident operator: the
 * assembled nine-point PML stencil of one sound-speed model at one frequency
 * and its block-tridiagonal factorization.  Every entry point takes plain
 * pointers and sizes; device pointers are HIP device memory, host pointers
 * are marked "host".  `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  All entry points return a wust_status; on failure
 * wust_last_error() returns a thread-local message and wust_status_id()
 * maps the status to the reference's error identifier ("WUST:InvalidInput",
 * "WUST:SingularSchurBlock", ...; Runtime/solver/block_lu_mex_utils.h:37-159).
 *
 * Array layouts follow the reference (MATLAB column-major):
 *   grid arrays  [Ny, Nx]       element (y, x)    at  y + Ny*x
 *   fields       [Ny, Nx, R]    element (y, x, r) at  y + Ny*x + Ny*Nx*r
 *   observations [Ntx, Nrx]     element (t, q)    at  t + Ntx*q
 * complex values are interleaved (re, im) pairs of double.
 * Node indices (tx_index / rx_index) are MATLAB 1-based linear indices
 * y + Ny*x + 1 (Runtime/matlab/wust_geometry.m:27).
 */
#ifndef WUST_HIP_H
#define WUST_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WUST_ABI_VERSION 3

typedef enum {
    WUST_OK = 0,
    WUST_INVALID_INPUT = 1,        /* WUST:InvalidInput */
    WUST_INVALID_OUTPUT = 2,       /* WUST:InvalidOutput */
    WUST_HIP_ERROR = 3,            /* WUST:CUDAError analogue (HIP runtime) */
    WUST_LIBRARY_ERROR = 4,        /* WUST:CUBLASError / WUST:CUSOLVERError analogue */
    WUST_SINGULAR_SCHUR_BLOCK = 5, /* WUST:SingularSchurBlock */
    WUST_NUMERICAL_FAILURE = 6,    /* WUST:NumericalFailure */
    WUST_RESOURCE_CLEANUP = 7,     /* WUST:ResourceCleanup */
    WUST_RUNTIME_ERROR = 8,        /* WUST:RuntimeError */
    WUST_NOT_CONVERGED = 9         /* iterative refinement missed its tolerance */
} wust_status;

/* Arithmetic of the factor and solves. */
typedef enum {
    /* Hand-written gfx950 kernels: complex-FP32 MFMA block factor and sweeps,
     * FP64 matrix-free residual, FP64 refinement accumulation. */
    WUST_MODE_REFINED = 0,
    /* The same complex-FP32 kernels without refinement. */
    WUST_MODE_C64 = 1,
    /* The reference's GPU algorithm on vendor libraries, unrefined: one
     * rocSOLVER getrf+getrs per Schur block and one rocBLAS cgemm plus a
     * tridiagonal kernel per grid column (the "plain complex-FP32 block LU"
     * of Runtime/solver/decompBlockLU.cu / applyBlockLU.cu). */
    WUST_MODE_PLAIN = 2
} wust_mode;

/* Flag OR-ed into the mode of wust_helmholtz_create (engine modes only):
 * invert every Schur block by Gauss-Jordan elimination.  Without it, blocks
 * that have an already-inverted neighbour in their chain are inverted by
 * Newton-Schulz iteration started from that neighbour's inverse, with a
 * device-side residual bound; if any block misses the bound the whole
 * factor is redone by Gauss-Jordan (wust_info.factor_fallback). */
#define WUST_FACTOR_GAUSS_JORDAN 0x100
/* Flag OR-ed into the mode of wust_helmholtz_create (engine modes only, no
 * reference counterpart): Newton-Schulz with the default round budgets.  Without
 * it, the round budget of each chain step is learned from the previous
 * factorization of an operator with the same shape, frequency and grid (a
 * process-wide profile), so the c64 factor -- and the rounding of everything
 * computed from it -- depends on which operators the process factored before.
 * With it the factor depends only on the operator. */
#define WUST_FACTOR_FIXED_BUDGET 0x200

typedef struct wust_helmholtz wust_helmholtz;

/* Grid, frequency and PML description: the scalar arguments of
 * HelmholtzSolver(x, y, vel, atten, f, signConvention, a0, L_PML, options)
 * (Runtime/solver/HelmholtzSolver.m:32).  Node coordinates are
 * x_i = x0 + i*dx, y_i = y0 + i*dy. */
typedef struct {
    int32_t ny, nx;
    double x0, dx, y0, dy;
    double frequency_hz;
    double sign_convention;     /* wust_oracle passes -1 */
    double pml_strength;        /* a0 */
    double pml_m;               /* L_PML [m] */
    double stencil_vmin, stencil_vmax;   /* options.stencil_bounds */
    int32_t dispersion;         /* 0 continuum, 1 kwave-ldr9 */
    double ldr9_time_step_s;    /* options.dispersion.time_step_s */
    double ldr9_reference_speed_mps;
} wust_grid_desc;

typedef struct {
    int32_t ny, nx, n, m, npad;    /* n = ny-2 block size, m = nx-2 blocks, npad = padded n */
    int32_t mode;
    double b, d, e;                /* stencil parameters */
    double assembly_ms, factor_ms; /* device time, hipEvent-timed */
    double min_pivot_ratio;        /* smallest |pivot| / max |column entry| seen in the factor */
    int64_t factor_bytes;          /* device bytes held by the factors */
    double inverse_residual;       /* largest bound on ||I - T_j X_j||_F over Newton-Schulz blocks */
    int32_t iterative_blocks;      /* Schur blocks inverted by Newton-Schulz */
    int32_t factor_fallback;       /* 1: Newton-Schulz missed its bound, factor redone by Gauss-Jordan;
                                      2: a pivot-free Gauss-Jordan block failed its residual check (or met a
                                         zero pivot), factor redone with partial pivoting (rocSOLVER getrf) */
    double gj_residual;            /* largest ||I - T_j X_j||_F over the Gauss-Jordan blocks */
} wust_info;

typedef struct {
    int32_t iterations;            /* refinement corrections applied (0 = one plain solve) */
    double initial_residual;       /* max over RHS of ||b - H u0|| / ||b|| after the first solve */
    double final_residual;         /* same after the last correction */
    double solve_ms;               /* device time of the whole call */
} wust_solve_stats;

/* ---- library ------------------------------------------------------------------ */
int wust_abi_version(void);
const char *wust_status_id(int status);
const char *wust_last_error(void);

/* Optimal nine-point mass coefficients, b = 5/6 fixed (host computation).
 * Replaces stencilOptParams(vmin, vmax, f, h, g) (Runtime/solver/stencilOptParams.m:1-38). */
int wust_stencil_opt_params(double vmin, double vmax, double f, double h, double g,
                            double *b, double *d, double *e);

/* ---- operator ------------------------------------------------------------------- */
/* Assemble the operator on the device from the sound speed (and optional
 * attenuation, may be NULL = 0), then factor it.  d_vel, d_atten: device
 * double [Ny, Nx].  Replaces the HelmholtzSolver constructor
 * (Runtime/solver/HelmholtzSolver.m:32-172), ninePointCoefficients
 * (Runtime/solver/ninePointCoefficients.m:1-74) and decompBlockLU
 * (Runtime/solver/decompBlockLU.cu:173-396).  The factor is held by the
 * handle (the reference returns it as invT). */
int wust_helmholtz_create(const wust_grid_desc *desc, const double *d_vel, const double *d_atten,
                          int mode, void *stream, wust_helmholtz **out);
int wust_helmholtz_destroy(wust_helmholtz *h);

/* Device memory of destroyed operators is kept (up to WUST_POOL_MB megabytes, default 98304)
 * for the next operator with the same sizes, and the inverse pages of the leading (PML) chain
 * steps of the last factorization per operator key (up to WUST_SEED_CACHE_MB, default 4096) seed
 * the next factorization with that key; wust_pool_trim frees both.  No reference counterpart (the
 * MEX gateways allocate and free per call). */
int wust_pool_trim(void);
int wust_helmholtz_info(const wust_helmholtz *h, wust_info *info);

/* The nine coefficient arrays of ninePointCoefficients in the reference
 * shapes, written as complex128 device arrays (any pointer may be NULL):
 * Dd [Ny,Nx]; Dl, Du [Ny-1,Nx]; Ld, Ud [Ny,Nx-1]; Ll, Lu, Ul, Uu [Ny-1,Nx-1]. */
int wust_helmholtz_coefficients(const wust_helmholtz *h, void *Dd, void *Dl, void *Du, void *Ld,
                                void *Ll, void *Lu, void *Ud, void *Ul, void *Uu, void *stream);

/* Solve H u = src (adjoint = 0) or H^H u = src (adjoint = 1) for R right-hand
 * sides.  d_src, d_out: device complex128 [Ny, Nx, R] with the operator's Ny
 * and Nx: the library reads and writes exactly Ny * Nx * R values and cannot
 * check a buffer's size (the reference rejects mismatched shapes,
 * Runtime/solver/applyBlockLU.cu:380-388; the Python and MEX layers check
 * them before calling).  Sources must vanish
 * on the outer Dirichlet ring, and the output is zero there (the reference
 * block LU keeps those degrees of freedom at zero,
 * Runtime/solver/applyBlockLU.cu:7-10).  In WUST_MODE_REFINED the FP64
 * residual is driven below `tol` (relative, max over RHS) within `max_iter`
 * corrections.  Replaces HelmholtzSolver.solve
 * (Runtime/solver/HelmholtzSolver.m:174-203) and applyBlockLU
 * (Runtime/solver/applyBlockLU.cu:291-481).  stats may be NULL. */
int wust_helmholtz_solve(wust_helmholtz *h, const void *d_src, void *d_out, int64_t R, int adjoint,
                         double tol, int max_iter, void *stream, wust_solve_stats *stats);

/* v = (C .* q') .* u, the second output of solve (HelmholtzSolver.m:197-199).
 * d_u, d_v: device complex128 [Ny, Nx, R]. */
int wust_helmholtz_virtual_sources(const wust_helmholtz *h, const void *d_u, void *d_v, int64_t R,
                                   void *stream);

/* Exact mass-stencil derivative M v (adjoint = 0, output border zeroed) or
 * M^T v (adjoint = 1, input border zeroed first).  Replaces massApply
 * (Runtime/solver/HelmholtzSolver.m:205-220).  Device complex128 [Ny,Nx,R]. */
int wust_helmholtz_mass_apply(const wust_helmholtz *h, const void *d_in, void *d_out, int64_t R,
                              int adjoint, void *stream);

/* ---- objective / derivatives ------------------------------------------------------ */
/* Loss, gradient and (retain != 0) curvature state of wust_oracle
 * (Runtime/matlab/wust_oracle.m:1-86) for the handle's model and frequency.
 *   tx_index, rx_index: host int64, 1-based node indices (n_tx, n_rx entries)
 *   d_Y:    device complex128 [n_tx, n_rx] observations of this frequency
 *   d_mask: device uint8 [n_tx, n_rx] (nonzero = valid channel)
 *   batch:  source batch size (cfg.source_batch_size); <= 0 means all
 *   d_grad: device double [Ny, Nx], written when need_gradient != 0
 *   loss:   host double
 * Source scale, residual and loss are FP64; refinement as in solve().  A non-finite loss
 * or gradient value is WUST_NUMERICAL_FAILURE ("Nonfinite objective/gradient"). */
int wust_oracle_eval(wust_helmholtz *h, const int64_t *tx_index, int64_t n_tx, const int64_t *rx_index,
                     int64_t n_rx, const void *d_Y, const uint8_t *d_mask, int64_t batch,
                     int need_gradient, int retain, double tol, int max_iter, double *loss,
                     void *d_grad, void *stream);

/* Fitted source scales a* (host complex128, n_tx entries) of the last
 * wust_oracle_eval call that retained state. */
int wust_oracle_scales(const wust_helmholtz *h, void *scales_host);

/* kappa = sum_b || -(H^{-1} M (v_b .* direction))[rx] a_b .* mask ||^2 with
 * the retained v and frozen a (Runtime/matlab/directional_curvature.m:1-26).
 * d_direction: device double [Ny, Nx]; d_mask: device uint8 [n_tx, n_rx]
 * (the receiver mask of this frequency); kappa: host double. */
int wust_directional_curvature(wust_helmholtz *h, const double *d_direction, const uint8_t *d_mask,
                               double tol, int max_iter, double *kappa, void *stream);

/* Drop the retained forward fields (the reference clears its cache after the
 * last curvature consumer, Runtime/matlab/wust_reconstruct.m:61). */
int wust_oracle_release(wust_helmholtz *h);

/* ---- NCG step around the oracle (FP64, device arrays, [Ny, Nx] grids) ------------------ */
/* ROI gradient and its filtered copy: roi = mask ? g : 0, z = mask ? F(roi) : 0,
 * with F the radial Butterworth low-pass of cutoff * 2 f / c0 cycles/m and the
 * given order (F = identity when cutoff == 0).  Grid spacing dx (along x,
 * columns) and dy (along y, rows).  Replaces wust_roi_gradient
 * (Runtime/matlab/wust_roi_gradient.m:1-10) and ringingRemovalFilt
 * (Runtime/solver/ringingRemovalFilt.m:1-45). */
int wust_roi_gradient(const double *d_g, const uint8_t *d_mask, int32_t ny, int32_t nx, double dx, double dy,
                      double c0, double frequency_hz, double cutoff, double order, double *d_roi, double *d_z,
                      void *stream);

/* Clipped PR/FR direction with the objective's slope: gz = <g, z>; beta =
 * clip(PR, 0, FR) on the same frequency (prev_gz > 0), else 0; direction =
 * -z + beta * prev_direction, restarted as -z when <g, direction> >= 0.
 * Outputs (host) beta, gz and slope = <g, direction>; WUST_NUMERICAL_FAILURE
 * for a non-finite direction or gz < 0.  prev_z / prev_direction may be NULL
 * when same_frequency == 0.  Replaces wust_ncg_direction
 * (Runtime/matlab/wust_ncg_direction.m:1-17). */
int wust_ncg_direction(const double *d_g, const double *d_z, const double *d_prev_z, double prev_gz,
                       const double *d_prev_direction, int same_frequency, int64_t count, double *d_direction,
                       double *beta, double *gz, double *slope, void *stream);

/* alpha = max(0, -slope / curvature) (0 when curvature == 0 and slope == 0);
 * WUST_NUMERICAL_FAILURE for non-finite input, negative curvature, or zero
 * curvature with a nonzero slope.  Host only.  Replaces
 * linearized_step_length (Runtime/matlab/linearized_step_length.m:1-11). */
int wust_linearized_step_length(double slope, double curvature, double *alpha);

/* Bounded slowness update: s = 1/c + step_damping * alpha * direction,
 * clipped to [1/min(c + max_update, hi), 1/max(c - max_update, lo)], c_new =
 * 1/s on the update mask and c elsewhere.  relative_update (host, may be
 * NULL) = ||1/c_new - 1/c|| / ||1/c|| over the mask.  Replaces
 * wust_project_slowness_update (Runtime/matlab/wust_project_slowness_update.m:1-10)
 * and the record field of Runtime/matlab/wust_reconstruct.m:65-66. */
int wust_project_slowness_update(const double *d_c, const double *d_direction, double alpha, double step_damping,
                                 double max_update_mps, double bound_lo_mps, double bound_hi_mps,
                                 const uint8_t *d_update_mask, int64_t count, double *d_c_new,
                                 double *relative_update, void *stream);

/* ---- RF -> frequency DTFT (data preparation) ------------------------------------------- */
/* pressure(tx, r, f) = single( sum_t exp(-2 pi i f_hz[f] time_s[t]) dt rf(t, r, tx) ) in FP64,
 * with every sample of a channel whose active[tx + n_tx r] == 0 treated as zero.  rf: device
 * [n_time, n_rx, n_tx] column-major, float32 (rf_dtype 0) or float64 (rf_dtype 1); time_s: device
 * [n_time]; frequencies_hz: host [n_freq], each in (0, 0.5/dt); active: device uint8 [n_tx, n_rx];
 * pressure: device complex64 [n_tx, n_rx, n_freq].  kernel_ms (host, may be NULL) receives the GEMM
 * launches' duration.  WUST_INVALID_INPUT for a non-finite sample on an active channel.  Replaces the
 * DTFT loop of wust_prepare (Runtime/matlab/wust_prepare.m:29-37). */
int wust_rf_dtft(const void *d_rf, int rf_dtype, int64_t n_time, int32_t n_rx, int32_t n_tx, const double *d_time_s,
                 double dt, const double *frequencies_hz, int32_t n_freq, const uint8_t *d_active, void *d_pressure,
                 float *kernel_ms, void *stream);

/* wust_rf_dtft without any device allocation: no reference counterpart.  For callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as wust_rf_dtft except kernel_ms (not offered); its work memory (the kernel
 * table of one pass over the RF and a status word, 8 MiB) is a static array that each GPU holds until
 * the process ends, so n_time is limited to WUST_DTFT_NOALLOC_MAX_TIME samples (WUST_INVALID_INPUT
 * beyond it).  Every call on a GPU shares that GPU's array: calls that overlap on one GPU (other
 * threads, other streams) race.  The call waits for its stream to read the status word. */
#define WUST_DTFT_NOALLOC_MAX_TIME 16384
int wust_rf_dtft_noalloc(const void *d_rf, int rf_dtype, int64_t n_time, int32_t n_rx, int32_t n_tx,
                         const double *d_time_s, double dt, const double *frequencies_hz, int32_t n_freq,
                         const uint8_t *d_active, void *d_pressure, void *stream);

/* ---- tooling: per-kernel-class HIP-event timing (no reference counterpart) ------------ */
/* Classes: 0 assembly, 1 Schur update, 2 Gauss-Jordan panel, 3 sweep operand,
 * 4 sweep GEMM, 5 FP64 residual, 6 source fit, 7 gradient, 8 plain factor step,
 * 9 plain solve step, 10 Newton-Schulz block-inverse GEMM.  When a class bit is
 * set, every launch of that class is timestamped by its dispatch packet
 * (hipExtLaunchKernelGGL events); wust_profile_read resolves them. */
int wust_profile_enable(uint32_t class_mask);
int wust_profile_read(int kernel_class, int64_t *count, double *total_ms);
int wust_profile_reset(void);
/* Work units counted with the profiled launches of a class since the last reset: for the sweep
 * GEMM class, block chain steps (one n x n by n x R complex product each), so that a caller can
 * price the class's time against the work actually executed. */
int wust_profile_units(int kernel_class, int64_t *units);
/* Time one launch in `stride` of each profiled class (default 1: every launch); resets the
 * per-class launch counters.  An event-timed launch's completion signal idles the device for
 * several microseconds before the next launch, so a run that is itself being timed samples. */
int wust_profile_stride(uint32_t stride);

#ifdef __cplusplus
}
#endif
#endif /* WUST_HIP_H */
