// NCG step kernels (FP64, device arrays, MATLAB [Ny, Nx] column-major grids).
#pragma once
#include "wust_common.h"

namespace wust {

// roi = mask ? g : 0;  z = mask ? filter(roi) : 0  (filter skipped when cutoff == 0)
void ncg_roi_gradient(const double *g, const uint8_t *mask, int ny, int nx, double dx, double dy, double c0,
                      double f, double cutoff, double order, double *roi, double *z, hipStream_t s);
// clipped PR/FR direction; returns beta, gz = <g, z> and slope = <g, direction>
void ncg_direction(const double *g, const double *z, const double *prev_z, double prev_gz, const double *prev_dir,
                   bool same_frequency, int64_t n, double *direction, double *beta, double *gz, double *slope,
                   hipStream_t s);
// bounded slowness update; rel_update (host, may be null) = |s_new - s_old| / |s_old| over the mask
void ncg_project(const double *c, const double *direction, double alpha, double step_damping, double max_update,
                 double lo, double hi, const uint8_t *mask, int64_t n, double *c_new, double *rel_update,
                 hipStream_t s);

}  // namespace wust
