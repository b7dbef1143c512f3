#pragma once
#include "wust_common.h"

namespace wust {

struct AssemblyParams {
    int ny, nx, npad;
    double x0, dx, y0, dy;
    double f, sgn, a0, L;
    double b, d, e;
    int dispersion;
    double ldr9_dt, ldr9_cr;
};

// ck2, vsm: [Ny*Nx] node arrays; w64: [9][m][n]; w32: [9][m][npad].
void launch_assembly(const AssemblyParams &p, const double *vel, const double *atten, cd *ck2, cd *vsm, cd *w64,
                     cf *w32, uint32_t *status, hipStream_t s);

void launch_export(int ny, int nx, const cd *w64, cd *Dd, cd *Dl, cd *Du, cd *Ld, cd *Ll, cd *Lu, cd *Ud, cd *Ul,
                   cd *Uu, hipStream_t s);

}  // namespace wust
