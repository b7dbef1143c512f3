#pragma once
#include "engine.h"

namespace wust {

// rocBLAS handle plus pivot/info scratch of the plain (vendor-library) path.
struct PlainLib {
    void *handle = nullptr;
    int *ipiv = nullptr;
    int *info = nullptr;
    void ensure(hipStream_t s, int n);
    ~PlainLib();
};

void factor_plain(Operator &op, PlainLib &lib, hipStream_t s);
// src, u: [R][m][npad] complex-FP32; tmp: [R][npad]
void solve_plain(const Operator &op, PlainLib &lib, const cf *src, cf *u, cf *tmp, int R, bool adjoint,
                 hipStream_t s);
void ext_to_plain(const cd *ext, cf *out, int ny, int nx, int npad, int R, hipStream_t s);
void plain_to_ext(const cf *in, cd *ext, int ny, int nx, int npad, int R, hipStream_t s);

}  // namespace wust

namespace wust {
void solve64_to_plain32(const cd *in, cf *out, int m, int npad, int R, int Rpad, hipStream_t s);
void plain32_to_solve64(const cf *in, cd *out, int m, int npad, int R, int Rpad, hipStream_t s);
}  // namespace wust
