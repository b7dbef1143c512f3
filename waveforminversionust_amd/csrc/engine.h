// Device state of one Helmholtz operator and the host drivers that act on it.
#pragma once
#include <vector>

#include "assembly.h"
#include "wust_common.h"

namespace wust {

// Device layouts (all column-major within a block):
//   w64   [9][m][n]        FP64 per-node stencil rows of the interior nodes
//   w32   [9][m][npad]     complex-FP32 block coefficients (ring couplings and padding zeroed)
//   pages [m][npad][npad]  inverse Schur blocks X_j (twisted: left chain j < p, right chain
//                          j > p, middle j = p); padding rows/cols hold the identity
//   solve fields  [m][npad][Rpad]   interior block x, row y, RHS r fastest
// How Schur blocks are inverted (FACTOR_AUTO: Newton-Schulz from the
// neighbouring inverse where a chain has one, Gauss-Jordan elsewhere and as
// the fallback; FACTOR_GJ: Gauss-Jordan for every block).
enum FactorMethod { FACTOR_AUTO = 0, FACTOR_GJ = 1 };

struct Operator {
    int ny = 0, nx = 0, n = 0, m = 0, npad = 0, p = 0;  // p: twisted meeting block
    int mode = 0;
    int factor_method = FACTOR_AUTO;
    bool fixed_budget = false;    // Newton-Schulz with the default budgets: no learned per-key profile
    double b = 0, d = 0, e = 0;
    AssemblyParams ap{};
    cd *ck2 = nullptr, *vsm = nullptr, *w64 = nullptr;
    cf *w32 = nullptr;
    cf *pages = nullptr;
    cf *work = nullptr;           // factor buffers: [2 chains][3][npad*npad] (Schur block T, two more)
    cf *pbuf = nullptr;           // diagonal-block inverses handed between panel launches: [2 chains][2][32*32]
    cf *nswork = nullptr;         // Newton-Schulz iterates and cubic operand: [2 chains][3][npad*npad]
    float *ns_norm = nullptr;     // ||E_r||_F^2 per chain and round: [2 chains][8]
    int *ns_rounds = nullptr;     // Newton-Schulz rounds used per block, + 64 if the last was forced: [m]
    int *colflags = nullptr;      // per block column: 1 when a right-hand side has a nonzero there: [m]
    float *gj_check = nullptr;    // ||I - T_j X_j||_F^2 of each Gauss-Jordan block (0 elsewhere): [m]
    uint32_t *status = nullptr;   // device status words: [0] flags, [2..3] pivot ratios (float),
                                  // [4] largest Newton-Schulz residual bound (float); wust_oracle_eval:
                                  // [6] a gradient value is not finite, [7] source-fit failure flags,
                                  // [8..9] the loss summed over batches (double)
    double assembly_ms = 0, factor_ms = 0, min_pivot_ratio = 1.0;
    double inverse_residual = 0;  // max over Newton-Schulz blocks of the bound on ||I - T_j X_j||_F
    int ns_blocks = 0;            // blocks inverted by Newton-Schulz
    int factor_fallback = 0;      // 1: Newton-Schulz missed its bound and every block was redone by Gauss-Jordan;
                                  // 2: a Gauss-Jordan block failed its check and every block was redone pivoted
    double gj_residual = 0;       // largest ||I - T_j X_j||_F over the Gauss-Jordan blocks
    int64_t factor_bytes = 0;

    const cf *coef(int dir, int block) const { return w32 + ((int64_t)dir * m + block) * npad; }
    cf *page(int j) const { return pages + (int64_t)j * npad * npad; }
};

struct PlainLib;

// ---- factor (twisted block-tridiagonal Schur recurrence + Gauss-Jordan inverses) ----
// lib: rocSOLVER handle for the pivoted refactorization when a pivot-free block fails its check
void factor_engine(Operator &op, PlainLib &lib, hipStream_t s);
// frees the cached leading inverse pages that seed later factorizations with the same key (entries
// held by a factorization in progress stay)
void seed_cache_trim();
// single-chain Schur block T = D_j - L_j X_{j-1} U_{j-1} (T = D_0 for j = 0), X = op.page
void launch_schur_left(const Operator &op, int j, cf *T, hipStream_t s);

// ---- complex-FP32 block solve: out = H^{-1} src (or H^{-H} src), solve layout ----
// src, out: [m][npad][Rpad] complex-FP32; out must not alias src; scratch: 2*npad*Rpad.
// Rpad must be a multiple of 16.
// work: [m][npad][Rpad] forward iterates + 4 operand blocks [npad][Rpad] (solve_work_elems)
// lead / trail: block columns 0 .. lead-1 and m-trail .. m-1 of src are zero, so the forward
// elimination of each chain starts at its first nonzero column (its iterates before it are
// exactly zero); each chain still computes at least its last step.
void solve_engine(const Operator &op, const cf *src, cf *out, cf *work, int Rpad, bool adjoint, hipStream_t s,
                  int lead = 0, int trail = 0);
inline int64_t solve_work_elems(const Operator &op, int Rpad) { return ((int64_t)op.m + 4) * op.npad * Rpad; }

}  // namespace wust
