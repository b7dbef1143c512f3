// Complex-FP32 block sweeps over grid columns, batched over all right-hand sides.
//
// The reference (Runtime/solver/applyBlockLU.cu:433-467) runs, per column, one
// tridiagonal kernel and one cuBLAS Cgemm3m (Ny x R x Ny), serially over x
// for the forward and backward passes.  Here each step is an operand pass
// (the tridiagonal couplings, evaluated once per element) followed by a
// complex GEMM on MFMA (v_mfma_f32_16x16x4f32, four real products per complex
// product, FP32 accumulation), and the twisted factorization lets the left and
// right chains of a step run in the same launches:
//   forward  left  y_j = X_j (b_j - L_j y_{j-1}),       j = 0..p-1
//            right z_j = X_j (b_j - U_j z_{j+1}),       j = m-1..p+1
//            mid   u_p = X_p (b_p - L_p y_{p-1} - U_p z_{p+1})
//            back  u_j = y_j - X_j U_j u_{j+1} (left),  u_j = z_j - X_j L_j u_{j-1} (right)
//   adjoint  the same with X_j^H and the conjugate-transposed couplings
//            (H^H has U_{j-1}^H below and L_{j+1}^H above the diagonal).
// Layout of every field block: [npad rows][Rpad RHS], RHS fastest, so the
// coupling loads and the output stores are coalesced across RHS.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <tuple>
#include <mutex>
#include <map>

#include "cgemm.h"
#include "engine.h"
#include "profile.h"

#ifndef WUST_SWEEP_FOLD
#define WUST_SWEEP_FOLD 1
#endif

namespace wust {

namespace {

// The sweep products accumulate each k-block on MFMA and fold it into FP64 sums (CgemmWave FOLD): a
// long FP32 accumulation over the block rows sets the refinement contraction of the whole solve.
constexpr bool SWEEP_FOLD = WUST_SWEEP_FOLD != 0;

struct SweepTerm {
    const cf *cn, *cc, *cs;   // tridiagonal row coefficients (block of the coupling)
    const cf *v;              // field block the coupling acts on
    int transposed;           // apply P^H instead of P
};
struct SweepTask {
    const cf *X;              // inverse page
    int xconj;                // use X^H (selects the kernel instantiation)
    const cf *src;            // source block or null
    SweepTerm t[2];
    int nterms;
    cf *B;                    // operand scratch block [npad][Rpad]
    cf *out;
    int accumulate;           // out += X B instead of out = X B
};
struct SweepParams {
    SweepTask task[2];
    int ntasks, npad, Rpad;
    int units, per_xcd;       // logical tiles and tiles per XCD slot
};

// B[k][r] = src[k][r] - sum_t (P_t v_t)[k][r]
__device__ __forceinline__ cf operand_b(const SweepTask &tk, int k, int r, int npad, int Rpad) {
    cf b = tk.src ? tk.src[(int64_t)k * Rpad + r] : mk(0.f, 0.f);
    for (int t = 0; t < tk.nterms; ++t) {
        const SweepTerm &st = tk.t[t];
        const cf *v = st.v + r;
        cf acc;
        if (!st.transposed) {
            acc = st.cc[k] * v[(int64_t)k * Rpad];
            if (k >= 1) cfma(acc, st.cn[k], v[(int64_t)(k - 1) * Rpad]);
            if (k + 1 < npad) cfma(acc, st.cs[k], v[(int64_t)(k + 1) * Rpad]);
        } else {
            acc = conjf_(st.cc[k]) * v[(int64_t)k * Rpad];
            if (k + 1 < npad) cfma(acc, conjf_(st.cn[k + 1]), v[(int64_t)(k + 1) * Rpad]);
            if (k >= 1) cfma(acc, conjf_(st.cs[k - 1]), v[(int64_t)(k - 1) * Rpad]);
        }
        b = b - acc;
    }
    return b;
}

// Operand pass: the tridiagonal couplings of the step, once per (k, r).
__global__ void __launch_bounds__(256) sweep_operand_kernel(SweepParams p) {
    const int64_t per = (int64_t)p.npad * p.Rpad;
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= per * p.ntasks) return;
    const int task = (int)(idx / per);
    const int64_t rem = idx - task * per;
    const int k = (int)(rem / p.Rpad), r = (int)(rem - (int64_t)k * p.Rpad);
    const SweepTask &tk = p.task[task];
    tk.B[rem] = operand_b(tk, k, r, p.npad, p.Rpad);
}

// GEMM pass: out (+)= op(X) B, op(X) = X or X^H, on tiles of SM x 16 (SM = 16 RB
// block rows x 16 RHS).  The 4 waves of a WG deal the k-blocks of 16
// round-robin and each accumulates the whole tile (CgemmWave: four real
// products per complex product, two k-blocks of operands in flight, straight
// from global memory in the MFMA lane layout); the partial tiles are summed
// through LDS.  Tiles
// are dealt to XCDs in contiguous runs (blockIdx % 8 is the XCD) with the RHS
// block fastest, so one XCD reads a band of rows of X (1/4 of a page at two
// chains on 8 XCDs) and the whole operand block, instead of whole pages.
template <int RB, int ALAY>
__global__ void __launch_bounds__(256) sweep_gemm_kernel(SweepParams p) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int L = xcd * p.per_xcd + local;
    if (L >= p.units) return;
    const int npad = p.npad, Rpad = p.Rpad;
    const int Mb = npad / (16 * RB), Nb = Rpad / 16;
    const int nb = L % Nb;
    const int mb = (L / Nb) % Mb;
    const int task = L / (Nb * Mb);
    const SweepTask &tk = p.task[task];
    const int i0 = mb * 16 * RB, j0 = nb * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;

    CgemmWave<RB, 1, ALAY, B_ROW, false, 2, SWEEP_FOLD> acc;
    acc.zero();
    const cf *A = (ALAY == A_COL) ? tk.X + i0 : tk.X + (int64_t)i0 * npad;
    acc.run(A, npad, tk.B + j0, Rpad, wave, npad / 16, 4, g, c);

    // waves 1..3 park their partial tiles; wave 0 sums and stores
    __shared__ float4 red[3][RB][2][64];
    if (wave > 0) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            float re[4], im[4];
            acc.result(rb, 0, re, im);
            red[wave - 1][rb][0][lane] = make_float4(re[0], re[1], re[2], re[3]);
            red[wave - 1][rb][1][lane] = make_float4(im[0], im[1], im[2], im[3]);
        }
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            float re[4], im[4];
            acc.result(rb, 0, re, im);
#pragma unroll
            for (int w = 0; w < 3; ++w) {
                const float4 u = red[w][rb][0][lane], v = red[w][rb][1][lane];
                re[0] += u.x; re[1] += u.y; re[2] += u.z; re[3] += u.w;
                im[0] += v.x; im[1] += v.y; im[2] += v.z; im[3] += v.w;
            }
            // lane holds rows i0 + 16 rb + 4 g + r of RHS column j0 + c
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                cf *o = tk.out + (int64_t)(i0 + 16 * rb + 4 * g + r) * Rpad + j0 + c;
                cf v = mk(re[r], im[r]);
                if (tk.accumulate) v = v + *o;
                *o = v;
            }
        }
    }
}

constexpr int SWEEP_RB = 1;   // 16-row tiles: 2 * 36 * 16 = 1152 WGs at 560^2 x 256 RHS

// ---- fused sweep step: GEMM + the next step's operand in the epilogue ---------------------
//
// The next step's operand B' = src' - P v needs rows k-1, k, k+1 of this step's result v.
// Each workgroup therefore computes a 16-row MFMA tile that overlaps its neighbours by one
// row on each side and owns its 14 middle rows: strip mb owns rows 14 mb .. 14 mb + 13 and
// computes rows i0 .. i0 + 15 with i0 = clamp(14 mb - 1, 0, npad - 16), so every owned row
// has both neighbours in the tile (or outside the block, where they are zero).  The owned
// rows of v are stored and B' is formed from the tile in LDS; the separate operand pass and
// its launch disappear, at the cost of 16/14 of the MFMA work (32-row tiles that own 30 rows
// cost 32/30, with half the workgroups).  Back substitution reads its base (y_j) from a
// different array than it writes (u_j), so halo rows are never read after a neighbour
// rewrote them.
constexpr int FUSED_OWN = 14;   // owned rows of a 16-row tile

struct FusedTask {
    const cf *X;              // inverse page
    const cf *B;              // operand block [npad][Rpad]
    const cf *base;           // null, or the block added to op(X) B
    cf *out;                  // result block
    cf *Bn;                   // null, or the next step's operand block
    const cf *nsrc;           // next step's source block (or null)
    const cf *cn, *cc, *cs;   // coupling acting on this step's result
    int transposed;
};
struct FusedParams {
    FusedTask task[2];
    int ntasks, n, npad, Rpad;
    int strips, units, per_xcd;
    cf *skpart;               // split-K variant: partial tiles [tile][SK_PARTS][32][32]
    int *skcount;             // split-K variant: arrivals per tile (a multiple of SK_PARTS between launches)
};

// KS waves share one 16 x 16 tile (k-blocks dealt round-robin, partial tiles summed in LDS),
// so a workgroup of four waves covers TPW = 4 / KS adjacent RHS tiles of one strip; DEPTH
// k-blocks of operands are in flight per wave.  The variant is picked per shape at run time.
template <int ALAY, int DEPTH, int KS>
__device__ __forceinline__ void sweep_fused_body(const FusedParams &p) {
    constexpr int TPW = 4 / KS;
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int L = xcd * p.per_xcd + local;
    if (L >= p.units) return;
    const int npad = p.npad, Rpad = p.Rpad;
    const int Ng = Rpad / 16 / TPW;
    const int ng = L % Ng;
    const int mb = (L / Ng) % p.strips;
    const int task = L / (Ng * p.strips);
    const FusedTask &tk = p.task[task];
    const int o0 = FUSED_OWN * mb;
    const int o1 = min(o0 + FUSED_OWN - 1, npad - 1);
    const int i0 = min(max(o0 - 1, 0), npad - 16);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
    const int tw = wave / KS, ks = wave % KS;
    const int j0 = (ng * TPW + tw) * 16;

    // epilogue operands (base rows of this lane's tile, and for this thread's output element the
    // next source and the three coupling coefficients) are loaded before the GEMM so that their
    // latency overlaps it instead of trailing it
    const int t = (threadIdx.x >> 4) & 15, cc = threadIdx.x & 15;
    const int k = i0 + t;
    const bool own = k >= o0 && k <= o1;
    cf pbase[4];
    if (tk.base && ks == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) pbase[r] = tk.base[(int64_t)(i0 + 4 * g + r) * Rpad + j0 + c];
    }
    cf pn = mk(0.f, 0.f), pc = mk(0.f, 0.f), ps = mk(0.f, 0.f), psrc[TPW];
    if (own && tk.Bn) {
        if (!tk.transposed) {
            pc = tk.cc[k];
            if (k >= 1) pn = tk.cn[k];
            if (k + 1 < npad) ps = tk.cs[k];
        } else {
            pc = conjf_(tk.cc[k]);
            if (k >= 1) pn = conjf_(tk.cs[k - 1]);
            if (k + 1 < npad) ps = conjf_(tk.cn[k + 1]);
        }
#pragma unroll
        for (int q = 0; q < TPW; ++q)
            psrc[q] = tk.nsrc ? tk.nsrc[(int64_t)k * Rpad + (ng * TPW + q) * 16 + cc] : mk(0.f, 0.f);
    }

    CgemmWave<1, 1, ALAY, B_ROW, false, DEPTH, SWEEP_FOLD> acc;
    acc.zero();
    const cf *A = (ALAY == A_COL) ? tk.X + i0 : tk.X + (int64_t)i0 * npad;
    acc.run(A, npad, tk.B + j0, Rpad, ks, npad / 16, KS, g, c);

    __shared__ float4 red[4][2][64];
    __shared__ cf tile[TPW][16][17];
    float re[4], im[4];
    acc.result(0, 0, re, im);
    if (KS > 1) {
        if (ks > 0) {
            red[wave][0][lane] = make_float4(re[0], re[1], re[2], re[3]);
            red[wave][1][lane] = make_float4(im[0], im[1], im[2], im[3]);
        }
        __syncthreads();
    }
    if (ks == 0) {
#pragma unroll
        for (int w = 1; w < KS; ++w) {
            const float4 u = red[wave + w][0][lane], v = red[wave + w][1][lane];
            re[0] += u.x; re[1] += u.y; re[2] += u.z; re[3] += u.w;
            im[0] += v.x; im[1] += v.y; im[2] += v.z; im[3] += v.w;
        }
        // lane holds tile rows 4 g + r of column j0 + c
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cf v = mk(re[r], im[r]);
            if (tk.base) v = v + pbase[r];
            tile[tw][4 * g + r][c] = v;
        }
    }
    __syncthreads();
    // epilogue, one element per thread and tile: tile row t, column j0' + cc of tile q
    if (!own) return;
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
        const int64_t off = (int64_t)k * Rpad + (ng * TPW + q) * 16 + cc;
        const cf v = tile[q][t][cc];
        tk.out[off] = v;
        if (!tk.Bn) continue;
        const cf vn = (k >= 1) ? tile[q][t - 1][cc] : mk(0.f, 0.f);        // row k - 1 (in the tile when k >= 1)
        const cf vs = (k + 1 < npad) ? tile[q][t + 1][cc] : mk(0.f, 0.f);  // row k + 1 (in the tile when < npad)
        // the same products and order as operand_b: P v = c_c v_k + c_n v_{k-1} + c_s v_{k+1}
        // (transposed: conj(cc_k) v_k + conj(cn_{k+1}) v_{k+1} + conj(cs_{k-1}) v_{k-1})
        cf a = pc * v;
        if (!tk.transposed) {
            if (k >= 1) cfma(a, pn, vn);
            if (k + 1 < npad) cfma(a, ps, vs);
        } else {
            if (k + 1 < npad) cfma(a, ps, vs);
            if (k >= 1) cfma(a, pn, vn);
        }
        tk.Bn[off] = psrc[q] - a;
    }
}

// The same step on TR-row tiles (TR a multiple of 16; TR - 2 rows owned, a wave accumulating
// TR / 16 row blocks so that each operand block it loads serves all of them).  Kept apart from
// the 16-row body so that each compiles to its own register allocation.
template <int ALAY, int DEPTH, int KS, int TR>
__device__ __forceinline__ void sweep_fused_body_tall(const FusedParams &p) {
    constexpr int TPW = 4 / KS, RB = TR / 16, OWN = TR - 2;
    constexpr int EPT = RB * TPW;   // epilogue elements per thread
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int L = xcd * p.per_xcd + local;
    if (L >= p.units) return;
    const int npad = p.npad, Rpad = p.Rpad;
    const int Ng = Rpad / 16 / TPW;
    const int ng = L % Ng;
    const int mb = (L / Ng) % p.strips;
    const int task = L / (Ng * p.strips);
    const FusedTask &tk = p.task[task];
    const int o0 = OWN * mb;
    const int o1 = min(o0 + OWN - 1, npad - 1);
    const int i0 = min(max(o0 - 1, 0), npad - TR);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
    const int tw = wave / KS, ks = wave % KS;
    const int j0 = (ng * TPW + tw) * 16;

    // epilogue operands (base rows of this lane's tile, and for this thread's output elements the
    // next source and the three coupling coefficients) are loaded before the GEMM so that their
    // latency overlaps it instead of trailing it.  Element q of this thread is tile row
    // (threadIdx.x / 16) + 16 (q % RB), column threadIdx.x % 16 of RHS tile q / RB.
    cf pbase[RB][4];
    if (tk.base && ks == 0) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) pbase[rb][r] = tk.base[(int64_t)(i0 + 16 * rb + 4 * g + r) * Rpad + j0 + c];
    }
    const int cc = threadIdx.x & 15;
    bool own[RB];
    cf pn[RB], pc[RB], ps[RB], psrc[EPT];
#pragma unroll
    for (int rr = 0; rr < RB; ++rr) {
        const int k = i0 + (threadIdx.x >> 4) + 16 * rr;
        own[rr] = k >= o0 && k <= o1;
        pn[rr] = pc[rr] = ps[rr] = mk(0.f, 0.f);
        if (own[rr] && tk.Bn) {
            if (!tk.transposed) {
                pc[rr] = tk.cc[k];
                if (k >= 1) pn[rr] = tk.cn[k];
                if (k + 1 < npad) ps[rr] = tk.cs[k];
            } else {
                pc[rr] = conjf_(tk.cc[k]);
                if (k >= 1) pn[rr] = conjf_(tk.cs[k - 1]);
                if (k + 1 < npad) ps[rr] = conjf_(tk.cn[k + 1]);
            }
#pragma unroll
            for (int tq = 0; tq < TPW; ++tq)
                psrc[rr + RB * tq] = tk.nsrc ? tk.nsrc[(int64_t)k * Rpad + (ng * TPW + tq) * 16 + cc] : mk(0.f, 0.f);
        }
    }

    CgemmWave<RB, 1, ALAY, B_ROW, false, DEPTH, SWEEP_FOLD> acc;
    acc.zero();
    const cf *A = (ALAY == A_COL) ? tk.X + i0 : tk.X + (int64_t)i0 * npad;
    acc.run(A, npad, tk.B + j0, Rpad, ks, npad / 16, KS, g, c);

    __shared__ float4 red[4][RB][2][64];
    __shared__ cf tile[TPW][TR][17];
    float re[RB][4], im[RB][4];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc.result(rb, 0, re[rb], im[rb]);
    if (KS > 1) {
        if (ks > 0) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                red[wave][rb][0][lane] = make_float4(re[rb][0], re[rb][1], re[rb][2], re[rb][3]);
                red[wave][rb][1][lane] = make_float4(im[rb][0], im[rb][1], im[rb][2], im[rb][3]);
            }
        }
        __syncthreads();
    }
    if (ks == 0) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
            for (int w = 1; w < KS; ++w) {
                const float4 u = red[wave + w][rb][0][lane], v = red[wave + w][rb][1][lane];
                re[rb][0] += u.x; re[rb][1] += u.y; re[rb][2] += u.z; re[rb][3] += u.w;
                im[rb][0] += v.x; im[rb][1] += v.y; im[rb][2] += v.z; im[rb][3] += v.w;
            }
            // lane holds tile rows 16 rb + 4 g + r of column j0 + c
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                cf v = mk(re[rb][r], im[rb][r]);
                if (tk.base) v = v + pbase[rb][r];
                tile[tw][16 * rb + 4 * g + r][c] = v;
            }
        }
    }
    __syncthreads();
    // epilogue: this thread's elements
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int rr = q % RB, tq = q / RB;
        if (!own[rr]) continue;
        const int t = (threadIdx.x >> 4) + 16 * rr, k = i0 + t;
        const int64_t off = (int64_t)k * Rpad + (ng * TPW + tq) * 16 + cc;
        const cf v = tile[tq][t][cc];
        tk.out[off] = v;
        if (!tk.Bn) continue;
        const cf vn = (k >= 1) ? tile[tq][t - 1][cc] : mk(0.f, 0.f);        // row k - 1 (in the tile when k >= 1)
        const cf vs = (k + 1 < npad) ? tile[tq][t + 1][cc] : mk(0.f, 0.f);  // row k + 1 (in the tile when < npad)
        // the same products and order as operand_b: P v = c_c v_k + c_n v_{k-1} + c_s v_{k+1}
        // (transposed: conj(cc_k) v_k + conj(cn_{k+1}) v_{k+1} + conj(cs_{k-1}) v_{k-1})
        cf a = pc[rr] * v;
        if (!tk.transposed) {
            if (k >= 1) cfma(a, pn[rr], vn);
            if (k + 1 < npad) cfma(a, ps[rr], vs);
        } else {
            if (k + 1 < npad) cfma(a, ps[rr], vs);
            if (k >= 1) cfma(a, pn[rr], vn);
        }
        tk.Bn[off] = psrc[q] - a;
    }
}

// ---- split-K step on 32 x 32 tiles with operands shared through LDS ------------------------------
//
// The 16-row steps above are bound by operand traffic from L2: each wave streams its own A (16 rows of
// the page) and B (16 RHS of the operand) over its K share, so every page row is read once per RHS
// tile and every operand row once per strip (~170 MB per step at 576 x 256 against ~8 MB of distinct
// data).  Here a workgroup's four waves compute the four 16 x 16 blocks of one 32-row (30 owned) x
// 32-RHS tile over one quarter of K, with each k-block of A (32 rows) and B (32 RHS) staged once in
// LDS and read by the two waves that share it: half the traffic per output element.  Splitting K
// over SK_PARTS workgroups keeps the launch at ~4.75 waves per SIMD (1216 workgroups at 558 rows and
// 256 RHS).  Each workgroup writes its FP32-rounded partial tile (the k-block fold's FP64 sums, as
// the other variants round them before their LDS reduction); the last of a tile's workgroups to
// arrive (an atomic count per tile) sums the SK_PARTS partials in part order in FP64, so the result
// does not depend on the arrival order, and runs the fused epilogue.  The workgroups of one tile
// are dealt to the same XCD (blockIdx % 8), whose L2 is their point of coherence: each waits for
// its partial's stores to complete before counting itself in, and the last one invalidates its L1
// before reading the others' partials.  No workgroup ever waits for another.
constexpr int SK_PARTS = 4;
constexpr int SK_TR = 32, SK_OWN = 30, SK_TC = 32;

__device__ __forceinline__ void wait_vmem_stores() {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#endif
}
__device__ __forceinline__ void invalidate_l1() {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("buffer_inv sc0\n\ts_waitcnt vmcnt(0)" ::: "memory");
#endif
}

template <int ALAY>
__device__ __forceinline__ void sk_load(const cf *X, const cf *B, int npad, int Rpad, int i0, int j0, int kb, int t,
                                        cf (&a)[2], cf (&b)[2]) {
    const int k0 = 16 * kb;
    if (ALAY == A_COL) {   // A[i][k] = X[k npad + i]: thread t takes rows 2 (t % 16) + {0, 1} of k = t / 16
        const cf *src = X + (int64_t)(k0 + (t >> 4)) * npad + i0 + 2 * (t & 15);
        a[0] = src[0];
        a[1] = src[1];
    } else {               // A[i][k] = conj(X[i npad + k]): thread t takes k = 2 (t % 8) + {0, 1} of row t / 8
        const float4 u = *reinterpret_cast<const float4 *>(X + (int64_t)(i0 + (t >> 3)) * npad + k0 + 2 * (t & 7));
        a[0] = mk(u.x, -u.y);
        a[1] = mk(u.z, -u.w);
    }
    const float4 v = *reinterpret_cast<const float4 *>(B + (int64_t)(k0 + (t >> 4)) * Rpad + j0 + 2 * (t & 15));
    b[0] = mk(v.x, v.y);
    b[1] = mk(v.z, v.w);
}

// LDS staging: As[k][row], Bs[k][col] of one k-block (k-major, so a lane's MFMA operand reads are
// contiguous over its 16 rows / columns)
template <int ALAY>
__device__ __forceinline__ void sk_stage(cf (*As)[SK_TR], cf (*Bs)[SK_TC], int t, const cf (&a)[2], const cf (&b)[2]) {
    if (ALAY == A_COL) {
        *reinterpret_cast<float4 *>(&As[t >> 4][2 * (t & 15)]) = make_float4(a[0].re, a[0].im, a[1].re, a[1].im);
    } else {
        As[2 * (t & 7)][t >> 3] = a[0];
        As[2 * (t & 7) + 1][t >> 3] = a[1];
    }
    *reinterpret_cast<float4 *>(&Bs[t >> 4][2 * (t & 15)]) = make_float4(b[0].re, b[0].im, b[1].re, b[1].im);
}

template <int ALAY>
__global__ void __launch_bounds__(256) sweep_splitk_kernel(FusedParams p) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int tile = xcd * p.per_xcd + (local / SK_PARTS);
    const int part = local % SK_PARTS;
    if (tile >= p.units) return;
    const int npad = p.npad, Rpad = p.Rpad;
    const int Ng = Rpad / SK_TC;
    const int ng = tile % Ng;
    const int mb = (tile / Ng) % p.strips;
    const int task = tile / (Ng * p.strips);
    const FusedTask &tk = p.task[task];
    const int o0 = SK_OWN * mb;
    const int o1 = min(o0 + SK_OWN - 1, npad - 1);
    const int i0 = min(max(o0 - 1, 0), npad - SK_TR);
    const int j0 = ng * SK_TC;
    const int nkb = npad / 16, per = (nkb + SK_PARTS - 1) / SK_PARTS;
    const int kb0 = min(part * per, nkb), kb1 = min(kb0 + per, nkb);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, g = lane >> 4, c = lane & 15;
    const int rb = wave >> 1, cb = wave & 1;   // this wave's 16 x 16 block of the 32 x 32 tile

    // one LDS block: the two k-block buffers of A and of B, later the 32 x 33 result tile
    __shared__ cf sk_lds[2 * 16 * SK_TR + 2 * 16 * SK_TC];
    __shared__ int last_flag;
    cf(*As)[16][SK_TR] = reinterpret_cast<cf(*)[16][SK_TR]>(sk_lds);
    cf(*Bs)[16][SK_TC] = reinterpret_cast<cf(*)[16][SK_TC]>(sk_lds + 2 * 16 * SK_TR);
    v4f p1 = {0.f, 0.f, 0.f, 0.f}, p3 = p1;
    double q1[4] = {0.0, 0.0, 0.0, 0.0}, q3[4] = {0.0, 0.0, 0.0, 0.0};
    cf ra[2], rbv[2];
    if (kb0 < kb1) {
        sk_load<ALAY>(tk.X, tk.B, npad, Rpad, i0, j0, kb0, t, ra, rbv);
        sk_stage<ALAY>(As[0], Bs[0], t, ra, rbv);
    }
    __syncthreads();
    for (int kb = kb0; kb < kb1; ++kb) {
        const int cur = (kb - kb0) & 1;
        const bool more = kb + 1 < kb1;
        if (more) sk_load<ALAY>(tk.X, tk.B, npad, Rpad, i0, j0, kb + 1, t, ra, rbv);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const cf x = As[cur][4 * g + s4][16 * rb + c], y = Bs[cur][4 * g + s4][16 * cb + c];
            p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x.re, y.re, p1, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(x.re, y.im, p3, 0, 0, 0);
            p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(-x.im, y.im, p1, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(x.im, y.re, p3, 0, 0, 0);
        }
        // fold the k-block into the FP64 sums
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            q1[r] += (double)p1[r];
            q3[r] += (double)p3[r];
        }
        p1 = v4f{0.f, 0.f, 0.f, 0.f};
        p3 = p1;
        if (more) sk_stage<ALAY>(As[cur ^ 1], Bs[cur ^ 1], t, ra, rbv);
        __syncthreads();
    }
    // this workgroup's partial: lane holds rows 16 rb + 4 g + r, column 16 cb + c
    cf *mine = p.skpart + ((int64_t)tile * SK_PARTS + part) * (SK_TR * SK_TC);
#pragma unroll
    for (int r = 0; r < 4; ++r)
        mine[(16 * rb + 4 * g + r) * SK_TC + 16 * cb + c] = mk((float)q1[r], (float)q3[r]);
    wait_vmem_stores();
    __syncthreads();
    if (t == 0) {
        const int old = __hip_atomic_fetch_add(p.skcount + tile, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last_flag = (old % SK_PARTS) == SK_PARTS - 1;
    }
    __syncthreads();
    if (!last_flag) return;
    invalidate_l1();
    // the tile: the partials summed in part order in FP64 (+ base), staged for the epilogue
    cf(*tile_s)[SK_TC + 1] = reinterpret_cast<cf(*)[SK_TC + 1]>(sk_lds);   // 32 x 33 cf within the LDS block
    const cf *parts = p.skpart + (int64_t)tile * SK_PARTS * (SK_TR * SK_TC);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = 16 * rb + 4 * g + r, col = 16 * cb + c;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int q = 0; q < SK_PARTS; ++q) {
            const cf v = parts[q * (SK_TR * SK_TC) + row * SK_TC + col];
            sr += (double)v.re;
            si += (double)v.im;
        }
        cf v = mk((float)sr, (float)si);
        if (tk.base) v = v + tk.base[(int64_t)(i0 + row) * Rpad + j0 + col];
        tile_s[row][col] = v;
    }
    __syncthreads();
    // epilogue: thread t handles column t % 32 of rows t / 32 + 8 q
    const int cc = t & 31;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = (t >> 5) + 8 * q, k = i0 + row;
        if (k < o0 || k > o1) continue;
        const int64_t off = (int64_t)k * Rpad + j0 + cc;
        const cf v = tile_s[row][cc];
        tk.out[off] = v;
        if (!tk.Bn) continue;
        cf pn = mk(0.f, 0.f), pc, ps = mk(0.f, 0.f);
        if (!tk.transposed) {
            pc = tk.cc[k];
            if (k >= 1) pn = tk.cn[k];
            if (k + 1 < npad) ps = tk.cs[k];
        } else {
            pc = conjf_(tk.cc[k]);
            if (k >= 1) pn = conjf_(tk.cs[k - 1]);
            if (k + 1 < npad) ps = conjf_(tk.cn[k + 1]);
        }
        const cf vn = (k >= 1) ? tile_s[row - 1][cc] : mk(0.f, 0.f);
        const cf vs = (k + 1 < npad) ? tile_s[row + 1][cc] : mk(0.f, 0.f);
        // the same products and order as operand_b
        cf a = pc * v;
        if (!tk.transposed) {
            if (k >= 1) cfma(a, pn, vn);
            if (k + 1 < npad) cfma(a, ps, vs);
        } else {
            if (k + 1 < npad) cfma(a, ps, vs);
            if (k >= 1) cfma(a, pn, vn);
        }
        const cf src = tk.nsrc ? tk.nsrc[off] : mk(0.f, 0.f);
        tk.Bn[off] = src - a;
    }
}

// Clears the split-K arrival counts (a fill kernel, so the clear is an ordinary launch).
__global__ void sk_clear_kernel(int *count, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) count[i] = 0;
}

// 1280 workgroups of four waves (40 strips x 16 RHS tiles x 2 chains at 558 rows and 256 RHS)
// need five resident waves per SIMD to run in one round; with two k-blocks in flight the step fits
// that register budget without spilling (deeper prefetch does not).  Launches that take several
// rounds anyway (larger blocks or more RHS) run faster with the unconstrained build, so both are
// autotuner variants.
template <int ALAY, int KS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
sweep_fused_kernel_d2(FusedParams p) {
    sweep_fused_body<ALAY, 2, KS>(p);
}
template <int ALAY, int DEPTH, int KS, int TR = 16>
__global__ void __launch_bounds__(256) sweep_fused_kernel(FusedParams p) {
    if constexpr (TR == 16) sweep_fused_body<ALAY, DEPTH, KS>(p);
    else sweep_fused_body_tall<ALAY, DEPTH, KS, TR>(p);
}

// Zeroes rows r0 .. npad-1 of the [npad][Rpad] blocks of two arrays (blockIdx.y selects the
// array): the padding rows that the fused steps' strips do not cover, which the steps read as
// operand rows (against zero columns of the pages) and as tile halos (against zero couplings),
// so they must hold finite values -- zeros, as every other padding row does.
__global__ void pad_rows_zero_kernel(cf *a, int64_t na, cf *b, int64_t nb, int r0, int npad, int Rpad) {
    cf *dst = blockIdx.y ? b : a;
    const int64_t count = blockIdx.y ? nb : na;
    const int64_t per = (int64_t)(npad - r0) * Rpad;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= per * count) return;
    const int64_t blkno = i / per, e = i % per;
    dst[blkno * npad * Rpad + (int64_t)r0 * Rpad + e] = mk(0.f, 0.f);
}

// Kernel variants of the fused step and their per-shape choice.  Variant v: (DEPTH, KS, tile
// rows) = (2, 4, 16) five waves, (3, 4, 16), (4, 2, 16), (2, 2, 16) five waves, (2, 4, 16),
// (2, 2, 16), (2, 4, 32), (3, 4, 32), (2, 4, 64); a variant with TPW tiles per workgroup needs
// Rpad / 16 to be a multiple of TPW.
constexpr int N_FUSED_VARIANTS = 10;
constexpr int SPLITK_VARIANT = 9;   // 32 x 32 tiles, K over SK_PARTS workgroups, operands through LDS
constexpr int FUSED_KS[N_FUSED_VARIANTS] = {4, 4, 2, 2, 4, 2, 4, 4, 4, 4};
constexpr int FUSED_DEPTH[N_FUSED_VARIANTS] = {2, 3, 4, 2, 2, 2, 2, 3, 2, 1};
constexpr bool FUSED_FIVE[N_FUSED_VARIANTS] = {true, false, false, true, false, false, false, false, false, false};
constexpr int FUSED_TR[N_FUSED_VARIANTS] = {16, 16, 16, 16, 16, 16, 32, 32, 64, 32};

// strips of variant v's tiles over n rows (each owns FUSED_TR - 2 rows), and the rows they cover
inline int fused_strips(int v, int n) { return (n + FUSED_TR[v] - 3) / (FUSED_TR[v] - 2); }
inline int fused_covered(int v, int n, int npad) { return std::min(fused_strips(v, n) * (FUSED_TR[v] - 2), npad); }

// Workgroup geometry of variant v for fp's shape: strips, units (tiles; for the split-K variant one tile
// per SK_PARTS workgroups) and units per XCD slot.
void fused_geometry(int v, FusedParams &fp) {
    fp.strips = fused_strips(v, fp.n);
    if (v == SPLITK_VARIANT) {
        fp.units = fp.ntasks * fp.strips * (fp.Rpad / SK_TC);
        fp.per_xcd = (fp.units + 7) / 8;
        return;
    }
    const int tpw = 4 / FUSED_KS[v];
    fp.units = fp.ntasks * fp.strips * (fp.Rpad / 16) / tpw;
    fp.per_xcd = (fp.units + 7) / 8;
}
inline int splitk_tiles(int n, int Rpad) { return 2 * ((n + SK_OWN - 1) / SK_OWN) * (Rpad / SK_TC); }

template <int ALAY>
void launch_fused_variant(int v, const FusedParams &fp, hipStream_t s, bool profiled) {
    const dim3 grid(8 * fp.per_xcd * (v == SPLITK_VARIANT ? SK_PARTS : 1)), block(256);
    auto go = [&](auto kernel) {
        if (profiled) launch_k(PC_SWEEP_GEMM, kernel, grid, block, 0, s, fp);
        else {
            hipLaunchKernelGGL(kernel, grid, block, 0, s, fp);
            WUST_LAUNCH_CHECK();
        }
    };
    switch (v) {
        case 1: go(sweep_fused_kernel<ALAY, 3, 4>); break;
        case 2: go(sweep_fused_kernel<ALAY, 4, 2>); break;
        case 3: go(sweep_fused_kernel_d2<ALAY, 2>); break;
        case 4: go(sweep_fused_kernel<ALAY, 2, 4>); break;
        case 5: go(sweep_fused_kernel<ALAY, 2, 2>); break;
        case 6: go(sweep_fused_kernel<ALAY, 2, 4, 32>); break;
        case 7: go(sweep_fused_kernel<ALAY, 3, 4, 32>); break;
        case 8: go(sweep_fused_kernel<ALAY, 2, 4, 64>); break;
        case SPLITK_VARIANT: go(sweep_splitk_kernel<ALAY>); break;
        default: go(sweep_fused_kernel_d2<ALAY, 4>); break;
    }
}

struct FusedKey {
    int npad, Rpad, adjoint;
    bool operator<(const FusedKey &o) const {
        return std::tie(npad, Rpad, adjoint) < std::tie(o.npad, o.Rpad, o.adjoint);
    }
};
std::mutex g_fused_mu;
std::map<FusedKey, int> g_fused_choice;

bool fused_eligible(int v, const FusedParams &fp) {
    if (v == SPLITK_VARIANT) return fp.skpart && fp.skcount && fp.Rpad % SK_TC == 0 && SK_TR <= fp.npad;
    return (fp.Rpad / 16) % (4 / FUSED_KS[v]) == 0 && FUSED_TR[v] <= fp.npad;
}

// Times every eligible variant on `fp` (a real step whose outputs the sweep rewrites before
// reading them) and returns the fastest; the choice is kept per (npad, Rpad, adjoint).
// WUST_SWEEP_VARIANT=v forces variant v where it is eligible (read per solve, for tests).
int fused_variant(const FusedParams &fp, bool adjoint, hipStream_t s) {
    if (const char *e = std::getenv("WUST_SWEEP_VARIANT"); e && e[0]) {
        const int v = std::atoi(e);
        if (v >= 0 && v < N_FUSED_VARIANTS && fused_eligible(v, fp)) return v;
    }
    static const bool tune = [] {
        const char *e = std::getenv("WUST_AUTOTUNE");
        return !(e && e[0] == '0');
    }();
    if (!tune) return 0;
    const FusedKey key{fp.npad, fp.Rpad, adjoint ? 1 : 0};
    {
        std::lock_guard<std::mutex> lk(g_fused_mu);
        auto it = g_fused_choice.find(key);
        if (it != g_fused_choice.end()) return it->second;
    }
    // Each eligible variant runs BURST back-to-back launches (as the sweep runs them) per round,
    // rounds interleaving the variants; a variant's time is its best round (the least disturbed).
    constexpr int ROUNDS = 3, BURST = 8;
    hipEvent_t e0, e1;
    WUST_HIP_CHECK(hipEventCreate(&e0));
    WUST_HIP_CHECK(hipEventCreate(&e1));
    FusedParams q[N_FUSED_VARIANTS];
    bool ok[N_FUSED_VARIANTS];
    float per_launch[N_FUSED_VARIANTS];
    auto run = [&](int v) {
        if (adjoint) launch_fused_variant<A_CONJT>(v, q[v], s, false);
        else launch_fused_variant<A_COL>(v, q[v], s, false);
    };
    for (int v = 0; v < N_FUSED_VARIANTS; ++v) {
        ok[v] = fused_eligible(v, fp);
        per_launch[v] = 1e30f;
        q[v] = fp;
        fused_geometry(v, q[v]);
        if (ok[v]) run(v);    // warms the code object and the caches
    }
    for (int round = 0; round < ROUNDS; ++round)
        for (int v = 0; v < N_FUSED_VARIANTS; ++v) {
            if (!ok[v]) continue;
            WUST_HIP_CHECK(hipEventRecord(e0, s));
            for (int b = 0; b < BURST; ++b) run(v);
            WUST_HIP_CHECK(hipEventRecord(e1, s));
            WUST_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            WUST_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            per_launch[v] = std::min(per_launch[v], ms / BURST);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int best = 0;
    for (int v = 0; v < N_FUSED_VARIANTS; ++v)
        if (ok[v] && (!ok[best] || per_launch[v] < per_launch[best])) best = v;
    if (const char *e = std::getenv("WUST_AUTOTUNE_LOG"); e && e[0] == '1') {
        std::fprintf(stderr, "sweep step variants, npad %d Rpad %d %s (us per launch):", fp.npad, fp.Rpad,
                     adjoint ? "adjoint" : "forward");
        for (int v = 0; v < N_FUSED_VARIANTS; ++v)
            if (ok[v]) std::fprintf(stderr, " [%d] %.2f", v, 1e3f * per_launch[v]);
        std::fprintf(stderr, "\n");
        std::fprintf(stderr, "sweep step kernel: npad %d Rpad %d %s -> %d-row tiles, depth %d, K split %d%s (%.2f us)\n",
                     fp.npad, fp.Rpad, adjoint ? "adjoint" : "forward", FUSED_TR[best], FUSED_DEPTH[best],
                     FUSED_KS[best], FUSED_FIVE[best] ? ", five waves" : "", 1e3f * per_launch[best]);
    }
    std::lock_guard<std::mutex> lk(g_fused_mu);
    g_fused_choice[key] = best;
    return best;
}

}  // namespace

static void solve_engine_unfused(const Operator &op, const cf *src, cf *out, cf *scratch, int Rpad, bool adjoint,
                                 hipStream_t s);

void solve_engine(const Operator &op, const cf *src, cf *out, cf *work, int Rpad, bool adjoint, hipStream_t s,
                  int lead, int trail) {
    static const bool unfused = [] {
        const char *e = std::getenv("WUST_SWEEP_UNFUSED");
        return e && e[0] == '1';
    }();
    const int m = op.m, p = op.p, npad = op.npad;
    const int64_t blk = (int64_t)npad * Rpad;
    cf *yf = work;                                 // forward iterates y_j / z_j: [m][npad][Rpad]
    cf *ops = work + (int64_t)m * blk;             // operand blocks: [2 chains][2 ping-pong]
    // split-K variant: partial tiles, then the arrival counts (cleared per solve)
    cf *sk_part = work + ((int64_t)m + 4) * blk;
    const int sk_tiles = splitk_tiles(op.n, Rpad);
    int *sk_count = reinterpret_cast<int *>(sk_part + (int64_t)sk_tiles * SK_PARTS * SK_TR * SK_TC);
    hipLaunchKernelGGL(sk_clear_kernel, dim3((sk_tiles + 255) / 256), dim3(256), 0, s, sk_count, sk_tiles);
    WUST_LAUNCH_CHECK();
    if (unfused) {
        solve_engine_unfused(op, src, out, ops, Rpad, adjoint, s);
        return;
    }
    auto S = [&](int j) { return src + j * blk; };
    auto Y = [&](int j) { return yf + j * blk; };
    auto O = [&](int j) { return out + j * blk; };
    auto OP = [&](int chain, int pp) { return ops + (int64_t)(2 * chain + pp) * blk; };
    int Ld[3] = {DNW, DW, DSW};
    int Ud[3] = {DNE, DE, DSE};
    // strips cover the n real rows only (1280 rather than 1344 workgroups at 558 rows, 256 RHS and
    // 16-row tiles: five per CU); the padding rows beyond what every variant's strips cover are
    // zeroed once per solve
    int covered = npad;
    for (int v = 0; v < N_FUSED_VARIANTS; ++v) covered = std::min(covered, fused_covered(v, op.n, npad));
    if (covered < npad) {
        const int64_t per = (int64_t)(npad - covered) * Rpad, most = per * (m + 4);
        hipLaunchKernelGGL(pad_rows_zero_kernel, dim3((unsigned)((most + 255) / 256), 2), dim3(256), 0, s, work,
                           (int64_t)m + 4, out, (int64_t)m, covered, npad, Rpad);
        WUST_LAUNCH_CHECK();
    }
    auto couple = [&](FusedTask &tk, const int dirs[3], int cblock, bool transposed) {
        tk.cn = op.coef(dirs[0], cblock);
        tk.cc = op.coef(dirs[1], cblock);
        tk.cs = op.coef(dirs[2], cblock);
        tk.transposed = transposed ? 1 : 0;
    };
    int variant = -1;
    auto launch = [&](std::vector<FusedTask> &tasks) {
        if (tasks.empty()) return;
        FusedParams fp{};
        fp.ntasks = (int)tasks.size();
        for (int i = 0; i < fp.ntasks; ++i) fp.task[i] = tasks[i];
        fp.n = op.n;
        fp.npad = npad;
        fp.Rpad = Rpad;
        fp.skpart = sk_part;
        fp.skcount = sk_count;
        fp.strips = fused_strips(0, op.n);
        if (variant < 0) variant = fused_variant(fp, adjoint, s);
        fused_geometry(variant, fp);
        prof_units(PC_SWEEP_GEMM, fp.ntasks);
        if (adjoint) launch_fused_variant<A_CONJT>(variant, fp, s, true);
        else launch_fused_variant<A_COL>(variant, fp, s, true);
    };
    const int nl = p, nr = m - 1 - p;
    const int nsteps = nl > nr ? nl : nr;
    // a chain's iterates before the first nonzero column of src are exactly zero: its
    // elimination starts there (l0 / r0 steps in), and those columns' back substitution has no
    // y_j term; each chain computes at least its last step, which the meeting block reads
    const int l0 = std::max(0, std::min(lead, nl - 1)), r0 = std::max(0, std::min(trail, nr - 1));
    const int nfwd = std::max(nl - l0, nr - r0);
    // elimination from both ends: y_j = X_j B_j, and B_{next} from y_j in the epilogue
    for (int k = 0; k < nfwd; ++k) {
        std::vector<FusedTask> tasks;
        if (int st = l0 + k; st < nl) {
            const int j = st;
            FusedTask tk{};
            tk.X = op.page(j);
            tk.B = (st == l0) ? S(j) : OP(0, st & 1);
            tk.out = Y(j);
            if (st + 1 < nl) {                       // next left block j + 1 (not the meeting block)
                tk.Bn = OP(0, (st + 1) & 1);
                tk.nsrc = S(j + 1);
                if (!adjoint) couple(tk, Ld, j + 1, false);    // - L_{j+1} y_j
                else couple(tk, Ud, j, true);                  // - U_j^H y_j
            }
            tasks.push_back(tk);
        }
        if (int st = r0 + k; st < nr) {
            const int j = m - 1 - st;
            FusedTask tk{};
            tk.X = op.page(j);
            tk.B = (st == r0) ? S(j) : OP(1, st & 1);
            tk.out = Y(j);
            if (st + 1 < nr) {                       // next right block j - 1
                tk.Bn = OP(1, (st + 1) & 1);
                tk.nsrc = S(j - 1);
                if (!adjoint) couple(tk, Ud, j - 1, false);    // - U_{j-1} z_j
                else couple(tk, Ld, j, true);                  // - L_j^H z_j
            }
            tasks.push_back(tk);
        }
        launch(tasks);
    }
    // meeting block u_p = X_p (b_p - L_p y_{p-1} - U_p z_{p+1}), and the operands of the first
    // back-substitution steps, u_p's couplings: operand passes + the plain GEMM
    {
        auto term = [&](SweepTask &tk, const int dirs[3], int cblock, const cf *v, bool transposed) {
            SweepTerm &t = tk.t[tk.nterms++];
            t.cn = op.coef(dirs[0], cblock);
            t.cc = op.coef(dirs[1], cblock);
            t.cs = op.coef(dirs[2], cblock);
            t.v = v;
            t.transposed = transposed ? 1 : 0;
        };
        auto run = [&](std::vector<SweepTask> &tasks, bool gemm) {
            SweepParams sp{};
            sp.ntasks = (int)tasks.size();
            for (int i = 0; i < sp.ntasks; ++i) sp.task[i] = tasks[i];
            sp.npad = npad;
            sp.Rpad = Rpad;
            const int64_t nops = (int64_t)sp.ntasks * blk;
            launch_k(PC_SWEEP_OPERAND, sweep_operand_kernel, dim3((unsigned)((nops + 255) / 256)), dim3(256), 0, s,
                     sp);
            if (!gemm) return;
            sp.units = sp.ntasks * (npad / (16 * SWEEP_RB)) * (Rpad / 16);
            sp.per_xcd = (sp.units + 7) / 8;
            prof_units(PC_SWEEP_GEMM, sp.ntasks);
            if (adjoint)
                launch_k(PC_SWEEP_GEMM, sweep_gemm_kernel<SWEEP_RB, A_CONJT>, dim3(8 * sp.per_xcd), dim3(256), 0, s,
                         sp);
            else
                launch_k(PC_SWEEP_GEMM, sweep_gemm_kernel<SWEEP_RB, A_COL>, dim3(8 * sp.per_xcd), dim3(256), 0, s, sp);
        };
        std::vector<SweepTask> mid(1);
        SweepTask &tk = mid[0];
        tk = SweepTask{};
        tk.X = op.page(p);
        tk.xconj = adjoint ? 1 : 0;
        tk.src = S(p);
        tk.B = OP(0, 0);
        tk.out = O(p);
        if (p >= 1) {
            if (!adjoint) term(tk, Ld, p, Y(p - 1), false);
            else term(tk, Ud, p - 1, Y(p - 1), true);
        }
        if (p + 1 <= m - 1) {
            if (!adjoint) term(tk, Ud, p, Y(p + 1), false);
            else term(tk, Ld, p + 1, Y(p + 1), true);
        }
        run(mid, true);
        std::vector<SweepTask> first;
        if (nl >= 1) {                                 // left j = p - 1: - X_j U_j u_p
            SweepTask t{};
            t.B = OP(0, 1);
            if (!adjoint) term(t, Ud, p - 1, O(p), false);
            else term(t, Ld, p, O(p), true);
            first.push_back(t);
        }
        if (nr >= 1) {                                 // right j = p + 1: - X_j L_j u_p
            SweepTask t{};
            t.B = OP(1, 1);
            if (!adjoint) term(t, Ld, p + 1, O(p), false);
            else term(t, Ud, p, O(p), true);
            first.push_back(t);
        }
        if (!first.empty()) run(first, false);
    }
    // back substitution outward: u_j = y_j + X_j B_j with B_j = - (coupling) u_{j +- 1}
    for (int st = 0; st < nsteps; ++st) {
        std::vector<FusedTask> tasks;
        if (st < nl) {
            const int j = p - 1 - st;
            FusedTask tk{};
            tk.X = op.page(j);
            tk.B = OP(0, (st + 1) & 1);
            tk.base = (j >= l0) ? Y(j) : nullptr;
            tk.out = O(j);
            if (st + 1 < nl) {                       // next left block j - 1
                tk.Bn = OP(0, st & 1);
                if (!adjoint) couple(tk, Ud, j - 1, false);    // - U_{j-1} u_j
                else couple(tk, Ld, j, true);                  // - L_j^H u_j
            }
            tasks.push_back(tk);
        }
        if (st < nr) {
            const int j = p + 1 + st;
            FusedTask tk{};
            tk.X = op.page(j);
            tk.B = OP(1, (st + 1) & 1);
            tk.base = (j <= m - 1 - r0) ? Y(j) : nullptr;
            tk.out = O(j);
            if (st + 1 < nr) {                       // next right block j + 1
                tk.Bn = OP(1, st & 1);
                if (!adjoint) couple(tk, Ld, j + 1, false);    // - L_{j+1} u_j
                else couple(tk, Ud, j, true);                  // - U_j^H u_j
            }
            tasks.push_back(tk);
        }
        launch(tasks);
    }
}

static void solve_engine_unfused(const Operator &op, const cf *src, cf *out, cf *scratch, int Rpad, bool adjoint,
                                 hipStream_t s) {
    const int m = op.m, p = op.p, npad = op.npad;
    const int64_t blk = (int64_t)npad * Rpad;
    auto S = [&](int j) { return src + j * blk; };
    auto O = [&](int j) { return out + j * blk; };
    int Ld[3] = {DNW, DW, DSW};
    int Ud[3] = {DNE, DE, DSE};
    auto term = [&](SweepTask &tk, const int dirs[3], int cblock, int vblock, bool transposed) {
        SweepTerm &t = tk.t[tk.nterms++];
        t.cn = op.coef(dirs[0], cblock);
        t.cc = op.coef(dirs[1], cblock);
        t.cs = op.coef(dirs[2], cblock);
        t.v = O(vblock);
        t.transposed = transposed ? 1 : 0;
    };
    auto launch = [&](std::vector<SweepTask> &tasks) {
        if (tasks.empty()) return;
        SweepParams sp{};
        sp.ntasks = (int)tasks.size();
        for (int i = 0; i < sp.ntasks; ++i) sp.task[i] = tasks[i];
        for (int i = 0; i < sp.ntasks; ++i) sp.task[i].B = scratch + i * blk;
        sp.npad = npad;
        sp.Rpad = Rpad;
        const int64_t nops = (int64_t)sp.ntasks * blk;
        launch_k(PC_SWEEP_OPERAND, sweep_operand_kernel, dim3((unsigned)((nops + 255) / 256)), dim3(256), 0, s, sp);
        sp.units = sp.ntasks * (npad / (16 * SWEEP_RB)) * (Rpad / 16);
        sp.per_xcd = (sp.units + 7) / 8;
        prof_units(PC_SWEEP_GEMM, sp.ntasks);
        if (adjoint)
            launch_k(PC_SWEEP_GEMM, sweep_gemm_kernel<SWEEP_RB, A_CONJT>, dim3(8 * sp.per_xcd), dim3(256), 0, s, sp);
        else
            launch_k(PC_SWEEP_GEMM, sweep_gemm_kernel<SWEEP_RB, A_COL>, dim3(8 * sp.per_xcd), dim3(256), 0, s, sp);
    };
    auto base = [&](int j, bool with_src, bool acc) {
        SweepTask tk{};
        tk.X = op.page(j);
        tk.xconj = adjoint ? 1 : 0;
        tk.src = with_src ? S(j) : nullptr;
        tk.nterms = 0;
        tk.out = O(j);
        tk.accumulate = acc ? 1 : 0;
        return tk;
    };
    const int nl = p, nr = m - 1 - p;
    const int nsteps = nl > nr ? nl : nr;
    // elimination from both ends
    for (int st = 0; st < nsteps; ++st) {
        std::vector<SweepTask> tasks;
        if (st < nl) {
            int j = st;
            SweepTask tk = base(j, true, false);
            if (j >= 1) {
                if (!adjoint) term(tk, Ld, j, j - 1, false);      // - L_j y_{j-1}
                else term(tk, Ud, j - 1, j - 1, true);             // - U_{j-1}^H y_{j-1}
            }
            tasks.push_back(tk);
        }
        if (st < nr) {
            int j = m - 1 - st;
            SweepTask tk = base(j, true, false);
            if (j <= m - 2) {
                if (!adjoint) term(tk, Ud, j, j + 1, false);      // - U_j z_{j+1}
                else term(tk, Ld, j + 1, j + 1, true);             // - L_{j+1}^H z_{j+1}
            }
            tasks.push_back(tk);
        }
        launch(tasks);
    }
    // meeting block
    {
        std::vector<SweepTask> tasks;
        SweepTask tk = base(p, true, false);
        if (p >= 1) {
            if (!adjoint) term(tk, Ld, p, p - 1, false);
            else term(tk, Ud, p - 1, p - 1, true);
        }
        if (p + 1 <= m - 1) {
            if (!adjoint) term(tk, Ud, p, p + 1, false);
            else term(tk, Ld, p + 1, p + 1, true);
        }
        tasks.push_back(tk);
        launch(tasks);
    }
    // back substitution outward
    for (int st = 0; st < nsteps; ++st) {
        std::vector<SweepTask> tasks;
        if (st < nl) {
            int j = p - 1 - st;
            SweepTask tk = base(j, false, true);
            if (!adjoint) term(tk, Ud, j, j + 1, false);           // - X_j U_j u_{j+1}
            else term(tk, Ld, j + 1, j + 1, true);                 // - X_j^H L_{j+1}^H u_{j+1}
            tasks.push_back(tk);
        }
        if (st < nr) {
            int j = p + 1 + st;
            SweepTask tk = base(j, false, true);
            if (!adjoint) term(tk, Ld, j, j - 1, false);           // - X_j L_j u_{j-1}
            else term(tk, Ud, j - 1, j - 1, true);                 // - X_j^H U_{j-1}^H u_{j-1}
            tasks.push_back(tk);
        }
        launch(tasks);
    }
}

}  // namespace wust
