// Complex-FP32 block sweeps over grid columns, batched over all right-hand sides.
//
// The reference (Runtime/solver/applyBlockLU.cu:433-467) runs, per column, one
// tridiagonal kernel and one cuBLAS Cgemm3m (Ny x R x Ny), serially over x
// for the forward and backward passes.  Here each step is an operand pass
// (the tridiagonal couplings, evaluated once per element) followed by a
// complex GEMM on MFMA (v_mfma_f32_16x16x4f32, three real products per complex
// product, FP32 accumulation per k-block of 16 folded into running sums), and the
// twisted factorization lets the left and
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
#ifndef WUST_SWEEP_M3
#define WUST_SWEEP_M3 1
#endif

namespace wust {

namespace {

// The sweep products accumulate each k-block on MFMA and fold it into running sums (CgemmWave FOLD):
// one long FP32 accumulation over the block rows would set the refinement contraction of the whole
// solve.  They take the 3M form (three real products per complex product): with the fold its extra
// rounding leaves the contraction where four products have it (at 560^2 x 256 RHS the relative
// residual after the first c64 solve is 3.7e-6 forward / 5.9e-6 adjoint).
constexpr bool SWEEP_FOLD = WUST_SWEEP_FOLD != 0;
constexpr bool SWEEP_M3 = WUST_SWEEP_M3 != 0;

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
            acc = cmul_fixed(st.cc[k], v[(int64_t)k * Rpad]);
            if (k >= 1) cfma(acc, st.cn[k], v[(int64_t)(k - 1) * Rpad]);
            if (k + 1 < npad) cfma(acc, st.cs[k], v[(int64_t)(k + 1) * Rpad]);
        } else {
            acc = cmul_fixed(conjf_(st.cc[k]), v[(int64_t)k * Rpad]);
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
// round-robin and each accumulates the whole tile (CgemmWave: three real
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

    CgemmWave<RB, 1, ALAY, B_ROW, SWEEP_M3, 2, SWEEP_FOLD> acc;
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
// its launch disappear, at the cost of 16/14 of the MFMA work.  Back substitution reads its
// base (y_j) from a different array than it writes (u_j), so halo rows are never read after a
// neighbour rewrote them.
constexpr int FUSED_OWN = 14;   // owned rows of a 16-row tile

// Stores of a sweep step's outputs (its result block and the next step's operand block) are
// non-temporal: no step reads them back through this XCD's L2 (the next step's workgroups on the
// other XCDs see them only after the kernel boundary's write-back, the back substitution and the
// residual much later), so they go out without holding L2 lines that the boundary must flush.
// 64.5-64.7 against 65.5-65.9 ms per evaluation at 560^2 x 256, same box.
__device__ __forceinline__ void store_stream(cf *p, cf v) {
    typedef float v2f_st __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2f_st{v.re, v.im}, reinterpret_cast<v2f_st *>(p));
}

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
};

// The four waves of a workgroup share one 16 x 16 tile (k-blocks dealt round-robin, partial tiles
// summed in LDS in wave order); DEPTH k-blocks of operands are in flight per wave.  BUF: buffer
// loads (scalar k offsets) instead of flat loads; LATE: the epilogue operands are loaded after the
// GEMM instead of before it (fewer registers live across the k loop).
template <int ALAY, int DEPTH, bool LATE = false, bool BUF = false>
__device__ __forceinline__ void sweep_fused_body(const FusedParams &p) {
    constexpr int KS = 4, TPW = 1, NW = 4;
    constexpr bool FOLD = SWEEP_FOLD;
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
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform (scalar k offsets)
    const int tw = wave / KS, ks = wave % KS;
    const int j0 = (ng * TPW + tw) * 16;

    // epilogue operands (base rows of this lane's tile, and for this thread's output element the
    // next source and the three coupling coefficients) are loaded before the GEMM so that their
    // latency overlaps it instead of trailing it
    const int t = (threadIdx.x >> 4) & 15, cc = threadIdx.x & 15;
    const int k = i0 + t;
    const bool own = k >= o0 && k <= o1;
    cf pbase[4];
    cf pn = mk(0.f, 0.f), pc = mk(0.f, 0.f), ps = mk(0.f, 0.f), psrc[TPW];
    auto load_epilogue = [&] {
        if (tk.base && ks == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) pbase[r] = tk.base[(int64_t)(i0 + 4 * g + r) * Rpad + j0 + c];
        }
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
    };
    if (!LATE) load_epilogue();

    CgemmWave<1, 1, ALAY, B_ROW, SWEEP_M3, DEPTH, FOLD> acc;
    acc.zero();
    const int64_t a_off = (ALAY == A_COL) ? i0 : (int64_t)i0 * npad;
    const cf *A = tk.X + a_off;
    if (BUF)
        acc.run_buf(A, npad, (int64_t)npad * npad - a_off, tk.B + j0, Rpad, (int64_t)npad * Rpad - j0, ks, npad / 16,
                    KS, g, c);
    else
        acc.run(A, npad, tk.B + j0, Rpad, ks, npad / 16, KS, g, c);
    if (LATE) load_epilogue();

    __shared__ float4 red[NW][2][64];
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
        store_stream(tk.out + off, v);
        if (!tk.Bn) continue;
        const cf vn = (k >= 1) ? tile[q][t - 1][cc] : mk(0.f, 0.f);        // row k - 1 (in the tile when k >= 1)
        const cf vs = (k + 1 < npad) ? tile[q][t + 1][cc] : mk(0.f, 0.f);  // row k + 1 (in the tile when < npad)
        // the same products and order as operand_b: P v = c_c v_k + c_n v_{k-1} + c_s v_{k+1}
        // (transposed: conj(cc_k) v_k + conj(cn_{k+1}) v_{k+1} + conj(cs_{k-1}) v_{k-1})
        cf a = cmul_fixed(pc, v);
        if (!tk.transposed) {
            if (k >= 1) cfma(a, pn, vn);
            if (k + 1 < npad) cfma(a, ps, vs);
        } else {
            if (k + 1 < npad) cfma(a, ps, vs);
            if (k >= 1) cfma(a, pn, vn);
        }
        store_stream(tk.Bn + off, psrc[q] - a);
    }
}

// Epilogue operands of a fused step for one thread: the base rows of its wave's tile (pbase, wave 0
// only), and for its output element (row k, column j0 + cc) the coupling of row k -- pc, then pa
// and pz in the order of their use (c_n[k], c_s[k]; transposed c_s[k-1], c_n[k+1]), raw -- and the
// next step's source.  Within a wave every lane loads them, with clamped indices, whatever it
// needs (an absent source is a zero-extent buffer, which reads as zeros): loads under divergent
// branches (only the owned rows need them) would make the compiler drain every load in flight, the
// GEMM's first operands included, at the branch joins.  Whether a wave loads them at all depends
// only on wave-uniform facts (wave 0 of a step with a base; a step with a next operand), so those
// branches do not diverge.  next() forms B' = src' - P v with the products and order of
// operand_b.
struct StepEpilogue {
    typedef float v2f_t __attribute__((ext_vector_type(2)));
    __amdgpu_buffer_rsrc_t rbase, rcc, rca, rcz, rsrc;
    uint32_t obase, osrc;
    int k, ka, kz, npad;
    bool tr, use_base, use_next;
    cf pbase[4], pc, pa, pz, psrc;

    __device__ __forceinline__ static __amdgpu_buffer_rsrc_t rsrc_or_zero(const cf *ptr, int64_t elems, bool use) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(ptr), (short)0, (use && ptr) ? (int)(elems * 8) : 0,
                                                 0x00020000);
    }
    // aux 0: cached; aux 2: non-temporal (the base and next-source blocks are read once per launch:
    // kept out of L2, they leave it to the pages and operand blocks that the strips share; 1.2-1.6 ms
    // per evaluation at 560^2 x 256, same box)
    template <int AUX = 0>
    __device__ __forceinline__ static cf ld(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
        const v2f_t u = __builtin_bit_cast(v2f_t, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, AUX));
        return mk(u.x, u.y);
    }
    __device__ __forceinline__ StepEpilogue(const FusedTask &tk, int npad_, int Rpad, int i0, int j0, int k_, int cc,
                                            bool wave0) {
        npad = npad_;
        k = k_;
        tr = tk.transposed != 0;
        use_base = wave0 && tk.base != nullptr;
        use_next = tk.Bn != nullptr;
        rbase = rsrc_or_zero(tk.base, (int64_t)npad * Rpad, use_base);
        rcc = rsrc_or_zero(tk.cc, npad, use_next);
        rca = rsrc_or_zero(tr ? tk.cs : tk.cn, npad, use_next);
        rcz = rsrc_or_zero(tr ? tk.cn : tk.cs, npad, use_next);
        rsrc = rsrc_or_zero(tk.nsrc, (int64_t)npad * Rpad, use_next);
        ka = tr ? max(k - 1, 0) : k;
        kz = tr ? min(k + 1, npad - 1) : k;
        obase = (uint32_t)(((int64_t)i0 * Rpad + j0) * 8);
        osrc = (uint32_t)(((int64_t)k * Rpad + j0 + cc) * 8);
    }
    // lane (g, c) of the tile loads base rows i0 + 4 g + r of column j0 + c
    __device__ __forceinline__ void load(int g, int c, int Rpad) {
        if (use_base) {
#pragma unroll
            for (int r = 0; r < 4; ++r) pbase[r] = ld<2>(rbase, obase + (uint32_t)(((4 * g + r) * Rpad + c) * 8));
        }
        if (use_next) {
            pc = ld(rcc, (uint32_t)k * 8);
            pa = ld(rca, (uint32_t)ka * 8);
            pz = ld(rcz, (uint32_t)kz * 8);
            psrc = ld<2>(rsrc, osrc);
        }
    }
    // B' = src' - P v for output element v with row neighbours vn (k - 1) and vs (k + 1)
    __device__ __forceinline__ cf next(cf v, cf vn, cf vs) const {
        cf a;
        if (!tr) {
            a = cmul_fixed(pc, v);
            if (k >= 1) cfma(a, pa, vn);
            if (k + 1 < npad) cfma(a, pz, vs);
        } else {
            a = cmul_fixed(conjf_(pc), v);
            if (k + 1 < npad) cfma(a, conjf_(pz), vs);
            if (k >= 1) cfma(a, conjf_(pa), vn);
        }
        return psrc - a;
    }
};

// The 16-row step with operands loaded 16 bytes per lane.  The MFMA wants lane (g, c) to hold row
// c of A (and column c of B) at k = 4 g + s, which no layout here stores contiguously per lane, so
// the flat loads above move 8 bytes per lane per instruction.  The per-CU load path costs about the
// same per instruction for 8 and 16 bytes per lane when a wave's addresses cover whole cache lines,
// so here each wave fetches its k-block (16 x 16 of A and of B, 2 KiB each) as four 16-byte-per-lane
// buffer loads of whole lines into registers, writes them to a wave-private LDS tile in [k][row]
// order (transposing A for the adjoint), and reads the MFMA operands back from LDS.  Each wave uses
// only its own LDS tile and LDS operations of one wave complete in order, so no barrier is needed;
// DEPTH k-blocks of loads are in flight.  Same K split, accumulation order and epilogue as the
// 16-row body with KS = 4, so the results are bit-identical to it.
// LDS_B: B staged through LDS too (else each lane loads its B operands directly, 8 bytes per lane:
// B's rows are contiguous along the MFMA lanes, so those loads already read whole lines).  The
// epilogue operands are loaded after the GEMM, or with MID right after the first k-block's operand
// loads, so that the GEMM hides their latency without their loads delaying its first MFMAs
// (loading them before the GEMM's first loads measured slower).
template <int ALAY, int DEPTH, bool LDS_B = true, bool MID = false>
__device__ __forceinline__ void sweep_fused_body_lds(const FusedParams &p) {
    constexpr int KS = 4, STR = 17;   // LDS row stride in complex elements (16 + 1 pad: no bank conflicts)
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int L = xcd * p.per_xcd + local;
    if (L >= p.units) return;
    const int npad = p.npad, Rpad = p.Rpad;
    const int Ng = Rpad / 16;
    const int ng = L % Ng;
    const int mb = (L / Ng) % p.strips;
    const int task = L / (Ng * p.strips);
    const FusedTask &tk = p.task[task];
    const int o0 = FUSED_OWN * mb;
    const int o1 = min(o0 + FUSED_OWN - 1, npad - 1);
    const int i0 = min(max(o0 - 1, 0), npad - 16);
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ks = wave;
    const int j0 = ng * 16;
    const int t = (threadIdx.x >> 4) & 15, cc = threadIdx.x & 15;
    const int k = i0 + t;
    const bool own = k >= o0 && k <= o1;

    // operand staging: per wave A[k][i] and B[k][j] of one k-block
    __shared__ cf ops_lds[4][LDS_B ? 2 : 1][16][STR];
    cf(*LA)[STR] = ops_lds[wave][0];
    cf(*LB)[STR] = ops_lds[wave][LDS_B ? 1 : 0];   // (unused without LDS_B)
    const int64_t a_off = (ALAY == A_COL) ? i0 : (int64_t)i0 * npad;
    const auto ra = __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(tk.X + a_off), (short)0,
                                                      (int)(((int64_t)npad * npad - a_off) * 8), 0x00020000);
    const auto rbuf = __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(tk.B + j0), (short)0,
                                                        (int)(((int64_t)npad * Rpad - j0) * 8), 0x00020000);
    // lane l fetches 16 bytes at line (l / 8) + 8 q, elements 2 (l % 8) + {0, 1} of the line; a line
    // is a k of A_COL and B (16 rows / columns), a row of A_CONJT (16 k)
    const int ln = lane >> 3, el = 2 * (lane & 7);
    const uint32_t va = (uint32_t)((ln * npad + el) * 8), vb = (uint32_t)((ln * Rpad + el) * 8);
    const uint32_t vbd = (uint32_t)((4 * g * Rpad + c) * 8);   // direct B: row 4 g + s, column c
    struct Stage {
        v4f a[2], b[2];   // b: B lines (LDS_B) or, as complex pairs, B[4 g + s][c] for s = 0..3
    };
    auto gload = [&](Stage &r, int k0) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t sa = (ALAY == A_COL) ? (uint32_t)((k0 + 8 * q) * npad * 8) : (uint32_t)((8 * q * npad + k0) * 8);
            r.a[q] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(ra, va, sa, 0));
            if (LDS_B) {
                r.b[q] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rbuf, vb, (uint32_t)((k0 + 8 * q) * Rpad * 8), 0));
            } else {
                typedef float v2f_t __attribute__((ext_vector_type(2)));
                const v2f_t u0 = __builtin_bit_cast(v2f_t, __builtin_amdgcn_raw_buffer_load_b64(rbuf, vbd, (uint32_t)((k0 + 2 * q) * Rpad * 8), 0));
                const v2f_t u1 = __builtin_bit_cast(v2f_t, __builtin_amdgcn_raw_buffer_load_b64(rbuf, vbd, (uint32_t)((k0 + 2 * q + 1) * Rpad * 8), 0));
                r.b[q] = v4f{u0.x, u0.y, u1.x, u1.y};
            }
        }
    };
    auto stage = [&](const Stage &r) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int ln8 = ln + 8 * q;
            if (ALAY == A_COL) {
                LA[ln8][el] = mk(r.a[q].x, r.a[q].y);
                LA[ln8][el + 1] = mk(r.a[q].z, r.a[q].w);
            } else {   // row ln8 of A, k = el, el + 1: conjugated into LA[k][row]
                LA[el][ln8] = mk(r.a[q].x, -r.a[q].y);
                LA[el + 1][ln8] = mk(r.a[q].z, -r.a[q].w);
            }
            if (LDS_B) {
                LB[ln8][el] = mk(r.b[q].x, r.b[q].y);
                LB[ln8][el + 1] = mk(r.b[q].z, r.b[q].w);
            }
        }
    };
    StepEpilogue ep(tk, npad, Rpad, i0, j0, k, cc, ks == 0);
    auto load_epilogue = [&] { ep.load(g, c, Rpad); };

    using Acc = CgemmWave<1, 1, ALAY, B_ROW, SWEEP_M3, DEPTH, SWEEP_FOLD>;
    Acc acc;
    acc.zero();
    {
        const int nkb = npad / 16;
        const int nk = (nkb - ks + KS - 1) / KS;
        const int last = ks + (nk - 1) * KS;
        auto k0 = [&](int u) {
            const int kb = ks + u * KS;
            return 16 * (kb < last ? kb : last);
        };
        Stage st[DEPTH + 1];
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) gload(st[u], k0(u));
        if (MID) {
            __builtin_amdgcn_sched_barrier(0);
            load_epilogue();
            __builtin_amdgcn_sched_barrier(0);
        }
        for (int t0 = 0; t0 < nk; t0 += DEPTH + 1) {
#pragma unroll
            for (int u = 0; u <= DEPTH; ++u) {
                gload(st[(u + DEPTH) % (DEPTH + 1)], k0(t0 + u + DEPTH));
                __builtin_amdgcn_sched_barrier(0);
                if (t0 + u < nk) {
                    stage(st[u]);
                    typename Acc::Ops o;
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        o.a[0][s] = LA[4 * g + s][c];
                        if (LDS_B) o.b[0][s] = LB[4 * g + s][c];
                    }
                    if (!LDS_B) {
                        const Stage &r = st[u];
                        o.b[0][0] = mk(r.b[0].x, r.b[0].y);
                        o.b[0][1] = mk(r.b[0].z, r.b[0].w);
                        o.b[0][2] = mk(r.b[1].x, r.b[1].y);
                        o.b[0][3] = mk(r.b[1].z, r.b[1].w);
                    }
                    // raised wave priority for the MFMA phase: a wave that has its operands issues
                    // its MFMAs ahead of the other waves' staging work, so the waves of a SIMD
                    // drift apart in phase instead of all waiting for the matrix core together
                    __builtin_amdgcn_s_setprio(1);
                    acc.mma(o);
                    __builtin_amdgcn_s_setprio(0);
                    acc.fold();
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }

    if (!MID) load_epilogue();

    __shared__ float4 red[4][2][64];
    __shared__ cf tile[16][17];
    float re[4], im[4];
    acc.result(0, 0, re, im);
    if (ks > 0) {
        red[wave][0][lane] = make_float4(re[0], re[1], re[2], re[3]);
        red[wave][1][lane] = make_float4(im[0], im[1], im[2], im[3]);
    }
    __syncthreads();
    if (ks == 0) {
#pragma unroll
        for (int w = 1; w < KS; ++w) {
            const float4 u = red[wave + w][0][lane], v = red[wave + w][1][lane];
            re[0] += u.x; re[1] += u.y; re[2] += u.z; re[3] += u.w;
            im[0] += v.x; im[1] += v.y; im[2] += v.z; im[3] += v.w;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cf v = mk(re[r], im[r]);
            if (tk.base) v = v + ep.pbase[r];
            tile[4 * g + r][c] = v;
        }
    }
    __syncthreads();
    if (!own) return;
    const int64_t off = (int64_t)k * Rpad + j0 + cc;
    const cf v = tile[t][cc];
    store_stream(tk.out + off, v);
    if (!tk.Bn) return;
    const cf vn = (k >= 1) ? tile[t - 1][cc] : mk(0.f, 0.f);
    const cf vs = (k + 1 < npad) ? tile[t + 1][cc] : mk(0.f, 0.f);
    store_stream(tk.Bn + off, ep.next(v, vn, vs));
}

// s_waitcnt immediate that waits until at most n vector memory operations are outstanding
// (vmcnt in bits 3:0 and 15:14; expcnt and lgkmcnt left at their no-wait maxima)
constexpr unsigned vmcnt_wait(int n) { return (unsigned)((n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8)); }

// The step with A fetched straight into LDS by buffer_load_dwordx4 ... lds: the load unit writes
// lane l's 16 bytes at the instruction's LDS base + 16 l, so the choice of which 16 bytes of the
// page each lane fetches lays the k-block out in LDS -- no registers and no LDS write instructions
// on the way, and deeper prefetch costs LDS only.  A k-block of A (16 x 16 complex, 2 KiB) is 128
// slots of 16 bytes, filled by two load instructions (slots 0-63, 64-127):
//   forward (A = X rows i0 .. i0+15, column-major: a slot holds rows 2r, 2r+1 at one k): slot
//     32 s + 8 g + r holds rows 2r, 2r+1 at k = 4 g + s.  Each 8-lane group of a load reads one
//     whole 128-byte line (one k), and MFMA step s reads complex 64 s + 16 g + c, 512 contiguous
//     bytes per wave.
//   adjoint (A = X^H, so row i of A is column i0 + i of X, k contiguous: a slot holds k-pair p of
//     one row): slot 8 i + (p ^ (i & 7)) holds k-pair p of row i.  Each load reads 8 whole lines
//     (rows), and lane (g, c) reads pieces (c, 2g) and (c, 2g + 1): the slots that eight
//     consecutive lanes read sit in eight different 16-byte bank groups.
// B is loaded directly into the MFMA lanes (8 bytes per lane; its rows are contiguous along them),
// or with DMA_B into a stage of its own the same way as the forward A (a k-line of B holds 16
// columns: slot 32 s + 8 g + r holds columns 2r, 2r+1 at k = 4 g + s), so that a k-block
// takes four load instructions, all of whole lines.  The operands are the same values in the same
// k order as the other variants', so the results are bit-identical to theirs.  Nothing orders
// these LDS writes against the wave's own later LDS reads, so the loop waits on the vector memory
// counter itself: a k-block is 2 + 4 (DMA_B: 2 + 2) loads issued in a fixed order, and when k-block
// u is read the loads of u + 1 .. u + DEPTH (and, for u = 0, the epilogue operands) are the only
// ones issued after it.  Each wave's stages are its own (no
// barrier); after the loop they hold the cross-wave partial tiles and wave 0's the output tile.
template <int ALAY, int DEPTH, bool DMA_B = false>
__device__ __forceinline__ void sweep_fused_body_dma(const FusedParams &p) {
    constexpr int KS = 4, NST = DEPTH + 1, VM_PER_KB = DMA_B ? 4 : 6;
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int L = xcd * p.per_xcd + local;
    if (L >= p.units) return;
    const int npad = p.npad, Rpad = p.Rpad;
    const int Ng = Rpad / 16;
    const int ng = L % Ng;
    const int mb = (L / Ng) % p.strips;
    const int task = L / (Ng * p.strips);
    const FusedTask &tk = p.task[task];
    const int o0 = FUSED_OWN * mb;
    const int o1 = min(o0 + FUSED_OWN - 1, npad - 1);
    const int i0 = min(max(o0 - 1, 0), npad - 16);
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ks = wave;
    const int j0 = ng * 16;
    const int t = (threadIdx.x >> 4) & 15, cc = threadIdx.x & 15;
    const int k = i0 + t;
    const bool own = k >= o0 && k <= o1;

    // one LDS array per stage and operand, so that the compiler sees that reading one does not
    // depend on the loads in flight into the others
    __shared__ float4 stg0[4][128], stg1[4][128], stg2[NST > 2 ? 4 : 1][128], stg3[NST > 3 ? 4 : 1][128];
    __shared__ float4 stb0[DMA_B ? 4 : 1][128], stb1[DMA_B ? 4 : 1][128], stb2[DMA_B && NST > 2 ? 4 : 1][128],
        stb3[DMA_B && NST > 3 ? 4 : 1][128];
    auto stage = [&](int st) -> float4 * {
        switch (st) {
            case 0: return stg0[wave];
            case 1: return stg1[wave];
            case 2: return stg2[NST > 2 ? wave : 0];
            default: return stg3[NST > 3 ? wave : 0];
        }
    };
    auto bstage = [&](int st) -> float4 * {
        const int w = DMA_B ? wave : 0;
        switch (st) {
            case 0: return stb0[w];
            case 1: return stb1[w];
            case 2: return stb2[DMA_B && NST > 2 ? w : 0];
            default: return stb3[DMA_B && NST > 3 ? w : 0];
        }
    };
    const int64_t a_off = (ALAY == A_COL) ? i0 : (int64_t)i0 * npad;
    const auto ra = __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(tk.X + a_off), (short)0,
                                                      (int)(((int64_t)npad * npad - a_off) * 8), 0x00020000);
    const auto rbuf = __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(tk.B + j0), (short)0,
                                                        (int)(((int64_t)npad * Rpad - j0) * 8), 0x00020000);
    // lane part of the A loads' byte offsets (the instruction and the k-block add wave-uniform parts)
    const uint32_t va = (ALAY == A_COL)
        ? (uint32_t)(((4 * ((lane >> 3) & 3) + (lane >> 5)) * npad + 2 * (lane & 7)) * 8)
        : (uint32_t)(((lane >> 3) * npad + 2 * ((lane & 7) ^ (lane >> 3))) * 8);
    const uint32_t vbd = (uint32_t)((4 * g * Rpad + c) * 8);   // B[4 g + s][c]
    // DMA_B: B's k-block in slots 128-255 in the forward A's layout (a k-line of B is 16 columns)
    const uint32_t vbl = (uint32_t)(((4 * ((lane >> 3) & 3) + (lane >> 5)) * Rpad + 2 * (lane & 7)) * 8);
    typedef float v2f_t __attribute__((ext_vector_type(2)));
    auto gload = [&](v4f (&b)[2], int st, int k0) {
        float4 *dst = stage(st);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t sa = (ALAY == A_COL) ? (uint32_t)((k0 + 2 * q) * npad * 8) : (uint32_t)((8 * q * npad + k0) * 8);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (__attribute__((address_space(3))) void *)(dst + 64 * q), 16,
                                                     va, sa, 0, 0);
        }
        if (DMA_B) {
#pragma unroll
            for (int q = 0; q < 2; ++q)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rbuf, (__attribute__((address_space(3))) void *)(bstage(st) + 64 * q),
                                                         16, vbl, (uint32_t)((k0 + 2 * q) * Rpad * 8), 0, 0);
        } else {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const v2f_t u0 = __builtin_bit_cast(v2f_t, __builtin_amdgcn_raw_buffer_load_b64(rbuf, vbd, (uint32_t)((k0 + 2 * q) * Rpad * 8), 0));
                const v2f_t u1 = __builtin_bit_cast(v2f_t, __builtin_amdgcn_raw_buffer_load_b64(rbuf, vbd, (uint32_t)((k0 + 2 * q + 1) * Rpad * 8), 0));
                b[q] = v4f{u0.x, u0.y, u1.x, u1.y};
            }
        }
    };
    StepEpilogue ep(tk, npad, Rpad, i0, j0, k, cc, ks == 0);
    auto load_epilogue = [&] { ep.load(g, c, Rpad); };

    using Acc = CgemmWave<1, 1, ALAY, B_ROW, SWEEP_M3, DEPTH, SWEEP_FOLD>;
    Acc acc;
    acc.zero();
    {
        const int nkb = npad / 16;
        const int nk = (nkb - ks + KS - 1) / KS;
        const int last = ks + (nk - 1) * KS;
        auto k0 = [&](int u) {
            const int kb = ks + u * KS;
            return 16 * (kb < last ? kb : last);
        };
        v4f bst[NST][2];
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) gload(bst[u], u, k0(u));
        __builtin_amdgcn_sched_barrier(0);
        load_epilogue();
        __builtin_amdgcn_sched_barrier(0);
        for (int t0 = 0; t0 < nk; t0 += NST) {
#pragma unroll
            for (int u = 0; u < NST; ++u) {
                gload(bst[(u + DEPTH) % NST], (u + DEPTH) % NST, k0(t0 + u + DEPTH));
                __builtin_amdgcn_sched_barrier(0);
                if (t0 + u < nk) {
                    __builtin_amdgcn_s_waitcnt(vmcnt_wait(VM_PER_KB * DEPTH));   // k-block t0 + u is in LDS
                    __builtin_amdgcn_sched_barrier(0);
                    typename Acc::Ops o;
                    if (ALAY == A_COL) {
                        const cf *la = reinterpret_cast<const cf *>(stage(u));
#pragma unroll
                        for (int s = 0; s < 4; ++s) o.a[0][s] = la[64 * s + 16 * g + c];
                    } else {
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const float4 w = stage(u)[8 * c + ((2 * g + h) ^ (c & 7))];
                            o.a[0][2 * h] = mk(w.x, -w.y);
                            o.a[0][2 * h + 1] = mk(w.z, -w.w);
                        }
                    }
                    if (DMA_B) {
                        const cf *lb = reinterpret_cast<const cf *>(bstage(u));
#pragma unroll
                        for (int s = 0; s < 4; ++s) o.b[0][s] = lb[64 * s + 16 * g + c];
                    } else {
                        const v4f(&b)[2] = bst[u];
                        o.b[0][0] = mk(b[0].x, b[0].y);
                        o.b[0][1] = mk(b[0].z, b[0].w);
                        o.b[0][2] = mk(b[1].x, b[1].y);
                        o.b[0][3] = mk(b[1].z, b[1].w);
                    }
                    __builtin_amdgcn_s_setprio(1);   // MFMA phase first (see CgemmWave::run)
                    acc.mma(o);
                    __builtin_amdgcn_s_setprio(0);
                    acc.fold();
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the loads past the last k-block (re-reads of it) may still be writing this wave's stages
        __builtin_amdgcn_s_waitcnt(vmcnt_wait(0));
        __builtin_amdgcn_sched_barrier(0);
    }

    // partial tiles of waves 1-3 in their own stage-1 buffers; after the barrier (every wave past its
    // loop) the output tile in the stage-0 buffers
    float re[4], im[4];
    acc.result(0, 0, re, im);
    if (ks > 0) {
        stg1[wave][lane] = make_float4(re[0], re[1], re[2], re[3]);
        stg1[wave][64 + lane] = make_float4(im[0], im[1], im[2], im[3]);
    }
    __syncthreads();
    cf(*tile)[17] = reinterpret_cast<cf(*)[17]>(&stg0[0][0]);
    if (ks == 0) {
#pragma unroll
        for (int w = 1; w < KS; ++w) {
            const float4 u = stg1[w][lane], v = stg1[w][64 + lane];
            re[0] += u.x; re[1] += u.y; re[2] += u.z; re[3] += u.w;
            im[0] += v.x; im[1] += v.y; im[2] += v.z; im[3] += v.w;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cf v = mk(re[r], im[r]);
            if (tk.base) v = v + ep.pbase[r];
            tile[4 * g + r][c] = v;
        }
    }
    __syncthreads();
    if (!own) return;
    const int64_t off = (int64_t)k * Rpad + j0 + cc;
    const cf v = tile[t][cc];
    store_stream(tk.out + off, v);
    if (!tk.Bn) return;
    const cf vn = (k >= 1) ? tile[t - 1][cc] : mk(0.f, 0.f);
    const cf vs = (k + 1 < npad) ? tile[t + 1][cc] : mk(0.f, 0.f);
    store_stream(tk.Bn + off, ep.next(v, vn, vs));
}

template <int ALAY, int DEPTH, bool LATE, bool BUF>
__global__ void __launch_bounds__(256) sweep_fused_kernel(FusedParams p) {
    sweep_fused_body<ALAY, DEPTH, LATE, BUF>(p);
}
template <int ALAY, int DEPTH, bool LDS_B, bool MID = false>
__global__ void __launch_bounds__(256) sweep_fused_kernel_lds(FusedParams p) {
    sweep_fused_body_lds<ALAY, DEPTH, LDS_B, MID>(p);
}
template <int ALAY, int DEPTH, bool DMA_B>
__global__ void __launch_bounds__(256) sweep_fused_kernel_dma(FusedParams p) {
    sweep_fused_body_dma<ALAY, DEPTH, DMA_B>(p);
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

// Kernel variants of the fused step and their per-shape choice.  All take 16-row tiles with K dealt
// over the four waves of a workgroup in the same order, so they give bit-identical results; they
// differ in how the operands reach the MFMA:
//   0, 2  A and B fetched as whole lines, 16 bytes per lane, staged through a wave-private LDS tile
//         (DEPTH 1 / 2)
//   1     A the same way, B loaded directly into the MFMA lanes (its rows are contiguous along them)
//   3     buffer loads of both straight into the MFMA lanes, epilogue operands loaded after the GEMM
//   4     flat loads of both, DEPTH 2 (the reference variant)
//   5, 6  variants 0 and 1 with the epilogue operands loaded right after the first k-block's operands
//         (their latency hidden by the GEMM, without delaying its first MFMAs)
//   7     A loaded straight into LDS (buffer loads to LDS), B direct, epilogue operands early
//         (deeper prefetch, DEPTH 2 and 3, was never faster at 560^2, 717^2 or 1088^2)
//   8     A and B loaded straight into LDS, epilogue operands early
constexpr int N_FUSED_VARIANTS = 9;
constexpr int FUSED_DEPTH[N_FUSED_VARIANTS] = {1, 1, 2, 1, 2, 1, 1, 1, 1};
const char *const FUSED_KIND[N_FUSED_VARIANTS] = {"A and B staged through LDS", "A staged through LDS, B direct",
                                                  "A and B staged through LDS", "buffer loads", "flat loads",
                                                  "A and B staged through LDS, epilogue operands early",
                                                  "A staged through LDS, B direct, epilogue operands early",
                                                  "A loaded into LDS, B direct, epilogue operands early",
                                                  "A and B loaded into LDS, epilogue operands early"};

// strips of 16-row tiles over n rows (each owns FUSED_OWN rows), and the rows they cover
inline int fused_strips(int n) { return (n + FUSED_OWN - 1) / FUSED_OWN; }
inline int fused_covered(int n, int npad) { return std::min(fused_strips(n) * FUSED_OWN, npad); }

// Workgroup geometry for fp's shape: strips, units (tiles) and units per XCD slot.
void fused_geometry(FusedParams &fp) {
    fp.strips = fused_strips(fp.n);
    fp.units = fp.ntasks * fp.strips * (fp.Rpad / 16);
    fp.per_xcd = (fp.units + 7) / 8;
}

template <int ALAY>
void launch_fused_variant(int v, const FusedParams &fp, hipStream_t s, bool profiled) {
    const dim3 grid(8 * fp.per_xcd), block(256);
    auto go = [&](auto kernel) {
        if (profiled) launch_k(PC_SWEEP_GEMM, kernel, grid, block, 0, s, fp);
        else {
            hipLaunchKernelGGL(kernel, grid, block, 0, s, fp);
            WUST_LAUNCH_CHECK();
        }
    };
    switch (v) {
        case 0: go(sweep_fused_kernel_lds<ALAY, 1, true>); break;
        case 1: go(sweep_fused_kernel_lds<ALAY, 1, false>); break;
        case 2: go(sweep_fused_kernel_lds<ALAY, 2, true>); break;
        case 3: go(sweep_fused_kernel<ALAY, 1, true, true>); break;
        case 5: go(sweep_fused_kernel_lds<ALAY, 1, true, true>); break;
        case 6: go(sweep_fused_kernel_lds<ALAY, 1, false, true>); break;
        case 7: go(sweep_fused_kernel_dma<ALAY, 1, false>); break;
        case 8: go(sweep_fused_kernel_dma<ALAY, 1, true>); break;
        default: go(sweep_fused_kernel<ALAY, 2, false, false>); break;
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

// Buffer-resource variants address a page and an operand block with 32-bit byte offsets (and a 32-bit
// extent); larger blocks take the flat-load variant.
bool fused_eligible(int v, const FusedParams &fp) {
    if (v < 0 || v >= N_FUSED_VARIANTS || fp.npad < 16) return false;
    const int64_t page_bytes = (int64_t)fp.npad * fp.npad * 8, block_bytes = (int64_t)fp.npad * fp.Rpad * 8;
    return v == 4 || (page_bytes <= INT32_MAX && block_bytes <= INT32_MAX);
}

// Times every eligible variant on `fp` (a real step whose outputs the sweep rewrites before
// reading them) and returns the fastest; the choice is kept per (npad, Rpad, adjoint).  Within a
// burst, launch b reads the pages b steps further along each task's chain (task 0 walks up from its
// page, task 1 down, as the left and right eliminations do), so each launch meets its pages in HBM
// as the sweep does, not in the L2 that the previous launch filled.
// WUST_SWEEP_VARIANT=v forces variant v where it is eligible (read per solve, for tests).
int fused_variant(const FusedParams &fp, bool adjoint, hipStream_t s, const Operator &op) {
    if (const char *e = std::getenv("WUST_SWEEP_VARIANT"); e && e[0]) {
        const int v = std::atoi(e);
        if (v >= 0 && v < N_FUSED_VARIANTS && fused_eligible(v, fp)) return v;
    }
    static const bool tune = [] {
        const char *e = std::getenv("WUST_AUTOTUNE");
        return !(e && e[0] == '0');
    }();
    if (!tune) return fused_eligible(0, fp) ? 0 : 4;
    const FusedKey key{fp.npad, fp.Rpad, adjoint ? 1 : 0};
    {
        std::lock_guard<std::mutex> lk(g_fused_mu);
        auto it = g_fused_choice.find(key);
        if (it != g_fused_choice.end()) return it->second;
    }
    // Each eligible variant runs BURST back-to-back launches (as the sweep runs them) per round,
    // rounds interleaving the variants; a variant's time is its best round (the least disturbed).
    constexpr int ROUNDS = 3, BURST = 8;
    const int64_t page_elems = (int64_t)op.npad * op.npad;
    auto walked = [&](const FusedParams &base, int b) {
        FusedParams w = base;
        for (int i = 0; i < w.ntasks; ++i) {
            const int64_t j0 = (w.task[i].X - op.pages) / page_elems;
            const int64_t j = std::min<int64_t>(std::max<int64_t>(j0 + (i == 0 ? b : -b), 0), op.m - 1);
            w.task[i].X = op.page((int)j);
        }
        return w;
    };
    hipEvent_t e0, e1;
    WUST_HIP_CHECK(hipEventCreate(&e0));
    WUST_HIP_CHECK(hipEventCreate(&e1));
    FusedParams q[N_FUSED_VARIANTS];
    bool ok[N_FUSED_VARIANTS];
    float per_launch[N_FUSED_VARIANTS];
    auto run = [&](int v, int b) {
        const FusedParams w = walked(q[v], b);
        if (adjoint) launch_fused_variant<A_CONJT>(v, w, s, false);
        else launch_fused_variant<A_COL>(v, w, s, false);
    };
    for (int v = 0; v < N_FUSED_VARIANTS; ++v) {
        ok[v] = fused_eligible(v, fp);
        per_launch[v] = 1e30f;
        q[v] = fp;
        fused_geometry(q[v]);
        if (ok[v]) run(v, 0);    // warms the code object
    }
    for (int round = 0; round < ROUNDS; ++round)
        for (int v = 0; v < N_FUSED_VARIANTS; ++v) {
            if (!ok[v]) continue;
            WUST_HIP_CHECK(hipEventRecord(e0, s));
            for (int b = 0; b < BURST; ++b) run(v, round * BURST + b);
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
        std::fprintf(stderr, "sweep step kernel: npad %d Rpad %d %s -> variant %d (%s, depth %d, %.2f us)\n", fp.npad,
                     fp.Rpad, adjoint ? "adjoint" : "forward", best, FUSED_KIND[best], FUSED_DEPTH[best],
                     1e3f * per_launch[best]);
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
    const int covered = fused_covered(op.n, npad);
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
        fused_geometry(fp);
        if (variant < 0) variant = fused_variant(fp, adjoint, s, op);
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
