// Complex-FP32 GEMM core on MFMA for one wave (device code, header only).
//
// A wave accumulates an (16 RB) x (16 CB) block of  sum_k A[i][k] B[k][j]
// over a list of k-blocks of 16.  Inside a k-block, lane (g = lane / 16,
// c = lane % 16) contracts k = k0 + 4g + s at MFMA step s (s = 0..3), which
// makes every operand layout below load whole cache lines per 16 lanes, and
// the four k of a lane contiguous where the layout allows it.
//
// Complex products (v_mfma_f32_16x16x4f32, FP32 accumulation) take either
// four real products (M3 = false):  re = ar br - ai bi,  im = ar bi + ai br,
// or three (M3 = true, the 3M form):
//   p1 = ar br,  p2 = ai bi,  p3 = (ar + ai)(br + bi);   re = p1 - p2,  im = p3 - p1 - p2
// The 3M form saves a quarter of the MFMA work; its imaginary part carries
// the rounding of |ar + ai||br + bi|-sized terms.  Over a whole reduction in
// one FP32 accumulator that costs the block solves a refinement correction;
// with FOLD (below) the contraction stays where the four-product form has it.
//
// Operand layouts (lda / ldb in complex elements):
//   A_COL    A[i][k] = P[k lda + i]          column-major (inverse pages, Schur blocks)
//   A_CONJT  A[i][k] = conj(P[i lda + k])    conjugate transpose of a column-major P
//   B_KCOL   B[k][j] = Q[j ldb + k]          column-major (k contiguous)
//   B_ROW    B[k][j] = Q[k ldb + j]          row-major (fields: RHS fastest)
//
// The k loop keeps DEPTH k-blocks of operands in flight (DEPTH + 1 register
// sets rotated by unrolling, no register copies), so a wave waits only for
// data requested DEPTH k-blocks earlier.
//
// FOLD = true: the MFMA accumulators hold one k-block (16 products per
// element) at a time; after every k-block its complex result (the 3M
// combination re = p1 - p2, im = p3 - (p1 + p2) formed per k-block) is added
// into separate FP32 running sums.  One FP32 accumulation over the whole
// reduction (576 products at 560^2) loses accuracy in proportion to its
// length, and the block solves' contraction in iterative refinement is set by
// exactly that rounding; the two-level sum (16 products on MFMA, then the
// k-block partials) keeps it short.  FP64 running sums measured the same
// contraction (3.7e-6 forward, 5.8e-6 adjoint at 560^2 x 256 RHS) at more
// VALU work and registers.
#pragma once
#include "wust_common.h"

namespace wust {

typedef float v4f __attribute__((ext_vector_type(4)));

enum { A_COL = 0, A_CONJT = 1 };
enum { B_KCOL = 0, B_ROW = 1 };

template <int RB, int CB, int ALAY, int BLAY, bool M3, int DEPTH = 2, bool FOLD = false>
struct CgemmWave {
    // operands of one k-block: a[rb][s], b[cb][s]
    struct Ops {
        cf a[RB][4], b[CB][4];
    };
    v4f p1[RB][CB], p2[RB][CB], p3[RB][CB];   // p2 unused without M3
    float q1[FOLD ? RB : 1][FOLD ? CB : 1][4], q3[FOLD ? RB : 1][FOLD ? CB : 1][4];   // running sums (FOLD)

    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int x = 0; x < RB; ++x)
#pragma unroll
            for (int y = 0; y < CB; ++y) p1[x][y] = p2[x][y] = p3[x][y] = v4f{0.f, 0.f, 0.f, 0.f};
        if constexpr (FOLD) {
#pragma unroll
            for (int x = 0; x < RB; ++x)
#pragma unroll
                for (int y = 0; y < CB; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) q1[x][y][r] = q3[x][y][r] = 0.f;
        }
    }

    // FOLD: add the k-block held by the MFMA accumulators into the running sums and clear them
    __device__ __forceinline__ void fold() {
        if constexpr (FOLD) {
#pragma unroll
            for (int x = 0; x < RB; ++x)
#pragma unroll
                for (int y = 0; y < CB; ++y) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (M3) {
                            // the 3M combination of this k-block, then into the running sums
                            const float a = p1[x][y][r], b = p2[x][y][r], d = p3[x][y][r];
                            q1[x][y][r] += a - b;
                            q3[x][y][r] += d - (a + b);
                        } else {
                            q1[x][y][r] += p1[x][y][r];
                            q3[x][y][r] += p3[x][y][r];
                        }
                    }
                    p1[x][y] = p2[x][y] = p3[x][y] = v4f{0.f, 0.f, 0.f, 0.f};
                }
        }
    }

    // A points at A[i0][0] of the layout, B at B[0][j0].
    __device__ __forceinline__ static void load(Ops &o, const cf *A, int64_t lda, const cf *B, int64_t ldb, int k0,
                                                int g, int c) {
        const int k = k0 + 4 * g;
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const int i = 16 * rb + c;
            if (ALAY == A_COL) {
#pragma unroll
                for (int s = 0; s < 4; ++s) o.a[rb][s] = A[(int64_t)(k + s) * lda + i];
            } else {
                const float4 *p = reinterpret_cast<const float4 *>(A + (int64_t)i * lda + k);
                const float4 u0 = p[0], u1 = p[1];
                o.a[rb][0] = mk(u0.x, -u0.y);
                o.a[rb][1] = mk(u0.z, -u0.w);
                o.a[rb][2] = mk(u1.x, -u1.y);
                o.a[rb][3] = mk(u1.z, -u1.w);
            }
        }
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            const int j = 16 * cb + c;
            if (BLAY == B_KCOL) {
                const float4 *p = reinterpret_cast<const float4 *>(B + (int64_t)j * ldb + k);
                const float4 u0 = p[0], u1 = p[1];
                o.b[cb][0] = mk(u0.x, u0.y);
                o.b[cb][1] = mk(u0.z, u0.w);
                o.b[cb][2] = mk(u1.x, u1.y);
                o.b[cb][3] = mk(u1.z, u1.w);
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) o.b[cb][s] = B[(int64_t)(k + s) * ldb + j];
            }
        }
    }

    __device__ __forceinline__ void mma(const Ops &o) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (M3) {
                float as[RB], bs[CB];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) as[rb] = o.a[rb][s].re + o.a[rb][s].im;
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) bs[cb] = o.b[cb][s].re + o.b[cb][s].im;
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb) {
                        p1[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(o.a[rb][s].re, o.b[cb][s].re, p1[rb][cb], 0, 0, 0);
                        p2[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(o.a[rb][s].im, o.b[cb][s].im, p2[rb][cb], 0, 0, 0);
                        p3[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(as[rb], bs[cb], p3[rb][cb], 0, 0, 0);
                    }
            } else {
                // p1 accumulates the real part, p3 the imaginary part
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb) {
                        const cf x = o.a[rb][s], y = o.b[cb][s];
                        p1[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.re, y.re, p1[rb][cb], 0, 0, 0);
                        p3[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.re, y.im, p3[rb][cb], 0, 0, 0);
                        p1[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(-x.im, y.im, p1[rb][cb], 0, 0, 0);
                        p3[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.im, y.re, p3[rb][cb], 0, 0, 0);
                    }
            }
        }
    }

    // Accumulate over k-blocks kb = kb0, kb0 + step, ... < kb_end, keeping
    // DEPTH k-blocks of operands in flight (DEPTH + 1 register sets rotated by
    // unrolling).  Loads are issued unconditionally (indices past the last
    // k-block re-read the last one and are not used), so the wait counts
    // before each MFMA block are static; only the MFMA blocks are guarded, by
    // wave-uniform branches.  sched_barrier(0) keeps the scheduler from moving
    // loads across the MFMA phases, which would make the wait counts cover
    // later loads.  The MFMA phase runs at raised wave priority: a wave whose
    // operands have arrived issues its MFMAs ahead of the other waves' loads
    // and folds, so the waves of a SIMD drift apart in phase and keep the
    // matrix core busy instead of all waiting for it at once.
    __device__ __forceinline__ void run(const cf *A, int64_t lda, const cf *B, int64_t ldb, int kb0, int kb_end,
                                        int step, int g, int c) {
        const int nk = (kb_end - kb0 + step - 1) / step;
        if (nk <= 0) return;
        const int last = kb0 + (nk - 1) * step;
        auto k0 = [&](int t) {
            const int kb = kb0 + t * step;
            return 16 * (kb < last ? kb : last);
        };
        Ops o[DEPTH + 1];
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) load(o[u], A, lda, B, ldb, k0(u), g, c);
        for (int t = 0; t < nk; t += DEPTH + 1) {
#pragma unroll
            for (int u = 0; u <= DEPTH; ++u) {
                load(o[(u + DEPTH) % (DEPTH + 1)], A, lda, B, ldb, k0(t + u + DEPTH), g, c);
                __builtin_amdgcn_sched_barrier(0);
                if (t + u < nk) {
                    __builtin_amdgcn_s_setprio(1);
                    mma(o[u]);
                    __builtin_amdgcn_s_setprio(0);
                    fold();
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }

    // ---- the same k loop over buffer-resource loads ----
    //
    // Each operand is addressed as a raw buffer (base and byte extent in scalar registers).  A lane's
    // part of an address is a 32-bit byte offset that stays fixed over the loop, the k-block's part a
    // wave-uniform scalar offset, so the loop spends no vector instructions on addresses (the flat
    // loads above need a 64-bit multiply-add per load), and a read past the operand's extent returns
    // zero instead of touching memory.  kb0 and step must be wave-uniform.
    struct BufOps {
        __amdgpu_buffer_rsrc_t ra, rb;
        uint32_t va[RB], vb[CB];        // lane byte offsets
        uint32_t sa, sb;                // byte stride of one k for the layouts strided in k (A_COL, B_ROW)
    };
    __device__ __forceinline__ static BufOps buf_setup(const cf *A, int64_t lda, int64_t a_elems, const cf *B,
                                                       int64_t ldb, int64_t b_elems, int g, int c) {
        BufOps r;
        r.ra = __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(A), (short)0, (int)(a_elems * 8), 0x00020000);
        r.rb = __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(B), (short)0, (int)(b_elems * 8), 0x00020000);
        r.sa = (uint32_t)(lda * 8);
        r.sb = (uint32_t)(ldb * 8);
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const uint32_t i = 16 * rb + c;
            r.va[rb] = (ALAY == A_COL) ? (uint32_t)((4 * g * lda + i) * 8) : (uint32_t)((i * lda + 4 * g) * 8);
        }
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            const uint32_t j = 16 * cb + c;
            r.vb[cb] = (BLAY == B_ROW) ? (uint32_t)((4 * g * ldb + j) * 8) : (uint32_t)((j * ldb + 4 * g) * 8);
        }
        return r;
    }
    // (the loaded vector is reinterpreted whole: this compiler narrows the load to one dword when the
    // elements of the returned integer vector are extracted one by one)
    typedef float v2f_ __attribute__((ext_vector_type(2)));
    __device__ __forceinline__ static cf ld2(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        const v2f_ u = __builtin_bit_cast(v2f_, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
        return mk(u.x, u.y);
    }
    __device__ __forceinline__ static float4 ld4(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        const v4f u = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
        return make_float4(u.x, u.y, u.z, u.w);
    }
    __device__ __forceinline__ static void load_buf(Ops &o, const BufOps &q, int k0) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            if (ALAY == A_COL) {
#pragma unroll
                for (int s = 0; s < 4; ++s) o.a[rb][s] = ld2(q.ra, q.va[rb], (uint32_t)(k0 + s) * q.sa);
            } else {
                const float4 u0 = ld4(q.ra, q.va[rb], (uint32_t)k0 * 8), u1 = ld4(q.ra, q.va[rb], (uint32_t)k0 * 8 + 16);
                o.a[rb][0] = mk(u0.x, -u0.y);
                o.a[rb][1] = mk(u0.z, -u0.w);
                o.a[rb][2] = mk(u1.x, -u1.y);
                o.a[rb][3] = mk(u1.z, -u1.w);
            }
        }
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            if (BLAY == B_KCOL) {
                const float4 u0 = ld4(q.rb, q.vb[cb], (uint32_t)k0 * 8), u1 = ld4(q.rb, q.vb[cb], (uint32_t)k0 * 8 + 16);
                o.b[cb][0] = mk(u0.x, u0.y);
                o.b[cb][1] = mk(u0.z, u0.w);
                o.b[cb][2] = mk(u1.x, u1.y);
                o.b[cb][3] = mk(u1.z, u1.w);
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) o.b[cb][s] = ld2(q.rb, q.vb[cb], (uint32_t)(k0 + s) * q.sb);
            }
        }
    }
    // A / B as in run(); a_elems / b_elems are the complex elements readable from A / B.
    __device__ __forceinline__ void run_buf(const cf *A, int64_t lda, int64_t a_elems, const cf *B, int64_t ldb,
                                            int64_t b_elems, int kb0, int kb_end, int step, int g, int c) {
        const int nk = (kb_end - kb0 + step - 1) / step;
        if (nk <= 0) return;
        const BufOps q = buf_setup(A, lda, a_elems, B, ldb, b_elems, g, c);
        const int last = kb0 + (nk - 1) * step;
        auto k0 = [&](int t) {
            const int kb = kb0 + t * step;
            return 16 * (kb < last ? kb : last);
        };
        Ops o[DEPTH + 1];
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) load_buf(o[u], q, k0(u));
        for (int t = 0; t < nk; t += DEPTH + 1) {
#pragma unroll
            for (int u = 0; u <= DEPTH; ++u) {
                load_buf(o[(u + DEPTH) % (DEPTH + 1)], q, k0(t + u + DEPTH));
                __builtin_amdgcn_sched_barrier(0);
                if (t + u < nk) {
                    __builtin_amdgcn_s_setprio(1);   // MFMA phase first (see run())
                    mma(o[u]);
                    __builtin_amdgcn_s_setprio(0);
                    fold();
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }

    // Complex result of block (rb, cb): lane holds rows 4g + r (r = 0..3), column c.
    __device__ __forceinline__ void result(int rb, int cb, float (&re)[4], float (&im)[4]) const {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (FOLD) {
                re[r] = q1[rb][cb][r];
                im[r] = q3[rb][cb][r];
            } else if (M3) {
                re[r] = p1[rb][cb][r] - p2[rb][cb][r];
                im[r] = p3[rb][cb][r] - p1[rb][cb][r] - p2[rb][cb][r];
            } else {
                re[r] = p1[rb][cb][r];
                im[r] = p3[rb][cb][r];
            }
        }
    }
};

}  // namespace wust
