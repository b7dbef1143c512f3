// Block-tridiagonal factorization over grid columns (complex FP32).
//
// The reference (Runtime/solver/decompBlockLU.cu:7-20,337-368) runs one serial
// Schur recurrence T_{j+1} = D_{j+1} - L_{j+1} T_j^{-1} U_j over all columns and
// inverts each block with a pivoted getrf + getrs.  Here:
//   * the recurrence runs from both ends at once ("twisted" factorization):
//     a left chain T^L_j = D_j - L_j X^L_{j-1} U_{j-1} for j < p and a right chain
//     T^R_j = D_j - U_j X^R_{j+1} L_{j+1} for j > p, meeting at block p with
//     T_p = D_p - L_p X^L_{p-1} U_{p-1} - U_p X^R_{p+1} L_{p+1}; the serial depth
//     halves and every step carries two independent inversions;
//   * only the interior (n = Ny-2) rows of each column are kept, which is what
//     the reference's interior-only getrf/getrs computes;
//   * each block is inverted by blocked Gauss-Jordan elimination (one launch
//     per 32-column panel, ping-ponging between two buffers); the 32x32
//     diagonal-block inverses on the critical path use a 2x2 split into 16x16
//     single-wave register-resident inversions and MFMA products, and the WG
//     that produces the next panel's diagonal tile inverts it for the next
//     launch.
//     The Schur blocks of the nine-point PML operator are column-diagonally
//     dominant (partial pivoting selects the diagonal at every step), so the
//     elimination is pivot-free.  That assumption is checked for every block:
//     after its inversion, one MFMA product forms E = I - T_j X_j and its
//     Frobenius norm; a zero or non-finite pivot, or ||E||_F above the accepted
//     bound, redoes the whole factorization with pivoted inverses (rocSOLVER
//     getrf + getrs per block, the reference's method), and a block that
//     getrf finds exactly singular raises WUST_SINGULAR_SCHUR_BLOCK, as the
//     reference does from getrf's info (Runtime/solver/decompBlockLU.cu:339-342,
//     373-374).  The smallest |pivot| / max|entry below it| inside each 16x16
//     pivot block is also reported.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>

#include <rocsolver/rocsolver.h>

#include "cgemm.h"
#include "engine.h"
#include "plain.h"
#include "profile.h"

namespace wust {

namespace {

constexpr int NB = 32;  // Gauss-Jordan panel width == tile edge

struct SchurTerm {
    const cf *X;                 // inverse page (npad x npad, column-major)
    const cf *Pn, *Pc, *Ps;      // left tridiagonal factor, row coefficients
    const cf *Qn, *Qc, *Qs;      // right tridiagonal factor, row coefficients
};
constexpr int NS_KMAX = 8;  // Newton-Schulz rounds per block, upper limit

struct SchurTask {
    const cf *Dn, *Dc, *Ds;
    SchurTerm t[2];
    int nterms;
    cf *T;
    float *clear_norms;   // non-null: NS_KMAX Newton-Schulz norm slots zeroed for this block
};
struct SchurParams {
    SchurTask task[2];
    int ntasks, n, npad;
};

// T = D - sum_t P_t X_t Q_t with tridiagonal P, Q (three-point stencils on X).  A thread takes one
// row and SCHUR_COLS consecutive columns: the 3 x (SCHUR_COLS + 2) page entries around them are
// loaded once, from clamped in-range indices (a stencil point that leaves the block has a zero
// coefficient and adds an exact zero), all issued together, and shared by the columns' stencils; the
// column coefficients are the same across the wave.  A workgroup covers 64 rows (one per lane) and
// four waves of column runs.  At 576^2 x 2 chains, runs of two columns took 6.6-7.0 us per launch
// against 7.6 us for one column per thread in 256-row workgroups (the third row block of 576 then
// idles three of its four waves), 7.9 us for runs of one, ~6.9 us for three or four and 8.8 us for
// eight (fewer waves in flight).
constexpr int SCHUR_COLS = 2;
constexpr int SCHUR_ROWS = 64;
dim3 schur_grid(int ld, int ntasks) {
    return dim3((ld + SCHUR_ROWS - 1) / SCHUR_ROWS, (ld + 4 * SCHUR_COLS - 1) / (4 * SCHUR_COLS), ntasks);
}
__global__ void __launch_bounds__(256) schur_kernel(SchurParams p) {
    const SchurTask &tk = p.task[blockIdx.z];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = blockIdx.x * SCHUR_ROWS + (threadIdx.x & 63);
    const int c0 = (blockIdx.y * 4 + wave) * SCHUR_COLS;
    if (tk.clear_norms && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < NS_KMAX) tk.clear_norms[threadIdx.x] = 0.f;
    const int n = p.n, ld = p.npad;
    if (r >= ld || c0 >= ld) return;
    const cf zero = mk(0.f, 0.f);
    cf val[SCHUR_COLS];
#pragma unroll
    for (int i = 0; i < SCHUR_COLS; ++i) {
        const int c = c0 + i;
        val[i] = zero;
        if (r < n && c < n) {
            if (r == c) val[i] = tk.Dc[r];
            else if (c == r - 1) val[i] = tk.Dn[r];
            else if (c == r + 1) val[i] = tk.Ds[r];
        }
    }
    if (r < n) {
        const int rows[3] = {max(r - 1, 0), r, min(r + 1, n - 1)};
        auto term = [&](const SchurTerm &st) {
            const cf pr[3] = {(r >= 1) ? st.Pn[r] : zero, st.Pc[r], (r + 1 < n) ? st.Ps[r] : zero};
            cf x[3][SCHUR_COLS + 2];   // columns c0 - 1 .. c0 + SCHUR_COLS, clamped to the block
#pragma unroll
            for (int q = 0; q < SCHUR_COLS + 2; ++q) {
                const int64_t col = (int64_t)ld * min(max(c0 - 1 + q, 0), n - 1);
#pragma unroll
                for (int a = 0; a < 3; ++a) x[a][q] = st.X[rows[a] + col];
            }
#pragma unroll
            for (int i = 0; i < SCHUR_COLS; ++i) {
                const int c = c0 + i;
                if (c >= n) continue;
                const cf qm = (c >= 1) ? st.Qs[c - 1] : zero;
                const cf q0 = st.Qc[c];
                const cf qp = (c + 1 < n) ? st.Qn[c + 1] : zero;
                cf acc = zero;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    cf xq = q0 * x[a][i + 1];
                    cfma(xq, qm, x[a][i]);
                    cfma(xq, qp, x[a][i + 2]);
                    cfma(acc, pr[a], xq);
                }
                val[i] = val[i] - acc;
            }
        };
        if (tk.nterms >= 1) term(tk.t[0]);
        if (tk.nterms >= 2) term(tk.t[1]);
    }
#pragma unroll
    for (int i = 0; i < SCHUR_COLS; ++i) {
        const int c = c0 + i;
        if (c >= ld) break;
        // rows and columns beyond the block carry the identity
        tk.T[r + (int64_t)ld * c] = (r >= n || c >= n) ? ((r == c) ? mk(1.f, 0.f) : zero) : val[i];
    }
}

// X = I over the whole padded page (the pivoted path's getrs then overwrites the n x n block; the
// padding keeps the identity that every page carries)
__global__ void __launch_bounds__(256) page_identity_kernel(cf *X, int ld) {
    const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r < ld) X[r + (int64_t)ld * c] = (r == c) ? mk(1.f, 0.f) : mk(0.f, 0.f);
}
void launch_identity(cf *X, int ld, hipStream_t s) {
    hipLaunchKernelGGL(page_identity_kernel, dim3((ld + 255) / 256, ld), dim3(256), 0, s, X, ld);
    WUST_LAUNCH_CHECK();
}

struct GjTask {
    const cf *src;
    cf *dst;
    const cf *pin;   // inverse of this panel's diagonal block (32x32, ld 32), or null
    cf *pout;        // where the next panel's diagonal-block inverse goes, or null
};
struct GjParams {
    GjTask task[2];
    int ntasks, npad, q;
    uint32_t *status;
    float *pivot_ratio;   // per task: running min |pivot| / max |entry below it| in the diagonal block
};


struct GjShared {
    cf Pd[NB][NB + 1];     // P, row-major [r][c]
    cf PT[NB][NB + 1];     // P transposed [c][r]
    cf AikT[NB][NB + 1];   // A_IK transposed [l][i]; later W^T
    cf Akj[NB][NB + 1];    // A_KJ row-major [l][j]
};

// ---------------------------------------------------------------------------
// 32x32 inverse by 2x2 blocks of 16 (the serial chain of the factor):
//   Ai = A^-1;  E = C Ai;  F = Ai B;  S = D - E B;  Si = S^-1;
//   X = [[Ai - F X21, -F Si], [X21 = -Si E, Si]]
// Two 16-pivot Gauss-Jordan inversions, each by one wave with the block in
// registers, and 16x16 complex products on MFMA, each product on one wave.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_lds_sync() {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#endif
}

__device__ __forceinline__ float bits_f(int v) { return __int_as_float(v); }
__device__ __forceinline__ int f_bits(float v) { return __float_as_int(v); }

// In-place pivot-free Gauss-Jordan inverse of the row-major 16x16 block M (ld)
// by ONE wave with the block in registers: lane (r = lane & 15, cg = lane >> 4)
// holds a[e] = M[r][cg + 4e].  Per pivot k the three operands arrive without
// an LDS round trip or barrier:
//   pivot M[k][k]        readlane of lane (k, k & 3)                (scalar)
//   row k, my columns    DPP row_newbcast:k inside my 16-lane row
//   column k, my row     ds_bpermute from lane (r, k & 3)
// With `monitor`, reports zero / non-finite pivots and min |pivot| / max
// |entry below it in its column|; the monitor keeps per-lane state (the lanes
// on column k below the pivot compare |pivot|^2 / |own entry|^2 with their best
// pair by cross multiplication) and reduces once, after the last pivot, so the
// serial pivot chain carries no shuffles, divides or roots.
struct Inv16Mon {
    float best_num = 1.f, best_den = 1.f, pmin = 1.f, pchk = 0.f;
};

template <int K>
__device__ __forceinline__ void inv16_reg_pivot(cf (&a)[4], int r, int cg, bool monitor, Inv16Mon &mon) {
    constexpr int kc = K & 3, ke = K >> 2;
    const cf pk = mk(bits_f(__builtin_amdgcn_readlane(f_bits(a[ke].re), kc * 16 + K)),
                     bits_f(__builtin_amdgcn_readlane(f_bits(a[ke].im), kc * 16 + K)));
    const int src = (kc * 16 + r) * 4;
    const cf ck = mk(bits_f(__builtin_amdgcn_ds_bpermute(src, f_bits(a[ke].re))),
                     bits_f(__builtin_amdgcn_ds_bpermute(src, f_bits(a[ke].im))));
    cf rk[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
        rk[e] = mk(bits_f(__builtin_amdgcn_update_dpp(0, f_bits(a[e].re), 0x150 + K, 0xf, 0xf, false)),
                   bits_f(__builtin_amdgcn_update_dpp(0, f_bits(a[e].im), 0x150 + K, 0xf, 0xf, false)));
    if (monitor) {
        const float a2 = abs2(pk);
        mon.pmin = fminf(mon.pmin, a2);
        mon.pchk += a2 * 0.f;
        const float e2 = abs2(a[ke]);
        const bool cand = (cg == kc) && (r > K) && (a2 * mon.best_den < mon.best_num * e2);
        mon.best_num = cand ? a2 : mon.best_num;
        mon.best_den = cand ? e2 : mon.best_den;
    }
    const cf piv = crecip_fast(pk);
    const cf ckp = ck * piv;
    const bool rowk = (r == K);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const bool colk = (e == ke) && (cg == kc);
        const cf g = a[e] - ckp * rk[e];
        const cf onrow = colk ? piv : rk[e] * piv;
        const cf offrow = colk ? -ckp : g;
        a[e] = rowk ? onrow : offrow;
    }
}

template <int... K>
__device__ __forceinline__ void inv16_reg_pivots(cf (&a)[4], int r, int cg, bool monitor, Inv16Mon &mon,
                                                 std::integer_sequence<int, K...>) {
    (inv16_reg_pivot<K>(a, r, cg, monitor, mon), ...);
}

__device__ __forceinline__ void inv16_reg(cf *M, int ld, bool monitor, uint32_t *status, float *ratio_slot) {
    const int lane = threadIdx.x & 63, r = lane & 15, cg = lane >> 4;
    cf a[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) a[e] = M[r * ld + cg + 4 * e];
    Inv16Mon mon;
    inv16_reg_pivots(a, r, cg, monitor, mon, std::make_integer_sequence<int, 16>{});
#pragma unroll
    for (int e = 0; e < 4; ++e) M[r * ld + cg + 4 * e] = a[e];
    wave_lds_sync();
    if (monitor) {
        float ratio = (mon.best_den > 0.f) ? sqrtf(mon.best_num / mon.best_den) : 1.f;
        bool bad = !(mon.pmin > 0.f) || (mon.pchk != mon.pchk);
        for (int off = 32; off > 0; off >>= 1) ratio = fminf(ratio, __shfl_xor(ratio, off, 64));
        bad = __any(bad);
        if (lane == 0) {
            if (bad) atomicOr(status, (uint32_t)ST_ZERO_PIVOT);
            if (ratio < 1.0f) atomicMin(reinterpret_cast<int *>(ratio_slot), __float_as_int(ratio));
        }
    }
}

// C = sgn * (A B) + (Cin ? Cin : 0) for row-major 16x16 blocks (ld), one wave.
// Cin may alias C (each lane reads the elements it later writes).
__device__ __forceinline__ void gemm16_wave(const cf *A, const cf *B, const cf *Cin, cf *C, int ld, float sgn) {
    const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
    v4f dr = {0.f, 0.f, 0.f, 0.f}, di = dr;
#pragma unroll
    for (int ks = 0; ks < 16; ks += 4) {
        const cf av = A[li * ld + ks + kq];   // MFMA A operand: row li, k = ks + kq
        const cf bv = B[(ks + kq) * ld + li]; // MFMA B operand: k, column li
        dr = __builtin_amdgcn_mfma_f32_16x16x4f32(av.re, bv.re, dr, 0, 0, 0);
        di = __builtin_amdgcn_mfma_f32_16x16x4f32(av.re, bv.im, di, 0, 0, 0);
        dr = __builtin_amdgcn_mfma_f32_16x16x4f32(-av.im, bv.im, dr, 0, 0, 0);
        di = __builtin_amdgcn_mfma_f32_16x16x4f32(av.im, bv.re, di, 0, 0, 0);
    }
    // lane holds rows 4 kq + r, column li
    cf out[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        out[r] = mk(sgn * dr[r], sgn * di[r]);
        if (Cin) out[r] = out[r] + Cin[(4 * kq + r) * ld + li];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) C[(4 * kq + r) * ld + li] = out[r];
    wave_lds_sync();
}

// Invert the row-major 32x32 block M (ld 33) in place; T is a 32x33 scratch
// block.  Wave 0 runs both 16x16 inversions; waves 0 and 1 share the
// products.  All threads of the workgroup must call (barriers inside).
__device__ void inv32_blocked(cf (*M)[NB + 1], cf (*T)[NB + 1], bool monitor, uint32_t *status, float *ratio_slot) {
    const int wave = threadIdx.x >> 6;
    cf *A = &M[0][0], *B = &M[0][16], *C = &M[16][0], *D = &M[16][16];
    cf *E = &T[0][0], *F = &T[0][16];
    const int ld = NB + 1;
    if (wave == 0) inv16_reg(A, ld, monitor, status, ratio_slot);   // Ai
    lds_barrier();
    if (wave == 0) {
        gemm16_wave(C, A, nullptr, E, ld, 1.f);          // E = C Ai
        gemm16_wave(E, B, D, D, ld, -1.f);               // S = D - E B
        inv16_reg(D, ld, monitor, status, ratio_slot);   // Si
    } else if (wave == 1) {
        gemm16_wave(A, B, nullptr, F, ld, 1.f);          // F = Ai B
    }
    lds_barrier();
    if (wave == 0) gemm16_wave(D, E, nullptr, C, ld, -1.f);        // X21 = -Si E
    else if (wave == 1) gemm16_wave(F, D, nullptr, B, ld, -1.f);   // X12 = -F Si
    lds_barrier();
    if (wave == 0) gemm16_wave(F, C, A, A, ld, -1.f);              // X11 = Ai - F X21
    lds_barrier();
}

// D^T quadrant (16 x 16) of D = A B over a 32-deep reduction, computed as
// B^T A^T on MFMA so that lane (li, kq) ends up holding D[row0 + li][col0 + 4 kq + r]
// (contiguous rows for a fixed column: column-major stores coalesce).
//   Bt(k, j): element B[k][col0 + j]   (LDS, k-major)
//   At(k, i): element A[row0 + i][k]   (LDS, k-major)
template <typename FB, typename FA>
__device__ __forceinline__ void mfma_quadrant(v4f &dr, v4f &di, FB Bt, FA At, int li, int kq) {
#pragma unroll
    for (int ks = 0; ks < NB; ks += 4) {
        const cf a = Bt(ks + kq, li);   // "A" operand of the MFMA: row = column index j of D
        const cf b = At(ks + kq, li);   // "B" operand: column = row index i of D
        dr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.re, b.re, dr, 0, 0, 0);
        dr = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.im, b.im, dr, 0, 0, 0);
        di = __builtin_amdgcn_mfma_f32_16x16x4f32(a.re, b.im, di, 0, 0, 0);
        di = __builtin_amdgcn_mfma_f32_16x16x4f32(a.im, b.re, di, 0, 0, 0);
    }
}

// One Gauss-Jordan panel step on a 32x32 output tile.  With K the panel block
// and P = A_KK^{-1}:  A_IJ <- A_IJ - A_IK P A_KJ,  A_IK <- -A_IK P,
// A_KJ <- P A_KJ,  A_KK <- P.  Reads src, writes dst (no in-place hazards).
// P comes from the previous launch (pin) when available; otherwise every WG
// inverts A_KK itself.  The WG that produces the next panel's diagonal tile
// also inverts it and publishes the inverse (pout), so the serial pivot chain
// of the next panel runs in one WG instead of all of them.
__global__ void __launch_bounds__(256) gj_panel_kernel(GjParams p) {
    const int T = p.npad / NB;
    const int b = blockIdx.x;
    const int task = b / (T * T);
    const int rem = b - task * T * T;
    const int tj = rem / T, ti = rem - tj * T;
    const GjTask &tk = p.task[task];
    const int64_t ld = p.npad;
    const int q = p.q;
    const int t = threadIdx.x;
    const int i = t & 31, jb = t >> 5;           // element ownership: (i, jb + 8e)
    const int lane = t & 63, wave = t >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int qr = wave & 1, qc = wave >> 1;     // output quadrant of this wave
    __shared__ GjShared sh;

    const cf *src = tk.src;
    const bool rowK = (ti == q), colK = (tj == q);
    // issue every global load first; they complete while the pivots run
    cf a[4], aik[4], akj[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int j = jb + 8 * e;
        a[e] = tk.pin ? tk.pin[i + NB * j] : src[(int64_t)(q * NB + i) + ld * (q * NB + j)];
        aik[e] = rowK ? mk(0.f, 0.f) : src[(int64_t)(ti * NB + i) + ld * (q * NB + j)];
        akj[e] = colK ? mk(0.f, 0.f) : src[(int64_t)(q * NB + i) + ld * (tj * NB + j)];
    }
    cf aij[4];
    const bool general = !rowK && !colK;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = ti * NB + qr * 16 + li, col = tj * NB + qc * 16 + 4 * kq + r;
        aij[r] = general ? src[(int64_t)row + ld * col] : mk(0.f, 0.f);
    }
    if (!tk.pin) {
        // first panel: every WG inverts the diagonal block itself
#pragma unroll
        for (int e = 0; e < 4; ++e) sh.Pd[i][jb + 8 * e] = a[e];
        lds_barrier();
        inv32_blocked(sh.Pd, sh.PT, rowK && colK, p.status, p.pivot_ratio + task);
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = sh.Pd[i][jb + 8 * e];
    }

    // stage P and the strips for the MFMA products
    lds_barrier();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int j = jb + 8 * e;
        sh.Pd[i][j] = a[e];
        sh.PT[j][i] = a[e];
        sh.AikT[j][i] = aik[e];   // element (i, l = j) of A_IK
        sh.Akj[i][j] = akj[e];    // element (l = i, j) of A_KJ
    }
    lds_barrier();

    v4f dr = {0.f, 0.f, 0.f, 0.f}, di = dr;
    cf out[4];
    if (rowK && colK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) out[r] = sh.Pd[qr * 16 + li][qc * 16 + 4 * kq + r];
    } else if (rowK) {
        // P A_KJ
        mfma_quadrant(dr, di, [&](int k, int j) { return sh.Akj[k][qc * 16 + j]; },
                      [&](int k, int ii) { return sh.PT[k][qr * 16 + ii]; }, li, kq);
#pragma unroll
        for (int r = 0; r < 4; ++r) out[r] = mk(dr[r], di[r]);
    } else if (colK) {
        // -(A_IK P)
        mfma_quadrant(dr, di, [&](int k, int j) { return sh.Pd[k][qc * 16 + j]; },
                      [&](int k, int ii) { return sh.AikT[k][qr * 16 + ii]; }, li, kq);
#pragma unroll
        for (int r = 0; r < 4; ++r) out[r] = mk(-dr[r], -di[r]);
    } else {
        // W = A_IK P (quadrant), shared through LDS as W^T, then A_IJ - W A_KJ
        mfma_quadrant(dr, di, [&](int k, int j) { return sh.Pd[k][qc * 16 + j]; },
                      [&](int k, int ii) { return sh.AikT[k][qr * 16 + ii]; }, li, kq);
        lds_barrier();   // every wave is done reading AikT
#pragma unroll
        for (int r = 0; r < 4; ++r) sh.AikT[qc * 16 + 4 * kq + r][qr * 16 + li] = mk(dr[r], di[r]);
        lds_barrier();
        v4f er = {0.f, 0.f, 0.f, 0.f}, ei = er;
        mfma_quadrant(er, ei, [&](int k, int j) { return sh.Akj[k][qc * 16 + j]; },
                      [&](int k, int ii) { return sh.AikT[k][qr * 16 + ii]; }, li, kq);
#pragma unroll
        for (int r = 0; r < 4; ++r) out[r] = mk(aij[r].re - er[r], aij[r].im - ei[r]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = ti * NB + qr * 16 + li, col = tj * NB + qc * 16 + 4 * kq + r;
        tk.dst[(int64_t)row + ld * col] = out[r];
    }
    // the next panel's diagonal tile: invert it here and publish the inverse
    if (tk.pout && ti == q + 1 && tj == q + 1) {
        // the serial inversion is the launch's critical path: let its waves win
        // issue arbitration against the other WGs sharing this CU
        __builtin_amdgcn_s_setprio(3);
        lds_barrier();
#pragma unroll
        for (int r = 0; r < 4; ++r) sh.Pd[qr * 16 + li][qc * 16 + 4 * kq + r] = out[r];
        lds_barrier();
        inv32_blocked(sh.Pd, sh.PT, true, p.status, p.pivot_ratio + task);
#pragma unroll
        for (int e = 0; e < 4; ++e) tk.pout[i + NB * (jb + 8 * e)] = sh.Pd[i][jb + 8 * e];
        __builtin_amdgcn_s_setprio(0);
    }
}

// ---------------------------------------------------------------------------
// Newton-Schulz block inverse.  Consecutive Schur blocks of one chain differ
// little away from the PML (T_j X_{j-1} is close to I), so X_j is reached from
// X_0 = X_{j-1} on MFMA instead of by a serial pivot chain.  Round r of a
// block, with iterate X_r and residual E_r = I - T X_r:
//   ||E_r||_F <= q            last update, quadratic:  X_{r+1} = X_r (I + E_r)
//   q < ||E_r||_F <= c        last update, cubic:      X_{r+1} = X_r (I + E_r + E_r^2)
//   otherwise                 X_{r+1} = X_r (I + E_r), then E_{r+1} = I - T X_{r+1}
// (I - T X_{r+1} = E_r^2, resp. E_r^3, so the residual after the last update
// is bounded by ||E_r||_F^2, resp. ||E_r||_F^3; q and c are set below).  The
// last round of the host's budget is final
// whatever the norm.  Each round is three launches -- residual, cubic
// operand, update -- and each launch decides on the device, from the norms
// of earlier launches, whether it has work; launches after a block's last
// update return at once.  Every product takes the 3M form (three real
// products per complex product); the residual, which sets the accuracy floor,
// sums its k-blocks in two levels (CgemmWave FOLD).
// ---------------------------------------------------------------------------
enum NsKind { NS_RESIDUAL = 0, NS_CUBIC = 1, NS_UPDATE = 2 };

struct NsTask {
    const cf *T;      // Schur block
    const cf *X;      // iterate X_r
    cf *E;            // residual E_r
    cf *M;            // cubic operand E_r + E_r^2
    cf *page;         // X_{r+1} when this round is the last
    cf *xnext;        // X_{r+1} otherwise
    float *norms;     // ||E_r||_F^2 per round (cleared by the Schur launch)
    int *rounds;      // this block's slot of the rounds-used record
};
struct NsParams {
    NsTask task[2];
    int ntasks, npad;
    int kind, r, R;   // launch kind, round, round budget
    float q2, c2;     // q^2, c^2 (squared Frobenius norms)
    float accept;     // largest acceptable residual bound after the last update
    uint32_t *status;
    float *resid_max;  // running max of that bound over the factor (float bits, atomicMax)
};

// Tile order inside one task: the four quadrants of the T x T tile grid one
// after another, so that a contiguous run of tiles (one XCD's share) touches
// half of the rows of A and half of the columns of B.
__device__ __forceinline__ void ns_tile(int r, int T, int &ti, int &tj) {
    const int h0 = (T + 1) / 2, h1 = T / 2;
    const int rows[4] = {h0, h0, h1, h1}, cols[4] = {h0, h1, h0, h1};
    const int r0[4] = {0, 0, h0, h0}, c0[4] = {0, h0, 0, h0};
    int q = 0;
    while (q < 3 && r >= rows[q] * cols[q]) {
        r -= rows[q] * cols[q];
        ++q;
    }
    ti = r0[q] + r / cols[q];
    tj = c0[q] + r % cols[q];
}

// One 32x32 output tile of  out = C0 + sgn A B  per workgroup (C0 = I when
// null); the four waves split the reduction (k-blocks of 16 dealt
// round-robin) and each accumulates the whole tile as 2x2 blocks of 16x16
// (CgemmWave); the partial tiles are summed through LDS.  UPDATE: the cubic
// operand and update launches (else the residual launch).
template <bool UPDATE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) ns_gemm_kernel(NsParams p) {
    const int T = p.npad / NB, per_task = T * T, G = p.ntasks * per_task;
    const int per_xcd = (G + 7) / 8;
    const int L = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (L >= G) return;
    const int task = L / per_task;
    const NsTask &tk = p.task[task];
    // a block whose previous round was its last is finished
    if (p.r > 0 && tk.norms[p.r - 1] <= p.c2) return;
    const cf *A, *B, *C0;
    float sgn;
    cf *out;
    float *norm_out = nullptr;
    bool last = false, cubic = false;
    if (p.kind == NS_RESIDUAL) {
        A = tk.T; B = tk.X; C0 = nullptr; sgn = -1.f; out = tk.E; norm_out = tk.norms + p.r;
    } else {
        const float n = tk.norms[p.r];
        last = (n <= p.c2) || (p.r == p.R - 1);
        cubic = last && (n > p.q2);
        if (p.kind == NS_CUBIC) {
            if (!cubic) return;
            A = tk.E; B = tk.E; C0 = tk.E; sgn = 1.f; out = tk.M;
        } else {
            A = tk.X; B = cubic ? tk.M : tk.E; C0 = tk.X; sgn = 1.f; out = last ? tk.page : tk.xnext;
        }
    }
    int ti, tj;
    ns_tile(L % per_task, T, ti, tj);

    __shared__ float4 red[4][4][2][64];   // [wave][block][half][lane]
    __shared__ float wsum[4];
    const int ld = p.npad, i0 = ti * NB, j0 = tj * NB;
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform (scalar k offsets)
    // every product takes the 3M form; the residual E = I - T X folds each k-block into running sums
    // (its accuracy sets the accuracy of the inverse that the update X (I + E) produces), with one
    // k-block of operands in flight so that three workgroups per CU hold a 576^2 product's 648 tiles
    // in one round; the updates, without the sums, keep two in flight
    CgemmWave<2, 2, A_COL, B_KCOL, true, (UPDATE ? 2 : 1), !UPDATE> acc;
    acc.zero();
    const int64_t nn = (int64_t)ld * ld;
    acc.run_buf(A + i0, ld, nn - i0, B + (int64_t)j0 * ld, ld, nn - (int64_t)j0 * ld, wave, ld / 16, 4, g, c);

    // reduce the four k-partials: wave w finishes block (w / 2, w % 2)
#pragma unroll
    for (int blk = 0; blk < 4; ++blk) {
        float r[4], im[4];
        acc.result(blk >> 1, blk & 1, r, im);
        red[wave][blk][0][lane] = make_float4(r[0], r[1], r[2], r[3]);
        red[wave][blk][1][lane] = make_float4(im[0], im[1], im[2], im[3]);
    }
    lds_barrier();
    float sr[4] = {0.f, 0.f, 0.f, 0.f}, si[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const float4 u = red[w][wave][0][lane], v = red[w][wave][1][lane];
        sr[0] += u.x; sr[1] += u.y; sr[2] += u.z; sr[3] += u.w;
        si[0] += v.x; si[1] += v.y; si[2] += v.z; si[3] += v.w;
    }
    // lane holds rows i0 + 16 rb + 4 g + r (r = 0..3) of column j0 + 16 cb + c
    const int rb = wave >> 1, cb = wave & 1;
    const int row = i0 + 16 * rb + 4 * g, col = j0 + 16 * cb + c;
    const int64_t off = (int64_t)col * ld + row;
    cf v[4];
    if (C0) {
        const float4 *src = reinterpret_cast<const float4 *>(C0 + off);
        const float4 u0 = src[0], u1 = src[1];
        v[0] = mk(u0.x, u0.y);
        v[1] = mk(u0.z, u0.w);
        v[2] = mk(u1.x, u1.y);
        v[3] = mk(u1.z, u1.w);
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = mk(row + r == col ? 1.f : 0.f, 0.f);
    }
    float n2 = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        v[r] = mk(v[r].re + sgn * sr[r], v[r].im + sgn * si[r]);
        n2 += abs2(v[r]);
    }
    float4 *dst = reinterpret_cast<float4 *>(out + off);
    dst[0] = make_float4(v[0].re, v[0].im, v[1].re, v[1].im);
    dst[1] = make_float4(v[2].re, v[2].im, v[3].re, v[3].im);

    if (norm_out) {
        for (int o = 32; o > 0; o >>= 1) n2 += __shfl_xor(n2, o, 64);
        if (lane == 0) wsum[wave] = n2;
        lds_barrier();
        if (threadIdx.x == 0) atomicAdd(norm_out, wsum[0] + wsum[1] + wsum[2] + wsum[3]);
    }
    if (p.kind == NS_UPDATE && last && L % per_task == 0 && threadIdx.x == 0) {
        // residual bound after this update: ||E_r||^2 (quadratic) or ||E_r||^3 (cubic)
        const float n = tk.norms[p.r];
        const float bound = cubic ? n * sqrtf(n) : n;
        if (!(bound <= p.accept)) atomicOr(p.status, (uint32_t)ST_INVERSE_UNCONVERGED);
        if (bound >= 0.f) atomicMax(reinterpret_cast<int *>(p.resid_max), __float_as_int(bound));
        // rounds this block used; + 64 when the budget ran out before ||E||^2 reached c^2
        // (a non-finite norm counts as ran out, so that a diverged block reaches Gauss-Jordan next time)
        *tk.rounds = (p.r + 1) + (!(n <= p.c2) ? 64 : 0);
    }
}

struct ChainStep {
    int block;      // block index j to produce
    int kind;       // 0: left chain, 1: right chain, 2: middle
};

void add_term(const Operator &op, SchurTerm &t, const cf *X, int pblock, int pdirs[3], int qblock,
              int qdirs[3]) {
    t.X = X;
    t.Pn = op.coef(pdirs[0], pblock);
    t.Pc = op.coef(pdirs[1], pblock);
    t.Ps = op.coef(pdirs[2], pblock);
    t.Qn = op.coef(qdirs[0], qblock);
    t.Qc = op.coef(qdirs[1], qblock);
    t.Qs = op.coef(qdirs[2], qblock);
}

}  // namespace

void launch_schur_left(const Operator &op, int j, cf *T, hipStream_t s) {
    int Ldirs[3] = {DNW, DW, DSW};
    int Udirs[3] = {DNE, DE, DSE};
    SchurParams sp{};
    sp.ntasks = 1;
    sp.n = op.n;
    sp.npad = op.npad;
    SchurTask &tk = sp.task[0];
    tk.Dn = op.coef(DN, j);
    tk.Dc = op.coef(DC, j);
    tk.Ds = op.coef(DS, j);
    tk.nterms = 0;
    tk.T = T;
    if (j >= 1) add_term(op, tk.t[tk.nterms++], op.page(j - 1), j, Ldirs, j - 1, Udirs);
    launch_k(PC_SCHUR, schur_kernel, schur_grid(op.npad, 1), dim3(256), 0, s, sp);
}

namespace {

// Newton-Schulz controls.  The residual after the last update of a block is
// E^2 (quadratic update) or E^3 (cubic), whose Frobenius norms are bounded by
// ||E||_F^2 and ||E||_F^3.  The quadratic update is taken when ||E||_F^2 <=
// 2.5e-5: the bound overstates ||E^2||_F by about ten on these blocks (E's
// energy is spread over many singular values; at 560^2 the interior blocks
// have ||E||_F^2 = 1.9-2.3e-5 and ||E^2||_F = 2.0-2.3e-6), which leaves them at
// the FP32 floor of the inverses (~2e-6) with one GEMM fewer than the cubic
// update, and the refined solves' contraction where the cubic update has it
// (3.9e-6 against 3.7e-6 forward).  The cubic update is taken when ||E||_F^3
// <= 2e-6.  A block whose bound after its last update exceeds 1e-4 sends the
// whole factor back to Gauss-Jordan.
constexpr float NS_Q2 = 2.5e-5f;
constexpr float NS_C2 = 1.5874e-4f;   // (2e-6)^(2/3)
constexpr float NS_ACCEPT = 1e-4f;
// Gauss-Jordan blocks (pivot-free) are accepted when ||I - T_j X_j||_F stays below this.  The FP32
// inverses of these blocks measure up to 1.4e-4 (560^2 metric case); an elimination that loses
// column dominance grows its entries and leaves residuals of order one.  A block between the two
// still refines (the FP64 refinement reports NotConverged if it does not).  WUST_GJ_ACCEPT
// overrides the bound (tests force the pivoted path with it).
constexpr float GJ_ACCEPT = 1e-2f;

// Newton-Schulz round budgets learned per operator shape and frequency: the rounds each
// chain step needed in the last factorization of that key (both chains' maximum; one more
// where a block's budget ran out first).  The next factorization with that key uses exactly
// that budget in the interior columns, where one cubic round suffices with a wide margin,
// and one spare round in the PML columns, whose convergence depends more on the model.  A
// budget that still proves too small costs accuracy within the accepted bound or, beyond
// it, the Gauss-Jordan fallback of that factorization; the key then keeps that
// factorization's profile, in which the steps that ran out of rounds need one more than
// they had, so that they reach Gauss-Jordan in the next factorization.
struct BudgetKey {
    int ny, nx, npad, dispersion;
    double f, dx, dy, a0, L;
    bool operator<(const BudgetKey &o) const {
        return std::tie(ny, nx, npad, dispersion, f, dx, dy, a0, L) <
               std::tie(o.ny, o.nx, o.npad, o.dispersion, o.f, o.dx, o.dy, o.a0, o.L);
    }
};
std::mutex g_budget_mu;
std::map<BudgetKey, std::vector<uint8_t>> g_budgets;

// Inverse pages of the leading chain steps of the last factorization per (device, operator key).
// Those steps (the PML columns at the start of each chain) have no usable neighbour seed and go to
// Gauss-Jordan; their Schur blocks depend only on the model over the columns from the chain's end
// up to the step, which consecutive evaluations of a reconstruction leave unchanged (the update
// mask keeps the PML fixed) or change little.  The next factorization with the key seeds
// Newton-Schulz for those steps with the cached page of the same block instead (two rounds at most;
// a seed that no longer converges sends that factorization to the Gauss-Jordan fallback and drops
// the entry).  An entry is held by one factorization at a time; the cache keeps at most
// WUST_SEED_CACHE_MB megabytes (default 4096), least recently used entries first out.
struct SeedEntry {
    cf *buf = nullptr;        // [steps pages: blocks 0 .. steps-1][steps pages: blocks m-steps .. m-1]
    int steps = 0;            // leading chain steps cached per chain
    size_t bytes = 0;
    bool busy = false;
    uint64_t used = 0;        // last use (cache clock)
};
std::mutex g_seed_mu;
std::map<std::pair<int, BudgetKey>, SeedEntry> g_seeds;
uint64_t g_seed_clock = 0;
size_t seed_cache_cap() {
    static const size_t cap = [] {
        const char *e = std::getenv("WUST_SEED_CACHE_MB");
        const long long mb = e ? std::atoll(e) : 4096;
        return (size_t)(mb > 0 ? mb : 0) << 20;
    }();
    return cap;
}
int current_device_id() {
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}
// releases a held entry on every exit path
struct SeedHold {
    SeedEntry *e = nullptr;
    ~SeedHold() {
        if (e) {
            std::lock_guard<std::mutex> lk(g_seed_mu);
            e->busy = false;
        }
    }
};
void seed_free(SeedEntry &e) {
    if (e.buf) (void)hipFree(e.buf);
    e.buf = nullptr;
    e.steps = 0;
    e.bytes = 0;
}

BudgetKey budget_key(const Operator &op) {
    const AssemblyParams &a = op.ap;
    return BudgetKey{a.ny, a.nx, a.npad, a.dispersion, a.f, a.dx, a.dy, a.a0, a.L};
}

struct FactorRun {
    Operator &op;
    hipStream_t s;
    bool allow_ns;
    int pml_cells;
    PlainLib *pivoted = nullptr;    // non-null: every block by rocSOLVER getrf + getrs (partial pivoting)
    int ns_blocks = 0;
    int update_cap = NS_KMAX;
    std::vector<uint8_t> learned;   // per chain step; empty: the default budget
    const cf *seeds = nullptr;      // cached pages of the leading steps (SeedEntry layout), or null
    int seed_steps = 0;
    int seeded_upto = 0;            // leading steps whose blocks came from the cached seeds
    bool seeded(int st) const { return seeds && st < seed_steps; }
    // left chain step st is block st, right chain step st is block m-1-st (both halves in block order)
    const cf *seed_page(int chain, int st) const {
        const int64_t i = chain == 0 ? st : 2 * (int64_t)seed_steps - 1 - st;
        return seeds + i * op.npad * op.npad;
    }
    // Newton-Schulz round budget of chain step `st`: the PML columns at the
    // start of each chain change fastest.
    // Newton-Schulz or Gauss-Jordan for chain step `st`.  A Gauss-Jordan block (npad / 32 panel
    // launches) costs more than 3 Newton-Schulz rounds (three GEMM launches each) at npad 576
    // (0.31 against 0.26 ms for both chains) and less at npad 1088 (0.84 against 1.08 ms), the
    // GEMMs growing as npad^3 and the latency-bound panels much more slowly.  So steps that needed
    // 4 or more rounds (3 or more from npad 800 on), or whose need is unknown, use Gauss-Jordan;
    // without a learned profile the second column of each chain (seeded from the first, the
    // slowest to converge: 7-9 rounds at 6 points per wavelength) does too.
    bool ns_step(int st) const {
        if (learned.empty()) return st >= 2;
        const int most = op.npad >= 800 ? 2 : 3;
        return st < (int)learned.size() && learned[st] >= 1 && learned[st] <= most;
    }

    int ns_budget(int st) const {
        const bool pml = st < pml_cells + 8;
        int b = pml ? NS_KMAX : 2;
        if (st < (int)learned.size() && learned[st] >= 1) b = std::min(learned[st] + (pml ? 1 : 0), NS_KMAX);
        return b < update_cap ? b : update_cap;
    }

    void run_steps(const std::vector<ChainStep> &steps, int st) {
        const int m = op.m, n = op.n, ld = op.npad;
        const int64_t pg = (int64_t)ld * ld;
        int Ldirs[3] = {DNW, DW, DSW};
        int Udirs[3] = {DNE, DE, DSE};
        float *ratio_d = reinterpret_cast<float *>(op.status + 2);
        std::vector<int> gj, ns;
        for (int c = 0; c < (int)steps.size(); ++c) {
            const bool use_ns =
                allow_ns && op.nswork && steps[c].kind != 2 && (seeded(st) || (st > 0 && ns_step(st)));
            (use_ns ? ns : gj).push_back(c);
        }
        SchurParams sp{};
        sp.ntasks = (int)steps.size();
        sp.n = n;
        sp.npad = ld;
        for (int c = 0; c < sp.ntasks; ++c) {
            int j = steps[c].block;
            SchurTask &tk = sp.task[c];
            tk.Dn = op.coef(DN, j);
            tk.Dc = op.coef(DC, j);
            tk.Ds = op.coef(DS, j);
            tk.nterms = 0;
            tk.T = op.work + (int64_t)c * 3 * pg;   // chain c's [T | ping | pong]
            tk.clear_norms = (allow_ns && op.ns_norm) ? op.ns_norm + c * NS_KMAX : nullptr;
            int kind = steps[c].kind;
            if ((kind == 0 || kind == 2) && j >= 1)  // - L_j X^L_{j-1} U_{j-1}
                add_term(op, tk.t[tk.nterms++], op.page(j - 1), j, Ldirs, j - 1, Udirs);
            if ((kind == 1 || kind == 2) && j + 1 <= m - 1)  // - U_j X^R_{j+1} L_{j+1}
                add_term(op, tk.t[tk.nterms++], op.page(j + 1), j, Udirs, j + 1, Ldirs);
        }
        launch_k(PC_SCHUR, schur_kernel, schur_grid(ld, sp.ntasks), dim3(256), 0, s, sp);

        if (pivoted) {
            // the reference's inversion: partial-pivot LU of T, then X = T^-1 by getrs against I
            rocblas_handle hb = (rocblas_handle)pivoted->handle;
            for (int c = 0; c < sp.ntasks; ++c) {
                const int j = steps[c].block;
                cf *T = sp.task[c].T;
                cf *X = op.page(j);
                if (rocsolver_cgetrf(hb, n, n, (rocblas_float_complex *)T, ld, pivoted->ipiv,
                                     pivoted->info + 2 * (j % 65536)) != rocblas_status_success)
                    throw Error(WUST_LIBRARY_ERROR, "rocsolver_cgetrf failed");
                launch_identity(X, ld, s);
                if (rocsolver_cgetrs(hb, rocblas_operation_none, n, n, (rocblas_float_complex *)T, ld, pivoted->ipiv,
                                     (rocblas_float_complex *)X, ld) != rocblas_status_success)
                    throw Error(WUST_LIBRARY_ERROR, "rocsolver_cgetrs failed");
            }
            return;
        }

        if (!gj.empty()) {
            const int Q = ld / NB;
            for (int q = 0; q < Q; ++q) {
                GjParams gp{};
                gp.ntasks = (int)gj.size();
                gp.npad = ld;
                gp.q = q;
                gp.status = op.status;
                gp.pivot_ratio = ratio_d;
                for (int t = 0; t < gp.ntasks; ++t) {
                    const int c = gj[t];
                    // T stays intact for the residual check: the panels ping-pong between the
                    // chain's two other buffers
                    cf *T = op.work + (int64_t)c * 3 * pg, *ping = T + pg, *pong = T + 2 * pg;
                    gp.task[t].src = (q == 0) ? T : ((q % 2 == 1) ? ping : pong);
                    gp.task[t].dst = (q == Q - 1) ? op.page(steps[c].block) : ((q % 2 == 0) ? ping : pong);
                    cf *pb = op.pbuf + (int64_t)c * 2 * NB * NB;
                    gp.task[t].pin = (q == 0) ? nullptr : pb + (q % 2) * NB * NB;
                    gp.task[t].pout = (q + 1 < Q) ? pb + ((q + 1) % 2) * NB * NB : nullptr;
                }
                launch_k(PC_GJ_PANEL, gj_panel_kernel, dim3(gp.ntasks * Q * Q), dim3(256), 0, s, gp);
            }
            // the pivot-free elimination's check: E = I - T X_j and ||E||_F^2 into the block's slot
            NsParams cp{};
            cp.ntasks = (int)gj.size();
            cp.npad = ld;
            cp.kind = NS_RESIDUAL;
            cp.r = 0;
            cp.R = 1;
            for (int t = 0; t < cp.ntasks; ++t) {
                const int c = gj[t];
                NsTask &tk = cp.task[t];
                cf *Tc = op.work + (int64_t)c * 3 * pg;
                tk.T = Tc;
                tk.X = op.page(steps[c].block);
                tk.E = Tc + 2 * pg;
                tk.norms = op.gj_check + steps[c].block;
            }
            const int Tt = ld / NB, G = cp.ntasks * Tt * Tt;
            launch_k(PC_NS_GEMM, ns_gemm_kernel<false>, dim3(8 * ((G + 7) / 8)), dim3(256), 0, s, cp);
        }
        if (!ns.empty()) {
            const int R = seeded(st) ? std::min(2, update_cap) : ns_budget(st);
            const int T = ld / NB;
            const int G = (int)ns.size() * T * T;
            const int grid_ns = 8 * ((G + 7) / 8);
            NsParams np{};
            np.ntasks = (int)ns.size();
            np.npad = ld;
            np.R = R;
            np.q2 = NS_Q2;
            np.c2 = NS_C2;
            np.accept = NS_ACCEPT;
            np.status = op.status;
            np.resid_max = reinterpret_cast<float *>(op.status + 4);
            auto set_round = [&](int r) {
                np.r = r;
                for (int t = 0; t < (int)ns.size(); ++t) {
                    const int c = ns[t];
                    const int j = steps[c].block;
                    NsTask &tk = np.task[t];
                    cf *w = op.work + (int64_t)c * 3 * pg;       // [T | E | -]
                    cf *xb = op.nswork + (int64_t)c * 3 * pg;    // [X_a | X_b | M]
                    tk.T = w;
                    tk.E = w + pg;
                    tk.M = xb + 2 * pg;
                    const cf *seed = seeded(st) ? seed_page(steps[c].kind, st)
                                                : op.page(steps[c].kind == 0 ? j - 1 : j + 1);
                    tk.X = (r == 0) ? seed : xb + ((r - 1) % 2) * pg;
                    tk.xnext = xb + (r % 2) * pg;
                    tk.page = op.page(j);
                    tk.norms = op.ns_norm + c * NS_KMAX;
                    tk.rounds = op.ns_rounds + j;
                }
            };
            set_round(0);
            np.kind = NS_RESIDUAL;
            launch_k(PC_NS_GEMM, ns_gemm_kernel<false>, dim3(grid_ns), dim3(256), 0, s, np);
            for (int r = 0; r < R; ++r) {
                set_round(r);
                np.kind = NS_CUBIC;
                launch_k(PC_NS_GEMM, ns_gemm_kernel<true>, dim3(grid_ns), dim3(256), 0, s, np);
                np.kind = NS_UPDATE;
                launch_k(PC_NS_GEMM, ns_gemm_kernel<true>, dim3(grid_ns), dim3(256), 0, s, np);
                if (r + 1 < R) {
                    set_round(r + 1);
                    np.kind = NS_RESIDUAL;
                    launch_k(PC_NS_GEMM, ns_gemm_kernel<false>, dim3(grid_ns), dim3(256), 0, s, np);
                }
            }
            ns_blocks += (int)ns.size();
        }
    }

    // After the seeded steps: a block whose Newton-Schulz iteration ran out of rounds from its cached
    // seed (the model changed where the seed came from) hands the chain steps from the first such step
    // on to the unseeded policy (their Schur blocks depend on it), so that a stale cache costs those
    // steps, not the whole factor.  Returns the step to resume from (seed_steps when every seeded
    // block converged).
    int check_seeds() {
        const int m = op.m;
        std::vector<int> used((size_t)m);
        WUST_HIP_CHECK(hipMemcpyAsync(used.data(), op.ns_rounds, sizeof(int) * m, hipMemcpyDeviceToHost, s));
        WUST_HIP_CHECK(hipStreamSynchronize(s));
        int bad = seed_steps;
        for (int st = 0; st < seed_steps && bad == seed_steps; ++st)
            if ((used[st] & 64) || (used[m - 1 - st] & 64) || used[st] == 0 || used[m - 1 - st] == 0) bad = st;
        if (bad == seed_steps) return bad;
        // only seeded Newton-Schulz steps ran so far: clear the flags and the largest bound that the
        // failed ones left, and the rounds records of the steps to redo
        const uint32_t init[6] = {0u, 0u, 0x3f800000u, 0x3f800000u, 0u, 0u};
        WUST_HIP_CHECK(hipMemcpyAsync(op.status, init, sizeof(init), hipMemcpyHostToDevice, s));
        const int redo = seed_steps - bad;
        WUST_HIP_CHECK(hipMemsetAsync(op.ns_rounds + bad, 0, sizeof(int) * redo, s));
        WUST_HIP_CHECK(hipMemsetAsync(op.ns_rounds + (m - seed_steps), 0, sizeof(int) * redo, s));
        ns_blocks -= 2 * redo;
        seeds = nullptr;
        seeded_upto = bad;
        if (const char *e = std::getenv("WUST_NS_LOG"); e && e[0] == '1')
            std::fprintf(stderr, "newton-schulz: the cached seed of chain step %d did not converge; steps %d..%d again "
                                 "without seeds\n", bad, bad, seed_steps - 1);
        return bad;
    }

    void run() {
        const int m = op.m, p = op.p;
        // chains: left j = 0..p-1, right j = m-1..p+1, then the middle block p
        int nl = p, nr = m - 1 - p;
        int nsteps = nl > nr ? nl : nr;
        for (int st = 0; st < nsteps; ++st) {
            std::vector<ChainStep> steps;
            if (st < nl) steps.push_back({st, 0});
            if (st < nr) steps.push_back({m - 1 - st, 1});
            run_steps(steps, st);
            if (seeds && st == seed_steps - 1) st = check_seeds() - 1;
        }
        run_steps({{p, 2}}, 0);
    }
};

// After a factorization: keep the pages of the leading steps that the next factorization with this key
// would give to Gauss-Jordan (under `profile`) or that it seeded (`seeded` steps; -1 after a fallback,
// which drops the entry), evicting least recently used entries beyond the cache's size.
void update_seed_cache(const Operator &op, SeedEntry &e, const std::vector<uint8_t> &profile, int seeded,
                       hipStream_t s) {
    if (seeded < 0 || profile.empty()) {
        std::lock_guard<std::mutex> lk(g_seed_mu);
        seed_free(e);
        return;
    }
    const int most = op.npad >= 800 ? 2 : 3;
    const int nsteps = std::min(op.p, op.m - 1 - op.p);
    int lead = 0;
    while (lead < nsteps && (lead == 0 || lead < seeded || lead >= (int)profile.size() || profile[lead] < 1 ||
                             profile[lead] > most))
        ++lead;
    const int64_t pg = (int64_t)op.npad * op.npad;
    const size_t bytes = sizeof(cf) * (size_t)(2 * lead) * (size_t)pg;
    {
        std::lock_guard<std::mutex> lk(g_seed_mu);
        if (e.bytes != bytes) seed_free(e);
        // room for this entry: drop other idle entries, least recently used first
        size_t total = bytes;
        for (auto &kv : g_seeds)
            if (&kv.second != &e) total += kv.second.bytes;
        while (total > seed_cache_cap()) {
            SeedEntry *lru = nullptr;
            for (auto &kv : g_seeds)
                if (&kv.second != &e && !kv.second.busy && kv.second.bytes && (!lru || kv.second.used < lru->used))
                    lru = &kv.second;
            if (!lru) break;
            total -= lru->bytes;
            seed_free(*lru);
        }
        if (total > seed_cache_cap() || lead == 0) return;
    }
    if (!e.buf) {
        if (hipMalloc(&e.buf, bytes) != hipSuccess) {
            (void)hipGetLastError();
            e.buf = nullptr;
            return;   // no room: factor without seeds next time
        }
        e.bytes = bytes;
    }
    // the leading pages of each chain are contiguous: blocks 0 .. lead-1 and m-lead .. m-1
    WUST_HIP_CHECK(hipMemcpyAsync(e.buf, op.page(0), sizeof(cf) * pg * lead, hipMemcpyDeviceToDevice, s));
    WUST_HIP_CHECK(hipMemcpyAsync(e.buf + (int64_t)lead * pg, op.page(op.m - lead), sizeof(cf) * pg * lead,
                                  hipMemcpyDeviceToDevice, s));
    // the entry is released when this returns: the copies land before another factorization reads it
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    e.steps = lead;
}

}  // namespace

void seed_cache_trim() {
    std::lock_guard<std::mutex> lk(g_seed_mu);
    for (auto it = g_seeds.begin(); it != g_seeds.end();) {
        if (it->second.busy) {
            ++it;
            continue;
        }
        seed_free(it->second);
        it = g_seeds.erase(it);
    }
}

void factor_engine(Operator &op, PlainLib &lib, hipStream_t s) {
    const uint32_t init[6] = {0u, 0u, 0x3f800000u, 0x3f800000u, 0u, 0u};   // flags, -, ratios 1.0f, residual 0
    WUST_HIP_CHECK(hipMemcpyAsync(op.status, init, sizeof(init), hipMemcpyHostToDevice, s));
    WUST_HIP_CHECK(hipMemsetAsync(op.gj_check, 0, sizeof(float) * op.m, s));
    const int pml_cells = (int)std::ceil(op.ap.L / op.ap.dx);
    FactorRun run{op, s, op.factor_method == FACTOR_AUTO, pml_cells};
    // WUST_NS_MAX_UPDATES caps the Newton-Schulz rounds (one update each) per
    // block (tuning; a cap too small for the PML blocks exercises the
    // Gauss-Jordan fallback)
    if (const char *cap = std::getenv("WUST_NS_MAX_UPDATES")) {
        const int c = std::atoi(cap);
        if (c >= 1 && c <= NS_KMAX) run.update_cap = c;
    }
    const char *fixed = std::getenv("WUST_NS_FIXED_BUDGET");
    const bool learn = !(fixed && fixed[0] == '1') && !op.fixed_budget;
    const BudgetKey key = budget_key(op);
    if (learn && run.allow_ns && op.ns_rounds) {
        std::lock_guard<std::mutex> lk(g_budget_mu);
        auto it = g_budgets.find(key);
        if (it != g_budgets.end()) run.learned = it->second;
    }
    // the cached leading pages of this key, when no other factorization holds them
    const bool seeding = learn && run.allow_ns && op.nswork && op.ns_rounds && seed_cache_cap() > 0;
    const std::pair<int, BudgetKey> skey{current_device_id(), key};
    SeedHold hold;
    if (seeding) {
        std::lock_guard<std::mutex> lk(g_seed_mu);
        auto it = g_seeds.find(skey);
        if (it == g_seeds.end()) it = g_seeds.emplace(skey, SeedEntry{}).first;
        if (!it->second.busy) {
            hold.e = &it->second;
            hold.e->busy = true;
            hold.e->used = ++g_seed_clock;
            run.seeds = hold.e->steps > 0 ? hold.e->buf : nullptr;
            run.seed_steps = run.seeds ? hold.e->steps : 0;
            run.seeded_upto = run.seed_steps;
        }
    }
    if (op.ns_rounds) WUST_HIP_CHECK(hipMemsetAsync(op.ns_rounds, 0, sizeof(int) * op.m, s));
    run.run();
    uint32_t st[6];
    std::vector<int> used;
    std::vector<uint8_t> profile;   // the profile left for the next factorization with this key
    if (learn && run.allow_ns && op.ns_rounds) {
        used.resize(op.m);
        WUST_HIP_CHECK(hipMemcpyAsync(used.data(), op.ns_rounds, sizeof(int) * op.m, hipMemcpyDeviceToHost, s));
    }
    WUST_HIP_CHECK(hipMemcpyAsync(st, op.status, sizeof(st), hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    if (!used.empty()) {
        // chain step st produced blocks st (left) and m - 1 - st (right)
        const int m = op.m, nsteps = std::max(op.p, m - 1 - op.p);
        std::vector<uint8_t> next((size_t)nsteps, 0);
        for (int j = 0; j < m; ++j) {
            if (j == op.p || used[j] == 0) continue;
            const int stp = j < op.p ? j : m - 1 - j;
            if (stp < run.seeded_upto) continue;   // rounds from a cached seed say nothing of the neighbour seed
            const bool forced = used[j] >= 64;
            int need = (used[j] & 63) + (forced ? 1 : 0);
            // a step that ran out of rounds needs at least one more than its profile gave it, even
            // when a cap held its rounds below that
            if (forced && stp < (int)run.learned.size() && run.learned[stp] >= 1)
                need = std::max<int>(need, run.learned[stp] + 1);
            need = std::min(need, NS_KMAX);
            next[stp] = (uint8_t)std::max<int>(next[stp], need);
        }
        for (int stp = 0; stp < run.seeded_upto && stp < nsteps; ++stp)
            next[stp] = stp < (int)run.learned.size() ? run.learned[stp] : 0;
        if (const char *e = std::getenv("WUST_NS_LOG"); e && e[0] == '1') {
            float bound;
            std::memcpy(&bound, st + 4, sizeof(bound));
            std::fprintf(stderr, "newton-schulz: %d x %d, f %.4g Hz, worst bound %.3g%s; rounds per step:", op.ny,
                         op.nx, op.ap.f, bound, (st[0] & ST_INVERSE_UNCONVERGED) ? " (unconverged)" : "");
            for (int j = 0; j < m; ++j)
                if (j != op.p && used[j] >= 64) std::fprintf(stderr, " [block %d forced at %d]", j, used[j] & 63);
            for (int i = 0; i < nsteps && i < 48; ++i) std::fprintf(stderr, " %d", (int)next[i]);
            std::fprintf(stderr, " ... (%d leading steps seeded from the previous factorization)\n", run.seeded_upto);
        }
        // An unconverged factorization keeps its profile too: the steps whose blocks ran out of
        // rounds (forced, one more than they had) or needed 4+ reach Gauss-Jordan in the next
        // factorization with this key, which seeds the later steps from correct inverses.  (Starting
        // again from the defaults would repeat a divergence that the defaults cause, e.g. the third
        // column of a chain at low frequency, and the full Gauss-Jordan fallback, at every
        // factorization.)
        std::lock_guard<std::mutex> lk(g_budget_mu);
        g_budgets[key] = next;
        profile = next;
    }
    op.factor_fallback = 0;
    if (st[0] & ST_INVERSE_UNCONVERGED) {
        // some block's Newton-Schulz iteration missed its bound: redo every block by Gauss-Jordan
        WUST_HIP_CHECK(hipMemcpyAsync(op.status, init, sizeof(init), hipMemcpyHostToDevice, s));
        WUST_HIP_CHECK(hipMemsetAsync(op.gj_check, 0, sizeof(float) * op.m, s));
        FactorRun again{op, s, false, pml_cells};
        again.run();
        WUST_HIP_CHECK(hipMemcpyAsync(st, op.status, sizeof(st), hipMemcpyDeviceToHost, s));
        WUST_HIP_CHECK(hipStreamSynchronize(s));
        op.factor_fallback = 1;
        run.ns_blocks = 0;
    }
    float r[3];
    std::memcpy(r, st + 2, sizeof(r));
    op.min_pivot_ratio = r[0] < r[1] ? r[0] : r[1];
    op.inverse_residual = r[2];
    op.ns_blocks = run.ns_blocks;
    // the pivot-free Gauss-Jordan blocks: ||I - T_j X_j||_F of each, and the zero / non-finite pivots
    std::vector<float> gj((size_t)op.m);
    WUST_HIP_CHECK(hipMemcpyAsync(gj.data(), op.gj_check, sizeof(float) * op.m, hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    float accept = GJ_ACCEPT;
    if (const char *e = std::getenv("WUST_GJ_ACCEPT"); e && e[0]) accept = (float)std::atof(e);
    double worst = 0;
    bool bad = (st[0] & (ST_ZERO_PIVOT | ST_NONFINITE_PIVOT)) != 0;
    for (float v : gj) {
        if (!(v >= 0.f) || !std::isfinite(v)) bad = true;
        else worst = std::max(worst, std::sqrt((double)v));
    }
    if (worst > accept) bad = true;
    op.gj_residual = worst;
    if (const char *e = std::getenv("WUST_NS_LOG"); e && e[0] == '1')
        std::fprintf(stderr, "gauss-jordan check: %d x %d, largest ||I - T X||_F %.3g%s\n", op.ny, op.nx, worst,
                     bad ? " -> pivoted refactorization" : "");
    if (bad) {
        // a pivot-free elimination lost accuracy (or met a zero pivot): every block again with partial
        // pivoting; a block that getrf finds exactly singular is WUST_SINGULAR_SCHUR_BLOCK
        lib.ensure(s, op.n);
        const int ninfo = std::min(op.m, 65536);
        WUST_HIP_CHECK(hipMemsetAsync(lib.info, 0, sizeof(int) * 2 * (size_t)ninfo, s));
        WUST_HIP_CHECK(hipMemcpyAsync(op.status, init, sizeof(init), hipMemcpyHostToDevice, s));
        FactorRun piv{op, s, false, pml_cells};
        piv.pivoted = &lib;
        piv.run();
        std::vector<int> info(2 * (size_t)ninfo);
        WUST_HIP_CHECK(hipMemcpyAsync(info.data(), lib.info, sizeof(int) * info.size(), hipMemcpyDeviceToHost, s));
        WUST_HIP_CHECK(hipStreamSynchronize(s));
        for (size_t i = 0; i < info.size(); i += 2)
            if (info[i] > 0)
                throw Error(WUST_SINGULAR_SCHUR_BLOCK, "getrf reported a singular Schur block (block " +
                                                           std::to_string(i / 2) + ")");
        op.factor_fallback = 2;
        op.ns_blocks = 0;
        op.min_pivot_ratio = 1.0;
        op.inverse_residual = 0;
    }
    if (hold.e) update_seed_cache(op, *hold.e, profile, op.factor_fallback == 0 ? run.seeded_upto : -1, s);
}

}  // namespace wust
