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
//     elimination is pivot-free; a zero or non-finite pivot is reported as
//     WUST_SINGULAR_SCHUR_BLOCK, and the smallest |pivot| / max|entry below
//     it| inside each 16x16 pivot block is tracked so that a loss of
//     dominance is visible.
#include <utility>

#include "engine.h"
#include "profile.h"

namespace wust {

namespace {

constexpr int NB = 32;  // Gauss-Jordan panel width == tile edge

struct SchurTerm {
    const cf *X;                 // inverse page (npad x npad, column-major)
    const cf *Pn, *Pc, *Ps;      // left tridiagonal factor, row coefficients
    const cf *Qn, *Qc, *Qs;      // right tridiagonal factor, row coefficients
};
struct SchurTask {
    const cf *Dn, *Dc, *Ds;
    SchurTerm t[2];
    int nterms;
    cf *T;
};
struct SchurParams {
    SchurTask task[2];
    int ntasks, n, npad;
};

// T = D - sum_t P_t X_t Q_t with tridiagonal P, Q (three-point stencils on X).
__global__ void __launch_bounds__(256) schur_kernel(SchurParams p) {
    const SchurTask &tk = p.task[blockIdx.z];
    int r = blockIdx.x * 256 + threadIdx.x;
    int c = blockIdx.y;
    if (r >= p.npad) return;
    const int n = p.n, ld = p.npad;
    cf val = mk(0.f, 0.f);
    if (r >= n || c >= n) {
        val = (r == c) ? mk(1.f, 0.f) : mk(0.f, 0.f);
        tk.T[r + (int64_t)ld * c] = val;
        return;
    }
    if (r == c) val = tk.Dc[r];
    else if (c == r - 1) val = tk.Dn[r];
    else if (c == r + 1) val = tk.Ds[r];
    for (int t = 0; t < tk.nterms; ++t) {
        const SchurTerm &st = tk.t[t];
        cf qm = (c >= 1) ? st.Qs[c - 1] : mk(0.f, 0.f);
        cf q0 = st.Qc[c];
        cf qp = (c + 1 < n) ? st.Qn[c + 1] : mk(0.f, 0.f);
        cf acc = mk(0.f, 0.f);
        for (int da = -1; da <= 1; ++da) {
            int a = r + da;
            if (a < 0 || a >= n) continue;
            cf pr = (da < 0) ? st.Pn[r] : (da == 0 ? st.Pc[r] : st.Ps[r]);
            const cf *Xa = st.X + a;
            cf xq = q0 * Xa[(int64_t)ld * c];
            if (c >= 1) cfma(xq, qm, Xa[(int64_t)ld * (c - 1)]);
            if (c + 1 < n) cfma(xq, qp, Xa[(int64_t)ld * (c + 1)]);
            cfma(acc, pr, xq);
        }
        val = val - acc;
    }
    tk.T[r + (int64_t)ld * c] = val;
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

typedef float v4f __attribute__((ext_vector_type(4)));

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
    dim3 grid((op.npad + 255) / 256, op.npad, 1);
    launch_k(PC_SCHUR, schur_kernel, grid, dim3(256), 0, s, sp);
}

void factor_engine(Operator &op, hipStream_t s) {
    const int m = op.m, n = op.n, ld = op.npad, p = op.p;
    const int64_t pg = (int64_t)ld * ld;
    int Ldirs[3] = {DNW, DW, DSW};
    int Udirs[3] = {DNE, DE, DSE};
    float *ratio_d = reinterpret_cast<float *>(op.status + 2);
    float one[2] = {1.0f, 1.0f};
    WUST_HIP_CHECK(hipMemcpyAsync(ratio_d, one, sizeof(one), hipMemcpyHostToDevice, s));
    auto run_steps = [&](const std::vector<ChainStep> &steps) {
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
            tk.T = op.work + (int64_t)c * 2 * pg;
            int kind = steps[c].kind;
            if ((kind == 0 || kind == 2) && j >= 1)  // - L_j X^L_{j-1} U_{j-1}
                add_term(op, tk.t[tk.nterms++], op.page(j - 1), j, Ldirs, j - 1, Udirs);
            if ((kind == 1 || kind == 2) && j + 1 <= m - 1)  // - U_j X^R_{j+1} L_{j+1}
                add_term(op, tk.t[tk.nterms++], op.page(j + 1), j, Udirs, j + 1, Ldirs);
        }
        dim3 grid((ld + 255) / 256, ld, sp.ntasks);
        launch_k(PC_SCHUR, schur_kernel, grid, dim3(256), 0, s, sp);
        const int Q = ld / NB;
        for (int q = 0; q < Q; ++q) {
            GjParams gp{};
            gp.ntasks = sp.ntasks;
            gp.npad = ld;
            gp.q = q;
            gp.status = op.status;
            gp.pivot_ratio = ratio_d;
            for (int c = 0; c < sp.ntasks; ++c) {
                cf *b0 = op.work + (int64_t)c * 2 * pg;
                cf *b1 = b0 + pg;
                gp.task[c].src = (q % 2 == 0) ? b0 : b1;
                gp.task[c].dst = (q == Q - 1) ? op.page(steps[c].block) : ((q % 2 == 0) ? b1 : b0);
                cf *pb = op.pbuf + (int64_t)c * 2 * NB * NB;
                gp.task[c].pin = (q == 0) ? nullptr : pb + (q % 2) * NB * NB;
                gp.task[c].pout = (q + 1 < Q) ? pb + ((q + 1) % 2) * NB * NB : nullptr;
            }
            launch_k(PC_GJ_PANEL, gj_panel_kernel, dim3(sp.ntasks * (ld / NB) * (ld / NB)), dim3(256), 0, s, gp);
        }
    };

    // chains: left j = 0..p-1, right j = m-1..p+1, then the middle block p
    int nl = p, nr = m - 1 - p;
    int nsteps = nl > nr ? nl : nr;
    for (int st = 0; st < nsteps; ++st) {
        std::vector<ChainStep> steps;
        if (st < nl) steps.push_back({st, 0});
        if (st < nr) steps.push_back({m - 1 - st, 1});
        run_steps(steps);
    }
    run_steps({{p, 2}});

    float r[2];
    WUST_HIP_CHECK(hipMemcpyAsync(r, ratio_d, sizeof(r), hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    op.min_pivot_ratio = r[0] < r[1] ? r[0] : r[1];
}

}  // namespace wust
