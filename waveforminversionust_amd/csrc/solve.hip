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
#include "engine.h"
#include "profile.h"

namespace wust {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

struct SweepTerm {
    const cf *cn, *cc, *cs;   // tridiagonal row coefficients (block of the coupling)
    const cf *v;              // field block the coupling acts on
    int transposed;           // apply P^H instead of P
};
struct SweepTask {
    const cf *X;              // inverse page
    int xconj;                // use X^H
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

// GEMM pass: out (+)= op(X) B on a 16 x 16 tile (16 block rows x 16 RHS).
// The 4 waves of a WG split the K range into quarters and stream their MFMA
// operands straight from global memory into registers in the MFMA lane layout
// (A: lane (li, kq) = X[row li][k kq], B: lane = B[k kq][col li]) -- no operand
// is shared between waves, so no LDS is needed in the main loop.  Operands are
// prefetched in two register batches of DB sub-steps (4 k each).  The four
// partial tiles are summed through LDS at the end.  16 x 16 tiles give
// 2 * 36 * 16 = 1152 WGs at the 560^2 x 256-RHS size (4.5 per CU), which
// balances the 256 CUs far better than 32 x 32 tiles (288 WGs).
constexpr int DB = 6;   // sub-steps per prefetch batch (24 k)

// 5 waves per SIMD keep all 1152 WGs of a 560^2 x 256-RHS step resident at once.
__global__ void __launch_bounds__(256, 5) sweep_gemm_kernel(SweepParams p) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int L = xcd * p.per_xcd + local;
    if (L >= p.units) return;
    const int npad = p.npad, Rpad = p.Rpad;
    const int Mb = npad / 16, Nb = Rpad / 16;
    const int nb = L % Nb;
    const int mb = (L / Nb) % Mb;
    const int task = L / (Nb * Mb);
    const SweepTask &tk = p.task[task];
    const int i0 = mb * 16, j0 = nb * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int kquarter = npad / 4;           // multiple of 8 (npad is a multiple of 32)
    const int nsub = kquarter / 4;           // sub-steps of 4 k in this wave's quarter
    const int kbeg = wave * kquarter;

    // lane-resolved operand pointers; sub-step s reads element offset s*4*stride
    const cf *ap;
    int64_t astep;
    if (!tk.xconj) {
        ap = tk.X + (int64_t)(i0 + li) + (int64_t)npad * (kbeg + kq);
        astep = 4 * (int64_t)npad;
    } else {
        ap = tk.X + (int64_t)(kbeg + kq) + (int64_t)npad * (i0 + li);
        astep = 4;
    }
    const cf *bp = tk.B + (int64_t)(kbeg + kq) * Rpad + j0 + li;
    const int64_t bstep = 4 * (int64_t)Rpad;
    const float asgn = tk.xconj ? -1.f : 1.f;   // conj(X) for the adjoint

    v4f ar = {0.f, 0.f, 0.f, 0.f}, ai = ar;
    cf a0[DB], b0[DB], a1[DB], b1[DB];
    auto load = [&](int s0, cf (&ra)[DB], cf (&rb)[DB]) {
#pragma unroll
        for (int u = 0; u < DB; ++u) {
            const int sidx = s0 + u;
            if (sidx < nsub) {
                ra[u] = ap[sidx * astep];
                rb[u] = bp[sidx * bstep];
            } else {
                ra[u] = mk(0.f, 0.f);
                rb[u] = mk(0.f, 0.f);
            }
        }
    };
    auto mma = [&](const cf (&ra)[DB], const cf (&rb)[DB]) {
#pragma unroll
        for (int u = 0; u < DB; ++u) {
            const float are = ra[u].re, aim = asgn * ra[u].im;
            ar = __builtin_amdgcn_mfma_f32_16x16x4f32(are, rb[u].re, ar, 0, 0, 0);
            ai = __builtin_amdgcn_mfma_f32_16x16x4f32(are, rb[u].im, ai, 0, 0, 0);
            ar = __builtin_amdgcn_mfma_f32_16x16x4f32(-aim, rb[u].im, ar, 0, 0, 0);
            ai = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, rb[u].re, ai, 0, 0, 0);
        }
    };
    load(0, a0, b0);
    for (int s0 = 0; s0 < nsub; s0 += 2 * DB) {
        load(s0 + DB, a1, b1);
        mma(a0, b0);
        if (s0 + DB >= nsub) break;
        load(s0 + 2 * DB, a0, b0);
        mma(a1, b1);
    }
    __shared__ cf red[3][16][17];
    if (wave > 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wave - 1][4 * kq + r][li] = mk(ar[r], ai[r]);
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * kq + r;
            cf sacc = mk(ar[r], ai[r]) + red[0][row][li] + red[1][row][li] + red[2][row][li];
            cf *o = tk.out + (int64_t)(i0 + row) * Rpad + j0 + li;
            if (tk.accumulate) sacc = sacc + *o;
            *o = sacc;
        }
    }
}

}  // namespace

void solve_engine(const Operator &op, const cf *src, cf *out, cf *scratch, int Rpad, bool adjoint, hipStream_t s) {
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
        sp.units = sp.ntasks * (npad / 16) * (Rpad / 16);
        sp.per_xcd = (sp.units + 7) / 8;
        launch_k(PC_SWEEP_GEMM, sweep_gemm_kernel, dim3(8 * sp.per_xcd), dim3(256), 0, s, sp);
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
