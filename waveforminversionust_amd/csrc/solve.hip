// Complex-FP32 block sweeps over grid columns, batched over all right-hand sides.
//
// The reference (Runtime/solver/applyBlockLU.cu:433-467) runs, per column, one
// tridiagonal kernel and one cuBLAS Cgemm3m (Ny x R x Ny), serially over x
// for the forward and backward passes.  Here each step is ONE launch that
// fuses the tridiagonal coupling into the operand load of a complex GEMM on
// MFMA (v_mfma_f32_16x16x4f32, four real products per complex product,
// FP32 accumulation), and the twisted factorization lets the left and right
// chains of a step run in the same launch:
//   forward  left  y_j = X_j (b_j - L_j y_{j-1}),       j = 0..p-1
//            right z_j = X_j (b_j - U_j z_{j+1}),       j = m-1..p+1
//            mid   u_p = X_p (b_p - L_p y_{p-1} - U_p z_{p+1})
//            back  u_j = y_j - X_j U_j u_{j+1} (left),  u_j = z_j - X_j L_j u_{j-1} (right)
//   adjoint  the same with X_j^H and the conjugate-transposed couplings
//            (H^H has U_{j-1}^H below and L_{j+1}^H above the diagonal).
// Layout of every field block: [npad rows][Rpad RHS], RHS fastest, so the
// coupling loads and the output stores are coalesced across RHS.
#include "engine.h"

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
    cf *out;
    int accumulate;           // out += X B instead of out = X B
};
struct SweepParams {
    SweepTask task[2];
    int ntasks, npad, Rpad;
    int units, per_xcd;       // logical tiles and tiles per XCD slot
};

__device__ __forceinline__ cf ldc(const cf *p) { return *p; }

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

// One WG = one 16x16 output tile (16 block rows x 16 RHS); its 4 waves split
// the K = npad reduction into quarters and combine through LDS.
__global__ void __launch_bounds__(256) sweep_kernel(SweepParams p) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int L = xcd * p.per_xcd + local;
    if (L >= p.units) return;
    const int Mb = p.npad / 16, Nb = p.Rpad / 16;
    const int nb = L % Nb;
    const int mb = (L / Nb) % Mb;
    const int task = L / (Nb * Mb);
    const SweepTask &tk = p.task[task];
    const int npad = p.npad, Rpad = p.Rpad;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, kk = lane >> 4;
    const int i0 = mb * 16, j0 = nb * 16;
    const int K4 = npad / 4;
    const int kbeg = wave * K4, kend = kbeg + K4;

    v4f accr = {0.f, 0.f, 0.f, 0.f}, acci = {0.f, 0.f, 0.f, 0.f};
    const cf *X = tk.X;
    for (int k0 = kbeg; k0 < kend; k0 += 4) {
        const int ka = k0 + kk;
        cf a = tk.xconj ? conjf_(X[(int64_t)ka + (int64_t)npad * (i0 + li)]) : X[(int64_t)(i0 + li) + (int64_t)npad * ka];
        cf b = operand_b(tk, ka, j0 + li, npad, Rpad);
        accr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.re, b.re, accr, 0, 0, 0);
        accr = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.im, b.im, accr, 0, 0, 0);
        acci = __builtin_amdgcn_mfma_f32_16x16x4f32(a.re, b.im, acci, 0, 0, 0);
        acci = __builtin_amdgcn_mfma_f32_16x16x4f32(a.im, b.re, acci, 0, 0, 0);
    }
    __shared__ cf red[4][16][17];
    for (int r = 0; r < 4; ++r) red[wave][4 * kk + r][li] = mk(accr[r], acci[r]);
    __syncthreads();
    const int t = threadIdx.x;
    const int row = t >> 4, col = t & 15;
    cf s = red[0][row][col] + red[1][row][col] + red[2][row][col] + red[3][row][col];
    cf *o = tk.out + (int64_t)(i0 + row) * Rpad + j0 + col;
    if (tk.accumulate) s = s + *o;
    *o = s;
}

}  // namespace

void solve_engine(const Operator &op, const cf *src, cf *out, int Rpad, bool adjoint, hipStream_t s) {
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
        sp.npad = npad;
        sp.Rpad = Rpad;
        sp.units = sp.ntasks * (npad / 16) * (Rpad / 16);
        sp.per_xcd = (sp.units + 7) / 8;
        hipLaunchKernelGGL(sweep_kernel, dim3(8 * sp.per_xcd), dim3(256), 0, s, sp);
        WUST_LAUNCH_CHECK();
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
