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
#include "cgemm.h"
#include "engine.h"
#include "profile.h"

namespace wust {

namespace {

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

    CgemmWave<RB, 1, ALAY, B_ROW, false> acc;
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
        sp.units = sp.ntasks * (npad / (16 * SWEEP_RB)) * (Rpad / 16);
        sp.per_xcd = (sp.units + 7) / 8;
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
