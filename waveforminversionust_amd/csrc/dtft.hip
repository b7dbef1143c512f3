// RF -> frequency DTFT of the data-preparation side (Runtime/matlab/wust_prepare.m:29-37):
//
//   P(tx, r, f) = single( sum_t exp(-2 pi i f t_s[t]) dt x(t, r, tx) ),
//
// with channels whose mask is false for every frequency zeroed before the sum.
// The reference forms the [nf x Nt] kernel in FP64 and multiplies it with the
// [Nt x nr] traces of one transmitter in FP64, then rounds to single.  Here
// the whole acquisition is one real FP64 GEMM on MFMA,
//
//   C[rows x N] = K[rows x Nt] . X[Nt x N],  N = nr * nt columns (t fastest in X),
//
// where K holds the real and imaginary parts of the kernel as separate rows:
// row tile T (16 rows) carries frequencies 8T..8T+7, rows 0-7 their real
// parts and rows 8-15 their imaginary parts.  v_mfma_f64_16x16x4f64 takes
// A[i][k] from lane i + 16 k and B[k][j] from lane j + 16 k, and leaves
// D[i][j] in lane j + 16 (i % 4), register i / 4 (measured with one-hot
// operands on gfx950).  Lane j + 16 g therefore holds rows g, g + 4, g + 8,
// g + 12: the real and imaginary parts of frequencies 8T + g and 8T + g + 4
// in registers (0, 2) and (1, 3), ready for complex64 stores.
//
// X is read once from HBM (the kernel is HBM-bound for double input and
// balanced with FP64 MFMA for single input at ~12 frequencies); K is a small
// L2-resident table staged through LDS one sample chunk at a time and shared
// by the eight waves of a workgroup.
#include <mutex>
#include <string>

#include "dtft.h"

namespace wust {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int DT_WAVES = 8;        // waves per workgroup, 16 channels each
constexpr int DT_KC = 64;          // samples per chunk (4 k-blocks of 16)
constexpr int DT_LDS_PAD = 2;      // doubles of padding per LDS row

// The (up to) 32 frequencies of one pass over the RF, passed by value.
struct FreqChunk {
    double f[32];
};

// K[row][t] for row tiles of 8 frequencies (re rows 0-7, im rows 8-15), zero past
// n_time and past n_freq.  theta = (-2 pi f) t in that order of rounding, as the
// reference's exp(-2i*pi*frequency_hz*time_s.') evaluates it.  status (or null) is
// cleared by the first thread, before the GEMM of the first pass reads it.
__global__ void dtft_table_kernel(const double *time_s, FreqChunk freqs, int n_freq, int64_t n_time,
                                  int64_t nt_pad, int rows, double dt, double *K, uint32_t *status) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (status && i == 0) *status = 0u;
    if (i >= (int64_t)rows * nt_pad) return;
    const int row = (int)(i / nt_pad);
    const int64_t t = i - (int64_t)row * nt_pad;
    const int tile = row >> 4, in = row & 15;
    const int f = tile * 8 + (in & 7);
    double v = 0.0;
    if (f < n_freq && t < n_time) {
        const double w = -2.0 * M_PI * freqs.f[f];
        const double th = w * time_s[t];
        double sn, cs;
        sincos(th, &sn, &cs);
        v = (in < 8 ? cs : sn) * dt;
    }
    K[i] = v;
}

// This lane's 16 samples of one chunk, t0 + 16 b + 4 kq + s (b = k-block, s = MFMA step), as
// stored (float or double); zero for an inactive channel or past n_time.  VEC: n_time is a
// multiple of 4, so each group of four is contiguous, aligned, and either all in range or not.
template <typename T, bool VEC>
__device__ __forceinline__ void load_chunk(const T *xc, int64_t t0, int kq, int64_t n_time, bool active, T (&v)[4][4]) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int64_t t = t0 + 16 * b + 4 * kq;
        if (VEC) {
            const bool in = active && t < n_time;
            const T *p = xc + (in ? t : 0);
            if (sizeof(T) == 4) {
                const float4 u = *reinterpret_cast<const float4 *>(p);
                v[b][0] = in ? (T)u.x : (T)0;
                v[b][1] = in ? (T)u.y : (T)0;
                v[b][2] = in ? (T)u.z : (T)0;
                v[b][3] = in ? (T)u.w : (T)0;
            } else {
                const double2 u0 = reinterpret_cast<const double2 *>(p)[0];
                const double2 u1 = reinterpret_cast<const double2 *>(p)[1];
                v[b][0] = in ? (T)u0.x : (T)0;
                v[b][1] = in ? (T)u0.y : (T)0;
                v[b][2] = in ? (T)u1.x : (T)0;
                v[b][3] = in ? (T)u1.y : (T)0;
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const bool in = active && t + s < n_time;
                const T u = xc[in ? t + s : 0];
                v[b][s] = in ? u : (T)0;
            }
        }
    }
}

// One wave: 16 channels x RT row tiles over all samples; a workgroup of DT_WAVES waves shares
// the kernel table, double-buffered in LDS (one barrier per chunk; the next chunk's table and
// samples are in flight while the current chunk's MFMAs run).
template <int RT, typename T, bool VEC>
__global__ void __launch_bounds__(64 * DT_WAVES) dtft_kernel(DtftParams p) {
    constexpr int ROWS = RT * 16;
    constexpr int NT = 64 * DT_WAVES;
    constexpr int PER_THREAD = ROWS * DT_KC / NT;   // table doubles staged by each thread
    static_assert(PER_THREAD * NT == ROWS * DT_KC, "table chunk must split evenly over the threads");
    __shared__ double ksh[2][ROWS][DT_KC + DT_LDS_PAD];
    const T *x = static_cast<const T *>(p.rf);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, kq = lane >> 4;
    const int64_t col = ((int64_t)blockIdx.x * DT_WAVES + wave) * 16 + c;
    const bool col_ok = col < p.ncols;
    const int64_t colc = col_ok ? col : 0;
    const int tx = (int)(colc / p.n_rx), rx = (int)(colc - (int64_t)tx * p.n_rx);
    const bool active = col_ok && p.active[tx + (int64_t)p.n_tx * rx] != 0;
    const T *xc = x + colc * p.n_time;
    // table staging: element q * NT + threadIdx.x of the [ROWS][DT_KC] chunk (coalesced rows)

    v4d acc[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) acc[r] = v4d{0.0, 0.0, 0.0, 0.0};
    bool bad = false;
    T xv[4][4], xn[4][4];
    double kr[PER_THREAD];
    // prologue: chunk 0
    load_chunk<T, VEC>(xc, 0, kq, p.n_time, active, xv);
#pragma unroll
    for (int q = 0; q < PER_THREAD; ++q) {
        const int e = q * NT + threadIdx.x, row = e / DT_KC, tt = e - row * DT_KC;
        ksh[0][row][tt] = p.K[(int64_t)row * p.nt_pad + tt];
    }
    __syncthreads();
    const int64_t nchunks = p.nt_pad / DT_KC;
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        const int cur = (int)(ch & 1);
        const bool more = ch + 1 < nchunks;
        if (more) {
            load_chunk<T, VEC>(xc, (ch + 1) * DT_KC, kq, p.n_time, active, xn);
#pragma unroll
            for (int q = 0; q < PER_THREAD; ++q) {
                const int e = q * NT + threadIdx.x, row = e / DT_KC, tt = e - row * DT_KC;
                kr[q] = p.K[(int64_t)row * p.nt_pad + (ch + 1) * DT_KC + tt];
            }
        }
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const double xd = (double)xv[b][s];
                bad |= !isfinite(xd);
#pragma unroll
                for (int r = 0; r < RT; ++r)
                    acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(ksh[cur][16 * r + c][16 * b + 4 * kq + s], xd,
                                                                  acc[r], 0, 0, 0);
            }
        if (more) {
#pragma unroll
            for (int q = 0; q < PER_THREAD; ++q) {
                const int e = q * NT + threadIdx.x, row = e / DT_KC, tt = e - row * DT_KC;
                ksh[cur ^ 1][row][tt] = kr[q];
            }
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int s = 0; s < 4; ++s) xv[b][s] = xn[b][s];
        }
        __syncthreads();
    }
    if (__any(bad) && lane == 0) atomicOr(p.status, 1u);
    // registers (q, q + 2) of lane c + 16 kq: real and imaginary parts of frequency 8 r + kq + 4 q
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int f = 8 * r + kq + 4 * q;
            if (col_ok && f < p.n_freq_chunk) {
                float2 v = make_float2((float)acc[r][q], (float)acc[r][q + 2]);
                reinterpret_cast<float2 *>(p.out)[tx + (int64_t)p.n_tx * (rx + (int64_t)p.n_rx * (p.f0 + f))] = v;
            }
        }
}

template <typename T, bool VEC>
void launch_rt(int rt, const DtftParams &p, dim3 grid, hipStream_t s) {
    switch (rt) {
        case 1: hipLaunchKernelGGL((dtft_kernel<1, T, VEC>), grid, dim3(64 * DT_WAVES), 0, s, p); break;
        case 2: hipLaunchKernelGGL((dtft_kernel<2, T, VEC>), grid, dim3(64 * DT_WAVES), 0, s, p); break;
        case 3: hipLaunchKernelGGL((dtft_kernel<3, T, VEC>), grid, dim3(64 * DT_WAVES), 0, s, p); break;
        default: hipLaunchKernelGGL((dtft_kernel<4, T, VEC>), grid, dim3(64 * DT_WAVES), 0, s, p); break;
    }
    WUST_LAUNCH_CHECK();
}

template <typename T>
struct Buf {
    T *p = nullptr;
    explicit Buf(int64_t n) { WUST_HIP_CHECK(hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1))); }
    ~Buf() {
        if (p) (void)hipFree(p);
    }
};

// Largest kernel table of one pass: 4 row tiles (32 frequencies) of nt_pad doubles.
int64_t table_elems(int64_t n_time) {
    const int64_t nt_pad = (n_time + DT_KC - 1) / DT_KC * DT_KC;
    return (int64_t)64 * nt_pad;
}

// The DTFT on caller-provided work memory: K (table_elems(n_time) doubles) and status (one word).
// Passes over the RF run back to back on the stream (each pass's table kernel follows the
// previous pass's GEMM in stream order, so one table buffer serves every pass); the status is
// read once at the end.
void dtft_run(const void *rf, bool is_double, int64_t n_time, int n_rx, int n_tx, const double *time_s, double dt,
              const double *freqs_host, int n_freq, const uint8_t *active, void *out, hipStream_t s,
              float *kernel_ms, double *Kbuf, uint32_t *status) {
    const int64_t nt_pad = (n_time + DT_KC - 1) / DT_KC * DT_KC;
    const int64_t ncols = (int64_t)n_rx * n_tx;
    const dim3 grid((unsigned)((ncols + 16 * DT_WAVES - 1) / (16 * DT_WAVES)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (kernel_ms) {
        WUST_HIP_CHECK(hipEventCreate(&e0));
        WUST_HIP_CHECK(hipEventCreate(&e1));
    }
    float total_ms = 0.f;
    // frequencies in chunks of 32 (four row tiles); each chunk is one pass over the RF
    for (int f0 = 0; f0 < n_freq; f0 += 32) {
        const int nf = n_freq - f0 < 32 ? n_freq - f0 : 32;
        const int rt = (nf + 7) / 8;
        const int rows = rt * 16;
        FreqChunk fc{};
        for (int i = 0; i < nf; ++i) fc.f[i] = freqs_host[f0 + i];
        const int64_t nk = (int64_t)rows * nt_pad;
        hipLaunchKernelGGL(dtft_table_kernel, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, time_s, fc,
                           nf, n_time, nt_pad, rows, dt, Kbuf, f0 == 0 ? status : nullptr);
        WUST_LAUNCH_CHECK();
        DtftParams p{};
        p.rf = rf;
        p.K = Kbuf;
        p.active = active;
        p.out = out;
        p.status = status;
        p.n_time = n_time;
        p.nt_pad = nt_pad;
        p.ncols = ncols;
        p.n_rx = n_rx;
        p.n_tx = n_tx;
        p.f0 = f0;
        p.n_freq_chunk = nf;
        if (kernel_ms) WUST_HIP_CHECK(hipEventRecord(e0, s));
        const bool vec = n_time % 4 == 0;
        if (is_double) {
            if (vec) launch_rt<double, true>(rt, p, grid, s);
            else launch_rt<double, false>(rt, p, grid, s);
        } else {
            if (vec) launch_rt<float, true>(rt, p, grid, s);
            else launch_rt<float, false>(rt, p, grid, s);
        }
        if (kernel_ms) {
            WUST_HIP_CHECK(hipEventRecord(e1, s));
            WUST_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            WUST_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            total_ms += ms;
        }
    }
    uint32_t st = 0;
    WUST_HIP_CHECK(hipMemcpyAsync(&st, status, sizeof(st), hipMemcpyDeviceToHost, s));
    WUST_HIP_CHECK(hipStreamSynchronize(s));
    if (kernel_ms) {
        *kernel_ms = total_ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (st) throw Error(WUST_INVALID_INPUT, "Nonfinite RF on an active channel");
}

// Work memory of the allocation-free entry point: the status word (its own 256-byte line) and one
// pass's kernel table, for n_time <= DTFT_NOALLOC_MAX_TIME.  One copy per GPU, held until the
// process ends; every call on a GPU shares it, so calls that overlap on one GPU race.  Host code
// never uses this name as an address: it resolves the device address per device once
// (hipGetSymbolAddress) and passes that to the kernels.
constexpr int64_t DTFT_NOALLOC_TABLE = (int64_t)64 * DTFT_NOALLOC_MAX_TIME;
static __device__ __align__(256) double scratch[32 + DTFT_NOALLOC_TABLE];

std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};

void *scratch_address() {
    int dev = 0;
    WUST_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw Error(WUST_RUNTIME_ERROR, "device index out of range for the DTFT scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(WUST_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

void rf_dtft(const void *rf, bool is_double, int64_t n_time, int n_rx, int n_tx, const double *time_s, double dt,
             const double *freqs_host, int n_freq, const uint8_t *active, void *out, hipStream_t s,
             float *kernel_ms) {
    Buf<uint32_t> status(1);
    Buf<double> K(table_elems(n_time));
    dtft_run(rf, is_double, n_time, n_rx, n_tx, time_s, dt, freqs_host, n_freq, active, out, s, kernel_ms, K.p,
             status.p);
}

void rf_dtft_noalloc(const void *rf, bool is_double, int64_t n_time, int n_rx, int n_tx, const double *time_s,
                     double dt, const double *freqs_host, int n_freq, const uint8_t *active, void *out,
                     hipStream_t s) {
    if (n_time > DTFT_NOALLOC_MAX_TIME)
        throw Error(WUST_INVALID_INPUT, "n_time " + std::to_string(n_time) + " exceeds the allocation-free limit of " +
                                            std::to_string(DTFT_NOALLOC_MAX_TIME) + " samples");
    double *base = static_cast<double *>(scratch_address());
    uint32_t *status = reinterpret_cast<uint32_t *>(base);
    dtft_run(rf, is_double, n_time, n_rx, n_tx, time_s, dt, freqs_host, n_freq, active, out, s, nullptr, base + 32,
             status);
}

}  // namespace wust
