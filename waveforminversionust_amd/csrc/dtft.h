// RF -> frequency DTFT (FP64 MFMA), MATLAB column-major layouts.
#pragma once
#include "wust_common.h"

namespace wust {

struct DtftParams {
    const void *rf;          // [n_time, n_rx, n_tx] float or double
    const double *K;         // [rows][nt_pad] kernel table (re / im rows per 8 frequencies)
    const uint8_t *active;   // [n_tx, n_rx]: channel used by any frequency
    void *out;               // complex64 [n_tx, n_rx, n_freq]
    uint32_t *status;        // bit 0: non-finite sample on an active channel
    int64_t n_time, nt_pad, ncols;
    int n_rx, n_tx, f0, n_freq_chunk;
};

// P(tx, r, f) = single(sum_t exp(-2 pi i f time_s[t]) dt rf(t, r, tx)) on active channels, 0 elsewhere.
// time_s: device [n_time]; freqs_host: host [n_freq].  kernel_ms (may be null) receives the summed
// duration of the DTFT GEMM launches.  Throws WUST_INVALID_INPUT for a non-finite active sample.
void rf_dtft(const void *rf, bool is_double, int64_t n_time, int n_rx, int n_tx, const double *time_s, double dt,
             const double *freqs_host, int n_freq, const uint8_t *active, void *out, hipStream_t s,
             float *kernel_ms);

// The same DTFT with its work memory (status word, kernel table) in a static per-device array:
// no device allocation.  n_time <= DTFT_NOALLOC_MAX_TIME, else WUST_INVALID_INPUT.
constexpr int64_t DTFT_NOALLOC_MAX_TIME = WUST_DTFT_NOALLOC_MAX_TIME;
void rf_dtft_noalloc(const void *rf, bool is_double, int64_t n_time, int n_rx, int n_tx, const double *time_s,
                     double dt, const double *freqs_host, int n_freq, const uint8_t *active, void *out,
                     hipStream_t s);

}  // namespace wust
