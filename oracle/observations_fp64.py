"""FP64 CPU restatement of the RF -> frequency DTFT (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU leg import this module, as the
checker of the GPU engine.

Pins (tests/test_dtft_oracle.py): the closed form of the DTFT of a sampled
sinusoid with a non-zero time origin (a geometric series), a brute-force
double loop on a small case, the masking rules (an inactive channel is zeroed
before the sum, so a non-finite sample there is ignored; a non-finite sample
on an active channel is rejected), and the equivalence with the reference's
own "direct" ingestion fixture (Runtime/tests/test_ingest_frequency.m:96-124).
"""
from __future__ import annotations

import numpy as np

from .helmholtz_fp64 import WustError


def rf_dtft(pressure, time_s, frequencies_hz, active, single=True):
    """P[tx, rx, f] = sum_t exp(-2 pi i f t) dt x(t, rx, tx) (Runtime/matlab/wust_prepare.m:29-37).

    ``pressure`` is [time, rx, tx] real; ``active`` [tx, rx] boolean.  The
    kernel and the products are FP64; ``single`` rounds the result to
    complex64 as the reference's ``single(kernel * traces)`` does.
    """
    x = np.asarray(pressure)
    if x.ndim == 2:
        x = x[:, :, None]
    t = np.asarray(time_s, np.float64).ravel()
    f = np.asarray(frequencies_hz, np.float64).ravel()
    dt = t[1] - t[0]
    kernel = np.exp(-2j * np.pi * f[:, None] * t[None, :]) * dt          # :30
    nt_time, nr, ntx = x.shape
    out = np.zeros((ntx, nr, f.size), np.complex64 if single else np.complex128)
    act = np.asarray(active, bool)
    for tx in range(ntx):                                                 # :31-37
        traces = x[:, :, tx].astype(np.float64)
        traces[:, ~act[tx]] = 0
        if not np.all(np.isfinite(traces)):
            raise WustError("WUST:InvalidInput", "Nonfinite RF on an active channel")
        p = kernel @ traces                                               # [nf, nr]
        out[tx] = p.T.astype(out.dtype)
    return out
