"""wust_rf_dtft_noalloc (the DTFT with its work memory in a static per-device array) against
wust_rf_dtft, through the C ABI: identical results, call after call, as the input shrinks."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from waveforminversionust_amd import _native as N  # noqa: E402

MAX_TIME = 16384   # WUST_DTFT_NOALLOC_MAX_TIME


def _inputs(n_time, n_rx, n_tx, n_freq, dtype, seed, nan_at=None):
    g = torch.Generator(device="cuda").manual_seed(seed)
    rf = torch.randn((n_tx, n_rx, n_time), generator=g, device="cuda", dtype=dtype)   # storage [tx][rx][t]
    dt = 25e-9
    t = torch.as_tensor(2e-6 + np.arange(n_time) * dt, device="cuda")
    freqs = np.ascontiguousarray(np.linspace(0.3e6, 1.2e6, n_freq))
    rng = np.random.default_rng(seed)
    active = rng.uniform(size=(n_tx, n_rx)) < 0.9
    if nan_at is not None:
        tx, rx = nan_at
        active[tx, rx] = True
        rf[tx, rx, n_time // 2] = float("nan")
    act = torch.as_tensor(active, device="cuda").to(torch.uint8).t().contiguous()    # MATLAB [tx, rx]
    return rf, t, dt, freqs, act


def _call(lib, noalloc, rf, t, dt, freqs, act, n_time, n_rx, n_tx):
    out = torch.full((freqs.size, n_rx, n_tx), complex("nan"), dtype=torch.complex64, device="cuda")
    args = [C.c_void_p(rf.data_ptr()), 1 if rf.dtype == torch.float64 else 0, n_time, n_rx, n_tx,
            C.c_void_p(t.data_ptr()), dt, freqs.ctypes.data_as(C.c_void_p), freqs.size, C.c_void_p(act.data_ptr()),
            C.c_void_p(out.data_ptr())]
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        st = lib.wust_rf_dtft_noalloc(*args, stream)
    else:
        st = lib.wust_rf_dtft(*args, None, stream)
    torch.cuda.synchronize()
    return st, out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_matches_allocating_entry_point(dtype):
    lib = N.load()
    # first call: 40 frequencies (two passes over the RF), then a smaller one reusing the same array
    for shape in [(4097, 37, 9, 40), (1000, 16, 5, 12), (64, 3, 2, 1)]:
        n_time, n_rx, n_tx, n_freq = shape
        rf, t, dt, freqs, act = _inputs(n_time, n_rx, n_tx, n_freq, dtype, sum(shape))
        st0, want = _call(lib, False, rf, t, dt, freqs, act, n_time, n_rx, n_tx)
        st1, got = _call(lib, True, rf, t, dt, freqs, act, n_time, n_rx, n_tx)
        assert st0 == 0 and st1 == 0, lib.wust_last_error()
        assert torch.isfinite(torch.view_as_real(want)).all()
        assert torch.equal(got, want), shape


def test_noalloc_largest_input_then_smaller():
    lib = N.load()
    for shape in [(MAX_TIME, 8, 4, 33), (300, 5, 3, 7)]:
        n_time, n_rx, n_tx, n_freq = shape
        rf, t, dt, freqs, act = _inputs(n_time, n_rx, n_tx, n_freq, torch.float32, 7)
        st0, want = _call(lib, False, rf, t, dt, freqs, act, n_time, n_rx, n_tx)
        st1, got = _call(lib, True, rf, t, dt, freqs, act, n_time, n_rx, n_tx)
        assert st0 == 0 and st1 == 0
        assert torch.equal(got, want), shape


def test_noalloc_rejects_and_reports():
    lib = N.load()
    # beyond the static array: rejected on the host before any device work
    rf, t, dt, freqs, act = _inputs(MAX_TIME + 1, 2, 1, 3, torch.float32, 3)
    st, _ = _call(lib, True, rf, t, dt, freqs, act, MAX_TIME + 1, 2, 1)
    assert N.STATUS_IDS[st] == "WUST:InvalidInput" and b"allocation-free" in lib.wust_last_error()
    # a non-finite sample on an active channel: the status word in the static array reports it, and the
    # next call starts from a cleared status
    rf, t, dt, freqs, act = _inputs(512, 4, 3, 5, torch.float32, 4, nan_at=(1, 2))
    st, _ = _call(lib, True, rf, t, dt, freqs, act, 512, 4, 3)
    assert N.STATUS_IDS[st] == "WUST:InvalidInput"
    rf, t, dt, freqs, act = _inputs(512, 4, 3, 5, torch.float32, 5)
    st0, want = _call(lib, False, rf, t, dt, freqs, act, 512, 4, 3)
    st1, got = _call(lib, True, rf, t, dt, freqs, act, 512, 4, 3)
    assert st0 == 0 and st1 == 0 and torch.equal(got, want)
