"""CPU pins of the FP64 DTFT oracle (oracle/observations_fp64.py)."""
import numpy as np
import pytest

from oracle import helmholtz_fp64 as H
from oracle import observations_fp64 as O


def test_closed_form_sinusoid():
    # x(t) = sin(w t) on t_j = t0 + j dt: sum_j e^{i a t_j} = e^{i a t0} (1 - e^{i a N dt}) / (1 - e^{i a dt})
    N, dt, t0 = 700, 1e-7, 7e-7
    t = t0 + np.arange(N) * dt
    w = 2 * np.pi * 0.37e6
    f = np.array([0.1e6, 0.35e6, 1.3e6])
    x = np.sin(w * t)[:, None, None]
    got = O.rf_dtft(x, t, f, np.ones((1, 1), bool), single=False)[0, 0]

    def geo(a):
        return np.exp(1j * a * t0) * (1 - np.exp(1j * a * N * dt)) / (1 - np.exp(1j * a * dt))

    want = np.array([(geo(w - 2 * np.pi * fk) - geo(-w - 2 * np.pi * fk)) / 2j * dt for fk in f])
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_brute_force_and_masking():
    rng = np.random.default_rng(3)
    T, nr, ntx = 37, 3, 2
    t = 2e-7 + np.arange(T) * 5e-8
    f = np.array([0.2e6, 1.1e6])
    x = rng.standard_normal((T, nr, ntx))
    active = np.array([[True, False, True], [True, True, True]])
    x[:, 1, 0] = np.nan                                   # inactive channel: ignored
    got = O.rf_dtft(x, t, f, active, single=False)
    for tx in range(ntx):
        for r in range(nr):
            for k, fk in enumerate(f):
                s = 0j
                if active[tx, r]:
                    for j in range(T):
                        s += np.exp(-2j * np.pi * fk * t[j]) * (t[1] - t[0]) * x[j, r, tx]
                assert abs(got[tx, r, k] - s) <= 1e-14 * max(1.0, abs(s))
    single = O.rf_dtft(x, t, f, active)
    assert single.dtype == np.complex64 and np.array_equal(single, got.astype(np.complex64))
    x[3, 0, 1] = np.inf
    with pytest.raises(H.WustError):
        O.rf_dtft(x, t, f, active)


def test_reference_direct_ingestion_fixture():
    # Runtime/tests/test_ingest_frequency.m:96-124: MATLAB reshape(sin((1:4096)*.17), 512, 4, 2)
    t = np.arange(512) * 1e-7 + 7e-7
    rf = np.sin(np.arange(1, 4097) * .17).reshape(2, 4, 512).transpose(2, 1, 0)
    f = np.array([1e5, 2e5, 3e5])
    a = O.rf_dtft(rf, t, f, np.ones((2, 4), bool))
    direct = np.zeros((2, 4, 3), np.complex64)
    for i in range(2):
        direct[i] = (np.exp(-2j * np.pi * f[:, None] * t[None, :]) @ rf[:, :, i] * np.mean(np.diff(t))).T
    assert np.linalg.norm(a - direct) / np.linalg.norm(a) < 1e-6
