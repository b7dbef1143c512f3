"""GPU parity of the RF -> frequency DTFT (wust_rf_dtft, FP64 MFMA) against the FP64 oracle,
and of wust_prepare against the reference's own RF assertions (Runtime/tests/test_ingest_frequency.m:96-141).

The kernel's result is the FP64 sum rounded to complex64, so each element is compared with the
oracle's unrounded FP64 value to within one float32 rounding (2^-23 relative) plus the FP64
summation-order difference.
"""
import numpy as np
import pytest
import torch

from oracle import observations_fp64 as O
from waveforminversionust_amd._native import WustError

pytestmark = pytest.mark.gpu

OB = pytest.importorskip("waveforminversionust_amd.observations")

GRID = dict(shape_yx=[7, 9], origin_yx_m=[.01, -.02], spacing_yx_m=[.001, .002], origin_kind="pixel_edge")
X = -.02 + (np.arange(9) + .5) * .002
Y = .01 + (np.arange(7) + .5) * .001
TX = np.array([[X[2], Y[2]], [X[5], Y[4]]])
RX = np.array([TX[0], [X[6], Y[1]], [X[3], Y[5]], [X[7], Y[3]]])


def _check(got, exact, x):
    got = got.cpu().numpy().astype(np.complex128)
    scale = np.max(np.abs(exact)) if exact.size else 1.0
    err = np.abs(got - exact)
    assert np.all(err <= 2.0 ** -23 * np.abs(exact) + 1e-13 * scale), np.max(err / (np.abs(exact) + scale * 1e-6))


@pytest.mark.parametrize("n_time,nr,ntx,nf,dtype", [
    (512, 4, 2, 3, np.float32), (1000, 37, 5, 12, np.float32), (257, 70, 3, 40, np.float64),
    (64, 16, 4, 8, np.float64), (4097, 9, 11, 17, np.float32)])
def test_dtft_matches_oracle(n_time, nr, ntx, nf, dtype):
    rng = np.random.default_rng(n_time + nr + ntx)
    dt = 2.5e-8
    t = 3.1e-6 + np.arange(n_time) * dt
    f = np.sort(rng.uniform(0.05, 0.45, nf)) / dt
    x = rng.standard_normal((n_time, nr, ntx)).astype(dtype)
    active = rng.random((ntx, nr)) > 0.2
    x[:, ~active.T] = np.nan                          # inactive channels may hold anything
    got = OB.rf_dtft(x, t, f, active)
    assert got.dtype == torch.complex64 and tuple(got.shape) == (ntx, nr, nf)
    exact = O.rf_dtft(x, t, f, active, single=False)
    _check(got, exact, x)
    assert np.all(got.cpu().numpy()[~active] == 0)


def test_dtft_rejects_nonfinite_active_sample():
    t = np.arange(100) * 1e-7
    x = np.ones((100, 3, 2), np.float32)
    x[10, 2, 1] = np.inf
    with pytest.raises(WustError) as ei:
        OB.rf_dtft(x, t, [1e5], np.ones((2, 3), bool))
    assert ei.value.identifier == "WUST:InvalidInput"
    act = np.ones((2, 3), bool)
    act[1, 2] = False
    OB.rf_dtft(x, t, [1e5], act)                      # the same sample on an inactive channel is ignored
    with pytest.raises(WustError):
        OB.rf_dtft(x, t, [6e6], act)                  # above the saved Nyquist


def _rf_cfg(**kw):
    cfg = dict(grid=GRID, frequencies_hz=[1e5, 2e5, 3e5], c_geom_mps=1500, phase_correction="none",
               mask=np.ones((2, 4), bool), data_units="instrument_units", measurement_provenance="fixture")
    cfg.update(kw)
    return cfg


def test_prepare_matches_direct_ingestion():
    t = np.arange(512) * 1e-7 + 7e-7
    pressure = np.sin(np.arange(1, 4097) * .17).reshape(2, 4, 512).transpose(2, 1, 0)     # [time, rx, tx]
    rf = dict(time_s=t, pressure=pressure, tx_xy_m=TX, rx_xy_m=RX)
    a = OB.wust_prepare(rf, _rf_cfg())
    assert "rf_config" not in a["preparation"] and "mask" not in a["preparation"]
    assert a["preparation"]["valid_mask_elements"] == np.count_nonzero(a["mask"])
    assert a["preparation"]["sampling_interval_s"] == t[1] - t[0]
    f = np.array([1e5, 2e5, 3e5])
    direct_p = np.zeros((3, 2, 4), np.complex64)                     # [frequency, tx, rx]
    for i in range(2):
        direct_p[:, i, :] = (np.exp(-2j * np.pi * f[:, None] * t[None, :]) @ pressure[:, :, i]
                             * np.mean(np.diff(t))).astype(np.complex64)
    direct = dict(schema="wust.frequency_input", schema_version=2, grid=GRID, tx_xy_m=TX, rx_xy_m=RX,
                  frequencies_hz=f, pressure=direct_p, pressure_axes="frequency,tx,rx", mask=np.ones((2, 4), bool),
                  mask_axes="tx,rx", fourier_sign=-1, pressure_type="total_pressure", data_units="instrument_units",
                  spectrum_normalization="dtft_dt", measurement_provenance="fixture")
    b = OB.wust_prepare_observations(direct)
    assert np.linalg.norm(a["Y"] - b["Y"]) / np.linalg.norm(a["Y"]) < 1e-6
    assert np.array_equal(a["tx_index"], b["tx_index"]) and np.array_equal(a["rx_index"], b["rx_index"])
    # time-of-flight phase correction keeps the derived energies consistent
    shifted = dict(rf, tx_xy_m=TX + 1e-4)
    phased = OB.wust_prepare(shifted, _rf_cfg(phase_correction="homogeneous-tof"))
    assert np.array_equal(phased["frequency_data_norm"],
                          np.sum(np.abs(phased["Y"].astype(np.complex128)) ** 2, axis=(0, 1)))
    assert phased["preparation"]["geometry_reference_speed_mps"] == 1500
    assert not np.allclose(phased["Y"], OB.wust_prepare(shifted, _rf_cfg())["Y"])
    # a masked channel may hold non-finite RF; a valid one may not
    bad = dict(rf, pressure=pressure.copy())
    bad["pressure"][:, 0, 0] = np.nan
    mask = np.ones((2, 4), bool)
    mask[0, 0] = False
    masked = OB.wust_prepare(bad, _rf_cfg(mask=mask))
    assert np.all(masked["Y"][0, 0, :] == 0) and not masked["mask"][0, 0, :].any()
    with pytest.raises(WustError):
        OB.wust_prepare(bad, _rf_cfg())
    with pytest.raises(WustError):
        OB.wust_prepare(rf, dict(_rf_cfg(), extra=1))
