"""Observation / geometry ingestion (host logic), after Runtime/tests/test_ingest_frequency.m:1-94.

The expected values are that test's own: the asymmetric 7x9 grid, the snapped
indices [17, 40] / [17, 44, 27, 53], the conjugation of a +1-sign input, the
frequency sort, masked values zeroed with the mask kept, the data norms of a
constant input, and its rejections.
"""
import numpy as np
import pytest

from waveforminversionust_amd import observations as O
from waveforminversionust_amd._native import WustError

GRID = dict(shape_yx=[7, 9], origin_yx_m=[.01, -.02], spacing_yx_m=[.001, .002], origin_kind="pixel_edge")
X = -.02 + (np.arange(9) + .5) * .002
Y = .01 + (np.arange(7) + .5) * .001
TX = np.array([[X[2], Y[2]], [X[5], Y[4]]])
RX = np.array([TX[0], [X[6], Y[1]], [X[3], Y[5]], [X[7], Y[3]]])


def _input():
    # MATLAB reshape(1:24, 3, 2, 4) is column-major: element (i, j, k) = 1 + i + 3 j + 6 k
    base = np.arange(1, 25, dtype=np.float32).reshape(4, 2, 3).transpose(2, 1, 0)
    p = (base + 1j * (base + 100)).astype(np.complex64)            # [frequency, tx, rx] = [3, 2, 4]
    mask = np.ones(p.shape, bool)
    mask[0, 1, 2] = False
    p[0, 1, 2] = np.nan
    p[1, 0, 0] = 0
    return dict(schema="wust.frequency_input", schema_version=2, grid=GRID, tx_xy_m=TX, rx_xy_m=RX,
                frequencies_hz=[3e5, 1e5, 2e5], pressure=p, pressure_axes="frequency,tx,rx", mask=mask,
                mask_axes="frequency,tx,rx", fourier_sign=1, real_pressure=True, pressure_type="total_pressure",
                data_units="instrument_units", spectrum_normalization="dtft_dt", measurement_provenance="fixture")


def _reject(value, ident="WUST:InvalidInput"):
    with pytest.raises(WustError) as ei:
        O.wust_prepare_observations(value)
    assert ei.value.identifier == ident


def _wire(obs):
    """The canonical record as an artifact carries it: derived coordinates and statistics dropped."""
    derived = {"x_m", "y_m", "tx_snapped_xy_m", "rx_snapped_xy_m", "shape_yx", "n_tx", "n_rx",
               "valid_receiver_counts", "frequency_data_norm"}
    return {k: v for k, v in obs.items() if k not in derived}


def test_asymmetric_ingestion():
    inp = _input()
    obs = O.wust_prepare_observations(inp)
    p, mask = inp["pressure"], inp["mask"]
    expected = p.copy()
    expected[~mask] = 0
    expected = np.conj(expected[[1, 2, 0]]).transpose(1, 2, 0)
    assert np.array_equal(obs["Y"], expected)
    assert np.array_equal(obs["mask"], mask[[1, 2, 0]].transpose(1, 2, 0))
    assert obs["mask"][0, 0, 0] and obs["Y"][0, 0, 0] == 0
    assert list(obs["tx_index"]) == [17, 40] and list(obs["rx_index"]) == [17, 44, 27, 53]
    assert obs["Y"].dtype == np.complex64 and list(obs["frequencies_hz"]) == [1e5, 2e5, 3e5]
    assert np.max(np.abs(obs["x_m"] - X)) < np.finfo(float).eps and np.max(np.abs(obs["y_m"] - Y)) < np.finfo(float).eps
    prep = obs["preparation"]
    assert prep["conversion"] == "conjugate" and list(prep["canonical_to_input_permutation_1based"]) == [2, 3, 1]
    assert prep["input_precision"] == "single" and prep["measurement_provenance"] == "fixture"
    assert obs["valid_receiver_counts"].shape == (2, 3) and obs["valid_receiver_counts"][1, 2] == 3
    # the snapped coordinates are the grid nodes of the indices
    ti = obs["tx_index"] - 1
    assert np.array_equal(obs["tx_snapped_xy_m"], np.stack([X[ti // 7], Y[ti % 7]], axis=1))


def test_canonical_round_trip_and_rejections():
    obs = O.wust_prepare_observations(_input())
    wire = _wire(obs)
    restored = O.wust_prepare_observations(wire)
    assert np.array_equal(restored["Y"], obs["Y"]) and np.array_equal(restored["mask"], obs["mask"])
    assert tuple(restored["shape_yx"]) == (7, 9) and restored["n_tx"] == 2 and restored["n_rx"] == 4
    assert np.array_equal(restored["valid_receiver_counts"], obs["valid_receiver_counts"])
    assert np.array_equal(restored["frequency_data_norm"], obs["frequency_data_norm"])
    assert restored["preparation"] is obs["preparation"]
    assert restored["spectrum_normalization"] == "dtft_dt"
    unrecorded = O.wust_prepare_observations({k: v for k, v in wire.items() if k != "preparation"})
    assert np.array_equal(unrecorded["Y"], obs["Y"])
    assert unrecorded["preparation"]["measurement_provenance"] == "unknown"
    _reject(dict(wire, frequencies_hz=obs["frequencies_hz"][::-1]))
    _reject(dict(wire, fourier_sign=1))
    _reject(dict(wire, tx_index=obs["tx_index"][::-1]))
    changed = dict(wire, Y=np.ones_like(obs["Y"]), mask=np.ones_like(obs["mask"]))
    changed = O.wust_prepare_observations(changed)
    assert list(changed["frequency_data_norm"]) == [8, 8, 8]
    m = np.ones_like(obs["mask"])
    m[0, 1:, 0] = False
    changed = O.wust_prepare_observations(dict(wire, Y=np.zeros_like(obs["Y"]), mask=m))
    assert np.all(changed["frequency_data_norm"] == 0) and changed["valid_receiver_counts"][0, 0] == 1


def test_reingestion_and_input_rejections():
    inp = _input()
    obs = O.wust_prepare_observations(inp)
    again = dict(inp, pressure=obs["Y"].transpose(2, 0, 1), mask=obs["mask"].transpose(2, 0, 1),
                 frequencies_hz=obs["frequencies_hz"], fourier_sign=-1)
    assert np.array_equal(O.wust_prepare_observations(again)["Y"], obs["Y"])
    _reject(dict(inp, fourier_sign=0))
    _reject(dict(inp, tx_xy_m=np.array([TX[0], TX[0] + 1e-6])))              # two TX on one node
    p = inp["pressure"].copy()
    p[1, 0, 0] = np.inf
    _reject(dict(inp, pressure=p))
    _reject(dict(inp, real_pressure=False))                                   # +1 sign needs real RF
    _reject(dict(inp, frequencies_hz=[1e5, 1e5, 2e5]))
    _reject(dict(inp, pressure_type="scattered"))
    _reject(dict(inp, data_units="Pa*s", spectrum_normalization="discrete_sum"))
    _reject(dict(inp, pressure_axes="rx,tx,frequency"))
    _reject(dict(inp, mask=np.zeros_like(inp["mask"])))                       # empty observations
    _reject(dict(inp, schema_version=1), "WUST:IncompatibleRequest")
    _reject(dict(inp, grid=dict(GRID, origin_kind="pixel_centre")))
    _reject(dict(inp, rx_xy_m=np.array([[1.0, 1.0]])))                        # outside the grid
    good = dict(inp, mask=np.ones((2, 4), bool), mask_axes="tx,rx")
    p = inp["pressure"].copy()
    p[0, 1, 2] = 1
    good["pressure"] = p
    assert O.wust_prepare_observations(good)["mask"].all()
    _reject(dict(good, mask=np.ones((4, 2), bool)))


def test_geometry_matches_workload_snapping():
    from waveforminversionust_amd import workloads as W
    case = W.runtime_test_case()
    grid = dict(shape_yx=[31, 31], origin_yx_m=[-.0155, -.0155], spacing_yx_m=[.001, .001], origin_kind="pixel_edge")
    geo = O.wust_geometry(grid, case.tx_xy, case.rx_xy)
    assert np.array_equal(geo["tx_index"], case.tx_index) and np.array_equal(geo["rx_index"], case.rx_index)
    assert np.allclose(geo["x_m"], case.x_m, rtol=0, atol=1e-15)
