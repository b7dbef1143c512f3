"""The wust.reconstruction artifact (waveforminversionust_amd/artifact.py) against the layout of
Runtime/matlab/wust_write_artifact.m / wust_read_artifact.m."""
import json

import numpy as np
import pytest

from waveforminversionust_amd import artifact as A
from waveforminversionust_amd._native import WustError


def _result(history=True):
    rng = np.random.default_rng(1)
    c = 1500 + rng.standard_normal((7, 9))
    records = [dict(step=k + 1, frequency_hz=f, loss_before_update=1.0 / (k + 1), alpha=0.5, beta=0.0,
                    relative_residual_before_update=float("nan") if k == 0 else 0.1,
                    relative_residual_status="zero_observation_energy" if k == 0 else "available",
                    stage_id=fi, loss_model_step=k)
               for k, (fi, f) in enumerate([(0, 1e5), (0, 1e5), (1, 1.2e5)])]
    cfg = dict(backend="gpu", bounds_mps=[1400.0, 1600.0], update_mask=np.ones((7, 9), bool), schedule=[0, 0, 1],
               device="cuda:0", planned_schedule_length=3)
    r = dict(schema="wust.reconstruction", schema_version=2, c_mps=c, records=records, config=cfg,
             executed_schedule=[0, 0, 1], completed_updates=3, planned_schedule_length=3,
             executed_frequencies_hz=[1e5, 1e5, 1.2e5], canonical_frequencies_hz=[1e5, 1.2e5],
             completion=dict(category="completion", reason="schedule_complete", converged=False),
             final_data_residual=None, wall_seconds=1.25)
    if history:
        r["history_mps"] = rng.standard_normal((7, 9, 4)).astype(np.float32)
    return r


@pytest.mark.parametrize("history", [True, False])
def test_manifest_layout_and_round_trip(tmp_path, history):
    r = _result(history)
    path = tmp_path / "manifest.json"
    A.wust_write_artifact(str(path), r)
    doc = json.loads(path.read_text())
    assert doc["schema"] == "wust.reconstruction" and doc["schema_version"] == 2
    assert doc["arrays_file"] == "arrays.npz" and (tmp_path / "arrays.npz").exists()
    assert doc["arrays"]["c_mps"] == dict(axes="y,x", shape=[7, 9], dtype="float64", units="m/s", dataset="/c_mps")
    if history:
        assert doc["arrays"]["history_mps"] == dict(axes="y,x,model_step", shape=[7, 9, 4], dtype="float32",
                                                    units="m/s", dataset="/history_mps")
    else:
        assert "history_mps" not in doc["arrays"]
    meta = doc["metadata"]
    assert "c_mps" not in meta and "schema" not in meta and "update_mask" not in meta["config"]
    assert meta["config"]["schedule"] == [1, 1, 2] and meta["executed_schedule"] == [1, 1, 2]   # MATLAB indices
    assert [rec["stage_id"] for rec in meta["records"]] == [1, 1, 2]
    assert meta["records"][0]["relative_residual_before_update"] is None                      # NaN -> null
    assert meta["final_data_residual"] is None and meta["completion"]["reason"] == "schedule_complete"
    back = A.wust_read_artifact(str(path), "wust.reconstruction")
    assert np.array_equal(back["c_mps"], r["c_mps"]) and back["c_mps"].dtype == np.float64
    if history:
        assert np.array_equal(back["history_mps"], r["history_mps"]) and back["history_mps"].dtype == np.float32
    assert back["schema"] == "wust.reconstruction" and back["records"][2]["frequency_hz"] == 1.2e5


def test_rejections(tmp_path):
    r = _result()
    path = tmp_path / "m.json"
    A.wust_write_artifact(str(path), r)
    with pytest.raises(WustError) as ei:            # the run directory already holds an artifact
        A.wust_write_artifact(str(tmp_path / "other.json"), r)
    assert ei.value.identifier == "WUST:InvalidInput"
    bad = dict(r)
    del bad["c_mps"]
    with pytest.raises(WustError):
        A.wust_write_artifact(str(tmp_path / "sub" / "m.json"), bad)
    with pytest.raises(WustError):
        A.wust_read_artifact(str(path), "wust.measurements")
    doc = json.loads(path.read_text())
    doc["arrays"]["c_mps"]["units"] = "km/s"
    (tmp_path / "m2.json").write_text(json.dumps(doc))
    with pytest.raises(WustError):
        A.wust_read_artifact(str(tmp_path / "m2.json"))
    bad = dict(r, c_mps=r["c_mps"].astype(np.float32))   # c_mps is float64 by contract
    d2 = tmp_path / "d2"
    d2.mkdir()
    with pytest.raises(WustError):
        A.wust_write_artifact(str(d2 / "m.json"), bad)


@pytest.mark.gpu
def test_reconstruction_artifact_on_gpu(tmp_path, runtime_case, runtime_obs):
    """A wust_reconstruct result (device tensors, history saved) written and read back unchanged."""
    pytest.importorskip("torch")
    from waveforminversionust_amd import reconstruct as R
    case = runtime_case
    X, Y = np.meshgrid(case.x_m, case.y_m)
    cfg = dict(case.cfg, step_damping=0.25, max_update_mps=12.0, bounds_mps=[1400.0, 1600.0], filter_cutoff=0.0,
               filter_order=4, update_mask=X ** 2 + Y ** 2 < 8e-3 ** 2, save_history=True,
               evaluate_final_residual=True)
    res = R.wust_reconstruct(runtime_obs, 1500 * np.ones((case.ny, case.nx)), cfg)
    path = tmp_path / "manifest.json"
    A.wust_write_artifact(str(path), res)
    back = A.wust_read_artifact(str(path), "wust.reconstruction")
    assert np.array_equal(back["c_mps"], res["c_mps"].cpu().numpy())
    assert np.array_equal(back["history_mps"], res["history_mps"].cpu().numpy())
    assert len(back["records"]) == len(res["records"]) == 2
    assert back["executed_schedule"] == [i + 1 for i in res["executed_schedule"]]
    assert back["final_data_residual"]["oracle_calls"] == 2
    for got, want in zip(back["records"], res["records"]):
        assert got["loss_before_update"] == want["loss_before_update"] and got["stage_id"] == want["stage_id"] + 1
