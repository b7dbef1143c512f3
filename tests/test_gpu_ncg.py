"""GPU parity of the NCG step (C ABI via waveforminversionust_amd.reconstruct) against the FP64 oracle.

Elementwise pieces and the filter are compared at 1e-12 relative on random
grids of odd and even sizes (the FFT bin conventions differ between them);
the reference's own unit-test values (Runtime/tests/test_runtime.m:170-190)
are checked exactly; the full reconstruction loop is compared with the
oracle's loop driven by the FP64 Helmholtz oracle on the 31x31 fixture, and
checked against the bookkeeping assertions of Runtime/tests/test_runtime.m:62-114.
"""
import numpy as np
import pytest
import torch

from oracle import helmholtz_fp64 as H
from oracle import ncg_fp64 as O
from waveforminversionust_amd._native import WustError

pytestmark = pytest.mark.gpu

R = pytest.importorskip("waveforminversionust_amd.reconstruct")


def _np(t):
    return t.detach().cpu().numpy() if torch.is_tensor(t) else np.asarray(t)


def _mask(ny, nx, border=5):
    m = np.ones((ny, nx), bool)
    m[:border, :] = m[-border:, :] = False
    m[:, :border] = m[:, -border:] = False
    return m


def _cfg(case, **kw):
    cfg = dict(case.cfg, step_damping=0.25, max_update_mps=12.0, bounds_mps=[1300.0, 1800.0], filter_cutoff=0.0,
               filter_order=4, frequency_schedule_hz=[1e5, 1e5, 1.2e5, 1.2e5], update_mask=_mask(31, 31))
    cfg.update(kw)
    return cfg


def _err_id(fn, *a):
    with pytest.raises(WustError) as ei:
        fn(*a)
    return ei.value.identifier


@pytest.mark.parametrize("ny,nx", [(31, 31), (30, 33), (33, 30), (8, 9)])
def test_roi_gradient_and_filter_match_oracle(ny, nx):
    rng = np.random.default_rng(ny * 100 + nx)
    x = np.arange(nx) * 1.0e-3
    y = np.arange(ny) * 0.8e-3
    g = rng.standard_normal((ny, nx))
    mask = _mask(ny, nx, border=2)
    for cutoff, order in [(0.5, 4), (0.9, 2), (0.0, 4)]:
        roi, z = R.wust_roi_gradient(g, mask, x, y, 1500.0, 1e5, cutoff, order)
        w_roi, w_z = O.wust_roi_gradient(g, mask, x, y, 1500.0, 1e5, cutoff, order)
        assert np.array_equal(_np(roi), w_roi)
        assert np.max(np.abs(_np(z) - w_z)) <= 1e-12 * np.max(np.abs(g))
        assert np.all(_np(z)[~mask] == 0)
        f = R.ringingRemovalFilt(x, y, g, 1500.0, 1e5, cutoff, order)
        w = O.ringing_removal_filt(x, y, g, 1500.0, 1e5, cutoff, order)
        assert np.max(np.abs(_np(f) - w)) <= 1e-12 * np.max(np.abs(g))
    assert np.array_equal(_np(R.ringingRemovalFilt(x, y, g, 1500.0, 1e5, 0, 4)), g)


def test_roi_gradient_reference_values():
    # exterior values never leak through the filter; -z descends (Runtime/tests/test_runtime.m:182-188)
    x = np.arange(-15, 16) * 1e-3
    mask = _mask(31, 31)
    g = np.ones((31, 31))
    g[~mask] = -1e10
    roi, z = R.wust_roi_gradient(g, mask, x, x, 1500, 1e5, .5, 4)
    g[~mask] = 1e20
    _, z2 = R.wust_roi_gradient(g, mask, x, x, 1500, 1e5, .5, 4)
    assert torch.equal(z, z2) and bool(torch.all(z[torch.as_tensor(~mask)] == 0))
    assert float(torch.sum(roi * -z)) < 0
    assert _err_id(R.wust_roi_gradient, g, mask, x, x, 1500, 1e5, -1.0, 4) == "WUST:InvalidInput"


def test_direction_reference_values():
    g, z, pz = np.array([1., 2.]), np.array([2., 1.]), np.array([1., .5])
    d, beta, gz = R.wust_ncg_direction(g, z, pz, 1.5, np.array([-1., -2.]), True)
    assert abs(beta - 4 / 3) < 1e-14 and gz == 4 and float(np.sum(g * _np(d))) < 0
    d, beta, _ = R.wust_ncg_direction(g, z, pz, 1.5, np.array([100., 100.]), True)
    assert beta == 0 and np.array_equal(_np(d), [-2., -1.])
    _, beta, _ = R.wust_ncg_direction(g, z, pz, 1.5, np.array([-1., -2.]), False)
    assert beta == 0
    assert _err_id(R.wust_ncg_direction, g, np.array([-2., -1.]), np.zeros(2), 0, np.zeros(2), False) \
        == "WUST:NumericalFailure"


@pytest.mark.parametrize("case", ["pr", "fr_clip", "restart", "new_frequency", "zero_prev"])
def test_direction_matches_oracle(case):
    rng = np.random.default_rng(7)
    shape = (31, 30)
    g = rng.standard_normal(shape)
    z = g + 0.3 * rng.standard_normal(shape)
    pz = 0.1 * z + 0.5 * rng.standard_normal(shape)
    pd = -pz + 0.1 * rng.standard_normal(shape)
    pgz, same = float(np.sum(g * z)) * 0.8, True
    if case == "fr_clip":
        pz, pd = -2 * z, -z                                # PR above FR: beta = FR
    elif case == "restart":
        pd = 50 * z                                        # the previous direction ascends
    elif case == "new_frequency":
        same = False
    elif case == "zero_prev":
        pgz = 0.0
    d, beta, gz, slope = R.wust_ncg_direction(g, z, pz, pgz, pd, same, return_slope=True)
    wd, wbeta, wgz = O.wust_ncg_direction(g, z, pz, pgz, pd, same)
    assert abs(gz - wgz) <= 1e-12 * abs(wgz)
    assert abs(beta - wbeta) <= 1e-12 * max(abs(wbeta), 1e-300)
    assert np.max(np.abs(_np(d) - wd)) <= 1e-12 * np.max(np.abs(wd))
    assert abs(slope - float(np.sum(g * wd))) <= 1e-12 * float(np.sum(np.abs(g * wd)))
    if case == "fr_clip":
        assert wbeta == pytest.approx(wgz / pgz, rel=1e-14)
    if case in ("restart", "new_frequency", "zero_prev"):
        assert beta == 0 and np.array_equal(_np(d), -z)
    if case == "pr":
        assert 0 < wbeta < wgz / pgz


def test_step_length():
    assert R.linearized_step_length(-2.0, 4.0) == 0.5
    assert R.linearized_step_length(3.0, 4.0) == 0.0
    assert R.linearized_step_length(0.0, 0.0) == 0.0
    for bad in [(1.0, 0.0), (np.nan, 1.0), (1.0, -1.0)]:
        assert _err_id(R.linearized_step_length, *bad) == "WUST:NumericalFailure"


def test_projection_matches_oracle_and_reference_values(runtime_case):
    cfg = _cfg(runtime_case)
    mask = cfg["update_mask"]
    init = 1500 * np.ones((31, 31))
    up = _np(R.wust_project_slowness_update(init, -np.ones((31, 31)), 1.0, cfg))
    assert np.all(up[mask] == 1512) and np.array_equal(up[~mask], init[~mask])
    assert _err_id(R.wust_project_slowness_update, init, np.full((31, 31), np.nan), 1.0, cfg) \
        == "WUST:NumericalFailure"
    rng = np.random.default_rng(3)
    c = 1450 + 100 * rng.random((31, 31))
    d = 2e-5 * rng.standard_normal((31, 31))               # some nodes clip at +-12 m/s, some not
    got, rel = R.wust_project_slowness_update(c, d, 0.7, cfg, return_relative_update=True)
    want = O.wust_project_slowness_update(c, d, 0.7, cfg)
    assert np.max(np.abs(_np(got) - want)) <= 1e-12 * 1500
    s0, s1 = 1 / c[mask], 1 / want[mask]
    assert rel == pytest.approx(np.linalg.norm(s1 - s0) / np.linalg.norm(s0), rel=1e-10)
    clipped = np.abs(want - c) >= 12 - 1e-9
    assert clipped[mask].any() and (~clipped[mask]).any()


def _gpu_vs_oracle(case, obs, cfg, schedule):
    init = 1500 * np.ones((31, 31))
    res = R.wust_reconstruct(obs, init, cfg)
    ocfg = dict(cfg, schedule=schedule)
    c, rec = O.wust_reconstruct(obs, init, ocfg, H.wust_oracle, H.directional_curvature)
    return res, c, rec


@pytest.mark.parametrize("damping,cutoff,repeats", [(0.25, 0.0, 2), (0.25, 0.5, 2), (1.0, 0.5, 3)])
def test_reconstruct_matches_oracle(runtime_case, runtime_obs, damping, cutoff, repeats):
    # (0.25, 0, 2) is the reference's own configuration; (1.0, 0.5, 3) engages the conjugate term (beta > 0)
    hz = [1e5] * repeats + [1.2e5] * repeats
    cfg = _cfg(runtime_case, step_damping=damping, filter_cutoff=cutoff, frequency_schedule_hz=hz)
    res, c, rec = _gpu_vs_oracle(runtime_case, runtime_obs, cfg, [0] * repeats + [1] * repeats)
    assert len(res["records"]) == 2 * repeats
    for got, want in zip(res["records"], rec):
        assert got["loss_before_update"] == pytest.approx(want["loss_before_update"], rel=1e-8)
        assert got["gz"] == pytest.approx(want["gz"], rel=1e-7)
        assert got["curvature"] == pytest.approx(want["curvature"], rel=1e-7)
        assert got["alpha"] == pytest.approx(want["alpha"], rel=1e-7)
        assert got["beta"] == pytest.approx(want["beta"], rel=1e-6, abs=1e-9)
        assert got["relative_slowness_update"] == pytest.approx(want["relative_slowness_update"], rel=1e-6)
        assert got["alpha"] == max(0.0, -got["slope"] / got["curvature"])
    assert np.max(np.abs(_np(res["c_mps"]) - c)) <= 1e-6
    assert rec[-1]["loss_before_update"] < rec[0]["loss_before_update"]
    if repeats == 3:
        assert max(r["beta"] for r in rec) > 0.1


def test_reconstruct_bookkeeping(runtime_case, runtime_obs):
    """Runtime/tests/test_runtime.m:62-114 on the GPU engine."""
    init = 1500 * np.ones((31, 31))
    cfg = _cfg(runtime_case)
    G = pytest.importorskip("waveforminversionust_amd.helmholtz")
    loss0 = G.wust_oracle(init, runtime_obs, 0, cfg, nargout=1)
    res = R.wust_reconstruct(runtime_obs, init, cfg)
    assert bool(torch.all(torch.isfinite(res["c_mps"])))
    assert res["completion"]["reason"] == "schedule_complete" and not res["completion"]["converged"]
    assert res["completed_updates"] == 4 and res["planned_schedule_length"] == 4
    assert res["final_data_residual"] is None and res["records"][-1]["loss_model_step"] == 3
    assert G.wust_oracle(res["c_mps"], runtime_obs, 0, cfg, nargout=1) < loss0
    short = R.wust_reconstruct(runtime_obs, init, dict(cfg, max_updates=1, save_history=True,
                                                       evaluate_final_residual=True))
    assert short["completion"]["reason"] == "caller_truncated_schedule"
    assert short["planned_schedule_length"] == 4 and short["completed_updates"] == 1
    assert tuple(short["history_mps"].shape) == (31, 31, 2)
    fdr = short["final_data_residual"]
    assert fdr["oracle_calls"] == 2 and fdr["forward_rhs"] == 3 * 2 and fdr["adjoint_rhs"] == 0
    assert np.isfinite(fdr["relative"])
    zero = R.wust_reconstruct(runtime_obs, init, dict(cfg, max_updates=0))
    assert np.array_equal(_np(zero["c_mps"]), init) and zero["records"] == []
    assert zero["completed_updates"] == 0 and zero["completion"]["reason"] == "zero_updates"
    zero_data = dict(runtime_obs, Y=np.zeros_like(runtime_obs["Y"]))
    res = R.wust_reconstruct(zero_data, init, dict(cfg, max_updates=1, evaluate_final_residual=True))
    assert np.array_equal(_np(res["c_mps"]), init)
    r = res["records"][0]
    assert r["alpha"] == 0 and r["absolute_residual_before_update"] == 0
    assert np.isnan(r["relative_residual_before_update"])
    assert r["relative_residual_status"] == "zero_observation_energy"
    assert res["final_data_residual"]["absolute"] == 0 and np.isnan(res["final_data_residual"]["relative"])
    default = R.wust_reconstruct(runtime_obs, init, {k: v for k, v in cfg.items() if k != "frequency_schedule_hz"})
    assert default["planned_schedule_length"] == 2 and default["executed_frequencies_hz"] == [1e5, 1.2e5]
    huge = dict(runtime_obs, Y=np.full_like(runtime_obs["Y"], np.finfo(np.float64).max / 2))
    assert _err_id(R.wust_reconstruct, huge, init, cfg) == "WUST:NumericalFailure"


def test_device_resident_observations(runtime_case, runtime_obs):
    """Observations placed once in HBM give bit-identical oracle results and are read in place (with
    the history-free factor, so that the second evaluation's factor is not seeded from the first)."""
    from waveforminversionust_amd import helmholtz as G
    from waveforminversionust_amd.observations import observations_to_device
    init = 1500 * np.ones((31, 31))
    dev = observations_to_device(runtime_obs)
    fixed = dict(runtime_case.cfg, factor="fixed")
    for fi in range(2):
        l1, g1 = G.wust_oracle(init, runtime_obs, fi, fixed)
        l2, g2 = G.wust_oracle(init, dev, fi, fixed)
        assert l1 == l2 and torch.equal(g1, g2)
        Yd, Md = G._obs_slice(dev, fi, g2.device)
        assert Yd.data_ptr() == dev["Y"][:, :, fi].data_ptr() and Md.data_ptr() == dev["mask"][:, :, fi].data_ptr()
    cfg = _cfg(runtime_case, max_updates=2, factor="fixed")
    a = R.wust_reconstruct(runtime_obs, init, cfg)
    b = R.wust_reconstruct(dev, init, cfg)
    assert torch.equal(a["c_mps"], b["c_mps"])
