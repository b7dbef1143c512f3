"""Configuration checks of wust_reconstruct (host logic only, no GPU).

Cases follow the rejections and schedule bookkeeping of
Runtime/tests/test_runtime.m:62-167 and prepare_reconstruction
(Runtime/matlab/wust_reconstruct.m:128-207).
"""
import numpy as np
import pytest

from waveforminversionust_amd._native import WustError
from waveforminversionust_amd import reconstruct as R


def _cfg(case):
    mask = np.ones((case.ny, case.nx), bool)
    mask[:5, :] = mask[-5:, :] = False
    mask[:, :5] = mask[:, -5:] = False
    return dict(case.cfg, step_damping=0.25, max_update_mps=12.0, bounds_mps=[1300.0, 1800.0], filter_cutoff=0.0,
                filter_order=4, frequency_schedule_hz=[1e5, 1e5, 1.2e5, 1.2e5], update_mask=mask)


def _reject(obs, init, cfg):
    with pytest.raises(WustError) as ei:
        R.prepare_reconstruction(obs, init, cfg)
    assert ei.value.identifier == "WUST:InvalidInput"


def test_schedule_bookkeeping(runtime_case, runtime_obs):
    init = 1500 * np.ones((31, 31))
    cfg = _cfg(runtime_case)
    p = R.prepare_reconstruction(runtime_obs, init, cfg)
    assert p["schedule"] == [0, 0, 1, 1] and p["planned_schedule_length"] == 4
    assert p["save_history"] is False and p["evaluate_final_residual"] is False
    p = R.prepare_reconstruction(runtime_obs, init, dict(cfg, max_updates=1))
    assert p["schedule"] == [0] and p["planned_schedule_length"] == 4
    p = R.prepare_reconstruction(runtime_obs, init, dict(cfg, max_updates=0))
    assert p["schedule"] == []
    default = {k: v for k, v in cfg.items() if k != "frequency_schedule_hz"}
    p = R.prepare_reconstruction(runtime_obs, init, default)
    assert p["schedule"] == [0, 1] and p["planned_schedule_length"] == 2
    assert cfg["update_mask"].dtype == bool and "schedule" not in cfg    # the caller's cfg is untouched


def test_config_rejections(runtime_case, runtime_obs):
    init = 1500 * np.ones((31, 31))
    cfg = _cfg(runtime_case)
    _reject(runtime_obs, init, {k: v for k, v in cfg.items() if k != "filter_order"})
    _reject(runtime_obs, init, dict(cfg, unknown_field=1))
    _reject(runtime_obs, init, dict(cfg, step_damping=0.0))
    _reject(runtime_obs, init, dict(cfg, filter_order=2.5))
    _reject(runtime_obs, init, dict(cfg, filter_cutoff=-1.0))
    _reject(runtime_obs, init, dict(cfg, bounds_mps=[1800.0, 1300.0]))
    _reject(runtime_obs, init, dict(cfg, backend="tpu"))
    _reject(runtime_obs, init, dict(cfg, wavenumber="kwave-ldr9"))                   # no dispersion facts
    _reject(runtime_obs, init, dict(cfg, frequency_schedule_hz=[1.1e5]))            # not an observed frequency
    _reject(runtime_obs, init, dict(cfg, max_updates=1.5))
    _reject(runtime_obs, init, dict(cfg, save_history=1))
    _reject(runtime_obs, 1500 * np.ones((31, 30)), cfg)                           # no resize
    _reject(runtime_obs, 1200 * np.ones((31, 31)), cfg)                           # outside the speed bounds
    pml = cfg["update_mask"].copy()
    pml[3, 15] = True                                                              # inside the 4 mm PML
    _reject(runtime_obs, init, dict(cfg, update_mask=pml))


def test_receiver_counts_and_boundary_sampling(runtime_case, runtime_obs):
    init = 1500 * np.ones((31, 31))
    cfg = _cfg(runtime_case)
    one = dict(cfg, frequency_schedule_hz=[1e5])
    # one valid receiver on TX 1 at the first frequency cannot fit a source scale
    mask = np.ones_like(runtime_obs["mask"])
    mask[0, 1:, 0] = False
    _reject(dict(runtime_obs, mask=mask), init, cfg)
    # the second frequency unused by the schedule is not checked, unless the final residual is asked for
    mask = np.ones_like(runtime_obs["mask"])
    mask[:, :, 1] = False
    bad = dict(runtime_obs, mask=mask)
    assert R.prepare_reconstruction(bad, init, one)["schedule"] == [0]
    _reject(bad, init, dict(one, evaluate_final_residual=True))
    _reject(bad, init, dict(one, evaluate_final_residual=True, max_updates=0))
    # a TX on the Dirichlet column is rejected; one cell in is accepted
    ny = 31
    tx = runtime_obs["tx_index"].copy()
    row = (tx[0] - 1) % ny
    _reject(dict(runtime_obs, tx_index=np.r_[row + 1, tx[1:]]), init, cfg)
    R.prepare_reconstruction(dict(runtime_obs, tx_index=np.r_[row + 1 + ny, tx[1:]]), init, one)
    # an RX on the boundary is allowed while no evaluated frequency uses it
    rx = runtime_obs["rx_index"].copy()
    row = (rx[0] - 1) % ny
    mask = np.ones_like(runtime_obs["mask"])
    mask[:, 0, 0] = False
    boundary = dict(runtime_obs, rx_index=np.r_[row + 1, rx[1:]], mask=mask)
    R.prepare_reconstruction(boundary, init, dict(one, max_updates=1))
    _reject(boundary, init, dict(one, evaluate_final_residual=True))


def test_directional_curvature_rejects_other_receivers():
    """directional_curvature checks rx_index against the receivers of the retained state before any
    device work (WUST:InvalidInput), and accepts the same set in another integer type."""
    import pytest as _pt

    from waveforminversionust_amd import helmholtz as G
    from waveforminversionust_amd._native import WustError

    class _NoSolver:
        def __getattr__(self, name):
            raise AssertionError("the check must come before any device work")

    state = G.OracleState(solver=_NoSolver(), scales=np.zeros(2, complex), forward_rhs=2, adjoint_rhs=2,
                          extra=dict(rx_index=np.array([5, 9, 13], np.int64)))
    with _pt.raises(WustError) as ei:
        G.directional_curvature(state, np.zeros((4, 4)), [5, 9, 14])
    assert ei.value.identifier == "WUST:InvalidInput"
    with _pt.raises(WustError):
        G.directional_curvature(state, np.zeros((4, 4)), [5, 9])
    with _pt.raises(AssertionError):   # same receivers (int32): passes the check, reaches the solver
        G.directional_curvature(state, np.zeros((4, 4)), np.array([5, 9, 13], np.int32))
