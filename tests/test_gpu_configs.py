"""The continuation stages and the 1088^2 reconstruction grid on the GPU (needs an MI355X).

Cases:
* continuation stages (eight frequencies 0.3-1.0 MHz, each on its own grid at 6 points per wavelength
  of 1400 m/s, 24-cell PML, 256 TX on the 70 mm ring, two-disc model): the smallest (0.3 MHz), a
  middle one (0.6 MHz) and the largest (1.0 MHz, 717^2 nodes, npad 736: the 32-row sweep tiles);
* the reconstruction grid: 1024^2 interior + 32-cell PML (1088^2 nodes), dx 0.125 mm, 512 TX on the
  60 mm ring, breast phantom, evaluated in source batches of 64 (cfg.source_batch_size), at the
  schedule's highest frequency (1.2 MHz).

The FP64 sparse LU cannot run per test at these sizes, so the checks are ones the mathematics fixes
(as at the metric size): the nine coefficient arrays against the oracle's FP64 assembly (<= 1e-13);
the refined forward solve's FP64 residual of every point source by an independent stencil apply
(<= the refinement tolerance); forward / adjoint duality over all transmitter x receiver pairs
(<= 5e-8); the gradient against Richardson-extrapolated central differences along a smooth direction
(<= 1e-5); and, at 1088^2, batch invariance of the loss (batches of 64 against one batch of 512,
Runtime/tests/test_runtime.m:57-61) and one wust_reconstruct step whose update obeys the projection
bounds and lowers the misfit.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import helmholtz_fp64 as H  # noqa: E402
from waveforminversionust_amd import helmholtz as G  # noqa: E402
from waveforminversionust_amd import reconstruct as Rc  # noqa: E402
from waveforminversionust_amd import workloads as W  # noqa: E402

from test_gpu_metric_size import apply_H  # noqa: E402

OPTS = dict(backend="gpu", stencil_bounds=[1400, 1700])
CASES = ["stage0.3", "stage0.6", "stage1.0", "breast1088"]


def _case(name):
    if name == "breast1088":
        case = W.breast_case()
        return case, 1.2e6, 64
    stages = W.continuation_cases()
    f = float(name[len("stage"):]) * 1e6
    case = min(stages, key=lambda c: abs(c.frequencies_hz[0] - f))
    return case, case.frequencies_hz[0], case.n_tx


@pytest.fixture(scope="module", params=CASES)
def big(request):
    case, f, batch = _case(request.param)
    s = G.HelmholtzSolver(case.x_m, case.y_m, case.c_true, None, f, -1, case.cfg["pml_strength"], case.cfg["pml_m"],
                          dict(OPTS, mode="refined", tol=1e-10))
    yield request.param, case, f, batch, s
    s.close()
    torch.cuda.empty_cache()


def _impulses(case, index, amplitudes=None):
    b = torch.zeros((case.nx * case.ny, index.size), dtype=torch.complex128, device="cuda")
    cols = torch.arange(index.size, device="cuda")
    b[torch.as_tensor(index - 1, device="cuda"), cols] = (1.0 if amplitudes is None else amplitudes)
    return b.reshape(case.nx, case.ny, -1).permute(1, 0, 2)


def test_grid_shape(big):
    name, case, f, batch, s = big
    info = s.info()
    print(f"{name}: {case.ny}x{case.nx} nodes, npad {info['npad']}, {case.n_tx} TX, f {f / 1e6:.2f} MHz, factor "
          f"{info['factor_ms']:.1f} ms, fallback {info['factor_fallback']}, gj residual {info['gj_residual']:.2e}")
    if name == "stage1.0":
        assert case.ny == 717 and info["npad"] == 736
    if name == "breast1088":
        assert case.ny == 1088 and case.n_tx == 512
    assert info["factor_fallback"] != 2


def test_coefficients_match_oracle(big):
    name, case, f, batch, s = big
    lo, hi = case.cfg["stencil_bounds"]
    t = H.pml_terms(case.x_m, case.y_m, case.c_true, np.zeros_like(case.c_true), f, -1, case.cfg["pml_strength"],
                    case.cfg["pml_m"])
    b, d, e = H.stencil_opt_params(lo, hi, f, t.h, t.g)
    want = H.nine_point_coefficients(t.A, t.B, t.C, t.k, b, d, e, t.h, t.g, case.nx, case.ny)
    co = s.coefficients()
    for k, v in want.items():
        got = co[k].cpu().numpy()
        assert np.max(np.abs(got - v)) <= 1e-13 * np.max(np.abs(v)), k


def test_refined_forward_residual(big):
    name, case, f, batch, s = big
    b = _impulses(case, case.tx_index)
    u, _, st = s.solve(b, False, return_stats=True)
    r = apply_H(s.coefficients(), u) - b
    per_rhs = torch.linalg.vector_norm(r, dim=(0, 1)) / torch.linalg.vector_norm(b, dim=(0, 1))
    worst = float(per_rhs.max())
    del u, r
    assert worst <= s.tol * 1.01, (worst, st)
    assert abs(worst - st["final_residual"]) <= 1e-12, (worst, st)


def test_adjoint_duality(big):
    name, case, f, batch, s = big
    rng = np.random.default_rng(31)
    amp = torch.as_tensor(rng.standard_normal(case.n_rx) + 1j * rng.standard_normal(case.n_rx), device="cuda")
    u, _ = s.solve(_impulses(case, case.tx_index), False)
    idx = torch.as_tensor(case.rx_index - 1, device="cuda")
    b_a = torch.zeros((case.nx * case.ny, case.n_rx), dtype=torch.complex128, device="cuda")
    cols = torch.arange(case.n_rx, device="cuda")
    b_a[idx, cols] = amp
    b_a[idx.roll(-1), cols] += amp / 2
    lam, _ = s.solve(b_a.reshape(case.nx, case.ny, -1).permute(1, 0, 2), True)
    s1 = lam.permute(1, 0, 2).reshape(-1, case.n_rx)[torch.as_tensor(case.tx_index - 1, device="cuda"), :]
    s2 = u.permute(1, 0, 2).reshape(-1, case.n_tx).conj().T @ b_a
    rel = float(torch.linalg.norm(s1 - s2) / torch.linalg.norm(s1))
    assert rel <= 5e-8, rel


def _obs_from_solver(case, f, s):
    u, _ = s.solve(_impulses(case, case.tx_index), False)
    Y = u.permute(1, 0, 2).reshape(-1, case.n_tx)[torch.as_tensor(case.rx_index - 1, device="cuda"), :].T
    return dict(x_m=case.x_m, y_m=case.y_m, shape_yx=(case.ny, case.nx), n_tx=case.n_tx, tx_index=case.tx_index,
                rx_index=case.rx_index, frequencies_hz=[f], Y=Y.contiguous()[:, :, None],
                mask=torch.ones((case.n_tx, case.n_rx, 1), dtype=torch.bool, device="cuda"))


def test_gradient_central_differences(big):
    name, case, f, batch, s = big
    obs = _obs_from_solver(case, f, s)
    cfg = dict(case.cfg, tol=1e-12, source_batch_size=batch)
    s0 = torch.full((case.ny, case.nx), 1 / 1500, dtype=torch.float64, device="cuda")
    loss0, g = G.wust_oracle(1 / s0, obs, 0, cfg)
    rng = np.random.default_rng(41)
    d = W.gaussian_filter2(rng.standard_normal((case.ny, case.nx)), 3e-3 / case.dx)
    d = torch.as_tensor(d / np.max(np.abs(d)), device="cuda") * s0
    an = float(torch.sum(g * d))

    def fd(h):
        lp = G.wust_oracle(1 / (s0 + h * d), obs, 0, cfg, nargout=1)
        lm = G.wust_oracle(1 / (s0 - h * d), obs, 0, cfg, nargout=1)
        return (lp - lm) / (2 * h)

    h = 1e-3
    rich = (4 * fd(h / 2) - fd(h)) / 3
    assert abs(rich - an) <= 1e-5 * abs(an), (rich, an)
    if name == "breast1088":
        # batch invariance: batches of 64 against one batch of all 512 transmitters
        l_all = G.wust_oracle(1 / s0, obs, 0, dict(cfg, source_batch_size=case.n_tx), nargout=1)
        assert abs(l_all - loss0) <= 1e-10 * abs(loss0), (l_all, loss0)


def test_reconstruct_step_1088():
    """One NCG step of wust_reconstruct on the 1088^2 grid at 0.3 MHz, 512 TX in batches of 64: the
    update stays within max_update_mps and the speed bounds, leaves the model unchanged outside the
    update mask, and lowers the misfit of that frequency."""
    case = W.breast_case()
    f = 0.3e6
    s = G.HelmholtzSolver(case.x_m, case.y_m, case.c_true, None, f, -1, case.cfg["pml_strength"], case.cfg["pml_m"],
                          dict(OPTS, tol=1e-10))
    obs = _obs_from_solver(case, f, s)
    s.close()
    X, Yg = np.meshgrid(case.x_m, case.y_m)
    mask = X ** 2 + Yg ** 2 < 55e-3 ** 2
    cfg = dict(case.cfg, step_damping=0.25, max_update_mps=12.0, bounds_mps=[1400.0, 1650.0], filter_cutoff=0.5,
               filter_order=4, update_mask=mask, frequency_schedule_hz=[f], tol=1e-10, source_batch_size=64,
               device="cuda")
    init = torch.full((case.ny, case.nx), 1500.0, dtype=torch.float64, device="cuda")
    res = Rc.wust_reconstruct(obs, init, cfg)
    rec = res["records"]
    assert len(rec) == 1
    c1 = res["c_mps"]
    dc = (c1 - init).cpu().numpy()
    assert np.all(dc[~mask] == 0)
    assert np.max(np.abs(dc)) <= 12.0 + 1e-9 and np.max(np.abs(dc)) > 0
    assert float(c1.min()) >= 1400.0 and float(c1.max()) <= 1650.0
    loss0 = rec[0]["loss_before_update"]
    loss1 = G.wust_oracle(c1, obs, 0, dict(case.cfg, tol=1e-10, source_batch_size=64), nargout=1)
    assert loss1 < loss0, (loss1, loss0)
