"""Solves and oracle evaluations on small and ragged grids (needs an MI355X).

The block sweeps and the factor have short-chain cases (one or two interior
columns: no chain steps, a single chain, a meeting block with one neighbour)
and padding cases (rows not a multiple of the 32-row page tile, RHS counts
not a multiple of 16).  Each grid is checked through the C ABI against the FP64
oracle's own operator: the refined solution's backward error in H (forward
and adjoint) and the loss / gradient of one oracle evaluation.
"""
import numpy as np
import pytest
import torch

from oracle import helmholtz_fp64 as H

pytestmark = pytest.mark.gpu

G = pytest.importorskip("waveforminversionust_amd.helmholtz")

OPTS = dict(backend="gpu", stencil_bounds=[1400, 1700])
SHAPES = [(3, 3), (3, 5), (5, 3), (4, 4), (7, 7), (9, 11), (12, 5), (33, 34), (40, 23)]


def _grid(ny, nx):
    dx = 1e-3
    x = (np.arange(nx) - (nx - 1) / 2) * dx
    y = (np.arange(ny) - (ny - 1) / 2) * dx * 1.1
    X, Y = np.meshgrid(x, y)
    c = 1500 + 25 * np.exp(-(X ** 2 + Y ** 2) / 4e-5)
    return x, y, c


@pytest.mark.parametrize("ny,nx", SHAPES)
@pytest.mark.parametrize("adjoint", [False, True])
def test_refined_solve_backward_error(ny, nx, adjoint):
    x, y, c = _grid(ny, nx)
    f, a0, L = 2.2e5, 10.0, 2e-3
    ref = H.HelmholtzFP64(x, y, c, np.zeros_like(c), f, -1, a0, L, dict(stencil_bounds=[1400, 1700]))
    s = G.HelmholtzSolver(x, y, c, None, f, -1, a0, L, dict(OPTS, mode="refined", tol=1e-13))
    R = 5
    rng = np.random.default_rng(ny * 100 + nx)
    b = np.zeros((ny, nx, R), np.complex128)
    b[1:-1, 1:-1] = rng.standard_normal((ny - 2, nx - 2, R)) + 1j * rng.standard_normal((ny - 2, nx - 2, R))
    u, _, st = s.solve(b, adjoint, return_stats=True)
    u = u.cpu().numpy()
    Hm = ref.H.conj().T if adjoint else ref.H
    interior = np.zeros((ny, nx), bool)
    interior[1:-1, 1:-1] = True
    bf = b.reshape(-1, R, order="F")
    res = (Hm @ u.reshape(-1, R, order="F") - bf).reshape(ny, nx, R, order="F")[interior]
    assert np.linalg.norm(res) <= 1e-12 * np.linalg.norm(bf)
    assert np.all(u[~interior] == 0)
    assert st["final_residual"] <= 1e-13


@pytest.mark.parametrize("ny,nx", [(7, 7), (9, 11), (33, 34), (40, 23)])
def test_oracle_evaluation_on_ragged_grids(ny, nx):
    x, y, c = _grid(ny, nx)
    f = 2.2e5
    iy = [1, ny - 2, ny // 2]
    ix = [1, nx // 2, nx - 2]
    tx = np.array([iy[k] + ny * ix[k] + 1 for k in range(3)])
    rx = np.array(sorted({iy[k] + ny * ix[(k + 1) % 3] + 1 for k in range(3)} | {ny // 2 + ny * 1 + 1}))
    cfg = dict(backend="gpu", wavenumber="continuum", pml_strength=10.0, pml_m=2e-3, stencil_bounds=[1400, 1700],
               source_batch_size=2, tol=1e-13)
    ref = H.HelmholtzFP64(x, y, c, np.zeros_like(c), f, -1, 10.0, 2e-3, dict(stencil_bounds=[1400, 1700]))
    src = np.zeros((ny * nx, tx.size), np.complex128)
    src[tx - 1, np.arange(tx.size)] = 1
    u, _ = ref.solve(src.reshape(ny, nx, -1, order="F"))
    Y = u.reshape(ny * nx, -1, order="F")[rx - 1, :].T[:, :, None] * (1 + 0.1j)
    rng = np.random.default_rng(ny + nx)
    Y = Y * (1 + 0.1 * (rng.standard_normal(Y.shape) + 1j * rng.standard_normal(Y.shape)))   # not a fitted multiple
    obs = dict(x_m=x, y_m=y, shape_yx=(ny, nx), n_tx=tx.size, tx_index=tx, rx_index=rx, frequencies_hz=[f], Y=Y,
               mask=np.ones(Y.shape, bool))
    init = 1480 * np.ones((ny, nx))
    lo, go = H.wust_oracle(init, obs, 0, cfg)
    lg, gg = G.wust_oracle(init, obs, 0, cfg)
    assert abs(lg - lo) <= 1e-9 * abs(lo)
    assert np.linalg.norm(gg.cpu().numpy() - go) <= 1e-8 * np.linalg.norm(go)


def test_pooled_buffers_reproduce_results():
    """An operator built on buffers of a destroyed one (the device pool) gives bit-identical results
    (history-free factors: no seeds or learned profile from the earlier operators)."""
    from waveforminversionust_amd import _native as N
    x, y, c = _grid(40, 23)
    f, a0, L = 2.2e5, 10.0, 2e-3
    rng = np.random.default_rng(3)
    b = np.zeros((40, 23, 7), np.complex128)
    b[1:-1, 1:-1] = rng.standard_normal((38, 21, 7)) + 1j * rng.standard_normal((38, 21, 7))
    N.check(N.load().wust_pool_trim())
    results = []
    for _ in range(3):
        s = G.HelmholtzSolver(x, y, c, None, f, -1, a0, L, dict(OPTS, mode="refined", tol=1e-13, factor="fixed"))
        results.append(s.solve(b, False)[0].clone())
        s.close()
        if len(results) == 2:
            N.check(N.load().wust_pool_trim())
    assert torch.equal(results[0], results[1]) and torch.equal(results[0], results[2])


@pytest.mark.parametrize("mode", ["refined", "c64"])
@pytest.mark.parametrize("adjoint", [False, True])
def test_zero_column_skip_is_exact(monkeypatch, mode, adjoint):
    """Right-hand sides confined to a band of grid columns (point sources and receivers leave the
    columns between the array and the domain edges empty): the sweeps skip the forward
    elimination over the empty leading / trailing columns, and the solution is bit-identical to
    the solve that runs every step."""
    ny, nx = 36, 70
    x, y, c = _grid(ny, nx)
    f, a0, L = 2.2e5, 10.0, 2e-3
    rng = np.random.default_rng(7)
    b = np.zeros((ny, nx, 9), np.complex128)
    b[1:-1, 20:27] = rng.standard_normal((ny - 2, 7, 9)) + 1j * rng.standard_normal((ny - 2, 7, 9))
    b[5, 50, 3] = 1.0                                    # one more source further right
    ref = H.HelmholtzFP64(x, y, c, np.zeros_like(c), f, -1, a0, L, dict(stencil_bounds=[1400, 1700]))
    s = G.HelmholtzSolver(x, y, c, None, f, -1, a0, L, dict(OPTS, mode=mode, tol=1e-13))
    got = s.solve(b, adjoint)[0].clone()
    monkeypatch.setenv("WUST_SKIP_ZERO_COLUMNS", "0")
    full = s.solve(b, adjoint)[0].clone()
    assert torch.equal(got, full)
    Hm = ref.H.conj().T if adjoint else ref.H
    u = got.cpu().numpy().reshape(-1, 9, order="F")
    bf = b.reshape(-1, 9, order="F")
    interior = np.zeros((ny, nx), bool)
    interior[1:-1, 1:-1] = True
    res = (Hm @ u - bf).reshape(ny, nx, 9, order="F")[interior]
    assert np.linalg.norm(res) <= (1e-12 if mode == "refined" else 1e-4) * np.linalg.norm(bf)


@pytest.mark.parametrize("where", ["zero", "first", "second", "before_meet", "meet", "after_meet", "last"])
def test_zero_column_skip_edges(monkeypatch, where):
    """The skip at its edges: an all-zero right-hand side, and right-hand sides confined to one
    block column at either end of the chains, next to the meeting block or at it (the chains keep
    at least their last step).  Bit-identical to the solve that runs every step, and solved."""
    ny, nx = 20, 19                                       # m = 17 interior columns, meeting block p = 8
    x, y, c = _grid(ny, nx)
    m, p = nx - 2, (nx - 2) // 2
    col = dict(zero=None, first=0, second=1, before_meet=p - 1, meet=p, after_meet=p + 1, last=m - 1)[where]
    rng = np.random.default_rng(11)
    b = np.zeros((ny, nx, 3), np.complex128)
    if col is not None:
        b[1:-1, col + 1] = rng.standard_normal((ny - 2, 3)) + 1j * rng.standard_normal((ny - 2, 3))
    f, a0, L = 2.2e5, 10.0, 2e-3
    ref = H.HelmholtzFP64(x, y, c, np.zeros_like(c), f, -1, a0, L, dict(stencil_bounds=[1400, 1700]))
    s = G.HelmholtzSolver(x, y, c, None, f, -1, a0, L, dict(OPTS, mode="refined", tol=1e-13))
    for adjoint in (False, True):
        monkeypatch.delenv("WUST_SKIP_ZERO_COLUMNS", raising=False)
        got = s.solve(b, adjoint)[0].clone()
        monkeypatch.setenv("WUST_SKIP_ZERO_COLUMNS", "0")
        full = s.solve(b, adjoint)[0].clone()
        assert torch.equal(got, full), (where, adjoint)
        u = got.cpu().numpy()
        if col is None:
            assert not np.any(u)
            continue
        Hm = ref.H.conj().T if adjoint else ref.H
        interior = np.zeros((ny, nx), bool)
        interior[1:-1, 1:-1] = True
        bf = b.reshape(-1, 3, order="F")
        res = (Hm @ u.reshape(-1, 3, order="F") - bf).reshape(ny, nx, 3, order="F")[interior]
        assert np.linalg.norm(res) <= 1e-12 * np.linalg.norm(bf), (where, adjoint)


@pytest.mark.parametrize("ny,nx", [(66, 40), (52, 37)])
@pytest.mark.parametrize("adjoint", [False, True])
def test_sweep_step_variants_agree(monkeypatch, ny, nx, adjoint):
    """Every kernel variant of the fused sweep step, forced in turn: they differ only in how the
    operands reach the MFMA (staged through LDS, loaded into LDS by the load unit, buffer or flat
    loads, prefetch depth) and deal K
    over the four waves in the same order, so they give bit-identical solutions, each solving the
    oracle's H to single-precision accuracy.  (52 rows: the last strip is clamped to the block end
    and the padding rows beyond the strips are zero.)"""
    x, y, c = _grid(ny, nx)
    f, a0, L = 2.2e5, 10.0, 2e-3
    R = 32
    rng = np.random.default_rng(ny + nx)
    b = np.zeros((ny, nx, R), np.complex128)
    b[1:-1, 1:-1] = rng.standard_normal((ny - 2, nx - 2, R)) + 1j * rng.standard_normal((ny - 2, nx - 2, R))
    ref = H.HelmholtzFP64(x, y, c, np.zeros_like(c), f, -1, a0, L, dict(stencil_bounds=[1400, 1700]))
    s = G.HelmholtzSolver(x, y, c, None, f, -1, a0, L, dict(OPTS, mode="c64"))
    Hm = ref.H.conj().T if adjoint else ref.H
    bf = b.reshape(-1, R, order="F")
    interior = np.zeros((ny, nx), bool)
    interior[1:-1, 1:-1] = True
    base = None
    for v in range(9):
        monkeypatch.setenv("WUST_SWEEP_VARIANT", str(v))
        u = s.solve(b, adjoint)[0].clone()
        if base is None:
            base = u
        else:
            assert torch.equal(u, base), v
        res = (Hm @ u.cpu().numpy().reshape(-1, R, order="F") - bf).reshape(ny, nx, R, order="F")[interior]
        assert np.linalg.norm(res) <= 1e-4 * np.linalg.norm(bf), v

