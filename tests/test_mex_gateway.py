"""The MATLAB MEX gateway (waveforminversionust_amd/matlab/wust_hip_mex.cpp), compiled against the stub
mex.h / matrix.h of tests/mex_stub and driven through mexFunction.

CPU tests: the gateway links and loads, validates its arguments before touching the device, and maps
failures to the reference's identifiers.  GPU tests: every command (create / solve / mass / oracle /
curvature / release / destroy) reproduces the Python binding over the same C ABI bit for bit, and the
reference's MEX-boundary rejections hold (real storage and rank-4 fields are WUST:InvalidInput,
Runtime/tests/test_runtime.m:213-240), as do stale handles and the refinement's WUST:NotConverged.
"""
import numpy as np
import pytest

from mex_stub.harness import Mex, MexCallError

OPTS = dict(backend="gpu", stencil_bounds=[1400, 1700])


@pytest.fixture(scope="module")
def mex():
    return Mex()


def _grid(ny=24, nx=19, dx=5e-4):
    x = (np.arange(nx) - nx / 2) * dx
    y = (np.arange(ny) - ny / 2) * dx
    X, Y = np.meshgrid(x, y)
    return x, y, 1500 + 30 * np.exp(-(X ** 2 + Y ** 2) / (0.2 * (nx * dx) ** 2))


def _create_args(x, y, c, f=2.2e5, mode="refined"):
    return ("create", x, y, c, np.zeros((0, 0)), f, -1.0, 10.0, 2e-3, np.array([1400.0, 1700.0]), mode,
            np.zeros((0, 0)))


def _raises(mex, ident, *args, nlhs=1):
    with pytest.raises(MexCallError) as ei:
        mex(nlhs, *args)
    assert ei.value.identifier == ident, str(ei.value)
    return str(ei.value)


# ---- CPU: validation before any device work ------------------------------------------------------
def test_unknown_command_and_usage(mex):
    _raises(mex, "WUST:InvalidInput", "no-such-command")
    _raises(mex, "WUST:InvalidInput", 3.0)


def test_create_rejects_invalid_input(mex):
    x, y, c = _grid()
    args = list(_create_args(x, y, c))
    bad_c = [c.astype(complex), np.stack([c, c], axis=2), c[:, :-1]]
    for bc in bad_c:   # complex storage, rank-3 medium, wrong shape
        a = list(args)
        a[3] = bc
        _raises(mex, "WUST:InvalidInput", *a)
    a = list(args)
    a[1] = np.r_[x[:-1], x[-1] + 1e-4]          # non-uniform grid
    assert "Uniform" in _raises(mex, "WUST:InvalidInput", *a)
    a = list(args)
    a[10] = "single"
    _raises(mex, "WUST:InvalidInput", *a)
    a = list(args)
    a[9] = np.array([1400.0])
    _raises(mex, "WUST:InvalidInput", *a)
    _raises(mex, "WUST:InvalidInput", *args[:-1])   # one argument short


def test_foreign_handles_rejected(mex):
    """A handle this gateway did not create (or already destroyed) never reaches the engine."""
    for h in (np.uint64(0), np.uint64(12345678)):
        for cmd in ("info", "release", "destroy"):
            assert "handle" in _raises(mex, "WUST:InvalidInput", cmd, np.array(h))
    _raises(mex, "WUST:InvalidInput", "info", 1.0)   # a double is not a handle


# ---- GPU: the commands against the Python binding -------------------------------------------------
@pytest.mark.gpu
def test_gateway_matches_python_binding(mex):
    torch = pytest.importorskip("torch")
    from waveforminversionust_amd import helmholtz as G
    from waveforminversionust_amd import workloads as W

    case = W.runtime_test_case()
    x, y, c = case.x_m, case.y_m, case.c_true
    f = case.frequencies_hz[0]
    ny, nx = c.shape
    from waveforminversionust_amd import _native as N
    G.HelmholtzSolver(x, y, c, None, f, -1, 10.0, 0.004, OPTS).close()   # both below use the learned factor policy
    # each create below starts without cached leading pages (the second would otherwise be seeded from
    # the first factorization, which changes its rounding)
    N.check(N.load().wust_pool_trim())
    h, info = mex(2, "create", x, y, c, np.zeros((0, 0)), f, -1.0, 10.0, 0.004, np.array([1400.0, 1700.0]), "refined",
                  np.zeros((0, 0)))
    assert h.dtype == np.uint64 and h.shape == (1, 1) and info.shape == (1, 9)
    N.check(N.load().wust_pool_trim())
    py = G.HelmholtzSolver(x, y, c, None, f, -1, 10.0, 0.004, OPTS)
    assert info[0, 2] == py.stencil_b and info[0, 3] == py.massD and info[0, 4] == py.massE
    rng = np.random.default_rng(9)
    src = np.zeros((ny, nx, 3), np.complex128)
    src[1:-1, 1:-1] = rng.standard_normal((ny - 2, nx - 2, 3)) + 1j * rng.standard_normal((ny - 2, nx - 2, 3))
    for adjoint in (False, True):
        u, v, st = mex(3, "solve", h, src, float(adjoint), 1e-12, 30.0)
        uw, vw, stw = py.solve(src, adjoint, tol=1e-12, max_iter=30, return_stats=True)
        assert np.array_equal(u, uw.cpu().numpy()) and np.array_equal(v, vw.cpu().numpy())
        # the residual norms are sums over workgroups (atomic order): equal to rounding
        assert st[0, 0] == stw["iterations"] and abs(st[0, 2] - stw["final_residual"]) <= 1e-12 * stw["final_residual"]
        m = mex(1, "mass", h, u, float(adjoint))
        assert np.array_equal(m, py.massApply(uw, adjoint).cpu().numpy())
    # a single right-hand side as a [ny, nx] matrix
    u1 = mex(1, "solve", h, src[:, :, 0], 0.0, 1e-12, 30.0)
    assert u1.shape == (ny, nx) and np.array_equal(u1, py.solve(src[:, :, :1], False, tol=1e-12)[0].cpu().numpy()[:, :, 0])
    py.close()

    # the oracle: loss, gradient, scales; curvature on the retained state; release
    from conftest import make_obs
    obs = make_obs(case)
    Y = obs["Y"][:, :, 0]
    mask = np.ones(Y.shape, bool)
    init = np.full((ny, nx), 1500.0)
    cfg = dict(case.cfg, tol=1e-12)
    G.wust_oracle(init, obs, 0, cfg, nargout=1)   # the factor policy of this operator, for both below
    N.check(N.load().wust_pool_trim())
    ho = mex(1, "create", x, y, init, np.zeros((0, 0)), f, -1.0, 10.0, 0.004, np.array([1400.0, 1700.0]), "refined",
             np.zeros((0, 0)))
    loss, g, scales = mex(3, "oracle", ho, case.tx_index.astype(float), case.rx_index.astype(float), Y, mask,
                          2.0, 1.0, 1.0, 1e-12, 30.0)
    N.check(N.load().wust_pool_trim())
    lw, gw, state = G.wust_oracle(init, obs, 0, cfg, nargout=3)
    assert loss.item() == lw and np.array_equal(g, gw.cpu().numpy())
    assert np.array_equal(scales.ravel(), state.scales)
    direction = rng.standard_normal((ny, nx))
    kappa = mex(1, "curvature", ho, direction, mask, 1e-12, 30.0)
    kw = G.directional_curvature(state, direction, case.rx_index, mask)
    assert kappa.item() == kw
    mex(0, "release", ho)
    _raises(mex, "WUST:InvalidInput", "curvature", ho, direction, mask, 1e-12, 30.0)   # no retained state
    mex(0, "destroy", ho)
    mex(0, "destroy", h)
    _raises(mex, "WUST:InvalidInput", "info", h)   # destroyed


@pytest.mark.gpu
def test_gateway_boundary_rejections(mex):
    """The reference's MEX boundary tests (Runtime/tests/test_runtime.m:213-240): real storage and
    rank-4 fields are WUST:InvalidInput; so are fields of the wrong grid; refinement that cannot reach its
    tolerance is WUST:NotConverged."""
    pytest.importorskip("torch")
    x, y, c = _grid()
    ny, nx = c.shape
    h = mex(1, *_create_args(x, y, c))
    good = np.zeros((ny, nx, 2), np.complex128)
    good[5, 5, :] = 1
    _raises(mex, "WUST:InvalidInput", "solve", h, good.real, 0.0, 1e-10, 30.0)              # real storage
    _raises(mex, "WUST:InvalidInput", "solve", h, good[:, :, :, None].repeat(2, 3), 0.0, 1e-10, 30.0)   # rank 4
    _raises(mex, "WUST:InvalidInput", "solve", h, good[:-1], 0.0, 1e-10, 30.0)               # wrong grid
    _raises(mex, "WUST:InvalidInput", "mass", h, good.real, 1.0)
    _raises(mex, "WUST:NotConverged", "solve", h, good, 0.0, 1e-20, 1.0)
    u = mex(1, "solve", h, good, 0.0, 1e-10, 30.0)                                           # still usable
    assert np.all(np.isfinite(u))
    mex(0, "destroy", h)
