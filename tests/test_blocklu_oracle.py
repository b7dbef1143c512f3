"""The FP64 block-LU baseline (oracle/blocklu_fp64.py) against the sparse-LU oracle: the block
recurrence solves the same H (and H^H on the interior) to FP64 accuracy, and its loss / gradient equal
the oracle's."""
import numpy as np

from oracle import blocklu_fp64 as B
from oracle import helmholtz_fp64 as H
from waveforminversionust_amd import workloads as W

from conftest import make_obs


def _grid(ny, nx, dx=5e-4):
    x = (np.arange(nx) - nx / 2) * dx
    y = (np.arange(ny) - ny / 2) * dx * 1.1
    X, Y = np.meshgrid(x, y)
    return x, y, 1500 + 40 * np.exp(-(X ** 2 + Y ** 2) / (0.2 * (nx * dx) ** 2))


def test_block_solves_match_sparse_H():
    x, y, c = _grid(23, 27)
    f, a0, L = 2.2e5, 10.0, 2e-3
    ref = H.HelmholtzFP64(x, y, c, np.zeros_like(c), f, -1, a0, L, dict(stencil_bounds=[1400, 1700]))
    s = B.BlockLUFP64(x, y, c, f, a0, L, dict(stencil_bounds=[1400, 1700]))
    rng = np.random.default_rng(3)
    b = np.zeros((23, 27, 4), np.complex128)
    b[1:-1, 1:-1] = rng.standard_normal((21, 25, 4)) + 1j * rng.standard_normal((21, 25, 4))
    bf = b.reshape(-1, 4, order="F")
    interior = np.zeros((23, 27), bool)
    interior[1:-1, 1:-1] = True
    for adjoint in (False, True):
        u = np.zeros_like(b)
        u[1:-1, 1:-1] = s.solve_interior(b[1:-1, 1:-1], adjoint)
        Hm = ref.H.conj().T if adjoint else ref.H
        res = (Hm @ u.reshape(-1, 4, order="F") - bf).reshape(23, 27, 4, order="F")[interior]
        assert np.linalg.norm(res) <= 1e-12 * np.linalg.norm(bf), adjoint
        if not adjoint:   # the forward solution is the sparse LU's (zero on the Dirichlet ring)
            want = ref.solve(b)[0]
            assert np.linalg.norm(u - want) <= 1e-8 * np.linalg.norm(want)


def test_blocklu_oracle_matches_sparse_oracle(runtime_case, runtime_obs):
    cfg = runtime_case.cfg
    init = 1500 * np.ones((runtime_case.ny, runtime_case.nx))
    for fi in range(len(runtime_case.frequencies_hz)):
        lo, go = H.wust_oracle(init, runtime_obs, fi, cfg)
        lb, gb = B.wust_oracle_blocklu(init, runtime_obs, fi, cfg)
        # the fixture's loss (1e-17) is a near-cancellation of predicted and observed data: two FP64
        # solvers agree to ~1e-10 in it (the GPU engine's smoke check sees the same 8e-11)
        assert abs(lb - lo) <= 1e-9 * abs(lo)
        assert np.linalg.norm(gb - go) <= 1e-8 * np.linalg.norm(go)
        assert B.wust_oracle_blocklu(init, runtime_obs, fi, cfg, nargout=1) == lb


def test_blocklu_oracle_ring_case():
    case = W.ring_case("ring40_12tx", 28, 6, 0.5e-3, 12, 6e-3, [3e5], "discs")
    obs = make_obs(case)
    init = 1500 * np.ones((case.ny, case.nx))
    cfg = dict(case.cfg, source_batch_size=5)
    lo, go = H.wust_oracle(init, obs, 0, cfg)
    lb, gb = B.wust_oracle_blocklu(init, obs, 0, cfg)
    assert abs(lb - lo) <= 1e-10 * abs(lo)
    assert np.linalg.norm(gb - go) <= 1e-9 * np.linalg.norm(go)
