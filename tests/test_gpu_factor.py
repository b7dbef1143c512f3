"""GPU checks of the two ways the engine inverts Schur blocks (needs an MI355X).

Blocks with an inverted neighbour in their chain are inverted by Newton-Schulz
iteration (factor='auto') except where Gauss-Jordan is cheaper (the second
column of each chain, and columns that needed 4 or more rounds, 3 or more from
npad 800 on, in an earlier factorization of the same operator key); the chain starts, the meeting block,
and the whole factor when an iteration misses its residual bound use
Gauss-Jordan elimination (factor='gauss-jordan' forces it everywhere).  A later
factorization with the same key seeds Newton-Schulz for the leading chain steps
with the previous factorization's pages of the same blocks instead.
Tolerances:
* the two methods give complex-FP32 solutions that agree to <= 1e-5
  relative (both are FP32 inverses with residuals near 1e-6);
* the Newton-Schulz residual bound reported by the factor is <= 2.5e-5;
* refined solves from either factor agree with each other to <= 1e-12;
* a forced fallback reproduces the Gauss-Jordan factor bit for bit.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from waveforminversionust_amd import workloads as W  # noqa: E402
from waveforminversionust_amd import helmholtz as G  # noqa: E402

OPTS = dict(backend="gpu", stencil_bounds=[1400, 1700])


@pytest.fixture(scope="module")
def ring():
    return W.cpu_reference_case()


def _solver(case, mode, factor):
    return G.HelmholtzSolver(case.x_m, case.y_m, case.c_true, None, case.frequencies_hz[0], -1,
                             case.cfg["pml_strength"], case.cfg["pml_m"], dict(OPTS, mode=mode, factor=factor))


def _sources(case, count):
    src = torch.zeros((case.ny * case.nx, count), dtype=torch.complex128, device="cuda")
    src[torch.as_tensor(case.tx_index[:count] - 1), torch.arange(count)] = 1
    return src.reshape(case.nx, case.ny, count).permute(1, 0, 2)


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("adjoint", [False, True])
def test_newton_schulz_matches_gauss_jordan(ring, adjoint):
    src = _sources(ring, 16)
    s_ns = _solver(ring, "c64", "auto")
    s_gj = _solver(ring, "c64", "gauss-jordan")
    i_ns, i_gj = s_ns.info(), s_gj.info()
    m = ring.nx - 2
    # (up to m - 1 blocks: the leading steps are seeded when an earlier test factored this key)
    assert (m - 3) // 2 <= i_ns["iterative_blocks"] <= m - 1 and i_ns["factor_fallback"] == 0
    # the reported bound: ||E||_F^2 of a quadratic last update (taken up to 2.5e-5) or ||E||_F^3 of a
    # cubic one; the solutions below check the inverses themselves
    assert 0 < i_ns["inverse_residual"] <= 2.5e-5
    assert i_gj["iterative_blocks"] == 0 and i_gj["factor_fallback"] == 0
    u_ns, _ = s_ns.solve(src, adjoint)
    u_gj, _ = s_gj.solve(src, adjoint)
    assert _rel(u_ns, u_gj) <= 1e-5


def test_learned_budgets_keep_the_factor(ring):
    # a second factorization of the same operator key runs on the learned step policy and
    # budgets; its solutions match the first one's to FP32 inverse accuracy
    src = _sources(ring, 8)
    first = _solver(ring, "c64", "auto")
    second = _solver(ring, "c64", "auto")
    i1, i2 = first.info(), second.info()
    assert i2["factor_fallback"] == 0 and 0 < i2["inverse_residual"] <= 2.5e-5
    assert i2["iterative_blocks"] <= i1["iterative_blocks"]
    u1, _ = first.solve(src, False)
    u2, _ = second.solve(src, False)
    assert _rel(u2, u1) <= 1e-5


def test_fixed_budget_factor_is_history_free(ring):
    """factor='fixed' (WUST_FACTOR_FIXED_BUDGET): Newton-Schulz with the default round budgets, so
    the factor depends only on the operator, not on the profile that earlier factorizations of the
    same shape and frequency left behind: two fixed-budget operators built before and after learned
    ('auto') factorizations give bit-identical c64 solutions in both directions, and agree with the
    learned-budget factor to FP32 inverse accuracy.  The sweep kernel variants all accumulate in
    the same order, so the per-shape variant choice does not enter either."""
    src = _sources(ring, 8)
    before = _solver(ring, "c64", "fixed")
    u_b = [before.solve(src, adj)[0].clone() for adj in (False, True)]
    before.close()
    learned = [_solver(ring, "c64", "auto") for _ in range(2)]   # the second runs on a learned profile
    u_l = learned[1].solve(src, False)[0]
    after = _solver(ring, "c64", "fixed")
    u_a = [after.solve(src, adj)[0] for adj in (False, True)]
    info = after.info()
    assert info["factor_fallback"] == 0 and info["iterative_blocks"] > 0
    assert torch.equal(u_a[0], u_b[0]) and torch.equal(u_a[1], u_b[1])
    assert _rel(u_l, u_a[0]) <= 1e-5


def test_newton_schulz_refined_agrees(ring):
    src = _sources(ring, 8)
    u_ns, _ = _solver(ring, "refined", "auto").solve(src, False, tol=1e-13)
    u_gj, _ = _solver(ring, "refined", "gauss-jordan").solve(src, False, tol=1e-13)
    assert _rel(u_ns, u_gj) <= 1e-12


def test_newton_schulz_fallback_is_gauss_jordan(ring, monkeypatch):
    # with the default budgets capped at one update, the PML columns cannot reach the bound
    monkeypatch.setenv("WUST_NS_MAX_UPDATES", "1")
    monkeypatch.setenv("WUST_NS_FIXED_BUDGET", "1")
    s_fb = _solver(ring, "c64", "auto")
    info = s_fb.info()
    assert info["factor_fallback"] == 1 and info["iterative_blocks"] == 0
    monkeypatch.delenv("WUST_NS_MAX_UPDATES")
    monkeypatch.delenv("WUST_NS_FIXED_BUDGET")
    s_gj = _solver(ring, "c64", "gauss-jordan")
    src = _sources(ring, 8)
    u_fb, _ = s_fb.solve(src, False)
    u_gj, _ = s_gj.solve(src, False)
    assert torch.equal(u_fb, u_gj)


def test_unconverged_factor_keeps_its_profile(ring, monkeypatch):
    # An operator key (its own frequency, so that no earlier test taught it) whose Newton-Schulz
    # blocks cannot converge under a one-update cap: the first factorization falls back to
    # Gauss-Jordan, and the profile it leaves moves the blocks that ran out of rounds to
    # Gauss-Jordan step by step, so a later factorization with the same key and cap converges
    # without the fallback (starting again from the default budgets would fall back every time).
    f = ring.frequencies_hz[0] * 1.0371
    monkeypatch.setenv("WUST_NS_MAX_UPDATES", "1")

    def make():
        return G.HelmholtzSolver(ring.x_m, ring.y_m, ring.c_true, None, f, -1, ring.cfg["pml_strength"],
                                 ring.cfg["pml_m"], dict(OPTS, mode="c64", factor="auto"))

    infos = [make().info()]
    assert infos[0]["factor_fallback"] == 1
    s = None
    for _ in range(6):
        s = make()
        infos.append(s.info())
        if infos[-1]["factor_fallback"] == 0:
            break
    assert infos[-1]["factor_fallback"] == 0, [i["factor_fallback"] for i in infos]
    assert infos[-1]["inverse_residual"] <= 2.5e-5   # (under this cap every block may end on Gauss-Jordan)
    monkeypatch.delenv("WUST_NS_MAX_UPDATES")
    s_gj = G.HelmholtzSolver(ring.x_m, ring.y_m, ring.c_true, None, f, -1, ring.cfg["pml_strength"],
                             ring.cfg["pml_m"], dict(OPTS, mode="c64", factor="gauss-jordan"))
    src = _sources(ring, 8)
    u, _ = s.solve(src, False)
    u_gj, _ = s_gj.solve(src, False)
    assert _rel(u, u_gj) <= 1e-5



def test_seeded_leading_steps(ring):
    """The leading (PML) chain steps of a factorization are seeded from the previous factorization with
    the same key: after a model change that leaves the PML nearly untouched (a smooth bump), the
    seeded factor inverts every block but the meeting one by Newton-Schulz without a fallback, and its
    solutions match a history-free factor of the same model (c64 to FP32 inverse accuracy, refined to
    1e-12); after a large change (the whole model 60 m/s faster) the seeds stop converging, the steps
    from the first such one are inverted again without seeds (no fallback of the whole factor), and
    the solutions match the history-free factor the same way; wust_pool_trim drops the seeds."""
    from waveforminversionust_amd import _native as N
    f = ring.frequencies_hz[0] * 0.9713     # an operator key of its own
    yy, xx = np.meshgrid(np.linspace(-1, 1, ring.ny), np.linspace(-1, 1, ring.nx), indexing="ij")
    bump = np.exp(-(xx ** 2 + yy ** 2) / 0.1)

    def make(c, mode, factor):
        return G.HelmholtzSolver(ring.x_m, ring.y_m, c, None, f, -1, ring.cfg["pml_strength"], ring.cfg["pml_m"],
                                 dict(OPTS, mode=mode, factor=factor))

    m = ring.nx - 2
    first = make(ring.c_true, "c64", "auto").info()
    make(ring.c_true, "c64", "auto").close()            # learns the profile, caches the leading pages
    c2 = ring.c_true * (1 + 2e-3 * bump)
    seeded = make(c2, "c64", "auto")
    info = seeded.info()
    assert info["factor_fallback"] == 0 and info["iterative_blocks"] == m - 1 > first["iterative_blocks"]
    assert 0 < info["inverse_residual"] <= 1e-4
    src = _sources(ring, 8)
    u_s = seeded.solve(src, False)[0]
    u_f = make(c2, "c64", "fixed").solve(src, False)[0]
    assert _rel(u_s, u_f) <= 1e-5
    r_s = make(c2, "refined", "auto").solve(src, True, tol=1e-13)[0]
    r_f = make(c2, "refined", "fixed").solve(src, True, tol=1e-13)[0]
    assert _rel(r_s, r_f) <= 1e-12
    c3 = ring.c_true + 60.0
    stale = make(c3, "c64", "auto")
    i3 = stale.info()
    assert i3["factor_fallback"] == 0 and i3["iterative_blocks"] < m - 1
    assert _rel(stale.solve(src, False)[0], make(c3, "c64", "fixed").solve(src, False)[0]) <= 1e-5
    r_s = make(c3, "refined", "auto").solve(src, False, tol=1e-13)[0]
    r_f = make(c3, "refined", "fixed").solve(src, False, tol=1e-13)[0]
    assert _rel(r_s, r_f) <= 1e-12
    N.check(N.load().wust_pool_trim())
    again = make(ring.c_true, "c64", "auto").info()
    assert again["iterative_blocks"] < m - 1
