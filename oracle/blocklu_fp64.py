"""FP64 block-tridiagonal LU on the host: the CPU baseline's multicore variant (test infrastructure).

The reference's GPU algorithm in FP64 on the CPU, with the grid columns as blocks of the
interior rows (Runtime/solver/decompBlockLU.cu:7-20,337-368 for the Schur recurrence
T_0 = D_0, T_j = D_j - L_j T_{j-1}^{-1} U_{j-1}, one explicit inverse per block;
Runtime/solver/applyBlockLU.cu:433-467 for the sweeps, forward and H^H).  The dense inverses and
the multi-RHS products run on the host BLAS with as many threads as it is given, so this is the
parallel counterpart of the sparse-LU oracle (scipy SuperLU, one thread) in
``oracle/helmholtz_fp64.py``, whose assembly, mass stencil and loss / gradient formulas it reuses.
Only bench.py's CPU-baseline leg and the tests use it; it is checked against the sparse-LU
oracle in tests/test_blocklu_oracle.py.

Block layout: block j is grid column x = j + 1, rows y = 1 .. Ny-2 (index i = y - 1).  A
tridiagonal coupling P is held as three vectors (c, lo, up): (P v)[i] = c[i] v[i] + lo[i] v[i-1]
+ up[i] v[i+1], terms leaving the block dropped (those nodes lie on the Dirichlet ring).
"""
from __future__ import annotations

import numpy as np

from oracle import helmholtz_fp64 as H


def _tri_apply(P, v):
    """P v for a tridiagonal coupling P = (c, lo, up) and v [n, R] (or [n, n])."""
    c, lo, up = P
    out = c[:, None] * v
    out[1:] += lo[1:, None] * v[:-1]
    out[:-1] += up[:-1, None] * v[1:]
    return out


def _tri_apply_h(P, v):
    """P^H v: (P^H v)[i] = conj(c[i]) v[i] + conj(lo[i+1]) v[i+1] + conj(up[i-1]) v[i-1]."""
    c, lo, up = P
    out = np.conj(c)[:, None] * v
    out[:-1] += np.conj(lo[1:])[:, None] * v[1:]
    out[1:] += np.conj(up[:-1])[:, None] * v[:-1]
    return out


def _right_tri(M, P):
    """M P for a dense M [n, n] and tridiagonal P: column k = c[k] M[:, k] + lo[k+1] M[:, k+1] + up[k-1] M[:, k-1]."""
    c, lo, up = P
    out = M * c[None, :]
    out[:, :-1] += M[:, 1:] * lo[None, 1:]
    out[:, 1:] += M[:, :-1] * up[None, :-1]
    return out


class BlockLUFP64:
    """FP64 block LU of the interior operator H_II of HelmholtzSolver (Runtime/solver/HelmholtzSolver.m:32-203),
    built from the sparse-LU oracle's own FP64 assembly."""

    def __init__(self, x, y, vel, f, a0, L_pml, options):
        vel = np.asarray(vel, np.float64)
        Ny, Nx = vel.shape
        t = H.pml_terms(np.asarray(x, np.float64), np.asarray(y, np.float64), vel, np.zeros_like(vel), f, -1, a0,
                        L_pml)
        lo, hi = np.asarray(options["stencil_bounds"], np.float64).ravel()
        b, d, e = H.stencil_opt_params(lo, hi, f, t.h, t.g)
        self.massD, self.massE = d, e
        co = H.nine_point_coefficients(t.A, t.B, t.C, t.k, b, d, e, t.h, t.g, Nx, Ny)
        self.virtual_source_multiplier = t.C * t.dqds
        self.Ny, self.Nx = Ny, Nx
        n, m = Ny - 2, Nx - 2
        ys = np.arange(1, Ny - 1)
        self.D, self.L, self.U = [], [], []
        for j in range(m):
            xx = j + 1
            self.D.append((co["Dd"][ys, xx], co["Dl"][ys - 1, xx], co["Du"][ys, xx]))
            self.L.append((co["Ld"][ys, xx - 1], co["Ll"][ys - 1, xx - 1], co["Lu"][ys, xx - 1]))
            self.U.append((co["Ud"][ys, xx], co["Ul"][ys - 1, xx], co["Uu"][ys, xx]))
        # Schur recurrence, one explicit inverse per block (decompBlockLU.cu:337-368)
        self.X = []
        eye = np.eye(n, dtype=np.complex128)
        for j in range(m):
            T = _tri_apply(self.D[j], eye)
            if j > 0:
                T -= _tri_apply(self.L[j], _right_tri(self.X[j - 1], self.U[j - 1]))
            self.X.append(np.linalg.inv(T))

    def solve_interior(self, b, adjoint=False):
        """H_II u = b (or H_II^H u = b) for b [n, m, R] (interior nodes), returns [n, m, R]."""
        m = self.Nx - 2
        y = [None] * m
        for j in range(m):
            bj = b[:, j, :]
            if not adjoint:
                if j > 0:
                    bj = bj - _tri_apply(self.L[j], y[j - 1])
                y[j] = self.X[j] @ bj
            else:
                if j > 0:
                    bj = bj - _tri_apply_h(self.U[j - 1], y[j - 1])
                y[j] = self.X[j].conj().T @ bj
        u = [None] * m
        u[m - 1] = y[m - 1]
        for j in range(m - 2, -1, -1):
            if not adjoint:
                u[j] = y[j] - self.X[j] @ _tri_apply(self.U[j], u[j + 1])
            else:
                u[j] = y[j] - self.X[j].conj().T @ _tri_apply_h(self.L[j + 1], u[j + 1])
        return np.stack(u, axis=1)

    def mass_apply(self, v, adjoint=False):
        """The sparse-LU oracle's exact mass-stencil derivative (Runtime/solver/HelmholtzSolver.m:205-220)."""
        return H.HelmholtzFP64.mass_apply(self, v, adjoint)


def wust_oracle_blocklu(c, obs, fi, cfg, nargout=2):
    """Loss (and gradient) of Runtime/matlab/wust_oracle.m:1-86 with block-LU solves; the same steps and
    formulas as ``oracle.helmholtz_fp64.wust_oracle`` (source batches of cfg['source_batch_size'])."""
    Ny, Nx = obs["shape_yx"]
    nt = obs["n_tx"]
    s = BlockLUFP64(obs["x_m"], obs["y_m"], c, obs["frequencies_hz"][fi], cfg["pml_strength"], cfg["pml_m"],
                    dict(stencil_bounds=cfg["stencil_bounds"]))
    tx = np.asarray(obs["tx_index"], np.int64) - 1
    rx = np.asarray(obs["rx_index"], np.int64) - 1
    need_gradient = nargout > 1
    grad = np.zeros((Ny, Nx)) if need_gradient else None
    loss = 0.0
    bs = int(cfg["source_batch_size"])
    for first in range(0, nt, bs):
        ids = np.arange(first, min(first + bs, nt))
        nb = ids.size
        src = np.zeros((Ny * Nx, nb), np.complex128)
        src[tx[ids], np.arange(nb)] = 1
        b = src.reshape(Ny, Nx, nb, order="F")[1:-1, 1:-1, :]
        u = np.zeros((Ny, Nx, nb), np.complex128)
        u[1:-1, 1:-1] = s.solve_interior(b, False)
        v = s.virtual_source_multiplier[:, :, None] * u
        uf = u.reshape(Ny * Nx, nb, order="F")
        mask = np.asarray(obs["mask"][ids, :, fi], np.float64).T
        observed = np.asarray(obs["Y"][ids, :, fi], np.complex128).T * mask
        predicted = uf[rx, :] * mask
        den = np.sum(np.abs(predicted) ** 2, axis=0)
        num = np.sum(np.conj(predicted) * observed, axis=0)
        if not (np.all(np.isfinite(den) & (den > np.finfo(np.float64).tiny)) and np.all(np.isfinite(num))):
            raise H.WustError("WUST:NumericalFailure", "Unusable source-scale denominator/numerator")
        scale = num / den
        resid = (predicted * scale - observed) * mask
        loss += 0.5 * float(np.sum(np.abs(resid) ** 2))
        if need_gradient:
            rhs = np.zeros((Ny * Nx, nb), np.complex128)
            rhs[rx, :] = resid
            lam = np.zeros((Ny, Nx, nb), np.complex128)
            lam[1:-1, 1:-1] = s.solve_interior(rhs.reshape(Ny, Nx, nb, order="F")[1:-1, 1:-1, :], True)
            lam = s.mass_apply(lam, True)
            grad += np.sum(-np.real(np.conj(v * scale[None, None, :]) * lam), axis=2)
    if nargout == 1:
        return loss
    return loss, grad
