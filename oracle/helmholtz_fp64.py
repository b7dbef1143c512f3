"""FP64 NumPy/SciPy restatement of the reference Helmholtz path (the oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py`` for who may import this.

Array conventions follow the reference: grids are ``[y, x]`` (rows = y),
multi-source fields are ``[y, x, source]`` and linear node indices are
MATLAB-style ``y + Ny*x`` (column-major), 1-based where they are passed in
as ``tx_index`` / ``rx_index`` (Runtime/matlab/wust_geometry.m:27).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class WustError(RuntimeError):
    """Error carrying one of the reference's ``WUST:*`` identifiers."""

    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.identifier = ident


# ---------------------------------------------------------------------------
# a1: optimal 9-point stencil mass coefficients
# ---------------------------------------------------------------------------
def stencil_opt_params(vmin, vmax, f, h, g):
    """Least-squares mass coefficients (d, e) with b = 5/6 fixed.

    Follows Runtime/solver/stencilOptParams.m:12-37: 100 angles in [0, pi/4]
    times 10 samples of points-per-wavelength G spaced uniformly in 1/G
    between vmin/(f h) and vmax/(f h); solve [S2 S3][d;e] = S4 - b S1 in the
    least-squares sense.
    """
    angle_samples, wavelength_samples = 100, 10
    gmin = vmin / (f * h)
    gmax = vmax / (f * h)
    m = np.arange(1, angle_samples + 1)
    n = np.arange(1, wavelength_samples + 1)
    theta = (m - 1) * np.pi / (4 * (angle_samples - 1))
    G = 1.0 / (1.0 / gmax + ((n - 1) / (wavelength_samples - 1)) * (1.0 / gmin - 1.0 / gmax))
    TH, GG = np.meshgrid(theta, G)
    P = np.cos(g * 2 * np.pi * np.cos(TH) / GG)
    Q = np.cos(2 * np.pi * np.sin(TH) / GG)
    S1 = (1 + 1 / g**2) * GG**2 * (1 - P - Q + P * Q)
    S2 = np.pi**2 * (2 - P - Q)
    S3 = 2 * np.pi**2 * (1 - P * Q)
    S4 = 2 * np.pi**2 + GG**2 * ((1 + 1 / g**2) * P * Q - P - Q / g**2)
    b = 5.0 / 6.0
    A = np.stack([S2.ravel(order="F"), S3.ravel(order="F")], axis=1)
    y = S4.ravel(order="F") - b * S1.ravel(order="F")
    params, *_ = np.linalg.lstsq(A, y, rcond=None)
    return b, float(params[0]), float(params[1])


# ---------------------------------------------------------------------------
# a3: PML stretch, staggered A/B/C, wavenumber and dq/ds
# ---------------------------------------------------------------------------
@dataclass
class PmlTerms:
    A: np.ndarray  # [Ny, Nx-1]
    B: np.ndarray  # [Ny-1, Nx]
    C: np.ndarray  # [Ny, Nx]
    k: np.ndarray  # [Ny, Nx]
    dqds: np.ndarray  # [Ny, Nx]  d(k^2)/d(slowness)
    h: float
    g: float


def pml_terms(x, y, vel, atten, f, sign_convention, a0, L_pml):
    """Complex speed, wavenumber and fine-mesh PML terms.

    Follows Runtime/solver/HelmholtzSolver.m:55-101 (velocity :67-68,
    k :71, sx/sy :74-82, ex/ey :83-84, A/B/C :85-90, q' :101).
    """
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    vel = np.asarray(vel, dtype=np.float64)
    atten = np.asarray(atten, dtype=np.float64)
    h = float(np.mean(np.diff(x)))
    gh = float(np.mean(np.diff(y)))
    g = gh / h
    Nx, Ny = x.size, y.size
    xmin, xmax, ymin, ymax = x.min(), x.max(), y.min(), y.max()
    sgn = float(np.sign(sign_convention))
    SI = atten / (2 * np.pi)
    velocity = 1.0 / (1.0 / vel + 1j * SI * sgn)
    k = 2 * np.pi * f / velocity
    xe = np.linspace(xmin, xmax, 2 * (Nx - 1) + 1)
    ye = np.linspace(ymin, ymax, 2 * (Ny - 1) + 1)
    Xe, Ye = np.meshgrid(xe, ye)
    xctr, xspan = (xmin + xmax) / 2, (xmax - xmin) / 2
    yctr, yspan = (ymin + ymax) / 2, (ymax - ymin) / 2
    sx = 2 * np.pi * a0 * f * (np.maximum(np.abs(Xe - xctr) - xspan + L_pml, 0) / L_pml) ** 2
    sy = 2 * np.pi * a0 * f * (np.maximum(np.abs(Ye - yctr) - yspan + L_pml, 0) / L_pml) ** 2
    ex = 1 + 1j * sx * sgn / (2 * np.pi * f)
    ey = 1 + 1j * sy * sgn / (2 * np.pi * f)
    A = (ey / ex)[0::2, 1::2]
    B = (ex / ey)[1::2, 0::2]
    C = (ex * ey)[0::2, 0::2]
    dqds = 8 * np.pi**2 * f**2 / velocity
    return PmlTerms(A=A, B=B, C=C, k=k, dqds=dqds, h=h, g=g)


# ---------------------------------------------------------------------------
# a2 (optional): k-Wave matched LDR9 wavenumber
# ---------------------------------------------------------------------------
_GL8_NODES = np.array([-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                       -0.1834346424956498, 0.1834346424956498, 0.5255324099163290,
                       0.7966664774136267, 0.9602898564975363])
_GL8_WEIGHTS = 0.5 * np.array([0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                               0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                               0.2223810344533745, 0.1012285362903763])


def kwave_ldr9_wavenumber_squared(c, f, h, b, d, e, dt, cr):
    """q_eff and dq/ds matching k-Wave's time dispersion.

    Follows Runtime/solver/kwaveLdr9WavenumberSquared.m:7-64: an 8-point
    Gauss-Legendre average over theta in [0, pi/4] of the stencil symbol,
    Q = -N/D with derivative by the chain rule through asin.
    (The quadrature nodes/weights are the standard 8-point Gauss-Legendre
    rule; they are a published table, kept identical.)
    """
    c = np.asarray(c, dtype=np.float64)
    if np.any(c <= 0) or f <= 0 or h <= 0 or dt <= 0 or cr <= 0:
        raise ValueError("All physical inputs must be positive.")
    omega = 2 * np.pi * f
    ts = math.sin(omega * dt / 2)
    z = (cr / c) * ts
    if np.any(np.abs(z) >= 1):
        raise ValueError("k-Wave local dispersion target is evanescent")
    K = (2 * h / (cr * dt)) * np.arcsin(z)
    theta = (_GL8_NODES + 1) * (np.pi / 8)
    N = np.zeros_like(K)
    D = np.zeros_like(K)
    Np = np.zeros_like(K)
    Dp = np.zeros_like(K)
    for j in range(theta.size):
        ct, st = math.cos(theta[j]), math.sin(theta[j])
        X, Y = K * ct, K * st
        cx, cy, sx, sy = np.cos(X), np.cos(Y), np.sin(X), np.sin(Y)
        lam = -4 * b + 2 * (2 * b - 1) * (cx + cy) + 4 * (1 - b) * cx * cy
        mass = 1 - d - e + (d / 2) * (cx + cy) + e * cx * cy
        lamp = -2 * (2 * b - 1) * (sx * ct + sy * st) - 4 * (1 - b) * (sx * ct * cy + cx * sy * st)
        massp = -(d / 2) * (sx * ct + sy * st) - e * (sx * ct * cy + cx * sy * st)
        w = _GL8_WEIGHTS[j]
        N += w * lam * mass
        D += w * mass * mass
        Np += w * (lamp * mass + lam * massp)
        Dp += w * (2 * mass * massp)
    Q = -N / D
    Qp = -(Np * D - N * Dp) / (D * D)
    q = Q / (h * h)
    dKds = (2 * h * ts) / (dt * np.sqrt(1 - z * z))
    dq = (Qp / (h * h)) * dKds
    if np.any(~np.isfinite(q) | (q <= 0)) or np.any(~np.isfinite(dq)):
        raise ValueError("LDR-9 produced invalid wavenumber or derivative.")
    return q, dq


# ---------------------------------------------------------------------------
# a4: the nine stencil coefficient arrays
# ---------------------------------------------------------------------------
def nine_point_coefficients(A, B, C, k, b, d, e, h, g, Nx, Ny):
    """Return dict of Dd, Dl, Du, Ld, Ll, Lu, Ud, Ul, Uu (complex128).

    Follows Runtime/solver/ninePointCoefficients.m:22-72 (shapes: Dd [Ny,Nx];
    Dl, Du [Ny-1,Nx]; Ld, Ud [Ny,Nx-1]; Ll, Lu, Ul, Uu [Ny-1,Nx-1]).
    """
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    C = np.asarray(C, np.complex128)
    k2 = np.asarray(k, np.complex128) ** 2
    h2, g2 = h * h, g * g
    Dd = np.zeros((Ny, Nx), np.complex128)
    Dl = np.zeros((Ny - 1, Nx), np.complex128)
    Du = np.zeros((Ny - 1, Nx), np.complex128)
    Ld = np.zeros((Ny, Nx - 1), np.complex128)
    Ll = np.zeros((Ny - 1, Nx - 1), np.complex128)
    Lu = np.zeros((Ny - 1, Nx - 1), np.complex128)
    Ud = np.zeros((Ny, Nx - 1), np.complex128)
    Ul = np.zeros((Ny - 1, Nx - 1), np.complex128)
    Uu = np.zeros((Ny - 1, Nx - 1), np.complex128)
    Dd[[0, Ny - 1], :] = 1
    Dd[1:Ny - 1, [0, Nx - 1]] = 1
    y = slice(1, Ny - 1)      # MATLAB 2:Ny-1
    ym = slice(0, Ny - 2)     # MATLAB 1:Ny-2 (= y-1)
    yp = slice(2, Ny)         # MATLAB 3:Ny   (= y+1)
    x = slice(1, Nx - 1)      # MATLAB 2:Nx-1
    xm = slice(0, Nx - 2)     # x-1
    xp = slice(2, Nx)         # x+1
    Dd[y, x] = (1 - d - e) * C[y, x] * k2[y, x] - b * (
        A[y, x] + A[y, xm] + B[y, x] / g2 + B[ym, x] / g2) / h2
    Dl[ym, x] = (b * B[ym, x] / g2 - ((1 - b) / 2) * (A[ym, x] + A[ym, xm])) / h2 \
        + (d / 4) * C[ym, x] * k2[ym, x]
    Du[y, x] = (b * B[y, x] / g2 - ((1 - b) / 2) * (A[yp, x] + A[yp, xm])) / h2 \
        + (d / 4) * C[yp, x] * k2[yp, x]
    xl = slice(0, Nx - 2)     # MATLAB 1:Nx-2
    Ld[y, xl] = (b * A[y, xl] - ((1 - b) / 2) * (B[y, xl] / g2 + B[ym, xl] / g2)) / h2 \
        + (d / 4) * C[y, xl] * k2[y, xl]
    Ll[ym, xl] = ((1 - b) / 2) * (A[ym, xl] + B[ym, xl] / g2) / h2 + (e / 4) * C[ym, xl] * k2[ym, xl]
    Lu[y, xl] = ((1 - b) / 2) * (A[yp, xl] + B[y, xl] / g2) / h2 + (e / 4) * C[yp, xl] * k2[yp, xl]
    xu = slice(1, Nx - 1)     # MATLAB 2:Nx-1
    xu1 = slice(2, Nx)        # xu+1
    Ud[y, xu] = (b * A[y, xu] - ((1 - b) / 2) * (B[y, xu1] / g2 + B[ym, xu1] / g2)) / h2 \
        + (d / 4) * C[y, xu1] * k2[y, xu1]
    Ul[ym, xu] = ((1 - b) / 2) * (A[ym, xu] + B[ym, xu1] / g2) / h2 + (e / 4) * C[ym, xu1] * k2[ym, xu1]
    Uu[y, xu] = ((1 - b) / 2) * (A[yp, xu] + B[y, xu1] / g2) / h2 + (e / 4) * C[yp, xu1] * k2[yp, xu1]
    return dict(Dd=Dd, Dl=Dl, Du=Du, Ld=Ld, Ll=Ll, Lu=Lu, Ud=Ud, Ul=Ul, Uu=Uu)


def assemble_sparse(coef, Ny, Nx):
    """Sparse H with identity Dirichlet rows (Runtime/solver/HelmholtzSolver.m:114-138).

    Row (y,x) interior couples to 9 nodes with the coefficient lookups of
    :125-128; linear index y + Ny*x (0-based here).
    """
    yy, xx = np.meshgrid(np.arange(1, Ny - 1), np.arange(1, Nx - 1), indexing="ij")
    yy = yy.ravel()
    xx = xx.ravel()
    row = yy + Ny * xx
    entries = [
        (0, 0, coef["Dd"][yy, xx]),
        (-1, 0, coef["Dl"][yy - 1, xx]),
        (1, 0, coef["Du"][yy, xx]),
        (0, -1, coef["Ld"][yy, xx - 1]),
        (-1, -1, coef["Ll"][yy - 1, xx - 1]),
        (1, -1, coef["Lu"][yy, xx - 1]),
        (0, 1, coef["Ud"][yy, xx]),
        (-1, 1, coef["Ul"][yy - 1, xx]),
        (1, 1, coef["Uu"][yy, xx]),
    ]
    rows, cols, vals = [], [], []
    for dy, dx, v in entries:
        rows.append(row)
        cols.append((yy + dy) + Ny * (xx + dx))
        vals.append(v)
    by, bx = np.meshgrid(np.arange(Ny), np.arange(Nx), indexing="ij")
    border = (by == 0) | (by == Ny - 1) | (bx == 0) | (bx == Nx - 1)
    brow = (by + Ny * bx)[border]
    rows.append(brow)
    cols.append(brow)
    vals.append(np.ones(brow.size, np.complex128))
    N = Nx * Ny
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N, N))


def mass_kernel(d, e):
    """3x3 mass-stencil kernel of Runtime/solver/HelmholtzSolver.m:210."""
    return np.array([[e / 4, d / 4, e / 4], [d / 4, 1 - d - e, d / 4], [e / 4, d / 4, e / 4]])


def conv_same_3x3(v, kern):
    """MATLAB ``convn(v, kern, 'same')`` over the first two axes, zero padding."""
    v = np.asarray(v)
    out = np.zeros_like(v)
    Ny, Nx = v.shape[0], v.shape[1]
    pad = np.zeros((Ny + 2, Nx + 2) + v.shape[2:], dtype=v.dtype)
    pad[1:-1, 1:-1] = v
    kr = kern[::-1, ::-1]  # convolution flips the kernel (it is symmetric anyway)
    for a in range(3):
        for b in range(3):
            if kr[a, b] != 0:
                out += kr[a, b] * pad[a:a + Ny, b:b + Nx]
    return out


# ---------------------------------------------------------------------------
# a5/a8/a9: the operator, its factorization, solve and mass derivative
# ---------------------------------------------------------------------------
class HelmholtzFP64:
    """CPU FP64 HelmholtzSolver (Runtime/solver/HelmholtzSolver.m:32-220, cpu backend).

    ``sign_convention`` and the other constructor arguments keep the
    reference's meaning; ``options`` needs ``stencil_bounds`` and optionally
    ``dispersion`` = dict(time_step_s, reference_speed_mps).
    """

    def __init__(self, x, y, vel, atten, f, sign_convention, a0, L_pml, options):
        if not isinstance(options, dict) or "stencil_bounds" not in options:
            raise WustError("WUST:InvalidInput", "Explicit backend and fixed stencil_bounds are required")
        bounds = np.asarray(options["stencil_bounds"], dtype=np.float64).ravel()
        if bounds.size != 2 or not np.all(np.isfinite(bounds)) or bounds[0] <= 0 or bounds[1] <= bounds[0]:
            raise WustError("WUST:InvalidInput", "stencil_bounds must be a positive ordered [min,max] pair")
        vel = np.asarray(vel, dtype=np.float64)
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        if vel.shape != (y.size, x.size) or min(vel.shape) < 3 or not np.all(np.isfinite(vel) & (vel > 0)):
            raise WustError("WUST:InvalidInput", "Expected a positive finite [ny,nx] medium")
        Ny, Nx = vel.shape
        self.Ny, self.Nx = Ny, Nx
        t = pml_terms(x, y, vel, atten, f, sign_convention, a0, L_pml)
        b, d, e = stencil_opt_params(bounds[0], bounds[1], f, t.h, t.g)
        self.b, self.massD, self.massE = b, d, e
        k = t.k
        dqds = t.dqds
        disp = options.get("dispersion")
        if disp:
            q, dq = kwave_ldr9_wavenumber_squared(vel, f, t.h, b, d, e,
                                                  disp["time_step_s"], disp["reference_speed_mps"])
            k = np.sqrt(q.astype(np.complex128))
            dqds = dq
        self.coef = nine_point_coefficients(t.A, t.B, t.C, k, b, d, e, t.h, t.g, Nx, Ny)
        self.virtual_source_multiplier = t.C * dqds
        self.H = assemble_sparse(self.coef, Ny, Nx)
        self.lu = spla.splu(self.H)
        self._lu_adj = None

    def solve(self, src, adjoint=False):
        """Return (u, v) with H u = src (or H^H u = src) and v = (C q') u (:174-203)."""
        src = np.asarray(src)
        squeeze = src.ndim == 2
        if squeeze:
            src = src[:, :, None]
        Ny, Nx, R = src.shape
        rhs = src.reshape(Ny * Nx, R, order="F").astype(np.complex128)
        if adjoint:
            # H^H gets its own factorization: SuperLU's trans='H' solve leaves
            # residuals near 1e-10 on these operators, far above FP64 round-off.
            if self._lu_adj is None:
                self._lu_adj = spla.splu(self.H.conj().T.tocsc())
            sol = self._lu_adj.solve(rhs)
        else:
            sol = self.lu.solve(rhs)
        u = sol.reshape(Ny, Nx, R, order="F")
        v = self.virtual_source_multiplier[:, :, None] * u
        if squeeze:
            return u[:, :, 0], v[:, :, 0]
        return u, v

    def mass_apply(self, v, adjoint=False):
        """Exact mass-stencil derivative (Runtime/solver/HelmholtzSolver.m:205-220)."""
        v = np.array(v, dtype=np.complex128, copy=True)
        kern = mass_kernel(self.massD, self.massE)
        if adjoint:
            v[[0, -1], ...] = 0
            v[:, [0, -1], ...] = 0
        out = conv_same_3x3(v, kern)
        if not adjoint:
            out[[0, -1], ...] = 0
            out[:, [0, -1], ...] = 0
        return out


# ---------------------------------------------------------------------------
# a10-a12: loss / gradient oracle; a13: directional curvature
# ---------------------------------------------------------------------------
@dataclass
class OracleState:
    solver: HelmholtzFP64
    virtual: list = field(default_factory=list)
    scales: list = field(default_factory=list)
    forward_rhs: int = 0
    adjoint_rhs: int = 0


def wust_oracle(c, obs, fi, cfg, nargout=2):
    """Source-projected least squares at one frequency (Runtime/matlab/wust_oracle.m:1-86).

    ``fi`` is a 0-based frequency index.  ``nargout`` mirrors MATLAB: 1 =
    loss only, 2 = loss and gradient, 3 = also retain state for curvature.
    Returns ``loss`` / ``(loss, g)`` / ``(loss, g, state)``.
    """
    need_gradient = nargout > 1
    retain = nargout > 2
    Ny, Nx = obs["shape_yx"]
    nt = obs["n_tx"]
    disp = cfg.get("dispersion") if cfg.get("wavenumber") == "kwave-ldr9" else None
    solver = HelmholtzFP64(obs["x_m"], obs["y_m"], np.asarray(c, np.float64), np.zeros((Ny, Nx)),
                           obs["frequencies_hz"][fi], -1, cfg["pml_strength"], cfg["pml_m"],
                           dict(stencil_bounds=cfg["stencil_bounds"], dispersion=disp))
    tx = np.asarray(obs["tx_index"], np.int64) - 1
    rx = np.asarray(obs["rx_index"], np.int64) - 1
    grad = np.zeros((Ny, Nx)) if need_gradient else None
    loss = 0.0
    state = OracleState(solver=solver)
    bs = int(cfg["source_batch_size"])
    for first in range(0, nt, bs):
        ids = np.arange(first, min(first + bs, nt))
        nb = ids.size
        src = np.zeros((Ny * Nx, nb), np.complex128)
        src[tx[ids], np.arange(nb)] = 1
        u, v = solver.solve(src.reshape(Ny, Nx, nb, order="F"), False)
        uf = u.reshape(Ny * Nx, nb, order="F")
        mask = np.asarray(obs["mask"][ids, :, fi], np.float64).T
        observed = np.asarray(obs["Y"][ids, :, fi], np.complex128).T
        predicted = uf[rx, :] * mask
        observed = observed * mask
        den = np.sum(np.abs(predicted) ** 2, axis=0)
        num = np.sum(np.conj(predicted) * observed, axis=0)
        if not (np.all(np.isfinite(den) & (den > np.finfo(np.float64).tiny)) and np.all(np.isfinite(num))):
            raise WustError("WUST:NumericalFailure", "Unusable source-scale denominator/numerator")
        scale = num / den
        if not np.all(np.isfinite(scale)):
            raise WustError("WUST:NumericalFailure", "Nonfinite source scale")
        resid = (predicted * scale - observed) * mask
        loss += 0.5 * float(np.sum(np.abs(resid) ** 2))
        if need_gradient:
            rhs = np.zeros((Ny * Nx, nb), np.complex128)
            rhs[rx, :] = resid
            lam, _ = solver.solve(rhs.reshape(Ny, Nx, nb, order="F"), True)
            lam = solver.mass_apply(lam, True)
            grad += np.sum(-np.real(np.conj(v * scale[None, None, :]) * lam), axis=2)
        if retain:
            state.virtual.append(v)
            state.scales.append(scale)
    if not np.isfinite(loss) or (need_gradient and not np.all(np.isfinite(grad))):
        raise WustError("WUST:NumericalFailure", "Nonfinite objective/gradient")
    state.forward_rhs = nt
    state.adjoint_rhs = nt if need_gradient else 0
    if nargout == 1:
        return loss
    if nargout == 2:
        return loss, grad
    return loss, grad, state


def directional_curvature(state, direction, rx_index, receiver_mask):
    """kappa = sum_b || -(H^-1 M(v_b . delta))[rx] a_b . mask ||^2 (Runtime/matlab/directional_curvature.m:1-26)."""
    rx = np.asarray(rx_index, np.int64) - 1
    receiver_mask = np.asarray(receiver_mask, np.float64)
    curv = 0.0
    offset = 0
    for v, a in zip(state.virtual, state.scales):
        nb = a.size
        ids = np.arange(offset, offset + nb)
        offset += nb
        src = state.solver.mass_apply(v * np.asarray(direction)[:, :, None], False)
        u, _ = state.solver.solve(src, False)
        uf = u.reshape(-1, nb, order="F")
        deriv = -uf[rx, :] * a[None, :] * receiver_mask[ids, :].T
        curv += float(np.sum(np.abs(deriv) ** 2))
    return curv


# ---------------------------------------------------------------------------
# geometry snapping (Runtime/matlab/wust_geometry.m:1-31)
# ---------------------------------------------------------------------------
def geometry(shape_yx, origin_yx_m, spacing_yx_m, tx_xy, rx_xy):
    """Pixel-centre coordinates and nearest-node 1-based indices, lowest index on ties."""
    ny, nx = int(shape_yx[0]), int(shape_yx[1])
    y = origin_yx_m[0] + (np.arange(ny) + 0.5) * spacing_yx_m[0]
    x = origin_yx_m[1] + (np.arange(nx) + 0.5) * spacing_yx_m[1]

    def snap(xy):
        xy = np.asarray(xy, np.float64)
        ix = np.argmin(np.abs(x[:, None] - xy[None, :, 0]), axis=0)
        iy = np.argmin(np.abs(y[:, None] - xy[None, :, 1]), axis=0)
        idx = iy + ny * ix + 1
        if np.unique(idx).size != idx.size:
            raise WustError("WUST:InvalidInput", "Distinct elements in one array collapse onto a grid sample")
        return idx

    return dict(x_m=x, y_m=y, tx_index=snap(tx_xy), rx_index=snap(rx_xy))
