"""Host-side mirror of the reference's Helmholtz interface, on the HIP engine.

Same names, argument meaning and error identifiers as the reference:

* :class:`HelmholtzSolver` -- ``HelmholtzSolver(x, y, vel, atten, f,
  signConvention, a0, L_PML, options)`` with ``solve(src, adjoint)`` and
  ``massApply(v, adjoint)`` (Runtime/solver/HelmholtzSolver.m:32-220).
* :func:`wust_oracle` -- loss / gradient / retained state of one frequency
  (Runtime/matlab/wust_oracle.m:1-86), ``nargout`` selecting the work.
* :func:`directional_curvature` (Runtime/matlab/directional_curvature.m).
* :class:`WustOracleFunction` / :func:`wust_loss` -- the oracle as a torch
  ``autograd.Function`` of slowness (loss forward, gradient backward).

Arrays follow the reference's MATLAB conventions: grids are ``[y, x]``,
fields ``[y, x, source]``, node indices 1-based ``y + Ny*x + 1``.  Inputs may
be NumPy arrays or torch tensors; results are torch tensors on the GPU
(``cuda:0`` unless the inputs live elsewhere), like the reference's GPU
backend returning gpuArrays.  Everything runs through the C ABI of
``_lib/libwust_hip.so``; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as N
from ._native import WustError

DEFAULT_TOL = 1e-10
DEFAULT_MAX_ITER = 30


def _stream(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dev(device=None):
    if device is None:
        if not torch.cuda.is_available():
            raise WustError("WUST:InvalidInput", "the HIP engine needs a GPU (backend 'gpu')")
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def _grid_to_dev(a, device, dtype=torch.float64):
    """[Ny, Nx] array -> contiguous device tensor in column-major order (shape (Nx, Ny))."""
    t = torch.as_tensor(a, device=device)
    t = t.to(dtype)
    return t.t().contiguous()


def _field_to_dev(a, device, ny, nx, message):
    """[Ny, Nx, R] (or [Ny, Nx]) complex -> contiguous complex128 (R, Nx, Ny).

    The engine reads Ny * Nx * R values from the result, so fields whose leading
    dimensions are not the operator's grid are rejected here with ``message``
    (Runtime/solver/applyBlockLU.cu:380-388 rejects them as WUST:InvalidInput)."""
    t = torch.as_tensor(a, device=device)
    if t.dim() == 2:
        t = t.unsqueeze(-1)
    if t.dim() != 3 or tuple(t.shape[:2]) != (ny, nx) or t.shape[2] < 1:
        raise WustError("WUST:InvalidInput", message)
    return t.to(torch.complex128).permute(2, 1, 0).contiguous()


def _check_shape(a, shape, message):
    if tuple(a.shape) != tuple(shape):
        raise WustError("WUST:InvalidInput", message)


def _field_from_dev(t, squeeze):
    out = t.permute(2, 1, 0)
    return out[:, :, 0] if squeeze else out


def _uniform(v, name):
    v = np.asarray(v, dtype=np.float64).ravel()
    if v.size < 3 or not np.all(np.isfinite(v)):
        raise WustError("WUST:InvalidInput", "Expected a positive finite [ny,nx] medium")
    dv = np.diff(v)
    h = float(np.mean(dv))
    if not (np.all(dv > 0) and np.max(np.abs(dv - h)) < h * 1e-7):
        raise WustError("WUST:InvalidInput", "Uniform increasing grids required")
    return float(v[0]), h


class HelmholtzSolver:
    """Nine-point PML Helmholtz operator and its factorization on the GPU.

    ``options`` must hold ``backend`` ('gpu') and ``stencil_bounds`` [min, max];
    optional ``dispersion`` = dict(time_step_s, reference_speed_mps) selects
    kwave-ldr9, ``mode`` in {'refined', 'c64', 'plain'} selects the arithmetic
    (default 'refined'), ``factor`` in {'auto', 'fixed', 'gauss-jordan'} how the
    engine inverts Schur blocks (default 'auto': Newton-Schulz from the neighbouring
    inverse with a Gauss-Jordan fallback, round budgets learned from the previous
    factorization of the same shape and frequency; 'fixed': the same with the
    default budgets, so that the factor depends only on the operator), and
    ``tol`` / ``max_iter`` the refinement stop.
    """

    def __init__(self, x, y, vel, atten, f, signConvention, a0, L_PML, options):
        if not isinstance(options, dict) or "backend" not in options or "stencil_bounds" not in options:
            raise WustError("WUST:InvalidInput", "Explicit backend and fixed stencil_bounds are required")
        bounds = np.asarray(options["stencil_bounds"], dtype=np.float64).ravel() \
            if options["stencil_bounds"] is not None else np.zeros(0)
        if bounds.size != 2 or not np.all(np.isfinite(bounds)) or bounds[0] <= 0 or bounds[1] <= bounds[0]:
            raise WustError("WUST:InvalidInput", "stencil_bounds must be a positive ordered [min,max] pair")
        if str(options["backend"]) not in ("cpu", "gpu"):
            raise WustError("WUST:InvalidInput", "backend must be cpu or gpu")
        if str(options["backend"]) != "gpu":
            raise WustError("WUST:InvalidInput", "this engine implements the gpu backend only")
        self.backend = "gpu"
        mode = options.get("mode", "refined")
        if mode not in N.MODES:
            raise WustError("WUST:InvalidInput", f"unknown mode {mode!r}")
        self.mode = mode
        factor = options.get("factor", "auto")
        if factor not in N.FACTORS:
            raise WustError("WUST:InvalidInput", f"unknown factor method {factor!r}")
        if factor != "auto" and mode == "plain":
            raise WustError("WUST:InvalidInput", "the plain mode has a single factor method")
        self.factor = factor
        self.tol = float(options.get("tol", DEFAULT_TOL))
        self.max_iter = int(options.get("max_iter", DEFAULT_MAX_ITER))
        device = _dev(options.get("device"))
        self.device = device
        vel_t = torch.as_tensor(vel)
        ny, nx = int(np.asarray(y).size), int(np.asarray(x).size)
        if tuple(vel_t.shape) != (ny, nx) or min(ny, nx) < 3:
            raise WustError("WUST:InvalidInput", "Expected a positive finite [ny,nx] medium")
        vel_d = _grid_to_dev(vel_t, device)
        if not bool(torch.all(torch.isfinite(vel_d) & (vel_d > 0))):
            raise WustError("WUST:InvalidInput", "Expected a positive finite [ny,nx] medium")
        att_d = None
        if atten is not None:
            att_t = torch.as_tensor(atten)
            if tuple(att_t.shape) != (ny, nx):
                raise WustError("WUST:InvalidInput", "Expected a positive finite [ny,nx] medium")
            if bool(torch.any(att_t != 0)):
                att_d = _grid_to_dev(att_t, device)
        x0, dx = _uniform(x, "x")
        y0, dy = _uniform(y, "y")
        desc = N.GridDesc()
        desc.ny, desc.nx = ny, nx
        desc.x0, desc.dx, desc.y0, desc.dy = x0, dx, y0, dy
        desc.frequency_hz = float(f)
        desc.sign_convention = float(signConvention)
        desc.pml_strength = float(a0)
        desc.pml_m = float(L_PML)
        desc.stencil_vmin, desc.stencil_vmax = float(bounds[0]), float(bounds[1])
        disp = options.get("dispersion")
        if disp:
            if att_d is not None or abs(dy / dx - 1) >= 1e-10:
                raise WustError("WUST:InvalidInput", "LDR9 requires zero attenuation and square grid cells")
            desc.dispersion = 1
            desc.ldr9_time_step_s = float(disp["time_step_s"])
            desc.ldr9_reference_speed_mps = float(disp["reference_speed_mps"])
        self.ny, self.nx = ny, nx
        self._lib = N.load()
        h = C.c_void_p()
        t0 = time.perf_counter()
        with torch.cuda.device(device):
            N.check(self._lib.wust_helmholtz_create(
                C.byref(desc), C.c_void_p(vel_d.data_ptr()),
                C.c_void_p(att_d.data_ptr() if att_d is not None else None),
                N.MODES[mode] | N.FACTORS[factor], _stream(device), C.byref(h)))
        self._h = h
        self.wall_seconds = time.perf_counter() - t0
        info = self.info()
        self.massD, self.massE = info["d"], info["e"]
        self.stencil_b = info["b"]
        self.assembly_seconds = info["assembly_ms"] / 1e3
        self.factor_seconds = info["factor_ms"] / 1e3
        self.last_solve_stats = None

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            N.check(self._lib.wust_helmholtz_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _handle(self):
        if not self._h:
            raise WustError("WUST:InvalidInput", "solver has been released")
        return self._h

    def info(self):
        inf = N.Info()
        N.check(self._lib.wust_helmholtz_info(self._handle(), C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in N.Info._fields_}

    # -- operator pieces -----------------------------------------------------------
    def coefficients(self):
        """The nine arrays of ninePointCoefficients (complex128 tensors, reference shapes)."""
        ny, nx = self.ny, self.nx
        shapes = dict(Dd=(ny, nx), Dl=(ny - 1, nx), Du=(ny - 1, nx), Ld=(ny, nx - 1), Ll=(ny - 1, nx - 1),
                      Lu=(ny - 1, nx - 1), Ud=(ny, nx - 1), Ul=(ny - 1, nx - 1), Uu=(ny - 1, nx - 1))
        bufs = {k: torch.zeros((s[1], s[0]), dtype=torch.complex128, device=self.device) for k, s in shapes.items()}
        with torch.cuda.device(self.device):
            N.check(self._lib.wust_helmholtz_coefficients(
                self._handle(), *[C.c_void_p(bufs[k].data_ptr()) for k in
                                  ("Dd", "Dl", "Du", "Ld", "Ll", "Lu", "Ud", "Ul", "Uu")],
                _stream(self.device)))
        return {k: v.t() for k, v in bufs.items()}

    def solve(self, src, adjoint=False, tol=None, max_iter=None, return_stats=False):
        """[u, v] = solve(src, adjoint): H u = src or H^H u = src; v = (C q') u."""
        squeeze = torch.as_tensor(src).dim() == 2
        s = _field_to_dev(src, self.device, self.ny, self.nx, "Inconsistent Block-LU input shapes.")
        R = s.shape[0]
        out = torch.empty_like(s)
        st = N.SolveStats()
        with torch.cuda.device(self.device):
            N.check(self._lib.wust_helmholtz_solve(
                self._handle(), C.c_void_p(s.data_ptr()), C.c_void_p(out.data_ptr()), R, int(bool(adjoint)),
                float(self.tol if tol is None else tol), int(self.max_iter if max_iter is None else max_iter),
                _stream(self.device), C.byref(st)))
            self.last_solve_stats = {k: getattr(st, k) for k, _ in N.SolveStats._fields_}
            v = torch.empty_like(out)
            N.check(self._lib.wust_helmholtz_virtual_sources(
                self._handle(), C.c_void_p(out.data_ptr()), C.c_void_p(v.data_ptr()), R, _stream(self.device)))
        u = _field_from_dev(out, squeeze)
        vv = _field_from_dev(v, squeeze)
        if return_stats:
            return u, vv, self.last_solve_stats
        return u, vv

    def massApply(self, v, adjoint=False):
        """Exact mass-stencil derivative (output border zeroed; adjoint zeroes the input border)."""
        squeeze = torch.as_tensor(v).dim() == 2
        s = _field_to_dev(v, self.device, self.ny, self.nx, "massApply expects an [ny, nx] or [ny, nx, n] field")
        out = torch.empty_like(s)
        with torch.cuda.device(self.device):
            N.check(self._lib.wust_helmholtz_mass_apply(
                self._handle(), C.c_void_p(s.data_ptr()), C.c_void_p(out.data_ptr()), s.shape[0],
                int(bool(adjoint)), _stream(self.device)))
        return _field_from_dev(out, squeeze)

    mass_apply = massApply


@dataclass
class OracleState:
    """Retained per-call state for :func:`directional_curvature` (the reference's ``cache``)."""

    solver: HelmholtzSolver
    scales: np.ndarray
    forward_rhs: int
    adjoint_rhs: int
    forward_seconds: float = 0.0
    adjoint_seconds: float = 0.0
    extra: dict = field(default_factory=dict)


def _obs_slice(obs, fi, device):
    Y = np.asarray(obs["Y"])[:, :, fi] if not torch.is_tensor(obs["Y"]) else obs["Y"][:, :, fi]
    mask = np.asarray(obs["mask"])[:, :, fi] if not torch.is_tensor(obs["mask"]) else obs["mask"][:, :, fi]
    Yd = torch.as_tensor(Y, device=device).to(torch.complex128).t().contiguous()   # [nrx, ntx] = col-major [ntx, nrx]
    Md = torch.as_tensor(mask, device=device).to(torch.uint8).t().contiguous()
    return Yd, Md


def _check_obs_slice(Yd, Md, ntx, nrx):
    # the engine reads n_tx * n_rx observations and mask entries
    _check_shape(Yd, (nrx, ntx), "obs.Y must be [n_tx, n_rx, n_freq]")
    _check_shape(Md, (nrx, ntx), "obs.mask must be [n_tx, n_rx, n_freq]")


def wust_oracle(c, obs, fi, cfg, nargout=2):
    """Loss / gradient / retained state at frequency index ``fi`` (0-based).

    Mirrors Runtime/matlab/wust_oracle.m: ``nargout`` 1 runs forward solves only,
    2 adds the adjoint solves and the slowness gradient, 3 also retains the
    forward fields and source scales for :func:`directional_curvature`.
    """
    need_gradient = nargout > 1
    retain = nargout > 2
    ny, nx = int(obs["shape_yx"][0]), int(obs["shape_yx"][1])
    disp = cfg.get("dispersion") if cfg.get("wavenumber") == "kwave-ldr9" else None
    options = dict(backend=cfg.get("backend", "gpu"), stencil_bounds=cfg["stencil_bounds"], dispersion=disp,
                   mode=cfg.get("mode", "refined"), tol=cfg.get("tol", DEFAULT_TOL),
                   max_iter=cfg.get("max_iter", DEFAULT_MAX_ITER), factor=cfg.get("factor", "auto"))
    if "device" in cfg:
        options["device"] = cfg["device"]
    solver = HelmholtzSolver(obs["x_m"], obs["y_m"], c, None, float(obs["frequencies_hz"][fi]), -1,
                             cfg["pml_strength"], cfg["pml_m"], options)
    dev = solver.device
    tx = np.ascontiguousarray(np.asarray(obs["tx_index"], dtype=np.int64).ravel())
    rx = np.ascontiguousarray(np.asarray(obs["rx_index"], dtype=np.int64).ravel())
    Yd, Md = _obs_slice(obs, fi, dev)
    _check_obs_slice(Yd, Md, tx.size, rx.size)
    grad = torch.empty((nx, ny), dtype=torch.float64, device=dev) if need_gradient else None
    loss = C.c_double()
    t0 = time.perf_counter()
    with torch.cuda.device(dev):
        N.check(solver._lib.wust_oracle_eval(
            solver._handle(), tx.ctypes.data_as(C.c_void_p), tx.size, rx.ctypes.data_as(C.c_void_p), rx.size,
            C.c_void_p(Yd.data_ptr()), C.c_void_p(Md.data_ptr()), int(cfg.get("source_batch_size", tx.size)),
            int(need_gradient), int(retain), float(solver.tol), int(solver.max_iter), C.byref(loss),
            C.c_void_p(grad.data_ptr() if grad is not None else None), _stream(dev)))
    elapsed = time.perf_counter() - t0
    if nargout == 1:
        return loss.value
    g = grad.t()   # the engine has checked loss and gradient for non-finite values
    if nargout == 2:
        return loss.value, g
    scales = np.zeros(tx.size, np.complex128)
    N.check(solver._lib.wust_oracle_scales(solver._handle(), scales.ctypes.data_as(C.c_void_p)))
    state = OracleState(solver=solver, scales=scales, forward_rhs=tx.size,
                        adjoint_rhs=tx.size if need_gradient else 0, forward_seconds=elapsed,
                        extra=dict(receiver_mask=Md, rx_index=rx))
    return loss.value, g, state


def directional_curvature(state, direction, rx_index=None, receiver_mask=None):
    """kappa = sum_b || -(H^-1 M(v_b . delta))[rx] a_b . mask ||^2 with frozen a.

    ``rx_index`` (may be None) must be the one the state was built with: the engine keeps the
    receivers of the retained evaluation, so a different set raises ``WUST:InvalidInput``;
    ``receiver_mask`` is ``obs.mask[:, :, fi]`` ([n_tx, n_rx]).
    """
    if rx_index is not None:
        kept = np.asarray(state.extra.get("rx_index", []), dtype=np.int64).ravel()
        given = np.asarray(rx_index, dtype=np.int64).ravel()
        if given.shape != kept.shape or not np.array_equal(given, kept):
            raise WustError("WUST:InvalidInput", "rx_index differs from the receivers of the retained oracle state")
    solver = state.solver
    dev = solver.device
    direction = torch.as_tensor(direction)
    _check_shape(direction, (solver.ny, solver.nx), "direction must be [ny, nx]")
    d = _grid_to_dev(direction, dev)
    if receiver_mask is None:
        Md = state.extra["receiver_mask"]
    else:
        Md = torch.as_tensor(receiver_mask, device=dev).to(torch.uint8).t().contiguous()
        _check_shape(Md, tuple(state.extra["receiver_mask"].shape), "receiver_mask must be [n_tx, n_rx]")
    kappa = C.c_double()
    with torch.cuda.device(dev):
        N.check(solver._lib.wust_directional_curvature(
            solver._handle(), C.c_void_p(d.data_ptr()), C.c_void_p(Md.data_ptr()), float(solver.tol),
            int(solver.max_iter), C.byref(kappa), _stream(dev)))
    return kappa.value


class WustOracleFunction(torch.autograd.Function):
    """Misfit of one frequency as a differentiable function of slowness.

    ``forward(ctx, slowness, obs, fi, cfg)`` evaluates :func:`wust_oracle` at
    c = 1 / slowness (loss and gradient in one engine call: forward solves,
    FP64 source fit, adjoint solves reusing the factors, stencil-derivative
    gradient) and returns the FP64 loss as a 0-d tensor; ``backward`` returns
    ``grad_output * g`` in the dtype and on the device of ``slowness``.  The
    slowness may be FP32 or FP64; the engine always works from its FP64 value.
    """

    @staticmethod
    def forward(ctx, slowness, obs, fi, cfg):
        s = slowness.detach()
        if s.dim() != 2 or not bool(torch.all(torch.isfinite(s) & (s > 0))):
            raise WustError("WUST:InvalidInput", "Expected a positive finite [ny,nx] slowness")
        loss, g = wust_oracle(1.0 / s.to(torch.float64), obs, fi, cfg, nargout=2)
        ctx.save_for_backward(g)
        ctx.in_dtype, ctx.in_device = slowness.dtype, slowness.device
        return torch.tensor(loss, dtype=torch.float64, device=slowness.device)

    @staticmethod
    def backward(ctx, grad_output):
        (g,) = ctx.saved_tensors
        grad = grad_output.to(device=g.device, dtype=torch.float64) * g
        return grad.to(device=ctx.in_device, dtype=ctx.in_dtype), None, None, None


def wust_loss(slowness, obs, fi, cfg):
    """``WustOracleFunction.apply``: differentiable misfit Phi_f(s) of frequency ``fi``."""
    return WustOracleFunction.apply(slowness, obs, fi, cfg)


def stencil_opt_params(vmin, vmax, f, h, g):
    """(b, d, e) of the optimal nine-point stencil (Runtime/solver/stencilOptParams.m)."""
    lib = N.load()
    b, d, e = C.c_double(), C.c_double(), C.c_double()
    N.check(lib.wust_stencil_opt_params(float(vmin), float(vmax), float(f), float(h), float(g),
                                        C.byref(b), C.byref(d), C.byref(e)))
    return b.value, d.value, e.value
