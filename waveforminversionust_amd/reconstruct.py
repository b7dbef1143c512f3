"""Host-side mirror of the reference's NCG reconstruction step, on the HIP engine.

Same names, argument meaning and error identifiers as the reference:

* :func:`ringingRemovalFilt` (Runtime/solver/ringingRemovalFilt.m:1-45)
* :func:`wust_roi_gradient` (Runtime/matlab/wust_roi_gradient.m:1-10)
* :func:`wust_ncg_direction` (Runtime/matlab/wust_ncg_direction.m:1-17)
* :func:`linearized_step_length` (Runtime/matlab/linearized_step_length.m:1-11)
* :func:`wust_project_slowness_update` (Runtime/matlab/wust_project_slowness_update.m:1-10)
* :func:`wust_reconstruct` (Runtime/matlab/wust_reconstruct.m:1-207): the
  frequency-continuation loop, with :func:`prepare_reconstruction` its
  configuration checks.

Grids are ``[y, x]`` (MATLAB order).  Inputs may be NumPy arrays or torch
tensors; results are float64 torch tensors on the GPU, stored column-major
(y fastest) as the C ABI takes them, so one NCG step never leaves the device
except for its scalars (gz, beta, slope, alpha, curvature).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _native as N
from ._native import WustError
from .helmholtz import _dev, _grid_to_dev, _stream, _uniform, directional_curvature, wust_oracle

_REQUIRED = ("backend", "bounds_mps", "max_update_mps", "step_damping", "source_batch_size", "pml_strength",
             "pml_m", "stencil_bounds", "wavenumber", "filter_cutoff", "filter_order", "update_mask")
_OPTIONAL = ("dispersion", "frequency_schedule_hz", "max_updates", "save_history", "evaluate_final_residual")
# engine controls of this implementation, passed through to wust_oracle
_ENGINE = ("mode", "tol", "max_iter", "factor", "device")


def _cm(a, device, dtype=torch.float64):
    """[Ny, Nx] array -> (Nx, Ny) contiguous device tensor (column-major [Ny, Nx])."""
    t = torch.as_tensor(a)
    if t.dim() != 2:
        raise WustError("WUST:InvalidInput", "Expected an [ny,nx] grid")
    if t.dtype == torch.bool:
        t = t.to(torch.uint8)
    return _grid_to_dev(t, device, dtype)


def _ptr(t):
    return C.c_void_p(t.data_ptr() if t is not None else None)


def _device_of(*arrays):
    for a in arrays:
        if torch.is_tensor(a) and a.is_cuda:
            return a.device
    return _dev()


# ---- single-step pieces ------------------------------------------------------------------

def _roi_dev(g_cm, mask_cm, ny, nx, dx, dy, c0, frequency_hz, cutoff, order, device):
    roi = torch.empty_like(g_cm)
    z = torch.empty_like(g_cm)
    with torch.cuda.device(device):
        N.check(N.load().wust_roi_gradient(_ptr(g_cm), _ptr(mask_cm), ny, nx, float(dx), float(dy), float(c0),
                                           float(frequency_hz), float(cutoff), float(order), _ptr(roi), _ptr(z),
                                           _stream(device)))
    return roi, z


def _filter_args(x_m, y_m, cutoff, order):
    if not (np.isscalar(cutoff) or np.ndim(cutoff) == 0) or not np.isfinite(cutoff) or cutoff < 0:
        raise WustError("WUST:InvalidInput", "cutoff must be a finite nonnegative scalar")
    if cutoff == 0:
        return 1.0, 1.0
    _, dx = _uniform(x_m, "x")
    _, dy = _uniform(y_m, "y")
    if not np.isfinite(order) or order <= 0:
        raise WustError("WUST:InvalidInput", "filter order must be positive")
    return dx, dy


def ringingRemovalFilt(xin, yin, img_in, c0, f, cutoff, order):
    """Radial Butterworth low-pass of cutoff * 2 f / c0 cycles/m (cutoff 0: unchanged copy)."""
    device = _device_of(img_in)
    img = _cm(img_in, device)
    ny, nx = img.shape[1], img.shape[0]
    if np.asarray(xin).size != nx or np.asarray(yin).size != ny:
        raise WustError("WUST:InvalidInput", "grid axes must match the image")
    dx, dy = _filter_args(xin, yin, cutoff, order)
    ones = torch.ones_like(img, dtype=torch.uint8)
    _, z = _roi_dev(img, ones, ny, nx, dx, dy, c0, f, cutoff, order, device)
    return z.t()


ringing_removal_filt = ringingRemovalFilt


def wust_roi_gradient(g, mask, x_m, y_m, c0, frequency_hz, cutoff, order):
    """(roi_gradient, z) = (M g, M F M g) with F the ringing-removal filter."""
    device = _device_of(g, mask)
    g_cm = _cm(g, device)
    m_cm = _cm(mask, device, torch.uint8)
    if m_cm.shape != g_cm.shape:
        raise WustError("WUST:InvalidInput", "mask must match the gradient grid")
    ny, nx = g_cm.shape[1], g_cm.shape[0]
    dx, dy = _filter_args(x_m, y_m, cutoff, order)
    roi, z = _roi_dev(g_cm, m_cm, ny, nx, dx, dy, c0, frequency_hz, cutoff, order, device)
    return roi.t(), z.t()


def _direction_dev(g_cm, z_cm, prev_z_cm, prev_gz, prev_dir_cm, same_frequency, device):
    d = torch.empty_like(z_cm)
    beta, gz, slope = C.c_double(), C.c_double(), C.c_double()
    with torch.cuda.device(device):
        N.check(N.load().wust_ncg_direction(_ptr(g_cm), _ptr(z_cm), _ptr(prev_z_cm), float(prev_gz),
                                            _ptr(prev_dir_cm), int(bool(same_frequency)), z_cm.numel(), _ptr(d),
                                            C.byref(beta), C.byref(gz), C.byref(slope), _stream(device)))
    return d, beta.value, gz.value, slope.value


def wust_ncg_direction(g, z, prev_z, prev_gz, previous_direction, same_frequency, return_slope=False):
    """(direction, beta, gz): clipped PR/FR on the same frequency, restarted as -z on ascent.

    ``return_slope`` adds <g, direction>, the objective's slope along it.
    """
    device = _device_of(g, z, prev_z, previous_direction)
    vector = torch.as_tensor(g).dim() == 1

    def cm(a):
        t = torch.as_tensor(a)
        return _cm(t.unsqueeze(1) if vector and t.dim() == 1 else t, device)

    g_cm, z_cm = cm(g), cm(z)
    use_prev = bool(same_frequency) and prev_gz > 0
    pz = cm(prev_z) if use_prev else None
    pd = cm(previous_direction) if use_prev else None
    for t in (z_cm, pz, pd):
        if t is not None and t.shape != g_cm.shape:
            raise WustError("WUST:InvalidInput", "NCG arrays must share one grid")
    d, beta, gz, slope = _direction_dev(g_cm, z_cm, pz, prev_gz, pd, same_frequency, device)
    d = d[0] if vector else d.t()
    if return_slope:
        return d, beta, gz, slope
    return d, beta, gz


def linearized_step_length(slope, curvature):
    """alpha = max(0, -slope / curvature); 0 for a zero curvature with a zero slope."""
    alpha = C.c_double()
    N.check(N.load().wust_linearized_step_length(float(slope), float(curvature), C.byref(alpha)))
    return alpha.value


def _project_dev(c_cm, d_cm, alpha, cfg, m_cm, device):
    lo, hi = (float(v) for v in np.asarray(cfg["bounds_mps"], dtype=np.float64).ravel())
    out = torch.empty_like(c_cm)
    rel = C.c_double()
    with torch.cuda.device(device):
        N.check(N.load().wust_project_slowness_update(
            _ptr(c_cm), _ptr(d_cm), float(alpha), float(cfg["step_damping"]), float(cfg["max_update_mps"]), lo, hi,
            _ptr(m_cm), c_cm.numel(), _ptr(out), C.byref(rel), _stream(device)))
    return out, rel.value


def wust_project_slowness_update(c, direction, alpha, cfg, return_relative_update=False):
    """Bounded slowness step s = 1/c + step_damping * alpha * direction, clipped per node."""
    device = _device_of(c, direction)
    c_cm, d_cm = _cm(c, device), _cm(direction, device)
    m_cm = _cm(cfg["update_mask"], device, torch.uint8)
    if d_cm.shape != c_cm.shape or m_cm.shape != c_cm.shape:
        raise WustError("WUST:InvalidInput", "NCG arrays must share one grid")
    out, rel = _project_dev(c_cm, d_cm, alpha, cfg, m_cm, device)
    if return_relative_update:
        return out.t(), rel
    return out.t()


# ---- the reconstruction loop ---------------------------------------------------------------

def _positive_scalar(v):
    return np.ndim(v) == 0 and np.isreal(v) and np.isfinite(v) and v > 0


def prepare_reconstruction(obs, initial, cfg):
    """Validated copy of ``cfg`` with the 0-based ``schedule`` (Runtime/matlab/wust_reconstruct.m:128-207)."""
    cfg = dict(cfg)
    names = set(cfg)
    if not set(_REQUIRED) <= names or not names <= set(_REQUIRED + _OPTIONAL + _ENGINE):
        raise WustError("WUST:InvalidInput", "Missing or unsupported reconstruction config fields")
    for name in ("pml_m", "pml_strength", "max_update_mps", "step_damping", "source_batch_size", "filter_order"):
        if not _positive_scalar(cfg[name]):
            raise WustError("WUST:InvalidInput", f"Expected positive finite {name}")
    fc = cfg["filter_cutoff"]
    if (cfg["source_batch_size"] != round(cfg["source_batch_size"]) or cfg["filter_order"] != round(cfg["filter_order"])
            or np.ndim(fc) != 0 or not np.isfinite(fc) or fc < 0):
        raise WustError("WUST:InvalidInput", "Invalid batch/filter controls")
    for name in ("bounds_mps", "stencil_bounds"):
        b = np.asarray(cfg[name], dtype=np.float64).ravel()
        if b.size != 2 or not np.all(np.isfinite(b)) or b[0] <= 0 or b[1] <= b[0]:
            raise WustError("WUST:InvalidInput", "Invalid speed bounds")
    if str(cfg["backend"]) not in ("cpu", "gpu") or str(cfg["wavenumber"]) not in ("continuum", "kwave-ldr9"):
        raise WustError("WUST:InvalidInput", "Unsupported backend/wavenumber")
    if cfg["wavenumber"] == "kwave-ldr9" and "dispersion" not in cfg:
        raise WustError("WUST:InvalidInput", "LDR9 requires explicit simulation dispersion facts")
    x_m = np.asarray(obs["x_m"], np.float64).ravel()
    y_m = np.asarray(obs["y_m"], np.float64).ravel()
    init = initial.detach().cpu().numpy() if torch.is_tensor(initial) else np.asarray(initial)
    lo, hi = np.asarray(cfg["bounds_mps"], np.float64).ravel()
    if (init.shape != (y_m.size, x_m.size) or np.iscomplexobj(init) or not np.all(np.isfinite(init))
            or not np.all((init >= lo) & (init <= hi))):
        raise WustError("WUST:InvalidInput",
                        "Initial map must match grid and speed bounds; no resize or scalar expansion")
    um = cfg["update_mask"]
    um = um.detach().cpu().numpy() if torch.is_tensor(um) else np.asarray(um)
    if um.shape != init.shape or not np.all((um == 0) | (um == 1)):
        raise WustError("WUST:InvalidInput", "Invalid update mask")
    um = um.astype(bool)
    cfg["update_mask"] = um
    X, Y = np.meshgrid(x_m, y_m)
    tol = 16 * np.spacing(np.max(np.abs(np.concatenate([x_m, y_m]))) + 1)
    L = float(cfg["pml_m"])
    interior = ((X > x_m.min() + L - tol) & (X < x_m.max() - L + tol)
                & (Y > y_m.min() + L - tol) & (Y < y_m.max() - L + tol))
    interior[[0, -1], :] = False
    interior[:, [0, -1]] = False
    if np.any(um & ~interior):
        raise WustError("WUST:InvalidInput", "Update mask enters PML/Dirichlet boundary")
    freqs = np.asarray(obs["frequencies_hz"], np.float64).ravel()
    full = freqs if "frequency_schedule_hz" not in cfg else np.asarray(cfg["frequency_schedule_hz"], np.float64)
    if full.ndim > 1 and min(full.shape) > 1 or not np.all(np.isfinite(full) & (full > 0)):
        raise WustError("WUST:InvalidInput", "frequency_schedule_hz must contain finite positive Hz")
    full = full.ravel()
    schedule = []
    for f in full:
        hit = np.flatnonzero(freqs == f)
        if hit.size == 0:
            raise WustError("WUST:InvalidInput", "Schedule frequencies must exactly match observations")
        schedule.append(int(hit[0]))
    cfg["planned_schedule_length"] = len(schedule)
    if cfg.get("max_updates") is not None:
        n = cfg["max_updates"]
        if not (np.ndim(n) == 0 and np.isfinite(n) and n >= 0 and n == round(n)):
            raise WustError("WUST:InvalidInput", "max_updates must be a nonnegative integer")
        schedule = schedule[:int(n)]
    cfg["schedule"] = schedule
    cfg.setdefault("evaluate_final_residual", False)
    if not isinstance(cfg["evaluate_final_residual"], (bool, np.bool_)):
        raise WustError("WUST:InvalidInput", "evaluate_final_residual must be boolean")
    evaluated = sorted(set(schedule))
    if cfg["evaluate_final_residual"]:
        evaluated = list(range(freqs.size))
    mask = obs["mask"]
    mask = mask.detach().cpu().numpy() if torch.is_tensor(mask) else np.asarray(mask)
    mask = mask.astype(bool)
    counts = mask.sum(axis=1)                                   # [n_tx, n_freq]
    if evaluated and not np.all(counts[:, evaluated] >= 2):
        raise WustError("WUST:InvalidInput",
                        "Source-scale fitting requires two valid receivers per evaluated TX/frequency")
    ny, nx = init.shape
    sampling = np.ones((ny, nx), bool)
    sampling[[0, -1], :] = False
    sampling[:, [0, -1]] = False
    samp = sampling.ravel(order="F")
    tx = np.asarray(obs["tx_index"], np.int64).ravel() - 1
    rx = np.asarray(obs["rx_index"], np.int64).ravel() - 1
    active_rx = mask[:, :, evaluated].any(axis=(0, 2)) if evaluated else np.zeros(rx.size, bool)
    if not (np.all(samp[tx]) and np.all(samp[rx[active_rx]])):
        raise WustError("WUST:InvalidInput", "TX and fitted RX must sample interior (non-Dirichlet) grid points")
    cfg.setdefault("save_history", False)
    if not isinstance(cfg["save_history"], (bool, np.bool_)):
        raise WustError("WUST:InvalidInput", "save_history must be boolean")
    return cfg


def _frequency_data_norm(obs):
    if "frequency_data_norm" in obs:
        return np.asarray(obs["frequency_data_norm"], np.float64).ravel()
    Y = obs["Y"]
    Y = Y.detach().cpu().numpy() if torch.is_tensor(Y) else np.asarray(Y)
    with np.errstate(over="ignore"):       # huge data gives inf, as the reference's sum does
        return np.sum(np.abs(Y.astype(np.complex128)) ** 2, axis=(0, 1))


def wust_reconstruct(obs, initial, cfg, progress=None):
    """Frequency-continuation FWI in slowness with clipped PR/FR NCG.

    Returns a dict with the reference's result fields (``c_mps`` as a torch
    [ny, nx] tensor on the GPU, ``records`` a list of dicts, completion and
    schedule bookkeeping, ``final_data_residual`` None when not evaluated).
    ``progress`` (optional callable) receives the reference's per-step line
    "FWI k/n f=... MHz residual(pre)=... wall=...s".
    """
    t_all = time.perf_counter()
    cfg = prepare_reconstruction(obs, initial, cfg)
    device = _dev(cfg.get("device"))
    x_m = np.asarray(obs["x_m"], np.float64).ravel()
    y_m = np.asarray(obs["y_m"], np.float64).ravel()
    ny, nx = y_m.size, x_m.size
    freqs = [float(f) for f in np.asarray(obs["frequencies_hz"], np.float64).ravel()]
    data_norm = _frequency_data_norm(obs)
    c = _cm(initial, device)                           # (nx, ny) column-major
    c0 = float(torch.mean(c).item())
    mask_cm = _cm(cfg["update_mask"], device, torch.uint8)
    dx, dy = _filter_args(x_m, y_m, cfg["filter_cutoff"], cfg["filter_order"])
    obs_mask = obs["mask"]
    schedule = cfg["schedule"]
    history = [c.t().to(torch.float32).clone()] if cfg["save_history"] else None
    prev_z, prev_gz, direction, last = None, 0.0, None, -1
    records = []
    for step, fi in enumerate(schedule):
        t_step = time.perf_counter()
        loss, g, state = wust_oracle(c.t(), obs, fi, cfg, nargout=3)
        g_cm = g.t().contiguous()
        roi, z = _roi_dev(g_cm, mask_cm, ny, nx, dx, dy, c0, freqs[fi], cfg["filter_cutoff"],
                          cfg["filter_order"], device)
        same = fi == last
        use_prev = same and prev_gz > 0
        direction, beta, gz, slope = _direction_dev(roi, z, prev_z if use_prev else None, prev_gz,
                                                    direction if use_prev else None, same, device)
        prev_z, prev_gz, last = z, gz, fi
        t_lin = time.perf_counter()
        rmask = obs_mask[:, :, fi]
        curvature = directional_curvature(state, direction.t(), obs["rx_index"], rmask)
        lin_seconds = time.perf_counter() - t_lin
        alpha = linearized_step_length(slope, curvature)
        c, rel = _project_dev(c, direction, alpha, cfg, mask_cm, device)
        if history is not None:
            history.append(c.t().to(torch.float32).clone())
        absolute = float(np.sqrt(2 * loss))
        relative, status = float("nan"), "zero_observation_energy"
        if data_norm[fi] > 0:
            relative, status = absolute / float(np.sqrt(data_norm[fi])), "available"
        solver = state.solver
        records.append(dict(
            step=step + 1, frequency_hz=freqs[fi], loss_before_update=loss,
            absolute_residual_before_update=absolute, relative_residual_before_update=relative,
            relative_residual_status=status, alpha=alpha, beta=beta, gz=gz, slope=slope, curvature=curvature,
            wall_seconds=time.perf_counter() - t_step, assembly_seconds=solver.assembly_seconds,
            factor_seconds=solver.factor_seconds, forward_seconds=state.forward_seconds,
            adjoint_seconds=state.adjoint_seconds, linearized_seconds=lin_seconds, fresh_factorizations=1,
            forward_rhs=state.forward_rhs, adjoint_rhs=state.adjoint_rhs,
            linearized_rhs=int(np.asarray(obs["tx_index"]).size), relative_slowness_update=rel,
            stage_id=fi, loss_model_step=step))
        if progress is not None:
            r = records[-1]
            progress(f"FWI {step + 1}/{len(schedule)} f={freqs[fi] / 1e6:.3f} MHz "
                     f"residual(pre)={r['relative_residual_before_update']:.6g} wall={r['wall_seconds']:.3f}s")
        del state, solver
    n = len(schedule)
    reason, category = "schedule_complete", "completion"
    if n == 0:
        reason, category = "zero_updates", "budget"
    elif n < cfg["planned_schedule_length"]:
        reason, category = "caller_truncated_schedule", "budget"
    result = dict(schema="wust.reconstruction", schema_version=2, c_mps=c.t(), records=records, config=cfg)
    if history is not None:
        result["history_mps"] = torch.stack(history, dim=2)
    result.update(executed_schedule=list(schedule), completed_updates=n,
                  planned_schedule_length=cfg["planned_schedule_length"],
                  executed_frequencies_hz=[freqs[i] for i in schedule], canonical_frequencies_hz=freqs,
                  completion=dict(category=category, reason=reason, converged=False), final_data_residual=None)
    if cfg["evaluate_final_residual"]:
        t_fin = time.perf_counter()
        final_loss = sum(wust_oracle(c.t(), obs, fi, cfg, nargout=1) for fi in range(len(freqs)))
        energy = float(np.sum(data_norm))
        relative, status = float("nan"), "zero_observation_energy"
        if energy > 0:
            relative, status = float(np.sqrt(2 * final_loss / energy)), "available"
        result["final_data_residual"] = dict(
            absolute=float(np.sqrt(2 * final_loss)), relative=relative, status=status, oracle_calls=len(freqs),
            forward_rhs=int(np.asarray(obs["tx_index"]).size) * len(freqs), adjoint_rhs=0,
            wall_seconds=time.perf_counter() - t_fin)
    result["wall_seconds"] = time.perf_counter() - t_all
    return result
