"""FP64 CPU restatement of the NCG driver step around the oracle (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU leg import this module, as the
checker of the GPU engine.  Each function cites the reference file it follows.

Pins (tests/test_ncg_oracle.py):
* ``wust_ncg_direction``, ``wust_project_slowness_update``, ``wust_roi_gradient``:
  the reference's own unit-test values (Runtime/tests/test_runtime.m:170-190:
  beta = 4/3, gz = 4, restart on ascent, beta = 0 across frequencies, the
  crossing projection to 1512 m/s, masked filtering invariance);
* ``ringing_removal_filt``: a brute-force DFT of the same formula on odd and
  even grids (pins the fftshift / ifftshift conventions), the zero-frequency
  gain 1 (constant images pass unchanged), linearity, and the limit of a huge
  cutoff (identity);
* ``linearized_step_length``: closed form and the reference's error cases;
* ``wust_reconstruct``: loss decrease on the 31x31 fixture and the
  identities alpha = max(0, -slope / kappa), c within bounds.
"""
from __future__ import annotations

import numpy as np

from .helmholtz_fp64 import WustError


def ringing_removal_filt(xin, yin, img_in, c0, f, cutoff, order):
    """Radial Butterworth low-pass in k-space (Runtime/solver/ringingRemovalFilt.m:1-45)."""
    if not (np.isscalar(cutoff) and np.isfinite(cutoff) and cutoff >= 0):
        raise WustError("WUST:InvalidInput", "cutoff must be a finite nonnegative scalar")
    img_in = np.asarray(img_in, dtype=np.float64)
    if cutoff == 0:
        return img_in.copy()
    xin = np.asarray(xin, np.float64).ravel()
    yin = np.asarray(yin, np.float64).ravel()
    nx, ny = xin.size, yin.size
    dx, dy = np.mean(np.diff(xin)), np.mean(np.diff(yin))
    kx = np.fft.fftshift((np.arange(nx) / nx) / dx)                  # :32-33
    kx[kx >= 1 / (2 * dx)] -= 1 / dx
    ky = np.fft.fftshift((np.arange(ny) / ny) / dy)                  # :34-35
    ky[ky >= 1 / (2 * dy)] -= 1 / dy
    KX, KY = np.meshgrid(kx, ky)
    kcut = (2 * f / c0) * cutoff                                     # :38-40
    filt = 1.0 / (1.0 + ((KX ** 2 + KY ** 2) / kcut ** 2) ** order)  # :43
    spec = np.fft.fftshift(np.fft.fft2(img_in))                      # :46
    return np.real(np.fft.ifft2(np.fft.ifftshift(filt * spec)))


def wust_roi_gradient(g, mask, x_m, y_m, c0, frequency_hz, cutoff, order):
    """M F M g (Runtime/matlab/wust_roi_gradient.m:1-10)."""
    mask = np.asarray(mask, bool)
    roi = np.array(g, dtype=np.float64, copy=True)
    roi[~mask] = 0
    z = roi.copy()
    if cutoff > 0:
        z = ringing_removal_filt(x_m, y_m, z, c0, frequency_hz, cutoff, order)
    z[~mask] = 0
    return roi, z


def wust_ncg_direction(g, z, prev_z, prev_gz, previous_direction, same_frequency):
    """Clipped PR/FR direction (Runtime/matlab/wust_ncg_direction.m:1-17)."""
    g = np.asarray(g, np.float64)
    z = np.asarray(z, np.float64)
    gz = float(np.sum(g * z))
    beta = 0.0
    if same_frequency and prev_gz > 0:
        beta_fr = gz / prev_gz
        beta_pr = float(np.sum(g * (z - np.asarray(prev_z, np.float64)))) / prev_gz
        beta = min(max(beta_pr, 0.0), beta_fr)
    direction = -z + beta * np.asarray(previous_direction, np.float64)
    if np.sum(g * direction) >= 0:
        direction = -z
        beta = 0.0
    if not (np.all(np.isfinite(direction)) and np.isfinite(gz) and gz >= 0):
        raise WustError("WUST:NumericalFailure", "Invalid preconditioned gradient")
    return direction, beta, gz


def linearized_step_length(slope, curvature):
    """alpha = max(0, -slope / kappa) (Runtime/matlab/linearized_step_length.m:1-11)."""
    if not (np.isfinite(slope) and np.isfinite(curvature) and curvature >= 0):
        raise WustError("WUST:NumericalFailure", "Nonfinite slope/curvature")
    if not (curvature > 0 or slope == 0):
        raise WustError("WUST:NumericalFailure", "Zero curvature with nonzero directional slope")
    return max(0.0, -slope / curvature) if curvature > 0 else 0.0


def wust_project_slowness_update(c, direction, alpha, cfg):
    """Bounded slowness step (Runtime/matlab/wust_project_slowness_update.m:1-10)."""
    c = np.asarray(c, np.float64)
    lo, hi = cfg["bounds_mps"]
    c_lo = np.maximum(c - cfg["max_update_mps"], lo)
    c_hi = np.minimum(c + cfg["max_update_mps"], hi)
    trial = 1.0 / c + cfg["step_damping"] * alpha * np.asarray(direction, np.float64)
    if not np.all(np.isfinite(trial)):
        raise WustError("WUST:NumericalFailure", "Nonfinite slowness trial")
    s_new = np.maximum(1.0 / c_hi, np.minimum(1.0 / c_lo, trial))
    c_new = 1.0 / s_new
    m = np.asarray(cfg["update_mask"], bool)
    c_new[~m] = c[~m]
    return c_new


def wust_reconstruct(obs, initial, cfg, oracle_fn, curvature_fn):
    """NCG steps of Runtime/matlab/wust_reconstruct.m:20-70 with injected oracle and curvature.

    ``oracle_fn(c, obs, fi, cfg, nargout=3)`` -> (loss, g, state) and
    ``curvature_fn(state, direction, rx_index, mask)`` -> kappa, so the same loop
    drives the FP64 oracle (here) and checks the GPU driver.  Returns (c, records).
    """
    c = np.array(initial, dtype=np.float64, copy=True)
    mask = np.asarray(cfg["update_mask"], bool)
    prev_z = np.zeros_like(c)
    prev_gz = 0.0
    direction = np.zeros_like(c)
    last = -1
    records = []
    for step, fi in enumerate(cfg["schedule"]):
        loss, g, state = oracle_fn(c, obs, fi, cfg, nargout=3)
        roi, z = wust_roi_gradient(g, mask, obs["x_m"], obs["y_m"], float(np.mean(initial)),
                                   obs["frequencies_hz"][fi], cfg["filter_cutoff"], cfg["filter_order"])
        direction, beta, gz = wust_ncg_direction(roi, z, prev_z, prev_gz, direction, fi == last)
        prev_z, prev_gz, last = z, gz, fi
        kappa = curvature_fn(state, direction, obs["rx_index"], np.asarray(obs["mask"])[:, :, fi])
        slope = float(np.sum(roi * direction))
        alpha = linearized_step_length(slope, kappa)
        old_s = 1.0 / c
        c = wust_project_slowness_update(c, direction, alpha, cfg)
        rel = np.linalg.norm(1.0 / c[mask] - old_s[mask]) / max(np.linalg.norm(old_s[mask]), np.finfo(float).tiny)
        records.append(dict(step=step + 1, frequency_hz=obs["frequencies_hz"][fi], loss_before_update=loss,
                            alpha=alpha, beta=beta, gz=gz, slope=slope, curvature=kappa,
                            relative_slowness_update=rel))
    return c, records
