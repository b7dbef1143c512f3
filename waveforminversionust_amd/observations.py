"""Observation and geometry ingestion in the engine's layout.

Host mirror of the reference's ingestion boundary, with its names, field names
and ``WUST:*`` identifiers:

* :func:`wust_geometry` (Runtime/matlab/wust_geometry.m:1-31) -- pixel-edge
  grid axes and nearest-node snapping of the transducer positions;
* :func:`wust_prepare_observations` (Runtime/matlab/wust_prepare_observations.m:1-99)
  -- declared axes, Fourier sign, frequency order, masks and the derived
  per-frequency counts and energies of the canonical ``wust.measurements``;
* :func:`wust_assert_schema` (Runtime/matlab/wust_assert_schema.m:1-5);
* :func:`wust_prepare` (Runtime/matlab/wust_prepare.m:1-83) -- RF [time, rx, tx]
  to canonical complex pressure: the dt-weighted DTFT runs on the GPU as one
  FP64 MFMA GEMM over all transmitters (``wust_rf_dtft``), followed by the
  ingestion above and the optional homogeneous time-of-flight phase
  correction for the snapping offsets;
* :func:`observations_to_device` -- the canonical observations placed once in
  HBM so that every oracle call reads its frequency slice in place: ``Y`` is
  stored as complex128 ``[frequency][rx][tx]`` (the column-major
  ``[tx, rx]`` page the engine's source fit reads) and exposed as the
  reference's ``[tx, rx, frequency]`` view.

Arrays are NumPy on the host (the reference's ingestion is host code);
``Y`` keeps the input precision (complex64 for single input, complex128 for
double).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from ._native import WustError

_UNITS = ("Pa*s", "instrument_units", "unknown")
_NORMALIZATIONS = ("dtft_dt", "discrete_sum", "unknown")
_GRID_FIELDS = {"shape_yx", "origin_yx_m", "spacing_yx_m", "origin_kind"}


def _host(a):
    return a.detach().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)


def wust_assert_schema(value, names):
    names = (names,) if isinstance(names, str) else tuple(names)
    if not (isinstance(value, dict) and value.get("schema") in names and value.get("schema_version") == 2):
        raise WustError("WUST:IncompatibleRequest", "Expected a supported WUST v2 schema")


def _snap(xy, x, y):
    xy = _host(xy)
    if (not np.issubdtype(xy.dtype, np.number) or np.iscomplexobj(xy) or xy.ndim != 2 or xy.shape[1] != 2
            or xy.shape[0] == 0 or not np.all(np.isfinite(xy))):
        raise WustError("WUST:InvalidInput", "Positions must be finite [N,2] physical [x,y]")
    xy = xy.astype(np.float64)
    if not np.all((xy[:, 0] >= x[0]) & (xy[:, 0] <= x[-1]) & (xy[:, 1] >= y[0]) & (xy[:, 1] <= y[-1])):
        raise WustError("WUST:InvalidInput", "Array outside grid sampling coordinates")
    ix = np.argmin(np.abs(x[:, None] - xy[None, :, 0]), axis=0)    # first index on ties
    iy = np.argmin(np.abs(y[:, None] - xy[None, :, 1]), axis=0)
    index = (iy + y.size * ix + 1).astype(np.int64)
    if np.unique(index).size != index.size:
        raise WustError("WUST:InvalidInput", "Distinct elements in one array collapse onto a grid sample")
    return index, np.stack([x[ix], y[iy]], axis=1)


def wust_geometry(grid, tx, rx):
    """Grid axes and 1-based snapped indices; positions are physical [x, y]."""
    if not isinstance(grid, dict) or set(grid) != _GRID_FIELDS:
        raise WustError("WUST:InvalidInput", "Explicit grid fields required")
    n = np.asarray(grid["shape_yx"], np.float64).ravel()
    origin = np.asarray(grid["origin_yx_m"], np.float64).ravel()
    spacing = np.asarray(grid["spacing_yx_m"], np.float64).ravel()
    if not (n.size == 2 and np.all(np.isfinite(n) & (n >= 3) & (n == np.round(n))) and origin.size == 2
            and np.all(np.isfinite(origin)) and spacing.size == 2 and np.all(np.isfinite(spacing) & (spacing > 0))
            and grid["origin_kind"] == "pixel_edge"):
        raise WustError("WUST:InvalidInput", "Invalid grid shape/origin/spacing")
    y = origin[0] + (np.arange(int(n[0])) + 0.5) * spacing[0]
    x = origin[1] + (np.arange(int(n[1])) + 0.5) * spacing[1]
    ti, ts = _snap(tx, x, y)
    ri, rs = _snap(rx, x, y)
    return dict(x_m=x, y_m=y, tx_index=ti, rx_index=ri, tx_xy_m=_host(tx), rx_xy_m=_host(rx),
                tx_snapped_xy_m=ts, rx_snapped_xy_m=rs)


def _layout(a, axes):
    if axes == "tx,rx,frequency":
        return a
    if axes == "frequency,tx,rx":
        return np.transpose(_three(a), (1, 2, 0))
    raise WustError("WUST:InvalidInput", f"Unsupported declared axes: {axes}")


def _three(a):
    """MATLAB's size(a, 1..3) view: trailing singleton axes appended."""
    while a.ndim < 3:
        a = a[..., None]
    return a


def wust_prepare_observations(input):
    """Canonical ``wust.measurements`` from a frequency input or a canonical record."""
    wust_assert_schema(input, ("wust.frequency_input", "wust.measurements"))
    canonical = input["schema"] == "wust.measurements"
    geo = wust_geometry(input["grid"], input["tx_xy_m"], input["rx_xy_m"])
    if canonical:
        if not (np.array_equal(np.asarray(input["tx_index"]).ravel(), geo["tx_index"])
                and np.array_equal(np.asarray(input["rx_index"]).ravel(), geo["rx_index"])):
            raise WustError("WUST:InvalidInput", "Canonical sampling indices disagree with geometry")
    f = np.asarray(input["frequencies_hz"], np.float64).ravel()
    if f.size == 0 or not np.all(np.isfinite(f) & (f > 0)) or np.unique(f).size != f.size:
        raise WustError("WUST:InvalidInput", "Frequencies must be finite, positive and unique")
    if input.get("pressure_type") != "total_pressure":
        raise WustError("WUST:InvalidInput", "Ingestion requires total complex pressure, not scattered fields or ratios")
    units, norm = input.get("data_units"), input.get("spectrum_normalization")
    if units not in _UNITS or norm not in _NORMALIZATIONS:
        raise WustError("WUST:InvalidInput", "Explicit supported units and normalization required")
    if units == "Pa*s" and norm != "dtft_dt":
        raise WustError("WUST:InvalidInput", "Pa*s requires the dt-weighted DTFT normalization")
    nt, nr, nf = geo["tx_index"].size, geo["rx_index"].size, f.size
    sign = np.asarray(input["fourier_sign"])
    if canonical:
        if not (sign.size == 1 and float(sign) == -1 and np.all(np.diff(f) > 0)):
            raise WustError("WUST:InvalidInput",
                            "Canonical observations require Fourier sign -1 and strictly increasing frequencies")
        Y, mask = _host(input["Y"]), _host(input["mask"])
        pressure_axes = mask_axes = "tx,rx,frequency"
    else:
        pressure_axes, mask_axes = input["pressure_axes"], input["mask_axes"]
        Y = _layout(_host(input["pressure"]), pressure_axes)
        mask = _host(input["mask"])
        if mask_axes == "tx,rx":
            if mask.shape != (nt, nr):
                raise WustError("WUST:InvalidInput", "Mask shape mismatch")
            mask = np.repeat(mask[:, :, None], nf, axis=2)
        else:
            mask = _layout(mask, mask_axes)
    if Y.dtype not in (np.float32, np.float64, np.complex64, np.complex128) or _three(Y).shape != (nt, nr, nf):
        raise WustError("WUST:InvalidInput", "Pressure axes/shape mismatch")
    single = Y.dtype in (np.float32, np.complex64)
    Y = _three(Y).astype(np.complex64 if single else np.complex128)     # a fresh copy
    mask = _three(mask)
    if mask.shape != (nt, nr, nf) or not np.all((mask == 0) | (mask == 1)):
        raise WustError("WUST:InvalidInput", "Mask must contain only booleans/0/1")
    mask = mask.astype(bool)
    if not (mask.any() and np.all(np.isfinite(Y[mask]))):
        raise WustError("WUST:InvalidInput", "Empty observations or nonfinite pressure on a valid channel")
    Y[~mask] = 0
    conversion = "none"
    order = np.arange(nf)
    if not canonical:
        if not (sign.size == 1 and float(sign) in (-1.0, 1.0)):
            raise WustError("WUST:InvalidInput", "Declare Fourier sign -1 or +1; unknown conventions are rejected")
        if float(sign) == 1:
            if input.get("real_pressure") is not True:
                raise WustError("WUST:InvalidInput", "Conjugation requires declared real time-domain pressure")
            Y = np.conj(Y)
            conversion = "conjugate"
        order = np.argsort(f, kind="stable")
        Y, mask = Y[:, :, order], mask[:, :, order]
    obs = dict(geo)
    obs.update(schema="wust.measurements", schema_version=2, Y=Y, mask=mask, frequencies_hz=f[order],
               fourier_sign=-1, shape_yx=(geo["y_m"].size, geo["x_m"].size), n_tx=nt, n_rx=nr,
               valid_receiver_counts=mask.sum(axis=1), grid=input["grid"], data_units=units,
               spectrum_normalization=norm, pressure_type="total_pressure")
    with np.errstate(over="ignore"):
        obs["frequency_data_norm"] = np.sum(np.abs(Y.astype(np.complex128)) ** 2, axis=(0, 1))
    if canonical and "preparation" in input:
        obs["preparation"] = input["preparation"]
    else:
        obs["preparation"] = dict(
            input_fourier_sign=float(sign), conversion=conversion, input_frequencies_hz=f,
            canonical_to_input_permutation_1based=order + 1, input_pressure_axes=pressure_axes,
            input_mask_axes=mask_axes, phase_correction="none", masked_values_policy="zero_after_preserving_mask",
            input_precision="single" if single else "double",
            measurement_provenance=input.get("measurement_provenance", "unknown"))
    return obs


def observations_to_device(obs, device=None):
    """Copy of ``obs`` with ``Y`` / ``mask`` resident in HBM in the engine's layout.

    ``Y`` becomes a complex128 tensor whose storage is ``[frequency][rx][tx]``
    (each frequency a column-major ``[tx, rx]`` page) viewed as
    ``[tx, rx, frequency]``; ``mask`` likewise as uint8.  The oracle then
    reads one frequency with no copy and no host transfer.
    """
    if device is None:
        if not torch.cuda.is_available():
            raise WustError("WUST:InvalidInput", "the HIP engine needs a GPU (backend 'gpu')")
        device = torch.device("cuda", torch.cuda.current_device())
    Y = _three(_host(obs["Y"])).astype(np.complex128)
    mask = _three(_host(obs["mask"])).astype(np.uint8)
    Yd = torch.from_numpy(np.ascontiguousarray(Y.transpose(2, 1, 0))).to(device).permute(2, 1, 0)
    Md = torch.from_numpy(np.ascontiguousarray(mask.transpose(2, 1, 0))).to(device).permute(2, 1, 0)
    out = dict(obs)
    out.update(Y=Yd, mask=Md)
    return out


_RF_CFG = {"grid", "frequencies_hz", "c_geom_mps", "phase_correction", "mask", "data_units",
           "measurement_provenance"}


def rf_dtft(pressure, time_s, frequencies_hz, active, device=None, return_kernel_ms=False):
    """Device DTFT: complex64 [tx, rx, frequency] from real RF [time, rx, tx] (float32 or float64).

    ``active`` is the [tx, rx] boolean channel mask (a channel with no valid
    frequency is zeroed before the sum).  Inputs may be host or device arrays;
    the result is a device tensor.
    """
    if device is None:
        device = pressure.device if torch.is_tensor(pressure) and pressure.is_cuda else torch.device("cuda", 0)
    x = torch.as_tensor(pressure, device=device)
    if x.dim() == 2:
        x = x.unsqueeze(-1)
    if x.dtype not in (torch.float32, torch.float64) or x.dim() != 3:
        raise WustError("WUST:InvalidInput", "RF must be real [time,rx,tx] pressure")
    n_time, n_rx, n_tx = x.shape
    t = torch.as_tensor(np.asarray(time_s, np.float64).ravel(), device=device)
    f = np.ascontiguousarray(np.asarray(frequencies_hz, np.float64).ravel())
    if t.numel() != n_time:
        raise WustError("WUST:InvalidInput", "RF must be real [time,rx,tx] pressure")
    xd = x.permute(2, 1, 0).contiguous()                 # storage [tx][rx][time]: MATLAB [time, rx, tx]
    act = torch.as_tensor(np.asarray(_host(active), bool), device=device)
    if tuple(act.shape) != (n_tx, n_rx):
        raise WustError("WUST:InvalidInput", "RF mask must be [tx,rx] or [tx,rx,frequency]")
    act = act.to(torch.uint8).t().contiguous()           # storage [rx][tx]: MATLAB [tx, rx]
    out = torch.empty((f.size, n_rx, n_tx), dtype=torch.complex64, device=device)
    ms = C.c_float()
    dt = float(t[1] - t[0])
    with torch.cuda.device(device):
        N.check(N.load().wust_rf_dtft(
            C.c_void_p(xd.data_ptr()), 1 if x.dtype == torch.float64 else 0, n_time, n_rx, n_tx,
            C.c_void_p(t.data_ptr()), dt, f.ctypes.data_as(C.c_void_p), f.size, C.c_void_p(act.data_ptr()),
            C.c_void_p(out.data_ptr()), C.byref(ms), C.c_void_p(torch.cuda.current_stream(device).cuda_stream)))
    P = out.permute(2, 1, 0)                              # [tx, rx, frequency] view
    return (P, ms.value) if return_kernel_ms else P


def _pair_distance(a, b):
    return np.hypot(a[:, 0][:, None] - b[:, 0][None, :], a[:, 1][:, None] - b[:, 1][None, :])


def wust_prepare(rf, cfg):
    """RF [time, rx, tx] -> canonical observations with an explicit grid.

    ``rf``: dict with ``time_s``, ``pressure`` ([time, rx, tx], real), ``tx_xy_m``,
    ``rx_xy_m`` (optionally ``simulation``); ``cfg``: the reference's fields
    (grid, frequencies_hz, c_geom_mps, phase_correction, mask, data_units,
    measurement_provenance).  No window and no normalization besides dt.
    """
    if not isinstance(cfg, dict) or set(cfg) != _RF_CFG:
        raise WustError("WUST:InvalidInput", "Missing or unsupported RF preparation fields")
    time_s = np.asarray(_host(rf["time_s"]), np.float64).ravel()
    if time_s.size < 2 or not np.all(np.isfinite(time_s)):
        raise WustError("WUST:InvalidInput", "Invalid sample times")
    dt = time_s[1] - time_s[0]
    if not (dt > 0 and np.all(np.abs(np.diff(time_s) - dt) <= dt * 1e-6)):
        raise WustError("WUST:InvalidInput", "Nonuniform time")
    f = np.asarray(cfg["frequencies_hz"], np.float64).ravel()
    if not np.all(np.isfinite(f) & (f > 0) & (f < 0.5 / dt)):
        raise WustError("WUST:InvalidInput", "Frequency outside saved Nyquist")
    nt, nr, nf = np.shape(rf["tx_xy_m"])[0], np.shape(rf["rx_xy_m"])[0], f.size
    pressure = rf["pressure"]
    shape = tuple(pressure.shape)
    shape3 = shape + (1,) * (3 - len(shape))
    if (torch.is_tensor(pressure) and pressure.is_complex()) or np.iscomplexobj(pressure) \
            or shape3 != (time_s.size, nr, nt):
        raise WustError("WUST:InvalidInput", "RF must be real [time,rx,tx] pressure")
    mask = np.asarray(_host(cfg["mask"]))
    if not np.all((mask == 0) | (mask == 1)):
        raise WustError("WUST:InvalidInput", "Mask must contain 0/1")
    mask = mask.astype(bool)
    if mask.shape == (nt, nr):
        mask = np.repeat(mask[:, :, None], nf, axis=2)
    if _three(mask).shape != (nt, nr, nf):
        raise WustError("WUST:InvalidInput", "RF mask must be [tx,rx] or [tx,rx,frequency]")
    mask = _three(mask)
    P = rf_dtft(pressure if torch.is_tensor(pressure) else np.asarray(pressure).reshape(shape3), time_s, f,
                mask.any(axis=2))
    inp = dict(schema="wust.frequency_input", schema_version=2, grid=cfg["grid"], pressure=P.cpu().numpy(),
               mask=mask, pressure_axes="tx,rx,frequency", mask_axes="tx,rx,frequency", frequencies_hz=f,
               tx_xy_m=rf["tx_xy_m"], rx_xy_m=rf["rx_xy_m"], fourier_sign=-1, pressure_type="total_pressure",
               data_units=cfg["data_units"], spectrum_normalization="dtft_dt",
               measurement_provenance=cfg["measurement_provenance"])
    obs = wust_prepare_observations(inp)
    if cfg["phase_correction"] not in ("none", "homogeneous-tof"):
        raise WustError("WUST:InvalidInput", "Unsupported phase policy")
    if cfg["phase_correction"] == "homogeneous-tof":
        c = cfg["c_geom_mps"]
        if not (np.ndim(c) == 0 and np.isfinite(c) and c > 0):
            raise WustError("WUST:InvalidInput", "c_geom_mps must be a positive finite scalar")
        measured = _pair_distance(_host(obs["tx_xy_m"]).astype(np.float64), _host(obs["rx_xy_m"]).astype(np.float64))
        snapped = _pair_distance(obs["tx_snapped_xy_m"], obs["rx_snapped_xy_m"])
        delay = (snapped - measured) / c
        phase = np.exp(-2j * np.pi * delay[:, :, None] * obs["frequencies_hz"][None, None, :])
        obs["Y"] = obs["Y"] * phase.astype(obs["Y"].dtype)     # single data times the rounded phase, as MATLAB
        obs["frequency_data_norm"] = np.sum(np.abs(obs["Y"].astype(np.complex128)) ** 2, axis=(0, 1))
    prep = dict(obs["preparation"])
    prep.update(sampling_interval_s=dt, mask_shape=mask.shape, valid_mask_elements=int(np.count_nonzero(mask)),
                phase_correction=cfg["phase_correction"])
    if cfg["phase_correction"] == "homogeneous-tof":
        prep["geometry_reference_speed_mps"] = cfg["c_geom_mps"]
    obs["preparation"] = prep
    obs["saved_dt_s"] = dt
    if "simulation" in rf:
        obs["simulation"] = rf["simulation"]
    return obs
