"""Reconstruction artifacts: the ``wust.reconstruction`` manifest of a ``wust_reconstruct`` result.

Mirrors Runtime/matlab/wust_write_artifact.m:1-89 and Runtime/matlab/wust_read_artifact.m:1-66 for
the schema this engine produces (array fields of Runtime/python/contracts.py:48-51):

* the manifest ``{schema, schema_version: 2, metadata, arrays_file, arrays}`` is JSON; ``metadata`` is
  the result without its arrays, ``schema`` and ``schema_version``, with ``config.update_mask``
  removed and the schedule, frequency lists and records as JSON arrays; a NaN becomes ``null``
  (MATLAB's jsonencode does the same);
* ``arrays`` holds one spec per array field, ``{axes, shape, dtype, units, dataset}`` with the
  dataset named ``/<field>``: ``c_mps`` (``y,x``, m/s, float64) and, when the result has it,
  ``history_mps`` (``y,x,model_step``, m/s, float32);
* index fields are written as the reference's MATLAB values, 1-based: ``config.schedule``,
  ``executed_schedule`` and each record's ``stage_id`` (the Python result holds them 0-based).

The reference writes the datasets to an HDF5 sibling ``arrays.h5`` through MATLAB's HDF5 library.
No HDF5 library is available to this package, so the datasets go to a NumPy archive sibling,
``arrays.npz``, under the same dataset names (without the leading ``/``), each an array of the
declared shape in the declared axis order (row-major), which is the layout the reference's permuted
HDF5 write produces; ``arrays_file`` names that archive.  As in the reference, the arrays are
written first and the manifest last (the manifest is the success index of the run directory), and
an existing manifest or arrays file is ``WUST:InvalidInput``.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from ._native import WustError

ARRAY_FILE = "arrays.npz"
SCHEMA_VERSION = 2

# array fields of the reconstruction schema (Runtime/python/contracts.py:48-51)
FIELDS = {
    "wust.reconstruction": {
        "c_mps": dict(axes=["y,x"], units="m/s", dtypes=["float64"], optional=False),
        "history_mps": dict(axes=["y,x,model_step"], units="m/s", dtypes=["float32"], optional=True),
    },
}
# prepared coordinates and statistics the reference rebuilds from the canonical arrays
_DERIVED = ("x_m", "y_m", "tx_snapped_xy_m", "rx_snapped_xy_m", "pressure_axes", "mask_axes", "shape_yx", "n_tx",
            "n_rx", "valid_receiver_counts", "frequency_data_norm")


def _host(a):
    try:
        import torch
        if torch.is_tensor(a):
            return a.detach().cpu().numpy()
    except ImportError:
        pass
    return np.asarray(a)


def _json_value(v):
    """JSON-ready copy: arrays to (nested) lists, NaN / inf to None, numpy scalars to Python."""
    if isinstance(v, dict):
        return {str(k): _json_value(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_json_value(x) for x in v]
    try:
        import torch
        if torch.is_tensor(v):
            v = v.detach().cpu().numpy()
    except ImportError:
        pass
    if isinstance(v, np.ndarray):
        return _json_value(v.tolist())
    if isinstance(v, (np.bool_, bool)):
        return bool(v)
    if isinstance(v, (np.integer,)):
        return int(v)
    if isinstance(v, (float, np.floating)):
        f = float(v)
        return f if math.isfinite(f) else None
    if v is None or isinstance(v, (str, int)):
        return v
    return str(v)   # e.g. the torch device an engine config names


def wust_write_artifact(path, result):
    """Write ``result`` (a ``wust_reconstruct`` result) as the manifest ``path`` plus its arrays sibling."""
    schema = result.get("schema")
    if schema not in FIELDS:
        raise WustError("WUST:InvalidInput", f"unsupported artifact schema {schema!r}")
    directory = os.path.dirname(os.path.abspath(path))
    arrays_path = os.path.join(directory, ARRAY_FILE)
    if os.path.exists(arrays_path) or os.path.exists(path):
        raise WustError("WUST:InvalidInput", "Output exists")
    meta = {k: v for k, v in result.items() if k not in _DERIVED}
    specs, data = {}, {}
    for key, field in FIELDS[schema].items():
        if key not in result or result[key] is None:
            if not field["optional"]:
                raise WustError("WUST:InvalidInput", f"Missing output array: {key}")
            continue
        a = _host(result[key])
        meta.pop(key, None)
        axes = result.get(f"{key}_axes", field["axes"][0])
        rank = len(axes.split(","))
        if a.ndim > rank or (a.ndim < rank and rank > 1):
            a = a.reshape(a.shape + (1,) * (rank - a.ndim))
        shape = [int(s) for s in a.shape[:rank]] if rank > 1 else [int(a.size)]
        dtype = {np.dtype(np.float64): "float64", np.dtype(np.float32): "float32", np.dtype(np.bool_): "bool",
                 np.dtype(np.int64): "int64", np.dtype(np.uint8): "uint8"}.get(a.dtype)
        if dtype is None or dtype not in field["dtypes"]:
            raise WustError("WUST:InvalidInput", f"Invalid dtype for {key}: {a.dtype}")
        specs[key] = dict(axes=axes, shape=shape, dtype=dtype, units=field["units"], dataset=f"/{key}")
        data[key] = np.ascontiguousarray(a.astype(np.uint8) if dtype == "bool" else a)
    if schema == "wust.reconstruction":
        cfg = dict(meta.get("config", {}))
        cfg.pop("update_mask", None)
        if "schedule" in cfg:
            cfg["schedule"] = [int(i) + 1 for i in np.asarray(cfg["schedule"]).ravel()]
        meta["config"] = cfg
        if "executed_schedule" in meta:
            meta["executed_schedule"] = [int(i) + 1 for i in np.asarray(meta["executed_schedule"]).ravel()]
        for key in ("executed_frequencies_hz", "canonical_frequencies_hz"):
            if key in meta:
                meta[key] = [float(f) for f in np.asarray(meta[key], np.float64).ravel()]
        meta["records"] = [dict(r, stage_id=int(r["stage_id"]) + 1) if "stage_id" in r else dict(r)
                           for r in meta.get("records", [])]
    meta.pop("schema", None)
    meta.pop("schema_version", None)
    doc = dict(schema=schema, schema_version=SCHEMA_VERSION, metadata=_json_value(meta), arrays_file=ARRAY_FILE,
               arrays=specs)
    # arrays close first; the manifest is the success index of the run directory
    with open(arrays_path, "xb") as fh:
        np.savez(fh, **{k: v for k, v in data.items()})
    with open(path, "x") as fh:
        json.dump(doc, fh, allow_nan=False)
    return path


def wust_read_artifact(path, expected=None):
    """Read a manifest written by :func:`wust_write_artifact`: the metadata dict with ``schema``,
    ``schema_version`` and the arrays restored, after the checks of Runtime/matlab/wust_read_artifact.m."""
    with open(path) as fh:
        doc = json.load(fh)
    schema = doc.get("schema")
    if (expected is not None and schema != expected) or schema not in FIELDS or doc.get("schema_version") != SCHEMA_VERSION:
        raise WustError("WUST:InvalidInput", "Unexpected artifact schema")
    fields = FIELDS[schema]
    names = set(doc.get("arrays", {}))
    required = {k for k, f in fields.items() if not f["optional"]}
    if not (required <= names <= set(fields)):
        raise WustError("WUST:InvalidInput", "Missing or unexpected array fields")
    arrays_file = doc["arrays_file"]
    if os.path.basename(arrays_file) != arrays_file:
        raise WustError("WUST:InvalidInput", "Array file must be a sibling filename")
    value = dict(doc["metadata"])
    value["schema"], value["schema_version"] = schema, doc["schema_version"]
    with np.load(os.path.join(os.path.dirname(os.path.abspath(path)), arrays_file), allow_pickle=False) as z:
        for key in names:
            spec, field = doc["arrays"][key], fields[key]
            if spec["axes"] not in field["axes"] or spec["units"] != field["units"] or spec["dtype"] not in field["dtypes"]:
                raise WustError("WUST:InvalidInput", f"Invalid axes/units/dtype: {key}")
            shape = [int(s) for s in spec["shape"]]
            if len(shape) != len(spec["axes"].split(",")) or any(s < 1 for s in shape):
                raise WustError("WUST:InvalidInput", "Invalid array shape")
            if spec["dataset"] != f"/{key}":
                raise WustError("WUST:InvalidInput", "Invalid dataset name")
            a = z[key]
            if list(a.shape) != shape or str(a.dtype) != {"bool": "uint8"}.get(spec["dtype"], spec["dtype"]):
                raise WustError("WUST:InvalidInput", "Stored array disagrees with its declared shape/dtype")
            value[key] = a.astype(bool) if spec["dtype"] == "bool" else a
    return value
