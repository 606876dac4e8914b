"""Strict, atomic restart files with the reference's npz schema.

Mirrors ``src/zhenmode/model/io/restart.py:17-227``: the same contract record
(``make_restart_contract``), the same archive keys (``state__{u,v,T,S,eta,ice}``,
``cumulative__*``, ``history__*``, ``metadata_json``, ``metadata_sha256``), the
same finite-only/no-pickle rule and the same rejection of any contract or byte
mismatch with ``ValueError``.  Device tensors are copied to host numpy arrays
in the reference layout (nx, ny, nz) before writing.  A contract embeds source
hashes, so files are exchangeable only between runs of the same code; the
schema and semantics are shared with the reference.
"""

from __future__ import annotations

import hashlib
import json
import os
import tempfile
from contextlib import contextmanager
from dataclasses import dataclass, fields, is_dataclass
from pathlib import Path, PurePosixPath

import numpy as np

SCHEMA_VERSION = 1
STATE_FIELDS = ("u", "v", "T", "S", "eta", "ice")


def _json(value):
    return json.dumps(value, sort_keys=True, separators=(",", ":"), allow_nan=False)


def _host(value):
    if hasattr(value, "detach"):
        return value.detach().cpu().numpy()
    return np.asarray(value)


def _array_descriptor(value):
    array = np.ascontiguousarray(_host(value))
    if array.dtype.kind not in "biuf" or not np.isfinite(array).all():
        raise ValueError("restart arrays must be finite real numeric arrays without pickle")
    return {"shape": list(array.shape), "dtype": array.dtype.str,
            "sha256": hashlib.sha256(array.tobytes(order="C")).hexdigest()}


def fingerprint(value):
    """Fingerprint effective values (reference restart.py:35-53)."""
    if is_dataclass(value):
        return {f.name: fingerprint(getattr(value, f.name)) for f in fields(value)}
    if hasattr(value, "_asdict"):
        return fingerprint(value._asdict())
    if isinstance(value, dict):
        if any(not isinstance(k, str) for k in value):
            raise ValueError("restart mappings require string keys")
        return {k: fingerprint(v) for k, v in value.items()}
    if isinstance(value, (tuple, list)):
        return [fingerprint(v) for v in value]
    if value is None or isinstance(value, (str, bool, int)):
        return value
    if isinstance(value, float):
        if not np.isfinite(value):
            raise ValueError("restart contract contains a nonfinite scalar")
        return value
    if isinstance(value, (np.integer,)):
        return int(value)
    if isinstance(value, (np.floating,)):
        return fingerprint(float(value))
    return _array_descriptor(value)


def sha256_file(path):
    digest = hashlib.sha256()
    with open(path, "rb") as stream:
        for block in iter(lambda: stream.read(1 << 20), b""):
            digest.update(block)
    return digest.hexdigest()


def effective_params(params):
    """The frozen, array- and scalar-valued part of a parameter namespace."""
    out = {}
    for name, value in sorted(vars(params).items()):
        if name == "solver" or callable(value):
            continue
        if isinstance(value, np.ndarray) or isinstance(value, (int, float, str, bool)) or value is None:
            out[name] = value
    return out


def make_restart_contract(grid, params, *, dtype, forcing, controls, code_paths, execution):
    """Reference restart.py:56-73."""
    dims = [int(grid.nx), int(grid.ny), int(grid.nz)]
    compute = np.dtype(dtype)
    if compute.name not in {"float32", "float64"}:
        raise ValueError("restart compute dtype must be float32 or float64")
    dt = float(params.dt)
    if not np.isfinite(dt) or dt <= 0.0:
        raise ValueError("restart dt must be finite and positive")
    eff = effective_params(params) if not hasattr(params, "_asdict") else params
    return {"schema_version": SCHEMA_VERSION,
            "state_family": "FD_point_samples_linear_free_surface",
            "state_shapes": {name: dims[:2] if name in {"eta", "ice"} else dims for name in STATE_FIELDS},
            "state_dtype": compute.str, "dt_s": dt,
            "grid": fingerprint(grid), "effective_params": fingerprint(eff),
            "forcing": fingerprint(forcing), "controls": fingerprint(controls),
            "sources": {name: sha256_file(path) for name, path in code_paths.items()},
            "execution": fingerprint(execution)}


def _integer(value, name):
    array = np.asarray(value)
    if array.shape != () or array.dtype.kind not in "iu" or int(array) < 0:
        raise ValueError(f"restart {name} must be a nonnegative integer scalar")
    return int(array)


def _numeric_mapping(values, name):
    if not isinstance(values, dict) or any(not isinstance(k, str) or not k or "__" in k for k in values):
        raise ValueError(f"restart {name} requires nonempty string array names without '__'")
    arrays = {k: np.array(_host(v), copy=True) for k, v in values.items()}
    for array in arrays.values():
        _array_descriptor(array)
    return arrays


def _validate_state(state, contract):
    if set(state) != set(STATE_FIELDS):
        raise ValueError("restart must contain all six state fields including ice")
    arrays = _numeric_mapping(state, "state")
    for name, array in arrays.items():
        if list(array.shape) != contract["state_shapes"][name]:
            raise ValueError(f"restart {name} shape differs from the frozen grid")
        if array.dtype.str != contract["state_dtype"]:
            raise ValueError(f"restart {name} dtype differs from the frozen compute dtype")
    return arrays


def _validate_outputs(outputs):
    if not isinstance(outputs, dict):
        raise ValueError("restart outputs must be a mapping")
    for name, digest in outputs.items():
        if (not isinstance(name, str) or "\\" in name or ":" in name
                or PurePosixPath(name).is_absolute() or ".." in PurePosixPath(name).parts
                or not name or not isinstance(digest, str) or len(digest) != 64
                or any(ch not in "0123456789abcdef" for ch in digest)):
            raise ValueError("restart output paths/hashes must be safe relative paths and SHA256")


@dataclass(frozen=True)
class RestartRecord:
    state: dict
    step: int
    elapsed_seconds: float
    counters: dict
    cumulative: dict
    history: dict
    outputs: dict


@contextmanager
def atomic_archive(path, *, replace=False):
    """Same-directory temporary file, fsync, then atomic publication."""
    path = Path(path)
    tmp = None
    try:
        with tempfile.NamedTemporaryFile(dir=path.parent, prefix=f".{path.name}.", suffix=".tmp",
                                         delete=False) as stream:
            tmp = Path(stream.name)
            yield stream
            stream.flush()
            os.fsync(stream.fileno())
        if replace:
            os.replace(tmp, path)
        else:
            os.link(tmp, path)
    finally:
        if tmp is not None and tmp.exists():
            tmp.unlink()


def save_restart(path, state, contract, *, step, counters, cumulative, history, outputs=None):
    """Reference restart.py:153-179."""
    if contract.get("schema_version") != SCHEMA_VERSION:
        raise ValueError("unsupported restart contract version")
    step = _integer(step, "step")
    counters = {k: _integer(v, k) for k, v in counters.items()}
    state = state._asdict() if hasattr(state, "_asdict") else state
    arrays = {"state": _validate_state(state, contract),
              "cumulative": _numeric_mapping(cumulative, "cumulative"),
              "history": _numeric_mapping(history, "history")}
    outputs = dict(outputs or {})
    _validate_outputs(outputs)
    elapsed = step * contract["dt_s"]
    if not np.isfinite(elapsed):
        raise ValueError("restart elapsed time is nonfinite")
    metadata = {"schema_version": SCHEMA_VERSION, "contract": contract, "step": step,
                "elapsed_seconds": elapsed, "counters": counters, "outputs": outputs,
                "arrays": {g: {k: _array_descriptor(v) for k, v in vals.items()} for g, vals in arrays.items()}}
    metadata_json = _json(metadata)
    payload = {f"{g}__{k}": v for g, vals in arrays.items() for k, v in vals.items()}
    payload.update(metadata_json=np.asarray(metadata_json),
                   metadata_sha256=np.asarray(hashlib.sha256(metadata_json.encode()).hexdigest()))
    with atomic_archive(path, replace=True) as stream:
        np.savez_compressed(stream, **payload)


def load_restart(path, expected_contract):
    """Reference restart.py:182-227: reject before building any device state."""
    with np.load(path, allow_pickle=False) as saved:
        if "metadata_json" not in saved or "metadata_sha256" not in saved:
            raise ValueError("unversioned checkpoint is not a verified restart; migration required")
        encoded, checksum = saved["metadata_json"], saved["metadata_sha256"]
        if encoded.shape != () or encoded.dtype.kind != "U" or checksum.shape != ():
            raise ValueError("restart metadata must be scalar JSON and checksum")
        metadata_json = str(encoded)
        if hashlib.sha256(metadata_json.encode()).hexdigest() != str(checksum):
            raise ValueError("restart metadata checksum mismatch")
        metadata = json.loads(metadata_json)
        if metadata.get("schema_version") != SCHEMA_VERSION:
            raise ValueError("unsupported restart schema version")
        if _json(metadata["contract"]) != _json(expected_contract):
            bad = [k for k in expected_contract if metadata["contract"].get(k) != expected_contract[k]]
            raise ValueError(f"restart contract mismatch: {', '.join(bad)}")
        step = _integer(metadata["step"], "step")
        elapsed = step * expected_contract["dt_s"]
        if not np.isfinite(elapsed) or metadata["elapsed_seconds"] != elapsed:
            raise ValueError("restart absolute time differs from step * frozen dt")
        counters = {k: _integer(v, k) for k, v in metadata["counters"].items()}
        _validate_outputs(metadata["outputs"])
        arrays, required = {}, {"metadata_json", "metadata_sha256"}
        if set(metadata["arrays"]) != {"state", "cumulative", "history"}:
            raise ValueError("restart array groups mismatch")
        for group, descriptors in metadata["arrays"].items():
            arrays[group] = {}
            for name, descriptor in descriptors.items():
                key = f"{group}__{name}"
                required.add(key)
                if key not in saved:
                    raise ValueError(f"restart missing array {key}")
                value = np.array(saved[key], copy=True)
                if _array_descriptor(value) != descriptor:
                    raise ValueError(f"restart array content/shape/dtype mismatch: {key}")
                arrays[group][name] = value
            _numeric_mapping(arrays[group], group)
        if set(saved.files) != required:
            raise ValueError("restart contains unexpected arrays")
    state = _validate_state(arrays["state"], expected_contract)
    return RestartRecord(state, step, elapsed, counters, arrays["cumulative"], arrays["history"],
                         metadata["outputs"])
