"""Restart files for the mode-split step.

Interface kept from the reference (``src/zhenmode/model/io/restart.py:17-227``)
so that the same tooling reads both: the contract record's keys, the archive
keys (``state__{u,v,T,S,eta,ice}``, ``cumulative__<name>``, ``history__<name>``,
``metadata_json``, ``metadata_sha256``), the metadata fields (schema version,
contract, step, elapsed seconds, counters, outputs and a shape/dtype/SHA-256
descriptor per array), the canonical JSON text that the checksum covers, and
the rejection rules (any contract, checksum, descriptor or array-set
difference, and any non-finite or non-numeric array, raises ``ValueError``;
nothing is unpickled).  The contract embeds the SHA-256 of the code files, so
an archive is accepted only by a run of the same code.

Design: every value that enters the contract is reduced to a canonical JSON
tree by a small type-dispatch table; an archive is assembled in memory,
written to a fresh exclusive file beside the target, flushed to disk and
renamed over the target (the directory is synced too); loading verifies the
metadata checksum first, then the contract, then each array against its
descriptor, before any array is returned.  Device tensors are copied to host
arrays in the reference layout (nx, ny, nz).
"""

from __future__ import annotations

import hashlib
import io
import json
import os
import secrets
from dataclasses import dataclass, fields, is_dataclass
from pathlib import Path, PurePosixPath

import numpy as np

SCHEMA_VERSION = 1
STATE_FIELDS = ("u", "v", "T", "S", "eta", "ice")
GROUPS = ("state", "cumulative", "history")
STATE_FAMILY = "FD_point_samples_linear_free_surface"


def canonical_json(tree):
    """The exact text the checksum covers: sorted keys, no spaces, no NaN."""
    return json.dumps(tree, sort_keys=True, separators=(",", ":"), allow_nan=False)


def host_array(value):
    """A C-contiguous host copy of a numpy array or a (device) torch tensor."""
    if hasattr(value, "detach"):
        value = value.detach().cpu().numpy()
    return np.ascontiguousarray(value)


def describe(value):
    """Shape, dtype string and SHA-256 of the C-order bytes of a finite real array."""
    array = host_array(value)
    if array.dtype.kind not in "biuf":
        raise ValueError(f"restart arrays must be real numeric, not dtype {array.dtype}")
    if array.dtype.kind == "f" and not np.all(np.isfinite(array)):
        raise ValueError("restart arrays must be finite")
    return {"shape": list(array.shape), "dtype": array.dtype.str,
            "sha256": hashlib.sha256(array.tobytes()).hexdigest()}


def _scalar_float(value):
    value = float(value)
    if not np.isfinite(value):
        raise ValueError("restart contract values must be finite")
    return value


def _mapping(value):
    bad = [key for key in value if not isinstance(key, str)]
    if bad:
        raise ValueError(f"restart contract mappings need string keys, got {bad[0]!r}")
    return {key: fingerprint(item) for key, item in value.items()}


# (test, conversion) pairs, tried in order; anything else is described as an array
_CANONICAL = (
    (lambda v: v is None or isinstance(v, (str, bool)), lambda v: v),
    (lambda v: isinstance(v, (int, np.integer)) and not isinstance(v, (bool, np.bool_)), int),
    (lambda v: isinstance(v, (float, np.floating)), _scalar_float),
    (lambda v: isinstance(v, dict), _mapping),
    (lambda v: is_dataclass(v) and not isinstance(v, type),
     lambda v: {f.name: fingerprint(getattr(v, f.name)) for f in fields(v)}),
    (lambda v: hasattr(v, "_asdict"), lambda v: _mapping(v._asdict())),
    (lambda v: isinstance(v, (list, tuple)), lambda v: [fingerprint(item) for item in v]),
)


def fingerprint(value):
    """Canonical JSON tree of a contract value: scalars as themselves,
    records and mappings by field, sequences by item, arrays by descriptor."""
    for test, convert in _CANONICAL:
        if test(value):
            return convert(value)
    return describe(value)


def sha256_file(path):
    digest = hashlib.sha256()
    with open(path, "rb") as stream:
        while block := stream.read(1 << 20):
            digest.update(block)
    return digest.hexdigest()


def effective_params(params):
    """The value-carrying part of a frozen parameter namespace: arrays, scalars
    and strings, by name (handles and callables left out)."""
    keep = (np.ndarray, int, float, str, bool, type(None))
    return {name: value for name, value in sorted(vars(params).items())
            if name != "solver" and isinstance(value, keep)}


def make_restart_contract(grid, params, *, dtype, forcing, controls, code_paths, execution):
    """The record a restart must match exactly (reference restart.py:56-73):
    state shapes and dtype, dt, the grid, the effective parameters, forcing,
    run controls, the code files' SHA-256 and the execution settings."""
    compute = np.dtype(dtype)
    if compute not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError(f"restart compute dtype must be float32 or float64, not {compute}")
    dt = float(params.dt)
    if not (np.isfinite(dt) and dt > 0.0):
        raise ValueError("restart dt must be finite and positive")
    volume = [int(grid.nx), int(grid.ny), int(grid.nz)]
    shapes = {name: volume[:2] if name in ("eta", "ice") else volume for name in STATE_FIELDS}
    values = params if hasattr(params, "_asdict") else effective_params(params)
    return {"schema_version": SCHEMA_VERSION, "state_family": STATE_FAMILY, "state_shapes": shapes,
            "state_dtype": compute.str, "dt_s": dt, "grid": fingerprint(grid),
            "effective_params": fingerprint(values), "forcing": fingerprint(forcing),
            "controls": fingerprint(controls),
            "sources": {name: sha256_file(path) for name, path in code_paths.items()},
            "execution": fingerprint(execution)}


@dataclass(frozen=True)
class RestartRecord:
    state: dict
    step: int
    elapsed_seconds: float
    counters: dict
    cumulative: dict
    history: dict
    outputs: dict


def _count(value, what):
    array = np.asarray(value)
    if array.ndim or array.dtype.kind not in "iu" or array < 0:
        raise ValueError(f"restart {what} must be one nonnegative integer")
    return int(array)


def _group(values, what):
    """Host copies of a named array group; names are non-empty and contain no
    '__' (the archive key separator)."""
    if not isinstance(values, dict):
        raise ValueError(f"restart {what} must be a mapping of named arrays")
    out = {}
    for name, value in values.items():
        if not isinstance(name, str) or not name or "__" in name:
            raise ValueError(f"restart {what} array names must be non-empty strings without '__'")
        out[name] = np.array(host_array(value), copy=True)
        describe(out[name])
    return out


def _state_group(state, contract):
    if hasattr(state, "_asdict"):
        state = state._asdict()
    if not isinstance(state, dict) or set(state) != set(STATE_FIELDS):
        raise ValueError("restart state must hold the six state fields u, v, T, S, eta and ice")
    arrays = _group(state, "state")
    for name, array in arrays.items():
        if list(array.shape) != contract["state_shapes"][name]:
            raise ValueError(f"restart state {name} has shape {array.shape}, the contract's grid differs")
        if array.dtype.str != contract["state_dtype"]:
            raise ValueError(f"restart state {name} has dtype {array.dtype}, the contract's compute dtype differs")
    return arrays


def _safe_outputs(outputs):
    for name, digest in outputs.items():
        parts = PurePosixPath(name).parts if isinstance(name, str) else ()
        relative = (bool(name) and isinstance(name, str) and "\\" not in name and ":" not in name
                    and not PurePosixPath(name).is_absolute() and ".." not in parts)
        hexdigest = isinstance(digest, str) and len(digest) == 64 and set(digest) <= set("0123456789abcdef")
        if not (relative and hexdigest):
            raise ValueError(f"restart output {name!r} needs a relative path and a SHA-256 hex digest")
    return outputs


def _publish(path, data):
    """Write ``data`` to a new exclusive file beside ``path``, sync it, rename it
    over ``path`` and sync the directory; a crash leaves at most the hidden
    temporary file, never a partial archive under the final name."""
    path = Path(path)
    directory = path.parent if str(path.parent) else Path(".")
    temporary = directory / f".{path.name}.{secrets.token_hex(6)}.part"
    fd = os.open(temporary, os.O_WRONLY | os.O_CREAT | os.O_EXCL, 0o644)
    try:
        with os.fdopen(fd, "wb") as stream:
            stream.write(data)
            stream.flush()
            os.fsync(stream.fileno())
        os.replace(temporary, path)
    finally:
        if temporary.exists():
            temporary.unlink()
    dir_fd = os.open(directory, os.O_RDONLY)
    try:
        os.fsync(dir_fd)
    finally:
        os.close(dir_fd)


def save_restart(path, state, contract, *, step, counters, cumulative, history, outputs=None):
    """Check everything, then publish the archive atomically."""
    if contract.get("schema_version") != SCHEMA_VERSION:
        raise ValueError(f"restart contract schema {contract.get('schema_version')!r} is not {SCHEMA_VERSION}")
    step = _count(step, "step")
    groups = {"state": _state_group(state, contract), "cumulative": _group(cumulative, "cumulative"),
              "history": _group(history, "history")}
    elapsed = step * contract["dt_s"]
    if not np.isfinite(elapsed):
        raise ValueError("restart elapsed time is not finite")
    metadata = {"schema_version": SCHEMA_VERSION, "contract": contract, "step": step,
                "elapsed_seconds": elapsed,
                "counters": {name: _count(value, name) for name, value in counters.items()},
                "outputs": _safe_outputs(dict(outputs or {})),
                "arrays": {g: {name: describe(a) for name, a in arrays.items()} for g, arrays in groups.items()}}
    text = canonical_json(metadata)
    members = {f"{g}__{name}": a for g, arrays in groups.items() for name, a in arrays.items()}
    members["metadata_json"] = np.asarray(text)
    members["metadata_sha256"] = np.asarray(hashlib.sha256(text.encode()).hexdigest())
    buffer = io.BytesIO()
    np.savez_compressed(buffer, **members)
    _publish(path, buffer.getvalue())


def _verified_metadata(archive):
    if "metadata_json" not in archive.files or "metadata_sha256" not in archive.files:
        raise ValueError("unversioned checkpoint (no checksummed metadata): not a verified restart")
    text, checksum = archive["metadata_json"], archive["metadata_sha256"]
    if text.ndim or checksum.ndim or text.dtype.kind != "U":
        raise ValueError("restart metadata must be one JSON string and one checksum")
    text = str(text)
    if hashlib.sha256(text.encode()).hexdigest() != str(checksum):
        raise ValueError("restart metadata checksum mismatch")
    metadata = json.loads(text)
    if metadata.get("schema_version") != SCHEMA_VERSION:
        raise ValueError(f"restart schema {metadata.get('schema_version')!r} is not {SCHEMA_VERSION}")
    return metadata


def load_restart(path, expected_contract):
    """Read and verify an archive; nothing is returned unless every check holds."""
    with np.load(path, allow_pickle=False) as archive:
        metadata = _verified_metadata(archive)
        saved = metadata["contract"]
        if canonical_json(saved) != canonical_json(expected_contract):
            keys = sorted(set(saved) | set(expected_contract))
            differing = [k for k in keys if canonical_json(saved.get(k)) != canonical_json(expected_contract.get(k))]
            raise ValueError("restart contract mismatch in: " + ", ".join(differing))
        step = _count(metadata["step"], "step")
        elapsed = step * expected_contract["dt_s"]
        if metadata["elapsed_seconds"] != elapsed:
            raise ValueError("restart elapsed time is not step x the contract's dt")
        counters = {name: _count(value, name) for name, value in metadata["counters"].items()}
        outputs = _safe_outputs(metadata["outputs"])
        if sorted(metadata["arrays"]) != sorted(GROUPS):
            raise ValueError("restart array groups mismatch")
        expected_keys = {"metadata_json", "metadata_sha256"}
        groups = {}
        for g in GROUPS:
            groups[g] = {}
            for name, descriptor in metadata["arrays"][g].items():
                key = f"{g}__{name}"
                expected_keys.add(key)
                if key not in archive.files:
                    raise ValueError(f"restart array {key} is missing")
                array = np.array(archive[key], copy=True)
                if describe(array) != descriptor:
                    raise ValueError(f"restart array {key} mismatch (content, shape or dtype)")
                groups[g][name] = array
        if set(archive.files) != expected_keys:
            raise ValueError("restart archive holds arrays its metadata does not list")
    state = _state_group(groups["state"], expected_contract)
    return RestartRecord(state, step, elapsed, counters, groups["cumulative"], groups["history"], outputs)
