"""Restart schema and rejection rules (reference tests/runtime/test_runtime_contracts.py,
src/zhenmode/model/io/restart.py:56-227); host-only, runs on the CPU."""

import numpy as np
import pytest

from zhenmode_amd import restart as R
from zhenmode_amd.cases import make_case
from zhenmode_amd.params import freeze
from zhenmode_amd.solver import JaxStateG


def _setup(tmp_path, dtype="float32"):
    case = make_case("g4deg")
    p = freeze(case.grid, case.physics, case.dt, dtype=dtype, **case.solver_kwargs)
    src = tmp_path / "code.py"
    src.write_text("print('step')\n")
    contract = R.make_restart_contract(case.grid, p, dtype=dtype, forcing=dict(wind="cos3"),
                                       controls=dict(days=1.0), code_paths={"step": src},
                                       execution=dict(device="mi355x"))
    nx, ny, nz = case.grid.nx, case.grid.ny, case.grid.nz
    rng = np.random.default_rng(0)
    dt = np.dtype(dtype)
    state = JaxStateG(*(rng.standard_normal(s).astype(dt) for s in [(nx, ny, nz)] * 4 + [(nx, ny)] * 2))
    return contract, state, src


def test_round_trip_is_byte_identical(tmp_path):
    contract, state, _ = _setup(tmp_path)
    path = tmp_path / "r.npz"
    R.save_restart(path, state, contract, step=12, counters=dict(rejected=0),
                   cumulative=dict(heat=np.zeros(3)), history=dict(max_u=np.arange(4.0)))
    rec = R.load_restart(path, contract)
    assert rec.step == 12 and rec.elapsed_seconds == 12 * contract["dt_s"]
    for name, value in zip(JaxStateG._fields, state):
        assert rec.state[name].tobytes() == np.asarray(value).tobytes()
    saved = np.load(path, allow_pickle=False)
    assert {"state__u", "state__ice", "cumulative__heat", "history__max_u", "metadata_json",
            "metadata_sha256"} <= set(saved.files)


def test_contract_mismatch_is_rejected(tmp_path):
    contract, state, src = _setup(tmp_path)
    path = tmp_path / "r.npz"
    R.save_restart(path, state, contract, step=1, counters={}, cumulative={}, history={})
    other = dict(contract)
    other["controls"] = {"days": 2.0}
    with pytest.raises(ValueError, match="controls"):
        R.load_restart(path, other)
    src.write_text("print('changed')\n")
    changed = dict(contract, sources={"step": R.sha256_file(src)})
    with pytest.raises(ValueError, match="sources"):
        R.load_restart(path, changed)


def test_nonfinite_and_wrong_dtype_states_are_refused(tmp_path):
    contract, state, _ = _setup(tmp_path)
    bad = state._replace(T=np.full_like(state.T, np.nan))
    with pytest.raises(ValueError, match="finite"):
        R.save_restart(tmp_path / "x.npz", bad, contract, step=1, counters={}, cumulative={}, history={})
    wrong = state._replace(u=state.u.astype(np.float64))
    with pytest.raises(ValueError, match="dtype"):
        R.save_restart(tmp_path / "x.npz", wrong, contract, step=1, counters={}, cumulative={}, history={})
    missing = {k: v for k, v in state._asdict().items() if k != "ice"}
    with pytest.raises(ValueError, match="six state fields"):
        R.save_restart(tmp_path / "x.npz", missing, contract, step=1, counters={}, cumulative={}, history={})


def test_tampered_archive_is_refused(tmp_path):
    contract, state, _ = _setup(tmp_path)
    path = tmp_path / "r.npz"
    R.save_restart(path, state, contract, step=3, counters={}, cumulative={}, history={})
    with np.load(path, allow_pickle=False) as saved:
        payload = {k: saved[k] for k in saved.files}
    payload["state__u"] = payload["state__u"] + np.float32(1.0)
    np.savez_compressed(path, **payload)
    with pytest.raises(ValueError, match="mismatch"):
        R.load_restart(path, contract)
    payload.pop("metadata_sha256")
    np.savez_compressed(path, **payload)
    with pytest.raises(ValueError, match="unversioned"):
        R.load_restart(path, contract)


def test_extra_arrays_unsafe_outputs_and_bad_counters_are_refused(tmp_path):
    contract, state, _ = _setup(tmp_path)
    path = tmp_path / "r.npz"
    R.save_restart(path, state, contract, step=2, counters=dict(accepted=2), cumulative={}, history={},
                   outputs={"snap/day1.npz": "0" * 64})
    assert R.load_restart(path, contract).outputs == {"snap/day1.npz": "0" * 64}
    assert sorted(p.name for p in tmp_path.iterdir()) == ["code.py", "r.npz"]   # no temporary left
    with np.load(path, allow_pickle=False) as saved:
        payload = {k: saved[k] for k in saved.files}
    payload["history__extra"] = np.zeros(2)
    np.savez_compressed(path, **payload)
    with pytest.raises(ValueError, match="does not list"):
        R.load_restart(path, contract)
    for outputs in ({"../escape.npz": "0" * 64}, {"/abs.npz": "0" * 64}, {"ok.npz": "xyz"}):
        with pytest.raises(ValueError, match="output"):
            R.save_restart(path, state, contract, step=1, counters={}, cumulative={}, history={}, outputs=outputs)
    for counters in ({"n": -1}, {"n": 1.5}):
        with pytest.raises(ValueError, match="nonnegative integer"):
            R.save_restart(path, state, contract, step=1, counters=counters, cumulative={}, history={})
    with pytest.raises(ValueError, match="'__'"):
        R.save_restart(path, state, contract, step=1, counters={}, cumulative={"a__b": np.zeros(1)}, history={})
