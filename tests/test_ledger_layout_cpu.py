"""The Python ledger layout (zhenmode_amd.budgets.LEDGER_FIELDS) matches the
ZM_LEDGER_* offsets of include/zhenmode.h, and the host-side dict helpers
keep the reference's accumulate semantics (budgets.py:38-67)."""

import re
from pathlib import Path

import numpy as np

from zhenmode_amd import _lib, budgets

HEADER = Path(__file__).resolve().parents[1] / "include" / "zhenmode.h"
NAMES = {
    "OBSERVED": "observed_change", "STAGES": "stage_changes", "SOURCES": "source_inputs",
    "DECOMPOSITION": "decomposition_residual", "BUDGET": "budget_residual",
    "ABS_DECOMPOSITION": "absolute_decomposition_residual", "ABS_BUDGET": "absolute_budget_residual",
    "CHANGE_SCALE": "change_scale", "PROCESSES": "nonlinear_process_changes",
    "PROCESS_SCALE": "nonlinear_process_scale", "ADVECTION_BOUNDARY": "advection_boundary_changes",
    "ADVECTION_BOUNDARY_RESIDUAL": "advection_boundary_residual",
    "ABS_ADVECTION_BOUNDARY_RESIDUAL": "absolute_advection_boundary_residual",
    "NONLINEAR_RESIDUAL": "nonlinear_accounting_residual",
    "ABS_NONLINEAR_RESIDUAL": "absolute_nonlinear_accounting_residual",
    "SURFACE_DISPLACEMENT": "surface_displacement_tracer_change",
    "PROJECTION_NORMS": "projection_transport_norm_squared", "PROJECTION_MAX": "projection_relative_residual_max",
    "TRANSPORT_CONSISTENCY": "transport_consistency_max", "TRANSPORT_AUDITED": "transport_audited_steps",
    "DRAG_ENERGY": "bottom_drag_reference_energy_loss_J", "DRAG_FACE_CHANGE": "bottom_drag_face_change_max_m2_per_s",
    "DRAG_HALVES": "bottom_drag_audited_halves",
}


def test_offsets_match_the_header():
    defs = dict(re.findall(r"#define ZM_LEDGER_(\w+) (\d+)", HEADER.read_text()))
    assert int(defs.pop("SIZE")) == budgets.LEDGER_SIZE == _lib.LEDGER_SIZE
    assert set(defs) == set(NAMES)
    for key, name in NAMES.items():
        assert budgets._OFFSETS[name][0] == int(defs[key]), name


def test_dict_round_trip_and_accumulate():
    flat = np.arange(budgets.LEDGER_SIZE, dtype=np.float64)
    d = budgets.ledger_dict(flat)
    assert d["stage_changes"].shape == (6, 3) and d["source_inputs"].shape == (9, 3)
    assert d["stage_changes"][0, 0] == 3.0 and d["projection_relative_residual_max"] == 98.0
    z = budgets.empty_budget()
    assert all(np.all(v == 0.0) for v in z.values())
    a = budgets.accumulate_budget(d, d)
    np.testing.assert_array_equal(a["observed_change"], 2 * d["observed_change"])
    assert a["projection_relative_residual_max"] == d["projection_relative_residual_max"]
