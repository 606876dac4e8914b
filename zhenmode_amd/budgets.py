"""Shadow-step stage ledger: ``make_budget_step`` on the GPU.

Mirrors ``src/zhenmode/model/diagnostics/budgets.py``: ``make_budget_step``
(``:244-270``) returns a step that advances the state exactly like the solver's
step and also returns the float64 audit of ``_StageRecorder.result``
(``:187-225``), ``empty_budget`` (``:38-62``) and ``accumulate_budget``
(``:64-67``) keep their meaning.  The audit is formed on the device
(``zm_budget_step``); ``raw=True`` returns it as a float64 device tensor of
``LEDGER_SIZE`` entries without synchronising, ``ledger_dict`` converts it, and
``accumulate_ledger`` adds device ledgers over an interval.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .solver import as_field, ptr

METRIC_NAMES = ("fixed_node_water_ice_enthalpy_J", "water_salt_kg", "eta_volume_m3")
STAGE_NAMES = ("linear_diffusion", "sponge", "nonlinear", "free_surface", "polar_cap_and_masks", "dynamic_ice")
SOURCE_NAMES = ("prescribed_heat", "bulk_heat", "coastal_bulk_heat", "temperature_restore",
                "salinity_restore", "prescribed_brine", "sponge", "ice_atmosphere_heat", "ice_brine")
NONLINEAR_PROCESS_NAMES = ("advection", "convection", "gm", "redi")
MAXIMUM_BUDGET_FIELDS = ("projection_relative_residual_max", "transport_consistency_max",
                         "bottom_drag_face_change_max_m2_per_s")

# (name, shape) in the order of the device ledger (ZM_LEDGER_* in include/zhenmode.h)
LEDGER_FIELDS = (
    ("observed_change", (3,)), ("stage_changes", (6, 3)), ("source_inputs", (9, 3)),
    ("decomposition_residual", (3,)), ("budget_residual", (3,)),
    ("absolute_decomposition_residual", (3,)), ("absolute_budget_residual", (3,)),
    ("change_scale", (3,)), ("nonlinear_process_changes", (4, 3)), ("nonlinear_process_scale", (3,)),
    ("advection_boundary_changes", (3,)), ("advection_boundary_residual", (3,)),
    ("absolute_advection_boundary_residual", (3,)), ("nonlinear_accounting_residual", (3,)),
    ("absolute_nonlinear_accounting_residual", (3,)), ("surface_displacement_tracer_change", (3,)),
    ("projection_transport_norm_squared", (2,)), ("projection_relative_residual_max", ()),
    ("transport_consistency_max", (4,)), ("transport_audited_steps", ()),
    ("bottom_drag_reference_energy_loss_J", ()), ("bottom_drag_face_change_max_m2_per_s", ()),
    ("bottom_drag_audited_halves", ()),
)
LEDGER_SIZE = sum(int(np.prod(shape)) for _, shape in LEDGER_FIELDS)

_OFFSETS = {}
_o = 0
for _name, _shape in LEDGER_FIELDS:
    _OFFSETS[_name] = (_o, _shape)
    _o += int(np.prod(_shape))


def _max_mask(device):
    mask = torch.zeros(LEDGER_SIZE, dtype=torch.bool, device=device)
    for name in MAXIMUM_BUDGET_FIELDS:
        off, shape = _OFFSETS[name]
        mask[off:off + int(np.prod(shape))] = True
    return mask


def ledger_dict(ledger):
    """Device or host ledger vector -> dict of float64 numpy arrays (synchronises)."""
    flat = ledger.detach().double().cpu().numpy() if isinstance(ledger, torch.Tensor) else np.asarray(ledger)
    return {name: (flat[off:off + int(np.prod(shape))].reshape(shape) if shape else np.float64(flat[off]))
            for name, (off, shape) in _OFFSETS.items()}


def empty_budget():
    """Zero for interval accumulation with accumulate_budget; not a closure claim."""
    return ledger_dict(np.zeros(LEDGER_SIZE))


def accumulate_budget(totals, interval):
    """Sum inventories and squared norms, but keep the worst residual ratio."""
    return {name: (np.maximum(totals[name], interval[name]) if name in MAXIMUM_BUDGET_FIELDS
                   else totals[name] + interval[name]) for name in totals}


def accumulate_ledger(total, interval):
    """accumulate_budget on device ledger vectors (no host synchronisation)."""
    mask = _max_mask(total.device)
    return torch.where(mask, torch.maximum(total, interval), total + interval)


def make_budget_step(solver):
    """Return ``advance(state, forcing=None, atmosphere=None, surface_pressure_pa=None, raw=False)``
    yielding ``(new state, audit)``; forcing is ``(tau_x, tau_y, Q_heat)`` and
    atmosphere the ``(nx, ny, 1)`` air temperature, both applied to the step
    and to its source audit."""
    if _lib.LEDGER_SIZE != LEDGER_SIZE:
        raise RuntimeError("ledger layout differs from the library's")

    def advance(state, forcing=None, atmosphere=None, surface_pressure_pa=None, raw=False):
        if surface_pressure_pa is not None:
            raise ValueError("surface_pressure_pa is not supported by the HIP step")
        s2, dt, dev = solver.shape2, solver.dtype, solver.device
        ptrs = [None, None, None, None]
        keep = []
        if forcing is not None:
            if len(forcing) != 3:
                raise ValueError("forcing must contain tau_x, tau_y and Q_heat")
            for q, a in enumerate(forcing):
                t = as_field(a, s2, dt, dev)
                keep.append(t)
                ptrs[q] = ptr(t)
        if atmosphere is not None:
            arr = atmosphere.reshape(s2) if hasattr(atmosphere, "reshape") else atmosphere
            t = as_field(arr, s2, dt, dev)
            keep.append(t)
            ptrs[3] = ptr(t)
        src = solver.coerce(state)
        dst = solver.new_state()
        ledger = torch.empty(LEDGER_SIZE, dtype=torch.float64, device=dev)
        rc = solver.lib.zm_budget_step(solver.handle, ctypes.byref(solver.zm_state(src)),
                                       ctypes.byref(solver.zm_state(dst)), *ptrs, ptr(ledger), solver._stream())
        _lib.check(rc, solver.handle)
        return dst, (ledger if raw else ledger_dict(ledger))

    return advance
