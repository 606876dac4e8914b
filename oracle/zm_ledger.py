"""NumPy restatement of the shadow-step stage ledger.

TEST INFRASTRUCTURE ONLY: checker for ``zm_budget_step``; only ``tests/`` may
import it.

Follows ``_StageRecorder`` and ``make_budget_step``
(``src/zhenmode/model/diagnostics/budgets.py:62-270``) for the legacy
mode-split step: the recorder hooks sit where ``_step_impl`` calls them
(``timestepping/step.py:138-161, 299-300, 365-368, 426-427, 452-456``) and
where ``_compute_tracer_tendency`` books its terms
(``dynamics/tendencies.py:211-214``).  In this configuration the sponge,
dynamic-ice, bottom-drag and transport-consistency entries stay zero (no
sponge, no dynamic ice, no split drag, no nodal column geometry).

Pinned in ``tests/test_oracle_ledger.py`` by the identities the ledger exists
to check: the stage changes telescope to the observed change; the nonlinear
stage equals its sources plus its process tendencies; the column integral of
the flux-form advection equals the surface advective flux; the free-surface
stage moves no heat or salt and its displacement is the surface-area integral
of the eta change; a planted wrong source interval breaks the nonlinear
accounting.
"""

from __future__ import annotations

import numpy as np

from . import zm_oracle as O

STAGE_NAMES = ("linear_diffusion", "sponge", "nonlinear", "free_surface", "polar_cap_and_masks", "dynamic_ice")
SOURCE_NAMES = ("prescribed_heat", "bulk_heat", "coastal_bulk_heat", "temperature_restore",
                "salinity_restore", "prescribed_brine", "sponge", "ice_atmosphere_heat", "ice_brine")
PROCESS_NAMES = ("advection", "convection", "gm", "redi")
MAXIMUM_FIELDS = ("projection_relative_residual_max", "transport_consistency_max",
                  "bottom_drag_face_change_max_m2_per_s")
RHO_ICE, LATENT = 917.0, 3.34e5


class Recorder:
    def __init__(self, p, source_interval=None):
        self.p = p
        self.cp = O.C_P
        self.area = p.dx_2d * p.dy
        self.surface_area = self.area * p.wet_mask
        self.volume = self.area[:, :, None] * p.dz_node * p.wet_mask_z
        self.interval = p.dt / 2.0 if source_interval is None else source_interval
        self.stages = {n: np.zeros(3) for n in STAGE_NAMES}
        self.sources = {n: np.zeros(3) for n in SOURCE_NAMES}
        self.processes = {n: np.zeros(3) for n in PROCESS_NAMES}
        self.change_scale = np.zeros(3)
        self.process_scale = np.zeros(3)
        self.advection_boundary = np.zeros(3)
        self.projection_norm_squared = np.zeros(2)
        self.projection_relative_residual_max = 0.0

    def difference(self, before, after):
        heat = O.RHO_0 * self.cp * (after.T - before.T) * self.volume
        latent = RHO_ICE * LATENT * (np.asarray(after.ice) - np.asarray(before.ice)) * self.surface_area
        salt = O.RHO_0 / 1000.0 * (after.S - before.S) * self.volume
        disp = (after.eta - before.eta) * self.surface_area
        change = np.array([heat.sum() - np.sum(latent), salt.sum(), disp.sum()])
        scale = np.array([np.abs(heat).sum() + np.abs(latent).sum(), np.abs(salt).sum(), np.abs(disp).sum()])
        return change, scale

    def stage(self, name, before, after):
        change, scale = self.difference(before, after)
        self.stages[name] = self.stages[name] + change
        self.change_scale = self.change_scale + scale

    def sponge_sources(self, before, decay):
        fraction = 1.0 - decay
        t = (self.p.T_clim_3d - before.T) * fraction
        s = (self.p.S_clim_3d - before.S) * fraction
        self.sources["sponge"] = self.sources["sponge"] + np.array(
            [O.RHO_0 * self.cp * np.sum(t * self.volume), O.RHO_0 / 1000.0 * np.sum(s * self.volume), 0.0])

    def surface_sources(self, tendencies):
        for index, tendency in enumerate(tendencies):
            factor = O.RHO_0 * self.cp if index < 4 else O.RHO_0 / 1000.0
            component = 0 if index < 4 else 1
            amount = np.sum(tendency * self.volume) * factor * self.interval
            self.sources[SOURCE_NAMES[index]][component] += amount

    def nonlinear_terms(self, tendencies):
        for name, (t, s) in zip(PROCESS_NAMES, tendencies):
            heat = O.RHO_0 * self.cp * self.interval * t * self.volume
            salt = O.RHO_0 / 1000.0 * self.interval * s * self.volume
            self.processes[name] = self.processes[name] + np.array([heat.sum(), salt.sum(), 0.0])
            self.process_scale = self.process_scale + np.array([np.abs(heat).sum(), np.abs(salt).sum(), 0.0])

    def advection_boundary_fluxes(self, top_T, top_S):
        heat = O.RHO_0 * self.cp * self.interval * np.sum(top_T * self.surface_area)
        salt = O.RHO_0 / 1000.0 * self.interval * np.sum(top_S * self.surface_area)
        self.advection_boundary = self.advection_boundary + np.array([heat, salt, 0.0])

    def column_projection(self, ux, vy, cx, cy):
        before = O.vertical_transport_iface(ux, vy, self.p)[..., 0]
        after = O.vertical_transport_iface(cx, cy, self.p)[..., 0]
        norms = np.array([np.sum(before ** 2 * self.surface_area), np.sum(after ** 2 * self.surface_area)])
        self.projection_norm_squared = self.projection_norm_squared + norms
        if norms[0] > 0.0:
            rel = float(np.sqrt(norms[1] / norms[0]))
        else:
            rel = 0.0 if norms[1] == 0.0 else np.inf
        self.projection_relative_residual_max = max(self.projection_relative_residual_max, rel)

    def surface_displacement_change(self, before, after):
        t = after.eta * after.T[..., 0] - before.eta * before.T[..., 0]
        s = after.eta * after.S[..., 0] - before.eta * before.S[..., 0]
        return np.array([O.RHO_0 * self.cp * np.sum(t * self.surface_area),
                          O.RHO_0 / 1000.0 * np.sum(s * self.surface_area), 0.0])

    def result(self, before, after):
        observed, _ = self.difference(before, after)
        stages = np.stack([self.stages[n] for n in STAGE_NAMES])
        sources = np.stack([self.sources[n] for n in SOURCE_NAMES])
        processes = np.stack([self.processes[n] for n in PROCESS_NAMES])
        decomposition = observed - stages.sum(axis=0)
        budget = observed - sources.sum(axis=0)
        boundary = self.processes["advection"] - self.advection_boundary
        nonlinear = self.stages["nonlinear"] - sources[:6].sum(axis=0) - processes.sum(axis=0)
        return {"observed_change": observed, "stage_changes": stages, "source_inputs": sources,
                "decomposition_residual": decomposition, "budget_residual": budget,
                "absolute_decomposition_residual": np.abs(decomposition),
                "absolute_budget_residual": np.abs(budget), "change_scale": self.change_scale,
                "nonlinear_process_changes": processes, "nonlinear_process_scale": self.process_scale,
                "advection_boundary_changes": self.advection_boundary,
                "advection_boundary_residual": boundary,
                "absolute_advection_boundary_residual": np.abs(boundary),
                "nonlinear_accounting_residual": nonlinear,
                "absolute_nonlinear_accounting_residual": np.abs(nonlinear),
                "surface_displacement_tracer_change": self.surface_displacement_change(before, after),
                "projection_transport_norm_squared": self.projection_norm_squared,
                "projection_relative_residual_max": np.float64(self.projection_relative_residual_max),
                "transport_consistency_max": np.zeros(4), "transport_audited_steps": np.float64(0.0),
                "bottom_drag_reference_energy_loss_J": np.float64(0.0),
                "bottom_drag_face_change_max_m2_per_s": np.float64(0.0),
                "bottom_drag_audited_halves": np.float64(0.0)}


def budget_step(state, p, source_interval=None):
    """One legacy mode-split step with the stage ledger (budgets.py:244-270 on
    step.py:341-458): returns (new state, audit dict)."""
    state = O.State(*(np.asarray(a, dtype=np.float64) for a in state))
    rec = Recorder(p, source_interval)
    dt_half = p.dt / 2.0
    s1 = O.linear_half_step(state, p, dt_half, rec=rec)
    s2 = O.explicit_full_step(s1, p, p.dt, rec=rec)
    rec.stage("nonlinear", s1, s2)
    s3 = O.linear_half_step(s2, p, dt_half, rec=rec)
    s4 = O.barotropic_subcycle(s3, p)
    rec.stage("free_surface", s3, s4)
    out = O.final_masks(s4, state, p)
    rec.stage("polar_cap_and_masks", s4, out)
    rec.stage("dynamic_ice", out, out)
    return out, rec.result(state, out)


def empty_budget():
    """budgets.py:38-62."""
    return {"observed_change": np.zeros(3), "stage_changes": np.zeros((6, 3)), "source_inputs": np.zeros((9, 3)),
            "decomposition_residual": np.zeros(3), "budget_residual": np.zeros(3),
            "absolute_decomposition_residual": np.zeros(3), "absolute_budget_residual": np.zeros(3),
            "change_scale": np.zeros(3), "nonlinear_process_changes": np.zeros((4, 3)),
            "nonlinear_process_scale": np.zeros(3), "advection_boundary_changes": np.zeros(3),
            "advection_boundary_residual": np.zeros(3), "absolute_advection_boundary_residual": np.zeros(3),
            "nonlinear_accounting_residual": np.zeros(3), "absolute_nonlinear_accounting_residual": np.zeros(3),
            "surface_displacement_tracer_change": np.zeros(3), "projection_transport_norm_squared": np.zeros(2),
            "projection_relative_residual_max": np.float64(0.0), "transport_consistency_max": np.zeros(4),
            "transport_audited_steps": np.float64(0.0), "bottom_drag_reference_energy_loss_J": np.float64(0.0),
            "bottom_drag_face_change_max_m2_per_s": np.float64(0.0), "bottom_drag_audited_halves": np.float64(0.0)}


def accumulate_budget(totals, interval):
    """budgets.py:64-67: sums, except the worst-ratio fields, which keep the maximum."""
    return {k: (np.maximum(totals[k], interval[k]) if k in MAXIMUM_FIELDS else totals[k] + interval[k])
            for k in totals}
