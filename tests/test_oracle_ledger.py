"""Pins the oracle's stage ledger (reference diagnostics/budgets.py:62-270) by
the identities it is built to check, on the land case with every process on."""

import numpy as np
import pytest

from oracle import zm_ledger as L
from oracle import zm_oracle as O
from tests.fixtures import land_case, oracle_params, perturbed_state


@pytest.fixture(scope="module", params=[False, True], ids=["plain", "projected"])
def audit(request):
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, dict(kwargs, project_adv_vel=request.param))
    s = perturbed_state(p, T, S)
    out, led = L.budget_step(s, p)
    return p, s, out, led


def test_ledger_step_is_the_step(audit):
    p, s, out, _ = audit
    ref = O.step(s, p)
    for f in ("u", "v", "T", "S", "eta"):
        np.testing.assert_array_equal(getattr(out, f), getattr(ref, f))


def test_stage_changes_telescope_to_the_observed_change(audit):
    _, _, _, led = audit
    scale = led["change_scale"]
    assert np.all(scale[:2] > 0.0)
    assert np.all(np.abs(led["decomposition_residual"]) <= 1e-12 * scale + 1e-300)
    np.testing.assert_allclose(led["stage_changes"].sum(axis=0), led["observed_change"], rtol=0,
                               atol=1e-12 * scale.max())


def test_nonlinear_stage_equals_sources_plus_processes(audit):
    _, _, _, led = audit
    scale = led["nonlinear_process_scale"]
    assert scale[0] > 0.0 and scale[1] > 0.0
    res = led["nonlinear_accounting_residual"]
    assert abs(res[0]) <= 1e-10 * scale[0] and abs(res[1]) <= 1e-10 * scale[1], (res, scale)
    # the sources are not negligible: the check has teeth
    assert np.abs(led["source_inputs"][:6, 0]).sum() > 1e-4 * scale[0]


def test_flux_form_advection_integrates_to_the_surface_flux(audit):
    _, _, _, led = audit
    res = led["advection_boundary_residual"]
    scale = led["nonlinear_process_scale"]
    assert abs(res[0]) <= 1e-10 * scale[0] and abs(res[1]) <= 1e-10 * scale[1], (res, scale)
    assert abs(led["nonlinear_process_changes"][0, 0]) > 1e-6 * scale[0]


def test_free_surface_moves_volume_only_and_inactive_stages_are_zero(audit):
    p, s, out, led = audit
    names = L.STAGE_NAMES
    fs = led["stage_changes"][names.index("free_surface")]
    assert fs[0] == 0.0 and fs[1] == 0.0
    area = p.dx_2d * p.dy * p.wet_mask
    np.testing.assert_allclose(fs[2], np.sum((out.eta - s.eta) * area), rtol=1e-12)
    for n in ("sponge", "dynamic_ice"):
        assert np.all(led["stage_changes"][names.index(n)] == 0.0)
    for n in ("sponge", "ice_atmosphere_heat", "ice_brine"):
        assert np.all(led["source_inputs"][L.SOURCE_NAMES.index(n)] == 0.0)
    assert np.all(led["nonlinear_process_changes"][2:] == 0.0)


def test_projection_entries(audit):
    p, _, _, led = audit
    if p.project_adv_vel:
        before, after = led["projection_transport_norm_squared"]
        assert before > 0.0 and after <= 1e-16 * before
        assert led["projection_relative_residual_max"] <= 1e-8
    else:
        assert np.all(led["projection_transport_norm_squared"] == 0.0)


def test_planted_wrong_source_interval_breaks_the_accounting():
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, kwargs)
    s = perturbed_state(p, T, S)
    _, led = L.budget_step(s, p, source_interval=p.dt)
    scale = led["nonlinear_process_scale"]
    assert abs(led["nonlinear_accounting_residual"][0]) > 1e-3 * scale[0]


def test_accumulate_keeps_worst_ratios_and_sums_the_rest():
    a, b = L.empty_budget(), L.empty_budget()
    a["projection_relative_residual_max"] = np.float64(0.3)
    b["projection_relative_residual_max"] = np.float64(0.1)
    a["observed_change"] = np.array([1.0, 2.0, 3.0])
    b["observed_change"] = np.array([0.5, 0.5, 0.5])
    t = L.accumulate_budget(a, b)
    assert t["projection_relative_residual_max"] == 0.3
    np.testing.assert_array_equal(t["observed_change"], [1.5, 2.5, 3.5])
