"""Whole-step checks of the oracle (CPU): golden vectors, fixed points and
conservation properties of the legacy mode-split step."""

from pathlib import Path

import numpy as np
import pytest

from oracle import zm_oracle as O
from tests.fixtures import land_case, oracle_params, perturbed_state, quiet_physics
from tests.golden.make_golden import g4deg_case
from zhenmode_amd.cases import make_case

GOLDEN = Path(__file__).resolve().parent / "golden"
FIELDS = ("u", "v", "T", "S", "eta")


def test_c0_golden_vectors_reproduce():
    case, p, s = g4deg_case()
    gold = np.load(GOLDEN / "g4deg_8steps.npz")
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(s, f), gold["init_" + f])
    s = O.step(s, p)
    for f in FIELDS:
        np.testing.assert_allclose(getattr(s, f), gold["step1_" + f], rtol=0, atol=1e-12 * max(1.0, np.abs(gold["step1_" + f]).max()))


def test_land_case_golden_stages_reproduce():
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, kwargs)
    s0 = perturbed_state(p, T, S)
    stages = {}
    O.step(s0, p, stages=stages)
    gold = np.load(GOLDEN / "land_1step.npz")
    for name, st in stages.items():
        for f in FIELDS:
            ref = gold[f"{name}_{f}"]
            np.testing.assert_allclose(getattr(st, f), ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_resting_uniform_ocean_is_a_fixed_point():
    # A resting ocean at T_ref, S_ref everywhere has no force: every stage
    # returns it (u, v, eta to round-off zero; T, S to a few ulps of the cap
    # blend wb*x + (1-wb)*x).
    case = make_case("g4deg")
    kw = dict(case.solver_kwargs)
    kw["forcing"] = None
    p = oracle_params(case.grid, case.physics, case.dt, kw)
    s = O.init_state(p)
    out = O.step(s, p)
    assert np.max(np.abs(out.u)) < 1e-12 and np.max(np.abs(out.v)) < 1e-12
    assert np.max(np.abs(out.eta)) < 1e-12
    assert np.max(np.abs(out.T - p.T_ref)) <= 8 * np.spacing(p.T_ref)
    assert np.max(np.abs(out.S - p.S_ref)) <= 8 * np.spacing(p.S_ref)


def test_volume_is_conserved_to_roundoff():
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, kwargs)
    s = perturbed_state(p, T, S)
    area = p.dx_2d * p.dy
    v0 = np.sum(area * s.eta)
    scale = np.sum(area * np.abs(s.eta))
    for _ in range(3):
        s = O.step(s, p)
    assert abs(np.sum(area * s.eta) - v0) <= 1e-11 * scale


def test_land_and_ghost_nodes_hold_their_prestep_values():
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, kwargs)
    s = perturbed_state(p, T, S)
    out = O.step(s, p)
    dry = p.wet_mask_z == 0.0
    for f in ("u", "T", "S"):
        np.testing.assert_array_equal(getattr(out, f)[dry], getattr(s, f)[dry])
    assert np.all(out.v[:, 0] == 0.0) and np.all(out.v[:, -1] == 0.0)
    np.testing.assert_array_equal(out.eta[p.wet_mask == 0.0], 0.0)


def test_c0_run_stays_bounded_over_eight_steps():
    gold = np.load(GOLDEN / "g4deg_8steps.npz")
    assert np.all(np.isfinite(gold["step8_u"]))
    assert np.max(np.abs(gold["step8_u"])) < 1.0
    assert np.max(np.abs(gold["step8_eta"])) < 1.0


def test_barotropic_fb_step_is_neutral_without_damping():
    # D22 (reference barotropic.py:154-161): forward-backward on an all-wet
    # f = 0 channel is neutral below the gravity-wave CFL.
    from zhenmode_amd.grid import all_wet_grid
    grid = all_wet_grid(nx=32, ny=16, nz=4)
    grid.f = np.zeros_like(grid.f)
    p = O.make_params(grid, quiet_physics(r_bot=0.0), 600.0, dt_bt=20.0, polar_cap_rows=0, polar_cap_taper=0)
    eta = np.zeros((32, 16))
    eta[16, 8] = 1.0
    ubt = np.zeros_like(eta)
    vbt = np.zeros_like(eta)
    area = p.dx_2d * p.dy
    z = np.zeros_like(eta)
    peak = []
    for _ in range(600):
        eta, ubt, vbt = O.free_surface_substep(eta, ubt, vbt, z, z, p)
        peak.append(np.max(np.abs(eta)))
    assert abs(np.sum(area * eta) - area[16, 8]) <= 1e-12 * area[16, 8]
    assert max(peak[300:]) <= 1.0 + 1e-9
    assert max(peak[300:]) > 1e-3  # undamped: the wave has not decayed away


@pytest.mark.parametrize("field", ["u", "T"])
def test_one_step_is_deterministic(field):
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, kwargs)
    s = perturbed_state(p, T, S)
    a = O.step(s, p)
    b = O.step(s, p)
    np.testing.assert_array_equal(getattr(a, field), getattr(b, field))
