"""Stage ledger on the GPU (zm_budget_step) against the oracle's recorder
(reference diagnostics/budgets.py:62-270), through the C ABI.

The budget step must advance the state bit for bit like the plain step, and
its audit must match oracle.zm_ledger.budget_step entry by entry.

Tolerances: every heat/salt/displacement entry of metric c within
tol * (change_scale[c] + nonlinear_process_scale[c]) of the oracle, tol = 1e-11
in float64 (float64 sums formed in another order) and 1e-4 in float32 (the
oracle steps the same float32-rounded input in float64; the float32 step
rounds every node to ~1e-7 of its value, about 1e-3 of one 600 s step's
change of a 20 degC node, which sums to ~3e-5 of the change scale on the
land case; change_scale, a sum of absolute node changes, keeps that rounding
without cancellation and gets 1e-3); the projection's before-norm to
1e-10 (float64) / 1e-4 (float32) relative, its after-norm and worst ratio
as round-off.  On the device ledger itself the stage decomposition closes to
1e-12 of the change scale in both dtypes (differences of the stored stage
states, summed in float64).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import zm_ledger as L  # noqa: E402
from tests.fixtures import land_case, oracle_params, perturbed_state  # noqa: E402
from zhenmode_amd import budgets  # noqa: E402
from zhenmode_amd.cases import make_case  # noqa: E402
from zhenmode_amd.solver import JaxStateG, make_solver_global, to_numpy_state  # noqa: E402

FIELDS = ("u", "v", "T", "S", "eta")
METRIC_KEYS = ("observed_change", "stage_changes", "source_inputs", "decomposition_residual", "budget_residual",
               "change_scale", "nonlinear_process_changes", "nonlinear_process_scale",
               "advection_boundary_changes", "advection_boundary_residual", "nonlinear_accounting_residual",
               "surface_displacement_tracer_change")


def land(dtype, projected):
    grid, physics, dt, kwargs, T, S = land_case()
    kw = dict(kwargs, project_adv_vel=projected)
    step, _, _ = make_solver_global(grid, physics, dt, dtype=dtype, mode_split=True, **kw)
    eps = float(np.finfo(np.float32 if dtype == "float32" else np.float64).eps)
    p = oracle_params(grid, physics, dt, dict(kw, projection_eps=eps))
    return step.solver, p, perturbed_state(p, T, S)


@pytest.mark.parametrize("projected", [False, True], ids=["plain", "projected"])
@pytest.mark.parametrize("dtype,tol", [("float64", 1e-11), ("float32", 1e-4)])
def test_ledger_matches_oracle(dtype, tol, projected):
    print(f"case land ledger {dtype} projected={projected}", flush=True)
    sol, p, s = land(dtype, projected)
    if dtype == "float32":
        s = s._replace(**{f: np.asarray(getattr(s, f)).astype(np.float32).astype(np.float64) for f in FIELDS})
    advance = budgets.make_budget_step(sol)
    new, got = advance(JaxStateG(*s))
    plain = sol.step(JaxStateG(*s))
    for f in FIELDS:
        assert torch.equal(getattr(new, f), getattr(plain, f)), f
    _, ref = L.budget_step(s, p)
    scale = ref["change_scale"] + ref["nonlinear_process_scale"]
    bad = []
    for key in METRIC_KEYS:
        g, r = np.asarray(got[key]), np.asarray(ref[key])
        t = tol * (10.0 if dtype == "float32" and key == "change_scale" else 1.0)
        err = np.max(np.abs(g - r) / (t * scale + 1e-300), axis=tuple(range(g.ndim - 1)) if g.ndim > 1 else None)
        print(key, np.array2string(np.atleast_1d(err), precision=3), flush=True)
        if not np.all(err <= 1.0):
            bad.append(key)
    assert not bad, bad
    for key in ("transport_consistency_max", "bottom_drag_reference_energy_loss_J", "transport_audited_steps"):
        assert np.all(np.asarray(got[key]) == 0.0)
    # the device ledger closes its own stage decomposition
    assert np.all(np.abs(got["decomposition_residual"]) <= 1e-12 * got["change_scale"] + 1e-300)
    nb, na = got["projection_transport_norm_squared"]
    rb, ra = ref["projection_transport_norm_squared"]
    if projected:
        assert abs(nb - rb) <= (1e-10 if dtype == "float64" else 1e-4) * rb
        assert na <= (1e-16 if dtype == "float64" else 1e-8) * nb
        assert got["projection_relative_residual_max"] <= (1e-8 if dtype == "float64" else 1e-4)
    else:
        assert nb == na == 0.0


def test_ledger_identities_and_device_accumulation_at_4_degrees():
    print("case g4deg ledger float32", flush=True)
    case = make_case("g4deg")
    step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype="float32",
                                             **case.solver_kwargs)
    sol = step.solver
    advance = budgets.make_budget_step(sol)
    state = init_state(case.T_init, case.S_init)
    total = torch.zeros(budgets.LEDGER_SIZE, dtype=torch.float64, device=sol.device)
    host_total = budgets.empty_budget()
    for _ in range(3):
        state, led = advance(state, raw=True)
        total = budgets.accumulate_ledger(total, led)
        host_total = budgets.accumulate_budget(host_total, budgets.ledger_dict(led))
        d = budgets.ledger_dict(led)
        scale = d["change_scale"]
        assert np.all(np.abs(d["decomposition_residual"]) <= 1e-12 * scale + 1e-300)
        ps = d["nonlinear_process_scale"]
        assert np.all(np.abs(d["advection_boundary_residual"][:2]) <= 1e-5 * ps[:2])
        print("accounting", d["nonlinear_accounting_residual"][:2] / ps[:2], flush=True)
    acc = budgets.ledger_dict(total)
    for k, v in host_total.items():
        np.testing.assert_allclose(acc[k], v, rtol=1e-15, atol=0.0, err_msg=k)
    ref = to_numpy_state(sol.step(init_state(case.T_init, case.S_init), nsteps=3))
    got = to_numpy_state(state)
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(got, f), getattr(ref, f))
