"""Device-side run-time helpers against the oracle: seasonal forcing blend,
monitor (classify_state codes and first-rejection bookkeeping), float64
budget sums and the ring-buffer run loop.

Tolerances: blend float64 <= 1e-14 absolute on O(1) fields, float32 <= 1e-6;
budget float64 <= 1e-13 relative (compensated sums of ~3000 cells); monitor
metrics exact (maxima of absolute values); run-loop states bit-equal to
separate steps.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import zm_oracle as O  # noqa: E402
from oracle import zm_runtime as R  # noqa: E402
from tests.fixtures import land_case, oracle_params, perturbed_state  # noqa: E402
from zhenmode_amd import runtime  # noqa: E402
from zhenmode_amd.cases import make_case  # noqa: E402
from zhenmode_amd.solver import JaxStateG, make_solver_global  # noqa: E402


def land_solver(dtype):
    grid, physics, dt, kwargs, T, S = land_case()
    step, _, _ = make_solver_global(grid, physics, dt, dtype=dtype, mode_split=True, **kwargs)
    p = oracle_params(grid, physics, dt, kwargs)
    return step.solver, p, perturbed_state(p, T, S)


@pytest.mark.parametrize("dtype,tol", [("float64", 1e-14), ("float32", 1e-6)])
def test_seasonal_blend_matches_oracle(dtype, tol):
    print(f"case land seasonal {dtype}", flush=True)
    sol, p, _ = land_solver(dtype)
    rng = np.random.default_rng(0)
    st = rng.standard_normal((12, 3, p.ny, p.nx))
    dev = torch.as_tensor(st, dtype=sol.dtype, device=sol.device)
    for day in (0.0, 1.3, 28.9, 29.5, 30.0, 45.0, 359.2, 400.0):
        got = runtime.seasonal_forcing(sol, dev, day).double().cpu().numpy()
        ref = R.interp_monthly_field(st, day)
        assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref))), day


def test_monitor_codes_and_first_rejection():
    print("case land monitor float64", flush=True)
    sol, p, s = land_solver("float64")
    mon = runtime.DeviceMonitor(sol)
    ref = R.classify_state(s)
    mon.classify(JaxStateG(*s), 1)
    st = mon.read()
    assert st["failure"] == 0 and st["finite"] == 1 and st["first_failure_step"] == -1
    assert st["max_u"] == ref["max_u"] and st["max_velocity"] == ref["max_velocity"]
    assert st["max_eta"] == ref["max_eta"]
    u = s.u.copy()
    u[3, 4, 0] = 11.0
    mon.classify(JaxStateG(u, *s[1:]), 2)          # rejected: velocity
    ice = np.array(s.ice, dtype=np.float64) + np.zeros((p.nx, p.ny))
    ice[0, 0] = np.nan
    mon.classify(JaxStateG(*s[:5], ice), 3)         # ignored: after the first rejection
    st = mon.read()
    assert st["failure"] == 3 and st["first_failure_step"] == 2 and st["last_step"] == 2
    assert st["peak_velocity"] == 11.0
    for kwargs, state, code in ((dict(), JaxStateG(*s[:5], ice), 1),
                                (dict(), s._replace(eta=np.full(s.eta.shape, 15.0)), 0),
                                (dict(inclusive=True), s._replace(eta=np.full(s.eta.shape, 15.0)), 4)):
        m = runtime.DeviceMonitor(sol, **kwargs)
        m.classify(JaxStateG(*state), 1)
        assert m.read()["failure"] == code == R.classify_state(state, **kwargs)["failure"]


def test_budget_matches_oracle():
    print("case land budget float64", flush=True)
    sol, p, s = land_solver("float64")
    got = runtime.budget(sol, JaxStateG(*s))
    ref = R.budget_diagnostics(s, p)
    for k, v in ref.items():
        g = getattr(got, k)
        assert abs(g - v) <= 1e-13 * max(abs(v), 1e-300), (k, g, v)


def test_run_loop_accepts_then_rejects_like_the_reference():
    print("case g4deg run loop float32", flush=True)
    case = make_case("g4deg")
    step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype="float32",
                                             **case.solver_kwargs)
    sol = step.solver
    s0 = init_state(case.T_init, case.S_init)
    ref = [s0]
    for _ in range(12):
        ref.append(sol.step(ref[-1]))
    res = runtime.run(lambda a, out: sol.step(a, out=out), sol, s0, 12, check_every=5)
    assert res.accepted_steps == 12 and res.rejected is None and res.status["failure"] == 0
    for f in ("u", "v", "T", "S", "eta"):
        assert torch.equal(getattr(res.state, f), getattr(ref[12], f)), f
    # a velocity limit crossed during the run: the first crossing is rejected
    speeds = [float(max(x.u.abs().max(), x.v.abs().max())) for x in ref]
    limit = 0.5 * (speeds[6] + speeds[7]) if speeds[7] > speeds[6] else None
    if limit is None or any(sp > limit for sp in speeds[:7]):
        limit = speeds[7] * (1 - 1e-6)
    first = next(n for n, sp in enumerate(speeds) if sp > limit)
    res = runtime.run(lambda a, out: sol.step(a, out=out), sol, s0, 12, check_every=4, velocity_limit=limit)
    assert res.status["failure"] == 3 and res.status["first_failure_step"] == first
    assert res.accepted_steps == first - 1
    for f in ("u", "v", "T", "S", "eta"):
        assert torch.equal(getattr(res.state, f), getattr(ref[first - 1], f)), f
        assert torch.equal(getattr(res.rejected, f), getattr(ref[first], f)), f
