"""Column-divergence projection on the GPU against the oracle
(reference dynamics/projection.py:64-127), through the C ABI.

zm_project_velocity is compared with oracle.project_column_divergence on the
land case for both preconditioners, with the 150-iteration default and with a
5-iteration cap that leaves a transport residual so both refinements run; the
float32 handle is compared with the oracle applied to the same
float32-rounded velocities with the float32 tolerance floors (its solver
arithmetic is float64, so only the rounding of the result to float32
differs).  Steps with project_adv_vel are compared with oracle.step, and the
graph replay with eager issue.

Tolerances:
  projection, float64: max |u_gpu - u_oracle| <= 1e-9 max |u_oracle - u_in|
    (the CG sums are formed in another order; up to 150 iterations amplify
    those round-off differences); float32: 2 ulp of max |u| plus 1e-9 of the
    correction.  Iteration counts equal (within one where the stopping test is
    met on an iteration, since its decision can flip on round-off).
  step: float64 1e-11 relative per field; float32 as the flag tests
    (u, v, eta 1e-4, T, S 2e-6 of max |field|).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import zm_oracle as O  # noqa: E402
from tests.fixtures import land_case, oracle_params, perturbed_state  # noqa: E402
from zhenmode_amd.cases import make_case  # noqa: E402
from zhenmode_amd.solver import JaxStateG, make_solver_global, to_numpy_state  # noqa: E402

EPS = {"float64": float(np.finfo(np.float64).eps), "float32": float(np.finfo(np.float32).eps)}
FIELDS = ("u", "v", "T", "S", "eta")


def setup(dtype, **opts):
    grid, physics, dt, kwargs, T, S = land_case()
    kw = dict(kwargs, project_adv_vel=True, **opts)
    step, _, _ = make_solver_global(grid, physics, dt, dtype=dtype, mode_split=True, **kw)
    p = oracle_params(grid, physics, dt, dict(kw, projection_eps=EPS[dtype]))
    return step.solver, p, perturbed_state(p, T, S), dt


def to_np(t):
    return t.double().cpu().numpy()


@pytest.mark.parametrize("niter", [150, 5])
@pytest.mark.parametrize("pre", ["none", "jacobi"])
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_projection_matches_oracle(dtype, pre, niter):
    print(f"case land projection {dtype} {pre} niter={niter}", flush=True)
    sol, p, s, dt = setup(dtype, projection_preconditioner=pre, projection_niter=niter)
    u_in, v_in = s.u, s.v
    if dtype == "float32":
        u_in, v_in = (a.astype(np.float32).astype(np.float64) for a in (s.u, s.v))
    uo, vo, its = sol.project_velocity(u_in, v_in, dt=dt, iterations=True)
    (ur, vr), its_ref = O.project_column_divergence(u_in, v_in, p, dt, return_iterations=True)
    its = [int(x) for x in its.cpu()]
    its_ref = list(its_ref) + [0] * (3 - len(its_ref))
    print("iterations", its, its_ref, flush=True)
    for a, b in zip(its, its_ref):
        assert a == b or (abs(a - b) <= 1 and b < niter), (its, its_ref)
    for got, ref, inp, name in ((uo, ur, u_in, "u"), (vo, vr, v_in, "v")):
        got = to_np(got)
        corr = np.max(np.abs(ref - inp))
        err = np.max(np.abs(got - ref))
        bound = 1e-9 * corr + (2 * EPS["float32"] * np.max(np.abs(ref)) if dtype == "float32" else 0.0)
        print(name, f"err {err:.2e} correction {corr:.2e}", flush=True)
        assert err <= bound, (name, err, bound)
        dry = p.wet_mask_z < 0.5
        np.testing.assert_array_equal(got[dry], inp[dry])
    if niter == 150:
        before = np.abs(O.column_divergence(u_in, v_in, p)).max()
        after = np.abs(O.column_divergence(to_np(uo), to_np(vo), p)).max()
        assert after <= (1e-9 if dtype == "float64" else 1e-4) * before, (after, before)


@pytest.mark.parametrize("dtype,tol", [("float64", dict(u=1e-11, v=1e-11, T=1e-11, S=1e-11, eta=1e-11)),
                                       ("float32", dict(u=1e-4, v=1e-4, T=2e-6, S=2e-6, eta=1e-4))])
@pytest.mark.parametrize("pre", ["none", "jacobi"])
def test_step_with_projection_matches_oracle(dtype, tol, pre):
    print(f"case land step project_adv_vel {dtype} {pre}", flush=True)
    sol, p, s, dt = setup(dtype, projection_preconditioner=pre)
    got = to_numpy_state(sol.step(JaxStateG(*s)))
    ref = O.step(s, p)
    for f in FIELDS:
        a, b = getattr(got, f), getattr(ref, f)
        e = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        print(f, f"{e:.1e}", flush=True)
        assert e <= tol[f], (f, e)
    # the projection changes the step
    q = oracle_params(*land_case()[:3], land_case()[3])
    plain = O.step(s, q)
    assert np.max(np.abs(plain.T - ref.T)) > 10 * tol["T"] * np.max(np.abs(ref.T))


def test_projection_graph_replays_equal_eager_issue():
    print("case g4deg step project_adv_vel float32 graph", flush=True)
    case = make_case("g4deg")
    step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype="float32",
                                             project_adv_vel=True, projection_preconditioner="jacobi",
                                             **case.solver_kwargs)
    sol = step.solver
    s0 = init_state(case.T_init, case.S_init)
    s1 = sol.step(s0)
    sol.set_graph(False)
    eager = [sol.step(s0), sol.step(s1)]
    its_eager = sol.projection_iterations()
    sol.set_graph(True)
    for rep in range(3):
        out = [sol.step(s0), sol.step(s1)]
        torch.cuda.synchronize()
        for a, b in zip(out, eager):
            for f in FIELDS:
                assert torch.equal(getattr(a, f), getattr(b, f)), (rep, f)
    assert sol.projection_iterations() == its_eager
    assert its_eager[0] > 0


def test_projection_removes_the_column_divergence_at_4_degrees():
    print("case g4deg projection float64 jacobi", flush=True)
    case = make_case("g4deg")
    # the step itself does not project; zm_project_velocity uses the handle's settings
    kw = dict(case.solver_kwargs, projection_preconditioner="jacobi", projection_niter=2000)
    step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype="float64", **kw)
    sol = step.solver
    s = to_numpy_state(sol.step(init_state(case.T_init, case.S_init), nsteps=3))
    p = oracle_params(case.grid, case.physics, case.dt, dict(kw, mode_split=True))
    uo, vo, its = sol.project_velocity(s.u, s.v, iterations=True)
    its = [int(x) for x in its.cpu()]
    before = np.abs(O.column_divergence(s.u, s.v, p)).max()
    after = np.abs(O.column_divergence(to_np(uo), to_np(vo), p)).max()
    print("iterations", its, f"divergence {before:.2e} -> {after:.2e}", flush=True)
    assert 0 < its[0] < 2000
    assert after <= 1e-9 * before
    dry = p.wet_mask_z < 0.5
    np.testing.assert_array_equal(to_np(uo)[dry], s.u[dry])
