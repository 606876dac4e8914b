"""The persistent barotropic sub-cycle (all substeps in one launch, blocks
synchronised with their two neighbours through progress counters) against the
same sub-cycle launched once per substep: the steps must be equal bit for bit,
in captured graphs (replayed several times, with eager work in between) and in
eager issue, and no bounded wait may have given up."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from tests.fixtures import land_case, oracle_params, perturbed_state  # noqa: E402
from zhenmode_amd.cases import make_case  # noqa: E402
from zhenmode_amd.solver import JaxStateG, make_solver_global  # noqa: E402

FIELDS = ("u", "v", "T", "S", "eta", "ice")


def _solver_and_state(name, dtype):
    if name.startswith("land"):
        ny = 21 if name == "land_ny21" else 20
        grid, physics, dt, kwargs, T, S = land_case(ny=ny)
        step, _, _ = make_solver_global(grid, physics, dt, dtype=dtype, mode_split=True, **kwargs)
        p = oracle_params(grid, physics, dt, kwargs)
        return step.solver, JaxStateG(*perturbed_state(p, T, S))
    case = make_case(name)
    step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype=dtype, **case.solver_kwargs)
    sol = step.solver
    return sol, sol.step(init_state(case.T_init, case.S_init))   # one step in: eta, ubt nonzero


def _equal(a, b):
    return all(torch.equal(getattr(a, f), getattr(b, f)) for f in FIELDS)


@pytest.mark.parametrize("name,dtype", [("g4deg", "float64"), ("g4deg", "float32"), ("land", "float64"),
                                        ("land_ny21", "float64"), ("land_ny21", "float32"),
                                        ("g05deg", "float32")])
def test_persistent_subcycle_equals_per_substep_launches(name, dtype):
    print(f"case {name} {dtype}: persistent barotropic sub-cycle vs one launch per substep", flush=True)
    sol, s0 = _solver_and_state(name, dtype)
    ny = sol.shape2[1]
    sol.set_barotropic_persistent(False)
    ref = [sol.step(s0)]
    for _ in range(2):
        ref.append(sol.step(ref[-1]))
    sol.set_barotropic_persistent(True)
    for graph in (True, False):
        sol.set_graph(graph)
        s = s0
        for k in range(3):
            s = sol.step(s)
            torch.cuda.synchronize()
            # eager work of the program between graph replays
            assert float(s.T.double().abs().max()) > 0.0
            assert _equal(s, ref[k]), (name, dtype, graph, k)
    sol.set_graph(True)
    assert sol.barotropic_timeouts() == 0
    print(f"  {(ny + 1) // 2} blocks, 3 steps equal in graph and eager issue, no timeouts", flush=True)
