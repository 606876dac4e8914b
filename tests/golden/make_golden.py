"""Write the golden vectors of the mode-split step with the NumPy oracle.

    python tests/golden/make_golden.py

Outputs (float64, reference layout (nx, ny, nz)):
  g4deg_8steps.npz   -- 90x40x12 case (4 degrees), initial state and the state
                     after 1 and 8 steps.
  land_1step.npz  -- 24x20x6 grid with land, shallow one-level columns and all
                     processes on: input state, every stage of one step.
The inputs are regenerated from the seeded case generators, so a fixture only
stores what a test compares against.
"""

from pathlib import Path

import numpy as np

from oracle import zm_oracle as O
from tests.fixtures import land_case, oracle_params, perturbed_state
from zhenmode_amd.cases import make_case

HERE = Path(__file__).resolve().parent


def g4deg_case():
    case = make_case("g4deg")
    p = oracle_params(case.grid, case.physics, case.dt, case.solver_kwargs)
    return case, p, O.init_state(p, case.T_init, case.S_init)


def main():
    case, p, s = g4deg_case()
    out = {"init_" + f: getattr(s, f) for f in ("u", "v", "T", "S", "eta")}
    for n in range(8):
        s = O.step(s, p)
        if n in (0, 7):
            out.update({f"step{n + 1}_" + f: getattr(s, f) for f in ("u", "v", "T", "S", "eta")})
    np.savez_compressed(HERE / "g4deg_8steps.npz", **out)

    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, kwargs)
    s0 = perturbed_state(p, T, S)
    stages = {}
    O.step(s0, p, stages=stages)
    out = {"in_" + f: getattr(s0, f) for f in ("u", "v", "T", "S", "eta")}
    for name, st in stages.items():
        out.update({f"{name}_" + f: getattr(st, f) for f in ("u", "v", "T", "S", "eta")})
    np.savez_compressed(HERE / "land_1step.npz", **out)


if __name__ == "__main__":
    main()
