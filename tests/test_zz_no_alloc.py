"""The allocation-free Thomas solve (zm_vertical_mix_noalloc) against
zm_vertical_mix, through the C ABI, on two calls in a row: a 1-degree grid
whose 57,600 columns take three column ranges of the static scratch, then a
smaller 4-degree grid in the other dtype; each with and without a column mask.
The results must be equal bit for bit (the same per-column arithmetic)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from zhenmode_amd import _lib  # noqa: E402
from zhenmode_amd.cases import make_case  # noqa: E402
from zhenmode_amd.solver import empty3, make_solver_global, ptr  # noqa: E402


def _mix(sol, entry, T, S, dt, kv, kc, only):
    To = empty3(*sol.shape3, sol.dtype, sol.device).fill_(-999.0)
    So = empty3(*sol.shape3, sol.dtype, sol.device).fill_(-999.0)
    rc = getattr(sol.lib, entry)(sol.handle, ptr(T), ptr(S), ptr(To), ptr(So), float(dt), float(kv), float(kc),
                                 ptr(only) if only is not None else None, sol._stream())
    _lib.check(rc, sol.handle)
    return To, So


def _check_case(name, nx, ny, dtype):
    case = make_case(name, nx=nx, ny=ny)
    g = case.grid
    print(f"case {name} {g.nx}x{g.ny}x{g.nz} {dtype}: zm_vertical_mix_noalloc vs zm_vertical_mix", flush=True)
    step, init_state, _ = make_solver_global(g, case.physics, case.dt, dtype=dtype, **case.solver_kwargs)
    sol = step.solver
    s = step(init_state(case.T_init, case.S_init))
    rng = np.random.default_rng(11)
    only = torch.as_tensor(rng.integers(0, 2, g.nx * g.ny, dtype=np.int32), device=sol.device)
    kv, kc, dt = 1e-4, 0.05, case.dt
    for mask in (None, only):
        # the allocation-free entry first, so it never depends on the other's allocation
        a = _mix(sol, "zm_vertical_mix_noalloc", s.T, s.S, dt, kv, kc, mask)
        b = _mix(sol, "zm_vertical_mix", s.T, s.S, dt, kv, kc, mask)
        torch.cuda.synchronize()
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), (name, dtype, mask is None)
        changed = int((a[0] != s.T).sum())
        print(f"  mask={'none' if mask is None else 'half'}: equal, {changed} T values changed", flush=True)
        assert changed > 0
    with pytest.raises(ValueError, match="alias"):
        _lib.check(sol.lib.zm_vertical_mix_noalloc(sol.handle, ptr(s.T), ptr(s.S), ptr(s.T), ptr(s.S), dt, kv, kc,
                                                   None, sol._stream()), sol.handle)


def test_noalloc_thomas_equals_the_allocating_entry_point_on_two_calls():
    _check_case("g1deg", 360, 160, "float64")      # three column ranges
    _check_case("g4deg", 90, 40, "float32")        # the second, smaller call: one range
