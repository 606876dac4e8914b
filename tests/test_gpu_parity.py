"""HIP step vs the oracle on the GPU, through the C ABI.

Tolerances (max |HIP - oracle| / max |oracle| per field):
  float64: 1e-11 per stage and over 8 steps (round-off only);
  float32: u, v 2e-4; T, S 1e-5; eta 2e-3 over 8 steps of the 4-degree case
  (the float32 state against the float64 oracle; eta and v are small fields
  whose relative error is set by the float32 cancellation in the momentum
  residual).
"""

import os
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import zm_oracle as O  # noqa: E402
from tests.fixtures import land_case, oracle_params, perturbed_state  # noqa: E402
from zhenmode_amd import _lib  # noqa: E402
from zhenmode_amd.cases import make_case  # noqa: E402
from zhenmode_amd.solver import JaxStateG, make_solver_global, to_numpy_state  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"
FIELDS = ("u", "v", "T", "S", "eta")
TOL32 = dict(u=2e-4, v=2e-4, T=1e-5, S=1e-5, eta=2e-3)
TOL64 = 1e-11


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def check(got, ref, tol, label):
    got = to_numpy_state(got)
    errs = {f: rel(getattr(got, f), getattr(ref, f) if hasattr(ref, f) else ref[f]) for f in FIELDS}
    print(label, " ".join(f"{f}:{e:.1e}" for f, e in errs.items()), flush=True)
    for f, e in errs.items():
        t = tol[f] if isinstance(tol, dict) else tol
        assert e <= t, (label, f, e, t)


def g4deg_solver(dtype, **extra):
    case = make_case("g4deg")
    step, init_state, diag, params, _ = make_solver_global(case.grid, case.physics, case.dt, return_params=True,
                                                           dtype=dtype, **case.solver_kwargs, **extra)
    return case, step, init_state, diag, params


@pytest.mark.parametrize("dtype,tol", [("float64", TOL64), ("float32", TOL32)])
def test_c0_eight_steps_match_golden_vectors(dtype, tol):
    print(f"case g4deg 90x40x12 {dtype}", flush=True)
    gold = np.load(GOLDEN / "g4deg_8steps.npz")
    case, step, init_state, _, _ = g4deg_solver(dtype)
    s = init_state(case.T_init, case.S_init)
    check(s, {f: gold["init_" + f] for f in FIELDS}, TOL64 if dtype == "float64" else 1e-7, "init")
    s = step(s)
    check(s, {f: gold["step1_" + f] for f in FIELDS}, tol, "step1")
    for _ in range(7):
        s = step(s)
    torch.cuda.synchronize()
    check(s, {f: gold["step8_" + f] for f in FIELDS}, tol, "step8")


def _land_solver(dtype):
    grid, physics, dt, kwargs, T, S = land_case()
    step, _, _, params, _ = make_solver_global(grid, physics, dt, return_params=True, dtype=dtype,
                                               mode_split=True, **kwargs)
    p = oracle_params(grid, physics, dt, kwargs)
    return step.solver, p, perturbed_state(p, T, S)


@pytest.mark.parametrize("dtype,tol", [("float64", TOL64), ("float32", dict(u=1e-4, v=1e-4, T=2e-6, S=2e-6, eta=1e-4))])
def test_land_case_every_stage_matches_golden(dtype, tol):
    print(f"case land 24x20x6 {dtype}", flush=True)
    sol, p, s0 = _land_solver(dtype)
    gold = np.load(GOLDEN / "land_1step.npz")
    g = {name: O.State(*(gold[f"{name}_{f}"] for f in FIELDS), np.zeros((p.nx, p.ny)))
         for name in ("L1", "N", "L2", "BT", "out")}
    check(sol.stage(_lib.STAGE_LINEAR, JaxStateG(*s0)), g["L1"], tol, "linear")
    check(sol.stage(_lib.STAGE_NONLINEAR, JaxStateG(*g["L1"])), g["N"], tol, "nonlinear")
    check(sol.stage(_lib.STAGE_LINEAR, JaxStateG(*g["N"])), g["L2"], tol, "linear2")
    bt = O.final_masks(g["BT"], g["L2"], p)
    check(sol.stage(_lib.STAGE_BAROTROPIC, JaxStateG(*g["L2"])), bt, tol, "barotropic")
    check(sol.step(JaxStateG(*s0)), g["out"], tol, "step")


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_residual_stages_match_oracle(dtype):
    print(f"case land residuals {dtype}", flush=True)
    sol, p, s0 = _land_solver(dtype)
    dT, dS = O.tracer_residual(s0.u, s0.v, s0.T, s0.S, p)
    du, dv = O.momentum_residual(s0.u, s0.v, s0.T, s0.S, s0.eta, p)
    # the stage returns the residual with the implicit bottom-drag strip of
    # the legacy RK2 already added (reference step.py:273-277)
    bm = p.bottom_mask * p.wet_mask_z
    du, dv = du + p.r_bot * s0.u * bm, dv + p.r_bot * s0.v * bm
    gt = to_numpy_state(sol.stage(_lib.STAGE_TRACER_RES, JaxStateG(*s0)))
    gm = to_numpy_state(sol.stage(_lib.STAGE_MOMENTUM_RES, JaxStateG(*s0)))
    wet = p.wet_mask_z > 0
    tol = 1e-10 if dtype == "float64" else 2e-3
    for got, ref, name in ((gt.T, dT, "dT"), (gt.S, dS, "dS"), (gm.u, du, "du"), (gm.v, dv, "dv")):
        e = rel(got[wet], ref[wet])
        print(name, f"{e:.1e}", flush=True)
        assert e <= tol, (name, e)


@pytest.mark.parametrize("nx", [720, 800, 480, 360, 77])
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_dealiased_momentum_residual_at_production_row_lengths(nx, dtype):
    """The x de-alias inside the momentum residual at the row lengths of the
    production grids (float64: Stockham FFT over radices 2, 3, 4, 5, a block's
    rows transformed together; float32: the two-factor register FFT; the last
    block partial in both) and at a length with other prime factors (four-step
    DFT), against the oracle's numpy FFT projection.  Float32 errors are about
    1e-5 (float32 rounding of the residual); a wrong root or index would give
    O(1)."""
    print(f"case land {nx}x6x3 momentum residual {dtype}", flush=True)
    grid, physics, dt, kwargs, T, S = land_case(nx=nx, ny=6, nz=3)
    kwargs = dict(kwargs, polar_cap_rows=0, polar_cap_taper=0)  # six rows leave no room for the cap bands
    step, _, _ = make_solver_global(grid, physics, dt, dtype=dtype, mode_split=True, **kwargs)
    p = oracle_params(grid, physics, dt, kwargs)
    s0 = perturbed_state(p, T, S)
    du, dv = O.momentum_residual(s0.u, s0.v, s0.T, s0.S, s0.eta, p)
    bm = p.bottom_mask * p.wet_mask_z
    du, dv = du + p.r_bot * s0.u * bm, dv + p.r_bot * s0.v * bm
    gm = to_numpy_state(step.solver.stage(_lib.STAGE_MOMENTUM_RES, JaxStateG(*s0)))
    wet = p.wet_mask_z > 0
    tol = 1e-10 if dtype == "float64" else 1e-4
    for got, ref, name in ((gm.u, du, "du"), (gm.v, dv, "dv")):
        e = rel(got[wet], ref[wet])
        print(name, f"{e:.1e}", flush=True)
        assert e <= tol, (name, e)


@pytest.mark.parametrize("dtype,block,chunk", [("float64", "", ""), ("float32", "", ""), ("float32", "3", "8"),
                                               ("float32", "2", "16"), ("float64", "4", "0")],
                         ids=["f64-stream4", "f32-stream4", "f32-stream3-chunk8", "f32-stream2-chunk16",
                              "f64-tile4"])
def test_blocked_nu_subcycle_equals_single_substeps(dtype, block, chunk):
    """The temporally blocked nu_h kernels (streaming chunks of `chunk` rows,
    or the register-tile kernel for chunk 0) give the single-substep bits:
    4-degree grid, 40 rows, so chunk boundaries and both walls are crossed."""
    print(f"case g4deg nu blocking {dtype} block={block or 'max'} chunk={chunk or 'default'}", flush=True)
    case = make_case("g4deg")
    outs = []
    for single in (True, False):
        env = {"ZHENMODE_NU_BLOCK": "1"} if single else {"ZHENMODE_NU_BLOCK": block, "ZHENMODE_NU_CHUNK": chunk}
        for k, v in env.items():
            if v:
                os.environ[k] = v
        try:
            step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype=dtype,
                                                     **case.solver_kwargs)
        finally:
            for k in env:
                os.environ.pop(k, None)
        if not single:
            assert step.solver.nu_substeps_per_launch() == (int(block) if block else 4)
        s = init_state(case.T_init, case.S_init)
        s = step(step(s))
        outs.append(to_numpy_state(s))
    for f in FIELDS:  # identical arithmetic per node: bit-equal
        np.testing.assert_array_equal(getattr(outs[1], f), getattr(outs[0], f), err_msg=f)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_blocked_nu_subcycle_with_land_velocities_off_rest(dtype):
    """Velocities on land need not be at rest (a restart may hold any values):
    with random u, v on every node of the southern half and zero in the
    northern half, the streaming kernel's all-land tiles take both paths
    (nonzero inputs: the march; all-zero inputs: the +0 shortcut), and one
    step still gives the single-substep bits."""
    print(f"case g4deg nu blocking, land velocities off rest, {dtype}", flush=True)
    case = make_case("g4deg")
    outs = []
    for single in (True, False):
        if single:
            os.environ["ZHENMODE_NU_BLOCK"] = "1"
        try:
            step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype=dtype,
                                                     **case.solver_kwargs)
        finally:
            os.environ.pop("ZHENMODE_NU_BLOCK", None)
        s = init_state(case.T_init, case.S_init)
        gen = torch.Generator().manual_seed(7)
        for f in (s.u, s.v):  # (nx, ny, nz) views
            x = 0.05 * torch.randn(f.shape, generator=gen, dtype=torch.float64)
            x[:, f.shape[1] // 2:, :] = 0.0
            f.copy_(x.to(f.dtype))
        outs.append(to_numpy_state(step(s)))
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(outs[1], f), getattr(outs[0], f), err_msg=f)


def test_graph_replays_match_eager_steps():
    print("case g4deg graph replay float32", flush=True)
    case, step, init_state, _, _ = g4deg_solver("float32")
    sol = step.solver
    s0 = init_state(case.T_init, case.S_init)
    sol.set_graph(False)
    eager = [s0]
    for _ in range(4):
        eager.append(sol.step(eager[-1]))
    sol.set_graph(True)
    a, b = s0, sol.new_state()
    c = sol.new_state()
    for n in range(4):
        sol.step(a, out=b)
        torch.cuda.synchronize()
        _ = (b.T * 2.0).sum().item()  # unrelated eager work between replays
        for f in ("u", "v", "T", "S", "eta"):
            assert torch.equal(getattr(b, f), getattr(eager[n + 1], f)), (n, f)
        a, b, c = b, c, a


def test_step_is_functional_and_step_dyn_matches_frozen_forcing():
    print("case g4deg functional float32", flush=True)
    case = make_case("g4deg")
    step, init_state, _, step_dyn = make_solver_global(case.grid, case.physics, case.dt, dtype="float32",
                                                       dynamic_forcing=True, **case.solver_kwargs)
    s0 = init_state(case.T_init, case.S_init)
    before = to_numpy_state(s0)
    s1 = step(s0)
    after = to_numpy_state(s0)
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(before, f), getattr(after, f))
    tau_x, tau_y, q = case.wind(0.0)
    s2 = step_dyn(s0, tau_x, tau_y, q)
    for f in FIELDS:
        assert torch.equal(getattr(s1, f), getattr(s2, f)), f
    s3 = step_dyn(s0, 2.0 * tau_x, tau_y, q)
    assert not torch.equal(s1.u, s3.u)


def test_diagnostics_match_oracle():
    print("case land diagnostics float64", flush=True)
    grid, physics, dt, kwargs, T, S = land_case()
    _, _, diag, params, _ = make_solver_global(grid, physics, dt, return_params=True, dtype="float64",
                                               mode_split=True, **kwargs)
    p = oracle_params(grid, physics, dt, kwargs)
    s = perturbed_state(p, T, S)
    got = [np.asarray(t.cpu()) for t in diag(JaxStateG(*s))]
    ref = O.diagnostics(s, p)
    for g_, r, name in zip(got, ref, ("rho", "p", "w")):
        assert rel(g_, r) <= 1e-12, name


def test_init_state_masks_land_and_ghost_like_the_reference():
    case, _, init_state, _, params = g4deg_solver("float64")
    s = to_numpy_state(init_state(case.T_init, case.S_init))
    p = oracle_params(case.grid, case.physics, case.dt, case.solver_kwargs)
    ref = O.init_state(p, case.T_init, case.S_init)
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(s, f), getattr(ref, f))


FLAG_SETS = [dict(conservative_kv=True), dict(localize_conv=True), dict(fct_adv=True), dict(monotone_adv=True),
             dict(conservative_kv=True, localize_conv=True, fct_adv=True)]


@pytest.mark.parametrize("flags", FLAG_SETS, ids=lambda f: "+".join(sorted(f)))
@pytest.mark.parametrize("dtype,tol", [("float64", TOL64), ("float32", dict(u=1e-4, v=1e-4, T=2e-6, S=2e-6, eta=1e-4))])
def test_production_flags_match_oracle(flags, dtype, tol):
    print(f"case land flags {sorted(flags)} {dtype}", flush=True)
    grid, physics, dt, kwargs, T, S = land_case()
    kw = dict(kwargs, **flags)
    step, _, _, params, _ = make_solver_global(grid, physics, dt, return_params=True, dtype=dtype,
                                               mode_split=True, **kw)
    sol = step.solver
    p = oracle_params(grid, physics, dt, kw)
    s0 = perturbed_state(p, T, S)
    assert np.any(O.unstable_interfaces(s0.T, s0.S, p))
    stages = {}
    ref = O.step(s0, p, stages=stages)
    check(sol.stage(_lib.STAGE_LINEAR, JaxStateG(*s0)), stages["L1"], tol, "linear")
    check(sol.stage(_lib.STAGE_NONLINEAR, JaxStateG(*stages["L1"])), stages["N"], tol, "nonlinear")
    check(sol.step(JaxStateG(*s0)), ref, tol, "step")
    dT, dS = O.tracer_residual(s0.u, s0.v, s0.T, s0.S, p)
    gt = to_numpy_state(sol.stage(_lib.STAGE_TRACER_RES, JaxStateG(*s0)))
    wet = p.wet_mask_z > 0.5
    rtol = 1e-10 if dtype == "float64" else 2e-3
    for got, r, name in ((gt.T, dT, "dT"), (gt.S, dS, "dS")):
        e = rel(got[wet], r[wet])
        print(name, f"{e:.1e}", flush=True)
        assert e <= rtol, (name, e)


@pytest.mark.parametrize("cap", [0, 2], ids=["nocap", "cap2"])
@pytest.mark.parametrize("dtype,tol", [("float64", TOL64), ("float32", dict(u=1e-4, v=1e-4, T=2e-6, S=2e-6, eta=1e-4))])
def test_odd_row_count_matches_oracle(cap, dtype, tol):
    """21 latitude rows: the barotropic substep kernel's last block owns one
    row (its second row and halo are clamped), with and without the polar cap."""
    print(f"case land ny=21 polar_cap_rows={cap} {dtype}", flush=True)
    grid, physics, dt, kwargs, T, S = land_case(ny=21)
    kw = dict(kwargs, polar_cap_rows=cap, polar_cap_taper=2 if cap else 0)
    step, _, _ = make_solver_global(grid, physics, dt, dtype=dtype, mode_split=True, **kw)
    p = oracle_params(grid, physics, dt, kw)
    s0 = perturbed_state(p, T, S)
    ref = O.step(s0, p)
    s = step.solver.step(JaxStateG(*s0))
    check(s, ref, tol, "step")
    for _ in range(2):
        ref = O.step(ref, p)
        s = step.solver.step(s)
    check(s, ref, tol if dtype == "float64" else {f: 3 * t for f, t in tol.items()}, "3 steps")


@pytest.mark.parametrize("nc", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("local", [False, True], ids=["column", "localized"])
@pytest.mark.parametrize("dtype,tol", [("float64", TOL64), ("float32", dict(u=1e-4, v=1e-4, T=2e-6, S=2e-6, eta=1e-4))])
def test_convective_subcycle_counts(nc, local, dtype, tol):
    """Convective sub-cycles of 1-4 substeps run in registers in one pass down
    the column, longer ones in LDS; kappa_conv is chosen so that
    ceil(kappa_conv dt / (0.4 min dz^2)) gives each count."""
    grid, physics, dt, kwargs, T, S = land_case()
    min_dz = float(np.min(-np.diff(grid.z)))
    physics = replace(physics, kappa_conv=(nc - 0.5) * 0.4 * min_dz ** 2 / dt)
    kw = dict(kwargs, localize_conv=True) if local else kwargs
    print(f"case land conv_nsub={nc} localize_conv={local} {dtype}", flush=True)
    p = oracle_params(grid, physics, dt, kw)
    assert p.conv_nsub == nc
    step, _, _ = make_solver_global(grid, physics, dt, dtype=dtype, mode_split=True, **kw)
    s0 = perturbed_state(p, T, S)
    assert O.convective_mask(s0.T, s0.S, p).any()
    ref = O.step(s0, p)
    s = step.solver.step(JaxStateG(*s0))
    check(s, ref, tol, "step")
    ref = O.step(ref, p)
    s = step.solver.step(s)
    check(s, ref, tol if dtype == "float64" else {f: 2 * t for f, t in tol.items()}, "2 steps")


@pytest.mark.parametrize("nz,nc", [(3, 2), (3, 4), (5, 3)])
def test_convective_kernels_on_few_levels(nz, nc):
    """Few levels leave some of the four level segments of the convective
    gate and sub-cycle kernels one level long or empty."""
    grid, physics, dt, kwargs, T, S = land_case(nz=nz)
    min_dz = float(np.min(-np.diff(grid.z)))
    physics = replace(physics, kappa_conv=(nc - 0.5) * 0.4 * min_dz ** 2 / dt)
    print(f"case land nz={nz} conv_nsub={nc} float64", flush=True)
    p = oracle_params(grid, physics, dt, kwargs)
    assert p.conv_nsub == nc
    step, _, _ = make_solver_global(grid, physics, dt, dtype="float64", mode_split=True, **kwargs)
    s0 = perturbed_state(p, T, S)
    assert O.convective_mask(s0.T, s0.S, p).any()
    check(step.solver.step(JaxStateG(*s0)), O.step(s0, p), TOL64, "step")


@pytest.mark.parametrize("flags", [{}, {"fct_adv": True}, {"monotone_adv": True},
                                   {"localize_conv": True, "conservative_kv": True}],
                         ids=["centred", "fct", "monotone", "localized-conservative"])
@pytest.mark.parametrize("dtype,tol", [("float64", 1e-11), ("float32", 1e-4)])
def test_terms_fn_matches_oracle(flags, dtype, tol):
    """terms_fn (_tracer_terms, tendencies.py:74-92) against the oracle: each of
    the advection, horizontal and vertical diffusion and convection planes
    relative to its own maximum; GM and Redi are zero.  float32 is compared
    with the oracle on the same float32-rounded state."""
    print(f"case land terms_fn {flags} {dtype}", flush=True)
    grid, physics, dt, kwargs, T, S = land_case()
    kw = dict(kwargs, **flags)
    step, _, _, params, terms_fn = make_solver_global(grid, physics, dt, return_params=True, dtype=dtype,
                                                      mode_split=True, **kw)
    p = oracle_params(grid, physics, dt, kw)
    s0 = perturbed_state(p, T, S)
    if dtype == "float32":
        s0 = O.State(*(np.asarray(a, dtype=np.float32).astype(np.float64) for a in s0))
    got = terms_fn(JaxStateG(*s0))
    torch.cuda.synchronize()
    got = got.double().cpu().numpy()
    ref = O.tracer_terms(s0.u, s0.v, s0.T, s0.S, p)
    assert got.shape == ref.shape
    for q, name in enumerate(("adv", "diff_h", "diff_v", "conv")):
        assert np.max(np.abs(ref[q])) > 0.0, name
        e = rel(got[q], ref[q])
        print(f"  {name}: {e:.1e}", flush=True)
        assert e <= tol, (name, e)
    assert not got[4:].any()
