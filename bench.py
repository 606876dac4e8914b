"""Benchmark of the mode-split step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--case g05deg] [--dtype float32]

Workload: the 0.5-degree plain FP32 step (720 x 320 x 50, dt = 900 s,
45 barotropic and 45 nu_h substeps, seasonal wind via step_dyn), synthetic
seeded inputs resident in HBM.  One "step" is one baroclinic step.  Prints one
JSON line: SYPD of the whole step, the HBM roofline of the dominant kernel
(the nu_h Laplacian substep) and a CPU baseline from the NumPy oracle.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SYPD at 0.5deg x 50L on 1 GPU: learned+mixed vs plain FP32; % HBM peak"
HBM_PEAK_GBS = 8000.0
VALU_SIMDS = 1024  # 256 CUs x 4 SIMDs
VALU_CLOCK_HZ = 2.4e9


def algorithmic_step_bytes(nx, ny, nz, n_nu, n_sub, word):
    """B_alg = (8 n_nu + 86) F3 + 12 n_sub F2 (compulsory traffic of one step)."""
    f3 = nx * ny * nz * word
    f2 = nx * ny * word
    return (8 * n_nu + 86) * f3 + 12 * n_sub * f2


def read_profile_value(name, key):
    """A ``key: value`` number from a committed rocprofv3 summary under profiles/, if present."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def pmc_summary(fetch_csv, write_csv, nsteps):
    """Markdown summary of two rocprofv3 --pmc runs of ``bench.py --profile-only``
    (one with FETCH_SIZE, one with WRITE_SIZE): the counters of the 1 GiB copy
    calibrate KiB per dispatch to bytes, and every zm:: kernel's dispatches are
    summed and divided by the number of steps the run made."""
    import csv
    from collections import OrderedDict

    def per_kernel(path):
        out = OrderedDict()
        for row in csv.DictReader(open(path)):
            out.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        return out

    fetch, write = per_kernel(fetch_csv), per_kernel(write_csv)
    copy_name = next(k for k in fetch if "copyBuffer" in k)
    cf = [x for x in fetch[copy_name] if x > 1e5]
    cw = [x for x in write[copy_name] if x > 1e5]
    rf, rw = 1048576.0 / (sum(cf) / len(cf)), 1048576.0 / (sum(cw) / len(cw))
    lines = [f"Calibration: the 1 GiB device copy ({len(cf)} dispatches) reads FETCH_SIZE {sum(cf) / len(cf):.1f} KiB "
             f"-> read factor {rf:.4f}; WRITE_SIZE {sum(cw) / len(cw):.1f} KiB -> write factor {rw:.4f}.", "",
             "| kernel | dispatches per step | FETCH KiB/dispatch | WRITE KiB/dispatch | MB read/step | MB written/step |",
             "|---|---:|---:|---:|---:|---:|"]
    rows, total_r, total_w = [], 0.0, 0.0
    for name, vals in fetch.items():
        if "zm::" not in name:
            continue
        wv = write[name]
        r = sum(vals) * rf * 1024.0 / nsteps
        w = sum(wv) * rw * 1024.0 / nsteps
        total_r, total_w = total_r + r, total_w + w
        short = name.split("(")[0].replace("void ", "")
        rows.append((r + w, f"| `{short}` | {len(vals) / nsteps:g} | {sum(vals) / len(vals):.1f} | "
                            f"{sum(wv) / len(wv):.1f} | {r / 1e6:.1f} | {w / 1e6:.1f} |"))
    lines += [row for _, row in sorted(rows, reverse=True)]
    lines += ["", f"Per step: {total_r / 1e6:.1f} MB read, {total_w / 1e6:.1f} MB written.",
              f"traffic_bytes_per_step: {total_r + total_w:.0f}"]
    return "\n".join(lines)


def cpu_baseline():
    """The oracle (NumPy float64, single-threaded) timed directly on this host.

    * the 4-degree case (90x40x12, dt 600 s): median of 3 steps after a warm-up step;
    * the 1-degree case (360x160x50, dt 1200 s): one step after a warm-up step;
    * the 0.5-degree case of the benchmark (720x320x50, dt 900 s): one step
      (NumPy is warm from the steps above).  ``value`` is the SYPD of this step.
    """
    from oracle import zm_oracle as O
    from zhenmode_amd.cases import make_case

    def oracle(name):
        case = make_case(name)
        kw = {k: v for k, v in case.solver_kwargs.items() if k not in ("mode_split", "use_scan")}
        p = O.make_params(case.grid, case.physics, case.dt, **kw)
        return case, p, O.init_state(p, case.T_init, case.S_init)

    def timed(p, s):
        t0 = time.perf_counter()
        s = O.step(s, p)
        return time.perf_counter() - t0, s

    direct = {}
    case, p, s = oracle("g4deg")
    _, s = timed(p, s)
    runs = []
    for _ in range(3):
        t, s = timed(p, s)
        runs.append(t)
    t4 = float(np.median(runs))
    direct["g4deg_90x40x12"] = {"s_per_step": t4, "sypd": case.dt / (365.0 * t4), "rule": "median of 3 after 1 warm-up"}
    case, p, s = oracle("g1deg")
    _, s = timed(p, s)
    t1, s = timed(p, s)
    direct["g1deg_360x160x50"] = {"s_per_step": t1, "sypd": case.dt / (365.0 * t1), "rule": "1 step after 1 warm-up"}
    case, p, s = oracle("g05deg")
    t05, _ = timed(p, s)
    direct["g05deg_720x320x50"] = {"s_per_step": t05, "sypd": case.dt / (365.0 * t05), "rule": "1 step"}
    return {"value": case.dt / (365.0 * t05), "unit": "SYPD", "cores": 1, "kind": "port",
            "host_cores_available": len(os.sched_getaffinity(0)),
            "sample": (f"one step of the oracle (NumPy float64 restatement of the reference step, single-threaded) "
                       f"on the benchmark's own 720x320x50 workload took {t05:.1f} s (dt {case.dt:g} s); direct "
                       f"timings at 90x40x12 and 360x160x50 in 'direct'"),
            "direct": direct}


def timed_run(step_fn, steps):
    """``steps`` calls of ``step_fn`` in one run: mean wall time per step
    (host clock around the run, synchronised on both sides) and the median of
    per-step hipEvent times on torch's current stream."""
    import torch

    events = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    events[0].record()
    for k in range(steps):
        step_fn()
        events[k + 1].record()
    torch.cuda.synchronize()
    wall = 1e3 * (time.perf_counter() - t0) / steps
    per = [events[k].elapsed_time(events[k + 1]) for k in range(steps)]
    return wall, float(np.median(per))


def run_loop_leg(solver, state, steps, check_every):
    """runtime.run (device monitor after every step, host read every
    check_every steps) against the same number of bare steps between two
    preallocated states (frozen forcing for both)."""
    import torch

    from zhenmode_amd.runtime import run

    a, b = solver.new_state(), solver.new_state()
    for x, y in zip(a, solver.coerce(state)):
        x.copy_(y)
    box = [a, b]

    def bare():
        solver.step(box[0], out=box[1])
        box[0], box[1] = box[1], box[0]

    for _ in range(3):
        bare()
    bare_ms, _ = timed_run(bare, steps)
    run(lambda s, out: solver.step(s, out=out), solver, box[0], check_every, check_every=check_every)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = run(lambda s, out: solver.step(s, out=out), solver, box[0], steps, check_every=check_every)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    return {"ms_per_step": ms, "bare_ms_per_step": bare_ms, "overhead_ms_per_step": ms - bare_ms,
            "check_every": check_every, "accepted_steps": res.accepted_steps, "failure": res.status["failure"]}


# the surrogate's training recipe and residual tolerances (the 30-day 0.45-degree
# evaluation of this recipe: 88 % of columns accepted, SST RMSE 0.031 degC)
TRAIN_COLUMNS, TRAIN_EPOCHS = 6_000_000, 30
TOL_T, TOL_S = 2e-3, 4e-4


def learned_leg(case, plain_ms, steps, warmup):
    """The learned+mixed step: the plain FP32 step without explicit tracer
    vertical mixing, followed by the BF16 MFMA column surrogate (FP64
    column-integral correction, FP64 Thomas fallback for rejected columns).
    The surrogate is trained here on columns of the same generator (outside
    the timed region)."""
    import torch

    from zhenmode_amd.learned import ColumnMLP, MixingStep
    from zhenmode_amd.surrogate import generate_columns, normalise, train

    t0 = time.perf_counter()
    X, Y = generate_columns(case, TRAIN_COLUMNS, spinup=48, every=8, log=lambda m: None)
    mlp = ColumnMLP(case.grid.nz, device=X.device)
    Xn, Yn = normalise(mlp, X, Y)
    del X, Y
    losses = train(mlp, Xn, Yn, epochs=TRAIN_EPOCHS, lr=2e-3, decay_every=8, log=lambda m: None)
    del Xn, Yn
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0

    def timed(stepper):
        # the plain step's rule: one run of `steps` steps, mean wall time per step
        box = [stepper.init_state(case.T_init, case.S_init)]
        for _ in range(warmup):
            box[0] = stepper(box[0])
        if stepper.mixer is not None:
            stepper.mixer.count.zero_()

        def one():
            box[0] = stepper(box[0])

        wall, _ = timed_run(one, steps)
        return wall, box[0]

    kw = case.solver_kwargs
    learned = MixingStep(case.grid, case.physics, case.dt, mode="learned", mlp=mlp, tol_T=TOL_T, tol_S=TOL_S,
                         dtype="float32", **kw)
    ms_l, st = timed(learned)
    # the fused surrogate launch (features, 4 BF16 MFMA layers, apply) over every column
    import ctypes as _ct
    from zhenmode_amd.learned import _p
    mx = learned.mixer
    sol = learned.solver
    src = sol.coerce(st)
    To, So = torch.empty_like(src.T), torch.empty_like(src.S)

    def fused():
        m = mx.mlp
        rc = sol.lib.zm_mlp_mix(sol.handle, _ct.byref(sol.zm_state(src)), None, None, None, _p(m.in_mean),
                                _p(m.in_inv_std), mx._wptrs, mx._bptrs, _p(m.out_mean), _p(m.out_std), case.dt,
                                mx.kappa_v, mx.kappa_conv, mx.tol_T, mx.tol_S, _p(To), _p(So), _p(mx.flags),
                                _p(mx.count), sol._stream())
        assert rc == 0

    fused()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        fused()
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / 10
    sizes = mlp.sizes
    rows = case.grid.nx * case.grid.ny
    flops = 2.0 * rows * sum(k * n for k, n in zip(sizes[:-1], sizes[1:]))
    wet_cols = int((case.grid.wet_mask > 0.5).sum())
    fallback = learned.fallback_count() / max(1, steps * wet_cols)
    thomas = MixingStep(case.grid, case.physics, case.dt, mode="thomas", dtype="float32", **kw)
    ms_t, _ = timed(thomas)
    return {"sypd": case.dt / (365.0 * ms_l * 1e-3), "ms_per_step": ms_l, "vs_plain_fp32": plain_ms / ms_l,
            "fallback_fraction": fallback, "tol_T_degC": TOL_T, "tol_S_psu": TOL_S,
            "thomas_mixing_ms_per_step": ms_t, "vs_thomas_mixing": ms_t / ms_l,
            "train_columns": TRAIN_COLUMNS, "train_epochs": TRAIN_EPOCHS,
            "surrogate_launch": {"kernel": "k_mlp_fused (features, 4 BF16 MFMA layers, apply + residual check)",
                                 "rows": rows, "ms": fwd_ms, "tflops": flops / (fwd_ms * 1e-3) / 1e12,
                                 "bound": "mfma", "peak_tflops_bf16_dense": 2500.0,
                                 "frac": flops / (fwd_ms * 1e-3) / 1e12 / 2500.0},
            "final_loss": losses[-1], "train_s": train_s,
            "note": "frozen forcing; the learned step replaces the explicit tracer vertical diffusion and "
                    "convection of the plain step by the implicit mixing operator's surrogate (explicit mixing "
                    "increment plus a learned correction scaled by the column's mixing scale, Thomas fallback)"}


def projection_leg(case, plain_ms, tiles, steps, warmup, dtype):
    """The same step with project_adv_vel (reference defaults: 150 CG
    iterations per solve, up to two transport-checked refinements): time per
    step, CG iterations, and the algorithmic bytes of one CG iteration
    (k_proj_A: nine float64 taps, r, M, p_old read, p_new and Ap written = 112
    B per column; k_proj_B: x, p, r, Ap, M read, x and r written = 56 B)."""
    import torch

    from zhenmode_amd.solver import make_solver_global

    step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype=dtype, project_adv_vel=True,
                                             **case.solver_kwargs)
    sol = step.solver
    sol.set_tiles(*tiles[:2])
    a = init_state(case.T_init, case.S_init)
    b = sol.new_state()
    for _ in range(warmup):
        sol.step(a, out=b)
        a, b = b, a
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sol.step(a, out=b)
        a, b = b, a
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    its = sol.projection_iterations()
    n2 = case.grid.nx * case.grid.ny
    per_it_us = 1e3 * (ms - plain_ms) / max(sum(its), 1)
    return {"ms_per_step": ms, "plain_ms_per_step": plain_ms, "cg_iterations_last_step": its,
            "us_per_cg_iteration": per_it_us, "cg_bytes_per_iteration": 168 * n2,
            "cg_GBps": 168 * n2 / (per_it_us * 1e-6) / 1e9,
            "sypd": case.dt / (365.0 * ms * 1e-3),
            "finite": bool(torch.isfinite(a.T).all())}


def ensemble_leg(steps, warmup):
    """24 uneven members (8 x 180x80x30, 8 x 360x160x50, 8 x 480x214x40) in one
    hipGraph launch per ensemble step; the launch shapes of each member size
    (nu_h substeps per launch, column-kernel levels per thread) are picked by
    the on-device autotuner (zm_autotune, eagerly issued steps) on the first
    member of that size and set on the others; results do not depend on them."""
    import torch

    from zhenmode_amd.cases import ensemble_members
    from zhenmode_amd.solver import Ensemble, make_solver_global

    members = ensemble_members()
    sols, states = [], []
    for m in members:
        step, init_state, _ = make_solver_global(m.grid, m.physics, m.dt, dtype="float32", **m.solver_kwargs)
        sols.append(step.solver)
        states.append(init_state(m.T_init, m.S_init))
    tiles = {}
    for sol, st in zip(sols, states):
        if sol.shape3 not in tiles:
            nu, seg, ms = sol.autotune(st, reps=3)
            tiles[sol.shape3] = (nu, seg, ms)
        sol.set_tiles(*tiles[sol.shape3][:2])
    ens = Ensemble(sols)
    # two preallocated state sets stepped into each other, so that the timed
    # steps replay the two captured graphs (one per pointer set)
    bufs = [[s.new_state() for s in sols] for _ in range(2)]
    a = ens.step(states, out=bufs[0])
    for w in range(max(warmup, 2)):
        a = ens.step(a, out=bufs[(w + 1) % 2])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(steps):
        a = ens.step(a, out=bufs[(n + 1 + max(warmup, 2)) % 2])
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    b = [s.step(x, out=y) for s, x, y in zip(sols, states, bufs[0])]
    for w in range(max(warmup, 2)):
        b = [s.step(x, out=y) for s, x, y in zip(sols, b, bufs[(w + 1) % 2])]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(steps):
        b = [s.step(x, out=y) for s, x, y in zip(sols, b, bufs[(n + 1 + max(warmup, 2)) % 2])]
    torch.cuda.synchronize()
    ms_sep = 1e3 * (time.perf_counter() - t0) / steps
    dt = members[0].dt
    cells = sum(int(np.prod(s.shape3)) for s in sols)
    finite = all(bool(torch.isfinite(x.T).all()) for x in a)
    return {"members": len(sols), "cells": cells, "ms_per_ensemble_step": ms,
            "member_sim_days_per_s": len(sols) * dt / 86400.0 / (ms * 1e-3),
            "aggregate_sypd": len(sols) * dt / (365.0 * ms * 1e-3),
            "ms_per_step_members_separately": ms_sep, "batching_speedup": ms_sep / ms,
            "tiles_autotuned": {f"{k[0]}x{k[1]}x{k[2]}": {"nu_block": v[0], "seg": v[1], "eager_ms": v[2]}
                                for k, v in tiles.items()},
            "finite": finite}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--case", default="g05deg")
    ap.add_argument("--dtype", default="float32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-projection", action="store_true")
    ap.add_argument("--tiles", default="4,50", help="nu_block,seg: launch shapes (nu_h substeps per launch, "
                    "column-kernel levels per thread)")
    ap.add_argument("--profile-only", action="store_true",
                    help="counter runs: only the warm-up and timed steps and the 1 GiB copy, then exit")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "NSTEPS"),
                    help="print the per-step HBM traffic summary of two --profile-only counter runs and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        print(pmc_summary(args.pmc_summary[0], args.pmc_summary[1], int(args.pmc_summary[2])))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU decomposition is a design "
                                   "(latitude bands + RCCL halo exchange) and is not launched", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from zhenmode_amd.cases import make_case
    from zhenmode_amd.solver import make_solver_global

    rank = int(os.environ.get("RANK", "0"))
    case = make_case(args.case)
    g = case.grid
    kwargs = dict(case.solver_kwargs)
    step, init_state, _, params, _, step_dyn = make_solver_global(
        g, case.physics, case.dt, return_params=True, dynamic_forcing=True, dtype=args.dtype, **kwargs)
    solver = step.solver
    dev = solver.device
    tdtype = solver.dtype
    a = init_state(case.T_init, case.S_init)
    b = solver.new_state()
    # launch shapes (results do not depend on them); 4 substeps per launch and
    # whole 50-level columns per thread are the fastest at 0.5 degrees
    nu_b, seg = (int(x) for x in args.tiles.split(","))
    solver.set_tiles(nu_b, min(seg, g.nz))
    tiles = (nu_b, min(seg, g.nz), None)
    # on-device seasonal wind: tau_x(t) = (0.1 + 0.05 sin(2 pi t / 365 d)) cos(3 lat)
    from zhenmode_amd.solver import empty2
    base = torch.as_tensor(np.cos(3.0 * np.radians(g.lat)), dtype=tdtype, device=dev)[:, None].expand(g.ny, g.nx).contiguous()
    tau_x = empty2(g.nx, g.ny, tdtype, dev)
    tau_store = tau_x.permute(1, 0)  # contiguous (ny, nx) storage
    tau_y = empty2(g.nx, g.ny, tdtype, dev)
    q = empty2(g.nx, g.ny, tdtype, dev)
    tau_y.zero_()
    q.zero_()

    def one(t_index, src, dst):
        t = t_index * case.dt
        amp = 0.1 + (0.05 * np.sin(2.0 * np.pi * t / (365.0 * 86400.0)) if case.seasonal_wind else 0.0)
        torch.mul(base, amp, out=tau_store)
        return solver.step_dyn(src, tau_x, tau_y, q, out=dst)

    box = [a, b, 0]

    def advance():
        one(box[2], box[0], box[1])
        box[0], box[1], box[2] = box[1], box[0], box[2] + 1

    for _ in range(args.warmup):
        advance()
    ms_per_step, ev_ms = timed_run(advance, args.steps)
    a = box[0]
    finite = bool(torch.isfinite(a.u).all() and torch.isfinite(a.T).all() and torch.isfinite(a.eta).all())

    # STREAM-like copy of 1 GiB between two fresh buffers: the measured HBM
    # reference, and the known-size kernel that calibrates the FETCH_SIZE /
    # WRITE_SIZE counters of a --pmc run of this script
    copy_src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    copy_dst = torch.empty_like(copy_src)
    torch.cuda.synchronize()
    ce0, ce1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    copy_dst.copy_(copy_src)
    ce0.record()
    for _ in range(5):
        copy_dst.copy_(copy_src)
    ce1.record()
    torch.cuda.synchronize()
    copy_ms = ce0.elapsed_time(ce1) / 5
    copy_gbps = 2.0 * copy_src.numel() * 4 / (copy_ms * 1e-3) / 1e9
    del copy_src, copy_dst
    if args.profile_only:
        print(json.dumps({"profile_only": True, "steps": args.warmup + args.steps, "ms_per_step": ms_per_step,
                          "copy_1GiB_GBps": copy_gbps, "finite": finite}), flush=True)
        return

    word = 4 if args.dtype == "float32" else 8
    n3 = g.nx * g.ny * g.nz
    # dominant kernel: the fused nu_h sub-cycle launch (h substeps per HBM pass).
    # Its algorithmic bytes per launch are the compulsory traffic of that fused
    # pass: read u, v once and write them once (4 n3 words); the unfused
    # equivalent (h passes) is reported beside it, labelled as such.
    nu_h = solver.nu_substeps_per_launch()
    nu_ms = solver.time_kernel(5 if nu_h > 1 else 0, 50)
    nu_bytes = 4 * n3 * word
    nu_single_ms = solver.time_kernel(0, 50)
    achieved = nu_bytes / (nu_ms * 1e-3) / 1e9
    profiled = args.case == "g05deg" and args.dtype == "float32"
    traffic = read_profile_value(f"nu_stream_{args.case}_pmc.txt", "traffic_bytes_per_launch") if profiled else None
    step_bytes = algorithmic_step_bytes(g.nx, g.ny, g.nz, params.n_nu, params.n_subcyc, word)
    step_traffic = (read_profile_value(f"step_{args.case}_hbm_traffic.md", "traffic_bytes_per_step")
                    if profiled else None)
    # vector issue: the launch's VALU wave64 instructions (SQ_INSTS_VALU) at one
    # instruction per SIMD every 4 cycles, 1024 SIMDs at 2.4 GHz
    valu = (read_profile_value(f"step_{args.case}_sq_counters.md", "valu_instructions_per_launch")
            if profiled and nu_h == 4 else None)
    valu_ms = valu * 4 / (VALU_SIMDS * VALU_CLOCK_HZ) * 1e3 if valu else None
    sypd = case.dt / (365.0 * ms_per_step * 1e-3)
    out = {
        "metric": METRIC, "value": sypd, "unit": "SYPD", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if word == 4 else "f64", "data": "synthetic",
        "config": {"workload": f"{args.case}: plain mode-split step {g.nx}x{g.ny}x{g.nz} "
                               f"(0.5 deg), dt={case.dt:g}s, n_sub={params.n_subcyc}, n_nu={params.n_nu}, "
                               f"seasonal wind via step_dyn, {args.dtype}",
                   "sim_days_per_s": case.dt / 86400.0 / (ms_per_step * 1e-3),
                   "median_event_ms_per_step": ev_ms, "finite": finite,
                   "tiles": {"nu_block": tiles[0], "seg": tiles[1]}},
        "roofline": {"bound": "hbm",
                     "kernel": (f"k_nu_stream ({nu_h} nu_h Laplacian substeps of u and v per HBM pass)" if nu_h > 1
                                else "k_nu_substep (one nu_h Laplacian substep of u and v)"),
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": nu_bytes, "launch_ms": nu_ms,
                     "measured_GBps": traffic / (nu_ms * 1e-3) / 1e9 if traffic else None,
                     "measured_frac": traffic / (nu_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "equivalent_unfused_GBps": nu_h * nu_bytes / (nu_ms * 1e-3) / 1e9,
                     "single_substep_launch_ms": nu_single_ms,
                     "single_substep_GBps": nu_bytes / (nu_single_ms * 1e-3) / 1e9,
                     "copy_1GiB_GBps": copy_gbps,
                     "valu_issue": ({"instructions_per_launch": valu, "issue_ms": valu_ms,
                                     "frac": valu_ms / nu_ms,
                                     "rate": "one wave64 VALU instruction per SIMD every 4 cycles, "
                                             "1024 SIMDs at 2.4 GHz"} if valu else None)},
        "step_roofline": {"algorithmic_bytes": step_bytes,
                          "algorithmic_GBps": step_bytes / (ms_per_step * 1e-3) / 1e9,
                          "algorithmic_frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "measured_traffic_bytes": step_traffic,
                          "measured_GBps": step_traffic / (ms_per_step * 1e-3) / 1e9 if step_traffic else None,
                          "measured_frac": (step_traffic / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS
                                            if step_traffic else None),
                          "note": "algorithmic bytes count every nu_h substep as a separate HBM pass; the "
                                  "fused kernel makes one pass per 4 substeps, so measured_frac (rocprofv3 "
                                  "FETCH_SIZE + WRITE_SIZE summed over the step's kernels) is the HBM fraction"},
    }
    # barotropic sub-cycle: one persistent launch of n_sub substeps against n_sub launches of one substep
    n_sub = int(params.n_subcyc)
    sub_ms = solver.time_kernel(4, 50)
    bt = {"n_sub": n_sub, "per_substep_launch_ms": sub_ms, "per_substep_launches_ms": n_sub * sub_ms,
          "algorithmic_bytes_per_substep": 12 * g.nx * g.ny * word}
    try:
        bt["persistent_ms"] = solver.time_kernel(15, 20)
        bt["speedup"] = n_sub * sub_ms / bt["persistent_ms"]
        bt["timeouts"] = solver.barotropic_timeouts()
    except ValueError as e:   # the grid's blocks would not all be resident on this device
        bt["persistent_ms"] = None
        bt["note"] = str(e)
    out["barotropic_subcycle"] = bt
    out["run_loop"] = run_loop_leg(solver, a, min(args.steps, 40), check_every=10)
    if not args.no_learned:
        out["learned_mixed"] = learned_leg(case, ms_per_step, min(args.steps, 20), 3)
    if not args.no_projection:
        out["projected_step"] = projection_leg(case, ms_per_step, solver.get_tiles(), min(args.steps, 10), 2,
                                               args.dtype)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(min(args.steps, 10), 2)
    if rank == 0 and not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
