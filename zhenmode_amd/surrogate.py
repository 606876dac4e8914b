"""Training and evaluation of the learned vertical-mixing surrogate.

    python -m zhenmode_amd.surrogate [--train-case g05deg] [--case g045deg] [--columns 2000000]
                                     [--epochs 8] [--days 30] [--out surrogate.pt] [--report report.json]

1. Columns: the plain model with the implicit Thomas mixing (``MixingStep``,
   float32) is spun up from the case's initial state; every ``every`` steps
   the state that enters the mixing step is sampled: its wet columns' features
   (``zm_column_features``) and the Thomas solution's departure from the
   explicit mixing step in units of the column's mixing scale
   (``zm_mix_targets``) are the training pairs.  ``--train-case`` picks the
   generator the columns come from (same 50 levels as the evaluation case).
2. Training: minibatch Adam on the project's own BF16 MFMA forward/backward
   kernels (``ColumnMLP.train_step``).
3. Evaluation (``--days``): the learned step in float32 and the Thomas step in
   float64 run side by side from the same initial state; the report holds the
   SST and SSS RMSE between them, the fallback fraction, and the change of the
   global heat and salt content relative to the float64 run.
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from .cases import make_case
from .learned import N_OUT_PAD, ColumnMLP, LearnedMixer, MixingStep, mixing_targets


def generate_columns(case, n_columns, dtype="float32", spinup=48, every=4, seed=0, log=print):
    """Returns (X_raw float32 [N, 256], dX float32 [N, 2 nz]) on the GPU."""
    stepper = MixingStep(case.grid, case.physics, case.dt, mode="thomas", dtype=dtype, **case.solver_kwargs)
    sol = stepper.solver
    probe = ColumnMLP(case.grid.nz, device=sol.device, seed=seed)  # identity normalisation
    feats = LearnedMixer(sol, probe, stepper.kappa_v, stepper.kappa_conv)
    wet = torch.as_tensor(case.grid.wet_mask.T.reshape(-1) > 0.5, device=sol.device)
    n2 = wet.numel()
    s = stepper.init_state(case.T_init, case.S_init)
    for _ in range(spinup):
        s = stepper(s)
    Xs, Ys, n = [], [], 0
    while n < n_columns:
        for _ in range(every - 1):
            s = stepper(s)
        s1 = sol.step(s)
        X = feats.features(s1)[:n2][wet].float()
        Y = mixing_targets(sol, s1, case.dt, stepper.kappa_v, stepper.kappa_conv)[wet]
        Xs.append(X)
        Ys.append(Y)
        n += X.shape[0]
        s = stepper.mix(s1)
    X, Y = torch.cat(Xs)[:n_columns], torch.cat(Ys)[:n_columns]
    log(f"generated {X.shape[0]} columns from {len(Xs)} snapshots")
    return X, Y


def normalise(mlp, X_raw, dX):
    mlp.fit_normalisation(X_raw, dX)
    Xn = ((X_raw - mlp.in_mean) * mlp.in_inv_std).to(torch.bfloat16)
    Yn = torch.zeros((dX.shape[0], N_OUT_PAD), dtype=torch.float32, device=dX.device)
    Yn[:, :mlp.n_out] = (dX - mlp.out_mean[:mlp.n_out]) / mlp.out_std[:mlp.n_out]
    return Xn, Yn


def train(mlp, Xn, Yn, epochs=8, batch=4096, lr=1e-3, seed=0, decay_every=3, log=print):
    """Minibatch Adam, learning rate halved every ``decay_every`` epochs;
    returns the per-epoch mean losses."""
    n = (Xn.shape[0] // batch) * batch
    g = torch.Generator(device="cpu").manual_seed(seed)
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(Xn.shape[0], generator=g)[:n].to(Xn.device)
        total = torch.zeros((), dtype=torch.float32, device=Xn.device)
        step_lr = lr * (0.5 ** (ep // decay_every))
        for b in range(0, n, batch):
            idx = perm[b:b + batch]
            total += mlp.train_step(Xn[idx].contiguous(), Yn[idx].contiguous(), lr=step_lr)[0]
        losses.append(float(total.item()) / (n // batch))
        log(f"epoch {ep}: loss {losses[-1]:.4e}")
    return losses


def heat_salt(state, grid, params):
    """Global heat (J, rho0 cp T dV) and salt (psu m^3) content in float64."""
    from .config import C_P, RHO_0
    T = state.T.double()
    S = state.S.double()
    dz = torch.as_tensor(np.asarray(params.dz_node), dtype=torch.float64, device=T.device)
    area = torch.as_tensor(grid.dx_2d * grid.dy, dtype=torch.float64, device=T.device)
    wet = torch.as_tensor(grid.wet_mask_3d, dtype=torch.float64, device=T.device)
    vol = area[:, :, None] * dz[None, None, :] * wet
    return float((T * vol).sum() * RHO_0 * C_P), float((S * vol).sum())


def _time_steps(stepper, state, n=20):
    for _ in range(3):
        state = stepper(state)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        state = stepper(state)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / n * 1e3


def evaluate(case, mlp, days, tols=((1e-4, 2e-5), (1e-3, 2e-4)), untrained=None, log=print):
    """Side-by-side runs from one initial state against the float64 Thomas step.

    Variants: the float32 Thomas step (the float32 floor), the learned step at
    each (tol_T, tol_S), and, when given, an untrained network at the first
    tolerance (what training gains).  Returns per-variant SST/SSS RMSE against
    the float64 run, fallback fraction, global heat/salt difference and ms/step.
    """
    steps = int(round(days * 86400.0 / case.dt))
    kw = case.solver_kwargs
    ref = MixingStep(case.grid, case.physics, case.dt, mode="thomas", dtype="float64", **kw)
    variants = {"thomas_f32": MixingStep(case.grid, case.physics, case.dt, mode="thomas", dtype="float32", **kw)}
    for tT, tS in tols:
        variants[f"learned_tolT{tT:g}"] = MixingStep(case.grid, case.physics, case.dt, mode="learned", mlp=mlp,
                                                     tol_T=tT, tol_S=tS, dtype="float32", **kw)
    if untrained is not None:
        tT, tS = tols[0]
        variants[f"untrained_tolT{tT:g}"] = MixingStep(case.grid, case.physics, case.dt, mode="learned",
                                                        mlp=untrained, tol_T=tT, tol_S=tS, dtype="float32", **kw)
    sr = ref.init_state(case.T_init, case.S_init)
    states = {k: v.init_state(case.T_init, case.S_init) for k, v in variants.items()}
    wet = torch.as_tensor(case.grid.wet_mask > 0.5, device=sr.T.device)
    n2w = int(wet.sum().item())
    t0 = time.perf_counter()
    for n in range(steps):
        sr = ref(sr)
        for k, v in variants.items():
            states[k] = v(states[k])
        if (n + 1) % max(1, steps // 10) == 0:
            log(f"step {n + 1}/{steps}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    hr, Sr = heat_salt(sr, case.grid, ref.solver.params)
    out = {"days": days, "steps": steps, "wall_s": wall, "reference": "thomas_f64",
           "ms_per_step": {"thomas_f64": _time_steps(ref, sr)}}
    for k, v in variants.items():
        st = states[k]
        sst = (st.T[:, :, 0].double() - sr.T[:, :, 0])[wet]
        sss = (st.S[:, :, 0].double() - sr.S[:, :, 0])[wet]
        h, S = heat_salt(st, case.grid, v.solver.params)
        out[k] = {"sst_rmse_degC": float(torch.sqrt((sst ** 2).mean())),
                  "sss_rmse_psu": float(torch.sqrt((sss ** 2).mean())),
                  "fallback_fraction": v.fallback_count() / max(1, steps * n2w),
                  "heat_rel_diff": (h - hr) / abs(hr), "salt_rel_diff": (S - Sr) / abs(Sr),
                  "finite": bool(torch.isfinite(st.T).all())}
        out["ms_per_step"][k] = _time_steps(v, st)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="g045deg", help="evaluation case")
    ap.add_argument("--train-case", default="g05deg", help="case whose plain runs generate the training columns")
    ap.add_argument("--tols", default="1e-4,2e-5;1e-3,2e-4", help="(tol_T,tol_S) pairs of the learned variants")
    ap.add_argument("--columns", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--decay-every", type=int, default=3, help="epochs per halving of the learning rate")
    ap.add_argument("--days", type=float, default=0.0)
    ap.add_argument("--spinup", type=int, default=48)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--out", default="")
    ap.add_argument("--load", default="", help="evaluate a network saved with --out instead of training one")
    ap.add_argument("--no-untrained", action="store_true", help="leave out the untrained-network variant")
    ap.add_argument("--report", default="")
    args = ap.parse_args()
    log = lambda m: print(m, flush=True)  # noqa: E731
    case = make_case(args.case)
    tcase = make_case(args.train_case)
    if tcase.grid.nz != case.grid.nz or not np.allclose(tcase.grid.z, case.grid.z):
        raise SystemExit("the training and evaluation cases must share their levels")
    tols = tuple(tuple(float(x) for x in pair.split(",")) for pair in args.tols.split(";"))
    t0 = time.perf_counter()
    if args.load:
        mlp = ColumnMLP(case.grid.nz, device=torch.device("cuda", torch.cuda.current_device()))
        mlp.load_state_dict(torch.load(args.load, weights_only=True))
        report = {"case": args.case, "loaded": args.load}
    else:
        X, Y = generate_columns(tcase, args.columns, spinup=args.spinup, every=args.every, log=log)
        t1 = time.perf_counter()
        mlp = ColumnMLP(case.grid.nz, device=X.device)
        Xn, Yn = normalise(mlp, X, Y)
        del X
        losses = train(mlp, Xn, Yn, epochs=args.epochs, batch=args.batch, lr=args.lr, decay_every=args.decay_every,
                       log=log)
        del Xn, Yn
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        report = {"case": args.case, "train_case": args.train_case, "columns": args.columns,
                  "generate_s": t1 - t0, "train_s": t2 - t1,
                  "epochs": args.epochs, "lr": args.lr, "decay_every": args.decay_every, "losses": losses}
    if args.out:
        torch.save(mlp.state_dict(), args.out)
    if args.days > 0:
        untrained = None
        if not args.no_untrained:
            untrained = ColumnMLP(case.grid.nz, device=mlp.device, seed=1)
            for name in ("in_mean", "in_inv_std", "out_mean", "out_std"):
                getattr(untrained, name).copy_(getattr(mlp, name))
        report["evaluation"] = evaluate(case, mlp, args.days, tols=tols, untrained=untrained, log=log)
    log(json.dumps(report))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
