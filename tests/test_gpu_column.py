"""HIP column kernels vs the oracle on the GPU, through the C ABI: the Thomas
mixing solve, the surrogate's features, GEMM epilogues, gradients, training
targets, the apply step (integral correction, residual check, fallback), the
fused surrogate launch and training.

Tolerances: Thomas float64 <= 1e-12 relative (float64 elimination both sides);
Thomas with float32 I/O <= 4e-7 relative (one float32 rounding of the result);
GEMM float32 accumulation <= 1e-5 relative to the output scale; bfloat16
outputs within one bfloat16 ulp (2^-7 relative); MLP gradients against the
float64 oracle on the same bfloat16 weights and inputs <= 3e-2 of each
layer's gradient scale (bfloat16 activations and back-propagated errors);
training targets and the apply step in float64 <= 1e-11 of the state's scale
(float64 on both sides, different summation order); the fused launch equals
the unfused sequence bit for bit.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import zm_column as C  # noqa: E402
from oracle import zm_oracle as O  # noqa: E402
from tests.fixtures import land_case, oracle_params  # noqa: E402
from zhenmode_amd import _lib  # noqa: E402
from zhenmode_amd.learned import ColumnMLP, LearnedMixer, N_IN, N_OUT_PAD, mixing_targets, pad_rows  # noqa: E402
from zhenmode_amd.solver import JaxStateG, make_solver_global, to_numpy_state  # noqa: E402


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def land_setup(dtype):
    grid, physics, dt, kwargs, T, S = land_case()
    step, init_state, _, params, _ = make_solver_global(grid, physics, dt, return_params=True, dtype=dtype,
                                                        mode_split=True, **kwargs)
    p = oracle_params(grid, physics, dt, kwargs)
    s = O.init_state(p, T, S)
    rng = np.random.default_rng(11)
    # make a few columns statically unstable so the convective term is exercised
    Tu = s.T.copy()
    Tu[:, :, 1] += 3.0 * (rng.uniform(size=Tu.shape[:2]) > 0.6)
    s = s._replace(T=np.where(p.wet_mask_z > 0.5, Tu, s.T))
    return step.solver, p, s


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_thomas_mixing_matches_oracle(dtype):
    print(f"case land thomas {dtype}", flush=True)
    sol, p, s = land_setup(dtype)
    np_dt = np.float64 if dtype == "float64" else np.float32
    T = s.T.astype(np_dt).astype(np.float64)
    S = s.S.astype(np_dt).astype(np.float64)
    dt, kv, kc = 3600.0, 1e-3, 0.5
    out = sol.vertical_mix(JaxStateG(s.u, s.v, T, S, s.eta, s.ice), dt, kv, kc)
    Tn, Sn = C.implicit_mix(T, S, p, dt, kv, kc)
    got = to_numpy_state(out)
    tol = 1e-12 if dtype == "float64" else 4e-7
    assert rel(got.T, Tn) <= tol
    assert rel(got.S, Sn) <= tol
    if dtype == "float64":
        for Xn, X in ((got.T, T), (got.S, S)):
            i0, i1 = C.column_integral(X, p), C.column_integral(Xn, p)
            assert np.max(np.abs(i1 - i0)) <= 1e-12 * np.max(np.abs(i0))


def test_thomas_mixing_only_touches_flagged_columns():
    print("case land thomas masked float64", flush=True)
    sol, p, s = land_setup("float64")
    nx, ny = p.nx, p.ny
    rng = np.random.default_rng(5)
    flags = (rng.uniform(size=(ny, nx)) > 0.5).astype(np.int32)
    only = torch.as_tensor(flags.reshape(-1), device=sol.device)
    src = sol.coerce(JaxStateG(*s))
    To, So = torch.full_like(src.T, -7.0), torch.full_like(src.S, -7.0)
    sol.vertical_mix(src, 3600.0, 1e-3, 0.5, only=only, out=(To, So))
    Tn, _ = C.implicit_mix(s.T, s.S, p, 3600.0, 1e-3, 0.5)
    got = To.cpu().numpy()
    sel = flags.T.astype(bool)  # (nx, ny)
    assert rel(got[sel], Tn[sel]) <= 1e-12
    assert np.all(got[~sel] == -7.0)


def test_column_features_match_oracle():
    print("case land features float32", flush=True)
    grid, physics, dt, kwargs, T, S = land_case()
    step, _, _, params, _ = make_solver_global(grid, physics, dt, return_params=True, dtype="float32",
                                               mode_split=True, **kwargs)
    sol = step.solver
    p = oracle_params(grid, physics, dt, kwargs)
    from tests.fixtures import perturbed_state
    s = perturbed_state(p, T, S)
    s32 = JaxStateG(*(np.asarray(f, dtype=np.float32).astype(np.float64) for f in s))
    mlp = ColumnMLP(p.nz, device=sol.device)
    mixer = LearnedMixer(sol, mlp, 1e-3, 0.5)
    X = mixer.features(JaxStateG(*s32)).float().cpu().numpy()
    tau_x, tau_y, q = kwargs["forcing"]
    fw = p.restore_coef_S * (p.S_ref_2d - s32.S[:, :, 0]) * p.wet_mask
    tau = np.hypot(np.float32(tau_x).astype(np.float64), np.float32(tau_y).astype(np.float64)) * p.wet_mask
    ref = C.bf16_round(C.column_features(s32.T, s32.S, s32.u, s32.v, p, np.float32(q) * p.wet_mask, fw, tau))
    n2, nz = p.nx * p.ny, p.nz
    assert X.shape[0] == pad_rows(n2)
    # u, v exactly; T, S differences (one float32 subtraction before the
    # bfloat16 rounding) and N^2 within one bfloat16 ulp
    np.testing.assert_array_equal(X[:n2, 2 * nz:4 * nz], ref[:, 2 * nz:4 * nz])
    np.testing.assert_allclose(X[:n2], ref, rtol=2.0 ** -7, atol=1e-30)
    assert np.all(X[n2:] == 0.0)


def bf16_tensor(a, dev):
    return torch.as_tensor(np.asarray(a, dtype=np.float32), device=dev).to(torch.bfloat16)


def test_gemm_epilogues_match_numpy():
    print("case gemm 256x128x96", flush=True)
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(0)
    M, N, K = 256, 128, 96
    A = C.bf16_round(rng.standard_normal((M, K)))
    B = C.bf16_round(rng.standard_normal((N, K)))
    bias = rng.standard_normal(N).astype(np.float32)
    aux = rng.standard_normal((M, N)).astype(np.float32)
    At, Bt = bf16_tensor(A, dev), bf16_tensor(B, dev)
    bt, auxt = torch.as_tensor(bias, device=dev), torch.as_tensor(aux, device=dev)
    ref = A @ B.T
    scale = np.max(np.abs(ref))
    from zhenmode_amd.learned import _p, _stream
    Cf = torch.empty((M, N), dtype=torch.float32, device=dev)
    _lib.check(lib.zm_gemm_bf16_nt(M, N, K, _p(At), _p(Bt), _lib.EPI_F32, None, None, _p(Cf), None, None, _stream()))
    assert np.max(np.abs(Cf.cpu().numpy() - ref)) <= 1e-5 * scale
    _lib.check(lib.zm_gemm_bf16_nt(M, N, K, _p(At), _p(Bt), _lib.EPI_BIAS_F32, _p(bt), None, _p(Cf), None, None,
                                   _stream()))
    assert np.max(np.abs(Cf.cpu().numpy() - (ref + bias))) <= 1e-5 * scale
    Cb = torch.empty((M, N), dtype=torch.bfloat16, device=dev)
    CT = torch.empty((N, M), dtype=torch.bfloat16, device=dev)
    _lib.check(lib.zm_gemm_bf16_nt(M, N, K, _p(At), _p(Bt), _lib.EPI_BIAS_GELU, _p(bt), None, _p(Cf), _p(Cb), _p(CT),
                                   _stream()))
    z = ref + bias
    assert np.max(np.abs(Cf.cpu().numpy() - z)) <= 1e-5 * scale
    g = C.gelu(z)
    got = Cb.float().cpu().numpy()
    assert np.all(np.abs(got - g) <= 2.0 ** -7 * np.abs(g) + 1e-5 * scale)
    np.testing.assert_array_equal(CT.float().cpu().numpy(), got.T)
    _lib.check(lib.zm_gemm_bf16_nt(M, N, K, _p(At), _p(Bt), _lib.EPI_GELU_GRAD, None, _p(auxt), None, _p(Cb), _p(CT),
                                   _stream()))
    gg = ref * C.gelu_grad(aux.astype(np.float64))
    got = Cb.float().cpu().numpy()
    assert np.all(np.abs(got - gg) <= 2.0 ** -7 * np.abs(gg) + 1e-5 * scale)
    np.testing.assert_array_equal(CT.float().cpu().numpy(), got.T)


def test_mlp_gradients_match_cpu_oracle():
    print("case mlp gradients batch 256", flush=True)
    dev = torch.device("cuda", 0)
    mlp = ColumnMLP(50, device=dev, seed=1)
    rng = np.random.default_rng(2)
    M = 256
    Xn = C.bf16_round(rng.standard_normal((M, N_IN)))
    Xn[:, 253:] = 0.0
    target = np.zeros((M, N_OUT_PAD), dtype=np.float32)
    target[:, :100] = rng.standard_normal((M, 100))
    # non-zero biases so that every term is exercised
    for b in mlp.b:
        b.copy_(torch.as_tensor(0.1 * rng.standard_normal(b.shape[0]), dtype=torch.float32))
    X = bf16_tensor(Xn, dev)
    loss, gw, gb, Y = mlp.gradients(X, torch.as_tensor(target, device=dev))
    Ws = [C.bf16_round(w.t().cpu().numpy()) for w in mlp.w]   # (in, out) as the oracle takes them
    bs = [b.cpu().numpy().astype(np.float64) for b in mlp.b]
    Yr, cache = C.mlp_forward(Xn, Ws, bs)
    mask = np.zeros_like(Yr)
    mask[:, :100] = 1.0
    dY = 2.0 * (Yr - target) * mask / (M * 100)
    dWs, dbs = C.mlp_backward(dY, Ws, cache)
    assert rel(Y.cpu().numpy()[:, :100], Yr[:, :100]) <= 3e-2
    ref_loss = np.sum(((Yr - target) * mask) ** 2) / (M * 100)
    assert abs(float(loss.item()) - ref_loss) <= 3e-2 * ref_loss
    for l in range(4):
        g = gw[l].cpu().numpy().T  # (in, out)
        e = rel(g, dWs[l])
        eb = rel(gb[l].cpu().numpy(), dbs[l])
        print(f"layer {l} dW {e:.2e} db {eb:.2e}", flush=True)
        assert e <= 3e-2, (l, e)
        assert eb <= 3e-2, (l, eb)


def test_learned_mixer_conserves_and_falls_back_to_thomas():
    print("case land learned mixer float64", flush=True)
    sol, p, s = land_setup("float64")
    mlp = ColumnMLP(p.nz, device=sol.device, seed=3)
    dt, kv, kc = 3600.0, 1e-3, 0.5
    thomas = to_numpy_state(sol.vertical_mix(JaxStateG(*s), dt, kv, kc))
    # every column rejected: the result is the Thomas solve, bit for bit
    m0 = LearnedMixer(sol, mlp, kv, kc, tol_T=0.0, tol_S=0.0)
    out0 = to_numpy_state(m0.apply(JaxStateG(*s), dt))
    np.testing.assert_array_equal(out0.T, thomas.T)
    np.testing.assert_array_equal(out0.S, thomas.S)
    nw = np.sum(p.wet_mask_z > 0.5, axis=-1)
    # a one-level column's corrected increment is zero, so its residual can be exactly zero
    assert int(np.sum(nw >= 2)) <= int(m0.count.item()) <= int(np.sum(nw >= 1))
    # nothing rejected: the surrogate's increments, with exact column integrals
    m1 = LearnedMixer(sol, mlp, kv, kc, tol_T=1e30, tol_S=1e30)
    out1 = to_numpy_state(m1.apply(JaxStateG(*s), dt))
    assert int(m1.count.item()) == 0
    for Xn, X in ((out1.T, s.T), (out1.S, s.S)):
        i0, i1 = C.column_integral(X, p), C.column_integral(Xn, p)
        assert np.max(np.abs(i1 - i0)) <= 1e-12 * np.max(np.abs(i0))
        dry = p.wet_mask_z < 0.5
        np.testing.assert_array_equal(Xn[dry], X[dry])


def test_training_fits_the_thomas_mixing_and_beats_the_untrained_network():
    print("case g4deg training float32", flush=True)
    from zhenmode_amd.cases import make_case
    case = make_case("g4deg")
    step, init_state, _ = make_solver_global(case.grid, case.physics, case.dt, dtype="float32",
                                             **case.solver_kwargs)
    sol = step.solver
    kv, kc, dt = 1e-4, 0.05, case.dt
    wet3 = case.grid.wet_mask_3d > 0.5

    def cooled(st, seed):
        """A state with the top level cooled by 0.05-0.5 degC in 40 % of the
        wet columns (convectively unstable columns, as surface cooling makes)."""
        x = to_numpy_state(st)
        rng = np.random.default_rng(seed)
        T = x.T.copy()
        cool = rng.uniform(0.05, 0.5, T.shape[:2]) * (rng.uniform(size=T.shape[:2]) < 0.4)
        T[:, :, 0] -= cool
        return JaxStateG(x.u, x.v, np.where(wet3, T, x.T), x.S, x.eta, x.ice)

    states = []
    s = init_state(case.T_init, case.S_init)
    for k in range(6):
        s = step(s)
        states.append(cooled(s, k))
    mlp = ColumnMLP(case.grid.nz, device=sol.device, seed=0)
    tol_T, tol_S = 1e-3, 2e-4
    mixer = LearnedMixer(sol, mlp, kv, kc, tol_T=tol_T, tol_S=tol_S)
    wet = torch.as_tensor(case.grid.wet_mask.T.reshape(-1) > 0.5, device=sol.device)
    n2 = wet.numel()
    Xs, Ys = [], []
    for st in states:
        X = mixer.features(st).clone()
        Xs.append(X[:n2][wet].float())
        Ys.append(mixing_targets(sol, st, dt, kv, kc)[wet])
    Xraw, Yraw = torch.cat(Xs), torch.cat(Ys)
    mlp.fit_normalisation(Xraw, Yraw)
    Xn = []
    for st in states:
        Xn.append(mixer.features(st).clone()[:n2][wet])
    Xn = torch.cat(Xn)
    Yn = torch.zeros((Xn.shape[0], N_OUT_PAD), dtype=torch.float32, device=sol.device)
    Yn[:, :mlp.n_out] = (Yraw - mlp.out_mean[:mlp.n_out]) / mlp.out_std[:mlp.n_out]
    rows = (Xn.shape[0] // 128) * 128
    g = torch.Generator(device="cpu").manual_seed(0)
    first = None
    for it in range(300):
        idx = torch.randperm(Xn.shape[0], generator=g)[:rows].to(sol.device)
        loss = float(mlp.train_step(Xn[idx].contiguous(), Yn[idx].contiguous(), lr=1e-3).item())
        first = loss if first is None else first
    print(f"loss {first:.3e} -> {loss:.3e}", flush=True)
    assert loss < 0.5 * first
    # on a held-out state: the trained network against an untrained one with the same normalisation
    st = cooled(step(s), 99)
    thomas = to_numpy_state(sol.vertical_mix(st, dt, kv, kc))
    untrained = ColumnMLP(case.grid.nz, device=sol.device, seed=1)
    for name in ("in_mean", "in_inv_std", "out_mean", "out_std"):
        getattr(untrained, name).copy_(getattr(mlp, name))
    other = LearnedMixer(sol, untrained, kv, kc, tol_T=tol_T, tol_S=tol_S)
    res = {}
    for name, mx in (("trained", mixer), ("untrained", other)):
        mx.count.zero_()
        out = to_numpy_state(mx.apply(st, dt))
        wm = case.grid.wet_mask_3d > 0.5
        res[name] = (int(mx.count.item()), float(np.max(np.abs(out.T - thomas.T)[wm])),
                     float(np.max(np.abs(out.S - thomas.S)[wm])))
    wet_cols = int(wet.sum().item())
    print(f"of {wet_cols} wet columns: trained fallback {res['trained'][0]}, untrained {res['untrained'][0]}; "
          f"max |dT|, |dS| vs Thomas: trained {res['trained'][1]:.2e}, {res['trained'][2]:.2e}", flush=True)
    # accepted columns are within the residual bound of the Thomas solution (plus float32 rounding)
    assert res["trained"][1] <= tol_T + 1e-4 and res["trained"][2] <= tol_S + 1e-4
    assert res["trained"][0] < res["untrained"][0]


def _apply_raw(sol, st, Y, omean, ostd, dt, kv, kc, tol_T, tol_S):
    """zm_mlp_apply without the fallback: (To, So, flags, count)."""
    import ctypes
    from zhenmode_amd.learned import _p
    src = sol.coerce(st)
    To, So = torch.empty_like(src.T), torch.empty_like(src.S)
    flags = torch.zeros(sol.shape2[0] * sol.shape2[1], dtype=torch.int32, device=sol.device)
    count = torch.zeros(1, dtype=torch.int32, device=sol.device)
    _lib.check(sol.lib.zm_mlp_apply(sol.handle, ctypes.byref(sol.zm_state(src)), _p(Y), Y.shape[1], _p(omean),
                                    _p(ostd), dt, kv, kc, tol_T, tol_S, _p(To), _p(So), _p(flags), _p(count),
                                    sol._stream()), sol.handle)
    return To, So, flags, count


def test_mixing_targets_match_oracle():
    print("case land targets float64", flush=True)
    sol, p, s = land_setup("float64")
    dt, kv, kc = 3600.0, 1e-3, 0.5
    got = mixing_targets(sol, JaxStateG(*s), dt, kv, kc).cpu().numpy()
    Tn, Sn = C.implicit_mix(s.T, s.S, p, dt, kv, kc)
    ref = C.surrogate_targets(s.T, s.S, Tn, Sn, p, dt, kv, kc)
    n2, nz = p.nx * p.ny, p.nz
    ref = np.transpose(ref, (1, 0, 2)).reshape(n2, 2 * nz)
    kap = C.mixing_kappa(s.T, s.S, p, kv, kc)
    for q, X in enumerate((s.T, s.S)):
        sc = np.transpose(C.mixing_scale(C.explicit_increment(X, kap, p, dt)), (1, 0)).reshape(n2)
        g, r = got[:, q * nz:(q + 1) * nz], ref[:, q * nz:(q + 1) * nz]
        # float32 rows: one rounding of the target, plus float64 round-off of the departure (target x scale)
        atol = 1e-11 * np.max(np.abs(X)) / np.where(sc > 0, sc, 1.0)[:, None]
        assert np.all(np.abs(g - r) <= 2.0 ** -23 * np.abs(r) + atol), q
        assert np.all(g[sc == 0.0] == 0.0)
    print(f"  targets max |t| {np.max(np.abs(ref)):.2f}", flush=True)


def test_learned_apply_matches_oracle():
    print("case land learned apply float64", flush=True)
    sol, p, s = land_setup("float64")
    dt, kv, kc = 3600.0, 1e-3, 0.5
    rng = np.random.default_rng(9)
    n2, nz = p.nx * p.ny, p.nz
    Y = torch.zeros((n2, N_OUT_PAD), dtype=torch.float32, device=sol.device)
    y = rng.standard_normal((n2, 2 * nz)).astype(np.float32)
    Y[:, :2 * nz] = torch.as_tensor(y, device=sol.device)
    om = (0.1 * rng.standard_normal(N_OUT_PAD)).astype(np.float32)
    os_ = rng.uniform(0.05, 0.5, N_OUT_PAD).astype(np.float32)
    omean, ostd = torch.as_tensor(om, device=sol.device), torch.as_tensor(os_, device=sol.device)
    tol_T, tol_S = 2e-2, 2e-3
    To, So, flags, count = _apply_raw(sol, JaxStateG(*s), Y, omean, ostd, dt, kv, kc, tol_T, tol_S)
    z = y.astype(np.float64) * os_[:2 * nz].astype(np.float64) + om[:2 * nz].astype(np.float64)
    z = np.transpose(z.reshape(p.ny, p.nx, 2 * nz), (1, 0, 2))
    Tr, Sr, fr, rT, rS = C.learned_apply(s.T, s.S, z, p, dt, kv, kc, tol_T, tol_S)
    got_T, got_S = To.cpu().numpy(), So.cpu().numpy()  # (nx, ny, nz) views of the device layout
    assert np.max(np.abs(got_T - Tr)) <= 1e-11 * np.max(np.abs(s.T))
    assert np.max(np.abs(got_S - Sr)) <= 1e-11 * np.max(np.abs(s.S))
    gf = flags.cpu().numpy().reshape(p.ny, p.nx).T.astype(bool)
    clear = (np.abs(rT - tol_T) > 1e-9 * tol_T) & (np.abs(rS - tol_S) > 1e-9 * tol_S)
    np.testing.assert_array_equal(gf[clear], fr[clear])
    assert int(count.item()) == int(gf.sum())
    print(f"  flagged {int(gf.sum())} of {int((p.wet_mask > 0.5).sum())} wet columns", flush=True)
    assert 0 < int(gf.sum()) < int((p.wet_mask > 0.5).sum())


def _random_surrogate(sol, nz, seed):
    rng = np.random.default_rng(seed)
    mlp = ColumnMLP(nz, device=sol.device, seed=seed)
    for b in mlp.b:
        b.copy_(torch.as_tensor(0.1 * rng.standard_normal(b.shape[0]), dtype=torch.float32))
    mlp.out_mean.copy_(torch.as_tensor(0.05 * rng.standard_normal(N_OUT_PAD), dtype=torch.float32))
    mlp.out_std.copy_(torch.as_tensor(rng.uniform(0.02, 0.2, N_OUT_PAD), dtype=torch.float32))
    return mlp


@pytest.mark.parametrize("case", ["land", "g4deg"])
def test_fused_surrogate_equals_the_unfused_kernels(case):
    print(f"case {case} fused surrogate float32", flush=True)
    if case == "land":
        sol, p, s = land_setup("float32")
        st = JaxStateG(*s)
        dt, kv, kc = 3600.0, 1e-3, 0.5
    else:
        from zhenmode_amd.cases import make_case
        c = make_case("g4deg")
        step, init_state, _ = make_solver_global(c.grid, c.physics, c.dt, dtype="float32", **c.solver_kwargs)
        sol = step.solver
        st = init_state(c.T_init, c.S_init)
        for _ in range(3):
            st = step(st)
        dt, kv, kc = c.dt, 1e-4, 0.05
    nz = sol.shape3[2]
    mlp = _random_surrogate(sol, nz, 5)
    tol_T, tol_S = (1e-3, 2e-4) if case == "land" else (1e-6, 2e-7)
    mixer = LearnedMixer(sol, mlp, kv, kc, tol_T=tol_T, tol_S=tol_S)
    # input normalisation fitted on this state's own features
    Xraw = mixer.features(st).float()[:sol.shape2[0] * sol.shape2[1]]
    mlp.in_mean.copy_(Xraw.mean(0))
    mlp.in_inv_std.copy_(1.0 / Xraw.std(0).clamp_min(1e-3))
    # unfused: features -> 4 GEMMs -> zm_mlp_apply
    X = mixer.features(st)
    Y, _ = mlp.forward(X)
    To_u, So_u, fl_u, cnt_u = _apply_raw(sol, st, Y, mlp.out_mean, mlp.out_std, dt, kv, kc, tol_T, tol_S)
    # fused launch (zm_mlp_mix), fallback switched off by comparing before it
    import ctypes
    from zhenmode_amd.learned import _p
    src = sol.coerce(st)
    To_f, So_f = torch.empty_like(src.T), torch.empty_like(src.S)
    fl_f = torch.full_like(fl_u, -1)
    cnt_f = torch.zeros(1, dtype=torch.int32, device=sol.device)
    _lib.check(sol.lib.zm_mlp_mix(sol.handle, ctypes.byref(sol.zm_state(src)), None, None, None, _p(mlp.in_mean),
                                  _p(mlp.in_inv_std), mixer._wptrs, mixer._bptrs, _p(mlp.out_mean), _p(mlp.out_std),
                                  dt, kv, kc, tol_T, tol_S, _p(To_f), _p(So_f), _p(fl_f), _p(cnt_f), sol._stream()),
               sol.handle)
    torch.cuda.synchronize()
    n_fl = int(fl_u.sum().item())
    print(f"  columns {fl_u.numel()}, flagged {n_fl}", flush=True)
    assert torch.equal(To_f, To_u) and torch.equal(So_f, So_u)
    assert torch.equal(fl_f, fl_u) and int(cnt_f.item()) == int(cnt_u.item()) == n_fl
    assert 0 < n_fl < fl_u.numel()
    # the mixer's float32 path (fused + fallback) against the unfused path with the same fallback
    mixer.count.zero_()
    a = mixer.apply(st, dt)
    b = mixer.apply_unfused(st, dt)
    assert torch.equal(a.T, b.T) and torch.equal(a.S, b.S)


def test_fused_surrogate_rejects_float64_handles():
    print("case land fused surrogate float64 rejected", flush=True)
    import ctypes
    from zhenmode_amd.learned import _p
    sol, p, s = land_setup("float64")
    mlp = _random_surrogate(sol, p.nz, 1)
    mixer = LearnedMixer(sol, mlp, 1e-3, 0.5)
    src = sol.coerce(JaxStateG(*s))
    To, So = torch.empty_like(src.T), torch.empty_like(src.S)
    rc = sol.lib.zm_mlp_mix(sol.handle, ctypes.byref(sol.zm_state(src)), None, None, None, _p(mlp.in_mean),
                            _p(mlp.in_inv_std), mixer._wptrs, mixer._bptrs, _p(mlp.out_mean), _p(mlp.out_std),
                            600.0, 1e-3, 0.5, 1e-3, 1e-3, _p(To), _p(So), _p(mixer.flags), _p(mixer.count),
                            sol._stream())
    assert rc == _lib.ZM_EINVAL
