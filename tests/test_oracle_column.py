"""Pins the column operators of the oracle (Thomas solve, implicit mixing,
integral correction, MLP forward/backward) with library routines, an
independent explicit operator, closed forms and finite differences."""

import numpy as np
import pytest
import scipy.linalg

from oracle import zm_column as C
from oracle import zm_oracle as O
from tests.fixtures import land_case, oracle_params


def random_system(rng, shape, n):
    a = -rng.uniform(0.0, 2.0, shape + (n,))
    c = -rng.uniform(0.0, 2.0, shape + (n,))
    a[..., 0] = 0.0
    c[..., -1] = 0.0
    b = 1.0 - a - c + rng.uniform(0.0, 0.5, shape + (n,))
    d = rng.standard_normal(shape + (n,))
    return a, b, c, d


def test_thomas_matches_solve_banded_and_dense_solve():
    rng = np.random.default_rng(0)
    a, b, c, d = random_system(rng, (7,), 50)
    x = C.thomas(a, b, c, d)
    for col in range(7):
        ab = np.zeros((3, 50))
        ab[0, 1:] = c[col, :-1]
        ab[1] = b[col]
        ab[2, :-1] = a[col, 1:]
        ref = scipy.linalg.solve_banded((1, 1), ab, d[col])
        np.testing.assert_allclose(x[col], ref, rtol=1e-12, atol=1e-12)
        dense = np.diag(b[col]) + np.diag(c[col, :-1], 1) + np.diag(a[col, 1:], -1)
        np.testing.assert_allclose(x[col], np.linalg.solve(dense, d[col]), rtol=1e-12, atol=1e-12)


def test_thomas_with_a_wrong_sign_fails_the_dense_check():
    """A planted mistake in the elimination (sign of a) is caught by the check above."""
    rng = np.random.default_rng(1)
    a, b, c, d = random_system(rng, (1,), 12)
    x_bad = C.thomas(-a, b, c, d)[0]
    dense = np.diag(b[0]) + np.diag(c[0, :-1], 1) + np.diag(a[0, 1:], -1)
    assert not np.allclose(x_bad, np.linalg.solve(dense, d[0]), rtol=1e-6)


@pytest.fixture(scope="module")
def land():
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, kwargs)
    s = O.init_state(p, T, S)
    return p, s


def test_implicit_mix_solves_the_explicit_flux_form_system(land):
    p, s = land
    dt = 3600.0
    Tn, Sn = C.implicit_mix(s.T, s.S, p, dt, 1e-3, 0.5)
    kap = C.mixing_kappa(s.T, s.S, p, 1e-3, 0.5)
    # residual with the oracle's own explicit interface-flux operator
    for Xn, X in ((Tn, s.T), (Sn, s.S)):
        flux = O.diffusive_interface_flux(Xn, 1.0, p) * kap
        wm = p.wet_mask_z
        res = Xn - dt * O.iface_flux_divergence(flux, p) - X
        assert np.max(np.abs(res * wm)) <= 1e-11 * np.max(np.abs(X))
    # dry levels held
    dry = p.wet_mask_z < 0.5
    np.testing.assert_array_equal(Tn[dry], s.T[dry])


def test_implicit_mix_conserves_column_integrals_and_constants(land):
    p, s = land
    Tn, Sn = C.implicit_mix(s.T, s.S, p, 86400.0, 1e-3, 1.0)
    for Xn, X in ((Tn, s.T), (Sn, s.S)):
        i0, i1 = C.column_integral(X, p), C.column_integral(Xn, p)
        assert np.max(np.abs(i1 - i0)) <= 1e-12 * np.max(np.abs(i0))
    const = np.full_like(s.T, 7.5)
    Tc, _ = C.implicit_mix(const, s.S, p, 86400.0, 1e-3, 1.0)
    np.testing.assert_allclose(Tc, const, rtol=0, atol=1e-12)
    T0, S0 = C.implicit_mix(s.T, s.S, p, 86400.0, 0.0, 0.0)
    np.testing.assert_array_equal(T0, s.T)


def test_implicit_mix_two_level_closed_form(land):
    """Two wet levels: the exchange decays by 1/(1 + dt k (1/h0 + 1/h1)/dzi)."""
    p, s = land
    nw = (p.wet_mask_z > 0.5).sum(axis=-1)
    i, j = np.argwhere(nw == 2)[0]
    dt, kv = 5000.0, 2e-3
    T = s.T.copy()
    T[i, j, :2] = (20.0, 10.0)
    S = s.S.copy()
    S[i, j, :2] = (35.0, 35.0)  # stable: no convective term
    Tn, _ = C.implicit_mix(T, S, p, dt, kv, 0.0)
    h0, h1, dzi = p.dz_node[0], p.dz_node[1], p.dz_iface[0]
    decay = 1.0 / (1.0 + dt * kv / dzi * (1.0 / h0 + 1.0 / h1))
    np.testing.assert_allclose(Tn[i, j, 0] - Tn[i, j, 1], 10.0 * decay, rtol=1e-12)


def test_integral_correction_hits_the_target(land):
    p, s = land
    rng = np.random.default_rng(2)
    dX = rng.standard_normal(s.T.shape) * p.wet_mask_z
    target = rng.standard_normal(s.T.shape[:2]) * p.wet_mask
    out = C.integral_correction(dX, target, p)
    np.testing.assert_allclose(C.column_integral(out, p), target, rtol=0, atol=1e-12)


def small_mlp(rng, sizes):
    Ws = [rng.standard_normal((k, n)) / np.sqrt(k) for k, n in zip(sizes[:-1], sizes[1:])]
    bs = [0.1 * rng.standard_normal(n) for n in sizes[1:]]
    return Ws, bs


def test_mlp_gradients_match_central_differences():
    rng = np.random.default_rng(3)
    Ws, bs = small_mlp(rng, [6, 5, 4, 3])
    X = rng.standard_normal((8, 6))
    target = rng.standard_normal((8, 3))
    Y, cache = C.mlp_forward(X, Ws, bs)
    dY = 2.0 * (Y - target) / Y.size
    dWs, dbs = C.mlp_backward(dY, Ws, cache)
    eps = 1e-6
    for l in range(3):
        for idx in [(0, 0), (1, 2), (Ws[l].shape[0] - 1, Ws[l].shape[1] - 1)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += eps
            Wm[l][idx] -= eps
            num = (C.mse_loss(C.mlp_forward(X, Wp, bs)[0], target) -
                   C.mse_loss(C.mlp_forward(X, Wm, bs)[0], target)) / (2 * eps)
            assert abs(num - dWs[l][idx]) <= 1e-7 * max(1.0, abs(num)), (l, idx, num, dWs[l][idx])
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][1] += eps
        bm[l][1] -= eps
        num = (C.mse_loss(C.mlp_forward(X, Ws, bp)[0], target) -
               C.mse_loss(C.mlp_forward(X, Ws, bm)[0], target)) / (2 * eps)
        assert abs(num - dbs[l][1]) <= 1e-7 * max(1.0, abs(num))


def test_gelu_closed_forms():
    x = np.array([-3.0, -1.0, 0.0, 1.0, 3.0])
    g = C.gelu(x)
    from scipy.stats import norm
    np.testing.assert_allclose(g, x * norm.cdf(x), rtol=1e-14, atol=1e-15)
    eps = 1e-6
    np.testing.assert_allclose(C.gelu_grad(x), (C.gelu(x + eps) - C.gelu(x - eps)) / (2 * eps), rtol=1e-8)


def test_bf16_rounding_is_round_to_nearest_even():
    # bf16 keeps 7 fraction bits: the spacing at 1.0 is 2^-7
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -9, 1.0 + 3 * 2.0 ** -8, -2.5])
    out = C.bf16_round(vals)
    assert out[0] == 1.0
    assert out[1] == 1.0                        # tie between 1 and 1 + 2^-7: to even
    assert out[2] == 1.0 + 2.0 ** -7            # above the half-way point rounds up
    assert out[3] == 1.0                        # below the half-way point rounds down
    assert out[4] == 1.0 + 2.0 ** -6            # tie between odd 1 + 2^-7 and even 1 + 2^-6
    assert out[5] == -2.5


def test_column_features_layout(land):
    p, s = land
    nx, ny, nz = s.T.shape
    q = np.arange(nx * ny, dtype=float).reshape(nx, ny)
    X = C.column_features(s.T, s.S, s.u, s.v, p, q, 2 * q, 3 * q)
    i, j = 5, 7
    row = X[j * nx + i]
    wm = p.wet_mask_z[i, j]
    nw = int(wm.sum())
    for q0, X in ((0, s.T[i, j]), (nz, s.S[i, j])):
        # interface differences down to the last wet interface, then zeros; surface value last
        for k in range(nz - 1):
            assert row[q0 + k] == (X[k] - X[k + 1] if k + 1 < nw else 0.0)
        assert row[q0 + nz - 1] == (X[0] if nw > 0 else 0.0)
    np.testing.assert_array_equal(row[2 * nz:3 * nz], s.u[i, j] * wm)
    assert row[5 * nz] == q[i, j] and row[5 * nz + 1] == 2 * q[i, j] and row[5 * nz + 2] == 3 * q[i, j]
    assert np.all(row[5 * nz + 3:] == 0.0)


# --------------------------------------------------------------------------
# the surrogate's increment formulation
# --------------------------------------------------------------------------
def _unstable_land(land):
    p, s = land
    rng = np.random.default_rng(11)
    Tu = s.T.copy()
    Tu[:, :, 1] += 3.0 * (rng.uniform(size=Tu.shape[:2]) > 0.6)
    return p, np.where(p.wet_mask_z > 0.5, Tu, s.T), s.S


def _flux_form_increment(X, kap, p, dt):
    """dt D_kappa X through zm_oracle's interface-flux operator."""
    flux = O.diffusive_interface_flux(X, 1.0, p) * kap
    return dt * O.iface_flux_divergence(flux, p) * p.wet_mask_z


def test_explicit_increment_matches_the_flux_form_and_telescopes(land):
    p, T, S = _unstable_land(land)
    dt = 3600.0
    kap = C.mixing_kappa(T, S, p, 1e-3, 0.5)
    for X in (T, S):
        e = C.explicit_increment(X, kap, p, dt)
        ref = _flux_form_increment(X, kap, p, dt)
        assert np.max(np.abs(e - ref)) <= 1e-12 * np.max(np.abs(ref))
        # the same operator as the implicit step's residual at X_new = X
        np.testing.assert_allclose(e, -C.mix_residual(X, X, kap, p, dt) * p.wet_mask_z, rtol=0,
                                   atol=1e-12 * np.max(np.abs(ref)))
        assert np.max(np.abs(C.column_integral(e, p))) <= 1e-12 * np.max(np.abs(C.column_integral(np.abs(e), p)))
    # planted mistake: the sub- and super-diagonals exchanged
    def swapped(X, kappa, p, dt):
        a, b, c = C.tridiagonal(kappa, p, dt)
        AX = b * X
        AX[..., 1:] += c[..., :-1] * X[..., :-1]
        AX[..., :-1] += a[..., 1:] * X[..., 1:]
        return np.where(p.wet_mask_z > 0.5, X - AX, 0.0)
    bad = swapped(T, kap, p, dt)
    assert np.max(np.abs(bad - _flux_form_increment(T, kap, p, dt))) > 1e-6 * np.max(np.abs(bad))


def test_explicit_increment_two_level_closed_form(land):
    """Two wet levels: e_0 = dt k (X1 - X0) / (dzi h0), e_1 = -dt k (X1 - X0) / (dzi h1)."""
    p, s = land
    nw = (p.wet_mask_z > 0.5).sum(axis=-1)
    i, j = np.argwhere(nw == 2)[0]
    dt, kv = 5000.0, 2e-3
    T = s.T.copy()
    T[i, j, :2] = (20.0, 10.0)
    S = s.S.copy()
    S[i, j, :2] = (35.0, 35.0)
    kap = C.mixing_kappa(T, S, p, kv, 0.0)
    e = C.explicit_increment(T, kap, p, dt)
    h0, h1, dzi = p.dz_node[0], p.dz_node[1], p.dz_iface[0]
    np.testing.assert_allclose(e[i, j, :2], (-10.0 * dt * kv / (dzi * h0), 10.0 * dt * kv / (dzi * h1)), rtol=1e-13)
    assert C.mixing_scale(e)[i, j] == pytest.approx(10.0 * dt * kv / (dzi * min(h0, h1)), rel=1e-13)


def test_exact_targets_fed_back_reproduce_the_thomas_solution(land):
    p, T, S = _unstable_land(land)
    dt, kv, kc = 3600.0, 1e-3, 0.5
    Tn, Sn = C.implicit_mix(T, S, p, dt, kv, kc)
    z = C.surrogate_targets(T, S, Tn, Sn, p, dt, kv, kc)
    assert np.all(np.isfinite(z))
    Ta, Sa, flags, rT, rS = C.learned_apply(T, S, z, p, dt, kv, kc, 1e-9, 1e-9)
    np.testing.assert_allclose(Ta, Tn, rtol=0, atol=1e-11 * np.max(np.abs(T)))
    np.testing.assert_allclose(Sa, Sn, rtol=0, atol=1e-11 * np.max(np.abs(S)))
    assert not flags.any() and max(rT.max(), rS.max()) <= 1e-10
    # zero outputs give the explicit step, whose residual is large in convecting columns
    _, _, flags0, rT0, _ = C.learned_apply(T, S, np.zeros_like(z), p, dt, kv, kc, 1e-3, 1e-3)
    assert flags0.any() and rT0.max() > 1e-2
    # planted mistake: the explicit increment left out of the apply step
    def without_e(T, S, z):
        nz = T.shape[-1]
        kap = C.mixing_kappa(T, S, p, kv, kc)
        e = C.explicit_increment(T, kap, p, dt)
        dX = C.integral_correction(C.mixing_scale(e)[..., None] * z[..., :nz] * p.wet_mask_z, 0.0, p)
        return np.where(p.wet_mask_z > 0.5, T + dX, T)
    assert np.max(np.abs(without_e(T, S, z) - Tn)) > 1e-3


def test_learned_apply_conserves_column_integrals(land):
    p, T, S = _unstable_land(land)
    rng = np.random.default_rng(4)
    z = rng.standard_normal(T.shape[:2] + (2 * p.nz,))
    Ta, Sa, _, _, _ = C.learned_apply(T, S, z, p, 3600.0, 1e-3, 0.5, 1e-3, 1e-3)
    for Xn, X in ((Ta, T), (Sa, S)):
        i0, i1 = C.column_integral(X, p), C.column_integral(Xn, p)
        assert np.max(np.abs(i1 - i0)) <= 1e-12 * np.max(np.abs(i0))
        dry = p.wet_mask_z < 0.5
        np.testing.assert_array_equal(Xn[dry], X[dry])
