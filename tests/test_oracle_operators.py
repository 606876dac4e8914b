"""Pin the oracle's operators (CPU).

The assertions restate the reference's own invariant tests on the same fixture
grids (file:line cited per test), plus closed forms and explicit index-loop
evaluations on tiny grids that catch a wrong shift direction, a dropped gate
or a transposed metric.
"""

import numpy as np
import pytest

from oracle import zm_oracle as O
from tests.fixtures import diffusion_fixture, land_case, land_grid, oracle_params, quiet_physics
from zhenmode_amd.config import R_EARTH
from zhenmode_amd.grid import all_wet_grid


# ---------------------------------------------------------------------------
# D1 diffusion (reference tests/fd/test_horizontal_tracer_diffusion.py)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("latitude_limit", [30.0, 65.0])
@pytest.mark.parametrize("land", [False, True])
@pytest.mark.parametrize("seed", [2718, 3141, 1618])
@pytest.mark.parametrize("biharmonic", [False, True])
def test_d1_closes_weighted_budget(latitude_limit, land, seed, biharmonic):
    # :19-28
    _, p, volume = diffusion_fixture(latitude_limit, land)
    tracer = np.random.default_rng(seed).uniform(5.0, 25.0, volume.shape)
    tend = -O.tracer_biharmonic(tracer, p) if biharmonic else O.tracer_diffusion(tracer, p)
    assert abs(np.sum(tend * volume)) <= 5e-13 * np.sum(np.abs(tend) * volume)


@pytest.mark.parametrize("biharmonic", [False, True])
def test_d1_does_not_read_closed_faces(biharmonic):
    # :31-39
    _, p, _ = diffusion_fixture(65.0, land=True)
    wet = p.wet_mask_z > 0.5
    tracer = np.random.default_rng(2718).uniform(5.0, 25.0, wet.shape)
    op = O.tracer_biharmonic if biharmonic else O.tracer_diffusion
    first = op(np.where(wet, tracer, -100.0), p)
    second = op(np.where(wet, tracer, 100.0), p)
    np.testing.assert_array_equal(first[wet], second[wet])
    np.testing.assert_array_equal(first[~wet], 0.0)


@pytest.mark.parametrize("land", [False, True])
@pytest.mark.parametrize("biharmonic", [False, True])
def test_d1_constants_semidefinite_self_adjoint(land, biharmonic):
    # :41-57
    _, p, volume = diffusion_fixture(65.0, land)
    rng = np.random.default_rng(2718)
    a, b = rng.normal(size=volume.shape), rng.normal(size=volume.shape)

    def op(f):
        return -O.tracer_biharmonic(f, p) if biharmonic else O.tracer_diffusion(f, p)

    ta, tb = op(a), op(b)
    np.testing.assert_array_equal(op(np.full(volume.shape, 20.0)), 0.0)
    assert np.sum(a * ta * volume) <= 0.0
    assert np.sum(a * tb * volume) == pytest.approx(np.sum(b * ta * volume), rel=5e-13)


def test_d1_safe_step_creates_no_extrema():
    # :59-66
    _, p, _ = diffusion_fixture(65.0, land=True)
    wet = p.wet_mask_z > 0.5
    tracer = np.random.default_rng(2718).uniform(5.0, 25.0, wet.shape)
    rate = p.kappa_h * float(np.max(p.inv_dx ** 2) + p.inv_dy ** 2)
    updated = tracer + (0.25 / rate) * O.tracer_diffusion(tracer, p)
    assert updated[wet].min() >= tracer[wet].min()
    assert updated[wet].max() <= tracer[wet].max()


@pytest.mark.parametrize("kappa_h,kappa_bi", [(100.0, 0.0), (0.0, 1e15), (100.0, 1e15)])
def test_linear_half_step_closes_heat_and_salt(kappa_h, kappa_bi):
    # :68-85 (mode-split half-step; nu_h = 0 so the sub-cycle is the identity)
    _, p, volume = diffusion_fixture(65.0, land=True)
    p.kappa_h, p.kappa_bi = kappa_h, kappa_bi
    p.n_subcyc = 10
    rng = np.random.default_rng(2718)
    T = rng.uniform(5.0, 25.0, volume.shape)
    S = rng.uniform(30.0, 37.0, volume.shape)
    z3 = np.zeros(volume.shape)
    state = O.State(z3, z3, T, S, np.zeros(volume.shape[:2]), np.zeros(volume.shape[:2]))
    out = O.linear_half_step(state, p, 30.0)
    for before, after in ((state.T, out.T), (state.S, out.S)):
        roundoff = 100.0 * np.finfo(np.float64).eps * np.sum(np.abs(before) * volume)
        assert abs(np.sum((after - before) * volume)) <= roundoff
        mean = np.sum(before * volume) / np.sum(volume)
        assert np.sum((after - mean) ** 2 * volume) <= np.sum((before - mean) ** 2 * volume)


def test_d1_second_order_spherical_accuracy():
    # :87-100
    errors = []
    for ny in (24, 48, 96):
        grid, p, volume = diffusion_fixture(50.0, ny=ny, nx=2 * ny)
        lon = 2.0 * np.pi * np.arange(grid.nx) / grid.nx
        tracer = np.broadcast_to(np.sin(lon)[:, None, None] * grid.cos_lat[None, :, None], volume.shape)
        exact = -2.0 * p.kappa_h * tracer / R_EARTH ** 2
        got = O.tracer_diffusion(np.array(tracer), p)
        sl = (slice(None), slice(4, -4), slice(None))
        errors.append(np.sqrt(np.sum((got[sl] - exact[sl]) ** 2 * volume[sl])
                              / np.sum(exact[sl] ** 2 * volume[sl])))
    assert errors[0] / errors[1] >= 3.5
    assert errors[1] / errors[2] >= 3.5


# ---------------------------------------------------------------------------
# vertical operators (reference tests/fd/test_vertical_bc.py:86-138)
# ---------------------------------------------------------------------------
def _column_params():
    grid = land_grid(nx=4, ny=6, nz=6)
    return grid, O.make_params(grid, quiet_physics(), 600.0, polar_cap_rows=0, polar_cap_taper=0)


def test_d2dz2_ignores_ghost_sentinel():
    _, p = _column_params()
    wet = p.wet_mask_z > 0.5
    rng = np.random.default_rng(1)
    T = rng.uniform(2.0, 20.0, wet.shape)
    a = O.d2_dz2(np.where(wet, T, 15.0), p)
    b = O.d2_dz2(np.where(wet, T, 1.0e6), p)
    np.testing.assert_array_equal(a[wet], b[wet])


def test_d2dz2_bottom_node_is_one_sided_and_deep_column_centered():
    _, p = _column_params()
    rng = np.random.default_rng(2)
    T = rng.uniform(2.0, 20.0, p.wet_mask_z.shape)
    got = O.d2_dz2(T, p)
    nwet = p.wet_mask_z.sum(-1).astype(int)
    checked_shallow = checked_deep = 0
    for i in range(p.nx):
        for j in range(p.ny):
            n = nwet[i, j]
            for k in range(1, p.nz - 1):
                hm, hp, den = p.d2z_hm[k - 1], p.d2z_hp[k - 1], p.d2z_denom[k - 1]
                if 2 <= n < p.nz and k == n - 1:
                    expected = hp * (T[i, j, k - 1] - T[i, j, k]) / den
                    assert np.isclose(got[i, j, k], expected, rtol=1e-13, atol=0.0)
                    checked_shallow += 1
                elif n == p.nz:
                    expected = ((T[i, j, k + 1] * hm + T[i, j, k - 1] * hp - T[i, j, k] * (hm + hp)) / den)
                    assert got[i, j, k] == expected
                    checked_deep += 1
    assert checked_shallow > 0 and checked_deep > 0


def test_vertical_metric_closed_forms():
    _, p = _column_params()
    z = p.z
    # linear profile: node-form second derivative vanishes in the interior
    lin = np.broadcast_to(3.0 + 0.01 * z, p.wet_mask_z.shape).copy()
    deep = p.wet_mask_z.sum(-1) == p.nz
    assert np.allclose(O.d2_dz2(lin, p)[deep][:, 1:-1], 0.0, atol=1e-15)
    # d_dz of a linear profile is its slope at every node
    assert np.allclose(O.d_dz(lin, p)[deep], 0.01, rtol=1e-12)
    # dz_node sums to the vertical span plus half the end gaps
    assert np.isclose(np.sum(p.dz_node), abs(z[-1] - z[0]) + 0.5 * (abs(z[1] - z[0]) + abs(z[-1] - z[-2])))


# ---------------------------------------------------------------------------
# polar cap (reference tests/fd/test_polar_cap.py:64-114)
# ---------------------------------------------------------------------------
def _cap_params(ncap=2, ntaper=3):
    grid = all_wet_grid()
    p = O.make_params(grid, quiet_physics(), 600.0, polar_cap_rows=ncap, polar_cap_taper=ntaper)
    return grid, p


def test_polar_cap_pole_rows_flat_and_interior_untouched():
    grid, p = _cap_params()
    nx, ny, nb = grid.nx, grid.ny, 5
    field = np.broadcast_to(np.where(np.arange(nx) % 2 == 0, 1.0, -1.0)[:, None], (nx, ny)).copy()
    out = O.apply_polar_cap(field, grid.wet_mask, p)
    for j in list(range(2)) + list(range(ny - 2, ny)):
        assert np.allclose(out[:, j], 0.0)
    assert np.array_equal(out[:, nb:ny - nb], field[:, nb:ny - nb])


def test_polar_cap_mirror_symmetric():
    grid, p = _cap_params()
    field = np.random.default_rng(0).standard_normal((grid.nx, grid.ny))
    out = O.apply_polar_cap(field, grid.wet_mask, p)
    flipped = O.apply_polar_cap(field[:, ::-1].copy(), grid.wet_mask, p)
    assert np.allclose(out, flipped[:, ::-1])


def test_polar_cap_taper_monotone():
    grid, p = _cap_params()
    nx, ny = grid.nx, grid.ny
    field = np.broadcast_to(np.where(np.arange(nx) % 2 == 0, 1.0, -1.0)[:, None], (nx, ny)).copy()
    amp = np.abs(O.apply_polar_cap(field, grid.wet_mask, p)[0, :])
    assert amp[0] == 0.0
    assert np.all(np.diff(amp[:5]) >= -1e-12)
    assert np.allclose(amp[5], 1.0)


def test_polar_cap_preserves_row_wet_sum_and_land():
    grid = land_grid()
    p = O.make_params(grid, quiet_physics(), 600.0, polar_cap_rows=2, polar_cap_taper=2)
    f = np.random.default_rng(4).standard_normal(p.wet_mask_z.shape)
    out = O.apply_polar_cap(f, p.wet_mask_z, p)
    wm = p.wet_mask_z
    assert np.allclose(np.sum(out * wm, axis=0), np.sum(f * wm, axis=0), rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(out[:, :4][wm[:, :4] == 0], 0.0)


# ---------------------------------------------------------------------------
# transport: Fz inverse, uniform-tracer identity, index-loop restatements
# ---------------------------------------------------------------------------
def _land_params():
    grid, physics, dt, kwargs, T, S = land_case()
    return grid, oracle_params(grid, physics, dt, kwargs), T, S


def test_fz_is_exact_inverse_of_layer_divergence():
    _, p, _, _ = _land_params()
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal(p.wet_mask_z.shape), rng.standard_normal(p.wet_mask_z.shape)
    Fz = O.vertical_transport_iface(u, v, p)
    assert np.all(Fz[..., -1] == 0.0)
    div = O.divergence_h(u, v, p)
    np.testing.assert_allclose(-(Fz[..., 1:] - Fz[..., :-1]) / p.dz_node, div, rtol=1e-9, atol=1e-12)


def test_uniform_tracer_has_zero_advective_tendency():
    # identity stated at transport.py:114-119 (pinned for limiter variants by
    # tests/fd/test_fct_advection.py:44-64)
    _, p, _, _ = _land_params()
    rng = np.random.default_rng(6)
    u, v = rng.standard_normal(p.wet_mask_z.shape), rng.standard_normal(p.wet_mask_z.shape)
    Fz = O.vertical_transport_iface(u, v, p)
    for value in (0.0, 7.5, 35.0):
        tend = O.advection_scalar(np.full(p.wet_mask_z.shape, value), u, v, Fz, p)
        assert np.max(np.abs(tend)) <= 1e-12 * max(value, 1.0) * np.max(np.abs(Fz)) / np.min(p.dz_node)


def _loop_laplacian(u, p):
    nx, ny, nz = u.shape
    wm = p.wet_mask_z
    sin = p.f[0, :] / (2.0 * O.OMEGA)
    out = np.zeros_like(u)
    for i in range(nx):
        ip, im = (i + 1) % nx, (i - 1) % nx
        for j in range(ny):
            jn, js = min(j + 1, ny - 1), max(j - 1, 0)
            for k in range(nz):
                c = u[i, j, k]
                gx = ((u[ip, j, k] - c) * wm[i, j, k] * wm[ip, j, k]
                      - (c - u[im, j, k]) * wm[i, j, k] * wm[im, j, k]) / p.dx_2d[i, j] ** 2
                gy = ((u[i, jn, k] - c) * wm[i, jn, k] * wm[i, j, k]
                      - (c - u[i, js, k]) * wm[i, j, k] * wm[i, js, k]) / p.dy ** 2
                metric = -(u[i, jn, k] - u[i, js, k]) / (2.0 * p.dy) * sin[j] / p.cos_lat[j] / R_EARTH
                out[i, j, k] = gx + gy + metric
    return out


def test_laplacian_matches_index_loop():
    _, p, _, _ = _land_params()
    u = np.random.default_rng(8).standard_normal(p.wet_mask_z.shape)
    np.testing.assert_allclose(O.laplacian_h(u, p), _loop_laplacian(u, p), rtol=1e-11, atol=1e-22)


def _loop_divergence(u, v, p):
    nx, ny, nz = u.shape
    wm = p.wet_mask_z
    out = np.zeros_like(u)
    cf = lambda j: 0.5 * (p.cos_lat[j] + p.cos_lat[(j + 1) % ny])  # noqa: E731
    for i in range(nx):
        ip, im = (i + 1) % nx, (i - 1) % nx
        for j in range(ny):
            for k in range(nz):
                fe = 0.5 * (u[i, j, k] + u[ip, j, k]) * wm[i, j, k] * wm[ip, j, k]
                fw = 0.5 * (u[im, j, k] + u[i, j, k]) * wm[im, j, k] * wm[i, j, k]
                fn = 0.0 if j == ny - 1 else 0.5 * (v[i, j, k] + v[i, j + 1, k]) * wm[i, j, k] * wm[i, j + 1, k] * cf(j)
                fs = 0.0 if j == 0 else 0.5 * (v[i, j - 1, k] + v[i, j, k]) * wm[i, j - 1, k] * wm[i, j, k] * cf(j - 1)
                out[i, j, k] = (fe - fw) / p.dx_2d[i, j] + (fn - fs) / (p.dy * p.cos_lat[j])
    return out


def test_divergence_matches_index_loop():
    _, p, _, _ = _land_params()
    rng = np.random.default_rng(9)
    u, v = rng.standard_normal(p.wet_mask_z.shape), rng.standard_normal(p.wet_mask_z.shape)
    np.testing.assert_allclose(O.divergence_h(u, v, p), _loop_divergence(u, v, p), rtol=1e-11, atol=1e-18)


def test_gradient_is_negative_adjoint_of_2d_divergence():
    # D3 (reference horizontal.py:99-105): area-weighted adjoint pair
    _, p, _, _ = _land_params()
    rng = np.random.default_rng(10)
    shape = p.wet_mask.shape
    u, v, eta = rng.standard_normal(shape), rng.standard_normal(shape), rng.standard_normal(shape)
    area = p.dx_2d * p.dy
    wm = p.wet_mask
    lhs = np.sum(area * eta * wm * O.divergence_conservative(u * wm, v * wm, p))
    gx, gy = O.gradient_conservative(eta * wm, p)
    rhs = -np.sum(area * (u * wm * gx + v * wm * gy))
    assert lhs == pytest.approx(rhs, rel=1e-11)


def test_2d_divergence_volume_conservation():
    _, p, _, _ = _land_params()
    rng = np.random.default_rng(11)
    shape = p.wet_mask.shape
    div = O.divergence_conservative(rng.standard_normal(shape) * p.wet_mask,
                                    rng.standard_normal(shape) * p.wet_mask, p)
    area = p.dx_2d * p.dy
    assert abs(np.sum(area * div)) <= 1e-12 * np.sum(area * np.abs(div))


def test_dealias_matches_direct_dft_projection():
    _, p, _, _ = _land_params()
    nx = p.nx
    f = np.random.default_rng(12).standard_normal((nx, 3, 2))
    keep = p.dealias_keep
    n = np.arange(nx)
    proj = np.zeros((nx, nx))
    for k in range(-(nx // 2), nx - nx // 2):
        if abs(k) <= keep:
            proj += np.cos(2.0 * np.pi * k * (n[:, None] - n[None, :]) / nx) / nx
    flon = np.einsum("ab,bjk->ajk", proj, f)
    fp = np.concatenate([flon[:, :1], flon[:, :1], flon, flon[:, -1:], flon[:, -1:]], axis=1)
    direct = (fp[:, :-4] + 4 * fp[:, 1:-3] + 6 * fp[:, 2:-2] + 4 * fp[:, 3:-1] + fp[:, 4:]) / 16.0
    np.testing.assert_allclose(O.dealias(f, p), direct, rtol=1e-12, atol=1e-13)


def test_dealias_projection_keeps_low_and_removes_high_modes():
    _, p, _, _ = _land_params()
    nx = p.nx
    x = 2.0 * np.pi * np.arange(nx) / nx
    const_rows = lambda a: np.broadcast_to(a[:, None, None], (nx, p.ny, 1)).copy()  # noqa: E731
    low = const_rows(np.cos(p.dealias_keep * x))
    high = const_rows(np.cos((p.dealias_keep + 1) * x))
    np.testing.assert_allclose(O.dealias(low, p), low, atol=1e-12)
    np.testing.assert_allclose(O.dealias(high, p), 0.0, atol=1e-12)


def test_trapezoid_pressure_exact_for_linear_density():
    # rho'(z) = b z on every wet node -> p_bc(z_k) = g * int_{z_k}^0 rho' dz = -g b z_k^2 / 2
    _, p, _, _ = _land_params()
    z = p.z
    T = np.broadcast_to(p.T_ref + 0.3 * z, p.wet_mask_z.shape).copy()
    S = np.full_like(T, p.S_ref)
    pbc = O.baroclinic_pressure(T, S, p)
    b = O.RHO_0 * (-O.ALPHA_T) * 0.3
    closed = -O.G_EARTH * b * z ** 2 / 2.0
    deep = p.wet_mask_z.sum(-1) == p.nz
    np.testing.assert_allclose(pbc[deep], np.broadcast_to(closed, pbc[deep].shape), rtol=1e-12, atol=1e-9)


def test_coriolis_rotation_is_exact_rotation():
    rng = np.random.default_rng(13)
    u, v = rng.standard_normal((5, 4, 3)), rng.standard_normal((5, 4, 3))
    f = rng.uniform(-1e-4, 1e-4, (5, 4))
    ur, vr = O.coriolis_rotation(u, v, f, 600.0)
    np.testing.assert_allclose(ur ** 2 + vr ** 2, u ** 2 + v ** 2, rtol=1e-13)
    back_u, back_v = O.coriolis_rotation(ur, vr, f, -600.0)
    np.testing.assert_allclose(back_u, u, atol=1e-13)
    np.testing.assert_allclose(back_v, v, atol=1e-13)


def test_surface_heat_matches_column_budget():
    # reference tests/fd/test_surface_energy.py:15-20 (depth=None branch)
    grid = all_wet_grid(nx=8, ny=8, nz=4)
    from dataclasses import replace
    grid = replace(grid, z=np.array([0.0, -5.0, -20.0, -50.0]), dz=np.array([5.0, 15.0, 30.0]))
    forcing = (np.zeros((8, 8)), np.zeros((8, 8)), np.full((8, 8), 100.0))
    p = O.make_params(grid, quiet_physics(), 3600.0, forcing=forcing, polar_cap_rows=0, polar_cap_taper=0)
    shape = (8, 8, 4)
    z3 = np.zeros(shape)
    dT, _, _ = O.tracer_tendency(z3, z3, np.full(shape, -1.8), np.full(shape, 35.0), p)
    flux = np.sum(dT * p.dz_node, axis=-1) * O.RHO_0 * O.C_P
    np.testing.assert_allclose(flux, 100.0, rtol=1e-12)


def test_convection_conserves_column_content_and_mixes_unstable_columns():
    _, p, T, S = _land_params()
    T = T.copy()
    T[..., 1] += 8.0  # denser-above-lighter inversions in every deep column
    gate = O.convective_mask(T, S, p)
    tend = O.conv_flux_tendency(T, gate, p.kappa_conv, p)
    col = np.sum(tend * p.dz_node, axis=-1)
    assert np.all(np.abs(col) <= 1e-12 * np.sum(np.abs(tend) * p.dz_node, axis=-1) + 1e-300)
    assert gate.sum() > 0


# ---------------------------------------------------------------------------
# sub-cycle counts (reference tests/fd/test_mode_split_smoke.py:72-80,
# tests/fd/test_nu_nsub_cfl.py:89-99)
# ---------------------------------------------------------------------------
def test_mode_split_counts():
    grid = all_wet_grid()
    physics = quiet_physics(nu_h=3e6, kappa_h=100.0, kappa_v=1e-5, kappa_conv=0.05)
    p = O.make_params(grid, physics, 3600.0, dt_bt=150.0, polar_cap_rows=0, polar_cap_taper=0)
    assert p.n_subcyc == 24
    assert p.conv_nsub == 1
    assert p.dt_bt == 150.0


def test_nu_nsub_cfl_is_15_at_one_degree():
    from dataclasses import replace
    nx, ny = 360, 120
    lat = np.linspace(-59.5, 59.5, ny)
    grid = replace(all_wet_grid(nx=8, ny=8, nz=4), nx=nx, ny=ny, lat=lat,
                   cos_lat=np.cos(np.radians(lat)),
                   dx_2d=np.broadcast_to(R_EARTH * np.radians(1.0) * np.cos(np.radians(lat)), (nx, ny)).copy(),
                   dy=R_EARTH * np.radians(1.0), f=np.zeros((nx, ny)), wet_mask=np.ones((nx, ny)),
                   wet_mask_3d=np.ones((nx, ny, 4)), z=np.array([0.0, -5.0, -20.0, -50.0]),
                   dz=np.array([5.0, 15.0, 30.0]), nz=4)
    p = O.make_params(grid, quiet_physics(nu_h=5e6), 3600.0, dt_bt=150.0, nu_nsub="cfl",
                      polar_cap_rows=0, polar_cap_taper=0)
    assert p.nu_nsub == 15
