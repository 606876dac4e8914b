"""Pins the oracle's production tracer/mixing options (conservative_kv,
localize_conv, fct_adv, monotone_adv) with the assertions of the reference's
own tests and the identities its docstrings state:

  * tests/fd/test_fct_advection.py:44-113 and test_monotone_advection.py:48-68
    (uniform tracer -> zero tendency; no overshoot of a sharp front at Courant
    0.9; the south wall does not read the north wall; a wet face does not read
    a land sentinel);
  * numerics/vertical.py:317-323 (the interface-flux form telescopes: the
    column integral of the tendency is zero for any field, the node form does
    not on a stretched grid);
  * physics/vertical.py:13-37 (with the per-interface gate only unstable
    interfaces carry a convective flux).
"""

from dataclasses import replace

import numpy as np
import pytest

from oracle import zm_oracle as O
from tests.fixtures import diffusion_grid, land_case, oracle_params, quiet_physics


def flag_params(grid, **flags):
    nx, ny = grid.nx, grid.ny
    physics = quiet_physics(kappa_h=100.0, kappa_v=1e-5)
    return O.make_params(grid, physics, 60.0, forcing=tuple(np.zeros((nx, ny)) for _ in range(3)),
                         polar_cap_rows=0, polar_cap_taper=0, **flags)


@pytest.mark.parametrize("flag", ["fct_adv", "monotone_adv"])
def test_uniform_tracer_has_zero_advective_tendency(flag):
    grid = diffusion_grid()
    p = flag_params(grid, **{flag: True})
    nx, ny, nz = p.nx, p.ny, p.nz
    lon, lat = np.arange(nx), np.arange(ny)
    u = 0.03 * np.cos(2 * np.pi * lon[:, None, None] / nx) + 0.01 * np.cos(np.pi * lat[None, :, None] / ny)
    v = 0.02 * np.sin(np.pi * lat[None, :, None] / ny)
    u, v = np.broadcast_to(u, (nx, ny, nz)).copy(), np.broadcast_to(v, (nx, ny, nz)).copy()
    Fz = O.vertical_transport_iface(u, v, p)
    for value in (20.0, 35.0):
        a = O.advection_scalar(np.full((nx, ny, nz), value), u, v, Fz, p)
        np.testing.assert_allclose(a, 0.0, atol=1e-12)


@pytest.mark.parametrize("flag", ["fct_adv", "monotone_adv"])
def test_sharp_front_stays_bounded_at_courant_0_9(flag):
    grid = diffusion_grid()
    p = flag_params(grid, **{flag: True})
    nx, ny, nz = p.nx, p.ny, p.nz
    dt = 0.9 * float(np.min(grid.dx_2d)) / 1.0
    T = np.full((nx, ny, nz), 10.0)
    T[nx // 2:] = 25.0
    u, v = np.ones((nx, ny, nz)), np.zeros((nx, ny, nz))
    Fz = O.vertical_transport_iface(u, v, p)
    Tn = T + O.advection_scalar(T, u, v, Fz, p) * dt
    assert Tn.min() >= 10.0 - 1e-12 and Tn.max() <= 25.0 + 1e-12
    assert np.max(Tn[:nx // 2]) > 10.0 + 1e-8


def test_centred_face_values_overshoot_where_the_limiter_does_not():
    """The planted check: the default centred scheme does overshoot the same front."""
    grid = diffusion_grid()
    p = flag_params(grid)
    nx, ny, nz = p.nx, p.ny, p.nz
    dt = 0.9 * float(np.min(grid.dx_2d))
    T = np.full((nx, ny, nz), 10.0)
    T[nx // 2:] = 25.0
    u, v = np.ones((nx, ny, nz)), np.zeros((nx, ny, nz))
    Tn = T + O.advection_scalar(T, u, v, O.vertical_transport_iface(u, v, p), p) * dt
    assert Tn.min() < 10.0 - 1e-6 or Tn.max() > 25.0 + 1e-6


def test_fct_south_wall_does_not_read_north_wall():
    grid = diffusion_grid()
    p = flag_params(grid, fct_adv=True)
    nx, ny, nz = p.nx, p.ny, p.nz
    v = np.zeros((nx, ny, nz))
    v[:, 1] = 1.0
    Fz = np.zeros((nx, ny, nz + 1))
    T = np.broadcast_to(10.0 + 2.0 * np.arange(ny)[None, :, None], (nx, ny, nz)).copy()
    a1, a2 = T.copy(), T.copy()
    a1[:, -1] = 0.0
    a2[:, -1] = 30.0
    r1 = O.advection_scalar(a1, np.zeros_like(v), v, Fz, p)
    r2 = O.advection_scalar(a2, np.zeros_like(v), v, Fz, p)
    np.testing.assert_array_equal(r1[:, :2], r2[:, :2])


def test_fct_wet_face_does_not_read_land_sentinel():
    grid = diffusion_grid()
    p = flag_params(grid, fct_adv=True)
    nx, ny, nz = p.nx, p.ny, p.nz
    wet = np.ones((nx, ny, nz))
    wet[0] = 0.0
    p.wet_mask_z = wet
    p.wet_mask = wet[:, :, 0]
    u = np.ones((nx, ny, nz))
    u[0] = 0.0
    Fz = np.zeros((nx, ny, nz + 1))
    T = np.full((nx, ny, nz), 20.0)
    T[1], T[2] = 10.0, 12.0
    a1, a2 = T.copy(), T.copy()
    a1[0], a2[0] = 0.0, 15.0
    r1 = O.advection_scalar(a1, u, np.zeros_like(u), Fz, p)
    r2 = O.advection_scalar(a2, u, np.zeros_like(u), Fz, p)
    np.testing.assert_array_equal(r1[1:3], r2[1:3])


@pytest.fixture(scope="module")
def land():
    grid, physics, dt, kwargs, T, S = land_case()
    return grid, physics, dt, kwargs, T, S


def test_conservative_kv_telescopes_and_the_node_form_does_not(land):
    grid, physics, dt, kwargs, T, S = land
    p_flux = oracle_params(grid, physics, dt, dict(kwargs, conservative_kv=True))
    p_node = oracle_params(grid, physics, dt, kwargs)
    rng = np.random.default_rng(4)
    X = rng.standard_normal(T.shape) * p_flux.wet_mask_z
    for p, telescopes in ((p_flux, True), (p_node, False)):
        tend = O.vertical_diffusion(X, p)
        col = np.sum(tend * p.dz_node * p.wet_mask_z, axis=-1)
        scale = np.max(np.abs(tend * p.dz_node))
        if telescopes:
            assert np.max(np.abs(col)) <= 1e-12 * scale
        else:
            assert np.max(np.abs(col)) > 1e-6 * scale
    np.testing.assert_allclose(O.vertical_diffusion(np.full_like(X, 3.0), p_flux), 0.0, atol=1e-15)


def test_localized_convection_moves_only_unstable_interfaces(land):
    grid, physics, dt, kwargs, T, S = land
    p_loc = oracle_params(grid, physics, dt, dict(kwargs, localize_conv=True))
    p_col = oracle_params(grid, physics, dt, kwargs)
    nw = np.sum(p_loc.wet_mask_z > 0.5, axis=-1)
    i, j = np.argwhere(nw == p_loc.nz)[0]
    Tc = np.broadcast_to(np.linspace(20.0, 5.0, p_loc.nz), T.shape).copy() * p_loc.wet_mask_z
    Sc = np.full(T.shape, 35.0)
    Tc[i, j, 3] = Tc[i, j, 4] - 1.0  # colder (denser) above level 4: interface 3+1/2 unstable
    unstable = O.unstable_interfaces(Tc, Sc, p_loc)
    assert unstable[i, j].sum() == 1 and unstable[i, j, 3]
    gate = O.convective_mask(Tc, Sc, p_loc)
    loc = O.conv_flux_tendency(Tc, gate, p_loc.kappa_conv, p_loc, unstable)
    col = O.conv_flux_tendency(Tc, gate, p_col.kappa_conv, p_col)
    changed = np.nonzero(np.abs(loc[i, j]) > 0)[0]
    np.testing.assert_array_equal(changed, [3, 4])
    assert np.count_nonzero(np.abs(col[i, j]) > 0) > 2
    for c in (loc, col):
        assert abs(np.sum(c[i, j] * p_loc.dz_node)) <= 1e-12 * np.max(np.abs(c[i, j] * p_loc.dz_node))


def test_resting_uniform_ocean_is_a_fixed_point_with_every_flag(land):
    grid, physics, dt, kwargs, T, S = land
    flags = dict(conservative_kv=True, localize_conv=True, fct_adv=True)
    quiet = replace(physics, kappa_conv=0.5)
    nx, ny = grid.nx, grid.ny
    kw = dict(polar_cap_rows=2, polar_cap_taper=2, dt_bt=dt / 12, **flags)
    p = O.make_params(grid, quiet, dt, forcing=tuple(np.zeros((nx, ny)) for _ in range(3)), **kw)
    s = O.init_state(p, np.full(T.shape, 12.0), np.full(T.shape, 35.0))
    out = O.step(s, p)
    for f in ("u", "v", "eta"):
        np.testing.assert_allclose(getattr(out, f), 0.0, atol=1e-13)
    np.testing.assert_allclose(out.T, s.T, atol=1e-12)
    np.testing.assert_allclose(out.S, s.S, atol=1e-12)
