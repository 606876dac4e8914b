"""Fixture grids shared by the oracle and GPU parity tests.

They restate the reference's test fixtures (same shapes, latitude bands, land
blocks and level sets) so that the reference's own assertions can be checked
against the oracle:
  * ``diffusion_fixture``: ``tests/support/fd/horizontal_diffusion.py:12-35``
  * ``surface_fixture``:   ``tests/support/fd/surface_energy.py:14-27``
  * ``all_wet_grid``:      ``tests/support/grid.py:12-35``
"""

from dataclasses import replace

import numpy as np

from oracle import zm_oracle as O
from zhenmode_amd.config import OMEGA, R_EARTH, PhysicsConfig
from zhenmode_amd.grid import all_wet_grid


def quiet_physics(**kw):
    base = dict(nu_h=0.0, nu_v=0.0, nu_bi=0.0, kappa_h=0.0, kappa_v=0.0, kappa_bi=0.0,
                kappa_conv=0.0, kappa_gm=0.0, kappa_redi=0.0)
    base.update(kw)
    return replace(PhysicsConfig(), **base)


def diffusion_grid(latitude_limit=30.0, land=False, ny=32, nx=24):
    grid = all_wet_grid(nx=nx, ny=ny, nz=4)
    lat = np.linspace(-latitude_limit, latitude_limit, ny)
    cos = np.cos(np.radians(lat))
    wet = np.ones((nx, ny, 4))
    if land:
        wet[4:9, ny // 2:ny // 2 + 5, :] = 0.0
        wet[12:17, ny // 4:ny // 4 + 5, 2:] = 0.0
    return replace(grid, lat=lat, cos_lat=cos,
                   dx_2d=np.broadcast_to(R_EARTH * np.radians(360.0 / nx) * cos, (nx, ny)).copy(),
                   dy=R_EARTH * np.radians(lat[1] - lat[0]),
                   f=np.broadcast_to(2.0 * OMEGA * np.sin(np.radians(lat)), (nx, ny)).copy(),
                   z=np.array([0.0, -5.0, -20.0, -50.0]), dz=np.array([5.0, 15.0, 30.0]),
                   wet_mask_3d=wet, wet_mask=wet[:, :, 0], ocean_mask=wet[:, :, 0].astype(bool),
                   land_mask=1.0 - wet[:, :, 0])


def diffusion_fixture(latitude_limit=30.0, land=False, ny=32, nx=24):
    grid = diffusion_grid(latitude_limit, land, ny, nx)
    physics = quiet_physics(kappa_h=100.0)
    p = O.make_params(grid, physics, 60.0, forcing=tuple(np.zeros((nx, ny)) for _ in range(3)),
                      polar_cap_rows=0, polar_cap_taper=0)
    volume = grid.dx_2d[:, :, None] * grid.dy * p.dz_node * grid.wet_mask_3d
    return grid, p, volume


def land_grid(nx=24, ny=20, nz=6, seed=7):
    """Small grid with land columns, shallow columns and wall rows."""
    rng = np.random.default_rng(seed)
    lat = np.linspace(-70.0, 70.0, ny)
    cos = np.cos(np.radians(lat))
    z = np.array([0.0, -10.0, -30.0, -70.0, -150.0, -300.0])[:nz]
    depth = rng.uniform(-60.0, 320.0, (nx, ny))
    depth[depth < 5.0] = 0.0
    depth[3:7, 5:9] = 0.0
    ocean = depth > 0.0
    wet3 = ((np.abs(z)[None, None, :] <= depth[:, :, None]) & ocean[:, :, None]).astype(float)
    grid = all_wet_grid(nx=nx, ny=ny, nz=nz)
    return replace(grid, lat=lat, cos_lat=cos,
                   dx_2d=np.broadcast_to(R_EARTH * np.radians(360.0 / nx) * cos, (nx, ny)).copy(),
                   dy=R_EARTH * np.radians(lat[1] - lat[0]),
                   f=np.broadcast_to(2.0 * OMEGA * np.sin(np.radians(lat)), (nx, ny)).copy(),
                   z=z, dz=-np.diff(z), depth=depth, wet_mask_3d=wet3, wet_mask=wet3[:, :, 0],
                   ocean_mask=ocean, land_mask=~ocean)


def land_case(seed=7, dt=600.0, ny=20, nx=24, nz=6):
    """A land_grid with every process of the legacy step switched on."""
    grid = land_grid(nx=nx, ny=ny, nz=nz, seed=seed)
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    physics = replace(PhysicsConfig(), nu_h=2.0e5, nu_v=1e-3, kappa_h=500.0, kappa_v=1e-4,
                      kappa_conv=0.5, nu_bi=1.0e15, kappa_bi=1.0e15, r_bot=2e-3)
    rng = np.random.default_rng(seed + 1)
    lat = np.radians(grid.lat)[None, :]
    tau_x = np.broadcast_to(0.1 * np.cos(3.0 * lat), (nx, ny)).copy()
    tau_y = 0.02 * rng.standard_normal((nx, ny))
    q = 50.0 * rng.standard_normal((nx, ny))
    zz = grid.z[None, None, :]
    T = 4.0 + 20.0 * np.exp(zz / 100.0) + 0.5 * rng.standard_normal((nx, ny, nz))
    S = 34.7 + 0.3 * np.tanh(zz / 100.0) + 0.05 * rng.standard_normal((nx, ny, nz))
    sst = 10.0 + 15.0 * np.cos(lat) ** 2 + 0.2 * rng.standard_normal((nx, ny))
    sss = 34.5 + 0.2 * rng.standard_normal((nx, ny))
    kwargs = dict(forcing=(tau_x, tau_y, q), T_atm=sst + 1.0, lambda_bulk=20.0,
                  S_ref_surf=sss, sss_restore_days=30.0,
                  coastal_restore_mask=grid.wet_mask.copy(), coastal_restore_days=20.0,
                  coastal_restore_T=sst, coastal_bulk_mask=grid.wet_mask.copy(), coastal_bulk_lambda=5.0,
                  coastal_kappa_h_mask=(rng.uniform(size=(nx, ny)) > 0.7).astype(float), coastal_kappa_h=300.0,
                  coastal_kappa_v_mask=(rng.uniform(size=(nx, ny)) > 0.7).astype(float), coastal_kappa_v=1e-3,
                  polar_cap_rows=2, polar_cap_taper=2, dt_bt=dt / 12)
    return grid, physics, dt, kwargs, T, S


def oracle_params(grid, physics, dt, kwargs):
    return O.make_params(grid, physics, dt, **{k: v for k, v in kwargs.items()
                                                 if k not in ("mode_split", "use_scan")})


def perturbed_state(p, T, S, seed=3, amp=0.05):
    """init_state plus random velocities and eta (so every term is exercised)."""
    rng = np.random.default_rng(seed)
    s = O.init_state(p, T, S)
    wm = p.wet_mask_z
    u = amp * rng.standard_normal(wm.shape) * wm
    v = amp * rng.standard_normal(wm.shape) * wm
    eta = 0.1 * rng.standard_normal(wm.shape[:2]) * p.wet_mask
    return O.State(u, v, s.T, s.S, eta, s.ice)


def surface_fixture(heat=100.0, **kwargs):
    """8x8x4 all-wet column set of ``tests/support/fd/surface_energy.py:14-27``
    (z = 0, -5, -20, -50 m; every diffusivity and viscosity zero; dt 3600 s;
    no polar cap) with T = -1.8, S = 35 and the state at rest.  Extra keyword
    arguments go to ``make_params`` (bulk heat, restoring)."""
    grid = all_wet_grid(nx=8, ny=8, nz=4)
    grid = replace(grid, z=np.array([0.0, -5.0, -20.0, -50.0]), dz=np.array([5.0, 15.0, 30.0]))
    forcing = (np.zeros((8, 8)), np.zeros((8, 8)), np.full((8, 8), float(heat)))
    p = O.make_params(grid, quiet_physics(), 3600.0, forcing=forcing, polar_cap_rows=0, polar_cap_taper=0,
                      **kwargs)
    shape = (8, 8, 4)
    state = O.State(np.zeros(shape), np.zeros(shape), np.full(shape, -1.8), np.full(shape, 35.0),
                    np.zeros((8, 8)), np.zeros((8, 8)))
    return grid, p, state


def planted(fn, old, new):
    """A copy of oracle function ``fn`` with the source text ``old`` replaced
    by ``new``: a planted mistake, used to show that a check fails on it.
    The copy resolves names in a snapshot of ``fn``'s module namespace."""
    import inspect
    import textwrap

    src = textwrap.dedent(inspect.getsource(fn))
    assert src.count(old) == 1, f"{old!r} must occur once in {fn.__name__}"
    namespace = dict(vars(inspect.getmodule(fn)))
    exec(compile(src.replace(old, new), f"<planted {fn.__name__}>", "exec"), namespace)
    return namespace[fn.__name__]
