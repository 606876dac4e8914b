"""Seeded synthetic cases for the mode-split step (no data files needed).

Four single-grid workloads and one ensemble, all on the reference's collocated lat-lon grid:

=======  ===========================  =====  ======  =====  ======================
name     grid (nx x ny x nz)          lat    dt [s]  n_sub  forcing
=======  ===========================  =====  ======  =====  ======================
g4deg    90 x 40 x 12, H = 4000 m     +-78      600     30  wind
g1deg    360 x 160 x 50, H = 5500 m   +-80     1200     40  wind + SST/SSS restore
g05deg   720 x 320 x 50, H = 5500 m   +-80      900     45  seasonal wind + restore
g045deg  800 x 356 x 50, H = 5500 m   +-80      900     45  seasonal wind + restore
=======  ===========================  =====  ======  =====  ======================

g045deg runs the reference's ``monotone_adv`` tracer advection: with the
centred scheme its T grows without bound at one 72 m deep coastal corner cell
from about day 7 (the plain and the Thomas-mixing steps alike); with
``monotone_adv``, ``fct_adv`` or ``project_adv_vel`` it stays bounded over 30
days.

The 24-member uneven ensemble (2, 1 and 0.75 degrees, dt = 1200 s, 40
substeps, per-member kappa_v draw) comes from ``ensemble_members``.

Levels: z_0 = 0 and spacings 10 r^k with r chosen so that z_{nz-1} = -H.
Bathymetry: 20 Gaussian elevation bumps on a flat abyssal plain of depth H,
``depth = clip(H - s * sum_i A_i exp(-d_i^2 / (2 w_i^2)), 0, H)`` with
A ~ U(500, 4000) m, w ~ U(10, 40) deg, centres uniform on the sphere within the
latitude band, s = H / 4000 (so every config shares one coastline pattern);
columns shallower than 50 m are land.
Initial state: T = 2 + 23 exp(z/500) + N(0, 0.05), S = 34.7 + 0.3 tanh(z/300) +
N(0, 0.01), u = v = eta = 0.  Wind: tau_x = 0.1 cos(3 lat) N/m^2.
Restoring targets (g1deg, g05deg, g045deg): SST* = -1 + 28 cos^2(lat) + N(0, 0.2),
SSS* = 34.7 + 0.5 cos(2 lat) + N(0, 0.02), 30-day timescale, applied through the
reference factory's ``S_ref_surf``/``sss_restore_days`` and
``coastal_restore_{mask,days,T}`` arguments (``factory.py:424-454``).
Physics: kappa_h = 100, nu_v = 1e-4, kappa_v = 1e-5, kappa_conv = 0.05,
r_bot = 1e-3, nu_h = 5e6 res^2 (reference ``runtime/cli.py:50-61``), and
nu_bi = kappa_bi = half the explicit biharmonic bound
``2 / (dt/2 * lambda^2)`` with lambda = 4 (1/dx^2 + 1/dy^2) taken over every
row that the polar cap does not fully average (rows ncap .. ny-1-ncap; the
taper rows keep part of their zonal structure, so they need the bound too).
This is the D31 bound of the reference's ``docs/decisions.md:125-127``.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from .config import PhysicsConfig
from .grid import build_global_grid

BATHYMETRY_SEED = 5
FIELD_SEED = 20261203


def stretched_levels(nz, depth, top=10.0):
    """z_0 = 0, z_k = -sum_{m<k} top r^m with r solved so that z_{nz-1} = -depth."""
    n = nz - 1
    lo, hi = 1.0 + 1e-12, 10.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        total = top * (r ** n - 1.0) / (r - 1.0)
        if total > depth:
            hi = r
        else:
            lo = r
    r = 0.5 * (lo + hi)
    spacing = top * r ** np.arange(n)
    spacing *= depth / spacing.sum()
    z = np.concatenate([[0.0], -np.cumsum(spacing)])
    z[-1] = -float(depth)
    return z


def gaussian_bathymetry(nx, ny, lat_max, depth_max, seed=BATHYMETRY_SEED, nbumps=20):
    rng = np.random.default_rng(seed)
    amp = rng.uniform(500.0, 4000.0, nbumps) * (depth_max / 4000.0)
    width = np.radians(rng.uniform(10.0, 40.0, nbumps))
    lon_c = rng.uniform(0.0, 2.0 * np.pi, nbumps)
    s = np.sin(np.radians(lat_max))
    lat_c = np.arcsin(rng.uniform(-s, s, nbumps))
    lat = np.radians(np.linspace(-lat_max, lat_max, ny))
    lon = np.radians((np.arange(nx) + 0.5) * 360.0 / nx)
    lo, la = np.meshgrid(lon, lat, indexing="ij")
    relief = np.zeros((nx, ny))
    for a, w, lc, phc in zip(amp, width, lon_c, lat_c):
        cosang = np.sin(la) * np.sin(phc) + np.cos(la) * np.cos(phc) * np.cos(lo - lc)
        ang = np.arccos(np.clip(cosang, -1.0, 1.0))
        relief += a * np.exp(-ang ** 2 / (2.0 * w ** 2))
    depth = np.clip(depth_max - relief, 0.0, depth_max)
    depth[depth < 50.0] = 0.0
    return depth


def biharmonic_bound(grid, dt, rows_capped):
    """Largest kappa_bi with kappa_bi dt/2 lambda^2 <= 2 on the uncapped rows."""
    j0, j1 = rows_capped, grid.ny - rows_capped
    inv_dx2 = float(np.max(1.0 / grid.dx_2d[:, j0:j1] ** 2))
    lam = 4.0 * (inv_dx2 + 1.0 / grid.dy ** 2)
    return 2.0 / (0.5 * dt * lam ** 2)


@dataclass
class Case:
    """Everything ``make_solver_global`` and a run loop need for one workload."""
    name: str
    grid: object
    physics: PhysicsConfig
    dt: float
    steps: int
    solver_kwargs: dict = field(default_factory=dict)
    T_init: np.ndarray = None
    S_init: np.ndarray = None
    seasonal_wind: bool = False
    land_fraction: float = 0.0

    def wind(self, t_seconds):
        """(tau_x, tau_y, q_heat) at time t; seasonal amplitude for g05deg/g045deg."""
        lat = np.radians(self.grid.lat)
        amp = 0.1 + (0.05 * np.sin(2.0 * np.pi * t_seconds / (365.0 * 86400.0)) if self.seasonal_wind else 0.0)
        nx, ny = self.grid.nx, self.grid.ny
        tau_x = np.broadcast_to((amp * np.cos(3.0 * lat))[None, :], (nx, ny)).copy()
        return tau_x, np.zeros((nx, ny)), np.zeros((nx, ny))


_SPECS = {
    "g4deg": dict(seed_offset=147, nx=90, ny=40, nz=12, lat=78.0, H=4000.0, dt=600.0, nsub=30, steps=8, restore=False, seasonal=False),
    "g1deg": dict(seed_offset=148, nx=360, ny=160, nz=50, lat=80.0, H=5500.0, dt=1200.0, nsub=40, steps=72, restore=True, seasonal=False),
    "g05deg": dict(seed_offset=149, nx=720, ny=320, nz=50, lat=80.0, H=5500.0, dt=900.0, nsub=45, steps=2880, restore=True, seasonal=True),
    "g045deg": dict(seed_offset=151, nx=800, ny=356, nz=50, lat=80.0, H=5500.0, dt=900.0, nsub=45, steps=2880, restore=True, seasonal=True,
                    monotone=True),
}


def make_case(name, *, nx=None, ny=None, nz=None, seed=FIELD_SEED):
    """Build a named case; nx/ny/nz override the sizes (same generator)."""
    spec = dict(_SPECS[name])
    nx = nx or spec["nx"]
    ny = ny or spec["ny"]
    nz = nz or spec["nz"]
    z = stretched_levels(nz, spec["H"])
    depth = gaussian_bathymetry(nx, ny, spec["lat"], spec["H"])
    grid = build_global_grid(z, depth, spec["lat"])
    res = 360.0 / nx
    dt = spec["dt"]
    ncap, ntaper = 2, 3
    bound = biharmonic_bound(grid, dt, ncap)
    physics = replace(PhysicsConfig(), nu_h=5.0e6 * res ** 2, nu_v=1.0e-4, kappa_h=100.0,
                      kappa_v=1.0e-5, kappa_conv=0.05, nu_bi=0.5 * bound, kappa_bi=0.5 * bound,
                      r_bot=1.0e-3)
    rng = np.random.default_rng(seed + spec["seed_offset"])
    zz = z[None, None, :]
    T = 2.0 + 23.0 * np.exp(zz / 500.0) + rng.normal(0.0, 0.05, (nx, ny, nz))
    S = 34.7 + 0.3 * np.tanh(zz / 300.0) + rng.normal(0.0, 0.01, (nx, ny, nz))
    case = Case(name=name, grid=grid, physics=physics, dt=dt, steps=spec["steps"],
                T_init=T, S_init=S, seasonal_wind=spec["seasonal"],
                land_fraction=float(np.mean(grid.wet_mask == 0.0)))
    kwargs = dict(forcing=case.wind(0.0), polar_cap_rows=ncap, polar_cap_taper=ntaper,
                  mode_split=True, dt_bt=dt / spec["nsub"], use_scan=True,
                  T_init=T, S_init=S)
    if spec.get("monotone"):
        # the reference's monotone (limited) tracer advection option: with the
        # centred scheme a 72 m deep coastal corner cell of this bathymetry
        # (i = 441, j = 187) gains heat without bound after about a week
        kwargs["monotone_adv"] = True
    if spec["restore"]:
        lat = np.radians(grid.lat)[None, :]
        sst = -1.0 + 28.0 * np.cos(lat) ** 2 + rng.normal(0.0, 0.2, (nx, ny))
        sss = 34.7 + 0.5 * np.cos(2.0 * lat) + rng.normal(0.0, 0.02, (nx, ny))
        kwargs.update(S_ref_surf=sss, sss_restore_days=30.0,
                      coastal_restore_mask=grid.wet_mask.copy(), coastal_restore_days=30.0,
                      coastal_restore_T=sst)
    case.solver_kwargs = kwargs
    return case


def ensemble_members(seed=FIELD_SEED):
    """24 uneven members: 8 x (180x80x30), 8 x (360x160x50), 8 x (480x214x40).

    Each member has its own T perturbation N(0, 0.1) and kappa_v drawn
    log-uniform in [1e-5, 1e-4]; all share dt = 1200 s and n_sub = 40.
    """
    rng = np.random.default_rng(seed + 3)
    members = []
    for nx, ny, nz in ((180, 80, 30), (360, 160, 50), (480, 214, 40)):
        for _ in range(8):
            z = stretched_levels(nz, 5500.0)
            grid = build_global_grid(z, gaussian_bathymetry(nx, ny, 80.0, 5500.0), 80.0)
            dt = 1200.0
            bound = biharmonic_bound(grid, dt, 2)
            kv = float(np.exp(rng.uniform(np.log(1e-5), np.log(1e-4))))
            physics = replace(PhysicsConfig(), nu_h=5.0e6 * (360.0 / nx) ** 2, kappa_v=kv,
                              nu_bi=0.5 * bound, kappa_bi=0.5 * bound)
            zz = z[None, None, :]
            T = 2.0 + 23.0 * np.exp(zz / 500.0) + rng.normal(0.0, 0.1, (nx, ny, nz))
            S = 34.7 + 0.3 * np.tanh(zz / 300.0) + rng.normal(0.0, 0.01, (nx, ny, nz))
            case = Case(name=f"e{len(members)}", grid=grid, physics=physics, dt=dt, steps=360,
                        T_init=T, S_init=S, land_fraction=float(np.mean(grid.wet_mask == 0.0)))
            case.solver_kwargs = dict(forcing=case.wind(0.0), polar_cap_rows=2, polar_cap_taper=3,
                                      mode_split=True, dt_bt=30.0, use_scan=True, T_init=T, S_init=S)
            members.append(case)
    return members
