"""Global lat-lon ocean grid: record type, construction and validation.

``GlobalOceanGrid`` carries the same fields as the reference
(``src/zhenmode/model/solver/geometry/grid.py:13-61``); ``build_global_grid``
follows the reference's metric and mask conventions
(``grid.py:291-380``): periodic longitude, dx = R cos(lat) dlon,
f = 2 Omega sin(lat), and a level k wet iff |z_k| <= depth in an ocean column.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .config import OMEGA, R_EARTH, finite_number, integer_count


@dataclass
class GlobalOceanGrid:
    lon: np.ndarray
    lat: np.ndarray
    dx_2d: np.ndarray
    dy: float
    cos_lat: np.ndarray
    f: np.ndarray
    z: np.ndarray
    dz: np.ndarray
    nz: int
    depth: np.ndarray
    wet_mask: np.ndarray
    ocean_mask: np.ndarray
    land_mask: np.ndarray
    wet_mask_3d: np.ndarray
    nx: int
    ny: int


def build_global_grid(z_levels, depth, lat_max, *, dy=None):
    """Metrics and masks for a bathymetry ``depth`` of shape (nx, ny).

    ``lat = linspace(-lat_max, lat_max, ny)``, ``lon_i = (i + 1/2) 360/nx``;
    ``dy`` defaults to the actual latitude spacing (as the reference's
    ``tests/support/grid.py:18-24`` fixtures do).
    """
    depth = np.asarray(depth, dtype=np.float64)
    nx, ny = depth.shape
    z = np.asarray(z_levels, dtype=np.float64)
    lat = np.linspace(-lat_max, lat_max, ny)
    lon = (np.arange(nx) + 0.5) * (360.0 / nx)
    cos_lat = np.cos(np.radians(lat))
    dx_2d = np.broadcast_to(R_EARTH * np.radians(360.0 / nx) * cos_lat, (nx, ny)).copy()
    if dy is None:
        dy = R_EARTH * np.radians(lat[1] - lat[0])
    f = np.broadcast_to(2.0 * OMEGA * np.sin(np.radians(lat))[None, :], (nx, ny)).copy()
    ocean = depth > 0.0
    wet_mask = ocean.astype(np.float64)
    wet3 = ((np.abs(z)[None, None, :] <= depth[:, :, None]) & ocean[:, :, None]).astype(np.float64)
    return GlobalOceanGrid(lon=lon, lat=lat, dx_2d=dx_2d, dy=float(dy), cos_lat=cos_lat, f=f,
                           z=z, dz=np.abs(np.diff(z)), nz=int(z.size), depth=depth,
                           wet_mask=wet_mask, ocean_mask=ocean, land_mask=~ocean,
                           wet_mask_3d=wet3, nx=int(nx), ny=int(ny))


def validate_grid(grid):
    """Reject a grid record whose sizes or metrics the step cannot use (the
    reference checks the same properties, grid.py:384-398): sizes are integers
    of at least 2, dy is positive, dx_2d has the (nx, ny) shape and positive
    finite entries, z has nz finite strictly descending nodes, and dz has the
    nz - 1 positive gaps between them."""
    for name in ("nx", "ny", "nz"):
        integer_count(name, getattr(grid, name), minimum=2)
    finite_number("dy", grid.dy, positive=True)
    dx, z, dz = (np.asarray(getattr(grid, name), dtype=np.float64) for name in ("dx_2d", "z", "dz"))
    checks = (
        ("dx_2d", dx.shape == (grid.nx, grid.ny) and bool(np.all(np.isfinite(dx))) and bool(np.all(dx > 0.0)),
         "an (nx, ny) array of positive finite spacings"),
        ("z", z.shape == (grid.nz,) and bool(np.all(np.isfinite(z))) and bool(np.all(z[1:] < z[:-1])),
         "nz finite nodes in strictly descending order"),
        ("dz", dz.shape == (grid.nz - 1,) and z.shape == (grid.nz,)
         and bool(np.all(np.abs(dz - (z[:-1] - z[1:])) <= 1e-12 * np.abs(z[:-1] - z[1:]))) and bool(np.all(dz > 0.0)),
         "the nz - 1 positive distances between consecutive z nodes"),
    )
    for name, ok, expected in checks:
        if not ok:
            raise ValueError(f"grid.{name} must be {expected}")


def all_wet_grid(nx=32, ny=32, nz=8):
    """Small fully-wet grid, same construction as the reference test fixture
    ``tests/support/grid.py:12-35`` (depth 4000 m, z = -linspace(50, 4000))."""
    lat = np.linspace(-30.0, 30.0, ny)
    lon = np.linspace(0.5, 359.5, nx)
    cos_lat = np.cos(np.radians(lat))
    dx_2d = np.broadcast_to(6.371e6 * cos_lat * np.radians(360.0 / nx), (nx, ny)).copy()
    dy = 6.371e6 * np.radians(abs(lat[1] - lat[0]))
    f = np.broadcast_to(2 * 7.2921e-5 * np.sin(np.radians(lat)), (nx, ny)).copy()
    z = -np.linspace(50.0, 4000.0, nz)
    return GlobalOceanGrid(lon=lon, lat=lat, dx_2d=dx_2d, dy=float(dy), cos_lat=cos_lat, f=f,
                           z=z, dz=-np.diff(z), nz=nz, depth=np.full((nx, ny), 4000.0),
                           wet_mask=np.ones((nx, ny)), ocean_mask=np.ones((nx, ny), dtype=bool),
                           land_mask=np.zeros((nx, ny)), wet_mask_3d=np.ones((nx, ny, nz)),
                           nx=nx, ny=ny)
