"""Host-built projection operator (zm_projection_operator) against the oracle.

The nine taps per column must reproduce the oracle's
-column_divergence(grad_c3d(psi)) * area (reference projection.py:85-87) as a
dense matrix, impulse by impulse, on grids with land, shallow columns, wall
rows and the periodic seam (including nx = 2 and 3, where the taps at
distance two wrap onto nearer columns); the diagonal must be the oracle's
_column_projection_diagonal (projection.py:10-50).  No device work.
"""

import ctypes

import numpy as np
import pytest

from oracle import zm_oracle as O
from tests.fixtures import land_case, land_grid, oracle_params, quiet_physics
from zhenmode_amd import _lib
from zhenmode_amd.params import freeze, make_config

# (dx, dy) of each tap: centre, east, west, east+2, west+2, north, south, north+2, south+2
OFFSETS = ((0, 0), (1, 0), (-1, 0), (2, 0), (-2, 0), (0, 1), (0, -1), (0, 2), (0, -2))


def host_operator(p):
    cfg, keep = make_config(p)
    n2 = p.nx * p.ny
    taps = np.zeros(9 * n2)
    diag = np.zeros(n2)
    lib = _lib.load()
    rc = lib.zm_projection_operator(ctypes.byref(cfg), taps.ctypes.data_as(ctypes.c_void_p),
                                    diag.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    del keep
    return taps.reshape(9, p.ny, p.nx), diag.reshape(p.ny, p.nx)


def dense_from_taps(taps):
    _, ny, nx = taps.shape
    A = np.zeros((ny * nx, ny * nx))
    for t, (ox, oy) in enumerate(OFFSETS):
        for j in range(ny):
            for i in range(nx):
                jj = j + oy
                if 0 <= jj < ny:
                    A[j * nx + i, jj * nx + (i + ox) % nx] += taps[t, j, i]
                else:
                    assert taps[t, j, i] == 0.0
    return A


def dense_from_oracle(q):
    n = q.nx * q.ny
    A = np.zeros((n, n))
    for col in range(n):
        j, i = divmod(col, q.nx)
        e = np.zeros((q.nx, q.ny))
        e[i, j] = 1.0
        A[:, col] = O.projection_matvec(e, q).T.reshape(-1)
    return A


def case_params(nx, ny, nz=6, seed=7):
    if nx == 24 and ny == 20:
        grid, physics, dt, kwargs, _, _ = land_case(seed=seed)
    else:
        grid = land_grid(nx=nx, ny=ny, nz=nz, seed=seed)
        physics, dt = quiet_physics(), 600.0
        kwargs = dict(forcing=tuple(np.zeros((nx, ny)) for _ in range(3)), polar_cap_rows=0, polar_cap_taper=0)
    p = freeze(grid, physics, dt, mode_split=True, **kwargs)
    q = oracle_params(grid, physics, dt, kwargs)
    return p, q


@pytest.mark.parametrize("nx,ny", [(24, 20), (3, 9), (2, 7), (5, 6)])
def test_taps_are_the_oracle_operator(nx, ny):
    p, q = case_params(nx, ny)
    taps, diag = host_operator(p)
    A = dense_from_taps(taps)
    B = dense_from_oracle(q)
    scale = np.abs(B).max()
    assert scale > 0.0
    np.testing.assert_allclose(A, B, rtol=0.0, atol=1e-13 * scale)
    np.testing.assert_allclose(diag, O.projection_diagonal(q).T, rtol=1e-13, atol=1e-300)
    if nx >= 5:  # without wrap-around aliasing the stated diagonal is the operator's
        np.testing.assert_allclose(np.diag(A), diag.reshape(-1), rtol=1e-12, atol=1e-14 * scale)


def test_operator_rows_vanish_on_land_and_taps_sum_to_zero():
    p, q = case_params(24, 20)
    taps, _ = host_operator(p)
    land = q.wet_mask.T == 0.0
    assert np.all(taps[:, land] == 0.0)
    # A kills constants: the taps of every row sum to zero
    scale = np.abs(taps).max()
    np.testing.assert_allclose(taps.sum(axis=0), 0.0, atol=1e-13 * scale)
