"""Pins the oracle's column-divergence projection (reference
dynamics/projection.py:10-127): the operator is the symmetric positive
semi-definite adjoint pair the docstring describes, its stated diagonal is the
operator's true diagonal, and the projection removes the column divergence
with a pure gradient correction."""

import numpy as np
import pytest

from oracle import zm_oracle as O
from tests.fixtures import land_case, oracle_params, perturbed_state


@pytest.fixture(scope="module")
def setup():
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, dict(kwargs, project_adv_vel=True, projection_niter=2000))
    s = perturbed_state(p, T, S)
    return p, s, dt


def test_operator_is_symmetric_psd_and_kills_constants(setup):
    p, s, dt = setup
    rng = np.random.default_rng(0)
    wet = p.wet_mask
    x, y = rng.standard_normal((2,) + wet.shape) * wet
    Ax, Ay = O.projection_matvec(x, p), O.projection_matvec(y, p)
    assert abs(np.vdot(Ax, y) - np.vdot(x, Ay)) <= 1e-12 * np.abs(Ax).max() * np.abs(y).sum()
    assert np.vdot(x, Ax) >= 0.0
    np.testing.assert_allclose(O.projection_matvec(np.ones_like(x), p), 0.0, atol=1e-15 * np.abs(Ax).max())


def test_stated_diagonal_is_the_true_diagonal(setup):
    p, s, dt = setup
    d = O.projection_diagonal(p)
    nx, ny = p.nx, p.ny
    brute = np.zeros((nx, ny))
    for i in range(nx):
        for j in range(ny):
            e = np.zeros((nx, ny))
            e[i, j] = 1.0
            brute[i, j] = O.projection_matvec(e, p)[i, j]
    np.testing.assert_allclose(d, brute, rtol=1e-12, atol=1e-15 * np.abs(brute).max())


def test_planted_sign_error_in_the_diagonal_is_caught(setup):
    """The check above fails for a diagonal that drops the cross term."""
    p, s, dt = setup
    wet = p.wet_mask_z
    pos_x = wet * O.east(wet)
    neg_x = wet * O.west(wet)
    wrong = 0.25 * p.inv_dx[..., :1] ** 2 * (pos_x ** 2 + neg_x ** 2)
    area = p.dx_2d * p.dy * p.wet_mask
    d_wrong = area * np.sum(wrong * p.dz_node, axis=-1)
    assert not np.allclose(d_wrong, O.projection_diagonal(p), rtol=1e-6)


@pytest.mark.parametrize("pre", ["none", "jacobi"])
def test_projection_removes_the_column_divergence(setup, pre):
    p, s, dt = setup
    p.projection_preconditioner = pre
    (u, v), its = O.project_column_divergence(s.u, s.v, p, dt, return_iterations=True)
    before = np.abs(O.column_divergence(s.u, s.v, p)).max()
    after = np.abs(O.column_divergence(u, v, p)).max()
    assert after <= 1e-9 * before, (after, before, its)
    # the correction is -dt g grad(psi) for some psi: zero on dry nodes
    dry = p.wet_mask_z < 0.5
    np.testing.assert_array_equal((u - s.u)[dry], 0.0)
    p.projection_preconditioner = "none"


def test_pcg_solves_a_dense_spd_system():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((30, 30))
    A = A @ A.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    x, its = O.pcg(lambda z: A @ z, b, 1e-13, 0.0, 500)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10)
    inv = 1.0 / np.diag(A)
    xj, _ = O.pcg(lambda z: A @ z, b, 1e-13, 0.0, 500, lambda r: r * inv)
    np.testing.assert_allclose(xj, np.linalg.solve(A, b), rtol=1e-10)


def test_resting_ocean_is_a_fixed_point_with_projection(setup):
    p, s, dt = setup
    s0 = O.init_state(p, np.full(s.T.shape, 12.0), np.full(s.T.shape, 35.0))
    q = oracle_params(*land_case()[:3], dict(land_case()[3], project_adv_vel=True,
                                             forcing=tuple(np.zeros(s.eta.shape) for _ in range(3)),
                                             lambda_bulk=0.0, sss_restore_days=0.0, coastal_restore_days=0.0,
                                             coastal_bulk_lambda=0.0))
    out = O.step(s0, q)
    np.testing.assert_allclose(out.u, 0.0, atol=1e-13)
    np.testing.assert_allclose(out.T, s0.T, atol=1e-12)
