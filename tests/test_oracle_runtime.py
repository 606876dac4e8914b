"""Closed-form pins of the run-time oracle (seasonal blend, classify, budget)."""

import numpy as np

from oracle import zm_oracle as O
from oracle import zm_runtime as R
from tests.fixtures import land_case, oracle_params


def stack(seed=0, shape=(3, 4)):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((12,) + shape)


def test_mid_month_returns_the_month_and_boundaries_average():
    s = stack()
    for m in range(12):
        np.testing.assert_array_equal(R.interp_monthly_field(s, 30.0 * m + 15.0), s[m])
        np.testing.assert_allclose(R.interp_monthly_field(s, 30.0 * m), 0.5 * (s[m] + s[(m - 1) % 12]),
                                   rtol=1e-15, atol=1e-15)
    np.testing.assert_array_equal(R.interp_monthly_field(s, 29.9, blend_days=0.0), s[0])


def test_blend_is_continuous_and_weights_sum_to_one():
    s = stack(1)
    for day in (60.0, 359.999, 720.0):
        a = R.interp_monthly_field(s, day - 1e-9)
        b = R.interp_monthly_field(s, day + 1e-9)
        np.testing.assert_allclose(a, b, atol=1e-8)
    const = np.full((12, 3, 4), 7.0)
    for day in np.linspace(0.0, 720.0, 97):
        np.testing.assert_allclose(R.interp_monthly_field(const, day), 7.0, rtol=1e-15)


def test_classify_codes():
    z = np.zeros((4, 3, 2))
    s = O.State(z.copy(), z.copy(), z + 10.0, z + 35.0, np.zeros((4, 3)), np.zeros((4, 3)))
    assert R.classify_state(s)["failure"] == 0
    u = z.copy()
    u[1, 1, 1] = -11.0
    assert R.classify_state(s._replace(u=u))["failure"] == 3
    e = np.zeros((4, 3))
    e[0, 0] = 15.0
    assert R.classify_state(s._replace(eta=e))["failure"] == 0
    assert R.classify_state(s._replace(eta=e), inclusive=True)["failure"] == 4
    T = s.T.copy()
    T[0, 0, 0] = np.nan
    assert R.classify_state(s._replace(T=T, u=u))["failure"] == 1


def test_budget_of_uniform_fields():
    grid, physics, dt, kwargs, T, S = land_case()
    p = oracle_params(grid, physics, dt, kwargs)
    s = O.init_state(p, np.full(T.shape, 4.0), np.full(T.shape, 34.0))
    b = R.budget_diagnostics(s, p)
    assert abs(b["mean_T"] - 4.0) < 1e-14 and abs(b["mean_S"] - 34.0) < 1e-13
    assert abs(b["heat_content_J"] - O.RHO_0 * O.C_P * 4.0 * b["total_volume_m3"]) <= 1e-12 * b["heat_content_J"]
    assert b["eta_volume_m3"] == 0.0
