"""NumPy restatements of the run-time helpers kept on the GPU.

TEST INFRASTRUCTURE ONLY: checker for ``zm_seasonal_forcing``, ``zm_monitor``
and ``zm_budget``; only ``tests/`` may import it.

* ``interp_monthly_field``: ``src/zhenmode/model/inputs/forcing/seasonal.py:86-103``
  (30-day months, blend window centred on month boundaries).
* ``classify_state``: ``runtime/monitor.py:20-30`` (codes 0, 1, 3, 4).
* ``budget_diagnostics``: ``diagnostics/snapshot.py:67-128`` (legacy node
  thickness, true wet mask).

Pinned in ``tests/test_oracle_runtime.py`` by closed forms: mid-month returns
the month's field, a boundary returns the mean of the two months, weights sum
to one, the blend is continuous across boundaries; classify codes on planted
states; budgets of uniform fields.
"""

from __future__ import annotations

import numpy as np

from .zm_oracle import C_P, RHO_0


def interp_monthly_field(stack, day, blend_days=5.0):
    month_len = 30.0
    mpos = day % month_len
    mi = int(np.floor(day / month_len)) % 12
    half = blend_days / 2.0
    if blend_days <= 0.0:
        return stack[mi]
    w_after = min(1.0, max(0.0, (mpos - (month_len - half)) / blend_days))
    w_before = min(1.0, max(0.0, (half - mpos) / blend_days))
    return (1.0 - w_after - w_before) * stack[mi] + w_after * stack[(mi + 1) % 12] + w_before * stack[(mi - 1) % 12]


def classify_state(state, velocity_limit=10.0, eta_limit=15.0, inclusive=False):
    max_u = float(np.max(np.abs(state.u)))
    velocity = max(max_u, float(np.max(np.abs(state.v))))
    eta = float(np.max(np.abs(state.eta)))
    finite = all(bool(np.all(np.isfinite(np.asarray(f)))) for f in state)
    bad_v = velocity >= velocity_limit if inclusive else velocity > velocity_limit
    bad_e = eta >= eta_limit if inclusive else eta > eta_limit
    failure = 1 if not finite else (3 if bad_v else (4 if bad_e else 0))
    return dict(max_u=max_u, max_velocity=velocity, max_eta=eta, finite=finite, failure=failure)


def budget_diagnostics(state, p, rho0=RHO_0, cp=C_P):
    area = p.dx_2d * p.dy
    wet3 = p.wet_mask_z
    volume = area[:, :, None] * p.dz_node[None, None, :] * wet3
    vol = volume.sum()
    surf = area * p.wet_mask
    return dict(mean_T=float((state.T * volume).sum() / vol), mean_S=float((state.S * volume).sum() / vol),
                mean_T_top=float((state.T[:, :, 0] * surf).sum() / surf.sum()),
                mean_S_top=float((state.S[:, :, 0] * surf).sum() / surf.sum()),
                heat_content_J=float((rho0 * cp * state.T * volume).sum()),
                salt_content_kg=float((rho0 * (state.S / 1000.0) * volume).sum()),
                total_volume_m3=float(vol), mean_depth_m=float(vol / area.sum()),
                eta_volume_m3=float((state.eta * surf).sum()))
