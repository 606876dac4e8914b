"""Physical constants and the physics parameter record.

Mirrors ``src/zhenmode/model/config.py:12-25`` (constants) and
``src/zhenmode/model/config.py:86-140`` (``PhysicsConfig``) so that callers of
the reference factory can pass the same objects.
"""

from __future__ import annotations

from dataclasses import dataclass
from numbers import Integral, Real

import numpy as np

G_EARTH = 9.81          # m/s^2
RHO_0 = 1025.0          # kg/m^3
ALPHA_T = 2.0e-4        # 1/degC
BETA_S = 7.6e-4         # 1/psu
C_P = 3992.0            # J/(kg degC)
R_EARTH = 6371.0e3      # m
OMEGA = 7.2921e-5       # rad/s


@dataclass(frozen=True)
class PhysicsConfig:
    """Turbulence, EOS and friction coefficients (reference config.py:86-140)."""
    nu_h: float = 100.0
    nu_v: float = 1.0e-4
    kappa_h: float = 100.0
    kappa_v: float = 1.0e-5
    kappa_conv: float = 0.05
    nu_bi: float = 1.0e12
    kappa_bi: float = 1.0e12
    kappa_gm: float = 0.0
    gm_slope_max: float = 0.01
    kappa_redi: float = 0.0
    T_ref: float = 15.0
    S_ref: float = 35.0
    tau_x: float = 0.0
    tau_y: float = 0.0
    Q_heat: float = 0.0
    cd: float = 2.5e-3
    r_bot: float = 1.0e-3
    bottom_friction: str = "linear"
    thermodynamics: str = "linear"


def _is_bool(value):
    return isinstance(value, (bool, np.bool_))


def checked_scalar(name, value, *, integer=False, above=None, at_least=None):
    """Return ``value`` if it is a real (or, with ``integer``, an integral)
    non-boolean finite number within the stated bounds; otherwise raise
    ``ValueError`` naming the argument, as the reference factory does for
    invalid configuration (config.py:157-170)."""
    kind = Integral if integer else Real
    if _is_bool(value) or not isinstance(value, kind) or not np.isfinite(value):
        raise ValueError(f"{name} must be {'an integer' if integer else 'a finite number'}, got {value!r}")
    if above is not None and not value > above:
        raise ValueError(f"{name} must be greater than {above}, got {value!r}")
    if at_least is not None and not value >= at_least:
        raise ValueError(f"{name} must be at least {at_least}, got {value!r}")
    return value


def finite_number(name, value, *, positive=False, nonnegative=False):
    """A finite real: > 0 with ``positive``, >= 0 with ``nonnegative``."""
    return checked_scalar(name, value, above=0.0 if positive else None, at_least=0.0 if nonnegative else None)


def integer_count(name, value, *, minimum=0):
    """An integer >= ``minimum``."""
    return checked_scalar(name, value, integer=True, at_least=minimum)
