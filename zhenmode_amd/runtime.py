"""Device-side run loop: seasonal forcing, monitor and budget without per-step host syncs.

* ``seasonal_forcing`` blends a 12-month stack of 2-D fields on the GPU with the
  30-day calendar of ``interp_monthly_field_jit``
  (``src/zhenmode/model/inputs/forcing/seasonal.py:86-103``).
* ``DeviceMonitor`` classifies states on the GPU like ``classify_state``
  (``runtime/monitor.py:20-30``) and keeps the running peaks and the
  first-rejection record of ``run_integration`` (``runtime/run_loop.py:53-96``)
  in device memory.
* ``budget`` returns the global diagnostics of ``compute_budget_diagnostics``
  (``diagnostics/snapshot.py:67-128``) from float64 compensated GPU sums.
* ``run`` advances a step function with the reference's accept/reject
  semantics, reading the device status only every ``check_every`` steps; the
  last ``check_every + 1`` states stay in a ring of device buffers, so the last
  accepted and the first rejected state are both available after a rejection.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from .config import C_P, RHO_0

MAX_U_BOUND = 10.0   # m/s, reference runtime/cli.py:21
ETA_BLOWUP_M = 15.0  # m, reference runtime/cli.py:27


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def seasonal_forcing(solver, stack, day, blend_days=5.0, out=None):
    """stack: device tensor (12, nfields, ny, nx) in the solver dtype -> (nfields, ny, nx)."""
    if stack.dim() != 4 or stack.shape[0] != 12 or tuple(stack.shape[2:]) != (solver.shape2[1], solver.shape2[0]):
        raise ValueError("stack must be (12, nfields, ny, nx)")
    if stack.dtype != solver.dtype or not stack.is_contiguous():
        raise ValueError("stack must be contiguous in the solver dtype")
    nf = stack.shape[1]
    out = torch.empty(stack.shape[1:], dtype=stack.dtype, device=stack.device) if out is None else out
    rc = solver.lib.zm_seasonal_forcing(solver.handle, _p(stack), nf, float(day), float(blend_days), _p(out),
                                        solver._stream())
    _lib.check(rc, solver.handle)
    return out


class DeviceMonitor:
    """classify_state on the device with run-loop first-rejection bookkeeping."""

    def __init__(self, solver, velocity_limit=MAX_U_BOUND, eta_limit=ETA_BLOWUP_M, inclusive=False):
        self.solver = solver
        self.velocity_limit, self.eta_limit, self.inclusive = float(velocity_limit), float(eta_limit), bool(inclusive)
        self.status = torch.zeros(ctypes.sizeof(_lib.ZmMonitorStatus), dtype=torch.uint8, device=solver.device)
        self.reset()

    def reset(self):
        s = self.solver
        _lib.check(s.lib.zm_monitor_reset(s.handle, _p(self.status), s._stream()), s.handle)

    def classify(self, state, step):
        s = self.solver
        src = s.coerce(state)
        rc = s.lib.zm_monitor(s.handle, ctypes.byref(s.zm_state(src)), int(step), self.velocity_limit,
                              self.eta_limit, int(self.inclusive), _p(self.status), s._stream())
        _lib.check(rc, s.handle)

    def read(self):
        """Copy the status to the host (synchronises)."""
        raw = bytes(self.status.cpu().numpy().tobytes())
        st = _lib.ZmMonitorStatus.from_buffer_copy(raw)
        return {f: getattr(st, f) for f, _ in _lib.ZmMonitorStatus._fields_}


@dataclass
class Budget:
    mean_T: float
    mean_S: float
    mean_T_top: float
    mean_S_top: float
    heat_content_J: float
    salt_content_kg: float
    total_volume_m3: float
    mean_depth_m: float
    eta_volume_m3: float


def budget(solver, state, rho0=RHO_0, cp=C_P):
    s = solver
    src = s.coerce(state)
    sums = torch.empty(9, dtype=torch.float64, device=s.device)
    _lib.check(s.lib.zm_budget(s.handle, ctypes.byref(s.zm_state(src)), _p(sums), s._stream()), s.handle)
    vol, tv, sv, a_wet, ta, sa, a_tot, ea, _ = (float(x) for x in sums.cpu())
    return Budget(mean_T=tv / vol, mean_S=sv / vol, mean_T_top=ta / a_wet, mean_S_top=sa / a_wet,
                  heat_content_J=rho0 * cp * tv, salt_content_kg=rho0 * sv / 1000.0, total_volume_m3=vol,
                  mean_depth_m=vol / a_tot, eta_volume_m3=ea)


@dataclass
class RunResult:
    state: object            # last accepted state
    accepted_steps: int
    status: dict             # final device monitor status
    rejected: object = None  # first rejected state, if any


def run(step, solver, state, n_steps, check_every=10, velocity_limit=MAX_U_BOUND, eta_limit=ETA_BLOWUP_M,
        inclusive=False):
    """Advance ``n_steps`` with ``step(state_in, out=state_out)``-style stepping
    into a device ring of ``check_every + 1`` states, classifying every proposed
    state on the device and reading the status once per ``check_every`` steps.
    Stops at the first rejected state, like ``run_integration``."""
    if check_every < 1:
        raise ValueError("check_every must be >= 1")
    mon = DeviceMonitor(solver, velocity_limit, eta_limit, inclusive)
    ring = [solver.new_state() for _ in range(check_every + 1)]
    for dst, src in zip(ring[0], solver.coerce(state)):  # the caller's state is never written
        dst.copy_(src)
    size = len(ring)
    mon.classify(ring[0], 0)
    st = mon.read()
    if st["failure"]:
        return RunResult(ring[0], 0, st, rejected=ring[0])
    n = 0
    while n < n_steps:
        take = min(check_every, n_steps - n)
        for q in range(take):
            src, dst = ring[(n + q) % size], ring[(n + q + 1) % size]
            step(src, out=dst)
            mon.classify(dst, n + q + 1)
        st = mon.read()
        if st["failure"]:
            f = st["first_failure_step"]
            return RunResult(ring[(f - 1) % size], f - 1, st, rejected=ring[f % size])
        n += take
    return RunResult(ring[n % size], n, st)
