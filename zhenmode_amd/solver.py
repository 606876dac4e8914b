"""``make_solver_global`` for the MI355X mode-split step.

Drop-in for the reference factory (``src/zhenmode/model/solver/factory.py:30-68``,
``_compile_solver`` at ``:721-799``): same arguments, same return arity
(``step, init_state, diagnostics`` [+ ``params, terms_fn``] [+ ``step_dyn``]),
same ``ValueError`` at construction for invalid settings, and a state record
``JaxStateG(u, v, T, S, eta, ice)`` (``solver/state.py:10``).

State fields are torch tensors on the GPU with the reference shapes
``(nx, ny, nz)`` / ``(nx, ny)``; their storage is i-innermost ``(nz, ny, nx)``
(the tensors are permuted views), which is the layout the HIP kernels use, so
stepping never transposes.  ``step`` is functional: it returns new tensors and
leaves its input untouched, like the jitted reference step.
"""

from __future__ import annotations

import ctypes
from collections import namedtuple

import numpy as np
import torch

from . import _lib
from .params import freeze, make_config

JaxStateG = namedtuple("JaxStateG", ["u", "v", "T", "S", "eta", "ice"])
JaxStateG.__new__.__defaults__ = (0.0,)
State = JaxStateG

_TORCH_DTYPE = {"float32": torch.float32, "float64": torch.float64}


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("the HIP mode-split step needs a ROCm GPU (torch.cuda is unavailable)")
    return torch.device("cuda", torch.cuda.current_device())


def empty3(nx, ny, nz, dtype, device):
    """A (nx, ny, nz) tensor whose storage is contiguous (nz, ny, nx)."""
    return torch.empty((nz, ny, nx), dtype=dtype, device=device).permute(2, 1, 0)


def empty2(nx, ny, dtype, device):
    return torch.empty((ny, nx), dtype=dtype, device=device).permute(1, 0)


def _is_internal(t, shape, dtype, device):
    return (isinstance(t, torch.Tensor) and tuple(t.shape) == shape and t.dtype == dtype
            and t.device == device and t.permute(*reversed(range(t.dim()))).is_contiguous())


def as_field(x, shape, dtype, device):
    """Return x as a reference-shaped tensor in the kernels' storage order."""
    if _is_internal(x, shape, dtype, device):
        return x
    if isinstance(x, torch.Tensor):
        src = x.detach().to(device=device, dtype=dtype)
    else:
        src = torch.as_tensor(np.asarray(x, dtype=np.float64)).to(device=device, dtype=dtype)
    if src.dim() == 0:
        src = src.expand(shape)
    if tuple(src.shape) != shape:
        raise ValueError(f"field shape {tuple(src.shape)} differs from {shape}")
    out = (empty3(*shape, dtype, device) if len(shape) == 3 else empty2(*shape, dtype, device))
    out.copy_(src)
    return out


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Solver:
    """One frozen parameter set on the GPU (a ``zm_handle``)."""

    def __init__(self, params):
        self.params = params
        self.lib = _lib.load()
        self.device = _device()
        self.dtype = _TORCH_DTYPE[params.dtype]
        cfg, keep = make_config(params)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            rc = self.lib.zm_create(ctypes.byref(cfg), ctypes.byref(handle))
        _lib.check(rc, None)
        del keep
        self.handle = handle
        nx, ny, nz = params.nx, params.ny, params.nz
        self.shape3, self.shape2 = (nx, ny, nz), (nx, ny)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.zm_destroy(h)
            self.handle = None

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def coerce(self, state):
        """Any JaxStateG-like record -> internal-layout tensors of the compute dtype."""
        s3, s2, dt, dev = self.shape3, self.shape2, self.dtype, self.device
        ice = state.ice if hasattr(state, "ice") else 0.0
        return JaxStateG(as_field(state.u, s3, dt, dev), as_field(state.v, s3, dt, dev),
                         as_field(state.T, s3, dt, dev), as_field(state.S, s3, dt, dev),
                         as_field(state.eta, s2, dt, dev), as_field(ice, s2, dt, dev))

    def new_state(self):
        s3, s2, dt, dev = self.shape3, self.shape2, self.dtype, self.device
        return JaxStateG(empty3(*s3, dt, dev), empty3(*s3, dt, dev), empty3(*s3, dt, dev),
                         empty3(*s3, dt, dev), empty2(*s2, dt, dev), empty2(*s2, dt, dev))

    @staticmethod
    def zm_state(s):
        return _lib.ZmState(*(t.data_ptr() for t in s))

    def step(self, state, nsteps=1, out=None):
        src = self.coerce(state)
        dst = self.new_state() if out is None else out
        rc = self.lib.zm_step(self.handle, ctypes.byref(self.zm_state(src)), ctypes.byref(self.zm_state(dst)),
                              int(nsteps), self._stream())
        _lib.check(rc, self.handle)
        return dst

    def step_dyn(self, state, tau_x, tau_y, q_heat, T_atm_3d=None, surface_pressure_pa=None, out=None):
        if surface_pressure_pa is not None:
            raise ValueError("surface_pressure_pa is not supported by the HIP step")
        src = self.coerce(state)
        dst = self.new_state() if out is None else out
        s2, dt, dev = self.shape2, self.dtype, self.device
        tx, ty, qh = (as_field(a, s2, dt, dev) for a in (tau_x, tau_y, q_heat))
        ta = None
        if T_atm_3d is not None:
            arr = T_atm_3d.reshape(s2) if hasattr(T_atm_3d, "reshape") else T_atm_3d
            ta = as_field(arr, s2, dt, dev)
        rc = self.lib.zm_step_dyn(self.handle, ctypes.byref(self.zm_state(src)), ctypes.byref(self.zm_state(dst)),
                                  ptr(tx), ptr(ty), ptr(qh), ptr(ta) if ta is not None else None,
                                  self._stream())
        _lib.check(rc, self.handle)
        return dst

    def diagnostics(self, state, surface_pressure_pa=None):
        if surface_pressure_pa is not None:
            raise ValueError("surface_pressure_pa is not supported by the HIP step")
        src = self.coerce(state)
        s3, dt, dev = self.shape3, self.dtype, self.device
        rho, pres, w = (empty3(*s3, dt, dev) for _ in range(3))
        rc = self.lib.zm_diagnostics(self.handle, ctypes.byref(self.zm_state(src)), ptr(rho), ptr(pres), ptr(w),
                                     self._stream())
        _lib.check(rc, self.handle)
        return rho, pres, w

    def tracer_terms(self, state):
        """_tracer_terms (tendencies.py:74-92): a (6, nx, ny, nz) device tensor of
        the temperature's [advection, horizontal diffusion, vertical diffusion,
        convection, GM, Redi] tendencies on wet nodes (GM and Redi zero)."""
        src = self.coerce(state)
        nx, ny, nz = self.shape3
        out = torch.empty((6, nz, ny, nx), dtype=self.dtype, device=self.device)
        rc = self.lib.zm_tracer_terms(self.handle, ctypes.byref(self.zm_state(src)), ptr(out), self._stream())
        _lib.check(rc, self.handle)
        return out.permute(0, 3, 2, 1)

    def stage(self, which, state):
        """One stage of the step; the fields the stage does not write (a
        residual stage writes only u, v or T, S) keep the input's values."""
        src = self.coerce(state)
        dst = self.new_state()
        for a, b in zip(dst, src):
            a.copy_(b)
        rc = self.lib.zm_stage(self.handle, int(which), ctypes.byref(self.zm_state(src)),
                               ctypes.byref(self.zm_state(dst)), self._stream())
        _lib.check(rc, self.handle)
        return dst

    def set_graph(self, enabled):
        _lib.check(self.lib.zm_set_graph(self.handle, int(bool(enabled))), self.handle)

    def nu_substeps_per_launch(self):
        return int(self.lib.zm_nu_substeps_per_launch(self.handle))

    def time_kernel(self, which, reps):
        ms = ctypes.c_double()
        _lib.check(self.lib.zm_time_kernel(self.handle, int(which), int(reps), ctypes.byref(ms)), self.handle)
        return ms.value

    def vertical_mix(self, state, dt, kappa_v, kappa_conv, only=None, out=None):
        """Implicit (backward-Euler) vertical mixing of T and S over dt, Thomas
        solve per column; returns a state whose T, S are the mixed fields and
        whose other fields are the input's (shared, not copied)."""
        src = self.coerce(state)
        s3, dt_, dev = self.shape3, self.dtype, self.device
        To = empty3(*s3, dt_, dev) if out is None else out[0]
        So = empty3(*s3, dt_, dev) if out is None else out[1]
        if only is not None and (only.dtype != torch.int32 or only.numel() != self.shape2[0] * self.shape2[1]):
            raise ValueError("only must be an int32 tensor with one entry per column")
        rc = self.lib.zm_vertical_mix(self.handle, ptr(src.T), ptr(src.S), ptr(To), ptr(So), float(dt),
                                      float(kappa_v), float(kappa_conv), ptr(only) if only is not None else None,
                                      self._stream())
        _lib.check(rc, self.handle)
        return src._replace(T=To, S=So)

    def project_velocity(self, u, v, dt=None, out=None, iterations=False):
        """_project_column_divergence(u, v, p, dt) (dynamics/projection.py:64-127):
        (u, v) minus dt g grad(psi) with the column divergence removed, using
        the handle's projection settings.  With ``iterations`` also returns
        the device int32[3] CG iteration counts (first solve, refinements)."""
        s3, dt_, dev = self.shape3, self.dtype, self.device
        uu, vv = as_field(u, s3, dt_, dev), as_field(v, s3, dt_, dev)
        uo = empty3(*s3, dt_, dev) if out is None else out[0]
        vo = empty3(*s3, dt_, dev) if out is None else out[1]
        its = torch.zeros(3, dtype=torch.int32, device=dev)
        rc = self.lib.zm_project_velocity(self.handle, ptr(uu), ptr(vv), ptr(uo), ptr(vo),
                                          float(self.params.dt if dt is None else dt), ptr(its), self._stream())
        _lib.check(rc, self.handle)
        return (uo, vo, its) if iterations else (uo, vo)

    def projection_iterations(self):
        """CG iterations of the last projection on this handle (synchronises)."""
        its = torch.zeros(3, dtype=torch.int32, device=self.device)
        _lib.check(self.lib.zm_projection_iterations(self.handle, ptr(its), self._stream()), self.handle)
        return [int(x) for x in its.cpu()]

    def set_barotropic_persistent(self, enabled):
        """Run the barotropic sub-cycle as one persistent launch (True) or one
        launch per substep (False); results are bit-identical."""
        _lib.check(self.lib.zm_set_barotropic_persistent(self.handle, int(bool(enabled))), self.handle)

    def barotropic_timeouts(self):
        """Bounded waits of the persistent sub-cycle that gave up since the handle
        was created (nonzero: the steps since then are invalid)."""
        n = ctypes.c_int32(0)
        _lib.check(self.lib.zm_barotropic_timeouts(self.handle, ctypes.byref(n)), self.handle)
        return int(n.value)

    def set_tiles(self, nu_block, seg):
        _lib.check(self.lib.zm_set_tiles(self.handle, int(nu_block), int(seg)), self.handle)

    def get_tiles(self):
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self.lib.zm_get_tiles(self.handle, ctypes.byref(a), ctypes.byref(b)), self.handle)
        return a.value, b.value

    def autotune(self, sample, reps=5):
        """Pick (nu_block, seg) by timing eagerly issued steps from ``sample``
        for every candidate (no graph capture); leaves the fastest set and
        returns (nu_block, seg, ms per eager step)."""
        src = self.coerce(sample)
        a, b, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        torch.cuda.synchronize(self.device)
        rc = self.lib.zm_autotune(self.handle, ctypes.byref(self.zm_state(src)), int(reps), ctypes.byref(a),
                                  ctypes.byref(b), ctypes.byref(ms))
        _lib.check(rc, self.handle)
        return a.value, b.value, ms.value

    def init_state(self, T_init=None, S_init=None):
        """Reference factory.py:766-788: zero velocity/eta/ice, land and ghost
        nodes set to the T_ref/S_ref sentinel."""
        p = self.params
        shape = self.shape3
        if T_init is not None:
            T = np.array(_to_numpy(T_init), dtype=np.float64)
            S = (np.array(_to_numpy(S_init), dtype=np.float64) if S_init is not None
                 else np.full_like(T, p.S_ref))
        else:
            T = np.full(shape, p.T_ref)
            S = np.full(shape, p.S_ref)
        # The reference casts the field to the state dtype before masking.
        np_dtype = np.float32 if p.dtype == "float32" else np.float64
        wm = p.wet_mask_z.astype(np_dtype)
        T = T.astype(np_dtype) * wm + (np_dtype(1.0) - wm) * np_dtype(p.T_ref)
        S = S.astype(np_dtype) * wm + (np_dtype(1.0) - wm) * np_dtype(p.S_ref)
        zero3 = np.zeros(shape)
        zero2 = np.zeros(self.shape2)
        return self.coerce(JaxStateG(zero3, zero3, T, S, zero2, zero2))


def _to_numpy(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def make_solver_global(grid, physics, dt, forcing=None, *, return_params=False, dynamic_forcing=False,
                       **kwargs):
    """Create the HIP mode-split solver; signature of the reference factory."""
    params = freeze(grid, physics, dt, forcing, **kwargs)
    solver = Solver(params)

    def step(state):
        return solver.step(state)

    def init_state(T_init=None, S_init=None):
        return solver.init_state(T_init, S_init)

    def diagnostics(state, surface_pressure_pa=None):
        return solver.diagnostics(state, surface_pressure_pa)

    def terms_fn(state):
        return solver.tracer_terms(state)

    def step_dyn(state, tau_x, tau_y, q_heat, T_atm_3d=None, surface_pressure_pa=None):
        return solver.step_dyn(state, tau_x, tau_y, q_heat, T_atm_3d, surface_pressure_pa)

    step.solver = solver
    params.solver = solver
    if dynamic_forcing:
        if return_params:
            return step, init_state, diagnostics, params, terms_fn, step_dyn
        return step, init_state, diagnostics, step_dyn
    if return_params:
        return step, init_state, diagnostics, params, terms_fn
    return step, init_state, diagnostics


class Ensemble:
    """Float32 solvers of different sizes stepped together (``zm_ensemble_*``)."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        self.lib = _lib.load()
        arr = (ctypes.c_void_p * len(self.solvers))(*(s.handle.value for s in self.solvers))
        h = ctypes.c_void_p()
        rc = self.lib.zm_ensemble_create(arr, len(self.solvers), ctypes.byref(h))
        _lib.check(rc, None)
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.zm_ensemble_destroy(h)
            self.handle = None

    def set_graph(self, enabled):
        _lib.check(self.lib.zm_ensemble_set_graph(self.handle, int(bool(enabled))), None)

    def step(self, states, out=None):
        if len(states) != len(self.solvers):
            raise ValueError("one state per member is required")
        src = [s.coerce(st) for s, st in zip(self.solvers, states)]
        dst = [s.new_state() for s in self.solvers] if out is None else list(out)
        n = len(self.solvers)
        a_in = (_lib.ZmState * n)(*(Solver.zm_state(x) for x in src))
        a_out = (_lib.ZmState * n)(*(Solver.zm_state(x) for x in dst))
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.solvers[0].device).cuda_stream)
        rc = self.lib.zm_ensemble_step(self.handle, a_in, a_out, stream)
        _lib.check(rc, None)
        return dst


def to_numpy_state(state):
    """Reference-layout float64 numpy copy of a state (for checks and restarts)."""
    return JaxStateG(*(np.asarray(_to_numpy(f), dtype=np.float64) for f in state))
