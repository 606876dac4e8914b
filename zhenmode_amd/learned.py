"""Learned vertical mixing: a per-column MLP that stands in for the implicit
mixing step, with its training loop, both on the project's own HIP kernels.

The network maps a column's features (T, S, u, v and N^2 on 50 levels plus
q_heat, the SSS-restoring salt flux and |tau|: 253 inputs, padded to 256) to
50-level T and S outputs (100 outputs, padded to 128): 253 -> 256 -> 256 ->
256 -> 100 with GELU, bfloat16 weights and activations, float32 accumulation
on MFMA, float32 master weights updated by Adam (``zm_adam``).

What the network predicts: per column and tracer X, with e the explicit
increment of the mixing operator over dt and s = max_k |e_k| the column's
mixing scale, the increment is dX = e + s z, where z (the network's output,
de-normalised per level) is the implicit solution's departure from the
explicit step in units of s.  Stable columns have a mixing scale orders of
magnitude below convectively unstable ones, so the network's error shrinks
with the increment it corrects.  The column integral of dX is shifted to zero
in float64 (heat and salt conserved exactly) and the implicit-system residual
of every column is checked; columns above tolerance fall back to the Thomas
solve.

``LearnedMixer.apply`` runs, on float32 handles, one fused launch
(``zm_mlp_mix``: features -> 4 MFMA layers -> apply, activations kept on
chip), then ``zm_vertical_mix`` on the flagged columns only (the Thomas
fallback).  ``apply_unfused`` (and float64 handles) run the same arithmetic as
features -> 4 GEMMs -> ``zm_mlp_apply``.  The reference has no learned
component; the mixing operator it replaces is the backward-Euler counterpart
of the reference's flux-form vertical operator
(``src/zhenmode/model/solver/numerics/vertical.py:296-332``).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

N_IN = 256      # padded feature count (253 used at nz = 50)
WIDTH = 256
N_OUT_PAD = 128
ROW_BLOCK = 128  # GEMM row granularity


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(rc, handle=None):
    _lib.check(rc, handle)


def pad_rows(n):
    return ((n + ROW_BLOCK - 1) // ROW_BLOCK) * ROW_BLOCK


class ColumnMLP:
    """Weights, normalisation and the forward/backward/Adam kernels' driver."""

    def __init__(self, nz, device=None, seed=0):
        if 5 * nz + 3 > N_IN:
            raise ValueError("the column MLP takes at most 50 levels")
        self.lib = _lib.load()
        self.nz = nz
        self.n_out = 2 * nz
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        dev = self.device
        self.sizes = [N_IN, WIDTH, WIDTH, WIDTH, N_OUT_PAD]
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.w, self.b = [], []
        for l, (k, n) in enumerate(zip(self.sizes[:-1], self.sizes[1:])):
            w = torch.randn((n, k), generator=g, dtype=torch.float64) * math.sqrt(2.0 / k)
            if l == 0:
                w[:, 5 * nz + 3:] = 0.0
            if l == len(self.sizes) - 2:
                w *= 0.1
                w[self.n_out:, :] = 0.0
            self.w.append(w.to(dtype=torch.float32, device=dev).contiguous())
            self.b.append(torch.zeros(n, dtype=torch.float32, device=dev))
        self.w_bf = [torch.empty_like(w, dtype=torch.bfloat16) for w in self.w]
        self.wT_bf = [torch.empty((w.shape[1], w.shape[0]), dtype=torch.bfloat16, device=dev) for w in self.w]
        self.refresh_bf16()
        self.m = [torch.zeros_like(w) for w in self.w]
        self.v = [torch.zeros_like(w) for w in self.w]
        self.bm = [torch.zeros_like(b) for b in self.b]
        self.bv = [torch.zeros_like(b) for b in self.b]
        self.t = 0
        self.in_mean = torch.zeros(N_IN, dtype=torch.float32, device=dev)
        self.in_inv_std = torch.ones(N_IN, dtype=torch.float32, device=dev)
        self.out_mean = torch.zeros(N_OUT_PAD, dtype=torch.float32, device=dev)
        self.out_std = torch.ones(N_OUT_PAD, dtype=torch.float32, device=dev)

    def refresh_bf16(self):
        for w, wb, wt in zip(self.w, self.w_bf, self.wT_bf):
            wb.copy_(w)
            wt.copy_(w.t())

    # ------------------------------------------------------------------ forward
    def inference_buffers(self, M):
        """Preallocated activations and output of forward() over M rows."""
        dev = self.device
        acts = [torch.empty((M, n), dtype=torch.bfloat16, device=dev) for n in self.sizes[1:-1]]
        return acts, torch.empty((M, N_OUT_PAD), dtype=torch.float32, device=dev)

    def forward(self, X, keep=False, bufs=None):
        """X: (M, 256) bf16, M a multiple of 128.  Returns Y (M, 128) float32
        and, with keep=True, the cache the backward pass needs.  bufs (from
        inference_buffers(M), keep=False only) receives the activations and Y,
        so that a forward pass allocates nothing."""
        M = X.shape[0]
        if M % ROW_BLOCK or X.dtype != torch.bfloat16 or tuple(X.shape[1:]) != (N_IN,) or not X.is_contiguous():
            raise ValueError("X must be a contiguous (M, 256) bfloat16 tensor with M a multiple of 128")
        if bufs is not None and (keep or bufs[1].shape[0] != M):
            raise ValueError("inference buffers are for keep=False over the same row count")
        dev = self.device
        acts, actsT, pre = [X], [], []
        h = X
        nl = len(self.w)
        for l in range(nl - 1):
            n = self.sizes[l + 1]
            a = bufs[0][l] if bufs is not None else torch.empty((M, n), dtype=torch.bfloat16, device=dev)
            z = torch.empty((M, n), dtype=torch.float32, device=dev) if keep else None
            aT = torch.empty((n, M), dtype=torch.bfloat16, device=dev) if keep else None
            _check(self.lib.zm_gemm_bf16_nt(M, n, self.sizes[l], _p(h), _p(self.w_bf[l]), _lib.EPI_BIAS_GELU,
                                            _p(self.b[l]), None, _p(z), _p(a), _p(aT), _stream()))
            acts.append(a)
            actsT.append(aT)
            pre.append(z)
            h = a
        Y = bufs[1] if bufs is not None else torch.empty((M, N_OUT_PAD), dtype=torch.float32, device=dev)
        _check(self.lib.zm_gemm_bf16_nt(M, N_OUT_PAD, self.sizes[-2], _p(h), _p(self.w_bf[-1]), _lib.EPI_BIAS_F32,
                                        _p(self.b[-1]), None, _p(Y), None, None, _stream()))
        return Y, ((acts, actsT, pre) if keep else None)

    # ----------------------------------------------------------------- backward
    def gradients(self, X, target):
        """Loss and parameter gradients for one batch (normalised target)."""
        M = X.shape[0]
        dev = self.device
        Y, (acts, actsT, pre) = self.forward(X, keep=True)
        XT = torch.empty((N_IN, M), dtype=torch.bfloat16, device=dev)
        _check(self.lib.zm_transpose_bf16(M, N_IN, _p(X), _p(XT), _stream()))
        actsT = [XT] + actsT
        dZ = torch.empty((M, N_OUT_PAD), dtype=torch.bfloat16, device=dev)
        dZT = torch.empty((N_OUT_PAD, M), dtype=torch.bfloat16, device=dev)
        loss = torch.zeros(1, dtype=torch.float32, device=dev)
        _check(self.lib.zm_mse_grad(M, N_OUT_PAD, self.n_out, _p(Y), _p(target), _p(dZ), _p(dZT), _p(loss),
                                    _stream()))
        gw, gb = [None] * len(self.w), [None] * len(self.w)
        for l in range(len(self.w) - 1, -1, -1):
            n, k = self.sizes[l + 1], self.sizes[l]
            gw[l] = torch.empty((n, k), dtype=torch.float32, device=dev)
            _check(self.lib.zm_gemm_bf16_nt(n, k, M, _p(dZT), _p(actsT[l]), _lib.EPI_F32, None, None, _p(gw[l]),
                                            None, None, _stream()))
            gb[l] = torch.empty(n, dtype=torch.float32, device=dev)
            _check(self.lib.zm_rowsum_bf16(n, M, _p(dZT), _p(gb[l]), _stream()))
            if l > 0:
                nd = torch.empty((M, k), dtype=torch.bfloat16, device=dev)
                ndT = torch.empty((k, M), dtype=torch.bfloat16, device=dev)
                _check(self.lib.zm_gemm_bf16_nt(M, k, n, _p(dZ), _p(self.wT_bf[l]), _lib.EPI_GELU_GRAD, None,
                                                _p(pre[l - 1]), None, _p(nd), _p(ndT), _stream()))
                dZ, dZT = nd, ndT
        return loss / (M * self.n_out), gw, gb, Y

    def adam_step(self, gw, gb, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        for l in range(len(self.w)):
            n, k = self.w[l].shape
            _check(self.lib.zm_adam(n, k, _p(self.w[l]), _p(gw[l]), _p(self.m[l]), _p(self.v[l]),
                                    _p(self.w_bf[l]), _p(self.wT_bf[l]), _p(self.b[l]), _p(gb[l]), _p(self.bm[l]),
                                    _p(self.bv[l]), lr, beta1, beta2, eps, self.t, _stream()))

    def train_step(self, X, target, lr=1e-3):
        loss, gw, gb, _ = self.gradients(X, target)
        self.adam_step(gw, gb, lr=lr)
        return loss

    # ---------------------------------------------------------- normalisation
    def fit_normalisation(self, X_raw, targets):
        """Per-feature mean / inverse std of raw features (float32 rows); per
        output mean and std of the scaled targets (rows x 2 nz), with a floor
        of 1e-6 so that levels whose target is always zero keep a negligible
        output scale."""
        mu = X_raw.mean(dim=0)
        sd = X_raw.std(dim=0)
        inv = torch.where(sd > 1e-12, 1.0 / sd.clamp_min(1e-30), torch.zeros_like(sd))
        self.in_mean.copy_(torch.where(sd > 1e-12, mu, torch.zeros_like(mu)))
        self.in_inv_std.copy_(inv)
        self.out_mean.zero_()
        self.out_std.fill_(1.0)
        self.out_mean[:self.n_out] = targets.mean(dim=0)
        self.out_std[:self.n_out] = targets.std(dim=0).clamp_min(1e-6)

    def state_dict(self):
        return {"w": [w.cpu() for w in self.w], "b": [b.cpu() for b in self.b],
                "in_mean": self.in_mean.cpu(), "in_inv_std": self.in_inv_std.cpu(),
                "out_mean": self.out_mean.cpu(), "out_std": self.out_std.cpu(), "nz": self.nz}

    def load_state_dict(self, d):
        for w, src in zip(self.w, d["w"]):
            w.copy_(src)
        for b, src in zip(self.b, d["b"]):
            b.copy_(src)
        for name in ("in_mean", "in_inv_std", "out_mean", "out_std"):
            getattr(self, name).copy_(d[name])
        self.refresh_bf16()


class LearnedMixer:
    """The surrogate mixing step on one solver handle (all columns at once)."""

    def __init__(self, solver, mlp, kappa_v, kappa_conv, tol_T=0.01, tol_S=0.002):
        self.solver = solver
        self.mlp = mlp
        self.kappa_v, self.kappa_conv = float(kappa_v), float(kappa_conv)
        self.tol_T, self.tol_S = float(tol_T), float(tol_S)
        nx, ny = solver.shape2
        self.n2 = nx * ny
        self.rows = pad_rows(self.n2)
        dev = solver.device
        self.X = torch.empty((self.rows, N_IN), dtype=torch.bfloat16, device=dev)
        self.flags = torch.zeros(self.n2, dtype=torch.int32, device=dev)
        self.count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.bufs = mlp.inference_buffers(self.rows)

    def features(self, state, tau_x=None, tau_y=None, q_heat=None, out=None):
        s = self.solver
        src = s.coerce(state)
        X = self.X if out is None else out
        rc = s.lib.zm_column_features(s.handle, ctypes.byref(s.zm_state(src)), _p(tau_x), _p(tau_y), _p(q_heat),
                                      _p(self.mlp.in_mean), _p(self.mlp.in_inv_std), _p(X), X.shape[0],
                                      s._stream())
        _lib.check(rc, s.handle)
        return X

    def _outputs(self, s, out):
        from .solver import empty3
        To = empty3(*s.shape3, s.dtype, s.device) if out is None else out[0]
        So = empty3(*s.shape3, s.dtype, s.device) if out is None else out[1]
        return To, So

    def apply_unfused(self, state, dt, tau_x=None, tau_y=None, q_heat=None, out=None):
        """Surrogate increments through features -> 4 GEMMs -> zm_mlp_apply
        (the same arithmetic as the fused launch), then the Thomas fallback."""
        s = self.solver
        src = s.coerce(state)
        X = self.features(src, tau_x, tau_y, q_heat)
        Y, _ = self.mlp.forward(X, bufs=self.bufs)
        To, So = self._outputs(s, out)
        rc = s.lib.zm_mlp_apply(s.handle, ctypes.byref(s.zm_state(src)), _p(Y), N_OUT_PAD, _p(self.mlp.out_mean),
                                _p(self.mlp.out_std), float(dt), self.kappa_v, self.kappa_conv, self.tol_T,
                                self.tol_S, _p(To), _p(So), _p(self.flags), _p(self.count), s._stream())
        _lib.check(rc, s.handle)
        return self._fallback(src, To, So, dt)

    def apply(self, state, dt, tau_x=None, tau_y=None, q_heat=None, out=None):
        """Mixed state (T, S into ``out`` when given); self.count accumulates
        the columns sent to the fallback.  Float32 handles take the fused
        launch, float64 handles the unfused sequence."""
        s = self.solver
        if s.dtype != torch.float32:
            return self.apply_unfused(state, dt, tau_x, tau_y, q_heat, out)
        src = s.coerce(state)
        To, So = self._outputs(s, out)
        m = self.mlp
        rc = s.lib.zm_mlp_mix(s.handle, ctypes.byref(s.zm_state(src)), _p(tau_x), _p(tau_y), _p(q_heat),
                              _p(m.in_mean), _p(m.in_inv_std), self._wptrs, self._bptrs, _p(m.out_mean),
                              _p(m.out_std), float(dt), self.kappa_v, self.kappa_conv, self.tol_T, self.tol_S,
                              _p(To), _p(So), _p(self.flags), _p(self.count), s._stream())
        _lib.check(rc, s.handle)
        return self._fallback(src, To, So, dt)

    def _fallback(self, src, To, So, dt):
        """Thomas solve on the flagged columns only."""
        s = self.solver
        rc = s.lib.zm_vertical_mix(s.handle, _p(src.T), _p(src.S), _p(To), _p(So), float(dt), self.kappa_v,
                                   self.kappa_conv, _p(self.flags), s._stream())
        _lib.check(rc, s.handle)
        return src._replace(T=To, S=So)

    @property
    def _wptrs(self):
        return (ctypes.c_void_p * 4)(*(w.data_ptr() for w in self.mlp.w_bf))

    @property
    def _bptrs(self):
        return (ctypes.c_void_p * 4)(*(b.data_ptr() for b in self.mlp.b))


def mixing_targets(solver, state, dt, kappa_v, kappa_conv):
    """Training targets (rows x 2 nz, float32, row = j*nx + i) of the Thomas
    step on ``state``: the Thomas solution's departure from the explicit
    mixing step in units of the column's mixing scale (``zm_mix_targets``)."""
    src = solver.coerce(state)
    mixed = solver.vertical_mix(src, dt, kappa_v, kappa_conv)
    nz = solver.shape3[2]
    n2 = solver.shape2[0] * solver.shape2[1]
    Y = torch.empty((n2, 2 * nz), dtype=torch.float32, device=solver.device)
    rc = solver.lib.zm_mix_targets(solver.handle, ctypes.byref(solver.zm_state(src)), _p(mixed.T), _p(mixed.S),
                                   float(dt), float(kappa_v), float(kappa_conv), _p(Y), 2 * nz, solver._stream())
    _lib.check(rc, solver.handle)
    return Y


class MixingStep:
    """One step with the tracer vertical mixing taken out of the explicit
    step and applied afterwards as a column operator over dt.

    ``mode="thomas"``: the implicit Thomas solve (float64 elimination).
    ``mode="learned"``: the surrogate with the Thomas fallback.
    The inner step is the HIP mode-split step built with kappa_v = 0 and
    kappa_conv = 0 (everything else unchanged); the mixing operator uses the
    physics' kappa_v and kappa_conv.
    """

    def __init__(self, grid, physics, dt, mode="thomas", mlp=None, tol_T=0.01, tol_S=0.002, **kwargs):
        from dataclasses import replace

        from .solver import make_solver_global

        if mode not in ("thomas", "learned"):
            raise ValueError("mode must be 'thomas' or 'learned'")
        if mode == "learned" and mlp is None:
            raise ValueError("the learned mixing step needs a trained ColumnMLP")
        self.dt = float(dt)
        self.kappa_v, self.kappa_conv = float(physics.kappa_v), float(physics.kappa_conv)
        inner = replace(physics, kappa_v=0.0, kappa_conv=0.0)
        kw = dict(kwargs)
        kw.pop("coastal_kappa_v", None)
        kw.pop("coastal_kappa_v_mask", None)
        step, init_state, diag, step_dyn = make_solver_global(grid, inner, dt, dynamic_forcing=True, **kw)
        self.solver = step.solver
        self.init_state = init_state
        self.mode = mode
        self.mixer = (LearnedMixer(self.solver, mlp, self.kappa_v, self.kappa_conv, tol_T, tol_S)
                      if mode == "learned" else None)
        self._bufs = None
        self._k = 0

    def mix(self, state, tau_x=None, tau_y=None, q_heat=None, out=None):
        if self.mode == "thomas":
            return self.solver.vertical_mix(state, self.dt, self.kappa_v, self.kappa_conv, out=out)
        return self.mixer.apply(state, self.dt, tau_x, tau_y, q_heat, out=out)

    def _buffers(self):
        """Two-deep ping-pong of step outputs and mixed T, S: the returned
        state stays valid until the second next call (the step's hipGraph is
        then replayed from a fixed set of pointers)."""
        if self._bufs is None:
            from .solver import empty3
            s = self.solver
            self._bufs = [(s.new_state(), (empty3(*s.shape3, s.dtype, s.device), empty3(*s.shape3, s.dtype, s.device)))
                          for _ in range(2)]
        k = self._k
        self._k ^= 1
        return self._bufs[k]

    def __call__(self, state):
        stepped, ts = self._buffers()
        return self.mix(self.solver.step(state, out=stepped), out=ts)

    def step_dyn(self, state, tau_x, tau_y, q_heat):
        stepped, ts = self._buffers()
        s1 = self.solver.step_dyn(state, tau_x, tau_y, q_heat, out=stepped)
        s = self.solver
        tx, ty, qh = (_as2(a, s) for a in (tau_x, tau_y, q_heat))
        return self.mix(s1, tx, ty, qh, out=ts)

    def fallback_count(self):
        return int(self.mixer.count.item()) if self.mixer is not None else 0


def _as2(a, solver):
    from .solver import as_field
    return as_field(a, solver.shape2, solver.dtype, solver.device)
