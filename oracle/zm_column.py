"""NumPy float64 column operators of the implicit / learned vertical-mixing step.

TEST INFRASTRUCTURE ONLY: this module is the checker for the HIP column
kernels.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the product never does.

The reference has no implicit vertical solve and no learned component; these
operators are new.  They are defined from the reference's own explicit
flux-form vertical operator (``_diffusive_interface_flux`` /
``_iface_flux_divergence``, ``src/zhenmode/model/solver/numerics/vertical.py:296-332``)
and its linear equation of state (``physics/eos.py:13-18``):

  mixing diffusivity at interface k+1/2 of a column with nw wet levels
      kappa = kappa_v + kappa_conv * [rho'_k > rho'_{k+1}]   for k + 1 < nw, else 0
  backward-Euler step  (I - dt D_kappa) X_new = X   for X in (T, S),
      D_kappa X[k] = (F[k-1/2] - F[k+1/2]) / dz_node[k],
      F[k+1/2] = -kappa_{k+1/2} (X[k+1] - X[k]) / dz_iface[k]   (zero-flux ends),
  solved per column with the Thomas algorithm; dry levels are held.

How each function is pinned (``tests/test_oracle_column.py``):
  * ``thomas``: ``scipy.linalg.solve_banded`` and dense ``numpy.linalg.solve``
    on random diagonally dominant systems;
  * ``implicit_mix``: the residual ``X_new - dt D(X_new) - X`` evaluated with
    the explicit operator of ``zm_oracle`` (an independent code path) is zero
    to round-off; constants are fixed points; the column integral
    ``sum dz_node X`` is conserved; kappa = 0 is the identity;
  * ``mlp_forward`` / ``mlp_backward``: central finite differences of the loss
    and a closed-form one-layer case;
  * ``integral_correction``: the corrected column integral equals the target
    by construction, checked in float64;
  * ``explicit_increment``: the flux form of ``mix_residual`` (e = -r(X, X)),
    telescoping column integrals and a two-level closed form;
  * ``surrogate_targets`` / ``learned_apply``: the exact targets fed back
    through the apply step reproduce the Thomas solution with zero residual,
    and a stable column's scale is dt kappa_v times its curvature.
"""

from __future__ import annotations

import math

import numpy as np

from .zm_oracle import density_anomaly

N_FEATURES = 253      # 5 x nz level features + 3 surface fluxes (nz = 50)
N_FEATURES_PAD = 256  # padded row length of the feature matrix
WIDTH = 256
N_OUT_PER_TRACER = 50


# --------------------------------------------------------------------------
# Thomas solve and the implicit mixing step
# --------------------------------------------------------------------------
def thomas(a, b, c, d):
    """Solve a[k] x[k-1] + b[k] x[k] + c[k] x[k+1] = d[k] along the last axis.

    a[..., 0] and c[..., -1] are ignored.  Float64 elimination.
    """
    a, b, c, d = (np.asarray(x, dtype=np.float64) for x in (a, b, c, d))
    n = d.shape[-1]
    cp = np.empty_like(d)
    dp = np.empty_like(d)
    cp[..., 0] = c[..., 0] / b[..., 0]
    dp[..., 0] = d[..., 0] / b[..., 0]
    for k in range(1, n):
        m = b[..., k] - a[..., k] * cp[..., k - 1]
        cp[..., k] = c[..., k] / m
        dp[..., k] = (d[..., k] - a[..., k] * dp[..., k - 1]) / m
    x = np.empty_like(d)
    x[..., n - 1] = dp[..., n - 1]
    for k in range(n - 2, -1, -1):
        x[..., k] = dp[..., k] - cp[..., k] * x[..., k + 1]
    return x


def mixing_kappa(T, S, p, kappa_v, kappa_conv):
    """(nx, ny, nz-1) interface diffusivities; zero at interfaces not wet on both sides."""
    wm = p.wet_mask_z
    wet_iface = (wm[..., :-1] > 0.5) & (wm[..., 1:] > 0.5)
    rho = density_anomaly(T, S, p)
    unstable = rho[..., :-1] > rho[..., 1:]
    kap = kappa_v + kappa_conv * unstable
    return np.where(wet_iface, kap, 0.0)


def tridiagonal(kappa, p, dt):
    """Coefficients (a, b, c) of I - dt D_kappa for every column."""
    nx, ny, nz1 = kappa.shape
    nz = nz1 + 1
    a = np.zeros((nx, ny, nz))
    c = np.zeros((nx, ny, nz))
    dzn, dzi = p.dz_node, p.dz_iface
    c[..., :-1] = -dt * kappa / (dzn[:-1] * dzi)
    a[..., 1:] = -dt * kappa / (dzn[1:] * dzi)
    b = 1.0 - a - c
    return a, b, c


def implicit_mix(T, S, p, dt, kappa_v, kappa_conv):
    """Backward-Euler vertical mixing of T and S over dt (dry levels held)."""
    kap = mixing_kappa(T, S, p, kappa_v, kappa_conv)
    a, b, c = tridiagonal(kap, p, dt)
    Tn = thomas(a, b, c, T)
    Sn = thomas(a, b, c, S)
    wet = p.wet_mask_z > 0.5
    return np.where(wet, Tn, T), np.where(wet, Sn, S)


def mix_residual(Xn, X, kappa, p, dt):
    """Xn - dt D_kappa(Xn) - X evaluated with explicit interface fluxes."""
    flux = -kappa * (Xn[..., 1:] - Xn[..., :-1]) / p.dz_iface
    zero = np.zeros_like(flux[..., :1])
    div = (np.concatenate([zero, flux], axis=-1) - np.concatenate([flux, zero], axis=-1)) / p.dz_node
    return Xn - dt * div - X


def column_integral(X, p):
    return np.sum(X * p.dz_node * p.wet_mask_z, axis=-1)


def integral_correction(dX, target, p):
    """Shift the wet levels of each column's increment so that
    sum(dz_node * dX) equals target exactly (uniform shift, float64)."""
    wm = p.wet_mask_z
    h = np.sum(p.dz_node * wm, axis=-1)
    err = column_integral(dX, p) - target
    shift = np.where(h > 0, err / np.where(h > 0, h, 1.0), 0.0)
    return (dX - shift[..., None]) * wm


# --------------------------------------------------------------------------
# the surrogate's increment formulation
# --------------------------------------------------------------------------
def explicit_increment(X, kappa, p, dt):
    """dt D_kappa X on wet levels (0 on dry levels), written through the
    tridiagonal matrix of the implicit step: e = X - (I - dt D) X, with
    (I - dt D) from ``tridiagonal``.  The flux form of ``mix_residual`` is an
    independent restatement of the same operator (e = -mix_residual(X, X))."""
    a, b, c = tridiagonal(kappa, p, dt)
    X = np.asarray(X, dtype=np.float64)
    AX = b * X
    AX[..., 1:] += a[..., 1:] * X[..., :-1]
    AX[..., :-1] += c[..., :-1] * X[..., 1:]
    return np.where(p.wet_mask_z > 0.5, X - AX, 0.0)


def mixing_scale(e):
    """Per-column max |e| over the levels (dry levels carry e = 0)."""
    return np.max(np.abs(e), axis=-1)


def surrogate_targets(T, S, Tn, Sn, p, dt, kappa_v, kappa_conv):
    """(nx, ny, 2 nz) targets: per tracer (X_new - X - e) / s, 0 where s = 0
    and on dry levels (X_new: the implicit solution of the same state)."""
    kap = mixing_kappa(T, S, p, kappa_v, kappa_conv)
    out = []
    for X, Xn in ((T, Tn), (S, Sn)):
        e = explicit_increment(X, kap, p, dt)
        s = mixing_scale(e)
        safe = np.where(s > 0, s, 1.0)[..., None]
        t = np.where((s > 0)[..., None] & (p.wet_mask_z > 0.5), (Xn - X - e) / safe, 0.0)
        out.append(t)
    return np.concatenate(out, axis=-1)


def learned_apply(T, S, z, p, dt, kappa_v, kappa_conv, tol_T, tol_S):
    """The surrogate's apply step from de-normalised outputs z (nx, ny, 2 nz):
    dX = e + s z per tracer, shifted so its column integral is zero, X_new =
    X + dX on wet levels; flags = columns whose implicit-system residual
    exceeds tol_T or tol_S.  Returns (T_new, S_new, flags, r_T, r_S)."""
    nz = T.shape[-1]
    kap = mixing_kappa(T, S, p, kappa_v, kappa_conv)
    wm = p.wet_mask_z
    new, res = [], []
    for X, zx in ((T, z[..., :nz]), (S, z[..., nz:])):
        e = explicit_increment(X, kap, p, dt)
        s = mixing_scale(e)
        dX = (e + s[..., None] * zx) * wm
        dX = integral_correction(dX, 0.0, p)
        Xn = np.where(wm > 0.5, X + dX, X)
        new.append(Xn)
        res.append(np.max(np.abs(mix_residual(Xn, X, kap, p, dt) * wm), axis=-1))
    flags = (res[0] > tol_T) | (res[1] > tol_S)
    return new[0], new[1], flags, res[0], res[1]


# --------------------------------------------------------------------------
# per-column MLP (253 -> 256 -> 256 -> 256 -> 100, GELU)
# --------------------------------------------------------------------------
def gelu(x):
    """x Phi(x) with the exact normal CDF (Hendrycks & Gimpel 2016)."""
    erf = np.vectorize(math.erf, otypes=[np.float64])
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    erf = np.vectorize(math.erf, otypes=[np.float64])
    cdf = 0.5 * (1.0 + erf(x / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)
    return cdf + x * pdf


def mlp_forward(X, Ws, bs):
    """Returns (Y, cache).  Ws[l] is (K_l, N_l), row-major like the device layout."""
    acts = [np.asarray(X, dtype=np.float64)]
    pre = []
    h = acts[0]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ np.asarray(W, dtype=np.float64) + np.asarray(b, dtype=np.float64)
        pre.append(z)
        if l + 1 < len(Ws):
            h = gelu(z)
            acts.append(h)
        else:
            h = z
    return h, (acts, pre)


def mse_loss(Y, target):
    return float(np.mean((Y - target) ** 2))


def mlp_backward(dY, Ws, cache):
    """Gradients (dWs, dbs) of a scalar loss with dL/dY = dY."""
    acts, pre = cache
    dWs, dbs = [None] * len(Ws), [None] * len(Ws)
    g = np.asarray(dY, dtype=np.float64)
    for l in range(len(Ws) - 1, -1, -1):
        dWs[l] = acts[l].T @ g
        dbs[l] = g.sum(axis=0)
        if l > 0:
            g = (g @ np.asarray(Ws[l], dtype=np.float64).T) * gelu_grad(pre[l - 1])
    return dWs, dbs


def bf16_round(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


# --------------------------------------------------------------------------
# column features (inputs of the surrogate)
# --------------------------------------------------------------------------
def column_features(T, S, u, v, p, q_heat, fw_flux, tau_mag):
    """(nx*ny, 256) raw features, columns in (j, i) order (i fastest).

    Per level: the T and S profiles as differences across each level's lower
    interface (X_k - X_{k+1}, 0 at and below the last wet level) with the
    surface value X_0 in the last slot; u, v; N^2 at the level's lower
    interface (g/rho0 * (rho'_{k+1} - rho'_k) / dz_iface[k], 0 at the last
    level and below the seafloor); then q_heat, the freshwater flux and |tau|;
    then three zero pads.  Dry levels contribute zeros.
    """
    nx, ny, nz = T.shape
    wm = p.wet_mask_z
    rho = density_anomaly(T, S, p)
    n2 = np.zeros_like(T)
    wet_iface = (wm[..., :-1] > 0.5) & (wm[..., 1:] > 0.5)
    n2[..., :-1] = np.where(wet_iface, 9.81 / 1025.0 * (rho[..., 1:] - rho[..., :-1]) / p.dz_iface, 0.0)

    def profile(X):
        out = np.zeros_like(X)
        out[..., :-1] = np.where(wet_iface, X[..., :-1] - X[..., 1:], 0.0)
        out[..., -1] = X[..., 0] * (wm[..., 0] > 0.5)
        return out

    blocks = [profile(T), profile(S), u * wm, v * wm, n2]
    X = np.zeros((ny, nx, N_FEATURES_PAD))
    for f, blk in enumerate(blocks):
        X[:, :, f * nz:(f + 1) * nz] = np.transpose(blk, (1, 0, 2))
    X[:, :, 5 * nz] = np.asarray(q_heat).T
    X[:, :, 5 * nz + 1] = np.asarray(fw_flux).T
    X[:, :, 5 * nz + 2] = np.asarray(tau_mag).T
    return X.reshape(ny * nx, N_FEATURES_PAD)
