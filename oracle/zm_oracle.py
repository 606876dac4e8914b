"""NumPy float64 restatement of ZhenMode's legacy mode-split time step.

TEST INFRASTRUCTURE ONLY: this module is the checker for the HIP step. Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product never does.

Scope: the default ("legacy") scheme of ``_step_impl`` with ``mode_split=True``
(reference ``src/zhenmode/model/solver/timestepping/step.py:341-458``), i.e.
column_geometry='legacy', process_time_scheme='legacy', linear EOS, linear
bottom drag, centered tracer advection, node-form vertical diffusion, no GM or
Redi, no dynamic ice. Arrays use the reference layout ``(nx, ny, nz)`` with
axis 0 periodic longitude, axis 1 latitude (index 0 = south), axis 2 depth
(index 0 = surface).

How each part is pinned (checked by ``tests/test_oracle_*.py``):
  * D1 diffusion / biharmonic: the reference's budget, self-adjointness,
    constant-preservation, ghost-independence and second-order accuracy
    assertions (``tests/fd/test_horizontal_tracer_diffusion.py:19-112``).
  * node-form d2/dz2 and ghost fill: the closed-form one-sided bottom stencil
    and sentinel independence (``tests/fd/test_vertical_bc.py:86-138``).
  * polar cap: checkerboard, mirror and monotone-taper assertions
    (``tests/fd/test_polar_cap.py:64-114``).
  * flux-form advection + Fz: a uniform tracer gives an exactly zero tendency
    for any velocity (identity stated in ``transport.py:114-119``).
  * de-alias: numpy.fft against an independent direct-DFT projection.
  * prefix sums, pressure, Laplacian, gradients, divergences: explicit
    index-loop evaluation on tiny grids, the trapezoid closed form for a linear
    density, and the exact-adjoint identity of divergence and gradient.
  * barotropic FB step: |lambda| = 1 gravity-wave neutrality and exact volume
    conservation on closed basins.
  * whole step: a resting uniform ocean is an exact fixed point; volume is
    conserved to round-off.
  * index loops written from the reference semantics, each with planted-
    mistake twins that must fail (``tests/test_oracle_pins.py``): the
    face-gated and layer-gated 3-D gradients, momentum advection with the
    ghost-filled -w d/dz, the vertical velocity, the centred flux-form tracer
    advection with its donor-cell vertical flux, the barotropic density
    forcing and the legacy barotropic velocity.
  * closed forms (same file): one forward-backward substep rotates (u, v) by
    -atan(f dt) and scales the speed by 1/sqrt(1 + (f dt)^2) times the drag
    1/(1 + r dt); the momentum residual of an eta field vanishes and that of
    a zonal shear flow is the bottom drag alone; the legacy convective
    sub-cycle's mean rate on two-level columns; the nu_h sub-cycle and the
    biharmonic on a zonal eigenmode; the barotropic projection identity; the
    prescribed-heat, bulk-heat and SST/SSS restoring integrals of one RK2
    step (``tests/fd/test_stage_budgets.py:48-82``).
  * production flags (conservative_kv, localize_conv, fct_adv,
    monotone_adv): the reference's FCT/monotone assertions and the flux-form
    telescoping identity (``tests/test_oracle_flags.py``).
  * column-divergence projection (project_adv_vel): operator symmetry,
    positive semi-definiteness and null constants, the stated diagonal against
    the brute-force diagonal (a planted dropped term fails), divergence
    removal to 1e-9, PCG against a dense solve, the resting-ocean fixed point
    (``tests/test_oracle_projection.py``).
  Parity unpinned: the bit-level agreement of the composite step with the
  reference JAX executable (XLA's FFT, cumsum association and transcendental
  rounding cannot be reproduced without importing JAX, which is absent).
"""

from __future__ import annotations

from collections import namedtuple
from types import SimpleNamespace

import numpy as np

# Physical constants (reference src/zhenmode/model/config.py:12-25).
G_EARTH = 9.81
RHO_0 = 1025.0
ALPHA_T = 2.0e-4
BETA_S = 7.6e-4
C_P = 3992.0
R_EARTH = 6371.0e3
OMEGA = 7.2921e-5

State = namedtuple("State", ["u", "v", "T", "S", "eta", "ice"])


# --------------------------------------------------------------------------
# index helpers: shifted copies along the periodic x axis and the closed y axis
# --------------------------------------------------------------------------
def east(a):
    """a[i+1] (periodic in x)."""
    return np.roll(a, -1, axis=0)


def west(a):
    """a[i-1] (periodic in x)."""
    return np.roll(a, 1, axis=0)


def north_edge(a):
    """a[min(j+1, ny-1)]: edge-replicated (mirror ghost = boundary value)."""
    return np.concatenate([a[:, 1:], a[:, -1:]], axis=1)


def south_edge(a):
    """a[max(j-1, 0)]: edge-replicated."""
    return np.concatenate([a[:, :1], a[:, :-1]], axis=1)


def north_wrap(a):
    """a[(j+1) % ny] (the reference's jnp.roll along latitude)."""
    return np.roll(a, -1, axis=1)


def south_wrap(a):
    return np.roll(a, 1, axis=1)


def _row(vec, ndim):
    """Broadcast a per-latitude vector along axis 1 of an ndim array."""
    shape = [1] * ndim
    shape[1] = -1
    return np.asarray(vec).reshape(shape)


# --------------------------------------------------------------------------
# parameters: restates geometry/fd_metrics.py:9-148 and factory.py:397-695
# --------------------------------------------------------------------------
def make_params(grid, physics, dt, forcing=None, *, T_atm=None, lambda_bulk=0.0,
                S_ref_surf=None, sss_restore_days=0.0,
                coastal_restore_mask=None, coastal_restore_days=0.0, coastal_restore_T=None,
                coastal_bulk_mask=None, coastal_bulk_lambda=0.0,
                coastal_kappa_h_mask=None, coastal_kappa_h=0.0,
                coastal_kappa_v_mask=None, coastal_kappa_v=0.0,
                sponge_days=0.0, sponge_cells=0, T_init=None, S_init=None,
                polar_cap_rows=2, polar_cap_taper=3, dt_bt=150.0, nu_nsub=None,
                ice_freeze_temp_c=-1.8, ice_salt_flux=0.0, conservative_kv=False, localize_conv=False,
                fct_adv=False, monotone_adv=False, project_adv_vel=False, projection_niter=None,
                projection_rtol=None, projection_preconditioner="none", projection_max_refinements=2,
                projection_eps=np.finfo(np.float64).eps):
    """Frozen parameter set of the legacy mode-split step (float64).

    The last four flags are the production tracer/mixing options of the
    reference factory (factory.py:49-51, 649-653)."""
    nx, ny, nz = int(grid.nx), int(grid.ny), int(grid.nz)
    p = SimpleNamespace(nx=nx, ny=ny, nz=nz, dt=float(dt))
    dx_2d = np.asarray(grid.dx_2d, dtype=np.float64)
    p.dx_2d = dx_2d
    p.dy = float(grid.dy)
    p.cos_lat = np.asarray(grid.cos_lat, dtype=np.float64)
    p.inv_dx = (1.0 / dx_2d)[:, :, None]
    p.inv_dy = 1.0 / p.dy
    p.inv_dx2 = (1.0 / dx_2d ** 2)[:, :, None]
    p.inv_dy2 = p.inv_dy ** 2
    p.f = np.asarray(grid.f, dtype=np.float64)
    p.wet_mask = np.asarray(grid.wet_mask, dtype=np.float64)
    p.wet_mask_z = np.asarray(grid.wet_mask_3d, dtype=np.float64)
    interior = np.ones(ny)
    interior[0] = interior[-1] = 0.0
    p.interior_mask_z = np.broadcast_to(interior[None, :], (nx, ny))[:, :, None].copy()

    z = np.asarray(grid.z, dtype=np.float64)
    step = np.diff(z)                                  # negative
    p.z = z
    p.dz_3d = np.asarray(grid.dz, dtype=np.float64)    # (nz-1,) positive
    p.dz_denom_interior = (z[2:] - z[1:-1]) + (z[1:-1] - z[:-2])
    p.dz_bnd_top = float(z[1] - z[0])
    p.dz_bnd_bot = float(z[-1] - z[-2])
    hm = np.abs(z[:-2] - z[1:-1])
    hp = np.abs(z[2:] - z[1:-1])
    p.d2z_hm, p.d2z_hp = hm, hp
    p.d2z_denom = hm * hp * (hm + hp) / 2.0
    p.d2z_h0_top = float(abs(z[1] - z[0]))
    p.d2z_h0_bot = float(abs(z[-1] - z[-2]))
    p.dz_surface = float(abs(z[0] - z[1]))
    p.dz_iface = np.abs(step)
    p.dz_node = np.concatenate([[abs(z[1] - z[0])],
                                0.5 * (np.abs(step)[:-1] + np.abs(step)[1:]),
                                [abs(z[-1] - z[-2])]])
    p.surface_mask = np.zeros(nz)
    p.surface_mask[0] = 1.0
    p.bottom_mask = np.zeros(nz)
    p.bottom_mask[-1] = 1.0

    p.nu_h, p.nu_v = float(physics.nu_h), float(physics.nu_v)
    p.kappa_h, p.kappa_v = float(physics.kappa_h), float(physics.kappa_v)
    p.kappa_conv = float(physics.kappa_conv)
    p.nu_bi, p.kappa_bi = float(physics.nu_bi), float(physics.kappa_bi)
    p.T_ref, p.S_ref = float(physics.T_ref), float(physics.S_ref)
    p.r_bot = float(physics.r_bot)

    if forcing is None:
        p.tau_x_2d = np.zeros((nx, ny))
        p.tau_y_2d = np.zeros((nx, ny))
        p.Q_heat_2d = np.zeros((nx, ny))
    else:
        p.tau_x_2d, p.tau_y_2d, p.Q_heat_2d = (np.array(a, dtype=np.float64) for a in forcing)

    p.H_sw = float(np.sum(p.dz_3d))
    p.dz_norm = p.dz_3d / p.H_sw

    if T_atm is not None and lambda_bulk > 0.0:
        p.T_atm = np.array(T_atm, dtype=np.float64)
        p.lambda_bulk = float(lambda_bulk)
    else:
        p.T_atm = np.zeros((nx, ny))
        p.lambda_bulk = 0.0
    if S_ref_surf is not None and sss_restore_days > 0.0:
        p.S_ref_2d = np.array(S_ref_surf, dtype=np.float64)
        p.restore_coef_S = 1.0 / (sss_restore_days * 86400.0)
    else:
        p.S_ref_2d = np.zeros((nx, ny))
        p.restore_coef_S = 0.0

    def masked(mask, coef):
        if mask is not None and coef > 0.0:
            return np.array(mask, dtype=np.float64) * coef
        return np.zeros((nx, ny))

    p.coastal_kappa_v_2d = masked(coastal_kappa_v_mask, coastal_kappa_v)
    p.coastal_kappa_h_2d = masked(coastal_kappa_h_mask, coastal_kappa_h)
    p.coastal_bulk_lambda_2d = masked(coastal_bulk_mask, coastal_bulk_lambda)
    if coastal_restore_mask is not None and coastal_restore_days > 0.0 and coastal_restore_T is not None:
        p.coastal_restore_coef_2d = np.array(coastal_restore_mask, dtype=np.float64) / (coastal_restore_days * 86400.0)
        p.coastal_restore_T_2d = np.array(coastal_restore_T, dtype=np.float64)
    else:
        p.coastal_restore_coef_2d = np.zeros((nx, ny))
        p.coastal_restore_T_2d = np.zeros((nx, ny))

    if sponge_days > 0.0 and sponge_cells > 0:
        r_max = 1.0 / (sponge_days * 86400.0)
        nc = int(sponge_cells)
        j = np.arange(ny)
        edge = np.minimum(j, ny - 1 - j)
        taper = np.where(edge < nc, 0.5 * (1.0 + np.cos(np.pi * edge / nc)), 0.0)
        p.sponge_rate_2d = np.broadcast_to((r_max * taper)[None, :], (nx, ny)).copy()
        wm3 = p.wet_mask_z
        if T_init is not None:
            T_clim = np.array(T_init, dtype=np.float64) * wm3 + (1.0 - wm3) * p.T_ref
            S_clim = (np.array(S_init, dtype=np.float64) if S_init is not None
                      else np.full_like(T_clim, p.S_ref))
            S_clim = S_clim * wm3 + (1.0 - wm3) * p.S_ref
        else:
            T_clim = np.full((nx, ny, nz), p.T_ref)
            S_clim = np.full((nx, ny, nz), p.S_ref)
        p.T_clim_3d, p.S_clim_3d = T_clim, S_clim
    else:
        p.sponge_rate_2d = np.zeros((nx, ny))
        p.T_clim_3d = np.zeros((nx, ny, nz))
        p.S_clim_3d = np.zeros((nx, ny, nz))
    p.sponge_rate = p.sponge_rate_2d[:, :, None]

    p.polar_cap_rows = int(polar_cap_rows)
    p.polar_cap_taper = int(polar_cap_taper)
    keep = max(1, int((nx // 2) * 2 / 3))
    p.dealias_keep = keep
    p.dealias_lon_mask = (np.abs(np.fft.fftfreq(nx) * nx) <= keep).astype(np.float64)

    p.n_subcyc = max(1, int(round(dt / dt_bt)))
    p.dt_bt = dt / p.n_subcyc
    if nu_nsub == "cfl":
        inv_dx2_max = float(np.max(dx_2d ** -2))
        nu_nsub = max(1, int(np.ceil(p.nu_h * (dt / 2.0) * (inv_dx2_max + 1.0 / p.dy ** 2) / 0.25)))
    elif nu_nsub is not None:
        nu_nsub = max(1, int(nu_nsub))
    p.nu_nsub = nu_nsub
    min_dz = float(np.min(p.dz_3d))
    p.conv_nsub = max(1, int(np.ceil(p.kappa_conv * dt / (0.4 * min_dz ** 2))))
    p.adv_nsub = max(1, int(np.ceil(dt * 4.0e-3 / (0.5 * min_dz))))
    p.ice_freeze_temp_c = float(ice_freeze_temp_c)
    p.ice_salt_flux = float(ice_salt_flux)
    p.conservative_kv = bool(conservative_kv)
    p.localize_conv = bool(localize_conv)
    p.fct_adv = bool(fct_adv)
    p.monotone_adv = bool(monotone_adv)
    # column-divergence projection of the stage-2 tracer velocity (factory.py:368-386)
    p.project_adv_vel = bool(project_adv_vel)
    p.projection_niter = 150 if projection_niter is None else int(projection_niter)
    p.projection_rtol = max(float(projection_rtol or 1e-12), 32.0 * float(projection_eps))
    p.projection_eps = float(projection_eps)
    p.projection_preconditioner = projection_preconditioner
    p.projection_max_refinements = int(projection_max_refinements)
    return p


def init_state(p, T_init=None, S_init=None):
    """init_state (reference factory.py:766-788): land/ghost -> T_ref, S_ref."""
    shape = (p.nx, p.ny, p.nz)
    u = np.zeros(shape)
    v = np.zeros(shape)
    eta = np.zeros(shape[:2])
    if T_init is not None:
        T = np.array(T_init, dtype=np.float64)
        S = np.array(S_init, dtype=np.float64) if S_init is not None else np.full_like(T, p.S_ref)
    else:
        T = np.full(shape, p.T_ref)
        S = np.full(shape, p.S_ref)
    wm = p.wet_mask_z
    T = T * wm + (1.0 - wm) * p.T_ref
    S = S * wm + (1.0 - wm) * p.S_ref
    return State(u, v, T, S, eta, np.zeros_like(eta))


# --------------------------------------------------------------------------
# horizontal operators (reference numerics/horizontal.py)
# --------------------------------------------------------------------------
def sin_lat_rows(p):
    """Signed sin(lat) per row, recovered from f[0, :] (horizontal.py:210)."""
    return p.f[0, :] / (2.0 * OMEGA)


def _laplacian_tables(p):
    """Face gates (products of 0/1 wet masks, so applying them as one product is
    exact) and the metric row of the legacy Laplacian, computed once per mask."""
    tables = getattr(p, "_laplacian_tables", None)
    if tables is None or tables[0] is not p.wet_mask_z:
        wm = p.wet_mask_z
        tables = (wm, wm * east(wm), wm * west(wm), north_edge(wm) * wm, wm * south_edge(wm),
                  _row(-((sin_lat_rows(p) / p.cos_lat) / R_EARTH), 3), 0.5 * p.inv_dy)
        p._laplacian_tables = tables
    return tables


def laplacian_h(u, p):
    """Legacy spherical Laplacian (horizontal.py:181-215): gated second
    differences in x and y (edge-replicated y ghost) minus (tan(lat)/R) du/dy."""
    _, gate_e, gate_w, gate_n, gate_s, minus_tan_over_r, half_inv_dy = _laplacian_tables(p)
    ue, uw = east(u), west(u)
    d2x = ((ue - u) * gate_e - (u - uw) * gate_w) * p.inv_dx2
    un, us = north_edge(u), south_edge(u)
    d2y = ((un - u) * gate_n - (u - us) * gate_s) * p.inv_dy2
    du_dy = (un - us) * half_inv_dy
    return d2x + d2y + du_dy * minus_tan_over_r


def cos_face_north(p):
    """cos at face j+1/2 with the reference's wrap at j = ny-1."""
    return 0.5 * (p.cos_lat + np.roll(p.cos_lat, -1))


def cos_face_south(p):
    return 0.5 * (p.cos_lat + np.roll(p.cos_lat, 1))


def diffusion_flux_form(tracer, kappa, p):
    """Wet-face flux-form diffusion D1 (horizontal.py:217-238)."""
    wet = p.wet_mask_z
    zonal = (0.5 * (kappa + east(kappa)) * (east(tracer) - tracer) * p.inv_dx * wet * east(wet))
    nt, nk, nw = north_edge(tracer), north_edge(kappa), north_edge(wet)
    merid = (0.5 * (kappa + nk) * (nt - tracer) * p.inv_dy * wet * nw
             * _row(cos_face_north(p), 3))
    merid[:, -1, :] = 0.0
    incoming = np.roll(merid, 1, axis=1)
    incoming[:, 0, :] = 0.0
    return ((zonal - west(zonal)) * p.inv_dx
            + (merid - incoming) * p.inv_dy / _row(p.cos_lat, 3))


def tracer_diffusion(tracer, p):
    kappa = np.broadcast_to((p.kappa_h + p.coastal_kappa_h_2d)[:, :, None], tracer.shape)
    return diffusion_flux_form(tracer, kappa, p)


def tracer_biharmonic(tracer, p):
    unit = np.ones(tracer.shape)
    return diffusion_flux_form(diffusion_flux_form(tracer, unit, p), unit, p)


def biharmonic_h(u, p):
    return laplacian_h(laplacian_h(u, p), p)


def divergence_h(u, v, p):
    """Face-flux layer divergence (horizontal.py:255-281)."""
    wm = p.wet_mask_z
    fx = 0.5 * (u + east(u)) * wm * east(wm)
    div_x = (fx - west(fx)) * p.inv_dx
    fy = (0.5 * (v + north_edge(v)) * wm * north_edge(wm) * _row(0.5 * (p.cos_lat + np.roll(p.cos_lat, -1)), 3))
    fy[:, -1] = 0.0
    div_y = (fy - np.roll(fy, 1, axis=1)) * p.inv_dy / _row(p.cos_lat, 3)
    return div_x + div_y


def gradient_conservative_3d(field, p):
    """Layer-gated adjoint gradient (horizontal.py:126-151). field may be (nx,ny,1)."""
    wm = p.wet_mask_z
    open_xp = wm * east(wm)
    open_xm = wm * west(wm)
    gx = p.inv_dx * 0.5 * ((east(field) - field) * open_xp + (field - west(field)) * open_xm)
    open_yp = wm * north_wrap(wm)
    open_yp[:, -1] = 0.0
    open_ym = wm * south_wrap(wm)
    open_ym[:, 0] = 0.0
    d_yp = (north_wrap(field) - field) * open_yp * _row(cos_face_north(p), 3)
    d_ym = (field - south_wrap(field)) * open_ym * _row(cos_face_south(p), 3)
    gy = (p.inv_dy / _row(p.cos_lat, 3)) * 0.5 * (d_yp + d_ym)
    return gx, gy


def gradient_face_gated_3d(field, p):
    """Face-gated centered gradient with mirror y ghost (horizontal.py:331-360)."""
    wm = p.wet_mask_z
    gx = p.inv_dx * 0.5 * ((east(field) - field) * (wm * east(wm))
                           + (field - west(field)) * (wm * west(wm)))
    fn, fs = north_edge(field), south_edge(field)
    wn, ws = north_edge(wm), south_edge(wm)
    gy = p.inv_dy * 0.5 * ((fn - field) * (wn * wm) + (field - fs) * (wm * ws))
    return gx, gy


def divergence_conservative(ubt, vbt, p):
    """2-D mass-conserving divergence (horizontal.py:69-97)."""
    wm = p.wet_mask
    uface = 0.5 * (ubt + east(ubt)) * (wm * east(wm))
    div_x = (uface - west(uface)) * p.inv_dx[..., 0]
    vopen = wm * north_wrap(wm)
    vopen[:, -1] = 0.0
    vface = 0.5 * (vbt + north_wrap(vbt)) * vopen
    fcos = vface * cos_face_north(p)[None, :]
    fin = south_wrap(fcos)
    fin[:, 0] = 0.0
    div_y = (fcos - fin) * p.inv_dy / p.cos_lat[None, :]
    return div_x + div_y


def gradient_conservative(eta, p):
    """2-D adjoint gradient (horizontal.py:99-124)."""
    wm = p.wet_mask
    inv_dx = p.inv_dx[..., 0]
    gx = inv_dx * 0.5 * ((east(eta) - eta) * (wm * east(wm)) + (eta - west(eta)) * (wm * west(wm)))
    open_yp = wm * north_wrap(wm)
    open_yp[:, -1] = 0.0
    open_ym = wm * south_wrap(wm)
    open_ym[:, 0] = 0.0
    d_yp = (north_wrap(eta) - eta) * open_yp * cos_face_north(p)[None, :]
    d_ym = (eta - south_wrap(eta)) * open_ym * cos_face_south(p)[None, :]
    gy = (p.inv_dy / p.cos_lat[None, :]) * 0.5 * (d_yp + d_ym)
    return gx, gy


def dealias(field, p):
    """2/3-rule x projection via numpy.fft plus [1,4,6,4,1]/16 y filter (horizontal.py:283-305)."""
    spec = np.fft.fft(field, axis=0) * p.dealias_lon_mask.reshape((-1,) + (1,) * (field.ndim - 1))
    flon = np.real(np.fft.ifft(spec, axis=0))
    fp = np.concatenate([flon[:, :1], flon[:, :1], flon, flon[:, -1:], flon[:, -1:]], axis=1)
    return (fp[:, :-4] + 4.0 * fp[:, 1:-3] + 6.0 * fp[:, 2:-2] + 4.0 * fp[:, 3:-1] + fp[:, 4:]) / 16.0


def polar_cap_weights(ncap, ntaper):
    """Pole-inward blend weights (horizontal.py:362-385)."""
    if ncap <= 0:
        return np.zeros(0)
    full = np.ones(ncap)
    if ntaper <= 0:
        return full
    x = np.linspace(0.0, np.pi / 2.0, ntaper + 2)[1:-1]
    return np.concatenate([full, np.cos(x) ** 2])


def apply_polar_cap(field, wm, p):
    """Legacy tapered wet-point zonal-mean cap (horizontal.py:387-451, :427-434)."""
    ncap = p.polar_cap_rows
    if ncap <= 0:
        return field
    nb = ncap + p.polar_cap_taper
    weights = polar_cap_weights(ncap, p.polar_cap_taper).astype(field.dtype)

    def band(f, w, wts):
        s = f * w
        count = np.maximum(np.sum(w, axis=0, keepdims=True), 1.0)
        zmean = np.broadcast_to(np.sum(s, axis=0, keepdims=True) / count, f.shape) * w
        wb = wts.reshape((1, nb) + (1,) * (field.ndim - 2))
        return wb * zmean + (1.0 - wb) * s

    south = band(field[:, :nb], wm[:, :nb], weights)
    north = band(field[:, -nb:], wm[:, -nb:], weights[::-1])
    return np.concatenate([south, field[:, nb:-nb], north], axis=1)


# --------------------------------------------------------------------------
# vertical operators (reference numerics/vertical.py)
# --------------------------------------------------------------------------
def kbot_index(p):
    """Deepest wet level per column; 0 for land columns (vertical.py:131-132)."""
    idx = (p.wet_mask_z > 0.5) * np.arange(p.nz)
    return np.max(idx, axis=-1)


def fill_ghost_bottom(u, p):
    """Replicate the deepest wet value into the ghost layers (vertical.py:122-135)."""
    kb = kbot_index(p)[..., None]
    kk = np.arange(p.nz)[None, None, :]
    bottom = np.take_along_axis(u, kb, axis=-1)
    return np.where(kk >= kb, bottom, u)


def d_dz(u, p):
    """One-sided ends, centered interior (vertical.py:267-272)."""
    inner = (u[..., 2:] - u[..., :-2]) / p.dz_denom_interior
    top = (u[..., 1:2] - u[..., 0:1]) / p.dz_bnd_top
    bot = (u[..., -1:] - u[..., -2:-1]) / p.dz_bnd_bot
    return np.concatenate([top, inner, bot], axis=-1)


def d2_dz2(u, p):
    """Node-form non-uniform second derivative with ghost fill (vertical.py:275-293)."""
    u = fill_ghost_bottom(u, p)
    inner = (u[..., 2:] * p.d2z_hm + u[..., :-2] * p.d2z_hp
             - u[..., 1:-1] * (p.d2z_hm + p.d2z_hp)) / p.d2z_denom
    top = 2.0 * (u[..., 1:2] - u[..., 0:1]) / (p.d2z_h0_top ** 2)
    bot = 2.0 * (u[..., -2:-1] - u[..., -1:]) / (p.d2z_h0_bot ** 2)
    return np.concatenate([top, inner, bot], axis=-1)


def iface_flux_divergence(flux, p):
    """(F[k-1/2] - F[k+1/2]) / dz_node with zero-flux ends (vertical.py:296-306)."""
    zero = np.zeros_like(flux[..., :1])
    up = np.concatenate([zero, flux], axis=-1)
    dn = np.concatenate([flux, zero], axis=-1)
    return (up - dn) / p.dz_node


def diffusive_interface_flux(tracer, kappa, p):
    t = fill_ghost_bottom(tracer, p)
    return -kappa * (t[..., 1:] - t[..., :-1]) / p.dz_iface


# --------------------------------------------------------------------------
# physics (reference physics/{eos,vertical,surface}.py)
# --------------------------------------------------------------------------
def density_anomaly(T, S, p):
    """Linear EOS (eos.py:13-18)."""
    return RHO_0 * (-ALPHA_T * (T - p.T_ref) + BETA_S * (S - p.S_ref))


def unstable_interfaces(T, S, p):
    """Per-interface gate: wet/wet with denser water above (physics/vertical.py:63-82)."""
    wm = p.wet_mask_z
    wet_iface = (wm[..., :-1] > 0.5) & (wm[..., 1:] > 0.5)
    rho = density_anomaly(T, S, p) * wm
    return (rho[..., :-1] > rho[..., 1:]) & wet_iface


def convective_mask(T, S, p):
    """Column gate: any unstable interface (physics/vertical.py:63-82)."""
    return np.any(unstable_interfaces(T, S, p), axis=-1, keepdims=True)


def conv_flux_tendency(tracer, gate, kappa, p, iface_gate=None):
    """Interface-flux convective mixing (physics/vertical.py:13-37); with
    iface_gate (localize_conv) a flux also needs its own interface unstable."""
    if kappa <= 0.0:
        return np.zeros_like(tracer)
    flux = diffusive_interface_flux(tracer, kappa, p)
    wm = p.wet_mask_z
    allow = gate & (wm[..., :-1] > 0.5) & (wm[..., 1:] > 0.5)
    if iface_gate is not None:
        allow = allow & iface_gate
    flux = np.where(allow, flux, 0.0)
    return iface_flux_divergence(flux, p) * wm


def d2_dz2_flux(tracer, kappa, p):
    """Interface-flux vertical diffusion (numerics/vertical.py:317-332)."""
    flux = diffusive_interface_flux(tracer, kappa, p)
    wm = p.wet_mask_z
    flux = np.where((wm[..., :-1] > 0.5) & (wm[..., 1:] > 0.5), flux, 0.0)
    return iface_flux_divergence(flux, p) * wm


def vertical_diffusion(tracer, p):
    """kappa_v vertical diffusion of a tracer (physics/vertical.py:45-53): node
    form by default, interface-flux form with conservative_kv."""
    kappa = p.kappa_v + p.coastal_kappa_v_2d[:, :, None]
    if getattr(p, "conservative_kv", False):
        if np.all(kappa == 0.0):
            return np.zeros_like(tracer)
        return d2_dz2_flux(tracer, kappa, p)
    return kappa * d2_dz2(tracer, p)


def vertical_momentum_diffusion(u, p):
    return p.nu_v * d2_dz2(u, p)


def surface_heat_weights(p):
    """surface_mask * wet / dz_surface (physics/surface.py:8-12)."""
    return p.surface_mask * p.wet_mask_z / p.dz_surface


# --------------------------------------------------------------------------
# transport and pressure (reference dynamics/{transport,pressure}.py)
# --------------------------------------------------------------------------
def vertical_transport_iface(u, v, p):
    """Fz (nx,ny,nz+1), downward positive, accumulated from the seafloor (transport.py:110-133)."""
    integrand = divergence_h(u, v, p) * p.dz_node
    acc = np.cumsum(integrand[..., ::-1], axis=-1)[..., ::-1]
    return np.concatenate([acc, np.zeros_like(acc[..., :1])], axis=-1)


def vertical_velocity(u, v, p):
    """w = -Fz[:nz] * wet (transport.py:149-166, dual_face_v1)."""
    return -vertical_transport_iface(u, v, p)[..., :-1] * p.wet_mask_z


def barotropic_velocity(u, v, p):
    """Legacy depth mean over all nodes with dz_norm weights (transport.py:135-143)."""
    ubt = np.sum(0.5 * (u[..., :-1] + u[..., 1:]) * p.dz_norm, axis=-1)
    vbt = np.sum(0.5 * (v[..., :-1] + v[..., 1:]) * p.dz_norm, axis=-1)
    return ubt, vbt


def advection_flux_form(u, v, w, p):
    """Advective-form momentum advection, de-aliased and masked (transport.py:169-197)."""
    du_dz = d_dz(fill_ghost_bottom(u, p), p)
    dv_dz = d_dz(fill_ghost_bottom(v, p), p)
    du_dx, du_dy = gradient_face_gated_3d(u, p)
    dv_dx, dv_dy = gradient_face_gated_3d(v, p)
    adv_u = -(u * du_dx + v * du_dy + w * du_dz)
    adv_v = -(u * dv_dx + v * dv_dy + w * dv_dz)
    return dealias(adv_u, p) * p.wet_mask_z, dealias(adv_v, p) * p.wet_mask_z


def limited_slope(T, wm, axis):
    """Minmod slope from wet neighbours only (transport.py:199-213); y uses
    edge replication at the walls, x is periodic."""
    if axis == 1:
        prev, nxt = south_edge(T), north_edge(T)
        pw, nw = south_edge(wm), north_edge(wm)
    else:
        prev, nxt = west(T), east(T)
        pw, nw = west(wm), east(wm)
    left, right = T - prev, nxt - T
    signs = np.sign(left) + np.sign(right)
    return 0.5 * signs * np.minimum(np.abs(left), np.abs(right)) * wm * pw * nw


def horizontal_tracer_fluxes(T, u, v, p):
    """Face-gated tracer fluxes (transport.py:224-290): centred face values by
    default, the minmod-limited upwind reconstruction with fct_adv, donor cell
    with monotone_adv (fct_adv takes precedence)."""
    wm = p.wet_mask_z
    ux = 0.5 * (u + east(u))
    vy = 0.5 * (v + north_edge(v))
    if getattr(p, "fct_adv", False):
        sx = limited_slope(T, wm, 0)
        tx = np.where(ux >= 0.0, T + 0.5 * sx, east(T) - 0.5 * east(sx))
        sy = limited_slope(T, wm, 1)
        ty = np.where(vy >= 0.0, T + 0.5 * sy, north_edge(T) - 0.5 * north_edge(sy))
    elif getattr(p, "monotone_adv", False):
        tx = np.where(ux >= 0.0, T, east(T))
        ty = np.where(vy >= 0.0, T, north_edge(T))
    else:
        tx = 0.5 * (T + east(T))
        ty = 0.5 * (T + north_edge(T))
    fx = ux * tx * (wm * east(wm))
    gate_y = wm * north_edge(wm)
    fy = vy * ty * gate_y * _row(cos_face_north(p), 3)
    fy[:, -1] = 0.0
    return fx, fy


def advection_scalar(T, u, v, Fz, p):
    """Flux-form tracer advection with donor-cell vertical flux (transport.py:365-475, legacy)."""
    wm = p.wet_mask_z
    fx, fy = horizontal_tracer_fluxes(T, u, v, p)
    fy_up = np.roll(fy, 1, axis=1)
    fy_up[:, 0] = 0.0
    hdiv = (fx - west(fx)) * p.inv_dx[..., 0:1] + (fy - fy_up) * p.inv_dy / _row(p.cos_lat, 3)
    t_deep = fill_ghost_bottom(T, p)[..., 1:]
    t_shallow = T[..., :-1]
    wet_iface = wm[..., :-1] * wm[..., 1:]
    f_int = Fz[..., 1:-1]
    flux_int = f_int * np.where(f_int > 0.0, t_shallow, t_deep) * wet_iface
    flux_top = Fz[..., :1] * T[..., :1]
    up = np.concatenate([flux_top, flux_int], axis=-1)
    dn = np.concatenate([flux_int, np.zeros_like(flux_int[..., :1])], axis=-1)
    div_z = (dn - up) / p.dz_node
    return -(hdiv + div_z) * wm


def baroclinic_pressure(T, S, p):
    """Trapezoid prefix sum of the masked density anomaly (pressure.py:32-42)."""
    rho = density_anomaly(T, S, p) * p.wet_mask_z
    dp = G_EARTH * (0.5 * (rho[..., :-1] + rho[..., 1:])) * p.dz_3d
    out = np.zeros_like(T)
    out[..., 1:] = np.cumsum(dp, axis=-1)
    return out


def hydrostatic_pressure(T, S, eta, p):
    """p = rho0 g eta + p_bc (pressure.py:62-74)."""
    return RHO_0 * G_EARTH * eta[:, :, None] + baroclinic_pressure(T, S, p)


def pressure_gradient(T, S, eta, p):
    gx, gy = gradient_conservative_3d(hydrostatic_pressure(T, S, eta, p), p)
    return -gx / RHO_0, -gy / RHO_0


def bt_rho_pgf(T, S, p):
    """Wet-interface depth-weighted baroclinic PGF / (rho0 H_sw) (pressure.py:144-180, legacy)."""
    pbc = baroclinic_pressure(T, S, p)
    gx, gy = gradient_conservative_3d(pbc, p)
    wm = p.wet_mask_z
    wet_iface = wm[..., :-1] * wm[..., 1:]
    lay_x = 0.5 * (gx[..., :-1] + gx[..., 1:])
    lay_y = 0.5 * (gy[..., :-1] + gy[..., 1:])
    fx = -np.sum(lay_x * wet_iface * p.dz_3d, axis=-1) / (RHO_0 * p.H_sw)
    fy = -np.sum(lay_y * wet_iface * p.dz_3d, axis=-1) / (RHO_0 * p.H_sw)
    return fx, fy


# --------------------------------------------------------------------------
# tendencies (reference dynamics/tendencies.py)
# --------------------------------------------------------------------------
def tracer_tendency(u, v, T, S, p, rec=None):
    """dT/dt, dS/dt, legacy defaults (tendencies.py:94-219); ``rec`` receives
    the surface sources, process tendencies and surface advective fluxes
    (tendencies.py:211-214)."""
    wm = p.wet_mask_z
    Fz = vertical_transport_iface(u, v, p)
    n_a = int(p.adv_nsub)
    if n_a > 1:
        dts = p.dt / n_a
        tT, tS = T, S
        accT = accS = 0.0
        for _ in range(n_a):
            aT = advection_scalar(tT, u, v, Fz, p)
            aS = advection_scalar(tS, u, v, Fz, p)
            accT, accS = accT + aT, accS + aS
            tT, tS = tT + aT * dts, tS + aS * dts
        adv_T, adv_S = accT / n_a, accS / n_a
    else:
        adv_T = advection_scalar(T, u, v, Fz, p)
        adv_S = advection_scalar(S, u, v, Fz, p)
    diff_h_T, diff_h_S = tracer_diffusion(T, p), tracer_diffusion(S, p)
    diff_v_T, diff_v_S = vertical_diffusion(T, p), vertical_diffusion(S, p)

    unstable = unstable_interfaces(T, S, p)
    gate = np.any(unstable, axis=-1, keepdims=True)
    iface_gate = unstable if getattr(p, "localize_conv", False) else None
    n_c = int(p.conv_nsub)
    if n_c > 1:
        kappa_c = p.kappa_conv / n_c
        h_c = p.dt / n_c
        tT, tS = T, S
        accT = accS = 0.0
        for _ in range(n_c):
            ttT = conv_flux_tendency(tT, gate, kappa_c, p, iface_gate)
            ttS = conv_flux_tendency(tS, gate, kappa_c, p, iface_gate)
            accT, accS = accT + ttT, accS + ttS
            tT, tS = tT + ttT * h_c, tS + ttS * h_c
        conv_T, conv_S = accT / n_c, accS / n_c
    else:
        conv_T = conv_flux_tendency(T, gate, p.kappa_conv, p, iface_gate)
        conv_S = conv_flux_tendency(S, gate, p.kappa_conv, p, iface_gate)

    heat_factor = surface_heat_weights(p) / (RHO_0 * C_P)
    sst = T[:, :, 0:1]
    heat_T = p.Q_heat_2d[:, :, None] * heat_factor
    bulk_T = p.lambda_bulk * (p.T_atm[:, :, None] - sst) * heat_factor
    cbulk_T = p.coastal_bulk_lambda_2d[:, :, None] * (p.T_atm[:, :, None] - sst) * heat_factor
    rest_S = p.restore_coef_S * (p.S_ref_2d[:, :, None] - S[:, :, 0:1]) * p.surface_mask
    ice_salt = p.ice_salt_flux * (sst <= p.ice_freeze_temp_c) * p.surface_mask
    rest_T = p.coastal_restore_coef_2d[:, :, None] * (p.coastal_restore_T_2d[:, :, None] - sst) * p.surface_mask

    dTdt = (adv_T + diff_h_T + diff_v_T + heat_T + bulk_T + cbulk_T + conv_T + rest_T) * wm
    dSdt = (adv_S + diff_h_S + diff_v_S + conv_S + rest_S + ice_salt) * wm
    if rec is not None:
        zero = np.zeros_like(T)
        rec.surface_sources((heat_T, bulk_T, cbulk_T, rest_T, rest_S, ice_salt))
        rec.nonlinear_terms(((adv_T, adv_S), (conv_T, conv_S), (zero, zero), (zero, zero)))
        rec.advection_boundary_fluxes(Fz[..., 0] * T[..., 0], Fz[..., 0] * S[..., 0])
    return dTdt, dSdt, (diff_h_T, diff_h_S, diff_v_T, diff_v_S)


def tracer_terms(u, v, T, S, p):
    """_tracer_terms (tendencies.py:74-92): the (6, nx, ny, nz) stack of the
    temperature's [advection, horizontal diffusion, vertical diffusion,
    convection (one pass with kappa_conv), GM, Redi] tendencies, wet-masked
    (GM and Redi are zero in this configuration)."""
    Fz = vertical_transport_iface(u, v, p)
    adv = advection_scalar(T, u, v, Fz, p)
    diff_h = tracer_diffusion(T, p)
    diff_v = vertical_diffusion(T, p)
    unstable = unstable_interfaces(T, S, p)
    gate = np.any(unstable, axis=-1, keepdims=True)
    conv = conv_flux_tendency(T, gate, p.kappa_conv, p, unstable if getattr(p, "localize_conv", False) else None)
    zero = np.zeros_like(T)
    return np.stack([t * p.wet_mask_z for t in (adv, diff_h, diff_v, conv, zero, zero)])


def tracer_residual(u, v, T, S, p, rec=None):
    """Tendency minus the L-step's linear diffusion, add-then-subtract (tendencies.py:227-245)."""
    dTdt, dSdt, (dhT, dhS, dvT, dvS) = tracer_tendency(u, v, T, S, p, rec=rec)
    dTdt = dTdt - dhT
    dSdt = dSdt - dhS
    dTdt = dTdt - dvT
    dSdt = dSdt - dvS
    return dTdt, dSdt


def momentum_tendency(u, v, T, S, eta, p):
    """du/dt, dv/dt and the shared terms (tendencies.py:39-72)."""
    wm = p.wet_mask_z
    w = vertical_velocity(u, v, p)
    adv_u, adv_v = advection_flux_form(u, v, w, p)
    f3 = p.f[:, :, None]
    pgf_x, pgf_y = pressure_gradient(T, S, eta, p)
    lap_u, lap_v = laplacian_h(u, p), laplacian_h(v, p)
    dvu, dvv = vertical_momentum_diffusion(u, p), vertical_momentum_diffusion(v, p)
    wind_factor = 1.0 / (RHO_0 * p.dz_surface)
    wind_u = p.tau_x_2d[:, :, None] * wind_factor * p.surface_mask
    wind_v = p.tau_y_2d[:, :, None] * wind_factor * p.surface_mask
    bot_u = -p.r_bot * u * p.bottom_mask
    bot_v = -p.r_bot * v * p.bottom_mask
    dudt = (adv_u + f3 * v + pgf_x + p.nu_h * lap_u + dvu + wind_u + bot_u) * wm
    dvdt = (adv_v + (-f3 * u) + pgf_y + p.nu_h * lap_v + dvv + wind_v + bot_v) * wm
    return dudt, dvdt, (lap_u, lap_v, dvu, dvv)


def momentum_residual(u, v, T, S, eta, p):
    """Tendency minus linear parts (tendencies.py:247-319, legacy mode-split branch)."""
    dudt, dvdt, (lap_u, lap_v, dvu, dvv) = momentum_tendency(u, v, T, S, eta, p)
    f3 = p.f[:, :, None]
    dudt = dudt - p.nu_h * lap_u
    dvdt = dvdt - p.nu_h * lap_v
    dudt = dudt - dvu
    dvdt = dvdt - dvv
    dudt = dudt - f3 * v
    dvdt = dvdt + f3 * u
    ex, ey = gradient_conservative_3d(eta[:, :, None], p)
    dudt = dudt + G_EARTH * ex
    dvdt = dvdt + G_EARTH * ey
    fx, fy = bt_rho_pgf(T, S, p)
    dudt = dudt - fx[:, :, None]
    dvdt = dvdt - fy[:, :, None]
    dudt = dudt - (p.tau_x_2d / (RHO_0 * p.H_sw))[:, :, None]
    dvdt = dvdt - (p.tau_y_2d / (RHO_0 * p.H_sw))[:, :, None]
    return dudt * p.wet_mask_z, dvdt * p.wet_mask_z


# --------------------------------------------------------------------------
# time stepping (reference timestepping/step.py)
# --------------------------------------------------------------------------
def coriolis_rotation(u, v, f, dt):
    """Exact rotation by f*dt on every node (step.py:42-54)."""
    angle = f[:, :, None] * dt
    c, s = np.cos(angle), np.sin(angle)
    return c * u + s * v, -s * u + c * v


def nu_subcycle(u, v, p, dt_half):
    """n_nu explicit Laplacian substeps (step.py:94-107)."""
    n_nu = max(1, int(p.n_subcyc if p.nu_nsub is None else p.nu_nsub))
    dts = dt_half / n_nu
    for _ in range(n_nu):
        u, v = (u + p.nu_h * laplacian_h(u, p) * dts,
                v + p.nu_h * laplacian_h(v, p) * dts)
    return u, v


def linear_half_step(state, p, dt_half, rec=None):
    """Mode-split linear half-step (step.py:76-181); ``rec`` records the
    linear_diffusion and sponge stages (step.py:138-161)."""
    u, v = nu_subcycle(state.u, state.v, p, dt_half)
    T = state.T + tracer_diffusion(state.T, p) * dt_half
    S = state.S + tracer_diffusion(state.S, p) * dt_half
    if p.nu_bi > 0.0:
        u = u - p.nu_bi * biharmonic_h(state.u, p) * dt_half
        v = v - p.nu_bi * biharmonic_h(state.v, p) * dt_half
    if p.kappa_bi > 0.0:
        T = T - p.kappa_bi * tracer_biharmonic(state.T, p) * dt_half
        S = S - p.kappa_bi * tracer_biharmonic(state.S, p) * dt_half
    u = u + vertical_momentum_diffusion(state.u, p) * dt_half
    v = v + vertical_momentum_diffusion(state.v, p) * dt_half
    T = T + vertical_diffusion(state.T, p) * dt_half
    S = S + vertical_diffusion(state.S, p) * dt_half
    wm = p.wet_mask_z
    u = u * wm + state.u * (1.0 - wm)
    v = v * wm + state.v * (1.0 - wm)
    T = T * wm + state.T * (1.0 - wm)
    S = S * wm + state.S * (1.0 - wm)
    if rec is not None:
        diffusion_state = State(u, v, T, S, state.eta, state.ice)
        rec.stage("linear_diffusion", state, diffusion_state)
    decay = np.exp(-p.sponge_rate * dt_half)
    u = u * decay
    v = v * decay
    T = p.T_clim_3d + (T - p.T_clim_3d) * decay
    S = p.S_clim_3d + (S - p.S_clim_3d) * decay
    if rec is not None:
        rec.stage("sponge", diffusion_state, State(u, v, T, S, state.eta, state.ice))
        rec.sponge_sources(diffusion_state, decay)
    u, v = coriolis_rotation(u, v, p.f, dt_half)
    return State(u, v, T, S, state.eta, state.ice)


def explicit_full_step(state, p, dt, rec=None):
    """Legacy tracer-first RK2 with implicit bottom-drag strip (step.py:249-320)."""
    u, v, T, S, eta = state.u, state.v, state.T, state.S, state.eta
    dT1, dS1 = tracer_residual(u, v, T, S, p, rec=rec)
    T_pred = T + dT1 * dt
    S_pred = S + dS1 * dt
    du1, dv1 = momentum_residual(u, v, T_pred, S_pred, eta, p)
    bm = p.bottom_mask * p.wet_mask_z
    du1 = du1 + p.r_bot * u * bm
    dv1 = dv1 + p.r_bot * v * bm
    u_pred = u + du1 * dt
    v_pred = v + dv1 * dt
    if getattr(p, "project_adv_vel", False):
        u_adv, v_adv = project_column_divergence(u_pred, v_pred, p, dt)
        if rec is not None:
            rec.column_projection(u_pred, v_pred, u_adv, v_adv)
    else:
        u_adv, v_adv = u_pred, v_pred
    dT2, dS2 = tracer_residual(u_adv, v_adv, T_pred, S_pred, p, rec=rec)
    T_new = T + 0.5 * (dT1 + dT2) * dt
    S_new = S + 0.5 * (dS1 + dS2) * dt
    du2, dv2 = momentum_residual(u, v, T_new, S_new, eta, p)
    du2 = du2 + p.r_bot * u * bm
    dv2 = dv2 + p.r_bot * v * bm
    u_new = u + 0.5 * (du1 + du2) * dt
    v_new = v + 0.5 * (dv1 + dv2) * dt
    decay = np.exp(-p.r_bot * dt * bm)
    return State(u_new * decay, v_new * decay, T_new, S_new, eta, state.ice)


# --------------------------------------------------------------------------
# column-divergence projection (reference dynamics/projection.py:64-127)
# --------------------------------------------------------------------------
def column_divergence(u, v, p):
    """sum_k div_h dz_node, masked (transport.py:45-57, legacy branch)."""
    return np.sum(divergence_h(u, v, p) * p.dz_node, axis=-1) * p.wet_mask


def projection_matvec(psi, p):
    """-column_divergence(grad_c3d(psi)) * area (projection.py:85-87)."""
    gx, gy = gradient_conservative_3d(psi[:, :, None], p)
    return -column_divergence(gx, gy, p) * (p.dx_2d * p.dy * p.wet_mask)


def projection_diagonal(p):
    """Exact diagonal of the projection operator, legacy branch (projection.py:10-50)."""
    wet = p.wet_mask_z
    pos_x = wet * east(wet)
    neg_x = wet * west(wet)
    diag_x = 0.25 * p.inv_dx[..., :1] ** 2 * ((pos_x - neg_x) ** 2 + pos_x ** 2 + neg_x ** 2)
    if p.nx <= 2:
        diag_x = np.zeros_like(wet)
    c = _row(p.cos_lat, 3)
    cn, cs = np.roll(c, -1, axis=1), np.roll(c, 1, axis=1)
    pos_y = wet * north_wrap(wet)
    pos_y[:, -1, :] = 0.0
    neg_y = wet * south_wrap(wet)
    neg_y[:, 0, :] = 0.0
    pos_y = pos_y * 0.5 * (c + cn)
    neg_y = neg_y * 0.5 * (c + cs)
    diag_y = 0.25 * p.inv_dy ** 2 * ((pos_y - neg_y) ** 2 / c ** 2 + pos_y ** 2 / (c * cn) + neg_y ** 2 / (c * cs))
    area = p.dx_2d * p.dy * p.wet_mask
    return area * np.sum((diag_x + diag_y) * p.dz_node, axis=-1)


def pcg(matvec, b, tol, atol, maxiter, precond=None):
    """Preconditioned conjugate gradients from x0 = 0 with the stopping rule of
    jax.scipy.sparse.linalg.cg: iterate while |r|^2 > max(tol^2 |b|^2, atol^2)
    and k < maxiter (the |r|^2 test uses r, not the preconditioned residual)."""
    M = precond if precond is not None else (lambda r: r)
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p_ = z.copy()
    gamma = float(np.vdot(r, z))
    atol2 = max(tol * tol * float(np.vdot(b, b)), atol * atol)
    k = 0
    while float(np.vdot(r, r)) > atol2 and k < maxiter:
        Ap = matvec(p_)
        alpha = gamma / float(np.vdot(p_, Ap))
        x = x + alpha * p_
        r = r - alpha * Ap
        z = M(r)
        gamma_new = float(np.vdot(r, z))
        p_ = z + (gamma_new / gamma) * p_
        gamma = gamma_new
        k += 1
    return x, k


def project_column_divergence(u, v, p, dt, return_iterations=False):
    """(u, v) minus dt g grad(psi) removing the column divergence, with the
    bounded transport-checked refinements (projection.py:64-127)."""
    g = G_EARTH
    area = p.dx_2d * p.dy * p.wet_mask
    rhs = -(column_divergence(u, v, p) / (dt * g)) * area
    tol = p.projection_rtol
    precond = None
    if p.projection_preconditioner == "jacobi":
        d = projection_diagonal(p)
        inv = 1.0 / np.where(d > 0.0, d, 1.0)
        precond = lambda r: r * inv  # noqa: E731
    matvec = lambda x: projection_matvec(x, p)  # noqa: E731
    psi, its = pcg(matvec, rhs, tol, 0.0, p.projection_niter, precond)
    iterations = [its]

    def correct(vel, pot):
        gx, gy = gradient_conservative_3d(pot[:, :, None], p)
        return vel[0] - (dt * g) * gx, vel[1] - (dt * g) * gy

    corrected = correct((u, v), psi)

    def transport_norm2(vel):
        top = vertical_transport_iface(vel[0], vel[1], p)[..., 0]
        return float(np.sum(top ** 2 * area))

    t_tol = max(1e-10, 32.0 * p.projection_eps)
    threshold = t_tol ** 2 * transport_norm2((u, v))
    floor = tol * float(np.linalg.norm(rhs))
    for _ in range(p.projection_max_refinements):
        if transport_norm2(corrected) > threshold:
            res = -(column_divergence(*corrected, p) / (dt * g)) * area
            delta, its = pcg(matvec, res, tol, floor, p.projection_niter, precond)
            iterations.append(its)
            corrected = correct(corrected, delta)
    if return_iterations:
        return corrected, iterations
    return corrected


def free_surface_substep(eta, ubt, vbt, Fx_rho, Fy_rho, p):
    """One forward-backward barotropic substep, mode-split legacy branch (barotropic.py:86-254)."""
    dt_bt = p.dt_bt
    Fx = np.zeros_like(ubt) + Fx_rho + p.tau_x_2d / (RHO_0 * p.H_sw)
    Fy = np.zeros_like(vbt) + Fy_rho + p.tau_y_2d / (RHO_0 * p.H_sw)
    drag = 1.0 / (1.0 + p.r_bot * dt_bt)
    wm = p.wet_mask
    div_bt = divergence_conservative(ubt * wm, vbt * wm, p)
    eta_new = (eta - dt_bt * p.H_sw * div_bt) * wm
    sw_decay = np.exp(-p.sponge_rate_2d * dt_bt)
    eta_new = filter_barotropic_eta(eta_new, p, dt_bt)
    gx, gy = gradient_conservative(eta_new, p)
    u_star = ubt + dt_bt * (-G_EARTH * gx + Fx)
    v_star = vbt + dt_bt * (-G_EARTH * gy + Fy)
    fd = p.f * dt_bt
    denom = 1.0 + fd * fd
    ubt_new = (u_star + fd * v_star) / denom
    vbt_new = (v_star - fd * u_star) / denom
    ubt_new = ubt_new * drag * wm * sw_decay
    vbt_new = vbt_new * drag * wm * sw_decay
    return eta_new, apply_polar_cap(ubt_new, wm, p), apply_polar_cap(vbt_new, wm, p)


def filter_barotropic_eta(eta, p, duration):
    """Sponge + volume refill + polar cap (barotropic.py:22-41); eta relaxation is off."""
    area = p.dx_2d * p.dy
    ocean_area = max(float(np.sum(area * p.wet_mask)), 1.0)
    decay = np.exp(-p.sponge_rate_2d * duration)
    damped = eta * decay
    dV = np.sum(area * (damped - eta))
    updated = damped + (-dV / ocean_area) * p.wet_mask
    return apply_polar_cap(updated, p.wet_mask, p)


def barotropic_subcycle(state, p):
    """Mode-split driver block (step.py:393-428): returns the new state and the 2-D terms."""
    Fx, Fy = bt_rho_pgf(state.T, state.S, p)
    ubt0, vbt0 = barotropic_velocity(state.u, state.v, p)
    eta, ubt, vbt = state.eta, ubt0, vbt0
    for _ in range(int(p.n_subcyc)):
        eta, ubt, vbt = free_surface_substep(eta, ubt, vbt, Fx, Fy, p)
    u = state.u + (ubt - ubt0)[:, :, None]
    v = state.v + (vbt - vbt0)[:, :, None]
    return State(u, v, state.T, state.S, eta, state.ice)


def final_masks(state, land, p):
    """Polar cap on u,v,T,S, hold land/ghost at pre-step values, close v walls (step.py:434-451)."""
    wm = p.wet_mask_z
    u = apply_polar_cap(state.u, wm, p)
    v = apply_polar_cap(state.v, wm, p)
    T = apply_polar_cap(state.T, wm, p)
    S = apply_polar_cap(state.S, wm, p)
    u = u * wm + land.u * (1.0 - wm)
    v = v * wm + land.v * (1.0 - wm)
    T = T * wm + land.T * (1.0 - wm)
    S = S * wm + land.S * (1.0 - wm)
    v = v * p.interior_mask_z
    return State(u, v, T, S, state.eta, state.ice)


def step(state, p, stages=None):
    """Strang L(dt/2) N(dt) L(dt/2) + barotropic subcycle + final masks (step.py:341-458).

    ``stages`` (a dict) receives the intermediate states when given.
    """
    state = State(*(np.asarray(a, dtype=np.float64) for a in state))
    land = state
    dt_half = p.dt / 2.0
    s1 = linear_half_step(state, p, dt_half)
    s2 = explicit_full_step(s1, p, p.dt)
    s3 = linear_half_step(s2, p, dt_half)
    s4 = barotropic_subcycle(s3, p)
    out = final_masks(s4, land, p)
    if stages is not None:
        stages.update(L1=s1, N=s2, L2=s3, BT=s4, out=out)
    return out


def with_forcing(p, tau_x, tau_y, q_heat, t_atm=None):
    """Dynamic forcing (factory.py:735-750): same params, runtime 2-D forcing."""
    q = SimpleNamespace(**vars(p))
    q.tau_x_2d = np.asarray(tau_x, dtype=np.float64)
    q.tau_y_2d = np.asarray(tau_y, dtype=np.float64)
    q.Q_heat_2d = np.asarray(q_heat, dtype=np.float64)
    if t_atm is not None:
        q.T_atm = np.asarray(t_atm, dtype=np.float64).reshape(p.nx, p.ny)
    return q


def diagnostics(state, p):
    """(rho, pressure, w) (factory.py:752-760)."""
    rho = RHO_0 + density_anomaly(state.T, state.S, p)
    return rho, hydrostatic_pressure(state.T, state.S, state.eta, p), vertical_velocity(state.u, state.v, p)


def budget(state, p):
    """Global heat/salt/volume inventories in float64 (diagnostics/snapshot.py:67-128)."""
    area = p.dx_2d * p.dy
    volume = area[:, :, None] * p.dz_node[None, None, :] * p.wet_mask_z
    return dict(heat=float(np.sum(RHO_0 * C_P * state.T * volume)),
                salt=float(np.sum(RHO_0 * (state.S / 1000.0) * volume)),
                volume=float(np.sum(area * state.eta * p.wet_mask)))


def sypd(dt, seconds_per_step):
    return dt / (365.0 * seconds_per_step)

