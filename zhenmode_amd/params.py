"""Host-side parameter freezing for the mode-split step (float64 numpy).

Restates the legacy branches of ``make_fd_params``
(``src/zhenmode/model/solver/geometry/fd_metrics.py:9-148``) and of the
derivations in ``make_solver_global`` (``src/zhenmode/model/solver/factory.py:397-695``),
validates the keyword arguments the way the reference does (``ValueError``),
and rejects options outside the legacy mode-split step with a ``ValueError``
that names them.  The result is handed to ``zm_create`` unchanged.
"""

from __future__ import annotations

import ctypes
import os
from types import SimpleNamespace

import numpy as np

from ._lib import ZmConfig
from .config import finite_number, integer_count
from .grid import validate_grid

# Options whose reference default is the only value this step implements.
LEGACY_ONLY = {
    "column_geometry": "legacy", "process_time_scheme": "legacy",
    "external_mode_scheme": "forward_backward", "meridional_boundary_scheme": "clamped_nodes",
    "pressure_continuity_scheme": "centered_second", "tracer_transport_scheme": "centered_second",
    "hydrostatic_surface_scheme": "reference_plane", "density_rotation_scheme": "split_sources",
    "vertical_tracer_scheme": "donor_cell", "hydrostatic_integral_scheme": "trapezoidal",
    "column_mean_scheme": "nodal_capacity", "tracer_flux_arithmetic": "legacy",
    "momentum_vertical_velocity_scheme": "dual_face_v1",
    "material_point_gradient_scheme": "minmod_v1",
    "momentum_tangent_filter_scheme": "binomial_v1",
}
OFF_ONLY = ("dynamic_ice", "freeze_adv_vel", "match_barotropic_transport")
# production tracer/mixing/projection options (reference factory.py:49-53, 649-653)
FLAGS = ("conservative_kv", "localize_conv", "fct_adv", "monotone_adv", "project_adv_vel")
PRECONDITIONERS = {"none": 0, "jacobi": 1}


def projection_options(dtype, projection_niter=None, projection_rtol=None, projection_preconditioner="none",
                       projection_max_refinements=2):
    """Validate and resolve the projection options like the reference factory
    (``factory.py:368-386``) and ``projection.py:55-62``: the iteration cap
    falls back to ``OCEAN_PAV_NITER`` and then 150, the CG tolerance and the
    transport tolerance get a floor of 32 eps of the compute dtype."""
    if projection_niter is None:
        env = os.environ.get("OCEAN_PAV_NITER")
        try:
            projection_niter = 150 if env is None else int(env)
        except ValueError as error:
            raise ValueError("projection_niter environment value must be an integer") from error
    if (isinstance(projection_niter, (bool, np.bool_)) or not isinstance(projection_niter, (int, np.integer))
            or projection_niter <= 0):
        raise ValueError("projection_niter must be a positive integer")
    if projection_rtol is not None and (not np.isfinite(projection_rtol) or projection_rtol <= 0.0):
        raise ValueError("projection_rtol must be finite and positive")
    if projection_preconditioner not in PRECONDITIONERS:
        raise ValueError("projection_preconditioner must be none or jacobi")
    if (isinstance(projection_max_refinements, (bool, np.bool_))
            or not isinstance(projection_max_refinements, (int, np.integer))
            or not 0 <= projection_max_refinements <= 2):
        raise ValueError("projection_max_refinements must be an integer from 0 to 2")
    eps = float(np.finfo(np.float32 if dtype == "float32" else np.float64).eps)
    return dict(projection_niter=int(projection_niter),
                projection_rtol=max(float(projection_rtol or 1e-12), 32.0 * eps),
                projection_preconditioner=projection_preconditioner,
                projection_max_refinements=int(projection_max_refinements),
                projection_transport_rtol=max(1e-10, 32.0 * eps))


def cap_weights(ncap, ntaper):
    """Pole-inward polar-cap blend weights (reference horizontal.py:362-385)."""
    if ncap <= 0:
        return np.zeros(0)
    if ntaper <= 0:
        return np.ones(ncap)
    ramp = np.cos(np.linspace(0.0, np.pi / 2.0, ntaper + 2)[1:-1]) ** 2
    return np.concatenate([np.ones(ncap), ramp])


def _masked(mask, coef, shape):
    if mask is not None and coef > 0.0:
        return np.asarray(mask, dtype=np.float64) * float(coef)
    return np.zeros(shape)


def freeze(grid, physics, dt, forcing=None, *, T_atm=None, lambda_bulk=0.0,
           mixed_layer_depth_m=None, mixed_layer_mask=None, mixed_layer_depth_2d=None,
           ice_freeze_temp_c=-1.8, ice_salt_flux=0.0, dynamic_ice=False, ice_insulation_scale_m=1.0,
           S_ref_surf=None, sss_restore_days=0.0,
           coastal_restore_mask=None, coastal_restore_days=0.0,
           coastal_bulk_mask=None, coastal_bulk_lambda=0.0,
           coastal_kappa_h_mask=None, coastal_kappa_h=0.0,
           coastal_kappa_v_mask=None, coastal_kappa_v=0.0, coastal_restore_T=None,
           sponge_days=0.0, sponge_cells=0, T_init=None, S_init=None,
           polar_cap_rows=2, polar_cap_taper=3, eta_relax_days=0.0, eta_relax_box=None,
           eta_relax_buffer=1.0, mode_split=False, dt_bt=150.0, nu_nsub=None, dtype="float64",
           use_scan=False, projection_niter=None, projection_rtol=None,
           projection_preconditioner="none", projection_max_refinements=2,
           eos_pressure_dbar=None, surface_pressure_pa=None, **schemes):
    """Validate the arguments and return the frozen parameter namespace."""
    finite_number("dt", dt, positive=True)
    finite_number("dt_bt", dt_bt, positive=True)
    unknown = set(schemes) - set(LEGACY_ONLY) - set(OFF_ONLY) - set(FLAGS)
    if unknown:
        raise TypeError(f"unexpected keyword arguments: {sorted(unknown)}")
    for name, default in LEGACY_ONLY.items():
        if schemes.get(name, default) != default:
            raise ValueError(f"{name}={schemes[name]!r} is not supported by the HIP mode-split step "
                             f"(only {default!r})")
    for name in OFF_ONLY:
        if schemes.get(name, False):
            raise ValueError(f"{name}=True is not supported by the HIP mode-split step")
    if dynamic_ice:
        raise ValueError("dynamic_ice=True is not supported by the HIP mode-split step")
    if not mode_split:
        raise ValueError("the HIP step implements mode_split=True only (the unsplit path is out of scope)")
    if dtype not in ("float32", "float64"):
        raise ValueError("dtype must be float32 or float64")
    projection = projection_options(dtype, projection_niter, projection_rtol, projection_preconditioner,
                                    projection_max_refinements)
    validate_grid(grid)
    if physics.thermodynamics != "linear":
        raise ValueError("only the linear equation of state is supported by the HIP step")
    if eos_pressure_dbar is not None:
        raise ValueError("linear EOS does not accept an unused TEOS pressure field")
    if physics.bottom_friction != "linear":
        raise ValueError("only linear bottom friction is supported by the HIP step")
    for name in ("nu_h", "nu_v", "nu_bi", "kappa_h", "kappa_v", "kappa_bi", "kappa_conv",
                 "kappa_gm", "kappa_redi", "r_bot", "cd"):
        finite_number(name, getattr(physics, name), nonnegative=True)
    if physics.kappa_gm > 0.0 or physics.kappa_redi > 0.0:
        raise ValueError("GM/Redi (kappa_gm, kappa_redi > 0) is not supported by the HIP step")
    for name, value in (("lambda_bulk", lambda_bulk), ("sss_restore_days", sss_restore_days),
                        ("coastal_restore_days", coastal_restore_days),
                        ("coastal_bulk_lambda", coastal_bulk_lambda),
                        ("coastal_kappa_h", coastal_kappa_h), ("coastal_kappa_v", coastal_kappa_v),
                        ("sponge_days", sponge_days), ("eta_relax_days", eta_relax_days)):
        finite_number(name, value, nonnegative=True)
    for name, value in (("polar_cap_rows", polar_cap_rows), ("polar_cap_taper", polar_cap_taper),
                        ("sponge_cells", sponge_cells)):
        integer_count(name, value)
    if nu_nsub is not None and not (isinstance(nu_nsub, str) and nu_nsub == "cfl"):
        integer_count("nu_nsub", nu_nsub, minimum=1)
    if polar_cap_rows > 0 and 2 * (polar_cap_rows + polar_cap_taper) > grid.ny:
        raise ValueError("polar cap bands must not overlap; reduce rows/taper or disable the cap")
    if sponge_days > 0.0 and sponge_cells > 0:
        raise ValueError("the lateral sponge is not supported by the HIP step")
    if eta_relax_days > 0.0 and eta_relax_box is not None:
        raise ValueError("semi-enclosed-sea eta relaxation is not supported by the HIP step")
    if surface_pressure_pa is not None:
        raise ValueError("surface_pressure_pa is not supported by the HIP step")
    if mixed_layer_depth_m or mixed_layer_depth_2d is not None:
        raise ValueError("mixed-layer heat deposition is not supported by the HIP step")

    nx, ny, nz = int(grid.nx), int(grid.ny), int(grid.nz)
    p = SimpleNamespace(nx=nx, ny=ny, nz=nz, dt=float(dt), dtype=dtype)
    dx = np.asarray(grid.dx_2d, dtype=np.float64)
    f = np.asarray(grid.f, dtype=np.float64)
    if not np.array_equal(dx, np.broadcast_to(dx[:1, :], dx.shape)):
        raise ValueError("the HIP step requires dx_2d constant along each latitude row")
    if not np.array_equal(f, np.broadcast_to(f[:1, :], f.shape)):
        raise ValueError("the HIP step requires f constant along each latitude row")
    wet3 = np.asarray(grid.wet_mask_3d, dtype=np.float64)
    wet2 = np.asarray(grid.wet_mask, dtype=np.float64)
    if wet3.shape != (nx, ny, nz) or not np.all((wet3 == 0.0) | (wet3 == 1.0)):
        raise ValueError("wet_mask_3d must be a 0/1 array of shape (nx, ny, nz)")
    if np.any(np.diff(wet3, axis=-1) > 0.0):
        raise ValueError("the HIP step requires wet levels contiguous from the surface")
    if not np.array_equal(wet2, wet3[..., 0]):
        raise ValueError("wet_mask must equal the surface level of wet_mask_3d")
    p.dx_2d, p.dy, p.cos_lat, p.f = dx, float(grid.dy), np.asarray(grid.cos_lat, dtype=np.float64), f
    p.wet_mask, p.wet_mask_z = wet2, wet3
    p.nwet = wet3.sum(axis=-1).astype(np.int32)
    p.inv_dx = (1.0 / dx)[:, :, None]
    p.inv_dy = 1.0 / p.dy

    z = np.asarray(grid.z, dtype=np.float64)
    gaps = np.abs(np.diff(z))
    p.z = z
    p.dz_3d = np.asarray(grid.dz, dtype=np.float64)
    p.dz_denom_interior = (z[1:-1] - z[:-2]) + (z[2:] - z[1:-1])
    p.dz_bnd_top = float(z[1] - z[0])
    p.dz_bnd_bot = float(z[-1] - z[-2])
    p.d2z_hm = np.abs(z[:-2] - z[1:-1])
    p.d2z_hp = np.abs(z[2:] - z[1:-1])
    p.d2z_denom = p.d2z_hm * p.d2z_hp * (p.d2z_hm + p.d2z_hp) / 2.0
    p.d2z_h0_top = float(abs(z[1] - z[0]))
    p.d2z_h0_bot = float(abs(z[-1] - z[-2]))
    p.dz_surface = float(abs(z[0] - z[1]))
    p.dz_iface = gaps
    p.dz_node = np.concatenate([[gaps[0]], 0.5 * (gaps[:-1] + gaps[1:]), [gaps[-1]]])
    p.H_sw = float(np.sum(p.dz_3d))
    p.dz_norm = p.dz_3d / p.H_sw

    for name in ("nu_h", "nu_v", "kappa_h", "kappa_v", "kappa_conv", "nu_bi", "kappa_bi",
                 "T_ref", "S_ref", "r_bot"):
        setattr(p, name, float(getattr(physics, name)))
    shape2 = (nx, ny)
    if forcing is None:
        p.tau_x_2d, p.tau_y_2d, p.Q_heat_2d = (np.zeros(shape2) for _ in range(3))
    else:
        p.tau_x_2d, p.tau_y_2d, p.Q_heat_2d = (np.array(a, dtype=np.float64).reshape(shape2) for a in forcing)
    if T_atm is not None and lambda_bulk > 0.0:
        p.T_atm, p.lambda_bulk = np.array(T_atm, dtype=np.float64).reshape(shape2), float(lambda_bulk)
    else:
        p.T_atm, p.lambda_bulk = np.zeros(shape2), 0.0
    if S_ref_surf is not None and sss_restore_days > 0.0:
        p.S_ref_2d = np.array(S_ref_surf, dtype=np.float64)
        p.restore_coef_S = 1.0 / (sss_restore_days * 86400.0)
    else:
        p.S_ref_2d, p.restore_coef_S = np.zeros(shape2), 0.0
    p.coastal_kappa_v_2d = _masked(coastal_kappa_v_mask, coastal_kappa_v, shape2)
    p.coastal_kappa_h_2d = _masked(coastal_kappa_h_mask, coastal_kappa_h, shape2)
    p.coastal_bulk_lambda_2d = _masked(coastal_bulk_mask, coastal_bulk_lambda, shape2)
    if coastal_restore_mask is not None and coastal_restore_days > 0.0 and coastal_restore_T is not None:
        p.coastal_restore_coef_2d = (np.array(coastal_restore_mask, dtype=np.float64)
                                     / (coastal_restore_days * 86400.0))
        p.coastal_restore_T_2d = np.array(coastal_restore_T, dtype=np.float64)
    else:
        p.coastal_restore_coef_2d, p.coastal_restore_T_2d = np.zeros(shape2), np.zeros(shape2)
    p.ice_freeze_temp_c, p.ice_salt_flux = float(ice_freeze_temp_c), float(ice_salt_flux)
    p.polar_cap_rows, p.polar_cap_taper = int(polar_cap_rows), int(polar_cap_taper)
    p.cap_weights = cap_weights(p.polar_cap_rows, p.polar_cap_taper)
    p.dealias_keep = max(1, int((nx // 2) * 2 / 3))
    # Same float64 expression as the reference (factory.py:529-533): for some
    # nx (90, 180, 360, 720) fftfreq(nx) * nx rounds the +-keep wavenumber just
    # above keep, so that wavenumber is dropped; the mask is kept as computed.
    p.dealias_mask = (np.abs(np.fft.fftfreq(nx) * nx) <= p.dealias_keep).astype(np.int32)
    p.mode_split = True
    p.n_subcyc = max(1, int(round(dt / dt_bt)))
    p.dt_bt = dt / p.n_subcyc
    if nu_nsub == "cfl":
        inv_dx2_max = float(np.max(dx ** -2))
        nu_nsub = max(1, int(np.ceil(p.nu_h * (dt / 2.0) * (inv_dx2_max + 1.0 / p.dy ** 2) / 0.25)))
    elif nu_nsub is not None:
        nu_nsub = max(1, int(nu_nsub))
    p.nu_nsub = nu_nsub
    p.n_nu = max(1, int(p.n_subcyc if nu_nsub is None else nu_nsub))
    for name in FLAGS:
        setattr(p, name, bool(schemes.get(name, False)))
    for name, value in projection.items():
        setattr(p, name, value)
    min_dz = float(np.min(p.dz_3d))
    p.conv_nsub = max(1, int(np.ceil(p.kappa_conv * dt / (0.4 * min_dz ** 2))))
    p.adv_nsub = max(1, int(np.ceil(dt * 4.0e-3 / (0.5 * min_dz))))
    if p.adv_nsub > 1:
        raise ValueError(f"adv_nsub={p.adv_nsub} > 1 (vertical advection sub-cycling at this dt and "
                         "surface spacing) is not supported by the HIP step")
    p.use_scan = bool(use_scan)
    return p


def to_internal2(a):
    """(nx, ny) reference layout -> contiguous (ny, nx)."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)


def make_config(p):
    """Build the ctypes ``zm_config``; returns (config, keepalive list)."""
    keep = []

    def dptr(a):
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
        if arr.size == 0:
            arr = np.zeros(1)
        keep.append(arr)
        return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    nwet = np.ascontiguousarray(p.nwet.T.astype(np.int32))
    keep.append(nwet)
    c = ZmConfig()
    c.nx, c.ny, c.nz = p.nx, p.ny, p.nz
    c.dtype_bytes = 4 if p.dtype == "float32" else 8
    c.dt, c.dt_bt = p.dt, p.dt_bt
    c.n_subcyc, c.n_nu, c.conv_nsub, c.adv_nsub = p.n_subcyc, p.n_nu, p.conv_nsub, p.adv_nsub
    c.polar_cap_rows, c.polar_cap_taper = p.polar_cap_rows, p.polar_cap_taper
    for name in ("nu_h", "nu_v", "kappa_h", "kappa_v", "kappa_conv", "nu_bi", "kappa_bi", "T_ref",
                 "S_ref", "r_bot", "lambda_bulk", "restore_coef_S", "ice_freeze_temp_c",
                 "ice_salt_flux", "dy", "H_sw", "dz_surface", "dz_bnd_top", "dz_bnd_bot",
                 "d2z_h0_top", "d2z_h0_bot"):
        setattr(c, name, float(getattr(p, name)))
    c.cos_lat = dptr(p.cos_lat)
    c.dx_row = dptr(p.dx_2d[0, :])
    c.f_row = dptr(p.f[0, :])
    c.dz_node, c.dz_iface, c.dz, c.dz_norm = (dptr(a) for a in (p.dz_node, p.dz_iface, p.dz_3d, p.dz_norm))
    c.d2z_hm, c.d2z_hp = dptr(p.d2z_hm), dptr(p.d2z_hp)
    c.d2z_denom, c.dz_denom_interior = dptr(p.d2z_denom), dptr(p.dz_denom_interior)
    c.cap_weights = dptr(p.cap_weights)
    mask = np.ascontiguousarray(p.dealias_mask.astype(np.int32))
    keep.append(mask)
    c.dealias_mask = mask.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    c.nwet = nwet.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    c.wet_mask = dptr(to_internal2(p.wet_mask))
    c.tau_x, c.tau_y = dptr(to_internal2(p.tau_x_2d)), dptr(to_internal2(p.tau_y_2d))
    c.q_heat, c.t_atm = dptr(to_internal2(p.Q_heat_2d)), dptr(to_internal2(p.T_atm))
    c.s_ref_surf = dptr(to_internal2(p.S_ref_2d))
    c.coastal_restore_coef = dptr(to_internal2(p.coastal_restore_coef_2d))
    c.coastal_restore_T = dptr(to_internal2(p.coastal_restore_T_2d))
    c.coastal_bulk_lambda = dptr(to_internal2(p.coastal_bulk_lambda_2d))
    c.kappa_h_2d = dptr(to_internal2(p.kappa_h + p.coastal_kappa_h_2d))
    c.kappa_v_2d = dptr(to_internal2(p.kappa_v + p.coastal_kappa_v_2d))
    c.conservative_kv = int(p.conservative_kv)
    c.localize_conv = int(p.localize_conv)
    c.fct_adv = int(p.fct_adv)
    c.monotone_adv = int(p.monotone_adv)
    c.project_adv_vel = int(p.project_adv_vel)
    c.projection_niter = p.projection_niter
    c.projection_preconditioner = PRECONDITIONERS[p.projection_preconditioner]
    c.projection_max_refinements = p.projection_max_refinements
    c.projection_rtol = p.projection_rtol
    c.projection_transport_rtol = p.projection_transport_rtol
    return c, keep
