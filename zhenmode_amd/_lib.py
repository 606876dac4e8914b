"""ctypes binding of ``libzhenmode_hip.so`` (declarations in include/zhenmode.h).

There is no CPU fallback: a missing or unloadable library raises immediately.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libzhenmode_hip.so"

ZM_OK, ZM_EINVAL, ZM_EHIP, ZM_ENOMEM = 0, 1, 2, 3
STAGE_LINEAR, STAGE_NONLINEAR, STAGE_BAROTROPIC, STAGE_TRACER_RES, STAGE_MOMENTUM_RES = range(5)
LEDGER_SIZE = 107  # ZM_LEDGER_SIZE: float64 entries written by zm_budget_step

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


class ZmConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("dtype_bytes", ctypes.c_int32),
        ("dt", ctypes.c_double), ("dt_bt", ctypes.c_double),
        ("n_subcyc", ctypes.c_int32), ("n_nu", ctypes.c_int32),
        ("conv_nsub", ctypes.c_int32), ("adv_nsub", ctypes.c_int32),
        ("polar_cap_rows", ctypes.c_int32), ("polar_cap_taper", ctypes.c_int32),
        ("nu_h", ctypes.c_double), ("nu_v", ctypes.c_double), ("kappa_h", ctypes.c_double),
        ("kappa_v", ctypes.c_double), ("kappa_conv", ctypes.c_double), ("nu_bi", ctypes.c_double),
        ("kappa_bi", ctypes.c_double), ("T_ref", ctypes.c_double), ("S_ref", ctypes.c_double),
        ("r_bot", ctypes.c_double), ("lambda_bulk", ctypes.c_double),
        ("restore_coef_S", ctypes.c_double), ("ice_freeze_temp_c", ctypes.c_double),
        ("ice_salt_flux", ctypes.c_double),
        ("dy", ctypes.c_double), ("H_sw", ctypes.c_double), ("dz_surface", ctypes.c_double),
        ("dz_bnd_top", ctypes.c_double), ("dz_bnd_bot", ctypes.c_double),
        ("d2z_h0_top", ctypes.c_double), ("d2z_h0_bot", ctypes.c_double),
        ("cos_lat", _D), ("dx_row", _D), ("f_row", _D),
        ("dz_node", _D), ("dz_iface", _D), ("dz", _D), ("dz_norm", _D),
        ("d2z_hm", _D), ("d2z_hp", _D), ("d2z_denom", _D), ("dz_denom_interior", _D),
        ("cap_weights", _D), ("dealias_mask", _I),
        ("nwet", _I), ("wet_mask", _D),
        ("tau_x", _D), ("tau_y", _D), ("q_heat", _D), ("t_atm", _D),
        ("s_ref_surf", _D), ("coastal_restore_coef", _D), ("coastal_restore_T", _D),
        ("coastal_bulk_lambda", _D), ("kappa_h_2d", _D), ("kappa_v_2d", _D),
        ("conservative_kv", ctypes.c_int32), ("localize_conv", ctypes.c_int32),
        ("fct_adv", ctypes.c_int32), ("monotone_adv", ctypes.c_int32),
        ("project_adv_vel", ctypes.c_int32), ("projection_niter", ctypes.c_int32),
        ("projection_preconditioner", ctypes.c_int32), ("projection_max_refinements", ctypes.c_int32),
        ("projection_rtol", ctypes.c_double), ("projection_transport_rtol", ctypes.c_double),
    ]


class ZmMonitorStatus(ctypes.Structure):
    _fields_ = [("max_u", ctypes.c_double), ("max_velocity", ctypes.c_double), ("max_eta", ctypes.c_double),
                ("peak_u", ctypes.c_double), ("peak_velocity", ctypes.c_double), ("peak_eta", ctypes.c_double),
                ("finite", ctypes.c_int32), ("failure", ctypes.c_int32),
                ("first_failure_step", ctypes.c_int64), ("last_step", ctypes.c_int64)]


class ZmState(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in ("u", "v", "T", "S", "eta", "ice")]


EXPORTS = ("zm_create", "zm_destroy", "zm_last_error", "zm_version", "zm_step", "zm_step_dyn",
           "zm_diagnostics", "zm_stage", "zm_set_graph", "zm_time_kernel", "zm_nu_substeps_per_launch",
           "zm_vertical_mix", "zm_column_features", "zm_mlp_apply", "zm_gemm_bf16_nt", "zm_mse_grad",
           "zm_rowsum_bf16", "zm_transpose_bf16", "zm_adam", "zm_set_tiles", "zm_get_tiles", "zm_autotune",
           "zm_ensemble_create", "zm_ensemble_step", "zm_ensemble_set_graph", "zm_ensemble_size",
           "zm_ensemble_destroy", "zm_seasonal_forcing", "zm_monitor_reset", "zm_monitor", "zm_budget",
           "zm_project_velocity", "zm_projection_iterations", "zm_projection_operator", "zm_budget_step",
           "zm_vertical_mix_noalloc", "zm_set_barotropic_persistent", "zm_barotropic_timeouts",
           "zm_tracer_terms", "zm_mix_targets", "zm_mlp_mix")

# epilogues of zm_gemm_bf16_nt
EPI_F32, EPI_BIAS_F32, EPI_BIAS_GELU, EPI_GELU_GRAD = range(4)

_lib = None


def load():
    """Load the library once; raise if it is absent (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m zhenmode_amd.build`")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp = ctypes.c_void_p
    lib.zm_create.argtypes = [ctypes.POINTER(ZmConfig), ctypes.POINTER(vp)]
    lib.zm_create.restype = ctypes.c_int
    lib.zm_destroy.argtypes = [vp]
    lib.zm_destroy.restype = ctypes.c_int
    lib.zm_last_error.argtypes = [vp]
    lib.zm_last_error.restype = ctypes.c_char_p
    lib.zm_version.argtypes = []
    lib.zm_version.restype = ctypes.c_int32
    lib.zm_step.argtypes = [vp, ctypes.POINTER(ZmState), ctypes.POINTER(ZmState), ctypes.c_int32, vp]
    lib.zm_step.restype = ctypes.c_int
    lib.zm_step_dyn.argtypes = [vp, ctypes.POINTER(ZmState), ctypes.POINTER(ZmState), vp, vp, vp, vp, vp]
    lib.zm_step_dyn.restype = ctypes.c_int
    lib.zm_diagnostics.argtypes = [vp, ctypes.POINTER(ZmState), vp, vp, vp, vp]
    lib.zm_diagnostics.restype = ctypes.c_int
    lib.zm_tracer_terms.argtypes = [vp, ctypes.POINTER(ZmState), vp, vp]
    lib.zm_tracer_terms.restype = ctypes.c_int
    lib.zm_stage.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ZmState), ctypes.POINTER(ZmState), vp]
    lib.zm_stage.restype = ctypes.c_int
    lib.zm_set_graph.argtypes = [vp, ctypes.c_int32]
    lib.zm_set_graph.restype = ctypes.c_int
    lib.zm_time_kernel.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    lib.zm_time_kernel.restype = ctypes.c_int
    lib.zm_nu_substeps_per_launch.argtypes = [vp]
    lib.zm_nu_substeps_per_launch.restype = ctypes.c_int32
    i64, i32, f32, f64 = ctypes.c_int64, ctypes.c_int32, ctypes.c_float, ctypes.c_double
    lib.zm_vertical_mix.argtypes = [vp, vp, vp, vp, vp, f64, f64, f64, vp, vp]
    lib.zm_vertical_mix_noalloc.argtypes = [vp, vp, vp, vp, vp, f64, f64, f64, vp, vp]
    lib.zm_set_barotropic_persistent.argtypes = [vp, i32]
    lib.zm_barotropic_timeouts.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
    lib.zm_column_features.argtypes = [vp, ctypes.POINTER(ZmState), vp, vp, vp, vp, vp, vp, i64, vp]
    lib.zm_mlp_apply.argtypes = [vp, ctypes.POINTER(ZmState), vp, i32, vp, vp, f64, f64, f64, f64, f64, vp, vp,
                                 vp, vp, vp]
    lib.zm_mix_targets.argtypes = [vp, ctypes.POINTER(ZmState), vp, vp, f64, f64, f64, vp, i32, vp]
    lib.zm_mlp_mix.argtypes = [vp, ctypes.POINTER(ZmState), vp, vp, vp, vp, vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                               vp, vp, f64, f64, f64, f64, f64, vp, vp, vp, vp, vp]
    lib.zm_gemm_bf16_nt.argtypes = [i64, i64, i64, vp, vp, i32, vp, vp, vp, vp, vp, vp]
    lib.zm_mse_grad.argtypes = [i64, i32, i32, vp, vp, vp, vp, vp, vp]
    lib.zm_rowsum_bf16.argtypes = [i64, i64, vp, vp, vp]
    lib.zm_transpose_bf16.argtypes = [i64, i64, vp, vp, vp]
    lib.zm_adam.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, f32, f32, f32, f32, i32, vp]
    lib.zm_seasonal_forcing.argtypes = [vp, vp, i32, f64, f64, vp, vp]
    lib.zm_monitor_reset.argtypes = [vp, vp, vp]
    lib.zm_monitor.argtypes = [vp, ctypes.POINTER(ZmState), i64, f64, f64, i32, vp, vp]
    lib.zm_budget.argtypes = [vp, ctypes.POINTER(ZmState), vp, vp]
    lib.zm_project_velocity.argtypes = [vp, vp, vp, vp, vp, f64, vp, vp]
    lib.zm_projection_iterations.argtypes = [vp, vp, vp]
    lib.zm_projection_operator.argtypes = [ctypes.POINTER(ZmConfig), vp, vp]
    lib.zm_budget_step.argtypes = [vp, ctypes.POINTER(ZmState), ctypes.POINTER(ZmState), vp, vp, vp, vp, vp, vp]
    for name in ("zm_seasonal_forcing", "zm_monitor_reset", "zm_monitor", "zm_budget", "zm_project_velocity",
                 "zm_projection_iterations", "zm_projection_operator", "zm_budget_step"):
        getattr(lib, name).restype = ctypes.c_int
    for name in ("zm_vertical_mix", "zm_vertical_mix_noalloc", "zm_set_barotropic_persistent",
                 "zm_barotropic_timeouts", "zm_column_features", "zm_mlp_apply", "zm_mix_targets", "zm_mlp_mix",
                 "zm_gemm_bf16_nt", "zm_mse_grad",
                 "zm_rowsum_bf16", "zm_transpose_bf16", "zm_adam", "zm_set_tiles", "zm_get_tiles", "zm_autotune",
           "zm_ensemble_create", "zm_ensemble_step", "zm_ensemble_set_graph", "zm_ensemble_size",
           "zm_ensemble_destroy", "zm_seasonal_forcing", "zm_monitor_reset", "zm_monitor", "zm_budget"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc, handle=None):
    """Raise the reference's exception types for a nonzero status."""
    if rc == ZM_OK:
        return
    msg = load().zm_last_error(handle).decode(errors="replace")
    if rc == ZM_EINVAL:
        raise ValueError(msg)
    if rc == ZM_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)
