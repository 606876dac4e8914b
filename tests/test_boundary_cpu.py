"""CPU checks of the drop-in boundary: the C ABI library, the host-side
parameter freezing and the reference's construction-time ValueErrors."""

import ctypes
import re
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from oracle import zm_oracle as O
from tests.fixtures import land_case, oracle_params
from zhenmode_amd import _lib
from zhenmode_amd.cases import make_case
from zhenmode_amd.params import freeze, make_config

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "zhenmode.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|const char\*)\s+(zm_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_symbols()
    assert set(declared) == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.zm_version() == 1


def test_create_rejects_invalid_config_without_touching_the_gpu():
    lib = _lib.load()
    cfg = _lib.ZmConfig()
    cfg.nx, cfg.ny, cfg.nz, cfg.dtype_bytes = 1, 4, 4, 4
    handle = ctypes.c_void_p()
    rc = lib.zm_create(ctypes.byref(cfg), ctypes.byref(handle))
    assert rc == _lib.ZM_EINVAL
    assert b">= 2" in lib.zm_last_error(None)
    with pytest.raises(ValueError):
        _lib.check(rc, None)


def test_config_struct_matches_header_field_order():
    text = (ROOT / "include" / "zhenmode.h").read_text()
    body = text[text.index("typedef struct zm_config {"):text.index("} zm_config;")]
    names = re.findall(r"(\w+)(?=\s*[,;])", re.sub(r"/\*.*?\*/", "", body, flags=re.S))
    assert names == [f[0] for f in _lib.ZmConfig._fields_]


def test_host_freeze_matches_oracle_parameters():
    grid, physics, dt, kwargs, _, _ = land_case()
    p = freeze(grid, physics, dt, mode_split=True, **kwargs)
    q = oracle_params(grid, physics, dt, kwargs)
    for name in ("n_subcyc", "dt_bt", "conv_nsub", "adv_nsub", "H_sw", "dz_surface", "dz_bnd_top",
                 "dz_bnd_bot", "d2z_h0_top", "d2z_h0_bot", "restore_coef_S", "lambda_bulk", "dealias_keep"):
        assert getattr(p, name) == getattr(q, name), name
    for name in ("dz_node", "dz_iface", "dz_norm", "d2z_hm", "d2z_hp", "d2z_denom", "dz_denom_interior",
                 "coastal_restore_coef_2d", "coastal_restore_T_2d", "coastal_bulk_lambda_2d",
                 "coastal_kappa_h_2d", "coastal_kappa_v_2d", "S_ref_2d", "tau_x_2d", "Q_heat_2d"):
        np.testing.assert_array_equal(getattr(p, name), getattr(q, name), err_msg=name)
    np.testing.assert_array_equal(p.cap_weights, O.polar_cap_weights(2, 2))
    np.testing.assert_array_equal(p.nwet, q.wet_mask_z.sum(-1))
    cfg, keep = make_config(p)
    assert cfg.nx == grid.nx and cfg.n_nu == p.n_subcyc and cfg.dtype_bytes == 8


@pytest.mark.parametrize("kwargs,match", [
    (dict(mode_split=False), "mode_split"),
    (dict(mode_split=True, dtype="float16"), "dtype"),
    (dict(mode_split=True, column_geometry="nodal_dual_v1"), "column_geometry"),
    (dict(mode_split=True, freeze_adv_vel=True), "freeze_adv_vel"),
    (dict(mode_split=True, projection_niter=0), "projection_niter"),
    (dict(mode_split=True, projection_niter=True), "projection_niter"),
    (dict(mode_split=True, projection_rtol=-1.0), "projection_rtol"),
    (dict(mode_split=True, projection_preconditioner="ilu"), "none or jacobi"),
    (dict(mode_split=True, projection_max_refinements=3), "0 to 2"),
    (dict(mode_split=True, sponge_days=10.0, sponge_cells=3), "sponge"),
    (dict(mode_split=True, polar_cap_rows=30, polar_cap_taper=3), "overlap"),
    (dict(mode_split=True, dt_bt=-1.0), "dt_bt"),
    (dict(mode_split=True, nu_nsub=0), "nu_nsub"),
    (dict(mode_split=True, surface_pressure_pa=np.zeros((90, 40))), "surface_pressure"),
])
def test_invalid_or_unsupported_options_raise_value_error(kwargs, match):
    case = make_case("g4deg")
    with pytest.raises(ValueError, match=match):
        freeze(case.grid, case.physics, case.dt, **kwargs)


def test_negative_coefficients_raise_like_the_reference():
    case = make_case("g4deg")
    with pytest.raises(ValueError, match="kappa_h"):
        freeze(case.grid, replace(case.physics, kappa_h=-1.0), case.dt, mode_split=True)
    with pytest.raises(ValueError, match="dt"):
        freeze(case.grid, case.physics, 0.0, mode_split=True)


def test_case_generator_is_seeded_and_near_thirty_percent_land():
    a, b = make_case("g4deg"), make_case("g4deg")
    np.testing.assert_array_equal(a.T_init, b.T_init)
    np.testing.assert_array_equal(a.grid.depth, b.grid.depth)
    assert 0.25 < a.land_fraction < 0.40
    assert a.grid.z[0] == 0.0 and a.grid.z[-1] == -4000.0 and a.grid.z[1] == pytest.approx(-10.0)
    g1deg = make_case("g1deg", nx=90, ny=40)
    assert g1deg.grid.z[-1] == -5500.0 and g1deg.grid.nz == 50


def test_production_flags_are_frozen_into_the_config():
    from zhenmode_amd.params import make_config
    case = make_case("g4deg")
    p = freeze(case.grid, case.physics, case.dt, conservative_kv=True, localize_conv=True, fct_adv=True,
               monotone_adv=False, **{k: v for k, v in case.solver_kwargs.items() if k != "mode_split"},
               mode_split=True)
    cfg, keep = make_config(p)
    assert (cfg.conservative_kv, cfg.localize_conv, cfg.fct_adv, cfg.monotone_adv) == (1, 1, 1, 0)


def test_projection_options_resolve_like_the_reference(monkeypatch):
    """factory.py:368-386 and projection.py:55-62: iteration cap from the
    argument, else OCEAN_PAV_NITER, else 150; tolerances floored at 32 eps."""
    case = make_case("g4deg")
    p = freeze(case.grid, case.physics, case.dt, mode_split=True, dtype="float32", project_adv_vel=True,
               projection_preconditioner="jacobi")
    eps32 = float(np.finfo(np.float32).eps)
    assert p.project_adv_vel and p.projection_niter == 150
    assert p.projection_rtol == 32.0 * eps32 and p.projection_transport_rtol == 32.0 * eps32
    cfg, _ = make_config(p)
    assert (cfg.project_adv_vel, cfg.projection_niter, cfg.projection_preconditioner) == (1, 150, 1)
    assert cfg.projection_max_refinements == 2 and cfg.projection_rtol == p.projection_rtol
    q = freeze(case.grid, case.physics, case.dt, mode_split=True, projection_rtol=1e-9)
    assert q.projection_rtol == 1e-9 and q.projection_transport_rtol == 1e-10 and not q.project_adv_vel
    monkeypatch.setenv("OCEAN_PAV_NITER", "40")
    assert freeze(case.grid, case.physics, case.dt, mode_split=True).projection_niter == 40
    assert freeze(case.grid, case.physics, case.dt, mode_split=True, projection_niter=7).projection_niter == 7
    monkeypatch.setenv("OCEAN_PAV_NITER", "many")
    with pytest.raises(ValueError, match="environment"):
        freeze(case.grid, case.physics, case.dt, mode_split=True)
