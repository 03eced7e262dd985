"""ctypes bindings to librdb_colphys.so (include/rdb_colphys.h).

Mirrors the reference's ``python/rdb/_ffi.py`` (library discovery order,
status codes, one prototype per exported symbol so a mismatched call raises
``ctypes.ArgumentError`` instead of corrupting memory), for the column-physics
library that replaces the vertical-mixing call site.

Library discovery order (first hit wins), after ``python/rdb/_ffi.py:8-22``:

  1. ``RDB_COLPHYS_LIB`` environment variable -- explicit path to the .so;
  2. the in-tree build, ``roundabout_amd/lib/librdb_colphys.so``;
  3. the dynamic-linker search path (``ctypes.CDLL("librdb_colphys.so")``).
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

# Status codes: same values as include/rdb_ocean.h:82-97.
OK = 0
ERR_CONFIG_PARSE = 1
ERR_CONFIG_VALIDATE = 2
ERR_SETUP = 3
ERR_IO = 5
ERR_BAD_HANDLE = 10
ERR_ALREADY_EXISTS = 11
ERR_NOT_INITIALISED = 12
ERR_NOT_FOUND = 13
ERR_BAD_SHAPE = 14
GUARD_TRIPPED = 1

PREC_FP64 = 0
PREC_SPLIT = 1
CLOSURE_PLAIN = 0
CLOSURE_SURROGATE = 1

_LIB_NAME = "librdb_colphys.so"
PACKAGE_DIR = Path(__file__).resolve().parent.parent
LIB_DIR = PACKAGE_DIR / "lib"


class LibraryNotFoundError(FileNotFoundError):
    """No usable librdb_colphys.so could be located (the HIP build is missing)."""


class RdbError(RuntimeError):
    """A non-zero status from the C ABI; ``status`` holds the code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"status {status}: {message}")
        self.status = status


class Config(ctypes.Structure):
    """ctypes mirror of ``rdb_colphys_config`` (field order is the ABI)."""

    _fields_ = [
        ("pp81_nu0", ctypes.c_double),
        ("pp81_nu_bg", ctypes.c_double),
        ("pp81_kappa_bg", ctypes.c_double),
        ("pp81_alpha", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("shear2_floor", ctypes.c_double),
        ("use_kpp", ctypes.c_int),
        ("pad0", ctypes.c_int),
        ("ri_crit", ctypes.c_double),
        ("c_vt2", ctypes.c_double),
        ("cs_nonlocal", ctypes.c_double),
        ("alpha_T", ctypes.c_double),
        ("beta_S", ctypes.c_double),
        ("p_ref", ctypes.c_double),
        ("cp", ctypes.c_double),
        ("has_heat", ctypes.c_int),
        ("has_salt", ctypes.c_int),
        ("kv_bg", ctypes.c_double),
        ("kt_bg", ctypes.c_double),
        ("ks_bg", ctypes.c_double),
        ("kv_max", ctypes.c_double),
        ("kd_max", ctypes.c_double),
        ("vmix_guard", ctypes.c_int),
        ("conv_enable", ctypes.c_int),
        ("kv_ml_invz2", ctypes.c_double),
        ("ddiff_enable", ctypes.c_int),
        ("kd_smooth_iterations", ctypes.c_int),
        ("use_harmonic", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("hmix_fixed", ctypes.c_double),
        ("conv_kd", ctypes.c_double),
        ("conv_prandtl", ctypes.c_double),
        ("conv_n2_thresh", ctypes.c_double),
        ("ddiff_strat_param_max", ctypes.c_double),
        ("ddiff_kappa_s", ctypes.c_double),
        ("ddiff_exp1", ctypes.c_double),
        ("ddiff_exp2", ctypes.c_double),
        ("ddiff_param1", ctypes.c_double),
        ("ddiff_param2", ctypes.c_double),
        ("ddiff_param3", ctypes.c_double),
        ("ddiff_mol_diff", ctypes.c_double),
        ("ddiff_use_k90", ctypes.c_int),
        ("epbl_enable", ctypes.c_int),
        ("epbl_mstar_scheme", ctypes.c_int),
        ("epbl_vstar_scheme", ctypes.c_int),
        ("epbl_combine", ctypes.c_int),
        ("epbl_mld_iteration", ctypes.c_int),
        ("epbl_mld_max_its", ctypes.c_int),
        ("epbl_mld_bisection", ctypes.c_int),
        ("epbl_mld_use_prev_guess", ctypes.c_int),
        ("epbl_use_lt", ctypes.c_int),
        ("epbl_lt_scheme", ctypes.c_int),
        ("epbl_tke_diags", ctypes.c_int),
    ] + [(n, ctypes.c_double) for n in (
        "epbl_mstar", "epbl_mstar_cap", "epbl_mstar_coef1", "epbl_c_ek", "epbl_mstar_conv_adj",
        "epbl_rh18_cn1", "epbl_rh18_cn2", "epbl_rh18_cn3", "epbl_rh18_cs1", "epbl_rh18_cs2",
        "epbl_nstar", "epbl_tke_decay", "epbl_wstar_ustar_coef", "epbl_vstar_scale_fac", "epbl_vstar_surf_fac",
        "epbl_von_karman", "epbl_ekman_scale_coef", "epbl_min_mix_len", "epbl_mixlen_exponent",
        "epbl_translay_scale", "epbl_mld_tol", "epbl_omega", "epbl_omega_frac", "epbl_ustar_min", "epbl_prandtl",
        "epbl_lt_enhance_coef", "epbl_lt_enhance_exp", "epbl_lt_max_enhance", "epbl_la_frac_hbl",
        "epbl_lt_lac1", "epbl_lt_lac2", "epbl_lt_lac3", "epbl_lt_lac4", "epbl_lt_lac5")]


class MlpConfig(ctypes.Structure):
    """ctypes mirror of ``rdb_colphys_mlp_config``."""

    _fields_ = [
        ("hidden", ctypes.c_int),
        ("n_hidden", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("epochs", ctypes.c_int),
        ("lr", ctypes.c_double),
        ("beta1", ctypes.c_double),
        ("beta2", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("seed", ctypes.c_int),
        ("guard_mode", ctypes.c_int),
        ("val_fraction", ctypes.c_double),
    ]


class MomentumConfig(ctypes.Structure):
    """ctypes mirror of ``rdb_colphys_momentum_config``."""

    _fields_ = [
        ("use_kv_source", ctypes.c_int),
        ("pad0", ctypes.c_int),
        ("K_v_momentum", ctypes.c_double),
        ("implicit_stress", ctypes.c_int),
        ("implicit_drag", ctypes.c_int),
        ("implicit_top_drag", ctypes.c_int),
        ("hvel_mom6", ctypes.c_int),
        ("hvel_harmonic", ctypes.c_int),
        ("hvel_upwind", ctypes.c_int),
        ("bbl_glue", ctypes.c_int),
        ("bbl_per_face", ctypes.c_int),
        ("hbbl_visc", ctypes.c_double),
        ("bbl_piston", ctypes.c_double),
        ("kv_bbl_bg", ctypes.c_double),
        ("zlevel_faces", ctypes.c_int),
        ("use_corner", ctypes.c_int),
        ("kv_corner_prandtl", ctypes.c_double),
        ("remnant", ctypes.c_int),
        ("remnant_only", ctypes.c_int),
    ]


GUARD_STRICT = 0
GUARD_SURFACE = 1
GUARD_RANGE = 2


class Counters(ctypes.Structure):
    """ctypes mirror of ``rdb_colphys_counters``."""

    _fields_ = [
        ("steps", ctypes.c_longlong),
        ("wet_columns", ctypes.c_longlong),
        ("land_columns", ctypes.c_longlong),
        ("last_kernel_ms", ctypes.c_double),
        ("sum_kernel_ms", ctypes.c_double),
        ("timed_launches", ctypes.c_longlong),
        ("algorithmic_bytes_per_step", ctypes.c_double),
        ("reference_bytes_per_step", ctypes.c_double),
        ("live_levels", ctypes.c_longlong),
        ("fallback_fraction", ctypes.c_double),
        ("block_threads", ctypes.c_int),
        ("order_mode", ctypes.c_int),
        ("lds_levels", ctypes.c_int),
        ("graph_replay", ctypes.c_int),
        ("mlp_forward_ms", ctypes.c_double),
        ("mlp_flop_per_step", ctypes.c_double),
    ]


# EPBL scheme tags (rdb_ocean_epbl.F90:68-104)
EPBL_MSTAR_CONSTANT, EPBL_MSTAR_OM4, EPBL_MSTAR_RH18 = 1, 2, 3
EPBL_VSTAR_CUBE_ROOT, EPBL_VSTAR_RH18 = 1, 2
EPBL_COMBINE_ADD, EPBL_COMBINE_MAX, EPBL_COMBINE_MAX_SHEAR = 1, 2, 3
EPBL_LT_RESCALE, EPBL_LT_ADDITIVE = 1, 2


def default_config() -> Config:
    """Reference defaults (src/core/rdb_config.F90:1743-1893), filled in Python
    so that hosts without the HIP library can still build a config."""
    c = Config()
    c.pp81_nu0 = 1.0e-2
    c.pp81_nu_bg = 1.0e-4
    c.pp81_kappa_bg = 1.0e-5
    c.pp81_alpha = 5.0
    c.rho0 = 1035.0
    c.shear2_floor = 1.0e-10
    c.use_kpp = 1
    c.ri_crit = 0.3
    c.c_vt2 = 1.8
    c.cs_nonlocal = 6.3
    c.alpha_T = 1.7e-4
    c.beta_S = 7.6e-4
    c.p_ref = 0.0
    c.cp = 3992.0
    c.has_heat = 1
    c.has_salt = 1
    c.kv_bg = 1.0e-4
    c.kt_bg = 1.0e-5
    c.ks_bg = 1.0e-5
    c.kv_max = 1.7976931348623157e308
    c.kd_max = 1.7976931348623157e308
    c.vmix_guard = 0
    c.conv_enable = 0
    c.kv_ml_invz2 = 0.0
    c.ddiff_enable = 0
    c.kd_smooth_iterations = 0
    c.use_harmonic = 0
    c.precision = PREC_FP64
    c.hmix_fixed = 20.0
    c.conv_kd = 1.0
    c.conv_prandtl = 1.0
    c.conv_n2_thresh = 0.0
    c.ddiff_strat_param_max = 2.55
    c.ddiff_kappa_s = 1.0e-4
    c.ddiff_exp1 = 1.0
    c.ddiff_exp2 = 3.0
    c.ddiff_param1 = 0.909
    c.ddiff_param2 = 4.6
    c.ddiff_param3 = -0.54
    c.ddiff_mol_diff = 1.5e-6
    c.ddiff_use_k90 = 0
    c.epbl_enable = 0
    c.epbl_mstar_scheme = EPBL_MSTAR_OM4
    c.epbl_vstar_scheme = EPBL_VSTAR_CUBE_ROOT
    c.epbl_combine = EPBL_COMBINE_ADD
    c.epbl_mld_iteration = 1
    c.epbl_mld_max_its = 20
    c.epbl_mld_bisection = 0
    c.epbl_mld_use_prev_guess = 0
    c.epbl_use_lt = 0
    c.epbl_lt_scheme = EPBL_LT_RESCALE
    c.epbl_tke_diags = 0
    c.epbl_mstar = 1.2
    c.epbl_mstar_cap = -1.0
    c.epbl_mstar_coef1 = 0.3
    c.epbl_c_ek = 0.085
    c.epbl_mstar_conv_adj = 0.0
    c.epbl_rh18_cn1 = 0.275
    c.epbl_rh18_cn2 = 8.0
    c.epbl_rh18_cn3 = -5.0
    c.epbl_rh18_cs1 = 0.2
    c.epbl_rh18_cs2 = 0.4
    c.epbl_nstar = 0.2
    c.epbl_tke_decay = 2.5
    c.epbl_wstar_ustar_coef = 1.0
    c.epbl_vstar_scale_fac = 1.0
    c.epbl_vstar_surf_fac = 1.2
    c.epbl_von_karman = 0.41
    c.epbl_ekman_scale_coef = 1.0
    c.epbl_min_mix_len = 0.0
    c.epbl_mixlen_exponent = 2.0
    c.epbl_translay_scale = 0.1
    c.epbl_mld_tol = 1.0
    c.epbl_omega = 7.2921e-5
    c.epbl_omega_frac = 0.0
    c.epbl_ustar_min = 1.0e-8
    c.epbl_prandtl = 1.0
    c.epbl_lt_enhance_coef = 0.447
    c.epbl_lt_enhance_exp = -1.33
    c.epbl_lt_max_enhance = 5.0
    c.epbl_la_frac_hbl = 0.04
    c.epbl_lt_lac1 = -0.87
    c.epbl_lt_lac2 = 0.0
    c.epbl_lt_lac3 = 0.0
    c.epbl_lt_lac4 = 0.95
    c.epbl_lt_lac5 = 0.95
    return c


# (name, restype, argtypes) for every symbol declared in include/rdb_colphys.h
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_hp = ctypes.POINTER(ctypes.c_void_p)
PROTOTYPES = [
    ("rdb_colphys_config_default", ctypes.c_int, [ctypes.POINTER(Config)]),
    ("rdb_colphys_working_precision", ctypes.c_int, []),
    ("rdb_colphys_create", ctypes.c_int,
     [ctypes.POINTER(Config), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
      _dp, _ip, _ip, _hp]),
    ("rdb_colphys_destroy", ctypes.c_int, [_hp]),
    ("rdb_colphys_set_field", ctypes.c_int,
     [_vp, ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_get_field", ctypes.c_int,
     [_vp, ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_eos", ctypes.c_int, [_vp]),
    ("rdb_colphys_coefficients", ctypes.c_int, [_vp, ctypes.c_double]),
    ("rdb_colphys_surface_flux", ctypes.c_int, [_vp, ctypes.c_double]),
    ("rdb_colphys_diffuse_tracers", ctypes.c_int, [_vp, ctypes.c_double]),
    ("rdb_colphys_nonlocal", ctypes.c_int, [_vp, ctypes.c_double]),
    ("rdb_colphys_step", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int]),
    ("rdb_colphys_synchronize", ctypes.c_int, [_vp]),
    ("rdb_colphys_autotune", ctypes.c_int, [_vp, ctypes.c_double]),
    ("rdb_colphys_set_precision", ctypes.c_int, [_vp, ctypes.c_int]),
    ("rdb_colphys_set_closure", ctypes.c_int, [_vp, ctypes.c_int]),
    ("rdb_colphys_heat_salt_totals", ctypes.c_int, [_vp, _dp, _dp]),
    ("rdb_colphys_set_timing", ctypes.c_int, [_vp, ctypes.c_int]),
    ("rdb_colphys_stats", ctypes.c_int, [_vp, ctypes.POINTER(Counters)]),
    ("rdb_colphys_last_error", ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    ("rdb_colphys_mlp_config_default", ctypes.c_int, [ctypes.POINTER(MlpConfig)]),
    ("rdb_colphys_mlp_collect", ctypes.c_int,
     [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]),
    ("rdb_colphys_mlp_train", ctypes.c_int,
     [_vp, ctypes.POINTER(MlpConfig), _fp, _fp, ctypes.c_int, _ip]),
    ("rdb_colphys_mlp_save", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("rdb_colphys_mlp_load", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("rdb_colphys_set_guard", ctypes.c_int, [_vp, ctypes.c_int]),
    ("rdb_colphys_ensemble_create", ctypes.c_int,
     [ctypes.POINTER(Config), ctypes.c_int, _ip, _ip, _ip, ctypes.c_int, ctypes.POINTER(_dp),
      ctypes.POINTER(_ip), _hp]),
    ("rdb_colphys_member_count", ctypes.c_int, [_vp, _ip]),
    ("rdb_colphys_member_set_field", ctypes.c_int,
     [_vp, ctypes.c_int, ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_member_get_field", ctypes.c_int,
     [_vp, ctypes.c_int, ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("rdb_mlp_eval", ctypes.c_int,
     [_ip, ctypes.c_int, _fp, ctypes.c_int, _fp, _fp]),
    ("rdb_mlp_loss_grad", ctypes.c_int,
     [_ip, ctypes.c_int, _fp, ctypes.c_int, _fp, _fp, _fp, _fp, _fp]),
    ("rdb_mlp_eval_t", ctypes.c_int,
     [_ip, ctypes.c_int, _fp, ctypes.c_int, _fp, _fp]),
    ("rdb_mlp_adam_steps", ctypes.c_int,
     [_ip, ctypes.c_int, _fp, ctypes.c_int, _fp, _fp, _fp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_double, _fp]),
    ("rdb_colphys_cbrt", ctypes.c_int, [_dp, _dp, ctypes.c_int]),
    ("rdb_colphys_safe_math", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
    ("rdb_colphys_momentum_config_default", ctypes.c_int, [ctypes.POINTER(MomentumConfig)]),
    ("rdb_colphys_momentum_config_from_string", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(MomentumConfig)]),
    ("rdb_colphys_config_from_string", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(Config)]),
    ("rdb_colphys_create_from_string", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _ip, _ip, _hp]),
    ("rdb_colphys_get_time", ctypes.c_int, [_vp, _dp]),
    ("rdb_colphys_get_step_count", ctypes.c_int, [_vp, _ip]),
    ("rdb_colphys_get_grid_info", ctypes.c_int, [_vp, _ip, _ip, _ip, _ip]),
    ("rdb_colphys_refresh_host", ctypes.c_int, [_vp]),
    ("rdb_colphys_get_tracer_count", ctypes.c_int, [_vp, _ip]),
    ("rdb_colphys_list_tracers", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    ("rdb_colphys_get_tracer_ptr", ctypes.c_int,
     [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), _ip, _ip, _ip, _ip]),
    ("rdb_colphys_set_h", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_set_u", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_set_v", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_set_wind", ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_set_heat_flux", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_set_salt_flux", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_set_tracer", ctypes.c_int,
     [_vp, ctypes.c_char_p, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("rdb_colphys_set_budgets", ctypes.c_int, [_vp, ctypes.c_int]),
    ("rdb_colphys_vdiff_momentum", ctypes.c_int, [_vp, ctypes.c_double, ctypes.POINTER(MomentumConfig)]),
    ("rdb_mlp_eval_noalloc", ctypes.c_int,
     [_ip, ctypes.c_int, _fp, ctypes.c_int, _fp, _fp]),
    # conservative vertical remap on device arrays (pointers as void *)
    ("rdb_remap_tracer_field", ctypes.c_int,
     [ctypes.c_int] * 4 + [_vp] * 4 + [ctypes.c_int] * 2 + [_vp] * 3),
    ("rdb_remap_face_velocity", ctypes.c_int,
     [ctypes.c_int] * 4 + [_vp] * 3 + [ctypes.c_int] * 5 + [_vp]),
    ("rdb_remap_columns", ctypes.c_int,
     [ctypes.c_int] * 3 + [_vp] * 4 + [ctypes.c_int] * 2 + [_vp]),
    ("rdb_remap_scan_preconditions", ctypes.c_int,
     [ctypes.c_int] * 3 + [_vp] * 2 + [ctypes.c_double, _ip, _dp, _dp, _vp]),
]
# pointer getters into host mirrors: (handle, void **ptr, extents..., gen)
for _name in ("h_layer", "u_face_x_layer", "v_face_y_layer", "rho_layer", "kv", "kt", "ks"):
    PROTOTYPES.append((f"rdb_colphys_get_{_name}_ptr", ctypes.c_int,
                       [_vp, ctypes.POINTER(ctypes.c_void_p), _ip, _ip, _ip, _ip]))
for _name in ("wet_t", "tau_x", "tau_y", "q_heat", "q_salt"):
    PROTOTYPES.append((f"rdb_colphys_get_{_name}_ptr", ctypes.c_int,
                       [_vp, ctypes.POINTER(ctypes.c_void_p), _ip, _ip, _ip]))


def _candidates():
    env = os.environ.get("RDB_COLPHYS_LIB")
    if env:
        yield Path(env)
    yield LIB_DIR / _LIB_NAME
    yield Path(_LIB_NAME)


_LIB = None


def library_path() -> Path | None:
    for cand in _candidates():
        if cand.is_absolute() and cand.exists():
            return cand
    return None


def load(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load the library once and attach every prototype.

    Raises ``LibraryNotFoundError`` when the HIP build is absent: there is no
    CPU fallback in the product path."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    errors = []
    cands = [Path(path)] if path is not None else list(_candidates())
    for cand in cands:
        try:
            lib = ctypes.CDLL(str(cand))
        except OSError as exc:
            errors.append(f"{cand}: {exc}")
            continue
        for name, res, args in PROTOTYPES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _LIB = lib
        return lib
    raise LibraryNotFoundError(
        "librdb_colphys.so not found or not loadable (build it with "
        "__graft_entry__.build()): " + "; ".join(errors))


def last_error(lib: ctypes.CDLL | None = None, idx: int = 0) -> str:
    lib = lib or load()
    cap = 1024
    buf = ctypes.create_string_buffer(cap)
    n = lib.rdb_colphys_last_error(idx, buf, cap)
    return buf.value.decode(errors="replace") if n > 0 else ""


def check(status: int, lib: ctypes.CDLL | None = None) -> int:
    """Raise ``RdbError`` with the ring's most recent message on a non-zero
    status (rdb_ocean.h:302-319 contract)."""
    if status != OK:
        raise RdbError(status, last_error(lib))
    return status
