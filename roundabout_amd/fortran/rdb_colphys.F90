!> ISO_C_BINDING interface to librdb_colphys.so (include/rdb_colphys.h).
!>
!> Drop-in for the vertical-mixing call site of the Roundabout dyn-core:
!> where rdb_ocean_dyn.F90 calls
!>   vmix_build_coefficients(...)           (rdb_ocean_dyn.F90:1473)
!>   vdiff_apply_tracers(..., kt, ks)       (rdb_ocean_dyn.F90:1856)
!>   vmix_apply_nonlocal_tendencies(...)    (rdb_ocean_dyn.F90:1859)
!> a host code built against this module calls rdb_colphys_step (fused) or the
!> staged entry points with the same state arrays.  Arrays are the reference
!> layout: column-major (i fastest), k = 1 at the bed; passed by address.
module rdb_colphys
   use, intrinsic :: iso_c_binding, only: c_int, c_double, c_float, c_char, c_ptr, c_null_ptr, &
                                          c_long_long, c_null_char, c_loc, c_associated
   implicit none
   private

   integer(c_int), parameter, public :: RDB_COLPHYS_OK = 0
   integer(c_int), parameter, public :: RDB_COLPHYS_ERR_CONFIG_VALIDATE = 2
   integer(c_int), parameter, public :: RDB_COLPHYS_ERR_SETUP = 3
   integer(c_int), parameter, public :: RDB_COLPHYS_ERR_BAD_HANDLE = 10
   integer(c_int), parameter, public :: RDB_COLPHYS_ERR_ALREADY_EXISTS = 11
   integer(c_int), parameter, public :: RDB_COLPHYS_ERR_NOT_FOUND = 13
   integer(c_int), parameter, public :: RDB_COLPHYS_ERR_BAD_SHAPE = 14
   integer(c_int), parameter, public :: RDB_COLPHYS_PREC_FP64 = 0
   integer(c_int), parameter, public :: RDB_COLPHYS_PREC_SPLIT = 1
   integer(c_int), parameter, public :: RDB_COLPHYS_GUARD_TRIPPED = 1
   integer(c_int), parameter, public :: RDB_COLPHYS_CLOSURE_PLAIN = 0
   integer(c_int), parameter, public :: RDB_COLPHYS_CLOSURE_SURROGATE = 1
   integer(c_int), parameter, public :: RDB_COLPHYS_GUARD_STRICT = 0
   integer(c_int), parameter, public :: RDB_COLPHYS_GUARD_SURFACE = 1
   integer(c_int), parameter, public :: RDB_COLPHYS_GUARD_RANGE = 2
   !> Remap methods (rdb_constants.F90:208-218)
   integer(c_int), parameter, public :: RDB_REMAP_PCM = 1, RDB_REMAP_PLM = 2, RDB_REMAP_PPM = 3, &
                                        RDB_REMAP_PPM_H4 = 4, RDB_REMAP_PQM = 5

   !> Mirror of rdb_colphys_config (field order is the ABI).
   type, bind(c), public :: rdb_colphys_config_t
      real(c_double) :: pp81_nu0, pp81_nu_bg, pp81_kappa_bg, pp81_alpha
      real(c_double) :: rho0, shear2_floor
      integer(c_int) :: use_kpp, pad0
      real(c_double) :: ri_crit, c_vt2, cs_nonlocal
      real(c_double) :: alpha_T, beta_S, p_ref
      real(c_double) :: cp
      integer(c_int) :: has_heat, has_salt
      real(c_double) :: kv_bg, kt_bg, ks_bg, kv_max, kd_max
      integer(c_int) :: vmix_guard
      integer(c_int) :: conv_enable
      real(c_double) :: kv_ml_invz2
      integer(c_int) :: ddiff_enable, kd_smooth_iterations
      integer(c_int) :: use_harmonic, precision
      ! contributors off at the defaults (KV_ML_INVZ2 depth, convection,
      ! double diffusion, EPBL)
      real(c_double) :: hmix_fixed, conv_kd, conv_prandtl, conv_n2_thresh
      real(c_double) :: ddiff_strat_param_max, ddiff_kappa_s, ddiff_exp1, ddiff_exp2
      real(c_double) :: ddiff_param1, ddiff_param2, ddiff_param3, ddiff_mol_diff
      integer(c_int) :: ddiff_use_k90
      integer(c_int) :: epbl_enable, epbl_mstar_scheme, epbl_vstar_scheme, epbl_combine
      integer(c_int) :: epbl_mld_iteration, epbl_mld_max_its, epbl_mld_bisection, epbl_mld_use_prev_guess
      integer(c_int) :: epbl_use_lt, epbl_lt_scheme, epbl_tke_diags
      real(c_double) :: epbl_mstar, epbl_mstar_cap, epbl_mstar_coef1, epbl_c_ek, epbl_mstar_conv_adj
      real(c_double) :: epbl_rh18_cn1, epbl_rh18_cn2, epbl_rh18_cn3, epbl_rh18_cs1, epbl_rh18_cs2
      real(c_double) :: epbl_nstar, epbl_tke_decay, epbl_wstar_ustar_coef, epbl_vstar_scale_fac
      real(c_double) :: epbl_vstar_surf_fac, epbl_von_karman, epbl_ekman_scale_coef, epbl_min_mix_len
      real(c_double) :: epbl_mixlen_exponent, epbl_translay_scale, epbl_mld_tol, epbl_omega
      real(c_double) :: epbl_omega_frac, epbl_ustar_min, epbl_prandtl
      real(c_double) :: epbl_lt_enhance_coef, epbl_lt_enhance_exp, epbl_lt_max_enhance, epbl_la_frac_hbl
      real(c_double) :: epbl_lt_lac1, epbl_lt_lac2, epbl_lt_lac3, epbl_lt_lac4, epbl_lt_lac5
   end type rdb_colphys_config_t

   !> Mirror of rdb_colphys_mlp_config (learned closure training).
   type, bind(c), public :: rdb_colphys_mlp_config_t
      integer(c_int) :: hidden, n_hidden, batch, epochs
      real(c_double) :: lr, beta1, beta2, eps
      integer(c_int) :: seed, guard_mode
      real(c_double) :: val_fraction
   end type rdb_colphys_mlp_config_t

   !> Mirror of rdb_colphys_momentum_config (vdiff_apply_momentum knobs).
   type, bind(c), public :: rdb_colphys_momentum_config_t
      integer(c_int) :: use_kv_source, pad0
      real(c_double) :: K_v_momentum
      integer(c_int) :: implicit_stress, implicit_drag, implicit_top_drag
      integer(c_int) :: hvel_mom6, hvel_harmonic, hvel_upwind, bbl_glue, bbl_per_face
      real(c_double) :: hbbl_visc, bbl_piston, kv_bbl_bg
      integer(c_int) :: zlevel_faces, use_corner
      real(c_double) :: kv_corner_prandtl
      integer(c_int) :: remnant, remnant_only
   end type rdb_colphys_momentum_config_t

   public :: rdb_colphys_momentum_config_default, rdb_colphys_vdiff_momentum
   public :: rdb_colphys_create_from_string, rdb_colphys_set_budgets, rdb_colphys_refresh_host
   public :: rdb_colphys_get_step_count, rdb_colphys_get_time
   public :: rdb_colphys_ensemble_create, rdb_colphys_member_count
   public :: rdb_colphys_member_set_field, rdb_colphys_member_get_field
   public :: colphys_member_set, colphys_member_get
   public :: rdb_remap_tracer_field, rdb_remap_face_velocity, rdb_remap_columns, rdb_remap_scan_preconditions
   public :: rdb_colphys_mlp_config_default, rdb_colphys_mlp_collect, rdb_colphys_mlp_train
   public :: rdb_colphys_mlp_save, rdb_colphys_mlp_load, rdb_colphys_set_guard
   public :: rdb_colphys_config_default, rdb_colphys_working_precision
   public :: rdb_colphys_create, rdb_colphys_destroy
   public :: rdb_colphys_set_field, rdb_colphys_get_field
   public :: rdb_colphys_eos, rdb_colphys_coefficients, rdb_colphys_surface_flux
   public :: rdb_colphys_diffuse_tracers, rdb_colphys_nonlocal, rdb_colphys_step
   public :: rdb_colphys_synchronize, rdb_colphys_autotune
   public :: rdb_colphys_set_precision, rdb_colphys_set_closure
   public :: rdb_colphys_heat_salt_totals, rdb_colphys_last_error
   public :: colphys_set, colphys_get, colphys_last_error

   interface
      integer(c_int) function rdb_colphys_config_default(cfg) bind(c, name="rdb_colphys_config_default")
         import :: c_int, rdb_colphys_config_t
         type(rdb_colphys_config_t), intent(out) :: cfg
      end function
      integer(c_int) function rdb_colphys_working_precision() bind(c, name="rdb_colphys_working_precision")
         import :: c_int
      end function
      integer(c_int) function rdb_colphys_create(cfg, nx, ny, nz, n_passive, wet, k_top, sort_key, handle) &
         bind(c, name="rdb_colphys_create")
         import :: c_int, c_double, c_ptr, rdb_colphys_config_t
         type(rdb_colphys_config_t), intent(in) :: cfg
         integer(c_int), value :: nx, ny, nz, n_passive
         real(c_double), intent(in) :: wet(*)
         type(c_ptr), value :: k_top, sort_key
         type(c_ptr), intent(out) :: handle
      end function
      integer(c_int) function rdb_colphys_destroy(handle) bind(c, name="rdb_colphys_destroy")
         import :: c_int, c_ptr
         type(c_ptr), intent(inout) :: handle
      end function
      integer(c_int) function rdb_colphys_set_field(handle, name, data, n1, n2, n3) &
         bind(c, name="rdb_colphys_set_field")
         import :: c_int, c_double, c_char, c_ptr
         type(c_ptr), value :: handle
         character(kind=c_char), intent(in) :: name(*)
         real(c_double), intent(in) :: data(*)
         integer(c_int), value :: n1, n2, n3
      end function
      integer(c_int) function rdb_colphys_get_field(handle, name, out, n1, n2, n3) &
         bind(c, name="rdb_colphys_get_field")
         import :: c_int, c_double, c_char, c_ptr
         type(c_ptr), value :: handle
         character(kind=c_char), intent(in) :: name(*)
         real(c_double), intent(out) :: out(*)
         integer(c_int), value :: n1, n2, n3
      end function
      integer(c_int) function rdb_colphys_eos(handle) bind(c, name="rdb_colphys_eos")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function
      integer(c_int) function rdb_colphys_coefficients(handle, dt) bind(c, name="rdb_colphys_coefficients")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         real(c_double), value :: dt
      end function
      integer(c_int) function rdb_colphys_surface_flux(handle, dt) bind(c, name="rdb_colphys_surface_flux")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         real(c_double), value :: dt
      end function
      integer(c_int) function rdb_colphys_diffuse_tracers(handle, dt) bind(c, name="rdb_colphys_diffuse_tracers")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         real(c_double), value :: dt
      end function
      integer(c_int) function rdb_colphys_nonlocal(handle, dt) bind(c, name="rdb_colphys_nonlocal")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         real(c_double), value :: dt
      end function
      integer(c_int) function rdb_colphys_step(handle, dt, n_steps) bind(c, name="rdb_colphys_step")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         real(c_double), value :: dt
         integer(c_int), value :: n_steps
      end function
      integer(c_int) function rdb_colphys_synchronize(handle) bind(c, name="rdb_colphys_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function
      integer(c_int) function rdb_colphys_autotune(handle, dt) bind(c, name="rdb_colphys_autotune")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         real(c_double), value :: dt
      end function
      integer(c_int) function rdb_colphys_set_precision(handle, mode) bind(c, name="rdb_colphys_set_precision")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: mode
      end function
      integer(c_int) function rdb_colphys_set_closure(handle, mode) bind(c, name="rdb_colphys_set_closure")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: mode
      end function
      integer(c_int) function rdb_colphys_heat_salt_totals(handle, heat, salt) &
         bind(c, name="rdb_colphys_heat_salt_totals")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         real(c_double), intent(out) :: heat, salt
      end function
      !> wet(m) and sort_key(m): C pointers to member m's masks (c_loc of the member arrays)
      integer(c_int) function rdb_colphys_ensemble_create(cfg, n_members, nx, ny, nz, n_passive, wet, sort_key, &
                                                          handle) bind(c, name="rdb_colphys_ensemble_create")
         import :: c_int, c_ptr, rdb_colphys_config_t
         type(rdb_colphys_config_t), intent(in) :: cfg
         integer(c_int), value :: n_members, n_passive
         integer(c_int), intent(in) :: nx(*), ny(*), nz(*)
         type(c_ptr), intent(in) :: wet(*)
         type(c_ptr), value :: sort_key
         type(c_ptr), intent(out) :: handle
      end function
      integer(c_int) function rdb_colphys_member_count(handle, n) bind(c, name="rdb_colphys_member_count")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), intent(out) :: n
      end function
      integer(c_int) function rdb_colphys_member_set_field(handle, member, name, data, n1, n2, n3) &
         bind(c, name="rdb_colphys_member_set_field")
         import :: c_int, c_double, c_char, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: member
         character(kind=c_char), intent(in) :: name(*)
         real(c_double), intent(in) :: data(*)
         integer(c_int), value :: n1, n2, n3
      end function
      integer(c_int) function rdb_colphys_member_get_field(handle, member, name, out, n1, n2, n3) &
         bind(c, name="rdb_colphys_member_get_field")
         import :: c_int, c_double, c_char, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: member
         character(kind=c_char), intent(in) :: name(*)
         real(c_double), intent(out) :: out(*)
         integer(c_int), value :: n1, n2, n3
      end function
      integer(c_int) function rdb_colphys_mlp_config_default(cfg) bind(c, name="rdb_colphys_mlp_config_default")
         import :: c_int, rdb_colphys_mlp_config_t
         type(rdb_colphys_mlp_config_t), intent(out) :: cfg
      end function
      integer(c_int) function rdb_colphys_mlp_collect(handle, dt, n_steps, stride, n_samples) &
         bind(c, name="rdb_colphys_mlp_collect")
         import :: c_int, c_double, c_ptr, c_long_long
         type(c_ptr), value :: handle
         real(c_double), value :: dt
         integer(c_int), value :: n_steps, stride
         integer(c_long_long), intent(out) :: n_samples
      end function
      integer(c_int) function rdb_colphys_mlp_train(handle, cfg, train_loss, val_loss, n_hist, n_done) &
         bind(c, name="rdb_colphys_mlp_train")
         import :: c_int, c_float, c_ptr, rdb_colphys_mlp_config_t
         type(c_ptr), value :: handle
         type(rdb_colphys_mlp_config_t), intent(in) :: cfg
         real(c_float), intent(out) :: train_loss(*), val_loss(*)
         integer(c_int), value :: n_hist
         integer(c_int), intent(out) :: n_done
      end function
      integer(c_int) function rdb_colphys_mlp_save(handle, path) bind(c, name="rdb_colphys_mlp_save")
         import :: c_int, c_char, c_ptr
         type(c_ptr), value :: handle
         character(kind=c_char), intent(in) :: path(*)
      end function
      integer(c_int) function rdb_colphys_mlp_load(handle, path) bind(c, name="rdb_colphys_mlp_load")
         import :: c_int, c_char, c_ptr
         type(c_ptr), value :: handle
         character(kind=c_char), intent(in) :: path(*)
      end function
      integer(c_int) function rdb_colphys_set_guard(handle, mode) bind(c, name="rdb_colphys_set_guard")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: mode
      end function
      integer(c_int) function rdb_colphys_momentum_config_default(cfg) &
         bind(c, name="rdb_colphys_momentum_config_default")
         import :: c_int, rdb_colphys_momentum_config_t
         type(rdb_colphys_momentum_config_t), intent(out) :: cfg
      end function
      !> vdiff_apply_momentum (rdb_ocean_vdiff.F90:455-718) on the handle's
      !> face velocities, with its kv or K_v_momentum.
      integer(c_int) function rdb_colphys_vdiff_momentum(handle, dt, cfg) bind(c, name="rdb_colphys_vdiff_momentum")
         import :: c_int, c_double, c_ptr, rdb_colphys_momentum_config_t
         type(c_ptr), value :: handle
         real(c_double), value :: dt
         type(rdb_colphys_momentum_config_t), intent(in) :: cfg
      end function
      !> rdb_ocean_create_from_string analogue: the config from namelist text.
      integer(c_int) function rdb_colphys_create_from_string(nml_text, nml_len, nx, ny, nz, n_passive, wet, &
                                                             k_top, sort_key, handle) &
         bind(c, name="rdb_colphys_create_from_string")
         import :: c_int, c_double, c_ptr, c_char
         character(kind=c_char), intent(in) :: nml_text(*)
         integer(c_int), value :: nml_len, nx, ny, nz, n_passive
         real(c_double), intent(in) :: wet(*)
         type(c_ptr), value :: k_top, sort_key
         type(c_ptr), intent(out) :: handle
      end function
      !> Keep heat/salt_budget_vdiff (read with colphys_get).
      integer(c_int) function rdb_colphys_set_budgets(handle, enable) bind(c, name="rdb_colphys_set_budgets")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: enable
      end function
      integer(c_int) function rdb_colphys_refresh_host(handle) bind(c, name="rdb_colphys_refresh_host")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function
      integer(c_int) function rdb_colphys_get_step_count(handle, n) bind(c, name="rdb_colphys_get_step_count")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), intent(out) :: n
      end function
      integer(c_int) function rdb_colphys_get_time(handle, t) bind(c, name="rdb_colphys_get_time")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         real(c_double), intent(out) :: t
      end function
      !> ocean_remap_tracer_field (rdb_ocean_remap.F90:236) on DEVICE arrays: pass
      !> c_loc of device-resident data (host_data use_device, or hipMalloc'd
      !> memory); c_out may be c_in (in place); budget / clamp_out may be
      !> c_null_ptr; stream c_null_ptr for the null stream.  Returns after
      !> enqueueing the kernels.
      integer(c_int) function rdb_remap_tracer_field(method, nx, ny, nz, h_old, h_new, c_in, c_out, bnd_extrap, &
                                                     nonunif, budget, clamp_out, stream) &
         bind(c, name="rdb_remap_tracer_field")
         import :: c_int, c_ptr
         integer(c_int), value :: method, nx, ny, nz
         type(c_ptr), value :: h_old, h_new, c_in, c_out
         integer(c_int), value :: bnd_extrap, nonunif
         type(c_ptr), value :: budget, clamp_out, stream
      end function
      !> remap_x_face_velocity (axis 0) / remap_y_face_velocity (axis 1)
      !> (rdb_ocean_remap.F90:967-1129) on device arrays, in place.
      integer(c_int) function rdb_remap_face_velocity(axis, nx, ny, nz, h_old, h_new, u, method, conserve_ke, &
                                                      zlevel_faces, bnd_extrap, nonunif, stream) &
         bind(c, name="rdb_remap_face_velocity")
         import :: c_int, c_ptr
         integer(c_int), value :: axis, nx, ny, nz
         type(c_ptr), value :: h_old, h_new, u
         integer(c_int), value :: method, conserve_ke, zlevel_faces, bnd_extrap, nonunif
         type(c_ptr), value :: stream
      end function
      !> remap_column (rdb_remap_column.F90:328) on ncol device columns (ncol, nz).
      integer(c_int) function rdb_remap_columns(method, ncol, nz, dz_old, dz_new, q_old, q_new, bnd_extrap, &
                                                nonunif, stream) bind(c, name="rdb_remap_columns")
         import :: c_int, c_ptr
         integer(c_int), value :: method, ncol, nz
         type(c_ptr), value :: dz_old, dz_new, q_old, q_new
         integer(c_int), value :: bnd_extrap, nonunif
         type(c_ptr), value :: stream
      end function
      !> ocean_remap_scan_preconditions (rdb_ocean_remap.F90:1348) on device arrays.
      integer(c_int) function rdb_remap_scan_preconditions(nx, ny, nz, h_old, h_new, rel_tol, n_bad, worst_rel, &
                                                           worst_neg, stream) &
         bind(c, name="rdb_remap_scan_preconditions")
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: nx, ny, nz
         type(c_ptr), value :: h_old, h_new
         real(c_double), value :: rel_tol
         integer(c_int), intent(out) :: n_bad
         real(c_double), intent(out) :: worst_rel, worst_neg
         type(c_ptr), value :: stream
      end function
      integer(c_int) function rdb_colphys_last_error(idx, buf, cap) bind(c, name="rdb_colphys_last_error")
         import :: c_int, c_char
         integer(c_int), value :: idx
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int), value :: cap
      end function
   end interface

contains

   !> Upload a 3-D field (n1, n2, n3) by name.
   integer(c_int) function colphys_set(handle, name, data) result(st)
      type(c_ptr), intent(in) :: handle
      character(len=*), intent(in) :: name
      real(c_double), intent(in) :: data(:, :, :)
      st = rdb_colphys_set_field(handle, trim(name)//c_null_char, data, &
                                 int(size(data, 1), c_int), int(size(data, 2), c_int), int(size(data, 3), c_int))
   end function colphys_set

   !> Download a 3-D field (n1, n2, n3) by name.
   integer(c_int) function colphys_get(handle, name, out) result(st)
      type(c_ptr), intent(in) :: handle
      character(len=*), intent(in) :: name
      real(c_double), intent(out) :: out(:, :, :)
      st = rdb_colphys_get_field(handle, trim(name)//c_null_char, out, &
                                 int(size(out, 1), c_int), int(size(out, 2), c_int), int(size(out, 3), c_int))
   end function colphys_get

   !> Upload a 3-D field (n1, n2, n3) of one ensemble member (0-based) by name.
   integer(c_int) function colphys_member_set(handle, member, name, data) result(st)
      type(c_ptr), intent(in) :: handle
      integer, intent(in) :: member
      character(len=*), intent(in) :: name
      real(c_double), intent(in) :: data(:, :, :)
      st = rdb_colphys_member_set_field(handle, int(member, c_int), trim(name)//c_null_char, data, &
                                        int(size(data, 1), c_int), int(size(data, 2), c_int), &
                                        int(size(data, 3), c_int))
   end function colphys_member_set

   !> Download a 3-D field (n1, n2, n3) of one ensemble member (0-based) by name.
   integer(c_int) function colphys_member_get(handle, member, name, out) result(st)
      type(c_ptr), intent(in) :: handle
      integer, intent(in) :: member
      character(len=*), intent(in) :: name
      real(c_double), intent(out) :: out(:, :, :)
      st = rdb_colphys_member_get_field(handle, int(member, c_int), trim(name)//c_null_char, out, &
                                        int(size(out, 1), c_int), int(size(out, 2), c_int), &
                                        int(size(out, 3), c_int))
   end function colphys_member_get

   !> Most recent error-ring message as a Fortran string.
   function colphys_last_error() result(msg)
      character(len=512) :: msg
      character(kind=c_char) :: buf(512)
      integer(c_int) :: n, i
      buf = ' '
      n = rdb_colphys_last_error(0_c_int, buf, 512_c_int)
      msg = ' '
      do i = 1, min(int(n), 511)
         msg(i:i) = buf(i)
      end do
   end function colphys_last_error

end module rdb_colphys
