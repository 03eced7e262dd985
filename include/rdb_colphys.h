/*
 * rdb_colphys.h -- C ABI of the MI355X column-physics engine (vertical mixing).
 *
 * This library replaces, for one call site, the reference's per-column
 * vertical-mixing chain of the Roundabout ocean model:
 *
 *   vmix_build_coefficients   src/core/ocean/dynamics/split_rk2/rdb_ocean_dyn.F90:1473-1634
 *     vmix_compute_pp81       src/parameterizations/vertical/rdb_ocean_vmix.F90:1007-1074
 *     vmix_apply_kpp_overlay  src/parameterizations/vertical/rdb_ocean_vmix.F90:1076-1516
 *     vmix_split_kd_heat_salt src/parameterizations/vertical/rdb_ocean_vmix.F90:1850-1934
 *     vmix_assemble           src/parameterizations/vertical/rdb_ocean_vmix.F90:2183-2385
 *   ocean_surface_flux_apply_tracers
 *                             src/parameterizations/vertical/rdb_ocean_surface_flux.F90:959-1142
 *   vdiff_apply_tracers       src/parameterizations/vertical/rdb_ocean_vdiff.F90:1107-1414
 *   vmix_apply_nonlocal_tendencies
 *                             src/parameterizations/vertical/rdb_ocean_vmix.F90:1605-1714
 *   eos_wright_impl           src/equation_of_state/rdb_eos.F90:909-959
 *
 * Conventions kept from the reference ABI (include/rdb_ocean.h):
 *   - every entry point returns an int status; 0 is success and the codes
 *     below carry the same numeric values as rdb_ocean.h:62-98;
 *   - the specific reason for a failure goes to an error ring read with
 *     rdb_colphys_last_error (snprintf semantics, rdb_ocean.h:302-314);
 *   - the handle is opaque, NULL on failure, and destroy is idempotent and
 *     NULL-safe (rdb_ocean.h:113-119);
 *   - at most one live handle per process (rdb_ocean_api.F90:23-29);
 *   - host arrays are Fortran column-major with i fastest, k = 1 at the bed
 *     and k = nz at the surface, interfaces k = 1..nz+1 (k = 1 bed,
 *     k = nz+1 surface), exactly the reference's multilayer_state layout
 *     (src/core/rdb_multilayer_state.F90:105-339);
 *   - the caller owns host arrays (copied from / into); the library owns the
 *     device buffers.  One HIP stream per handle, single host thread, no
 *     re-entrancy.  Calls that return host-visible data synchronise first.
 *
 * No torch types, no HIP types: plain pointers and sizes only.
 */
#ifndef RDB_COLPHYS_H
#define RDB_COLPHYS_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: identical numeric values to include/rdb_ocean.h:82-97. */
#define RDB_COLPHYS_OK 0
#define RDB_COLPHYS_ERR_CONFIG_PARSE 1
#define RDB_COLPHYS_ERR_CONFIG_VALIDATE 2
#define RDB_COLPHYS_ERR_SETUP 3
#define RDB_COLPHYS_ERR_IO 5
#define RDB_COLPHYS_ERR_BAD_HANDLE 10
#define RDB_COLPHYS_ERR_ALREADY_EXISTS 11
#define RDB_COLPHYS_ERR_NOT_INITIALISED 12
#define RDB_COLPHYS_ERR_NOT_FOUND 13
#define RDB_COLPHYS_ERR_BAD_SHAPE 14
/* vmix_assemble's guard status (rdb_ocean_vmix.F90:2248-2251, 2265-2277). */
#define RDB_COLPHYS_GUARD_TRIPPED 1

/* Precision of the diffusivity assembly (split precision: FP32 assembly, FP64 solve). */
#define RDB_COLPHYS_PREC_FP64 0  /* every stage in FP64                       */
#define RDB_COLPHYS_PREC_SPLIT 1 /* EOS..clip in FP32, Thomas + totals in FP64 */

/* Closure used for kv/kt/ks. */
#define RDB_COLPHYS_CLOSURE_PLAIN 0     /* PP81 + KPP (the reference)              */
#define RDB_COLPHYS_CLOSURE_SURROGATE 1 /* MLP with per-column plain fallback      */

/*
 * Knob set of the path.  Field names follow ocean_vmix_t
 * (rdb_ocean_vmix.F90:198-611), eos_t (rdb_eos.F90:537-554),
 * ocean_surface_flux_t (rdb_ocean_surface_flux.F90:203-215) and
 * ocean_vdiff_t.  rdb_colphys_config_default() fills the reference defaults
 * (src/core/rdb_config.F90:1743-1827).
 */
typedef struct rdb_colphys_config {
  /* PP81 interior closure (rdb_ocean_vmix.F90:231-262) */
  double pp81_nu0;      /* 1e-2  m^2/s */
  double pp81_nu_bg;    /* 1e-4  m^2/s */
  double pp81_kappa_bg; /* 1e-5  m^2/s */
  double pp81_alpha;    /* 5     */
  double rho0;          /* 1035  kg/m^3 (vmix, eos and surface flux share it) */
  double shear2_floor;  /* 1e-10 s^-2 */
  /* KPP overlay (rdb_ocean_vmix.F90:311-348) */
  int use_kpp;          /* 1 */
  int pad0;
  double ri_crit;       /* 0.3 */
  double c_vt2;         /* 1.8 */
  double cs_nonlocal;   /* 6.3 */
  /* EOS handle: dimensional linear sensitivities used by KPP's B0, and the
   * scalar reference pressure of the Wright EOS (rdb_eos.F90:537-554). */
  double alpha_T;       /* 1.7e-4 kg/m^3/degC */
  double beta_S;        /* 7.6e-4 kg/m^3/psu  */
  double p_ref;         /* 0 Pa */
  /* surface flux slot (rdb_ocean_surface_flux.F90:203-215) */
  double cp;            /* 3992 J/kg/K */
  int has_heat;         /* 1: deposit Q_heat */
  int has_salt;         /* 1: deposit Q_salt */
  /* vmix_assemble gate (rdb_ocean_vmix.F90:352-380) */
  double kv_bg;         /* 1e-4 */
  double kt_bg;         /* 1e-5 */
  double ks_bg;         /* 1e-5 */
  double kv_max;        /* huge */
  double kd_max;        /* huge */
  int vmix_guard;       /* 0 */
  /* contributors that are no-ops at the defaults (their knobs follow
   * below); kd_smooth_iterations > 0 needs a horizontal halo and is refused
   * at create (RDB_COLPHYS_ERR_CONFIG_VALIDATE) */
  int conv_enable;      /* 0: convective adjustment (rdb_ocean_vmix.F90:1750-1799) */
  double kv_ml_invz2;   /* 0: kv += K (hmix/z)^2 in the top hmix_fixed (:1716-1748) */
  int ddiff_enable;     /* 0: double diffusion in the heat/salt split (:1936-2028) */
  int kd_smooth_iterations; /* 0 */
  /* vdiff (rdb_ocean_vdiff.F90 face_thick :1928-1962) */
  int use_harmonic;     /* 0: arithmetic face thickness */
  int precision;        /* RDB_COLPHYS_PREC_* */
  /* ---- contributors off at the defaults; when on, the step runs the
   * staged kernels (ocean_vmix_t, rdb_ocean_vmix.F90:305-544) ---- */
  double hmix_fixed;            /* 20 m: depth scale of kv_ml_invz2 */
  double conv_kd;               /* 1 m^2/s: convective adjustment diffusivity */
  double conv_prandtl;          /* 1 */
  double conv_n2_thresh;        /* 0 s^-2 */
  double ddiff_strat_param_max; /* 2.55: R_rho cutoff of salt fingering */
  double ddiff_kappa_s;         /* 1e-4 m^2/s */
  double ddiff_exp1;            /* 1 */
  double ddiff_exp2;            /* 3 */
  double ddiff_param1;          /* 0.909 */
  double ddiff_param2;          /* 4.6 */
  double ddiff_param3;          /* -0.54 */
  double ddiff_mol_diff;        /* 1.5e-6 m^2/s */
  int ddiff_use_k90;            /* 0 */
  /* ---- EPBL (ocean_epbl_t, rdb_ocean_epbl.F90:148-330): replaces the
   * KPP overlay when on (use_kpp is then ignored, as the reference's
   * configure_ocean_epbl turns it off) ---- */
  int epbl_enable;              /* 0 */
  int epbl_mstar_scheme;        /* RDB_EPBL_MSTAR_OM4 */
  int epbl_vstar_scheme;        /* RDB_EPBL_VSTAR_CUBE_ROOT */
  int epbl_combine;             /* RDB_EPBL_COMBINE_ADD */
  int epbl_mld_iteration;       /* 1 */
  int epbl_mld_max_its;         /* 20 */
  int epbl_mld_bisection;       /* 0 */
  int epbl_mld_use_prev_guess;  /* 0 */
  int epbl_use_lt;              /* 0: Langmuir enhancement (LF17) */
  int epbl_lt_scheme;           /* RDB_EPBL_LT_RESCALE */
  int epbl_tke_diags;           /* 0: keep the column TKE budget terms */
  double epbl_mstar;            /* 1.2 (constant scheme) */
  double epbl_mstar_cap;        /* -1: off */
  double epbl_mstar_coef1;      /* 0.3 */
  double epbl_c_ek;             /* 0.085 */
  double epbl_mstar_conv_adj;   /* 0 */
  double epbl_rh18_cn1, epbl_rh18_cn2, epbl_rh18_cn3; /* 0.275, 8, -5 */
  double epbl_rh18_cs1, epbl_rh18_cs2;                /* 0.2, 0.4 */
  double epbl_nstar;            /* 0.2 */
  double epbl_tke_decay;        /* 2.5 */
  double epbl_wstar_ustar_coef; /* 1 */
  double epbl_vstar_scale_fac;  /* 1 */
  double epbl_vstar_surf_fac;   /* 1.2 */
  double epbl_von_karman;       /* 0.41 */
  double epbl_ekman_scale_coef; /* 1 */
  double epbl_min_mix_len;      /* 0 m */
  double epbl_mixlen_exponent;  /* 2 */
  double epbl_translay_scale;   /* 0.1 */
  double epbl_mld_tol;          /* 1 m */
  double epbl_omega;            /* 7.2921e-5 1/s */
  double epbl_omega_frac;       /* 0 */
  double epbl_ustar_min;        /* 1e-8 m/s */
  double epbl_prandtl;          /* 1 */
  double epbl_lt_enhance_coef;  /* 0.447 */
  double epbl_lt_enhance_exp;   /* -1.33 */
  double epbl_lt_max_enhance;   /* 5 */
  double epbl_la_frac_hbl;      /* 0.04 */
  double epbl_lt_lac1, epbl_lt_lac2, epbl_lt_lac3, epbl_lt_lac4, epbl_lt_lac5; /* -0.87, 0, 0, 0.95, 0.95 */
} rdb_colphys_config;

/* EPBL scheme tags (rdb_ocean_epbl.F90:68-104). */
#define RDB_EPBL_MSTAR_CONSTANT 1
#define RDB_EPBL_MSTAR_OM4 2
#define RDB_EPBL_MSTAR_RH18 3
#define RDB_EPBL_VSTAR_CUBE_ROOT 1
#define RDB_EPBL_VSTAR_RH18 2
#define RDB_EPBL_COMBINE_ADD 1
#define RDB_EPBL_COMBINE_MAX 2
#define RDB_EPBL_COMBINE_MAX_SHEAR 3 /* without kappa-shear: exactly ADD */
#define RDB_EPBL_LT_RESCALE 1
#define RDB_EPBL_LT_ADDITIVE 2

typedef struct rdb_colphys rdb_colphys;

/* Per-handle measurement counters (filled by rdb_colphys_stats). */
typedef struct rdb_colphys_counters {
  long long steps;             /* column steps are steps * wet_columns     */
  long long wet_columns;       /* packed wet columns                       */
  long long land_columns;      /* columns outside the wet mask             */
  double last_kernel_ms;       /* HIP-event time per step, last replay     */
  double sum_kernel_ms;        /* sum of timed step launches               */
  long long timed_launches;    /* number of launches in sum_kernel_ms      */
  double algorithmic_bytes_per_step; /* sum over wet columns of B_col(L_c), L_c = live levels */
  double reference_bytes_per_step;   /* B_col(nz) * wet_columns (all levels)  */
  long long live_levels;       /* sum of live levels over wet columns      */
  double fallback_fraction;    /* surrogate closure: share of wet columns
                                  whose guard tripped in the last step      */
  int block_threads;           /* autotuned launch shape                   */
  int order_mode;              /* 0 natural order, 1 sorted by depth       */
  int lds_levels;              /* autotuned live levels kept in LDS        */
  int graph_replay;            /* 1: the last rdb_colphys_step replayed its
                                  steps from a captured hipGraph (the fused
                                  step, the ensemble, the learned closure);
                                  0: it launched the staged kernels          */
  double mlp_forward_ms;       /* surrogate closure, timing on: HIP-event
                                  time of the network's forward GEMMs in the
                                  last step                                 */
  double mlp_flop_per_step;    /* 2 * sum of layer (in x out) * wet columns */
} rdb_colphys_counters;

/* ---- configuration ------------------------------------------------------ */

/* Fill *cfg with the reference defaults. */
int rdb_colphys_config_default(rdb_colphys_config *cfg);

/* Bytes per working-precision float of the state arrays (always 8);
 * mirrors rdb_working_precision (rdb_ocean.h:300). */
int rdb_colphys_working_precision(void);

/* ---- lifecycle (mirrors rdb_ocean_create_from_string / destroy,
 *      include/rdb_ocean.h:108-119) ---------------------------------------- */

/*
 * Build a plan for an nx_tot x ny_tot grid of columns with nz layers and
 * n_passive passive tracers (diffused with ks, like MOM6's Kd_salt rule,
 * rdb_ocean_vdiff.F90:1114-1131).
 *   wet      (nx_tot, ny_tot)  cell wet mask (1 wet, 0 land), as ms%wet_mask;
 *   k_top    (nx_tot, ny_tot)  static first-live scan start (ms%k_top), or
 *                              NULL for nz everywhere;
 *   sort_key (nx_tot, ny_tot)  per-column depth key (e.g. live-layer count)
 *                              used only to order wet columns so that a
 *                              wavefront holds columns of similar depth, or
 *                              NULL for natural order.
 * Columns with wet <= 0.5 are "land": the library keeps them out of the
 * per-step launch and refreshes them whenever one of their inputs changes.
 */
int rdb_colphys_create(const rdb_colphys_config *cfg, int nx_tot, int ny_tot,
                       int nz, int n_passive, const double *wet,
                       const int *k_top, const int *sort_key,
                       rdb_colphys **handle_out);

/* Idempotent: NULL or an already-destroyed handle is a no-op success. */
int rdb_colphys_destroy(rdb_colphys **handle);

/* ---- create from namelist text (mirrors rdb_ocean_create_from_string,
 *      include/rdb_ocean.h:108-109) ---------------------------------------
 * nml_text need not be NUL-terminated; nml_len is authoritative.  Read:
 *   &ocean_vmix_nml  (src/core/config/rdb_config_register_vmix.F90:738-870):
 *       use_kpp, pp81_nu0, pp81_nu_bg (also kv_bg), pp81_kappa_bg (also
 *       kt_bg and ks_bg), pp81_alpha, shear2_floor, kpp_ri_crit,
 *       kpp_cs_nonlocal, kpp_c_vt2, kv_max, kd_max, vmix_guard,
 *       harmonic_visc, kv_ml_invz2, hmix_fixed, kd_smooth_iterations;
 *       switches of
 *       branches this engine does not run (use_closure = .false.,
 *       interior_closure = 'none', direct_stress, bkgnd_profile,
 *       bkgnd_henyey, bkgnd_prandtl != -1, buoyancy_coeffs = 'eos') return
 *       RDB_COLPHYS_ERR_CONFIG_VALIDATE; the driver's time-step and
 *       advection knobs are accepted and have no effect;
 *   &ocean_eos_nml   eos ('wright' only), c_p, p_ref;
 *   &ocean_conv_nml  (:226-250) enable, kd_conv, prandtl_conv, n2_thresh;
 *   &ocean_ddiff_nml (:252-292) enable, strat_param_max, kappa_ddiff_s,
 *       ddiff_exp1, ddiff_exp2, param1..param3, mol_diff, use_k90;
 *   &ocean_epbl_nml  (:294-393) every key of the group (mstar_scheme,
 *       vel_scale_scheme, combine and lt_scheme by name);
 *   &ocean_ic_nml    rho_0;
 *   &ocean_vdiff_nml (rdb_colphys_momentum_config_from_string) the momentum
 *       switches implicit_stress, implicit_drag, implicit_top_drag,
 *       hvel_mom6, hvel_harmonic, hvel_upwind, bbl_glue, hbbl_visc,
 *       bbl_piston, accel_visc_rem (-> remnant).
 * Other groups of a model namelist are skipped; an unknown key in a group
 * read here is RDB_COLPHYS_ERR_CONFIG_PARSE.  Unlisted knobs keep the
 * reference defaults.  The grid, wet mask, k_top and sort key are given as
 * for rdb_colphys_create. */
int rdb_colphys_config_from_string(const char *nml_text, int nml_len, rdb_colphys_config *cfg);
int rdb_colphys_create_from_string(const char *nml_text, int nml_len, int nx_tot, int ny_tot, int nz,
                                   int n_passive, const double *wet, const int *k_top, const int *sort_key,
                                   rdb_colphys **handle_out);

/* ---- queries (mirror rdb_ocean_get_time / get_step_count /
 *      get_grid_info, include/rdb_ocean.h:283-293) ------------------------
 * Time is the sum of dt over the steps taken (rdb_colphys_step and the
 * plain steps of rdb_colphys_mlp_collect); nghost is 0 (the extents are the
 * caller's full arrays). */
int rdb_colphys_get_time(rdb_colphys *h, double *t_out);
int rdb_colphys_get_step_count(rdb_colphys *h, int *step_out);
int rdb_colphys_get_grid_info(rdb_colphys *h, int *nx, int *ny, int *nz, int *nghost);

/* ---- host pointers with a generation (mirror rdb_ocean_refresh_host and
 *      the rdb_ocean_get_*_ptr getters, include/rdb_ocean.h:321-406) ------
 * Every getter returns a pointer into a full-extent host mirror of the
 * field (reference layout, double) with its extents and generation = the
 * handle's step count at the call, WITHOUT copying from the device (the
 * first request of a field creates and fills its mirror).  When the
 * generation differs from the one last seen, call rdb_colphys_refresh_host
 * before reading: it copies every mirror older than the device state (and
 * is free when all are current).  rho_layer is the Wright EOS of the
 * current T, S and h.  The pointers stay valid until destroy. */
int rdb_colphys_refresh_host(rdb_colphys *h);
int rdb_colphys_get_h_layer_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *nz, int *gen);
int rdb_colphys_get_u_face_x_layer_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *nz, int *gen);
int rdb_colphys_get_v_face_y_layer_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *nz, int *gen);
int rdb_colphys_get_rho_layer_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *nz, int *gen);
int rdb_colphys_get_kv_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *nz, int *gen);
int rdb_colphys_get_kt_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *nz, int *gen);
int rdb_colphys_get_ks_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *nz, int *gen);
int rdb_colphys_get_wet_t_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *gen);
int rdb_colphys_get_tau_x_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *gen);
int rdb_colphys_get_tau_y_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *gen);
int rdb_colphys_get_q_heat_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *gen);
int rdb_colphys_get_q_salt_ptr(rdb_colphys *h, void **ptr, int *nx, int *ny, int *gen);

/* ---- tracers by name (mirror rdb_ocean_get_tracer_count /
 *      list_tracers / get_tracer_ptr, include/rdb_ocean.h:393-406) -------
 * The registry is temperature, salinity, tracer_1 .. tracer_<n_passive>;
 * list_tracers has snprintf semantics and returns the full name length. */
int rdb_colphys_get_tracer_count(rdb_colphys *h, int *count_out);
int rdb_colphys_list_tracers(rdb_colphys *h, int idx, char *buf, int cap);
int rdb_colphys_get_tracer_ptr(rdb_colphys *h, const char *name, int name_len, void **ptr, int *nx, int *ny,
                               int *nz, int *gen);

/* ---- narrow setters (mirror rdb_ocean_set_h/u/v/wind/heat_flux/
 *      salt_flux/tracer, include/rdb_ocean.h:408-446) ---------------------
 * Each pushes only what it changes.  Shapes as the reference's: nx, ny, nz
 * are the cell extents, face arrays carry the +1 in their staggered
 * dimension. */
int rdb_colphys_set_h(rdb_colphys *h, const double *h_data, int nx, int ny, int nz);
int rdb_colphys_set_u(rdb_colphys *h, const double *u_data, int nx, int ny, int nz);
int rdb_colphys_set_v(rdb_colphys *h, const double *v_data, int nx, int ny, int nz);
int rdb_colphys_set_wind(rdb_colphys *h, const double *taux_data, const double *tauy_data, int nx, int ny);
int rdb_colphys_set_heat_flux(rdb_colphys *h, const double *q_data, int nx, int ny);
int rdb_colphys_set_salt_flux(rdb_colphys *h, const double *q_data, int nx, int ny);
int rdb_colphys_set_tracer(rdb_colphys *h, const char *name, int name_len, const double *data, int nx, int ny,
                           int nz);

/* ---- state transfer (mirrors the rdb_ocean_set_* / get_*_ptr pairs,
 *      include/rdb_ocean.h:338-446, with explicit copies) -------------------
 *
 * Field names and host shapes (n1, n2, n3):
 *   "h_layer"        (nx, ny, nz)     layer thickness, m
 *   "u_face_x_layer" (nx+1, ny, nz)   west-face x velocity, m/s
 *   "v_face_y_layer" (nx, ny+1, nz)   south-face y velocity, m/s
 *   "temperature"    (nx, ny, nz)     degC
 *   "salinity"       (nx, ny, nz)     psu
 *   "tracer_<n>"     (nx, ny, nz)     passive tracer n = 1..n_passive
 *   "tau_x"          (nx+1, ny, 1)    east-face wind stress, N/m^2
 *   "tau_y"          (nx, ny+1, 1)    north-face wind stress, N/m^2
 *   "stress_shelf"   (nx, ny, 1)      ice-shelf stress magnitude (0 default)
 *   "Q_heat"         (nx, ny, 1)      W/m^2
 *   "Q_salt"         (nx, ny, 1)      salt flux
 *   "bl_depth"       (nx, ny, 1)      KPP boundary-layer depth (lagged seed)
 *   "kv" "kt" "ks"   (nx, ny, nz+1)   get only (set allowed for restart kv)
 *   "gamma_t" "gamma_s" (nx, ny, nz+1) get only
 *   "rho_layer"      (nx, ny, nz)     get; set overrides the EOS output
 *   "stress_mag"     (nx, ny, 1)      get only
 *   "b0"             (nx, ny, 1)      get only
 *   "kbl"            (nx, ny, 1)      get only; KPP crossing layer (0 none,
 *                                     -1 where the surrogate closure was used),
 *                                     stored as double
 *   "coriolis"       (nx, ny, 1)      Coriolis parameter f (surrogate input)
 *   "wet_t"          (nx, ny, 1)      get only; the wet mask given at create
 * A shape mismatch returns RDB_COLPHYS_ERR_BAD_SHAPE, an unknown name
 * RDB_COLPHYS_ERR_NOT_FOUND.
 */
int rdb_colphys_set_field(rdb_colphys *h, const char *name, const double *data,
                          int n1, int n2, int n3);
int rdb_colphys_get_field(rdb_colphys *h, const char *name, double *out,
                          int n1, int n2, int n3);

/* ---- compute ------------------------------------------------------------- */

/* A1: rho_layer from (T, S, h) with the Wright EOS (rdb_eos.F90:909-959). */
int rdb_colphys_eos(rdb_colphys *h);

/* A3-A10: PP81 -> KPP overlay -> split -> assemble on the device rho_layer;
 * writes kv, kt, ks, gamma_t/s, bl_depth, b0, kbl
 * (vmix_build_coefficients, rdb_ocean_dyn.F90:1473-1634).  Returns
 * RDB_COLPHYS_GUARD_TRIPPED when cfg.vmix_guard is on and a negative or NaN
 * diffusivity reached the gate. */
int rdb_colphys_coefficients(rdb_colphys *h, double dt);

/* A2: surface heat/salt deposit into the first live layer
 * (ocean_surface_flux_apply_tracers, rdb_ocean_surface_flux.F90:959-1070). */
int rdb_colphys_surface_flux(rdb_colphys *h, double dt);

/* A11-A13: backward-Euler Thomas solve of T (kt), S and passives (ks)
 * (vdiff_apply_tracers, rdb_ocean_vdiff.F90:1107-1226). */
int rdb_colphys_diffuse_tracers(rdb_colphys *h, double dt);

/* A14: KPP non-local tendencies on T and S
 * (vmix_apply_nonlocal_tendencies, rdb_ocean_vmix.F90:1605-1692). */
int rdb_colphys_nonlocal(rdb_colphys *h, double dt);

/* A1-A14 fused: n_steps column steps in the order of apply_deferred_thermo
 * (rdb_ocean_dyn.F90:3836-3856): EOS -> coefficients -> surface flux ->
 * tracer diffusion -> non-local.  One launch per step over all wet columns,
 * replayed from a captured hipGraph. */
int rdb_colphys_step(rdb_colphys *h, double dt, int n_steps);

/* The vdiff budget slots heat_budget_vdiff and salt_budget_vdiff
 * (src/core/rdb_multilayer_state.F90:403-410): with enable = 1 every tracer
 * solve (fused step or rdb_colphys_diffuse_tracers) adds the per-cell
 * content change h*c_new - h*c_old of T and S to them, exactly as
 * apply_factored_tracer does (rdb_ocean_vdiff.F90:1397-1405).  Enabling
 * allocates them zeroed; they are read and reset through get/set_field
 * "heat_budget_vdiff" / "salt_budget_vdiff" (nx, ny, nz).  Off by default
 * (the step then moves no budget bytes); single grids only. */
int rdb_colphys_set_budgets(rdb_colphys *h, int enable);

/* Block until the handle's stream is idle. */
int rdb_colphys_synchronize(rdb_colphys *h);

/* Choose the launch shape and column order by timing candidates on a copy
 * of the current state (the state is restored afterwards). */
int rdb_colphys_autotune(rdb_colphys *h, double dt);

/* RDB_COLPHYS_PREC_FP64 or RDB_COLPHYS_PREC_SPLIT. */
int rdb_colphys_set_precision(rdb_colphys *h, int mode);

/* RDB_COLPHYS_CLOSURE_PLAIN or RDB_COLPHYS_CLOSURE_SURROGATE. */
int rdb_colphys_set_closure(rdb_colphys *h, int mode);

/* A17: order-invariant column integrals sum(h*T), sum(h*S) over wet columns
 * (integer-bin reproducible sum, after compute_total_tracer_efp,
 * src/core/ocean/diag/rdb_ocean_console_stats.F90:1120). */
int rdb_colphys_heat_salt_totals(rdb_colphys *h, double *heat, double *salt);

/* Counters and timings; enable_timing brackets every rdb_colphys_step
 * graph replay with HIP events on the handle's stream (the bracketed region
 * holds only the step kernels), accumulated into sum_kernel_ms. */
int rdb_colphys_set_timing(rdb_colphys *h, int enable_timing);
int rdb_colphys_stats(rdb_colphys *h, rdb_colphys_counters *out);

/* ---- momentum vertical viscosity ----------------------------------------
 * Backward-Euler vertical viscosity of the face velocities u_face_x_layer and
 * v_face_y_layer: vdiff_apply_momentum -> diffuse_velocity_columns_impl
 * (src/parameterizations/vertical/rdb_ocean_vdiff.F90:455-718, 1416-2115),
 * the other consumer of kv.  Knob names follow ocean_vdiff_t
 * (rdb_ocean_vdiff.F90:77-280); rdb_colphys_momentum_config_default fills its
 * defaults.  The face-located inputs of the folds are fields of the handle
 * (rdb_colphys_set_field):
 *   "lambda_bot_u" (nx+1, ny, 1), "lambda_bot_v" (nx, ny+1, 1)  bottom drag rate, 1/s
 *   "lambda_top_u", "lambda_top_v"                              ice-shelf top drag rate
 *   "cover_u", "cover_v"                                        face ice cover (0 open, 1 covered)
 *   "kv_bbl_u", "kv_bbl_v", "bbl_thick_u", "bbl_thick_v"         per-face BBL (bbl_per_face)
 *   "k_top_u", "k_bot_u", "k_top_v", "k_bot_v"                  ms%k_top_u etc., 1-based
 *                                                               layer indices stored as double
 *                                                               (default nz and 1)
 *   "kv_corner" (nx+1, ny+1, nz+1)                              corner viscosity (use_corner)
 * the wind stress is "tau_x"/"tau_y".  Outputs: the face velocities, and
 * with `remnant` the viscous remnant "visc_rem_u" (nx+1, ny, nz) and
 * "visc_rem_v" (nx, ny+1, nz) (get only). */
typedef struct rdb_colphys_momentum_config {
  int use_kv_source;     /* 1: the closure's kv (vmix); 0: K_v_momentum everywhere (1) */
  int pad0;
  double K_v_momentum;   /* scalar viscosity, m^2/s (0) */
  int implicit_stress;   /* fold tau/rho0 into the top row (0) */
  int implicit_drag;     /* fold lambda_bot into the bed row (0; needs lambda_bot_u/v) */
  int implicit_top_drag; /* fold lambda_top and mask the wind by 1 - cover (0; needs lambda_top, cover) */
  int hvel_mom6;         /* MOM6 face thickness and h_shear (0) */
  int hvel_harmonic;     /* MOM6 HARMONIC_VISC branch (0) */
  int hvel_upwind;       /* near-bed upwind blend of the harmonic branch (1) */
  int bbl_glue;          /* MOM6 bottom-boundary-layer coupling (0; needs hvel_mom6) */
  int bbl_per_face;      /* 1: kv_bbl/bbl_thick fields; 0: bbl_piston*hbbl_visc over hbbl_visc (0) */
  double hbbl_visc;      /* 10 m */
  double bbl_piston;     /* 3e-4 m/s */
  double kv_bbl_bg;      /* 1e-4 m^2/s */
  int zlevel_faces;      /* z-level partial-step faces: min thickness, cut couplings (0) */
  int use_corner;        /* add kv_corner_prandtl * the corner viscosity (0) */
  double kv_corner_prandtl; /* 1 */
  int remnant;           /* also solve for the viscous remnant visc_rem_u/v (0) */
  int remnant_only;      /* the remnant without changing u and v (0) */
} rdb_colphys_momentum_config;

int rdb_colphys_momentum_config_default(rdb_colphys_momentum_config *cfg);
/* The &ocean_vdiff_nml switches over the defaults (see
 * rdb_colphys_config_from_string). */
int rdb_colphys_momentum_config_from_string(const char *nml_text, int nml_len, rdb_colphys_momentum_config *cfg);

/* One backward-Euler step of vertical viscosity on the face velocities with
 * the handle's kv (use_kv_source) or K_v_momentum, h_layer and wet mask; the
 * face velocities the next column step reads are updated with it.  The
 * tracer diffusivity option use_harmonic of the handle's config applies. */
int rdb_colphys_vdiff_momentum(rdb_colphys *h, double dt, const rdb_colphys_momentum_config *cfg);

/* ---- learned closure (no counterpart in the reference) -----------------
 * A multilayer perceptron predicts, per wet column, log10 kt and log10 kv at
 * the interior interfaces and the KPP boundary-layer depth from the column
 * state and forcing (T, S, centred u, v as differences to the live level
 * above and log10 h at every level; |tau|, Q_heat, Q_salt, Coriolis f,
 * the convective velocity scale of the lagged boundary layer, column depth).
 * BF16 weights and activations on MFMA with FP32 accumulation.  Its kv/kt go
 * through the same assembly clip and FP64 tridiagonal solve as the plain
 * path; the non-local transport uses the predicted depth.  A column falls
 * back to the plain PP81 + KPP path when its guard trips. */
#define RDB_COLPHYS_GUARD_STRICT 0  /* any live interface with N^2 < 0, or an input outside the training range */
#define RDB_COLPHYS_GUARD_SURFACE 1 /* N^2 < 0 below the surface layer, or an input out of range */
#define RDB_COLPHYS_GUARD_RANGE 2   /* an input outside the training range only */

typedef struct rdb_colphys_mlp_config {
  int hidden;          /* 256 hidden width */
  int n_hidden;        /* 4 hidden layers (GELU) */
  int batch;           /* 8192 */
  int epochs;          /* passes over the training columns */
  double lr;           /* Adam 1e-3 */
  double beta1, beta2, eps; /* 0.9, 0.999, 1e-8 */
  int seed;            /* initialisation and shuffling */
  int guard_mode;      /* RDB_COLPHYS_GUARD_* used at inference */
  double val_fraction; /* held-out share of the collected columns (0.05) */
} rdb_colphys_mlp_config;

int rdb_colphys_mlp_config_default(rdb_colphys_mlp_config *cfg);

/* Run n_steps plain steps, recording (features before, coefficients after)
 * of every stride-th wet column; *n_samples returns the collected total.
 * The handle's state advances by n_steps. */
int rdb_colphys_mlp_collect(rdb_colphys *h, double dt, int n_steps, int stride, long long *n_samples);

/* Normalise the collected columns and train with Adam on the GPU (forward,
 * backward and update kernels of this library).  Per epoch the mean training
 * and held-out losses (weighted MSE of the normalised targets) go to
 * train_loss / val_loss (n_hist entries at most). */
int rdb_colphys_mlp_train(rdb_colphys *h, const rdb_colphys_mlp_config *cfg, float *train_loss, float *val_loss,
                          int n_hist, int *n_epochs_done);

/* Network + normalisation statistics to / from a file: the features' mean,
 * std and training range (the guard's range test) and the targets' mean,
 * std and training range (a surrogate step clamps every prediction into
 * it).  Files of the previous format, without the targets' range, load and
 * are not clamped.  A later mlp_train call continues from the network and
 * the Adam moments it holds (so successive calls can lower the learning
 * rate). */
int rdb_colphys_mlp_save(rdb_colphys *h, const char *path);
int rdb_colphys_mlp_load(rdb_colphys *h, const char *path);

/* RDB_COLPHYS_GUARD_* for subsequent surrogate steps. */
int rdb_colphys_set_guard(rdb_colphys *h, int guard_mode);

/* ---- learned closure: the MLP on host arrays --------------------------
 * No counterpart in the reference: these run the BF16 MFMA network of the
 * surrogate closure (widths[0] inputs, widths[1..n-2] GELU hidden layers,
 * widths[n-1] linear outputs) on a host batch, so that its kernels can be
 * checked against an FP64 CPU oracle.  params is the unpadded FP32 layout
 * W_0 (widths[1] x widths[0], row-major), b_0, W_1, b_1, ...; x is
 * (batch x widths[0]), y and t and w are (batch x widths[n-1]), row-major.
 * rdb_mlp_loss_grad returns the weighted mean-squared error
 * sum(w (y - t)^2) / sum(w) and its gradient in the params layout. */
int rdb_mlp_eval(const int *widths, int n_widths, const float *params, int batch, const float *x, float *y);
int rdb_mlp_loss_grad(const int *widths, int n_widths, const float *params, int batch, const float *x,
                      const float *t, const float *w, float *loss, float *grad);
/* rdb_mlp_eval with the outputs output-major, yt (widths[n-1] x batch): the
 * layout the surrogate closure's fused inference kernel writes for the
 * column step. */
int rdb_mlp_eval_t(const int *widths, int n_widths, const float *params, int batch, const float *x, float *yt);
/* n_steps training steps of the learned closure on one batch (forward,
 * backward, gradient reduction, Adam with bias correction, BF16 refresh of
 * the weights and their transposes), starting from params with zero Adam
 * moments; params_out receives the updated parameters. */
int rdb_mlp_adam_steps(const int *widths, int n_widths, const float *params, int batch, const float *x,
                       const float *t, const float *w, int n_steps, double lr, double beta1, double beta2,
                       double eps, float *params_out);

/* rdb_mlp_eval without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and results as rdb_mlp_eval; the weights,
 * the padded input and output and the activations live in one 80 MiB
 * scratch array per GPU that the library holds from load to process end
 * (no hipMalloc, hipFree or other allocating routine is called).  A network
 * and batch whose working set exceeds it (for the 381-256x4-149 closure
 * network, batches above about 20000 rows) return RDB_COLPHYS_ERR_BAD_SHAPE
 * before any work.  Runs on the legacy null stream and returns after the
 * outputs are copied back.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (several host threads) race. */
int rdb_mlp_eval_noalloc(const int *widths, int n_widths, const float *params, int batch, const float *x,
                         float *y);

/* ---- batched ensemble (no counterpart in the reference) -----------------
 * Several independent ocean members, each with its own grid (nx x ny x nz),
 * wet mask and optional depth sort key, in one handle.  rdb_colphys_step
 * advances all of them with one persistent launch per step (captured in the
 * hipGraph like the single-grid step); the learned closure, when enabled,
 * runs one network over every member's columns (its levels are the largest
 * member nz, with the members' surfaces aligned) and the guard sends columns
 * back to the plain path per column.  Fields are reached per member; the
 * single-grid entry points (set_field/get_field and the staged routines)
 * refuse an ensemble handle with RDB_COLPHYS_ERR_BAD_HANDLE.  Totals and
 * counters cover all members.  Counts as the one live handle. */
int rdb_colphys_ensemble_create(const rdb_colphys_config *cfg, int n_members, const int *nx, const int *ny,
                                const int *nz, int n_passive, const double *const *wet, const int *const *sort_key,
                                rdb_colphys **out);
int rdb_colphys_member_count(rdb_colphys *h, int *n_members);
int rdb_colphys_member_set_field(rdb_colphys *h, int member, const char *name, const double *data, int n1, int n2,
                                 int n3);
int rdb_colphys_member_get_field(rdb_colphys *h, int member, const char *name, double *out, int n1, int n2, int n3);

/* ---- device routines on host arrays, for verification ------------------ */

/* A7: the device cube root of the KPP path on n host arguments (y[i] =
 * cbrt(x[i])), the routine rdb_sm_cbrt restates
 * (src/shared_module_utilities/rdb_safe_math_pow.inc:107-142), so that it can
 * be checked against the reference's golden bit table
 * (tests/test_safe_math_golden.inc:557-606). */
int rdb_colphys_cbrt(const double *x, double *y, int n);

/* The device exp / log / pow / erfc of the EPBL and double-diffusion paths
 * on n host arguments (fn: RDB_SM_EXP, RDB_SM_LOG, RDB_SM_POW with y the
 * exponents, RDB_SM_ERFC; y may be NULL otherwise): the position-independent
 * rdb_sm_exp / rdb_sm_log / rdb_sm_pow / rdb_sm_erfc
 * (src/shared_module_utilities/rdb_safe_math_core.inc, _pow.inc, _erfc.inc),
 * so that they can be checked against the reference's golden bit tables
 * (tests/test_safe_math_golden.inc: G_EXP, G_LOG, G_POW, G_POWBIG, G_ERFC). */
#define RDB_SM_EXP 0
#define RDB_SM_LOG 1
#define RDB_SM_POW 2
#define RDB_SM_ERFC 3
int rdb_colphys_safe_math(int fn, const double *x, const double *y, double *out, int n);

/* ---- conservative vertical (ALE) remap, on device arrays --------------- */

/* Reconstruction methods (src/core/rdb_constants.F90:208-218).  Any other
 * code runs PLM, as remap_column's default branch does
 * (src/ALE/rdb_remap_column.F90:371-372). */
#define RDB_REMAP_PCM 1
#define RDB_REMAP_PLM 2
#define RDB_REMAP_PPM 3
#define RDB_REMAP_PPM_H4 4
#define RDB_REMAP_PQM 5

/* These entry points take DEVICE pointers (a Fortran caller passes its
 * device-resident arrays, e.g. under host_data use_device) in the natural
 * layout: (nx, ny, nz), i fastest, k = 1 at the bed.  They enqueue their
 * kernels on `stream` (a hipStream_t; NULL = the null stream) and return
 * without waiting, except rdb_remap_scan_preconditions, which returns its
 * scalars.  The library keeps one grow-only work buffer set per device:
 * calls on one device must be ordered (one stream, or synchronised between
 * streams).  The thicknesses must be non-negative with matching column
 * totals (the reference's precondition, rdb_remap_column.F90:120-166);
 * rdb_remap_scan_preconditions checks it. */

/* ocean_remap_tracer_field (src/ALE/rdb_ocean_remap.F90:236-319): remap one
 * tracer concentration from h_old to h_new, concentration form: donor
 * column over the fillers (h <= 1.5e-4 m), the column remap with `method`,
 * each filler's excess added back, the target bounded by its contributors'
 * stencil range, an empty target taking the mean of the source cell at its
 * bottom interface.  c_out may equal c_in (in place, as the reference).
 * bnd_extrap / nonunif: &vcoord_nml remap_boundary_extrap /
 * remap_nonuniform_weights.  budget (optional, accumulated): the per-cell
 * h_new*c_new - h_old*c_old increment of the heat/salt_budget_remap slots;
 * clamp_out (optional, overwritten): h_new*(c_bounded - c_raw), the
 * c_new_out/clamp_out test hooks of the reference. */
int rdb_remap_tracer_field(int method, int nx, int ny, int nz, const double *h_old, const double *h_new,
                           const double *c_in, double *c_out, int bnd_extrap, int nonunif, double *budget,
                           double *clamp_out, void *stream);

/* remap_x_face_velocity (axis 0, u is (nx+1, ny, nz)) and
 * remap_y_face_velocity (axis 1, v is (nx, ny+1, nz))
 * (src/ALE/rdb_ocean_remap.F90:967-1129), in place: face thicknesses are the
 * mean of the two abutting cells (the wall faces take their cell's), or,
 * with zlevel_faces, their minimum with closed layers dropped; conserve_ke
 * rescales the baroclinic anomaly (rescale_anomaly_ke, :1131-1180). */
int rdb_remap_face_velocity(int axis, int nx, int ny, int nz, const double *h_old, const double *h_new, double *u,
                            int method, int conserve_ke, int zlevel_faces, int bnd_extrap, int nonunif,
                            void *stream);

/* remap_column (src/ALE/rdb_remap_column.F90:328-375) on ncol independent
 * columns, arrays (ncol, nz) with the column index fastest.  q_new may
 * equal q_old. */
int rdb_remap_columns(int method, int ncol, int nz, const double *dz_old, const double *dz_new, const double *q_old,
                      double *q_new, int bnd_extrap, int nonunif, void *stream);

/* ocean_remap_scan_preconditions (src/ALE/rdb_ocean_remap.F90:1348-1405):
 * the number of columns with a negative thickness or a column-total
 * mismatch above rel_tol, the largest relative mismatch and the most
 * negative thickness (0 if none).  Synchronises `stream`. */
int rdb_remap_scan_preconditions(int nx, int ny, int nz, const double *h_old, const double *h_new, double rel_tol,
                                 int *n_bad, double *worst_rel, double *worst_neg, void *stream);

/* ---- error ring (mirrors rdb_ocean_last_error, rdb_ocean.h:302-314) ---- */
int rdb_colphys_last_error(int idx, char *buf, int cap);

#ifdef __cplusplus
}
#endif

#endif /* RDB_COLPHYS_H */
