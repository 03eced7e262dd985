/*
 * colphys_oracle.h -- CPU FP64 restatement of the reference vertical-mixing
 * path.  TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the
 * timed CPU baseline.  Nothing in the product path links or calls it.
 *
 * Arrays use the reference layout: Fortran column-major, i fastest,
 * 0-based here: IDX(i, j, k) = i + nx*(j + ny*k), k = 0 is the bed layer.
 * Interface arrays have nz+1 levels (k = 0 bed, k = nz surface).
 */
#ifndef COLPHYS_ORACLE_H
#define COLPHYS_ORACLE_H

#include "../include/rdb_colphys.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_GRAVITY 9.80665
#define ORC_H_VANISHED 1.5e-4
#define ORC_H_DIV_EPS 1.0e-20

double orc_cbrt(double x);

void orc_eos_wright(int nx, int ny, int nz, const double *h, const double *S,
                    const double *T, double *rho, double rho0, double p_ref);

void orc_stress_mag(int nx, int ny, const double *tau_x, const double *tau_y,
                    double *stress_mag);

void orc_pp81(const rdb_colphys_config *cfg, int nx, int ny, int nz,
              const double *u, const double *v, const double *h,
              const double *rho, double *kv, double *kt);

void orc_kpp_overlay(const rdb_colphys_config *cfg, int nx, int ny, int nz,
                     const double *u, const double *v, const double *h,
                     const double *rho, const double *stress_mag,
                     const double *stress_shelf, const double *Q_heat,
                     const double *Q_salt, double *bl_depth, double *kv,
                     double *kt, double *gamma_t, double *gamma_s, double *b0,
                     int *kbl);

void orc_split(int nx, int ny, int nz, const double *kt, double *ks);

int orc_assemble(const rdb_colphys_config *cfg, int nx, int ny, int nz,
                 double *kv, double *kt, double *ks);

void orc_surface_flux(const rdb_colphys_config *cfg, int nx, int ny, int nz,
                      double dt, double *T, double *S, const double *h,
                      const double *wet, const double *Q_heat,
                      const double *Q_salt, const int *k_top);

void orc_build_factorize(int nx, int ny, int nz, int use_harmonic, double dt,
                         const double *K, const double *h, double *a,
                         double *b, double *c);

void orc_apply_factored(int nx, int ny, int nz, double *x, const double *a,
                        const double *b, const double *c, double *rhs);

void orc_vdiff_tracers(int nx, int ny, int nz, int use_harmonic, double dt,
                       const double *kt, const double *ks, const double *h,
                       double *T, double *S, int n_passive, double **passive);

void orc_vdiff_tracers_budget(int nx, int ny, int nz, int use_harmonic, double dt, const double *kt,
                              const double *ks, const double *h, double *T, double *S, int n_passive,
                              double **passive, double *heat_budget, double *salt_budget);

void orc_nonlocal(int nx, int ny, int nz, double dt, const double *gamma_t,
                  const double *gamma_s, const double *h, double *T,
                  double *S);

/* Whole grid state for one handle-equivalent. */
typedef struct orc_state {
  int nx, ny, nz, n_passive;
  double *h, *u, *v, *T, *S;
  double **passive;
  double *stress_mag, *stress_shelf, *Q_heat, *Q_salt, *wet;
  int *k_top;
  double *rho, *kv, *kt, *ks, *gamma_t, *gamma_s, *bl_depth, *b0;
  int *kbl;
  /* scratch for the tracer solve */
  double *a, *b, *c, *rhs;
  /* heat/salt_budget_vdiff slots (NULL: not kept) */
  double *heat_budget, *salt_budget;
  /* EPBL (epbl_oracle.c; read only when cfg->epbl_enable): |f| at centres
   * [nx*ny] (NULL: 0), the persistent MLD, kd_int [nz+1], La, B0 and the six
   * TKE budget planes [6*nx*ny] (NULL unless cfg->epbl_tke_diags) */
  double *coriolis, *mld, *kd_epbl, *la, *b0_epbl, *tke;
} orc_state;

/* One column step in the order of apply_deferred_thermo
 * (rdb_ocean_dyn.F90:3836-3856).  Returns the assemble guard status. */
int orc_column_step(const rdb_colphys_config *cfg, orc_state *st, double dt);

/* Column integrals sum(h*c) over wet cells in long double (checker only). */
void orc_totals(int nx, int ny, int nz, const double *h, const double *c,
                const double *wet, double *total);

int orc_num_threads(void);

/* Momentum vertical viscosity (mom_oracle.c): the switches of
 * diffuse_velocity_columns_impl (rdb_ocean_vdiff.F90:1416-1926). */
typedef struct orc_mom_opts {
  int do_stress, do_drag, do_top, solve_momentum, do_remnant;
  int hvel_mom6, hvel_harmonic, hvel_upwind, bbl_glue, zlevel_faces, do_corner, use_harmonic;
  double rho0, hbbl_visc, kv_bbl_bg, kv_prandtl;
} orc_mom_opts;

void orc_diffuse_velocity_columns(const orc_mom_opts *o, int nu, int nv, int nz, double dt, double *u_face,
                                  const double *h_layer, const double *kv_centre, const double *wet_cell,
                                  int x_face, int nx_cells, int ny_cells, const double *tau_face,
                                  const double *lambda_bot, const double *lambda_top, const double *cover_face,
                                  double *visc_rem_out, const double *kv_bbl_face, const double *bbl_thick_face,
                                  const double *kv_corner, const int *k_top_face, const int *k_bot_face,
                                  double *c_diag, double *rhs);

/* remap_oracle.c: the conservative vertical remap (src/ALE/
 * rdb_remap_column.F90, src/ALE/rdb_ocean_remap.F90).  Column arrays are
 * 0-based, index 0 at the bed. */
int orc_remap_preconditions_ok(int nz, const double *dz_old, const double *dz_new, double rel_tol);
void orc_remap_column(int method, int nz, const double *dz_old, const double *dz_new, const double *q_old,
                      double *q_new, int bnd_extrap, int nonunif);
void orc_pqm_end_value_h4(const double *dz, const double *u, double *csys);
void orc_vl_column_conc(int nz, const double *h, const double *c_in, double *c_out);
void orc_remap_tracer_column(int method, int nz, const double *h_old, const double *h_new, double *c,
                             int bnd_extrap, int nonunif, double *c_new_out, double *clamp_out);
void orc_remap_tracer_field(int method, int nx, int ny, int nz, const double *h_old, const double *h_new,
                            double *c, int bnd_extrap, int nonunif, double *budget, double *c_new_out,
                            double *clamp_out);
void orc_remap_face_velocity(int axis, int nx, int ny, int nz, const double *h_old, const double *h_new,
                             double *u, int method, int conserve_ke, int zlevel_faces, int bnd_extrap,
                             int nonunif);
void orc_remap_scan_preconditions(int nx, int ny, int nz, const double *h_old, const double *h_new, double rel_tol,
                                  int *n_bad, double *worst_rel, double *worst_neg);

/* epbl_oracle.c: the contributors that are off at the defaults (EPBL,
 * KV_ML_INVZ2, convective adjustment, double diffusion) and the
 * position-independent math they use. */
double orc_sm_exp_core(double x, double c);
double orc_sm_exp(double x);
double orc_sm_log(double x);
void orc_sm_log_dd(double x, double *lh, double *ll);
double orc_sm_pow(double x, double y);
double orc_sm_erfc(double x);
void orc_specvol_derivs(double T, double S, double p, double *dsv_dt, double *dsv_ds);
double orc_epbl_find_mstar(const rdb_colphys_config *c, double b0, double ustar, double bld, double absf);
double orc_epbl_mixlen_shape(double z_depth, double mld_guess, double translay, double expo, int shaped);
void orc_epbl_lf17_wave_state(double ustar_w, double rho_ocn, double *u10, double *ustokes, double *kphil);
double orc_epbl_lf17_la(double ustar_w, double zsl, double ustokes, double kphil);
double orc_epbl_lt_enhance(const rdb_colphys_config *c, double la, double b0, double ustar, double bld, double absf,
                           double mstar);
void orc_epbl_column(const rdb_colphys_config *c, int nz, double dt, const double *h, const double *T,
                     const double *S, double tau_tot, double Q_heat, double Q_salt, double wet, double f_centre,
                     int eos_linear, double *kd_int, double *mld, double *b0_out, double *la_out, double *tke);
void orc_epbl_compute_merge(const rdb_colphys_config *c, int nx, int ny, int nz, double dt, const double *h,
                            const double *T, const double *S, const double *stress_mag, const double *stress_shelf,
                            const double *Q_heat, const double *Q_salt, const double *wet, const double *f_centre,
                            double *kd_int, double *mld, double *b0, double *la, double *tke, double *kv,
                            double *kt);
void orc_vmix_invz2(const rdb_colphys_config *c, int nx, int ny, int nz, const double *h, double *kv);
void orc_vmix_convection(const rdb_colphys_config *c, int nx, int ny, int nz, const double *h, const double *rho,
                         const double *bl_depth, double *kv, double *kt);
void orc_ddiff_iface(const rdb_colphys_config *c, double hu, double hl, double Tu, double Tl, double Su, double Sl,
                     double *kd_t, double *kd_s);
void orc_vmix_split_ddiff(const rdb_colphys_config *c, int nx, int ny, int nz, const double *h, const double *T,
                          const double *S, double *kt, double *ks);

#ifdef __cplusplus
}
#endif

#endif
