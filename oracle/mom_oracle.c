/*
 * mom_oracle.c -- CPU FP64 restatement of the reference's backward-Euler
 * vertical viscosity of the face velocities (vdiff_apply_momentum ->
 * diffuse_velocity_columns_impl).  TEST INFRASTRUCTURE ONLY: the checker of
 * the HIP momentum kernel.  Nothing in the product path links or calls it.
 *
 * Follows src/parameterizations/vertical/rdb_ocean_vdiff.F90:
 *   diffuse_velocity_columns_impl  :1416-1926
 *   face_thick                     :1928-1962
 *   vdiff_face_layer               :1964-2065
 *   vdiff_interface_coupling       :2067-2115
 * and rdb_blf_is_live (src/shared_module_utilities/rdb_bed_live_face.inc),
 * operation for operation in the reference's left-to-right order (built
 * without FMA contraction).
 *
 * Layout: Fortran column-major, 0-based.  A face array of nu x nv faces has
 * face (i, j) of layer k at i + nu*(j + nv*k); cell arrays are nx x ny
 * (x nz or nz+1); the corner array is (nx+1) x (ny+1) x (nz+1).  k = 0 is
 * the bed.  k_top_face / k_bot_face are 0-based layer indices here.
 */
#include <math.h>
#include <stddef.h>

#include "colphys_oracle.h"

static double mfmax(double a, double b) { return (a > b) ? a : b; }
static double mfmin(double a, double b) { return (a < b) ? a : b; }

/* face_thick (rdb_ocean_vdiff.F90:1928-1962) */
static double mface_thick(double ha, double hb, int use_harmonic) {
  if (use_harmonic) {
    double sum_h = ha + hb;
    if (sum_h > 1.0e-30) return 2.0 * ha * hb / sum_h;
    return 0.5 * sum_h;
  }
  return 0.5 * (ha + hb);
}

/* vdiff_face_layer (rdb_ocean_vdiff.F90:1964-2065): face thickness and
 * normalised height above the bed of layer k (0-based), advancing the
 * upward accumulators. */
static void mface_layer(double hl_c, double hr_c, double u_k, int k, int kb, int hvel_mom6, int hvel_harmonic,
                        int hvel_upwind, int zlevel_faces, double d_min, double i_hbbl, double *zacc,
                        double *zcol_l, double *zcol_r, double *hvel, double *zint) {
  const double EPS_HVEL = 1.0e-30;
  *zint = 0.0;
  if (hvel_mom6 && !hvel_harmonic) {
    double h_harm = 2.0 * hl_c * hr_c / (hl_c + hr_c + EPS_HVEL);
    double h_arith = 0.5 * (hl_c + hr_c);
    double h_delta = hr_c - hl_c;
    *zcol_l = *zcol_l + hl_c;
    *zcol_r = *zcol_r + hr_c;
    if (k >= kb) *zacc = *zacc + h_harm;
    double z_clear = mfmax(*zcol_l, *zcol_r) + d_min;
    *zint = mfmax(*zacc, z_clear) * i_hbbl;
    *hvel = h_arith;
    if (u_k * h_delta > 0.0) {
      double z2 = *zint;
      double botfn = 1.0 / (1.0 + 0.09 * z2 * z2 * z2 * z2 * z2 * z2);
      *hvel = (1.0 - botfn) * h_arith + botfn * h_harm;
    }
  } else if (hvel_mom6) {
    double h_harm = 2.0 * hl_c * hr_c / (hl_c + hr_c + EPS_HVEL);
    double h_arith = 0.5 * (hl_c + hr_c);
    double h_delta = hr_c - hl_c;
    *hvel = h_harm;
    if (hvel_upwind && u_k * h_delta < 0.0) {
      double z2 = *zacc;
      double botfn = 1.0 / (1.0 + 0.09 * z2 * z2 * z2 * z2 * z2 * z2);
      *hvel = (1.0 - botfn) * h_harm + botfn * h_arith;
    }
    if (k >= kb) *zacc = *zacc + h_harm * i_hbbl;
    *zint = *zacc;
  } else {
    *hvel = 0.5 * (hl_c + hr_c);
  }
  if (zlevel_faces) *hvel = mfmin(hl_c, hr_c);
}

/* vdiff_interface_coupling (rdb_ocean_vdiff.F90:2067-2115) */
static void minterface_coupling(double hv_lo, double hv_hi, double zint_lo, double dt, int hvel_mom6,
                                int use_harmonic, int bbl_glue, double kv_bbl, double kv_bbl_bg, double bbl_thick,
                                int zlevel_faces, int do_corner, double kv_cnr, double kv_l, double kv_r,
                                double *num, double *dz, int *cut) {
  const double HV = ORC_H_VANISHED;
  double hf_lo = mfmax(hv_lo, HV);
  double hf_hi = mfmax(hv_hi, HV);
  if (hvel_mom6)
    *dz = 0.5 * (hf_lo + hf_hi);
  else
    *dz = mface_thick(hf_lo, hf_hi, use_harmonic);
  double nu_f = 0.5 * (kv_l + kv_r);
  if (do_corner) nu_f = nu_f + kv_cnr;
  if (bbl_glue) {
    double botfn_int = 1.0 / (1.0 + 0.09 * zint_lo * zint_lo * zint_lo * zint_lo * zint_lo * zint_lo);
    nu_f = mfmax(0.0, nu_f + (kv_bbl - kv_bbl_bg) * botfn_int);
    if (*dz > bbl_thick) *dz = (1.0 - botfn_int) * *dz + botfn_int * bbl_thick;
  }
  *num = dt * nu_f;
  *cut = zlevel_faces && (hv_lo <= HV || hv_hi <= HV);
}

void orc_diffuse_velocity_columns(const orc_mom_opts *o, int nu, int nv, int nz, double dt, double *u_face,
                                  const double *h_layer, const double *kv_centre, const double *wet_cell,
                                  int x_face, int nx_cells, int ny_cells, const double *tau_face,
                                  const double *lambda_bot, const double *lambda_top, const double *cover_face,
                                  double *visc_rem_out, const double *kv_bbl_face, const double *bbl_thick_face,
                                  const double *kv_corner, const int *k_top_face, const int *k_bot_face,
                                  double *c_diag, double *rhs) {
  const double HV = ORC_H_VANISHED;
  const double EPS_HVEL = 1.0e-30, STRESS_H_MIN = 1.0e-3;
  const double inv_rho0 = 1.0 / o->rho0;
  const size_t nf = (size_t)nu * (size_t)nv;
  const size_t nc = (size_t)nx_cells * (size_t)ny_cells;
  const size_t ncr = (size_t)(nx_cells + 1) * (size_t)(ny_cells + 1);
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < nv; ++j)
    for (int i = 0; i < nu; ++i) {
      const size_t f = (size_t)i + (size_t)nu * (size_t)j;
      int i_left, i_right, j_below, j_above, i_c2, j_c2;
      if (x_face) {
        i_left = (i - 1 > 0) ? i - 1 : 0;
        i_right = (i < nx_cells - 1) ? i : nx_cells - 1;
        j_below = j;
        j_above = j;
        i_c2 = i;
        j_c2 = j + 1;
      } else {
        i_left = i;
        i_right = i;
        j_below = (j - 1 > 0) ? j - 1 : 0;
        j_above = (j < ny_cells - 1) ? j : ny_cells - 1;
        i_c2 = i + 1;
        j_c2 = j;
      }
      const size_t cl = (size_t)i_left + (size_t)nx_cells * (size_t)j_below;
      const size_t cr = (size_t)i_right + (size_t)nx_cells * (size_t)j_above;
      const size_t cn1 = (size_t)i + (size_t)(nx_cells + 1) * (size_t)j;
      const size_t cn2 = (size_t)i_c2 + (size_t)(nx_cells + 1) * (size_t)j_c2;
      const int kb = k_bot_face[f];
      const int ktop = k_top_face[f];
      double kv_bbl = 0.0, bbl_thick = o->hbbl_visc;
      if (o->bbl_glue) {
        kv_bbl = kv_bbl_face[f];
        bbl_thick = bbl_thick_face[f];
      }
      const double i_hbbl = 1.0 / (bbl_thick + EPS_HVEL);
      double zacc = 0.0, zcol_l = 0.0, zcol_r = 0.0;
      if (o->hvel_mom6 && !o->hvel_harmonic) {
        for (int k = 0; k < nz; ++k) {
          zcol_l = zcol_l - h_layer[cl + nc * k];
          zcol_r = zcol_r - h_layer[cr + nc * k];
        }
      }
      const double d_min = mfmin(-zcol_l, -zcol_r);
      int kb_live = kb;
      for (int k = kb; k < nz; ++k) {
        int live = o->zlevel_faces ? (mfmin(h_layer[cl + nc * k], h_layer[cr + nc * k]) > HV) : 1;
        if (live) {
          kb_live = k;
          break;
        }
      }
      double hv_0, hv_p, zi_0, zi_p;
      mface_layer(h_layer[cl], h_layer[cr], u_face[f], 0, kb, o->hvel_mom6, o->hvel_harmonic, o->hvel_upwind,
                  o->zlevel_faces, d_min, i_hbbl, &zacc, &zcol_l, &zcol_r, &hv_p, &zi_p);
      hv_0 = 0.0;
      zi_0 = 0.0;
      double num_t = 0.0, dz_t = 1.0, num_b, dz_b;
      int cut_t = 0, cut_b;
      double cp_prev = 0.0, dp_prev = 0.0, gp_prev = 0.0, cp = 0.0, dp = 0.0, gp = 0.0;
      for (int k = 0; k < nz; ++k) {
        hv_0 = hv_p;
        zi_0 = zi_p;
        num_b = num_t;
        dz_b = dz_t;
        cut_b = cut_t;
        if (k < nz - 1) {
          mface_layer(h_layer[cl + nc * (k + 1)], h_layer[cr + nc * (k + 1)], u_face[f + nf * (k + 1)], k + 1, kb,
                      o->hvel_mom6, o->hvel_harmonic, o->hvel_upwind, o->zlevel_faces, d_min, i_hbbl, &zacc,
                      &zcol_l, &zcol_r, &hv_p, &zi_p);
          if (k >= kb_live && k <= ktop) {
            double kv_cnr = 0.0;
            if (o->do_corner)
              kv_cnr = o->kv_prandtl * (0.5 * (kv_corner[cn1 + ncr * (k + 1)] + kv_corner[cn2 + ncr * (k + 1)]));
            minterface_coupling(hv_0, hv_p, zi_0, dt, o->hvel_mom6, o->use_harmonic, o->bbl_glue, kv_bbl,
                                o->kv_bbl_bg, bbl_thick, o->zlevel_faces, o->do_corner, kv_cnr,
                                kv_centre[cl + nc * (k + 1)], kv_centre[cr + nc * (k + 1)], &num_t, &dz_t, &cut_t);
          }
        }
        double a_k = 0.0, b_k = 1.0, c_k = 0.0, alpha, beta;
        const double hf_k = mfmax(hv_0, HV);
        if (k == kb_live && k <= ktop) {
          alpha = 0.0;
          if (kb_live < nz - 1) {
            alpha = num_t / (hf_k * dz_t);
            if (cut_t) alpha = 0.0;
          }
          c_k = -alpha;
          b_k = 1.0 + alpha;
          if (o->bbl_glue) b_k = b_k + dt * kv_bbl / (hf_k * (mfmin(0.5 * hv_0, bbl_thick) + EPS_HVEL));
          if (o->do_drag && !o->bbl_glue) b_k = b_k + dt * lambda_bot[f];
        }
        if (k == ktop && k > kb_live) {
          beta = num_b / (hf_k * dz_b);
          if (cut_b) beta = 0.0;
          a_k = -beta;
          b_k = 1.0 + beta;
        }
        if (k > kb_live && k < ktop) {
          beta = num_b / (hf_k * dz_b);
          alpha = num_t / (hf_k * dz_t);
          if (cut_b) beta = 0.0;
          if (cut_t) alpha = 0.0;
          a_k = -beta;
          c_k = -alpha;
          b_k = 1.0 + alpha + beta;
        }
        double r_k = u_face[f + nf * k];
        if (k == ktop) {
          if (o->do_top) b_k = b_k + dt * lambda_top[f];
          if (o->do_stress) {
            double wind_open = 1.0;
            if (o->do_top) wind_open = 1.0 - cover_face[f];
            r_k = r_k + dt * tau_face[f] * wind_open * mfmin(wet_cell[cl], wet_cell[cr]) * inv_rho0 /
                            mfmax(hf_k, STRESS_H_MIN);
          }
        }
        if (k == 0) {
          cp = c_k / b_k;
          dp = r_k / b_k;
          gp = 1.0 / b_k;
        }
        if (k > 0) {
          double denom = b_k - a_k * cp_prev;
          cp = c_k / denom;
          dp = (r_k - a_k * dp_prev) / denom;
          gp = (1.0 - a_k * gp_prev) / denom;
        }
        c_diag[f + nf * k] = cp;
        if (o->solve_momentum) rhs[f + nf * k] = dp;
        if (o->do_remnant) visc_rem_out[f + nf * k] = gp;
        cp_prev = cp;
        dp_prev = dp;
        gp_prev = gp;
      }
      double u_next = dp_prev;
      double g_next = mfmin(gp_prev, 1.0);
      if (o->solve_momentum) u_face[f + nf * (nz - 1)] = u_next;
      if (o->do_remnant) visc_rem_out[f + nf * (nz - 1)] = g_next;
      for (int k = nz - 2; k >= 0; --k) {
        double cpk = c_diag[f + nf * k];
        if (o->solve_momentum) {
          u_next = rhs[f + nf * k] - cpk * u_next;
          u_face[f + nf * k] = u_next;
        }
        if (o->do_remnant) {
          g_next = mfmin(visc_rem_out[f + nf * k] - cpk * g_next, 1.0);
          visc_rem_out[f + nf * k] = g_next;
        }
      }
    }
}
