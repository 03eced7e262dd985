/*
 * remap_oracle.c -- CPU FP64 restatement of the reference's conservative
 * vertical (ALE) column remap and the field passes around it.  TEST
 * INFRASTRUCTURE ONLY: the checker of the HIP remap kernels.  Nothing in the
 * product path links or calls it.
 *
 * Follows src/ALE/rdb_remap_column.F90:
 *   boundary_half_jump             :83-118
 *   remap_column_preconditions_ok  :120-166
 *   plm_slope_nonuniform           :168-225
 *   ppm_edge_two_cell              :227-249
 *   ppm_jump_nonuniform            :251-285
 *   ppm_edge_nonuniform            :287-326
 *   remap_column (dispatch)        :328-375
 *   remap_column_pcm               :377-430
 *   remap_column_plm               :432-549
 *   remap_column_ppm               :551-778
 *   remap_column_ppm_h4            :780-1058
 *   pqm_solve_diag_dominant        :1060-1097
 *   pqm_end_value_h4               :1099-1170
 *   remap_column_pqm               :1172-1600
 * and src/ALE/rdb_ocean_remap.F90:
 *   ocean_remap_tracer_field       :236-319
 *   remap_bound_column             :321-452
 *   remap_add_filler_excess        :454-489
 *   remap_x/y_face_velocity        :967-1129
 *   rescale_anomaly_ke             :1131-1180
 *   ocean_remap_scan_preconditions :1348-1405
 *   ocean_remap_tracer_column      :1425-1457
 * with rdb_vl_column_conc (src/shared_module_utilities/
 * rdb_vanished_layer.inc:105-153), operation for operation in the
 * reference's left-to-right order (built without FMA contraction).
 *
 * Layout: Fortran column-major, 0-based: IDX(i, j, k) = i + nx*(j + ny*k),
 * k = 0 is the bed layer.  Column arrays are 0-based, z[0] = 0 at the bed.
 *
 * Pinning (tests/test_oracle_remap.py): the reference's remap tests
 * restated, its PPM_H4 golden vector, and the polynomial exactness,
 * conservation, monotonicity and range properties the methods guarantee.
 * Parity unpinned: two branches of the PQM limiter, the van Leer edge
 * replacement of bound_edge_values (:1362-1367) and the edge average of
 * check_discontinuous_edge_values (:1378-1382); replacing either with a
 * plain clamp keeps every property the tests can state, so only the
 * restatement's reading of the reference vouches for them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "colphys_oracle.h"

#define RH_NEGLECT 1.0e-30
#define RH_MIN_FRAC 1.0e-5
#define RH_RATIO_FLOOR 1.0e-12
#define RPQM_MIN_FRAC 1.0e-6

static double rmax(double a, double b) { return (a > b) ? a : b; }
static double rmin(double a, double b) { return (a < b) ? a : b; }
/* Fortran SIGN(a, b): |a| carrying the sign of b. */
static double rsign(double a, double b) { return copysign(fabs(a), b); }

/* boundary_half_jump (rdb_remap_column.F90:83-118) */
static double half_jump(double h_self, double h_nbr, double dq_up) {
  double d = dq_up * h_self / rmax(h_self + h_nbr, RH_NEGLECT);
  return rsign(rmin(fabs(d), fabs(dq_up)), d);
}

int orc_remap_preconditions_ok(int nz, const double *dz_old, const double *dz_new, double rel_tol) {
  int ok = 1;
  double s_old = 0.0, s_new = 0.0;
  for (int k = 0; k < nz; ++k) {
    if (dz_old[k] < 0.0) ok = 0;
    if (dz_new[k] < 0.0) ok = 0;
    s_old = s_old + dz_old[k];
    s_new = s_new + dz_new[k];
  }
  if (fabs(s_new - s_old) > rel_tol * rmax(fabs(s_old), fabs(s_new))) ok = 0;
  return ok;
}

/* plm_slope_nonuniform (:168-225) */
static double plm_slope_nu(double h_l, double h_c, double h_r, double q_l, double q_c, double q_r) {
  double sig_l = q_c - q_l;
  double sig_r = q_r - q_c;
  double sig_c = (h_c / (h_l + h_c + h_r + ORC_H_DIV_EPS)) *
                 ((2.0 * h_l + h_c) / (h_c + h_r + ORC_H_DIV_EPS) * sig_r +
                  (h_c + 2.0 * h_r) / (h_l + h_c + ORC_H_DIV_EPS) * sig_l);
  if (sig_l * sig_r > 0.0) {
    double q_min = rmin(rmin(q_l, q_c), q_r);
    double q_max = rmax(rmax(q_l, q_c), q_r);
    return 0.5 * rsign(rmin(fabs(sig_c), 2.0 * rmin(q_c - q_min, q_max - q_c)), sig_c);
  }
  return 0.0;
}

/* ppm_edge_two_cell (:227-249) */
static double edge_two_cell(double h_l, double h_r, double q_l, double q_r) {
  return (q_l * h_r + q_r * h_l) / (h_l + h_r + ORC_H_DIV_EPS);
}

/* ppm_jump_nonuniform (:251-285) */
static double jump_nu(double h_l, double h_c, double h_r, double q_l, double q_c, double q_r) {
  return (h_c / (h_l + h_c + h_r + ORC_H_DIV_EPS)) *
         ((2.0 * h_l + h_c) / (h_c + h_r + ORC_H_DIV_EPS) * (q_r - q_c) +
          (h_c + 2.0 * h_r) / (h_l + h_c + ORC_H_DIV_EPS) * (q_c - q_l));
}

/* ppm_edge_nonuniform (:287-326) */
static double edge_nu(double h0, double h1, double h2, double h3, double q1, double q2, double dq1,
                      double dq2) {
  double dq21 = q2 - q1;
  double i_h12 = 1.0 / (h1 + h2 + ORC_H_DIV_EPS);
  double wa = (h0 + h1) / (2.0 * h1 + h2 + ORC_H_DIV_EPS);
  double wb = (h2 + h3) / (h1 + 2.0 * h2 + ORC_H_DIV_EPS);
  return q1 + h1 * i_h12 * dq21 +
         (2.0 * h1 * h2 * i_h12 * (wa - wb) * dq21 - h1 * wa * dq2 + h2 * wb * dq1) /
             (h0 + h1 + h2 + h3 + ORC_H_DIV_EPS);
}

static void interfaces(int nz, const double *dz_old, const double *dz_new, double *z_old, double *z_new) {
  z_old[0] = 0.0;
  z_new[0] = 0.0;
  for (int k = 1; k <= nz; ++k) {
    z_old[k] = z_old[k - 1] + dz_old[k - 1];
    z_new[k] = z_new[k - 1] + dz_new[k - 1];
  }
}

/* remap_column_pcm (:377-430) */
static void remap_pcm(int nz, const double *dz_old, const double *dz_new, const double *q_old, double *q_new) {
  double z_old[nz + 1], z_new[nz + 1];
  interfaces(nz, dz_old, dz_new, z_old, z_new);
  int ko_start = 1;
  for (int k = 1; k <= nz; ++k) {
    if (dz_new[k - 1] <= 0.0) {
      q_new[k - 1] = 0.0;
      continue;
    }
    double integral = 0.0;
    for (int ko = ko_start; ko <= nz; ++ko) {
      double z_lo = rmax(z_new[k - 1], z_old[ko - 1]);
      double z_hi = rmin(z_new[k], z_old[ko]);
      double overlap = z_hi - z_lo;
      if (overlap <= 0.0) {
        if (z_old[ko] > z_new[k]) break;
        continue;
      }
      integral = integral + q_old[ko - 1] * overlap;
      if (z_old[ko] <= z_new[k]) ko_start = ko;
    }
    q_new[k - 1] = integral / dz_new[k - 1];
  }
}

/* remap_column_plm (:432-549) */
static void remap_plm(int nz, const double *dz_old, const double *dz_new, const double *q_old, double *q_new,
                      int be, int nu) {
  if (nz == 1) {
    q_new[0] = q_old[0];
    return;
  }
  double z_old[nz + 1], z_new[nz + 1], slope[nz];
  interfaces(nz, dz_old, dz_new, z_old, z_new);
  slope[0] = 0.0;
  if (nu) {
    for (int k = 2; k <= nz - 1; ++k)
      slope[k - 1] = plm_slope_nu(dz_old[k - 2], dz_old[k - 1], dz_old[k], q_old[k - 2], q_old[k - 1], q_old[k]);
  } else {
    for (int k = 2; k <= nz - 1; ++k) {
      double dq_l = q_old[k - 1] - q_old[k - 2];
      double dq_r = q_old[k] - q_old[k - 1];
      if (dq_l * dq_r > 0.0)
        slope[k - 1] = 0.5 * rsign(rmin(fabs(dq_l), fabs(dq_r)), dq_l);
      else
        slope[k - 1] = 0.0;
    }
  }
  slope[nz - 1] = 0.0;
  if (be) {
    slope[0] = half_jump(dz_old[0], dz_old[1], q_old[1] - q_old[0]);
    slope[nz - 1] = half_jump(dz_old[nz - 1], dz_old[nz - 2], q_old[nz - 1] - q_old[nz - 2]);
  }
  int ko_start = 1;
  for (int k = 1; k <= nz; ++k) {
    if (dz_new[k - 1] <= 0.0) {
      q_new[k - 1] = 0.0;
      continue;
    }
    double integral = 0.0;
    for (int ko = ko_start; ko <= nz; ++ko) {
      double z_lo = rmax(z_new[k - 1], z_old[ko - 1]);
      double z_hi = rmin(z_new[k], z_old[ko]);
      double overlap = z_hi - z_lo;
      if (overlap <= 0.0) {
        if (z_old[ko] > z_new[k]) break;
        continue;
      }
      if (dz_old[ko - 1] > 0.0) {
        double xi_lo = (z_lo - z_old[ko - 1]) / dz_old[ko - 1];
        double xi_hi = (z_hi - z_old[ko - 1]) / dz_old[ko - 1];
        integral = integral + overlap * (q_old[ko - 1] + slope[ko - 1] * (xi_lo + xi_hi - 1.0));
      } else {
        integral = integral + q_old[ko - 1] * overlap;
      }
      if (z_old[ko] <= z_new[k]) ko_start = ko;
    }
    q_new[k - 1] = integral / dz_new[k - 1];
  }
}

/* Steps 2-3 shared verbatim by remap_column_ppm (:680-777) and
 * remap_column_ppm_h4 (:957-1057): the Colella-Woodward limiter, the
 * optional boundary closure and the parabolic overlap integral.  q_L/q_R are
 * the raw edges (1-based cell index k at [k-1]). */
static void ppm_limit_integrate(int nz, const double *dz_old, const double *dz_new, const double *q_old,
                                double *q_new, double *q_L, double *q_R, int be) {
  double z_old[nz + 1], z_new[nz + 1], q6[nz];
  interfaces(nz, dz_old, dz_new, z_old, z_new);
  for (int k = 1; k <= nz; ++k) {
    double q = q_old[k - 1];
    double q_min = q, q_max = q;
    if (k > 1) {
      q_min = rmin(q_min, q_old[k - 2]);
      q_max = rmax(q_max, q_old[k - 2]);
    }
    if (k < nz) {
      q_min = rmin(q_min, q_old[k]);
      q_max = rmax(q_max, q_old[k]);
    }
    q_L[k - 1] = rmax(q_min, rmin(q_max, q_L[k - 1]));
    q_R[k - 1] = rmax(q_min, rmin(q_max, q_R[k - 1]));
    double dq = q_R[k - 1] - q_L[k - 1];
    double dq_l = q - q_L[k - 1];
    double dq_r = q_R[k - 1] - q;
    if (dq_l * dq_r <= 0.0) {
      q_L[k - 1] = q;
      q_R[k - 1] = q;
    } else {
      q6[k - 1] = 6.0 * q - 3.0 * (q_L[k - 1] + q_R[k - 1]);
      if (fabs(q6[k - 1]) > fabs(dq)) {
        if (q6[k - 1] * dq > 0.0)
          q_L[k - 1] = 3.0 * q - 2.0 * q_R[k - 1];
        else
          q_R[k - 1] = 3.0 * q - 2.0 * q_L[k - 1];
      }
    }
    q6[k - 1] = 6.0 * q - 3.0 * (q_L[k - 1] + q_R[k - 1]);
  }
  if (be) {
    double d = half_jump(dz_old[0], dz_old[1], q_old[1] - q_old[0]);
    q_L[0] = q_old[0] - d;
    q_R[0] = q_old[0] + d;
    q6[0] = 0.0;
    d = half_jump(dz_old[nz - 1], dz_old[nz - 2], q_old[nz - 1] - q_old[nz - 2]);
    q_L[nz - 1] = q_old[nz - 1] - d;
    q_R[nz - 1] = q_old[nz - 1] + d;
    q6[nz - 1] = 0.0;
  }
  int ko_start = 1;
  for (int k = 1; k <= nz; ++k) {
    if (dz_new[k - 1] <= 0.0) {
      q_new[k - 1] = 0.0;
      continue;
    }
    double integral = 0.0;
    for (int ko = ko_start; ko <= nz; ++ko) {
      double z_lo = rmax(z_new[k - 1], z_old[ko - 1]);
      double z_hi = rmin(z_new[k], z_old[ko]);
      double overlap = z_hi - z_lo;
      if (overlap <= 0.0) {
        if (z_old[ko] > z_new[k]) break;
        continue;
      }
      if (dz_old[ko - 1] > 0.0) {
        double xi_lo = (z_lo - z_old[ko - 1]) / dz_old[ko - 1];
        double xi_hi = (z_hi - z_old[ko - 1]) / dz_old[ko - 1];
        integral = integral + dz_old[ko - 1] * ((xi_hi - xi_lo) * q_L[ko - 1] +
                                                0.5 * (xi_hi * xi_hi - xi_lo * xi_lo) *
                                                    (q_R[ko - 1] - q_L[ko - 1] + q6[ko - 1]) -
                                                (xi_hi * xi_hi * xi_hi - xi_lo * xi_lo * xi_lo) * q6[ko - 1] / 3.0);
      } else {
        integral = integral + q_old[ko - 1] * overlap;
      }
      if (z_old[ko] <= z_new[k]) ko_start = ko;
    }
    q_new[k - 1] = integral / dz_new[k - 1];
  }
}

/* remap_column_ppm (:551-778) */
static void remap_ppm(int nz, const double *dz_old, const double *dz_new, const double *q_old, double *q_new,
                      int be, int nu) {
  if (nz == 1) {
    q_new[0] = q_old[0];
    return;
  }
  if (nz == 2) {
    remap_plm(nz, dz_old, dz_new, q_old, q_new, be, nu);
    return;
  }
  double q_L[nz], q_R[nz], dq_cw[nz];
  q_L[0] = q_old[0];
  q_R[nz - 1] = q_old[nz - 1];
  if (nu) {
    for (int k = 2; k <= nz - 1; ++k)
      dq_cw[k - 1] = jump_nu(dz_old[k - 2], dz_old[k - 1], dz_old[k], q_old[k - 2], q_old[k - 1], q_old[k]);
    double edge = edge_two_cell(dz_old[0], dz_old[1], q_old[0], q_old[1]);
    q_R[0] = edge;
    q_L[1] = edge;
    edge = edge_two_cell(dz_old[nz - 2], dz_old[nz - 1], q_old[nz - 2], q_old[nz - 1]);
    q_R[nz - 2] = edge;
    q_L[nz - 1] = edge;
    for (int k = 2; k <= nz - 2; ++k) {
      edge = edge_nu(dz_old[k - 2], dz_old[k - 1], dz_old[k], dz_old[k + 1], q_old[k - 1], q_old[k], dq_cw[k - 1],
                     dq_cw[k]);
      q_R[k - 1] = edge;
      q_L[k] = edge;
    }
  } else {
    q_R[0] = 0.5 * (q_old[0] + q_old[1]);
    q_L[nz - 1] = 0.5 * (q_old[nz - 2] + q_old[nz - 1]);
    for (int k = 2; k <= nz - 1; ++k) {
      double edge = 0.5 * (q_old[k - 1] + q_old[k]);
      if (k + 2 <= nz)
        edge = (7.0 / 12.0) * (q_old[k - 1] + q_old[k]) - (1.0 / 12.0) * (q_old[k - 2] + q_old[k + 1]);
      q_R[k - 1] = edge;
      q_L[k] = edge;
    }
    q_L[1] = q_R[0];
  }
  ppm_limit_integrate(nz, dz_old, dz_new, q_old, q_new, q_L, q_R, be);
}

/* The non-uniform three-cell (H3) quadratic of remap_column_ppm_h4
 * (:851-889 and :893-927): cell-average fit through cells a, b, c (bed
 * side first) evaluated at z_eval_sel = 1 (the a|b interface, z1) or 2 (the
 * b|c interface, z2). */
static double h3_edge(double h1, double h2, double h3, double qa, double qb, double qc, int at_z2) {
  if (h1 <= 0.0 || h2 <= 0.0 || h3 <= 0.0) {
    double hf = RH_MIN_FRAC * rmax(RH_NEGLECT, h1 + h2 + h3);
    h1 = rmax(h1, hf);
    h2 = rmax(h2, hf);
    h3 = rmax(h3, hf);
  }
  double z1 = h1, z2 = h1 + h2, z3 = h1 + h2 + h3;
  double m11 = (z1 * z1 * z1) / (3.0 * h1);
  double m12 = (z1 * z1) / (2.0 * h1);
  double m13 = 1.0;
  double m21 = (z2 * z2 * z2 - z1 * z1 * z1) / (3.0 * h2);
  double m22 = (z2 * z2 - z1 * z1) / (2.0 * h2);
  double m23 = 1.0;
  double m31 = (z3 * z3 * z3 - z2 * z2 * z2) / (3.0 * h3);
  double m32 = (z3 * z3 - z2 * z2) / (2.0 * h3);
  double m33 = 1.0;
  double det = m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31);
  double ca = (qa * (m22 * m33 - m23 * m32) - m12 * (qb * m33 - m23 * qc) + m13 * (qb * m32 - m22 * qc)) / det;
  double cb = (m11 * (qb * m33 - m23 * qc) - qa * (m21 * m33 - m23 * m31) + m13 * (m21 * qc - qb * m31)) / det;
  double cc = (m11 * (m22 * qc - qb * m32) - m12 * (m21 * qc - qb * m31) + qa * (m21 * m32 - m22 * m31)) / det;
  double z = at_z2 ? z2 : z1;
  return ca * z * z + cb * z + cc;
}

/* The interior H4 edge of remap_column_ppm_h4 (:932-955) between cells k
 * and k+1 from the stencil thicknesses h0..h3 and means q(k-1..k+2). */
static double h4_edge(double h0, double h1, double h2, double h3, double qm1, double q0, double qp1, double qp2) {
  double h_sum = h0 + h1 + h2 + h3;
  if (h0 + h1 <= 0.0 || h1 + h2 <= 0.0 || h2 + h3 <= 0.0) {
    double hf = RH_MIN_FRAC * rmax(RH_NEGLECT, h_sum);
    h0 = rmax(h0, hf);
    h1 = rmax(h1, hf);
    h2 = rmax(h2, hf);
    h3 = rmax(h3, hf);
  }
  double h01 = h0 + h1, h12 = h1 + h2, h23 = h2 + h3;
  double h012 = h0 + h1 + h2, h123 = h1 + h2 + h3, h0123 = h0 + h1 + h2 + h3;
  double f1 = h01 * h23 / h12;
  double f2 = h2 * q0 + h1 * qp1;
  double f3 = 1.0 / h012 + 1.0 / h123;
  double et1 = f1 * f2 * f3;
  double et2 = (h2 * h23 / (h012 * h01)) * ((h0 + 2.0 * h1) * q0 - h1 * qm1);
  double et3 = (h1 * h01 / (h123 * h23)) * ((2.0 * h2 + h3) * qp1 - h2 * qp2);
  return (et1 + et2 + et3) / h0123;
}

/* remap_column_ppm_h4 (:780-1058) */
static void remap_ppm_h4(int nz, const double *dz_old, const double *dz_new, const double *q_old, double *q_new,
                         int be, int nu) {
  if (nz == 1) {
    q_new[0] = q_old[0];
    return;
  }
  if (nz == 2) {
    remap_plm(nz, dz_old, dz_new, q_old, q_new, be, nu);
    return;
  }
  double q_L[nz], q_R[nz];
  q_L[0] = q_old[0];
  q_R[nz - 1] = q_old[nz - 1];
  double e = h3_edge(dz_old[0], dz_old[1], dz_old[2], q_old[0], q_old[1], q_old[2], 0);
  q_R[0] = e;
  q_L[1] = e;
  e = h3_edge(dz_old[nz - 3], dz_old[nz - 2], dz_old[nz - 1], q_old[nz - 3], q_old[nz - 2], q_old[nz - 1], 1);
  q_R[nz - 2] = e;
  q_L[nz - 1] = e;
  for (int k = 2; k <= nz - 2; ++k) {
    e = h4_edge(dz_old[k - 2], dz_old[k - 1], dz_old[k], dz_old[k + 1], q_old[k - 2], q_old[k - 1], q_old[k],
                q_old[k + 1]);
    q_R[k - 1] = e;
    q_L[k] = e;
  }
  ppm_limit_integrate(nz, dz_old, dz_new, q_old, q_new, q_L, q_R, be);
}

/* pqm_solve_diag_dominant (:1060-1097); arrays 0-based over n unknowns. */
static void pqm_solve(int n, const double *al, const double *ac, const double *au, const double *r, double *x) {
  double c1[n];
  double i_pivot = 1.0 / (ac[0] + au[0]);
  double d1 = ac[0] * i_pivot;
  c1[0] = au[0] * i_pivot;
  x[0] = r[0] * i_pivot;
  for (int k = 1; k < n - 1; ++k) {
    double denom_t1 = ac[k] + d1 * al[k];
    i_pivot = 1.0 / (denom_t1 + au[k]);
    d1 = denom_t1 * i_pivot;
    c1[k] = au[k] * i_pivot;
    x[k] = (r[k] - al[k] * x[k - 1]) * i_pivot;
  }
  i_pivot = 1.0 / (ac[n - 1] + d1 * al[n - 1]);
  x[n - 1] = (r[n - 1] - al[n - 1] * x[n - 2]) * i_pivot;
  for (int k = n - 2; k >= 0; --k) x[k] = x[k] - c1[k] * x[k + 1];
}

/* pqm_end_value_h4 (:1099-1170) */
void orc_pqm_end_value_h4(const double *dz, const double *u, double *csys) {
  double h1 = dz[0], h2 = dz[1], h3 = dz[2], h4 = dz[3];
  if ((h2 + h3) < RPQM_MIN_FRAC * h1) h3 = RPQM_MIN_FRAC * h1 - h2;
  if ((h3 + h4) < RPQM_MIN_FRAC * h1) h4 = RPQM_MIN_FRAC * h1 - h3;
  double h12 = h1 + h2, h23 = h2 + h3, h34 = h3 + h4;
  double h123 = h12 + h3, h234 = h2 + h34, h1234 = h12 + h34;
  double i_denb3 = 1.0 / (h123 * h12 * h23);
  double i_h12 = (h123 * h23) * i_denb3;
  double i_h23 = (h12 * h123) * i_denb3;
  double i_h123 = (h12 * h23) * i_denb3;
  double i_denom = 1.0 / (h1234 * (h234 * h34));
  double i_h34 = (h1234 * h234) * i_denom;
  double i_h234 = (h1234 * h34) * i_denom;
  double i_h1234 = (h234 * h34) * i_denom;
  (void)i_h34;
  double w11 = -h1 * (i_h1234 + i_h123 + i_h12);
  double w21 = h1 * h12 * (i_h234 * i_h1234 + i_h23 * (i_h234 + i_h123));
  double w31 = -h1 * h12 * h123 * i_denom;
  double w12 = 2.0 * (i_h12 * (1.0 + (h1 + h12) * (i_h1234 + i_h123)) + h1 * i_h1234 * i_h123);
  double w22 = -2.0 * ((h1 * h12 * i_h1234) * (i_h23 * (i_h234 + i_h123)) +
                       (h1 + h12) * (i_h1234 * i_h234 + i_h23 * (i_h234 + i_h123)));
  double w32 = 2.0 * ((h1 + h12) * h123 + h1 * h12) * i_denom;
  double w13 = -3.0 * i_h12 * i_h123 * (1.0 + i_h1234 * ((h1 + h12) + h123));
  double w23 = 3.0 * i_h23 * (i_h123 + i_h1234 * ((h1 + h12) + h123) * (i_h123 + i_h234));
  double w33 = -3.0 * ((h1 + h12) + h123) * i_denom;
  double w14 = 4.0 * i_h1234 * i_h123 * i_h12;
  double w24 = -4.0 * i_h1234 * (i_h23 * (i_h123 + i_h234));
  double w34 = 4.0 * i_denom;
  double du1 = u[1] - u[0], du2 = u[2] - u[1], du3 = u[3] - u[2];
  csys[0] = ((u[0] + (w11 * du1)) + (w21 * du2)) + (w31 * du3);
  csys[1] = ((w12 * du1) + (w22 * du2)) + (w32 * du3);
  csys[2] = ((w13 * du1) + (w23 * du2)) + (w33 * du3);
  csys[3] = ((w14 * du1) + (w24 * du2)) + (w34 * du3);
}

/* The W&A inflexion test of one root x of the quartic's curvature
 * (remap_column_pqm :1433-1474): flags the side to collapse onto. */
static int pqm_root_flags(double x, double eco, double dco, double cco, double b, double slope, double sigma_l,
                          double sigma_r, int *infl_l, int *infl_r) {
  double grad = 4.0 * eco * (x * x * x) + 3.0 * dco * (x * x) + 2.0 * cco * x + b;
  if (grad * slope < 0.0) {
    if (fabs(sigma_l) < fabs(sigma_r))
      *infl_l = 1;
    else
      *infl_r = 1;
    return 1;
  }
  return 0;
}

/* remap_column_pqm (:1172-1600) */
static void remap_pqm(int nz, const double *dz_old, const double *dz_new, const double *q_old, double *q_new,
                      int be, int nu) {
  if (nz == 1) {
    q_new[0] = q_old[0];
    return;
  }
  if (nz < 5) {
    remap_ppm(nz, dz_old, dz_new, q_old, q_new, be, nu);
    return;
  }
  const int np1 = nz + 1;
  double z_old[nz + 1], z_new[nz + 1];
  double ev_l[nz], ev_r[nz], es_l[nz], es_r[nz];
  double pa[nz], pb[nz], pc[nz], pd[nz], pe[nz];
  double tri_l[np1], tri_c[np1], tri_u[np1], tri_b[np1], tri_x[np1];
  double dz4[4], u4[4], csys[4];
  interfaces(nz, dz_old, dz_new, z_old, z_new);
  memset(tri_l, 0, sizeof tri_l);

  /* Step 1: implicit-h4 edge values.  Row k+1 (1-based) sits at [k]. */
  for (int k = 1; k <= nz - 1; ++k) {
    double h0 = rmax(dz_old[k - 1], RH_NEGLECT);
    double h1 = rmax(dz_old[k], RH_NEGLECT);
    if (fabs(h0) < RH_RATIO_FLOOR * fabs(h1)) h0 = RH_RATIO_FLOOR * h1;
    if (fabs(h1) < RH_RATIO_FLOOR * fabs(h0)) h1 = RH_RATIO_FLOOR * h0;
    double i_h2 = 1.0 / ((h0 + h1) * (h0 + h1));
    double alpha = (h1 * h1) * i_h2;
    double beta = (h0 * h0) * i_h2;
    double abmix = (h0 * h1) * i_h2;
    double aco = 2.0 * alpha * (alpha + 2.0 * beta + 3.0 * abmix);
    double bco = 2.0 * beta * (beta + 2.0 * alpha + 3.0 * abmix);
    tri_l[k] = alpha;
    tri_c[k] = 2.0 * abmix;
    tri_u[k] = beta;
    tri_b[k] = aco * q_old[k - 1] + bco * q_old[k];
  }
  for (int k = 0; k < 4; ++k) {
    dz4[k] = rmax(RH_NEGLECT, dz_old[k]);
    u4[k] = q_old[k];
  }
  orc_pqm_end_value_h4(dz4, u4, csys);
  tri_b[0] = csys[0];
  tri_c[0] = 1.0;
  tri_u[0] = 0.0;
  for (int k = 0; k < 4; ++k) {
    dz4[k] = rmax(RH_NEGLECT, dz_old[nz - 1 - k]);
    u4[k] = q_old[nz - 1 - k];
  }
  orc_pqm_end_value_h4(dz4, u4, csys);
  tri_b[np1 - 1] = csys[0];
  tri_c[np1 - 1] = 1.0;
  tri_l[np1 - 1] = 0.0;
  pqm_solve(np1, tri_l, tri_c, tri_u, tri_b, tri_x);
  ev_l[0] = tri_x[0];
  for (int k = 2; k <= nz; ++k) {
    ev_l[k - 1] = tri_x[k - 1];
    ev_r[k - 2] = tri_x[k - 1];
  }
  ev_r[nz - 1] = tri_x[np1 - 1];

  /* Step 2: implicit-h3 edge slopes. */
  for (int k = 1; k <= nz - 1; ++k) {
    double h0 = rmax(dz_old[k - 1], RH_NEGLECT);
    double h1 = rmax(dz_old[k], RH_NEGLECT);
    double i_h = 1.0 / (h0 + h1);
    h0 = h0 * i_h;
    h1 = h1 * i_h;
    double h0h1 = h0 * h1, h0_2 = h0 * h0, h1_2 = h1 * h1;
    double h0_3 = h0_2 * h0, h1_3 = h1_2 * h1;
    double i_d = 1.0 / (4.0 * h0h1 * (h0 + h1) + h1_3 + h0_3);
    tri_l[k] = (h1 * ((h0_2 + h0h1) - h1_2)) * i_d;
    tri_c[k] = 2.0 * ((h0_2 + h1_2) * (h0 + h1)) * i_d;
    tri_u[k] = (h0 * ((h1_2 + h0h1) - h0_2)) * i_d;
    tri_b[k] = 12.0 * (h0h1 * i_d) * ((q_old[k] - q_old[k - 1]) * i_h);
  }
  for (int k = 0; k < 4; ++k) {
    dz4[k] = rmax(RH_NEGLECT, dz_old[k]);
    u4[k] = q_old[k];
  }
  orc_pqm_end_value_h4(dz4, u4, csys);
  tri_b[0] = csys[1];
  tri_c[0] = 1.0;
  tri_u[0] = 0.0;
  for (int k = 0; k < 4; ++k) {
    dz4[k] = rmax(RH_NEGLECT, dz_old[nz - 1 - k]);
    u4[k] = q_old[nz - 1 - k];
  }
  orc_pqm_end_value_h4(dz4, u4, csys);
  tri_b[np1 - 1] = -csys[1];
  tri_c[np1 - 1] = 1.0;
  tri_l[np1 - 1] = 0.0;
  pqm_solve(np1, tri_l, tri_c, tri_u, tri_b, tri_x);
  es_l[0] = tri_x[0];
  for (int k = 2; k <= nz; ++k) {
    es_l[k - 1] = tri_x[k - 1];
    es_r[k - 2] = tri_x[k - 1];
  }
  es_r[nz - 1] = tri_x[np1 - 1];

  /* Step 3a: bound_edge_values. */
  for (int k = 1; k <= nz; ++k) {
    int kl = k - 1 > 1 ? k - 1 : 1, kr = k + 1 < nz ? k + 1 : nz;
    double u_l = q_old[kl - 1], u_c = q_old[k - 1], u_r = q_old[kr - 1];
    double h_l = dz_old[kl - 1], h_c = dz_old[k - 1], h_r = dz_old[kr - 1];
    double slope_x_h = 0.0;
    if (((h_l + h_r) + 2.0 * h_c) > 0.0) {
      double sigma_l = (u_c - u_l);
      double sigma_c = (u_r - u_l) * (h_c / ((h_l + h_r) + 2.0 * h_c));
      double sigma_r = (u_r - u_c);
      if ((sigma_l * sigma_r) > 0.0)
        slope_x_h = rsign(rmin(rmin(fabs(sigma_l), fabs(sigma_c)), fabs(sigma_r)), sigma_c);
    }
    if ((u_l - ev_l[k - 1]) * (ev_l[k - 1] - u_c) < 0.0)
      ev_l[k - 1] = u_c - rsign(rmin(fabs(slope_x_h), fabs(ev_l[k - 1] - u_c)), slope_x_h);
    if ((u_r - ev_r[k - 1]) * (ev_r[k - 1] - u_c) < 0.0)
      ev_r[k - 1] = u_c + rsign(rmin(fabs(slope_x_h), fabs(ev_r[k - 1] - u_c)), slope_x_h);
    ev_l[k - 1] = rmax(rmin(ev_l[k - 1], rmax(u_l, u_c)), rmin(u_l, u_c));
    ev_r[k - 1] = rmax(rmin(ev_r[k - 1], rmax(u_r, u_c)), rmin(u_r, u_c));
  }
  /* Step 3b: check_discontinuous_edge_values. */
  for (int k = 1; k <= nz - 1; ++k) {
    if ((ev_l[k] - ev_r[k - 1]) * (q_old[k] - q_old[k - 1]) < 0.0) {
      double u0_avg = 0.5 * (ev_r[k - 1] + ev_l[k]);
      u0_avg = rmax(rmin(u0_avg, rmax(q_old[k - 1], q_old[k])), rmin(q_old[k - 1], q_old[k]));
      ev_r[k - 1] = u0_avg;
      ev_l[k] = u0_avg;
    }
  }
  /* Step 3c: interior cells. */
  for (int k = 2; k <= nz - 1; ++k) {
    int infl_l = 0, infl_r = 0;
    double u0_l = ev_l[k - 1], u0_r = ev_r[k - 1];
    double u1_l = es_l[k - 1], u1_r = es_r[k - 1];
    double h_l = dz_old[k - 2], h_c = dz_old[k - 1], h_r = dz_old[k];
    double u_l = q_old[k - 2], u_c = q_old[k - 1], u_r = q_old[k];
    double sigma_l = 2.0 * (u_c - u_l) / (h_c + RH_NEGLECT);
    double sigma_c = 2.0 * (u_r - u_l) / (h_l + 2.0 * h_c + h_r + RH_NEGLECT);
    double sigma_r = 2.0 * (u_r - u_c) / (h_c + RH_NEGLECT);
    double slope;
    if ((sigma_l * sigma_r) > 0.0)
      slope = rsign(rmin(rmin(fabs(sigma_l), fabs(sigma_c)), fabs(sigma_r)), sigma_c);
    else
      slope = 0.0;
    if (u1_l * slope <= 0.0) u1_l = slope;
    if (u1_r * slope <= 0.0) u1_r = slope;
    if ((u0_r - u_c) * (u_c - u0_l) <= 0.0) {
      u0_l = u_c;
      u0_r = u_c;
      u1_l = 0.0;
      u1_r = 0.0;
      infl_l = -1;
      infl_r = -1;
    }
    if (infl_l == 0 && infl_r == 0) {
      double b = h_c * u1_l;
      double cco = 30.0 * u_c - 12.0 * u0_r - 18.0 * u0_l + 1.5 * h_c * (u1_r - 3.0 * u1_l);
      double dco = -60.0 * u_c + h_c * (6.0 * u1_l - 4.0 * u1_r) + 28.0 * u0_r + 32.0 * u0_l;
      double eco = 30.0 * u_c + 2.5 * h_c * (u1_r - u1_l) - 15.0 * (u0_l + u0_r);
      double alpha1 = 6.0 * eco, alpha2 = 3.0 * dco, alpha3 = cco;
      double rho = alpha2 * alpha2 - 4.0 * alpha1 * alpha3;
      if ((alpha1 != 0.0) && (rho >= 0.0)) {
        double sqrt_rho = sqrt(rho);
        double x1 = 0.5 * (-alpha2 - sqrt_rho) / alpha1;
        double x2 = 0.5 * (-alpha2 + sqrt_rho) / alpha1;
        int in1 = (x1 >= 0.0) && (x1 <= 1.0), in2 = (x2 >= 0.0) && (x2 <= 1.0);
        if (in1 && in2) {
          double g1 = 4.0 * eco * (x1 * x1 * x1) + 3.0 * dco * (x1 * x1) + 2.0 * cco * x1 + b;
          double g2 = 4.0 * eco * (x2 * x2 * x2) + 3.0 * dco * (x2 * x2) + 2.0 * cco * x2 + b;
          if ((g1 * slope < 0.0) || (g2 * slope < 0.0)) {
            if (fabs(sigma_l) < fabs(sigma_r))
              infl_l = 1;
            else
              infl_r = 1;
          }
        } else if (in1) {
          pqm_root_flags(x1, eco, dco, cco, b, slope, sigma_l, sigma_r, &infl_l, &infl_r);
        } else if (in2) {
          pqm_root_flags(x2, eco, dco, cco, b, slope, sigma_l, sigma_r, &infl_l, &infl_r);
        }
      }
      if ((alpha1 == 0.0) && (alpha2 != 0.0)) {
        double x1 = -alpha3 / alpha2;
        if ((x1 >= 0.0) && (x1 <= 1.0)) pqm_root_flags(x1, eco, dco, cco, b, slope, sigma_l, sigma_r, &infl_l, &infl_r);
      }
    }
    if (infl_l == 1) {
      u1_l = (10.0 * u_c - 2.0 * u0_r - 8.0 * u0_l) / (3.0 * h_c + RH_NEGLECT);
      u1_r = (-10.0 * u_c + 6.0 * u0_r + 4.0 * u0_l) / (h_c + RH_NEGLECT);
      if (u1_l * slope < 0.0) {
        u1_l = 0.0;
        u0_r = 5.0 * u_c - 4.0 * u0_l;
        u1_r = 20.0 * (u_c - u0_l) / (h_c + RH_NEGLECT);
      } else if (u1_r * slope < 0.0) {
        u1_r = 0.0;
        u0_l = (5.0 * u_c - 3.0 * u0_r) / 2.0;
        u1_l = 10.0 * (-u_c + u0_r) / (3.0 * h_c + RH_NEGLECT);
      }
    } else if (infl_r == 1) {
      u1_r = (-10.0 * u_c + 8.0 * u0_r + 2.0 * u0_l) / (3.0 * h_c + RH_NEGLECT);
      u1_l = (10.0 * u_c - 4.0 * u0_r - 6.0 * u0_l) / (h_c + RH_NEGLECT);
      if (u1_l * slope < 0.0) {
        u1_l = 0.0;
        u0_r = (5.0 * u_c - 3.0 * u0_l) / 2.0;
        u1_r = 10.0 * (u_c - u0_l) / (3.0 * h_c + RH_NEGLECT);
      } else if (u1_r * slope < 0.0) {
        u1_r = 0.0;
        u0_l = 5.0 * u_c - 4.0 * u0_r;
        u1_l = 20.0 * (-u_c + u0_r) / (h_c + RH_NEGLECT);
      }
    }
    ev_l[k - 1] = u0_l;
    ev_r[k - 1] = u0_r;
    es_l[k - 1] = u1_l;
    es_r[k - 1] = u1_r;
  }
  ev_l[0] = q_old[0];
  ev_r[0] = q_old[0];
  es_l[0] = 0.0;
  es_r[0] = 0.0;
  ev_l[nz - 1] = q_old[nz - 1];
  ev_r[nz - 1] = q_old[nz - 1];
  es_l[nz - 1] = 0.0;
  es_r[nz - 1] = 0.0;
  if (be) {
    double d = half_jump(dz_old[0], dz_old[1], q_old[1] - q_old[0]);
    ev_l[0] = q_old[0] - d;
    ev_r[0] = q_old[0] + d;
    es_l[0] = 2.0 * d / rmax(dz_old[0], RH_NEGLECT);
    es_r[0] = es_l[0];
    d = half_jump(dz_old[nz - 1], dz_old[nz - 2], q_old[nz - 1] - q_old[nz - 2]);
    ev_l[nz - 1] = q_old[nz - 1] - d;
    ev_r[nz - 1] = q_old[nz - 1] + d;
    es_l[nz - 1] = 2.0 * d / rmax(dz_old[nz - 1], RH_NEGLECT);
    es_r[nz - 1] = es_l[nz - 1];
  }
  /* Step 4: per-cell quartic coefficients. */
  for (int k = 0; k < nz; ++k) {
    double h_c = dz_old[k], u0_l = ev_l[k], u0_r = ev_r[k], u1_l = es_l[k], u1_r = es_r[k], u_c = q_old[k];
    pa[k] = u0_l;
    pb[k] = h_c * u1_l;
    pc[k] = 30.0 * u_c - 12.0 * u0_r - 18.0 * u0_l + 1.5 * h_c * (u1_r - 3.0 * u1_l);
    pd[k] = -60.0 * u_c + h_c * (6.0 * u1_l - 4.0 * u1_r) + 28.0 * u0_r + 32.0 * u0_l;
    pe[k] = 30.0 * u_c + 2.5 * h_c * (u1_r - u1_l) - 15.0 * (u0_l + u0_r);
  }
  /* Step 5: conservative quartic overlap integration. */
  int ko_start = 1;
  for (int k = 1; k <= nz; ++k) {
    if (dz_new[k - 1] <= 0.0) {
      q_new[k - 1] = 0.0;
      continue;
    }
    double integral = 0.0;
    for (int ko = ko_start; ko <= nz; ++ko) {
      double z_lo = rmax(z_new[k - 1], z_old[ko - 1]);
      double z_hi = rmin(z_new[k], z_old[ko]);
      double overlap = z_hi - z_lo;
      if (overlap <= 0.0) {
        if (z_old[ko] > z_new[k]) break;
        continue;
      }
      if (dz_old[ko - 1] > 0.0) {
        double xl = (z_lo - z_old[ko - 1]) / dz_old[ko - 1];
        double xh = (z_hi - z_old[ko - 1]) / dz_old[ko - 1];
        int m = ko - 1;
        integral = integral + overlap * (pa[m] + pb[m] * 0.5 * (xl + xh) +
                                         pc[m] * (1.0 / 3.0) * (xl * xl + xh * xh + xl * xh) +
                                         pd[m] * 0.25 * ((xl * xl + xh * xh) * (xl + xh)) +
                                         pe[m] * 0.2 * ((xh * xh * xh + xl * xl * xl) * (xl + xh) + xl * xl * xh * xh));
      } else {
        integral = integral + q_old[ko - 1] * overlap;
      }
      if (z_old[ko] <= z_new[k]) ko_start = ko;
    }
    q_new[k - 1] = integral / dz_new[k - 1];
  }
}

/* remap_column (:328-375) */
void orc_remap_column(int method, int nz, const double *dz_old, const double *dz_new, const double *q_old,
                      double *q_new, int bnd_extrap, int nonunif) {
  switch (method) {
    case RDB_REMAP_PCM: remap_pcm(nz, dz_old, dz_new, q_old, q_new); break;
    case RDB_REMAP_PPM: remap_ppm(nz, dz_old, dz_new, q_old, q_new, bnd_extrap, nonunif); break;
    case RDB_REMAP_PPM_H4: remap_ppm_h4(nz, dz_old, dz_new, q_old, q_new, bnd_extrap, nonunif); break;
    case RDB_REMAP_PQM: remap_pqm(nz, dz_old, dz_new, q_old, q_new, bnd_extrap, nonunif); break;
    default: remap_plm(nz, dz_old, dz_new, q_old, q_new, bnd_extrap, nonunif); break;
  }
}

/* rdb_vl_column_conc (rdb_vanished_layer.inc:105-153) */
void orc_vl_column_conc(int nz, const double *h, const double *c_in, double *c_out) {
  int k_top = -1;
  for (int k = nz - 1; k >= 0; --k)
    if (h[k] > ORC_H_VANISHED) {
      k_top = k;
      break;
    }
  if (k_top < 0) {
    for (int k = 0; k < nz; ++k) c_out[k] = 0.0;
    return;
  }
  int k_don = k_top;
  for (int k = nz - 1; k >= 0; --k) {
    if (h[k] > ORC_H_VANISHED) {
      c_out[k] = c_in[k];
      k_don = k;
    } else {
      c_out[k] = c_in[k_don];
    }
  }
}

/* remap_add_filler_excess (rdb_ocean_remap.F90:454-489) */
static void add_filler_excess(int nz, const double *h_old, const double *h_new, const double *c_don,
                              const double *c_own, double *c_new) {
  double zo_lo = 0.0;
  for (int m = 0; m < nz; ++m) {
    double zo_hi = zo_lo + h_old[m];
    if (h_old[m] > 0.0 && h_old[m] <= ORC_H_VANISHED) {
      double s = c_own[m] - c_don[m];
      if (s != 0.0) {
        double zn_lo = 0.0;
        for (int k = 0; k < nz; ++k) {
          double zn_hi = zn_lo + h_new[k];
          double overlap = rmin(zn_hi, zo_hi) - rmax(zn_lo, zo_lo);
          if (overlap > 0.0 && h_new[k] > 0.0) c_new[k] = c_new[k] + s * overlap / h_new[k];
          zn_lo = zn_hi;
        }
      }
    }
    zo_lo = zo_hi;
  }
}

/* remap_bound_column (rdb_ocean_remap.F90:321-452) */
static void bound_column(int nz, const double *h_old, const double *h_new, const double *c_old,
                         const double *c_own, double *c_new, double *clamp_amt, int bnd_extrap) {
  double z_old[nz], z_new[nz];
  double zacc_o = 0.0, zacc_n = 0.0;
  for (int k = 0; k < nz; ++k) {
    zacc_o = zacc_o + h_old[k];
    zacc_n = zacc_n + h_new[k];
    z_old[k] = zacc_o;
    z_new[k] = zacc_n;
  }
  int ko0 = 1; /* 1-based as the reference */
  for (int k = 1; k <= nz; ++k) {
    double zn_hi = z_new[k - 1];
    double zn_lo = (k > 1) ? z_new[k - 2] : 0.0;
    while (ko0 <= nz) {
      if (z_old[ko0 - 1] > zn_lo) break;
      ko0 = ko0 + 1;
    }
    double c_raw = c_new[k - 1];
    int n_hit = 0;
    double c_lo = 0.0, c_hi = 0.0;
    if (h_new[k - 1] > 0.0) {
      int ko = ko0;
      double zo_lo = 0.0;
      if (ko > 1 && ko <= nz) zo_lo = z_old[ko - 2];
      while (ko <= nz) {
        if (zo_lo >= zn_hi) break;
        double zo_hi = z_old[ko - 1];
        double overlap = rmin(zn_hi, zo_hi) - rmax(zn_lo, zo_lo);
        if (overlap > 0.0 && h_old[ko - 1] > 0.0) {
          int km = (ko - 1 > 1 ? ko - 1 : 1), kp = (ko + 1 < nz ? ko + 1 : nz);
          double cs_lo = rmin(c_old[km - 1], rmin(c_old[ko - 1], c_old[kp - 1]));
          double cs_hi = rmax(c_old[km - 1], rmax(c_old[ko - 1], c_old[kp - 1]));
          if (bnd_extrap && nz >= 2) {
            if (ko == 1) {
              cs_lo = rmin(cs_lo, 2.0 * c_old[0] - c_old[1]);
              cs_hi = rmax(cs_hi, 2.0 * c_old[0] - c_old[1]);
            } else if (ko == nz) {
              cs_lo = rmin(cs_lo, 2.0 * c_old[nz - 1] - c_old[nz - 2]);
              cs_hi = rmax(cs_hi, 2.0 * c_old[nz - 1] - c_old[nz - 2]);
            }
          }
          if (h_old[ko - 1] <= ORC_H_VANISHED) {
            cs_lo = cs_lo + (c_own[ko - 1] - c_old[ko - 1]);
            cs_hi = cs_hi + (c_own[ko - 1] - c_old[ko - 1]);
          }
          if (n_hit == 0) {
            c_lo = cs_lo;
            c_hi = cs_hi;
          } else {
            c_lo = rmin(c_lo, cs_lo);
            c_hi = rmax(c_hi, cs_hi);
          }
          n_hit = n_hit + 1;
        }
        zo_lo = zo_hi;
        ko = ko + 1;
      }
    }
    if (n_hit > 0) {
      if (isfinite(c_raw)) c_new[k - 1] = rmax(c_lo, rmin(c_hi, c_raw));
    } else {
      c_new[k - 1] = c_own[(ko0 < nz ? ko0 : nz) - 1];
    }
    clamp_amt[k - 1] = h_new[k - 1] * (c_new[k - 1] - c_raw);
  }
}

/* The per-column pipeline of ocean_remap_tracer_field (:296-317) and
 * ocean_remap_tracer_column (:1425-1457): donor column, remap, filler
 * excess, bounding.  c is read as c_old and overwritten with c_new. */
void orc_remap_tracer_column(int method, int nz, const double *h_old, const double *h_new, double *c,
                             int bnd_extrap, int nonunif, double *c_new_out, double *clamp_out) {
  double c_don[nz], c_new[nz], clamp[nz];
  orc_vl_column_conc(nz, h_old, c, c_don);
  orc_remap_column(method, nz, h_old, h_new, c_don, c_new, bnd_extrap, nonunif);
  add_filler_excess(nz, h_old, h_new, c_don, c, c_new);
  bound_column(nz, h_old, h_new, c_don, c, c_new, clamp, bnd_extrap);
  for (int k = 0; k < nz; ++k) {
    c[k] = c_new[k];
    if (c_new_out) c_new_out[k] = c_new[k];
    if (clamp_out) clamp_out[k] = clamp[k];
  }
}

/* ocean_remap_tracer_field (rdb_ocean_remap.F90:236-319) */
void orc_remap_tracer_field(int method, int nx, int ny, int nz, const double *h_old, const double *h_new,
                            double *c, int bnd_extrap, int nonunif, double *budget, double *c_new_out,
                            double *clamp_out) {
  const long plane = (long)nx * ny;
#pragma omp parallel for schedule(static)
  for (long col = 0; col < plane; ++col) {
    double ho[nz], hn[nz], co[nz], cn[nz], cl[nz];
    for (int k = 0; k < nz; ++k) {
      ho[k] = h_old[col + plane * k];
      hn[k] = h_new[col + plane * k];
      co[k] = c[col + plane * k];
      cn[k] = co[k];
    }
    orc_remap_tracer_column(method, nz, ho, hn, cn, bnd_extrap, nonunif, NULL, cl);
    for (int k = 0; k < nz; ++k) {
      long idx = col + plane * k;
      if (budget) {
        double p_new = hn[k] * cn[k];
        double p_old = ho[k] * co[k];
        budget[idx] = budget[idx] + (p_new - p_old);
      }
      if (c_new_out) c_new_out[idx] = cn[k];
      if (clamp_out) clamp_out[idx] = cl[k];
      c[idx] = cn[k];
    }
  }
}

/* rescale_anomaly_ke (rdb_ocean_remap.F90:1131-1180) */
static void rescale_ke(int nz, const double *ho, const double *hn, const double *uo, double *un) {
  double h_old_sum = 0.0, h_new_sum = 0.0, mom_old = 0.0, mom_new = 0.0;
  for (int k = 0; k < nz; ++k) {
    h_old_sum = h_old_sum + ho[k];
    h_new_sum = h_new_sum + hn[k];
    mom_old = mom_old + ho[k] * uo[k];
    mom_new = mom_new + hn[k] * un[k];
  }
  if (h_old_sum <= ORC_H_VANISHED || h_new_sum <= ORC_H_VANISHED) return;
  double u_bar_old = mom_old / h_old_sum;
  double u_bar_new = mom_new / h_new_sum;
  double ke_old = 0.0, ke_new = 0.0;
  for (int k = 0; k < nz; ++k) {
    double anom = uo[k] - u_bar_old;
    ke_old = ke_old + 0.5 * ho[k] * anom * anom;
    anom = un[k] - u_bar_new;
    ke_new = ke_new + 0.5 * hn[k] * anom * anom;
  }
  if (ke_new <= 0.0 || ke_old <= 0.0) return;
  double scale_fac = sqrt(ke_old / ke_new);
  if (scale_fac > 1.25) scale_fac = 1.25;
  if (scale_fac < 0.0) scale_fac = 0.0;
  for (int k = 0; k < nz; ++k) un[k] = u_bar_new + scale_fac * (un[k] - u_bar_new);
}

/* remap_x_face_velocity (:967-1049) with axis = 0 and remap_y_face_velocity
 * (:1051-1129) with axis = 1.  The face array is (nx+1) x ny x nz (axis 0)
 * or nx x (ny+1) x nz (axis 1), remapped in place. */
void orc_remap_face_velocity(int axis, int nx, int ny, int nz, const double *h_old, const double *h_new,
                             double *u, int method, int conserve_ke, int zlevel_faces, int bnd_extrap,
                             int nonunif) {
  const int nu = axis == 0 ? nx + 1 : nx, nv = axis == 0 ? ny : ny + 1;
  const int nfirst = axis == 0 ? nu : nv;
  const long plane = (long)nx * ny, fplane = (long)nu * nv;
#pragma omp parallel for schedule(static)
  for (long f = 0; f < fplane; ++f) {
    const int I = (int)(f % nu), J = (int)(f / nu);
    const int along = axis == 0 ? I : J; /* 0-based face index along the axis */
    long cell_lo, cell_hi;               /* the two cells either side (0-based) */
    if (axis == 0) {
      cell_lo = (I > 0 ? I - 1 : 0) + (long)nx * J;
      cell_hi = (I < nx ? I : nx - 1) + (long)nx * J;
    } else {
      cell_lo = I + (long)nx * (J > 0 ? J - 1 : 0);
      cell_hi = I + (long)nx * (J < ny ? J : ny - 1);
    }
    const int wall = (along == 0) || (along == nfirst - 1);
    double ho[nz], hn[nz], uo[nz], un[nz];
    for (int k = 0; k < nz; ++k) {
      const double oa = h_old[cell_lo + plane * k], ob = h_old[cell_hi + plane * k];
      const double na = h_new[cell_lo + plane * k], nb = h_new[cell_hi + plane * k];
      if (wall) {
        ho[k] = along == 0 ? ob : oa;
        hn[k] = along == 0 ? nb : na;
      } else {
        ho[k] = 0.5 * (oa + ob);
        hn[k] = 0.5 * (na + nb);
      }
      if (zlevel_faces) {
        if (wall) {
          ho[k] = along == 0 ? ob : oa;
          hn[k] = along == 0 ? nb : na;
        } else {
          ho[k] = rmin(oa, ob);
          hn[k] = rmin(na, nb);
        }
        if (ho[k] <= ORC_H_VANISHED || hn[k] <= ORC_H_VANISHED) {
          ho[k] = 0.0;
          hn[k] = 0.0;
        }
      }
      uo[k] = u[f + fplane * k];
    }
    orc_remap_column(method, nz, ho, hn, uo, un, bnd_extrap, nonunif);
    if (conserve_ke) rescale_ke(nz, ho, hn, uo, un);
    for (int k = 0; k < nz; ++k) u[f + fplane * k] = un[k];
  }
}

/* ocean_remap_scan_preconditions (rdb_ocean_remap.F90:1348-1405) */
void orc_remap_scan_preconditions(int nx, int ny, int nz, const double *h_old, const double *h_new, double rel_tol,
                                  int *n_bad, double *worst_rel, double *worst_neg) {
  const long plane = (long)nx * ny;
  int nb = 0;
  double wr = 0.0, wn = 0.0;
  for (long col = 0; col < plane; ++col) {
    double s_old = 0.0, s_new = 0.0, hmin = 0.0;
    for (int k = 0; k < nz; ++k) {
      s_old = s_old + h_old[col + plane * k];
      s_new = s_new + h_new[col + plane * k];
      hmin = rmin(hmin, rmin(h_old[col + plane * k], h_new[col + plane * k]));
    }
    double rel = fabs(s_new - s_old) / rmax(rmax(fabs(s_old), fabs(s_new)), ORC_H_DIV_EPS);
    wr = rmax(wr, rel);
    wn = rmin(wn, hmin);
    if (hmin < 0.0 || rel > rel_tol) nb = nb + 1;
  }
  *n_bad = nb;
  *worst_rel = wr;
  *worst_neg = wn;
}
