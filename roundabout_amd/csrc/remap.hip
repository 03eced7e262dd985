// remap.hip -- conservative vertical (ALE) remap of tracer concentrations
// and face velocities.
//
//   rdb_remap_tracer_field   ocean_remap_tracer_field
//                            (src/ALE/rdb_ocean_remap.F90:236-319): donor
//                            column (rdb_vanished_layer.inc:105-153), the
//                            column remap, filler excess (:454-489),
//                            bounding (:321-452), budget increment
//   rdb_remap_face_velocity  remap_x/y_face_velocity (:967-1129) with
//                            rescale_anomaly_ke (:1131-1180)
//   rdb_remap_columns        remap_column (src/ALE/rdb_remap_column.F90:
//                            328-375) over a batch of columns
//   rdb_remap_scan_preconditions  ocean_remap_scan_preconditions (:1348-1405)
//
// One lane per column.  Fields are in the reference's natural layout (i
// fastest, level-major), so the 64 lanes of a wavefront read 64 neighbouring
// columns of one level as one coalesced access.  The column remap is a merge
// walk of the target cells against the source cells, both bottom-up: the
// source cells sit in a four-cell register window (s-1 .. s+2) that slides
// up the column, each source cell's reconstruction (PLM slope, PPM/PPM_H4
// limited parabola) is formed from that window once, when the walk reaches
// it, and every target adds the overlap integrals of the source cells that
// overlap it.  So h_old, h_new and c are each read once from HBM and c_new
// written once; there are no per-column arrays.  PQM, whose edge values and
// slopes come from two tridiagonal solves over the whole column, first runs
// a per-column pass that leaves the quartic coefficients of every cell in a
// level-major work array, which the walk then reads like the window.
//
// Bitwise equality with the reference's loops.  For non-negative
// thicknesses the interface stacks are non-decreasing, so the cells with a
// positive overlap of a target form one contiguous run; the reference sums
// exactly those, from 0, in increasing source index, and its scan pointer
// and early exit only skip cells that cannot overlap (the argument of
// ocean_remap_tracer_field_stream, rdb_ocean_remap.F90:668-685).  The walk
// adds the same terms in the same order with the same expressions, the
// interfaces are the same running sums from the bed, and the library is
// built without FMA contraction; the bounding visits the same contributors
// with the same stencil ranges.  A filler's excess (remap_add_filler_excess)
// must be added after the division by h_new, in source order, so a target
// with a filler contributor re-walks its run of source cells once.
// Negative thicknesses violate the reference's precondition (it creates
// mass there); the results for them are not defined here either.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/rdb_colphys.h"
#include "colphys_handle.h"

namespace remap {

namespace {

constexpr double HV = 1.5e-4;           // H_VANISHED (rdb_constants.F90:67)
constexpr double EPS = 1.0e-20;         // H_DIV_EPS (rdb_constants.F90:85)
constexpr double NEG = 1.0e-30;         // H_NEGLECT (rdb_remap_column.F90:68)
constexpr double MIN_FRAC = 1.0e-5;     // H_MIN_FRAC
constexpr double RATIO_FLOOR = 1.0e-12; // H_RATIO_FLOOR
constexpr double PQM_MIN = 1.0e-6;      // PQM_MIN_FRAC

// The methods of rdb_constants.F90:208-218 plus the single-layer identity
// every reconstructing method reduces to at nz = 1.
enum { M_PCM = 1, M_PLM = 2, M_PPM = 3, M_H4 = 4, M_PQM = 5, M_IDENT = 6 };

constexpr int TPB = 256;

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
// Fortran SIGN(a, b)
__device__ __forceinline__ double dsign(double a, double b) { return copysign(fabs(a), b); }

// boundary_half_jump (rdb_remap_column.F90:83-118)
__device__ double half_jump(double h_self, double h_nbr, double dq_up) {
  double d = dq_up * h_self / dmax(h_self + h_nbr, NEG);
  return dsign(dmin(fabs(d), fabs(dq_up)), d);
}

// plm_slope_nonuniform (:168-225)
__device__ double plm_slope_nu(double h_l, double h_c, double h_r, double q_l, double q_c, double q_r) {
  const double sig_l = q_c - q_l;
  const double sig_r = q_r - q_c;
  const double sig_c = (h_c / (h_l + h_c + h_r + EPS)) *
                       ((2.0 * h_l + h_c) / (h_c + h_r + EPS) * sig_r + (h_c + 2.0 * h_r) / (h_l + h_c + EPS) * sig_l);
  if (sig_l * sig_r > 0.0) {
    const double q_min = dmin(dmin(q_l, q_c), q_r);
    const double q_max = dmax(dmax(q_l, q_c), q_r);
    return 0.5 * dsign(dmin(fabs(sig_c), 2.0 * dmin(q_c - q_min, q_max - q_c)), sig_c);
  }
  return 0.0;
}

// ppm_edge_two_cell (:227-249)
__device__ double edge_two_cell(double h_l, double h_r, double q_l, double q_r) {
  return (q_l * h_r + q_r * h_l) / (h_l + h_r + EPS);
}

// ppm_jump_nonuniform (:251-285)
__device__ double jump_nu(double h_l, double h_c, double h_r, double q_l, double q_c, double q_r) {
  return (h_c / (h_l + h_c + h_r + EPS)) *
         ((2.0 * h_l + h_c) / (h_c + h_r + EPS) * (q_r - q_c) + (h_c + 2.0 * h_r) / (h_l + h_c + EPS) * (q_c - q_l));
}

// ppm_edge_nonuniform (:287-326)
__device__ double edge_nu(double h0, double h1, double h2, double h3, double q1, double q2, double dq1, double dq2) {
  const double dq21 = q2 - q1;
  const double i_h12 = 1.0 / (h1 + h2 + EPS);
  const double wa = (h0 + h1) / (2.0 * h1 + h2 + EPS);
  const double wb = (h2 + h3) / (h1 + 2.0 * h2 + EPS);
  return q1 + h1 * i_h12 * dq21 +
         (2.0 * h1 * h2 * i_h12 * (wa - wb) * dq21 - h1 * wa * dq2 + h2 * wb * dq1) / (h0 + h1 + h2 + h3 + EPS);
}

// The H3 boundary quadratic of remap_column_ppm_h4 (:851-927): the cell-mean
// fit through three cells, evaluated at their first (at_z2 = false) or
// second interface.
__device__ double h3_edge(double h1, double h2, double h3, double qa, double qb, double qc, bool at_z2) {
  if (h1 <= 0.0 || h2 <= 0.0 || h3 <= 0.0) {
    const double hf = MIN_FRAC * dmax(NEG, h1 + h2 + h3);
    h1 = dmax(h1, hf);
    h2 = dmax(h2, hf);
    h3 = dmax(h3, hf);
  }
  const double z1 = h1, z2 = h1 + h2, z3 = h1 + h2 + h3;
  const double m11 = (z1 * z1 * z1) / (3.0 * h1);
  const double m12 = (z1 * z1) / (2.0 * h1);
  const double m13 = 1.0;
  const double m21 = (z2 * z2 * z2 - z1 * z1 * z1) / (3.0 * h2);
  const double m22 = (z2 * z2 - z1 * z1) / (2.0 * h2);
  const double m23 = 1.0;
  const double m31 = (z3 * z3 * z3 - z2 * z2 * z2) / (3.0 * h3);
  const double m32 = (z3 * z3 - z2 * z2) / (2.0 * h3);
  const double m33 = 1.0;
  const double det = m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31);
  const double ca = (qa * (m22 * m33 - m23 * m32) - m12 * (qb * m33 - m23 * qc) + m13 * (qb * m32 - m22 * qc)) / det;
  const double cb = (m11 * (qb * m33 - m23 * qc) - qa * (m21 * m33 - m23 * m31) + m13 * (m21 * qc - qb * m31)) / det;
  const double cc = (m11 * (m22 * qc - qb * m32) - m12 * (m21 * qc - qb * m31) + qa * (m21 * m32 - m22 * m31)) / det;
  const double z = at_z2 ? z2 : z1;
  return ca * z * z + cb * z + cc;
}

// The interior H4 edge of remap_column_ppm_h4 (:932-955).
__device__ double h4_edge(double h0, double h1, double h2, double h3, double qm1, double q0, double qp1, double qp2) {
  const double h_sum = h0 + h1 + h2 + h3;
  if (h0 + h1 <= 0.0 || h1 + h2 <= 0.0 || h2 + h3 <= 0.0) {
    const double hf = MIN_FRAC * dmax(NEG, h_sum);
    h0 = dmax(h0, hf);
    h1 = dmax(h1, hf);
    h2 = dmax(h2, hf);
    h3 = dmax(h3, hf);
  }
  const double h01 = h0 + h1, h12 = h1 + h2, h23 = h2 + h3;
  const double h012 = h0 + h1 + h2, h123 = h1 + h2 + h3, h0123 = h0 + h1 + h2 + h3;
  const double f1 = h01 * h23 / h12;
  const double f2 = h2 * q0 + h1 * qp1;
  const double f3 = 1.0 / h012 + 1.0 / h123;
  const double et1 = f1 * f2 * f3;
  const double et2 = (h2 * h23 / (h012 * h01)) * ((h0 + 2.0 * h1) * q0 - h1 * qm1);
  const double et3 = (h1 * h01 / (h123 * h23)) * ((2.0 * h2 + h3) * qp1 - h2 * qp2);
  return (et1 + et2 + et3) / h0123;
}

// pqm_end_value_h4 (:1099-1170): value (csys0) and slope (csys1) of the
// one-sided fit of four cells starting at the boundary.
__device__ void end_value_h4(const double dz[4], const double u[4], double& csys0, double& csys1) {
  double h1 = dz[0], h2 = dz[1], h3 = dz[2], h4 = dz[3];
  if ((h2 + h3) < PQM_MIN * h1) h3 = PQM_MIN * h1 - h2;
  if ((h3 + h4) < PQM_MIN * h1) h4 = PQM_MIN * h1 - h3;
  const double h12 = h1 + h2, h23 = h2 + h3, h34 = h3 + h4;
  const double h123 = h12 + h3, h234 = h2 + h34, h1234 = h12 + h34;
  const double i_denb3 = 1.0 / (h123 * h12 * h23);
  const double i_h12 = (h123 * h23) * i_denb3;
  const double i_h23 = (h12 * h123) * i_denb3;
  const double i_h123 = (h12 * h23) * i_denb3;
  const double i_denom = 1.0 / (h1234 * (h234 * h34));
  const double i_h234 = (h1234 * h34) * i_denom;
  const double i_h1234 = (h234 * h34) * i_denom;
  const double w11 = -h1 * (i_h1234 + i_h123 + i_h12);
  const double w21 = h1 * h12 * (i_h234 * i_h1234 + i_h23 * (i_h234 + i_h123));
  const double w31 = -h1 * h12 * h123 * i_denom;
  const double w12 = 2.0 * (i_h12 * (1.0 + (h1 + h12) * (i_h1234 + i_h123)) + h1 * i_h1234 * i_h123);
  const double w22 = -2.0 * ((h1 * h12 * i_h1234) * (i_h23 * (i_h234 + i_h123)) +
                             (h1 + h12) * (i_h1234 * i_h234 + i_h23 * (i_h234 + i_h123)));
  const double w32 = 2.0 * ((h1 + h12) * h123 + h1 * h12) * i_denom;
  const double du1 = u[1] - u[0], du2 = u[2] - u[1], du3 = u[3] - u[2];
  csys0 = ((u[0] + (w11 * du1)) + (w21 * du2)) + (w31 * du3);
  csys1 = ((w12 * du1) + (w22 * du2)) + (w32 * du3);
}

// ---------------------------------------------------------------------------
// Column accessors.  A lane binds one column; src_h / dst_h are the source
// and target thickness of level k, val its field value.

// A field in the natural layout (or a batch of raw columns): element k of
// column col at col + stride*k.
struct CellCols {
  const double* ho;
  const double* hn;
  const double* c;
  long stride;
  long col;
  __device__ void bind(long i) { col = i; }
  __device__ double src_h(int k) const { return ho[col + stride * k]; }
  __device__ double dst_h(int k) const { return hn[col + stride * k]; }
  __device__ double val(int k) const { return c[col + stride * k]; }
};

// A face column of remap_x_face_velocity (axis 0: (nx+1) x ny faces) or
// remap_y_face_velocity (axis 1: nx x (ny+1)): face thicknesses from the two
// abutting cells, the wall faces take their one cell's.
struct FaceCols {
  const double* ho;
  const double* hn;
  const double* u;
  long plane;  // nx*ny
  long fplane; // faces per level
  int nx, ny, axis, zlevel;
  long col, lo, hi;
  int wall;
  __device__ void bind(long f) {
    col = f;
    const int nu = axis == 0 ? nx + 1 : nx;
    const int I = (int)(f % nu), J = (int)(f / nu);
    const int along = axis == 0 ? I : J;
    const int nlast = axis == 0 ? nx : ny;
    if (axis == 0) {
      lo = (I > 0 ? I - 1 : 0) + (long)nx * J;
      hi = (I < nx ? I : nx - 1) + (long)nx * J;
    } else {
      lo = I + (long)nx * (J > 0 ? J - 1 : 0);
      hi = I + (long)nx * (J < ny ? J : ny - 1);
    }
    wall = (along == 0) || (along == nlast);
  }
  __device__ void pair(int k, double& o, double& n) const {
    const double oa = ho[lo + plane * k], na = hn[lo + plane * k];
    if (wall) {  // lo == hi: the one abutting cell
      o = oa;
      n = na;
    } else {
      const double ob = ho[hi + plane * k], nb = hn[hi + plane * k];
      if (zlevel) {
        o = dmin(oa, ob);
        n = dmin(na, nb);
      } else {
        o = 0.5 * (oa + ob);
        n = 0.5 * (na + nb);
      }
    }
    if (zlevel && (o <= HV || n <= HV)) {
      o = 0.0;
      n = 0.0;
    }
  }
  __device__ double src_h(int k) const {
    double o, n;
    pair(k, o, n);
    return o;
  }
  __device__ double dst_h(int k) const {
    double o, n;
    pair(k, o, n);
    return n;
  }
  __device__ double val(int k) const { return u[col + fplane * k]; }
};

// One source cell of the window: thickness, the donor-substituted value the
// reconstruction runs on (rdb_vl_column_conc) and the stored value.
struct Win {
  double h, cd, co;
};

// PQM work arrays, each (nz+1) levels x ncol, level-major.
enum { P_CD = 0, P_XV, P_C1V, P_XS, P_C1S, P_EVL, P_EVR, P_ESL, P_ESR, P_NARR };
// The quartic coefficients overwrite arrays that are dead by then.
enum { P_A = P_XV, P_B = P_C1V, P_C = P_XS, P_D = P_C1S, P_E = P_CD };

// ---------------------------------------------------------------------------
// PQM reconstruction (remap_column_pqm :1226-1539) of one column into the
// work arrays: the donor column, the two implicit tridiagonal solves for the
// edge values and slopes, the W&A limiter and the quartic coefficients.
template <class Cols, bool FIELD>
__global__ __launch_bounds__(TPB) void k_pqm_coeffs(Cols cols, long ncol, int nz, int be, double* __restrict__ W,
                                                    long arr) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  cols.bind(col);
  double* const cd = W + P_CD * arr;
  double* const xv = W + P_XV * arr;
  double* const c1v = W + P_C1V * arr;
  double* const xs = W + P_XS * arr;
  double* const c1s = W + P_C1S * arr;
  double* const evl = W + P_EVL * arr;
  double* const evr = W + P_EVR * arr;
  double* const esl = W + P_ESL * arr;
  double* const esr = W + P_ESR * arr;
#define AT(a, k) (a)[col + ncol * (long)(k)]
  // donor-substituted column (rdb_vl_column_conc)
  if (FIELD) {
    int k_top = -1;
    for (int k = nz - 1; k >= 0; --k)
      if (cols.src_h(k) > HV) {
        k_top = k;
        break;
      }
    if (k_top < 0) {
      for (int k = 0; k < nz; ++k) AT(cd, k) = 0.0;
    } else {
      double cdon = cols.val(k_top);
      for (int k = nz - 1; k >= 0; --k) {
        if (cols.src_h(k) > HV) {
          cdon = cols.val(k);
          AT(cd, k) = cdon;
        } else {
          AT(cd, k) = cdon;
        }
      }
    }
  } else {
    for (int k = 0; k < nz; ++k) AT(cd, k) = cols.val(k);
  }
  // one-sided four-cell end fits at the bed and (reversed) at the surface
  double dz4[4], u4[4], bed_v, bed_s, top_v, top_s;
  for (int m = 0; m < 4; ++m) {
    dz4[m] = dmax(NEG, cols.src_h(m));
    u4[m] = AT(cd, m);
  }
  end_value_h4(dz4, u4, bed_v, bed_s);
  for (int m = 0; m < 4; ++m) {
    dz4[m] = dmax(NEG, cols.src_h(nz - 1 - m));
    u4[m] = AT(cd, nz - 1 - m);
  }
  end_value_h4(dz4, u4, top_v, top_s);
  // forward sweep of pqm_solve_diag_dominant for both systems; row 0 is the
  // bed end (diagonal 1, upper 0), rows 1..nz-1 the interior edges
  double d1v = 1.0 * (1.0 / (1.0 + 0.0)), d1s = d1v;
  {
    const double ip = 1.0 / (1.0 + 0.0);
    AT(c1v, 0) = 0.0 * ip;
    AT(c1s, 0) = 0.0 * ip;
    AT(xv, 0) = bed_v * ip;
    AT(xs, 0) = bed_s * ip;
  }
  double xv_prev = AT(xv, 0), xs_prev = AT(xs, 0);
  double h_lo = cols.src_h(0), q_lo = AT(cd, 0);
  for (int k = 1; k <= nz - 1; ++k) {
    const double h_hi = cols.src_h(k), q_hi = AT(cd, k);
    // edge values (:1266-1281)
    {
      double h0 = dmax(h_lo, NEG);
      double h1 = dmax(h_hi, NEG);
      if (fabs(h0) < RATIO_FLOOR * fabs(h1)) h0 = RATIO_FLOOR * h1;
      if (fabs(h1) < RATIO_FLOOR * fabs(h0)) h1 = RATIO_FLOOR * h0;
      const double i_h2 = 1.0 / ((h0 + h1) * (h0 + h1));
      const double alpha = (h1 * h1) * i_h2;
      const double beta = (h0 * h0) * i_h2;
      const double abmix = (h0 * h1) * i_h2;
      const double aco = 2.0 * alpha * (alpha + 2.0 * beta + 3.0 * abmix);
      const double bco = 2.0 * beta * (beta + 2.0 * alpha + 3.0 * abmix);
      const double al = alpha, ac = 2.0 * abmix, au = beta, r = aco * q_lo + bco * q_hi;
      const double denom = ac + d1v * al;
      const double ip = 1.0 / (denom + au);
      d1v = denom * ip;
      AT(c1v, k) = au * ip;
      xv_prev = (r - al * xv_prev) * ip;
      AT(xv, k) = xv_prev;
    }
    // edge slopes (:1310-1327)
    {
      double h0 = dmax(h_lo, NEG);
      double h1 = dmax(h_hi, NEG);
      const double i_h = 1.0 / (h0 + h1);
      h0 = h0 * i_h;
      h1 = h1 * i_h;
      const double h0h1 = h0 * h1, h0_2 = h0 * h0, h1_2 = h1 * h1;
      const double h0_3 = h0_2 * h0, h1_3 = h1_2 * h1;
      const double i_d = 1.0 / (4.0 * h0h1 * (h0 + h1) + h1_3 + h0_3);
      const double al = (h1 * ((h0_2 + h0h1) - h1_2)) * i_d;
      const double ac = 2.0 * ((h0_2 + h1_2) * (h0 + h1)) * i_d;
      const double au = (h0 * ((h1_2 + h0h1) - h0_2)) * i_d;
      const double r = 12.0 * (h0h1 * i_d) * ((q_hi - q_lo) * i_h);
      const double denom = ac + d1s * al;
      const double ip = 1.0 / (denom + au);
      d1s = denom * ip;
      AT(c1s, k) = au * ip;
      xs_prev = (r - al * xs_prev) * ip;
      AT(xs, k) = xs_prev;
    }
    h_lo = h_hi;
    q_lo = q_hi;
  }
  // last row: the surface end (lower 0, diagonal 1), then back-substitution
  {
    const double ipv = 1.0 / (1.0 + d1v * 0.0);
    xv_prev = (top_v - 0.0 * xv_prev) * ipv;
    AT(xv, nz) = xv_prev;
    const double ips = 1.0 / (1.0 + d1s * 0.0);
    xs_prev = (-top_s - 0.0 * xs_prev) * ips;
    AT(xs, nz) = xs_prev;
  }
  for (int k = nz - 1; k >= 0; --k) {
    xv_prev = AT(xv, k) - AT(c1v, k) * xv_prev;
    AT(xv, k) = xv_prev;
    xs_prev = AT(xs, k) - AT(c1s, k) * xs_prev;
    AT(xs, k) = xs_prev;
  }
  // 3a. bound_edge_values (:1346-1370)
  for (int k = 0; k < nz; ++k) {
    const int kl = k > 0 ? k - 1 : 0, kr = k < nz - 1 ? k + 1 : nz - 1;
    const double u_l = AT(cd, kl), u_c = AT(cd, k), u_r = AT(cd, kr);
    const double h_l = cols.src_h(kl), h_c = cols.src_h(k), h_r = cols.src_h(kr);
    double ev_l = AT(xv, k), ev_r = AT(xv, k + 1);
    double slope_x_h = 0.0;
    if (((h_l + h_r) + 2.0 * h_c) > 0.0) {
      const double sigma_l = (u_c - u_l);
      const double sigma_c = (u_r - u_l) * (h_c / ((h_l + h_r) + 2.0 * h_c));
      const double sigma_r = (u_r - u_c);
      if ((sigma_l * sigma_r) > 0.0) slope_x_h = dsign(dmin(dmin(fabs(sigma_l), fabs(sigma_c)), fabs(sigma_r)), sigma_c);
    }
    if ((u_l - ev_l) * (ev_l - u_c) < 0.0) ev_l = u_c - dsign(dmin(fabs(slope_x_h), fabs(ev_l - u_c)), slope_x_h);
    if ((u_r - ev_r) * (ev_r - u_c) < 0.0) ev_r = u_c + dsign(dmin(fabs(slope_x_h), fabs(ev_r - u_c)), slope_x_h);
    ev_l = dmax(dmin(ev_l, dmax(u_l, u_c)), dmin(u_l, u_c));
    ev_r = dmax(dmin(ev_r, dmax(u_r, u_c)), dmin(u_r, u_c));
    AT(evl, k) = ev_l;
    AT(evr, k) = ev_r;
    AT(esl, k) = AT(xs, k);
    AT(esr, k) = AT(xs, k + 1);
  }
  // 3b. check_discontinuous_edge_values (:1375-1384)
  for (int k = 0; k < nz - 1; ++k) {
    const double qa = AT(cd, k), qb = AT(cd, k + 1);
    const double er = AT(evr, k), el = AT(evl, k + 1);
    if ((el - er) * (qb - qa) < 0.0) {
      double u0_avg = 0.5 * (er + el);
      u0_avg = dmax(dmin(u0_avg, dmax(qa, qb)), dmin(qa, qb));
      AT(evr, k) = u0_avg;
      AT(evl, k + 1) = u0_avg;
    }
  }
  // 3c. interior cells (:1388-1505)
  for (int k = 1; k <= nz - 2; ++k) {
    int infl_l = 0, infl_r = 0;
    double u0_l = AT(evl, k), u0_r = AT(evr, k);
    double u1_l = AT(esl, k), u1_r = AT(esr, k);
    const double h_l = cols.src_h(k - 1), h_c = cols.src_h(k), h_r = cols.src_h(k + 1);
    const double u_l = AT(cd, k - 1), u_c = AT(cd, k), u_r = AT(cd, k + 1);
    const double sigma_l = 2.0 * (u_c - u_l) / (h_c + NEG);
    const double sigma_c = 2.0 * (u_r - u_l) / (h_l + 2.0 * h_c + h_r + NEG);
    const double sigma_r = 2.0 * (u_r - u_c) / (h_c + NEG);
    double slope;
    if ((sigma_l * sigma_r) > 0.0)
      slope = dsign(dmin(dmin(fabs(sigma_l), fabs(sigma_c)), fabs(sigma_r)), sigma_c);
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
      const double b = h_c * u1_l;
      const double cco = 30.0 * u_c - 12.0 * u0_r - 18.0 * u0_l + 1.5 * h_c * (u1_r - 3.0 * u1_l);
      const double dco = -60.0 * u_c + h_c * (6.0 * u1_l - 4.0 * u1_r) + 28.0 * u0_r + 32.0 * u0_l;
      const double eco = 30.0 * u_c + 2.5 * h_c * (u1_r - u1_l) - 15.0 * (u0_l + u0_r);
      const double alpha1 = 6.0 * eco, alpha2 = 3.0 * dco, alpha3 = cco;
      const double rho = alpha2 * alpha2 - 4.0 * alpha1 * alpha3;
      bool flag = false;
      if ((alpha1 != 0.0) && (rho >= 0.0)) {
        const double sqrt_rho = sqrt(rho);
        const double x1 = 0.5 * (-alpha2 - sqrt_rho) / alpha1;
        const double x2 = 0.5 * (-alpha2 + sqrt_rho) / alpha1;
        const bool in1 = (x1 >= 0.0) && (x1 <= 1.0), in2 = (x2 >= 0.0) && (x2 <= 1.0);
        const double g1 = 4.0 * eco * (x1 * x1 * x1) + 3.0 * dco * (x1 * x1) + 2.0 * cco * x1 + b;
        const double g2 = 4.0 * eco * (x2 * x2 * x2) + 3.0 * dco * (x2 * x2) + 2.0 * cco * x2 + b;
        if (in1 && in2)
          flag = (g1 * slope < 0.0) || (g2 * slope < 0.0);
        else if (in1)
          flag = g1 * slope < 0.0;
        else if (in2)
          flag = g2 * slope < 0.0;
        if (flag) {
          if (fabs(sigma_l) < fabs(sigma_r))
            infl_l = 1;
          else
            infl_r = 1;
        }
      }
      if ((alpha1 == 0.0) && (alpha2 != 0.0)) {
        const double x1 = -alpha3 / alpha2;
        if ((x1 >= 0.0) && (x1 <= 1.0)) {
          const double g1 = 4.0 * eco * (x1 * x1 * x1) + 3.0 * dco * (x1 * x1) + 2.0 * cco * x1 + b;
          if (g1 * slope < 0.0) {
            if (fabs(sigma_l) < fabs(sigma_r))
              infl_l = 1;
            else
              infl_r = 1;
          }
        }
      }
    }
    if (infl_l == 1) {
      u1_l = (10.0 * u_c - 2.0 * u0_r - 8.0 * u0_l) / (3.0 * h_c + NEG);
      u1_r = (-10.0 * u_c + 6.0 * u0_r + 4.0 * u0_l) / (h_c + NEG);
      if (u1_l * slope < 0.0) {
        u1_l = 0.0;
        u0_r = 5.0 * u_c - 4.0 * u0_l;
        u1_r = 20.0 * (u_c - u0_l) / (h_c + NEG);
      } else if (u1_r * slope < 0.0) {
        u1_r = 0.0;
        u0_l = (5.0 * u_c - 3.0 * u0_r) / 2.0;
        u1_l = 10.0 * (-u_c + u0_r) / (3.0 * h_c + NEG);
      }
    } else if (infl_r == 1) {
      u1_r = (-10.0 * u_c + 8.0 * u0_r + 2.0 * u0_l) / (3.0 * h_c + NEG);
      u1_l = (10.0 * u_c - 4.0 * u0_r - 6.0 * u0_l) / (h_c + NEG);
      if (u1_l * slope < 0.0) {
        u1_l = 0.0;
        u0_r = (5.0 * u_c - 3.0 * u0_l) / 2.0;
        u1_r = 10.0 * (u_c - u0_l) / (3.0 * h_c + NEG);
      } else if (u1_r * slope < 0.0) {
        u1_r = 0.0;
        u0_l = 5.0 * u_c - 4.0 * u0_r;
        u1_l = 20.0 * (-u_c + u0_r) / (h_c + NEG);
      }
    }
    AT(evl, k) = u0_l;
    AT(evr, k) = u0_r;
    AT(esl, k) = u1_l;
    AT(esr, k) = u1_r;
  }
  // boundary cells (:1507-1532), then the quartic of every cell (:1535-1549);
  // the last coefficient overwrites the donor column, so the mean below is
  // carried in a register
  double u_below = 0.0;
  for (int k = 0; k < nz; ++k) {
    const double h_c = cols.src_h(k), u_c = AT(cd, k);
    double u0_l = AT(evl, k), u0_r = AT(evr, k), u1_l = AT(esl, k), u1_r = AT(esr, k);
    if (k == 0 || k == nz - 1) {
      u0_l = u_c;
      u0_r = u_c;
      u1_l = 0.0;
      u1_r = 0.0;
      if (be) {
        const double d = (k == 0) ? half_jump(h_c, cols.src_h(1), AT(cd, 1) - u_c)
                                  : half_jump(h_c, cols.src_h(nz - 2), u_c - u_below);
        u0_l = u_c - d;
        u0_r = u_c + d;
        u1_l = 2.0 * d / dmax(h_c, NEG);
        u1_r = u1_l;
      }
    }
    AT(W + P_A * arr, k) = u0_l;
    AT(W + P_B * arr, k) = h_c * u1_l;
    AT(W + P_C * arr, k) = 30.0 * u_c - 12.0 * u0_r - 18.0 * u0_l + 1.5 * h_c * (u1_r - 3.0 * u1_l);
    AT(W + P_D * arr, k) = -60.0 * u_c + h_c * (6.0 * u1_l - 4.0 * u1_r) + 28.0 * u0_r + 32.0 * u0_l;
    AT(W + P_E * arr, k) = 30.0 * u_c + 2.5 * h_c * (u1_r - u1_l) - 15.0 * (u0_l + u0_r);
    u_below = u_c;
  }
#undef AT
}

// ---------------------------------------------------------------------------
// The merge walk.  FIELD: the concentration-form tracer pipeline of
// ocean_remap_tracer_field (donor column, filler excess, bounding, budget);
// otherwise the plain remap_column of the face and raw-column passes.
template <class Cols, int M, bool FIELD>
__global__ __launch_bounds__(TPB) void k_walk(Cols cols, long ncol, int nz, int be, int nu,
                                              const double* __restrict__ W, long arr, double* __restrict__ out,
                                              double* __restrict__ budget, double* __restrict__ clamp) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  cols.bind(col);

  // Source cells arrive through a two-deep prefetch of their raw (h, c)
  // (loaded two advances before the window takes them), so the load
  // latency overlaps the reconstruction of the cells in between.
  struct Raw {
    double h, c;
  };
  auto raw = [&](int m) -> Raw {
    if (m >= nz) return Raw{0.0, 0.0};
    return Raw{cols.src_h(m), cols.val(m)};
  };
  // donor lookahead: the nearest live layer at or above the load front
  int la = -1;
  double la_c = 0.0, last_live = 0.0;
  bool have_live = false;
  auto finish = [&](int m, Raw r) -> Win {
    Win w{r.h, r.c, r.c};
    if (m >= nz) return w;
    if (FIELD) {
      if (w.h > HV) {
        last_live = w.co;
        have_live = true;
      } else {
        if (la < m) {
          // scan up for the next live layer, eight independent loads a round
          la = nz;
          for (int j0 = m + 1; j0 < nz && la == nz; j0 += 8) {
            double hh[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) hh[u] = (j0 + u < nz) ? cols.src_h(j0 + u) : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (la == nz && hh[u] > HV) la = j0 + u;
          }
          if (la < nz) la_c = cols.val(la);
        }
        // the nearest live layer above; none above: the topmost live one
        // (the last live layer loaded); no live layer: 0
        w.cd = (la < nz) ? la_c : (have_live ? last_live : 0.0);
      }
    }
    return w;
  };
  // the per-layer donor of a filler, for the rare re-walk
  auto donor = [&](int j) -> double {
    for (int m = j + 1; m < nz; ++m)
      if (cols.src_h(m) > HV) return cols.val(m);
    for (int m = j - 1; m >= 0; --m)
      if (cols.src_h(m) > HV) return cols.val(m);
    return 0.0;
  };

  Win w0{0.0, 0.0, 0.0};
  Win w1 = finish(0, raw(0));
  Win w2 = finish(1, raw(1));
  Win w3 = finish(2, raw(2));
  Raw pa = raw(3), pb = raw(4);
  const double cd0 = w1.cd;
  int s = 0;
  double zs_lo = 0.0, zs_hi = 0.0 + w1.h;
  // reconstruction of cell s: PLM slope in ra; PPM/H4 (q_L, q_R, q6) in
  // (ra, rb, r6); PQM coefficients in (ra, rb, r6, rd, re)
  double ra = 0.0, rb = 0.0, r6 = 0.0, rd = 0.0, re = 0.0, carry = 0.0;

  auto recon = [&]() {
    if (M == M_PLM) {
      double slope;
      if (s == 0) {
        slope = be ? half_jump(w1.h, w2.h, w2.cd - w1.cd) : 0.0;
      } else if (s == nz - 1) {
        slope = be ? half_jump(w1.h, w0.h, w1.cd - w0.cd) : 0.0;
      } else if (nu) {
        slope = plm_slope_nu(w0.h, w1.h, w2.h, w0.cd, w1.cd, w2.cd);
      } else {
        const double dq_l = w1.cd - w0.cd, dq_r = w2.cd - w1.cd;
        slope = (dq_l * dq_r > 0.0) ? 0.5 * dsign(dmin(fabs(dq_l), fabs(dq_r)), dq_l) : 0.0;
      }
      ra = slope;
    } else if (M == M_PPM || M == M_H4) {
      const double q = w1.cd;
      const double raw_l = (s == 0) ? q : carry;
      double raw_r;
      if (s == nz - 1) {
        raw_r = q;
      } else if (M == M_PPM) {
        if (nu) {
          if (s == 0 || s == nz - 2)
            raw_r = edge_two_cell(w1.h, w2.h, w1.cd, w2.cd);
          else
            raw_r = edge_nu(w0.h, w1.h, w2.h, w3.h, w1.cd, w2.cd, jump_nu(w0.h, w1.h, w2.h, w0.cd, w1.cd, w2.cd),
                            jump_nu(w1.h, w2.h, w3.h, w1.cd, w2.cd, w3.cd));
        } else {
          if (s == 0 || s == nz - 2)
            raw_r = 0.5 * (w1.cd + w2.cd);
          else
            raw_r = (7.0 / 12.0) * (w1.cd + w2.cd) - (1.0 / 12.0) * (w0.cd + w3.cd);
        }
      } else {
        if (s == 0)
          raw_r = h3_edge(w1.h, w2.h, w3.h, w1.cd, w2.cd, w3.cd, false);
        else if (s == nz - 2)
          raw_r = h3_edge(w0.h, w1.h, w2.h, w0.cd, w1.cd, w2.cd, true);
        else
          raw_r = h4_edge(w0.h, w1.h, w2.h, w3.h, w0.cd, w1.cd, w2.cd, w3.cd);
      }
      carry = raw_r;
      // Colella-Woodward limiter (:680-724)
      double q_min = q, q_max = q;
      if (s > 0) {
        q_min = dmin(q_min, w0.cd);
        q_max = dmax(q_max, w0.cd);
      }
      if (s < nz - 1) {
        q_min = dmin(q_min, w2.cd);
        q_max = dmax(q_max, w2.cd);
      }
      double qL = dmax(q_min, dmin(q_max, raw_l));
      double qR = dmax(q_min, dmin(q_max, raw_r));
      const double dq = qR - qL, dq_l = q - qL, dq_r = qR - q;
      if (dq_l * dq_r <= 0.0) {
        qL = q;
        qR = q;
      } else {
        const double t6 = 6.0 * q - 3.0 * (qL + qR);
        if (fabs(t6) > fabs(dq)) {
          if (t6 * dq > 0.0)
            qL = 3.0 * q - 2.0 * qR;
          else
            qR = 3.0 * q - 2.0 * qL;
        }
      }
      double q6 = 6.0 * q - 3.0 * (qL + qR);
      // boundary closure (:735-746)
      if (be && (s == 0 || s == nz - 1)) {
        const double d = (s == 0) ? half_jump(w1.h, w2.h, w2.cd - w1.cd) : half_jump(w1.h, w0.h, w1.cd - w0.cd);
        qL = q - d;
        qR = q + d;
        q6 = 0.0;
      }
      ra = qL;
      rb = qR;
      r6 = q6;
    } else if (M == M_PQM) {
      const long at = col + ncol * (long)s;
      ra = W[P_A * arr + at];
      rb = W[P_B * arr + at];
      r6 = W[P_C * arr + at];
      rd = W[P_D * arr + at];
      re = W[P_E * arr + at];
    }
  };
  auto advance = [&]() {
    ++s;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = finish(s + 2, pa);
    pa = pb;
    pb = raw(s + 4);
    zs_lo = zs_hi;
    zs_hi = zs_lo + w1.h;
    if (s < nz) recon();
  };
  recon();

  // target thicknesses arrive two targets ahead
  double hn_a = cols.dst_h(0), hn_b = (nz > 1) ? cols.dst_h(1) : 0.0;
  double zn_hi = 0.0;
  for (int k = 0; k < nz; ++k) {
    const double hn = hn_a;
    hn_a = hn_b;
    hn_b = (k + 2 < nz) ? cols.dst_h(k + 2) : 0.0;
    const long at = col + ncol * (long)k;
    const double b_old = (FIELD && budget) ? budget[at] : 0.0;
    const double zn_lo = zn_hi;
    zn_hi = zn_lo + hn;
    // the first source cell whose top lies above the target's bottom
    while (s < nz && zs_hi <= zn_lo) advance();
    const double fallback = (s < nz) ? w1.co : w0.co;  // c_own(min(ko0, nz))
    const int s_first = s;
    const double z_first = zs_lo;
    double integral = 0.0, c_lo = 0.0, c_hi = 0.0;
    int n_hit = 0;
    // filler excesses s*overlap of the contributors, in source order; more
    // than two in one target (never on a z* or z-level grid) re-walk
    double ex0 = 0.0, ex1 = 0.0;
    int n_ex = 0;
    if (hn > 0.0 && s < nz) {
      for (;;) {
        const double z_lo = dmax(zn_lo, zs_lo);
        const double z_hi = dmin(zn_hi, zs_hi);
        const double ov = z_hi - z_lo;
        if (ov > 0.0) {
          if (M == M_PCM) {
            integral = integral + w1.cd * ov;
          } else if (M == M_PLM) {
            const double xl = (z_lo - zs_lo) / w1.h, xh = (z_hi - zs_lo) / w1.h;
            integral = integral + ov * (w1.cd + ra * (xl + xh - 1.0));
          } else if (M == M_PPM || M == M_H4) {
            const double xl = (z_lo - zs_lo) / w1.h, xh = (z_hi - zs_lo) / w1.h;
            integral = integral + w1.h * ((xh - xl) * ra + 0.5 * (xh * xh - xl * xl) * (rb - ra + r6) -
                                          (xh * xh * xh - xl * xl * xl) * r6 / 3.0);
          } else if (M == M_PQM) {
            const double xl = (z_lo - zs_lo) / w1.h, xh = (z_hi - zs_lo) / w1.h;
            integral = integral + ov * (ra + rb * 0.5 * (xl + xh) + r6 * (1.0 / 3.0) * (xl * xl + xh * xh + xl * xh) +
                                        rd * 0.25 * ((xl * xl + xh * xh) * (xl + xh)) +
                                        re * 0.2 * ((xh * xh * xh + xl * xl * xl) * (xl + xh) + xl * xl * xh * xh));
          }
          if (FIELD) {
            // range of the contributor's reconstruction stencil (:402-427)
            const double a = (s > 0) ? w0.cd : w1.cd;
            const double c = (s < nz - 1) ? w2.cd : w1.cd;
            double lo = dmin(a, dmin(w1.cd, c)), hi = dmax(a, dmax(w1.cd, c));
            if (be && nz >= 2) {
              if (s == 0) {
                const double m = 2.0 * w1.cd - w2.cd;
                lo = dmin(lo, m);
                hi = dmax(hi, m);
              } else if (s == nz - 1) {
                const double m = 2.0 * w1.cd - w0.cd;
                lo = dmin(lo, m);
                hi = dmax(hi, m);
              }
            }
            if (w1.h <= HV) {
              const double sx = w1.co - w1.cd;
              lo = lo + sx;
              hi = hi + sx;
              if (sx != 0.0) {
                if (n_ex == 0)
                  ex0 = sx * ov;
                else if (n_ex == 1)
                  ex1 = sx * ov;
                ++n_ex;
              }
            }
            if (n_hit == 0) {
              c_lo = lo;
              c_hi = hi;
            } else {
              c_lo = dmin(c_lo, lo);
              c_hi = dmax(c_hi, hi);
            }
          }
          ++n_hit;
        }
        if (!(zs_hi < zn_hi) || s + 1 >= nz) break;
        advance();
      }
    }
    double q;
    if (M == M_IDENT)
      q = cd0;
    else
      q = (hn > 0.0) ? integral / hn : 0.0;
    if (!FIELD) {
      __builtin_nontemporal_store(q, &out[at]);
      continue;
    }
    // remap_add_filler_excess: each filler contributor's own excess, after
    // the division, in source order
    if (n_ex <= 2) {
      if (n_ex > 0) q = q + ex0 / hn;
      if (n_ex > 1) q = q + ex1 / hn;
    } else {
      double zj_lo = z_first;
      for (int j = s_first; j <= s; ++j) {
        const double hj = cols.src_h(j);
        const double zj_hi = zj_lo + hj;
        if (hj > 0.0 && hj <= HV) {
          const double sx = cols.val(j) - donor(j);
          if (sx != 0.0) {
            const double ov = dmin(zn_hi, zj_hi) - dmax(zn_lo, zj_lo);
            if (ov > 0.0) q = q + sx * ov / hn;
          }
        }
        zj_lo = zj_hi;
      }
    }
    const double c_raw = q;
    double cn;
    if (n_hit > 0)
      cn = isfinite(c_raw) ? dmax(c_lo, dmin(c_hi, c_raw)) : c_raw;
    else
      cn = fallback;
    __builtin_nontemporal_store(cn, &out[at]);
    if (clamp) clamp[at] = hn * (cn - c_raw);
    if (budget) {
      // the source cell of level k: from the window when it is there
      double h_k, c_k;
      const int d = k - s;
      if (d == -1) {
        h_k = w0.h;
        c_k = w0.co;
      } else if (d == 0) {
        h_k = w1.h;
        c_k = w1.co;
      } else if (d == 1) {
        h_k = w2.h;
        c_k = w2.co;
      } else if (d == 2) {
        h_k = w3.h;
        c_k = w3.co;
      } else {
        h_k = cols.src_h(k);
        c_k = cols.val(k);
      }
      const double p_new = hn * cn;
      const double p_old = h_k * c_k;
      budget[at] = b_old + (p_new - p_old);
    }
  }
}

// rescale_anomaly_ke (rdb_ocean_remap.F90:1131-1180) of a remapped face
// column: u holds u_old on entry and the result on exit, un the remap.
__global__ __launch_bounds__(TPB) void k_face_ke(FaceCols cols, long ncol, int nz, const double* __restrict__ un,
                                                 double* __restrict__ u) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  cols.bind(col);
  double h_old_sum = 0.0, h_new_sum = 0.0, mom_old = 0.0, mom_new = 0.0;
  for (int k = 0; k < nz; ++k) {
    double o, n;
    cols.pair(k, o, n);
    const long at = col + ncol * (long)k;
    h_old_sum = h_old_sum + o;
    h_new_sum = h_new_sum + n;
    mom_old = mom_old + o * u[at];
    mom_new = mom_new + n * un[at];
  }
  double scale_fac = -1.0;  // < 0: leave the remap as it is
  double u_bar_new = 0.0;
  if (!(h_old_sum <= HV || h_new_sum <= HV)) {
    const double u_bar_old = mom_old / h_old_sum;
    u_bar_new = mom_new / h_new_sum;
    double ke_old = 0.0, ke_new = 0.0;
    for (int k = 0; k < nz; ++k) {
      double o, n;
      cols.pair(k, o, n);
      const long at = col + ncol * (long)k;
      double anom = u[at] - u_bar_old;
      ke_old = ke_old + 0.5 * o * anom * anom;
      anom = un[at] - u_bar_new;
      ke_new = ke_new + 0.5 * n * anom * anom;
    }
    if (!(ke_new <= 0.0 || ke_old <= 0.0)) {
      scale_fac = sqrt(ke_old / ke_new);
      if (scale_fac > 1.25) scale_fac = 1.25;
      if (scale_fac < 0.0) scale_fac = 0.0;
    }
  }
  for (int k = 0; k < nz; ++k) {
    const long at = col + ncol * (long)k;
    u[at] = (scale_fac >= 0.0) ? u_bar_new + scale_fac * (un[at] - u_bar_new) : un[at];
  }
}

// ocean_remap_scan_preconditions (:1348-1405): per-block partial reductions.
__global__ __launch_bounds__(TPB) void k_scan(const double* __restrict__ ho, const double* __restrict__ hn, long ncol,
                                              int nz, double rel_tol, int* __restrict__ p_bad,
                                              double* __restrict__ p_rel, double* __restrict__ p_neg) {
  __shared__ int s_bad[TPB];
  __shared__ double s_rel[TPB], s_neg[TPB];
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int bad = 0;
  double rel = 0.0, hmin = 0.0;
  if (col < ncol) {
    double s_old = 0.0, s_new = 0.0;
    for (int k = 0; k < nz; ++k) {
      const double a = ho[col + ncol * (long)k], b = hn[col + ncol * (long)k];
      s_old = s_old + a;
      s_new = s_new + b;
      hmin = dmin(hmin, dmin(a, b));
    }
    rel = fabs(s_new - s_old) / dmax(dmax(fabs(s_old), fabs(s_new)), EPS);
    bad = (hmin < 0.0 || rel > rel_tol) ? 1 : 0;
  }
  s_bad[threadIdx.x] = bad;
  s_rel[threadIdx.x] = rel;
  s_neg[threadIdx.x] = hmin;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_bad[threadIdx.x] += s_bad[threadIdx.x + w];
      s_rel[threadIdx.x] = dmax(s_rel[threadIdx.x], s_rel[threadIdx.x + w]);
      s_neg[threadIdx.x] = dmin(s_neg[threadIdx.x], s_neg[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    p_bad[blockIdx.x] = s_bad[0];
    p_rel[blockIdx.x] = s_rel[0];
    p_neg[blockIdx.x] = s_neg[0];
  }
}

// ---------------------------------------------------------------------------
// host side

// Work buffers, one set per device, grown on demand and kept to process end.
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
};
constexpr int MAX_DEV = 64;
std::mutex g_mu;
DevBuf g_out[MAX_DEV], g_pqm[MAX_DEV], g_part[MAX_DEV];

int fail(int status, const char* what, hipError_t e) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  return rdb_colphys_push_error(status, buf);
}

int ensure(DevBuf* set, size_t bytes, void** p) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap: hipGetDevice", e);
  if (dev < 0 || dev >= MAX_DEV) return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "remap: device index");
  std::lock_guard<std::mutex> lk(g_mu);
  DevBuf& b = set[dev];
  if (b.n < bytes) {
    // a buffer an earlier call's kernels may still read is freed only after they end
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap: hipDeviceSynchronize", e);
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap: hipMalloc of a work buffer", e);
    b.n = bytes;
  }
  *p = b.p;
  return RDB_COLPHYS_OK;
}

// The method remap_column dispatches to for nz layers: an unknown code is
// PLM (rdb_remap_column.F90:371-372); every reconstructing method is the
// identity at nz = 1; PQM below five layers is PPM (:1245-1248), PPM and
// PPM_H4 at two layers are PLM (:608-612, :830-834).
int effective_method(int method, int nz) {
  int m = (method >= M_PCM && method <= M_PQM) ? method : M_PLM;
  if (m == M_PCM) return M_PCM;
  if (nz == 1) return M_IDENT;
  if (m == M_PQM && nz < 5) m = M_PPM;
  if ((m == M_PPM || m == M_H4) && nz == 2) m = M_PLM;
  return m;
}

bool overlaps(const void* a, const void* b, size_t bytes) {
  const char* x = static_cast<const char*>(a);
  const char* y = static_cast<const char*>(b);
  return x < y + bytes && y < x + bytes;
}

template <class Cols, bool FIELD>
int launch(const Cols& cols, long ncol, int nz, int method, int be, int nu, double* out, double* budget,
           double* clamp, hipStream_t st) {
  const int m = effective_method(method, nz);
  const unsigned grid = (unsigned)((ncol + TPB - 1) / TPB);
  double* W = nullptr;
  long arr = 0;
  if (m == M_PQM) {
    arr = (long)(nz + 1) * ncol;
    void* p = nullptr;
    int rc = ensure(g_pqm, sizeof(double) * (size_t)arr * P_NARR, &p);
    if (rc) return rc;
    W = static_cast<double*>(p);
    hipLaunchKernelGGL((k_pqm_coeffs<Cols, FIELD>), dim3(grid), dim3(TPB), 0, st, cols, ncol, nz, be, W, arr);
  }
  switch (m) {
#define RDB_REMAP_CASE(MM)                                                                                        \
  case MM:                                                                                                        \
    hipLaunchKernelGGL((k_walk<Cols, MM, FIELD>), dim3(grid), dim3(TPB), 0, st, cols, ncol, nz, be, nu, W, arr, out, \
                       budget, clamp);                                                                            \
    break;
    RDB_REMAP_CASE(M_PCM)
    RDB_REMAP_CASE(M_PLM)
    RDB_REMAP_CASE(M_PPM)
    RDB_REMAP_CASE(M_H4)
    RDB_REMAP_CASE(M_PQM)
    RDB_REMAP_CASE(M_IDENT)
#undef RDB_REMAP_CASE
    default:
      break;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap: kernel launch", e);
  return RDB_COLPHYS_OK;
}

bool dims_ok(long a, long b, long c) { return a > 0 && b > 0 && c > 0 && a * b * c < (1L << 40); }

}  // namespace

}  // namespace remap

using namespace remap;

extern "C" {

int rdb_remap_tracer_field(int method, int nx, int ny, int nz, const double* h_old, const double* h_new,
                           const double* c_in, double* c_out, int bnd_extrap, int nonunif, double* budget,
                           double* clamp_out, void* stream) {
  if (!dims_ok(nx, ny, nz)) return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_tracer_field: extents");
  if (!h_old || !h_new || !c_in || !c_out)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_tracer_field: NULL field");
  const long ncol = (long)nx * ny;
  const size_t bytes = sizeof(double) * (size_t)ncol * nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* out = c_out;
  const bool in_place = overlaps(c_in, c_out, bytes);
  if (in_place) {
    void* p = nullptr;
    int rc = ensure(g_out, bytes, &p);
    if (rc) return rc;
    out = static_cast<double*>(p);
  }
  CellCols cols{h_old, h_new, c_in, ncol, 0};
  int rc = launch<CellCols, true>(cols, ncol, nz, method, bnd_extrap ? 1 : 0, nonunif ? 1 : 0, out, budget,
                                  clamp_out, st);
  if (rc) return rc;
  if (in_place) {
    const hipError_t e = hipMemcpyAsync(c_out, out, bytes, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap_tracer_field: copy-back", e);
  }
  return RDB_COLPHYS_OK;
}

int rdb_remap_face_velocity(int axis, int nx, int ny, int nz, const double* h_old, const double* h_new, double* u,
                            int method, int conserve_ke, int zlevel_faces, int bnd_extrap, int nonunif,
                            void* stream) {
  if (axis != 0 && axis != 1) return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_face_velocity: axis");
  if (!dims_ok(nx, ny, nz)) return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_face_velocity: extents");
  if (!h_old || !h_new || !u) return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_face_velocity: NULL");
  const long nfu = axis == 0 ? nx + 1 : nx, nfv = axis == 0 ? ny : ny + 1;
  const long ncol = nfu * nfv;
  const size_t bytes = sizeof(double) * (size_t)ncol * nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  void* p = nullptr;
  int rc = ensure(g_out, bytes, &p);
  if (rc) return rc;
  double* un = static_cast<double*>(p);
  FaceCols cols{h_old, h_new, u, (long)nx * ny, ncol, nx, ny, axis, zlevel_faces ? 1 : 0, 0, 0, 0, 0};
  rc = launch<FaceCols, false>(cols, ncol, nz, method, bnd_extrap ? 1 : 0, nonunif ? 1 : 0, un, nullptr, nullptr, st);
  if (rc) return rc;
  if (conserve_ke) {
    hipLaunchKernelGGL(k_face_ke, dim3((unsigned)((ncol + TPB - 1) / TPB)), dim3(TPB), 0, st, cols, ncol, nz, un, u);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap_face_velocity: KE rescale launch", e);
  } else {
    const hipError_t e = hipMemcpyAsync(u, un, bytes, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap_face_velocity: copy-back", e);
  }
  return RDB_COLPHYS_OK;
}

int rdb_remap_columns(int method, int ncol, int nz, const double* dz_old, const double* dz_new, const double* q_old,
                      double* q_new, int bnd_extrap, int nonunif, void* stream) {
  if (!dims_ok(ncol, 1, nz)) return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_columns: extents");
  if (!dz_old || !dz_new || !q_old || !q_new)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_columns: NULL");
  const size_t bytes = sizeof(double) * (size_t)ncol * nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* out = q_new;
  const bool in_place = overlaps(q_old, q_new, bytes);
  if (in_place) {
    void* p = nullptr;
    int rc = ensure(g_out, bytes, &p);
    if (rc) return rc;
    out = static_cast<double*>(p);
  }
  CellCols cols{dz_old, dz_new, q_old, (long)ncol, 0};
  int rc = launch<CellCols, false>(cols, ncol, nz, method, bnd_extrap ? 1 : 0, nonunif ? 1 : 0, out, nullptr,
                                   nullptr, st);
  if (rc) return rc;
  if (in_place) {
    const hipError_t e = hipMemcpyAsync(q_new, out, bytes, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap_columns: copy-back", e);
  }
  return RDB_COLPHYS_OK;
}

int rdb_remap_scan_preconditions(int nx, int ny, int nz, const double* h_old, const double* h_new, double rel_tol,
                                 int* n_bad, double* worst_rel, double* worst_neg, void* stream) {
  if (!dims_ok(nx, ny, nz)) return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_scan: extents");
  if (!h_old || !h_new || !n_bad || !worst_rel || !worst_neg)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "remap_scan: NULL");
  const long ncol = (long)nx * ny;
  const long nblk = (ncol + TPB - 1) / TPB;
  hipStream_t st = static_cast<hipStream_t>(stream);
  void* p = nullptr;
  int rc = ensure(g_part, (sizeof(int) + 2 * sizeof(double)) * (size_t)nblk, &p);
  if (rc) return rc;
  double* p_rel = static_cast<double*>(p);
  double* p_neg = p_rel + nblk;
  int* p_bad = reinterpret_cast<int*>(p_neg + nblk);
  hipLaunchKernelGGL(k_scan, dim3((unsigned)nblk), dim3(TPB), 0, st, h_old, h_new, ncol, nz, rel_tol, p_bad, p_rel,
                     p_neg);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(RDB_COLPHYS_ERR_SETUP, "remap_scan: launch", e);
  std::vector<double> rel(nblk), neg(nblk);
  std::vector<int> bad(nblk);
  if ((e = hipMemcpyAsync(rel.data(), p_rel, sizeof(double) * nblk, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipMemcpyAsync(neg.data(), p_neg, sizeof(double) * nblk, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipMemcpyAsync(bad.data(), p_bad, sizeof(int) * nblk, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return fail(RDB_COLPHYS_ERR_SETUP, "remap_scan: read-back", e);
  int nb = 0;
  double wr = 0.0, wn = 0.0;
  for (long b = 0; b < nblk; ++b) {
    nb += bad[b];
    wr = rel[b] > wr ? rel[b] : wr;
    wn = neg[b] < wn ? neg[b] : wn;
  }
  *n_bad = nb;
  *worst_rel = wr;
  *worst_neg = wn;
  return RDB_COLPHYS_OK;
}

}  // extern "C"
