/*
 * colphys_oracle.c -- CPU FP64 restatement of the reference vertical-mixing
 * column path (Roundabout), op for op.
 *
 * TEST INFRASTRUCTURE.  This file is the checker that the HIP library is
 * compared against, and the timed CPU baseline.  It is imported only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; the
 * product library never links it.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (no fast-math), matching the
 * reference's CPU FMA policy (cmake/compiler_flags.cmake:335-362).  Every
 * expression keeps the reference's left-to-right Fortran evaluation order;
 * no FMA contraction anywhere, so each function is reproducible bit for bit.
 *
 * Parity pins (each function is checked by tests/test_oracle_*.py):
 *   orc_cbrt            - the reference's 48 golden bit rows
 *                         (tests/test_safe_math_golden.inc:557-606) and exact
 *                         cubes (tests/test_safe_math.F90:460-464)
 *   orc_eos_wright      - known answers (tests/test_ocean_wright_eos.F90),
 *                         vanished layer -> rho0
 *   orc_pp81            - closed forms (tests/test_ocean_pp81.F90)
 *   orc_kpp_overlay     - unstratified h_b = H, shape function, analytic
 *                         convective w* / B0 (tests/test_ocean_kpp.F90,
 *                         tests/test_ocean_buoyancy_flux.F90:250-346),
 *                         V_t^2 monotonicity, collapsed top layer
 *   orc_assemble        - default pass-through, floors/ceilings, guard
 *   orc_build_factorize / orc_apply_factored / orc_vdiff_tracers
 *                       - dense LU brute force, conservation of sum(h~ c),
 *                         constancy, two-pass == one-pass when kt == ks
 *   orc_nonlocal        - telescoping column sum, zero under B0 >= 0
 *   orc_surface_flux    - closed form deposit on the first live layer
 *   orc_column_step     - composition of the above in the reference order
 */
#include "colphys_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, k) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * (size_t)(k)))
#define IDXU(i, j, k) ((size_t)(i) + (size_t)(nx + 1) * ((size_t)(j) + (size_t)ny * (size_t)(k)))
#define IDXV(i, j, k) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)(ny + 1) * (size_t)(k)))
#define IDX2(i, j) ((size_t)(i) + (size_t)nx * (size_t)(j))

/* Fortran max/min on finite reals. */
static inline double fmx(double a, double b) { return (a > b) ? a : b; }
static inline double fmn(double a, double b) { return (a < b) ? a : b; }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------
 * Cube root, Kahan (1991) / Muller (2016 ch. 5): exponent split modulo 3, a
 * degree-4 seed on the mantissa part m in [1/2, 4), two Halley steps
 * y <- y + y (m - y^3) / (2 y^3 + m).  Follows rdb_sm_cbrt
 * (src/shared_module_utilities/rdb_safe_math_pow.inc:107-142).  The seed
 * polynomial's coefficients must equal the reference's table
 * (rdb_safe_math_consts.inc:104-106) for the result to agree bit for bit.
 * ---------------------------------------------------------------------- */
static const double CBRT_SEED[5] = {0.5189978361037301813, 0.6542973117201534849,
                                    -0.2158001895775201462, 0.04437845996234061896,
                                    -0.003660165019409192949};

double orc_cbrt(double x) {
  if (isnan(x) || !isfinite(x)) return x;
  double ax = fabs(x);
  if (ax < 2.2250738585072014e-308) return copysign(0.0, x); /* tiny(1d0) */
  int e;
  double f = frexp(ax, &e); /* ax = f * 2^e, f in [1/2, 1): Fortran fraction/exponent */
  int rem = ((e % 3) + 3) % 3;
  int q = (e - rem) / 3;
  double m = ldexp(f, rem);
  double y = CBRT_SEED[4];
  for (int i = 3; i >= 0; --i) y = y * m + CBRT_SEED[i];
  for (int it = 0; it < 2; ++it) {
    double y3 = (y * y) * y;
    y = y + y * ((m - y3) / ((y3 + y3) + m));
  }
  y = ldexp(y, q);
  if (x < 0.0) y = -y;
  return y;
}

/* ------------------------------------------------------------------------
 * A1 Wright (1997) reduced-range EOS at scalar p_ref; vanished layer -> rho0.
 * rdb_eos.F90:909-959, coefficients :203-217.
 * ---------------------------------------------------------------------- */
#define WA0 7.057924e-4
#define WA1 3.480336e-7
#define WA2 -1.112733e-7
#define WB0 5.790749e8
#define WB1 3.516535e6
#define WB2 -4.002714e4
#define WB3 2.084372e2
#define WB4 5.944068e5
#define WB5 -9.643486e3
#define WC0 1.704853e5
#define WC1 7.904722e2
#define WC2 -7.984422e0
#define WC3 5.140652e-2
#define WC4 -2.302158e2
#define WC5 -3.079464e0

static inline double eos_wright_point(double T, double S, double h, double rho0, double p_ref) {
  if (h > ORC_H_VANISHED) {
    double Tsq = T * T;
    double Tcu = Tsq * T;
    double alpha0 = WA0 + WA1 * T + WA2 * S;
    double p0 = WB0 + WB1 * T + WB2 * Tsq + WB3 * Tcu + WB4 * S + WB5 * S * T;
    double lambda = WC0 + WC1 * T + WC2 * Tsq + WC3 * Tcu + WC4 * S + WC5 * S * T;
    double ppp0 = p_ref + p0;
    double denom = lambda + alpha0 * ppp0;
    return ppp0 / denom;
  }
  return rho0;
}

void orc_eos_wright(int nx, int ny, int nz, const double *h, const double *S,
                    const double *T, double *rho, double rho0, double p_ref) {
  size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n; ++p) rho[p] = eos_wright_point(T[p], S[p], h[p], rho0, p_ref);
}

/* A16 |tau| at cell centres: rdb_ocean_surface_stress.F90:858-875. */
void orc_stress_mag(int nx, int ny, const double *tau_x, const double *tau_y,
                    double *stress_mag) {
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double tx = 0.5 * (tau_x[(size_t)i + (size_t)(nx + 1) * j] + tau_x[(size_t)i + 1 + (size_t)(nx + 1) * j]);
      double ty = 0.5 * (tau_y[(size_t)i + (size_t)nx * j] + tau_y[(size_t)i + (size_t)nx * (j + 1)]);
      stress_mag[IDX2(i, j)] = sqrt(tx * tx + ty * ty);
    }
}

/* ------------------------------------------------------------------------
 * A3 PP81: rdb_ocean_vmix.F90:1007-1074.
 * ---------------------------------------------------------------------- */
void orc_pp81(const rdb_colphys_config *cfg, int nx, int ny, int nz,
              const double *u, const double *v, const double *h,
              const double *rho, double *kv, double *kt) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      kv[IDX(i, j, 0)] = 0.0;
      kt[IDX(i, j, 0)] = 0.0;
      kv[IDX(i, j, nz)] = 0.0;
      kt[IDX(i, j, nz)] = 0.0;
      for (int k = 1; k < nz; ++k) { /* Fortran k = 2..nz */
        double u_km1 = 0.5 * (u[IDXU(i, j, k - 1)] + u[IDXU(i + 1, j, k - 1)]);
        double v_km1 = 0.5 * (v[IDXV(i, j, k - 1)] + v[IDXV(i, j + 1, k - 1)]);
        double u_k = 0.5 * (u[IDXU(i, j, k)] + u[IDXU(i + 1, j, k)]);
        double v_k = 0.5 * (v[IDXV(i, j, k)] + v[IDXV(i, j + 1, k)]);
        double dz_face = 0.5 * (h[IDX(i, j, k - 1)] + h[IDX(i, j, k)]);
        if (dz_face <= 0.0) {
          kv[IDX(i, j, k)] = cfg->pp81_nu_bg;
          kt[IDX(i, j, k)] = cfg->pp81_kappa_bg;
          continue;
        }
        double du_dz = (u_k - u_km1) / dz_face;
        double dv_dz = (v_k - v_km1) / dz_face;
        double shear2 = fmx(du_dz * du_dz + dv_dz * dv_dz, cfg->shear2_floor);
        double n2 = -(ORC_GRAVITY * (rho[IDX(i, j, k)] - rho[IDX(i, j, k - 1)]) / (cfg->rho0 * dz_face));
        double ri = n2 / shear2;
        double denom = 1.0 + cfg->pp81_alpha * fmx(ri, 0.0);
        double ri_factor = 1.0 / (denom * denom);
        kv[IDX(i, j, k)] = cfg->pp81_nu_bg + cfg->pp81_nu0 * ri_factor;
        kt[IDX(i, j, k)] = cfg->pp81_kappa_bg + cfg->pp81_nu0 * ri_factor / denom;
      }
    }
}

/* rdb_vl_first_live: rdb_vanished_layer.inc:66-103 (0-based kstart). */
static inline int first_live(const double *h, int nx, int ny, int i, int j, int kstart) {
  for (int k = kstart; k >= 0; --k)
    if (h[IDX(i, j, k)] > ORC_H_VANISHED) return k;
  return kstart;
}

/* A6 kpp_surface_buoyancy_flux: rdb_ocean_vmix.F90:1518-1559. */
static inline double kpp_b0(double alpha_T, double beta_S, double rho0, double qT, double qS) {
  return (ORC_GRAVITY / rho0) * (alpha_T * qT - beta_S * qS);
}

/* ------------------------------------------------------------------------
 * A4/A5 KPP overlay: rdb_ocean_vmix.F90:1121-1516 (shortwave inactive,
 * constant buoyancy coefficients).  kbl receives the 1-based layer at which
 * the bulk-Ri crossing fired (0 when the sweep reached the bed).
 * ---------------------------------------------------------------------- */
void orc_kpp_overlay(const rdb_colphys_config *cfg, int nx, int ny, int nz,
                     const double *u, const double *v, const double *h,
                     const double *rho, const double *stress_mag,
                     const double *stress_shelf, const double *Q_heat,
                     const double *Q_salt, double *bl_depth, double *kv,
                     double *kt, double *gamma_t, double *gamma_s, double *b0,
                     int *kbl) {
  const double rho0 = cfg->rho0;
  size_t n3 = (size_t)nx * ny * (nz + 1);
  for (size_t p = 0; p < n3; ++p) {
    gamma_t[p] = 0.0;
    gamma_s[p] = 0.0;
  }
  /* Pass 1: boundary-layer depth (:1287-1426). */
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int k_sfc = first_live(h, nx, ny, i, j, nz - 1);
      double u_ref = 0.5 * (u[IDXU(i, j, k_sfc)] + u[IDXU(i + 1, j, k_sfc)]);
      double v_ref = 0.5 * (v[IDXV(i, j, k_sfc)] + v[IDXV(i, j + 1, k_sfc)]);
      double b_ref = -(ORC_GRAVITY * rho[IDX(i, j, k_sfc)] / rho0);
      double d_centre_ref = 0.5 * h[IDX(i, j, k_sfc)];
      double shelf = stress_shelf ? stress_shelf[IDX2(i, j)] : 0.0;
      double tau_mag = stress_mag[IDX2(i, j)] + shelf;
      double u_star = sqrt(tau_mag / rho0);
      double h_b_lagged = bl_depth[IDX2(i, j)];
      double qT = Q_heat[IDX2(i, j)] / (rho0 * cfg->cp);
      double qS = Q_salt[IDX2(i, j)] / rho0;
      double B0 = kpp_b0(cfg->alpha_T, cfg->beta_S, rho0, qT, qS);
      double wstar3_lagged = fmx(0.0, -B0) * h_b_lagged;
      double cb = orc_cbrt(wstar3_lagged);
      double w_s_col = sqrt(u_star * u_star + cb * cb);

      double ri_prev = 0.0, d_prev = d_centre_ref, h_b = 0.0;
      int crossing = 0, kcross = 0;
      double d_centre_k = d_centre_ref;
      for (int k = k_sfc - 1; k >= 0; --k) {
        if (crossing) break; /* the reference keeps looping without effect */
        d_centre_k = d_centre_k + 0.5 * (h[IDX(i, j, k + 1)] + h[IDX(i, j, k)]);
        double u_k = 0.5 * (u[IDXU(i, j, k)] + u[IDXU(i + 1, j, k)]);
        double v_k = 0.5 * (v[IDXV(i, j, k)] + v[IDXV(i, j + 1, k)]);
        double b_k = -(ORC_GRAVITY * rho[IDX(i, j, k)] / rho0);
        double delta_d = d_centre_k - d_centre_ref;
        double delta_b = b_ref - b_k;
        double n_brunt = sqrt(fmx(0.0, delta_b / delta_d));
        double vt2 = (cfg->c_vt2 / cfg->ri_crit) * d_centre_k * n_brunt * w_s_col;
        double du = u_k - u_ref, dv = v_k - v_ref;
        double shear2 = fmx(du * du + dv * dv + vt2, cfg->shear2_floor);
        double ri_bulk = delta_b * delta_d / shear2;
        if (ri_bulk >= cfg->ri_crit) {
          double denom = ri_bulk - ri_prev;
          double frac = (fabs(denom) > 1.0e-12) ? (cfg->ri_crit - ri_prev) / denom : 0.5;
          h_b = d_prev + frac * (d_centre_k - d_prev);
          crossing = 1;
          kcross = k + 1;
        }
        ri_prev = ri_bulk;
        d_prev = d_centre_k;
      }
      if (!crossing) {
        h_b = 0.0;
        for (int k = 0; k <= k_sfc; ++k) h_b = h_b + h[IDX(i, j, k)];
      }
      bl_depth[IDX2(i, j)] = h_b;
      if (kbl) kbl[IDX2(i, j)] = kcross;
    }
  /* Pass 2: overlay inside the boundary layer (:1428-1515). */
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int k_sfc = first_live(h, nx, ny, i, j, nz - 1);
      double shelf = stress_shelf ? stress_shelf[IDX2(i, j)] : 0.0;
      double tau_mag = stress_mag[IDX2(i, j)] + shelf;
      double u_star = sqrt(tau_mag / rho0);
      double h_b = bl_depth[IDX2(i, j)];
      double qT = Q_heat[IDX2(i, j)] / (rho0 * cfg->cp);
      double qS = Q_salt[IDX2(i, j)] / rho0;
      double B0 = kpp_b0(cfg->alpha_T, cfg->beta_S, rho0, qT, qS);
      b0[IDX2(i, j)] = B0;
      int destab = (B0 < 0.0);
      double wstar3 = fmx(0.0, -B0) * h_b;
      double w_star = orc_cbrt(wstar3);
      double w_s = sqrt(u_star * u_star + w_star * w_star);
      if (h_b > 0.0 && w_s > 0.0) {
        double d_running = 0.0;
        for (int k = k_sfc; k >= 0; --k) {
          d_running = d_running + h[IDX(i, j, k)];
          double d_face_k = d_running;
          if (k > 0 && d_face_k < h_b) {
            double sigma = d_face_k / h_b;
            double g_shape = sigma * (1.0 - sigma) * (1.0 - sigma);
            double kv_kpp = h_b * w_s * g_shape;
            kv[IDX(i, j, k)] = fmx(kv[IDX(i, j, k)], kv_kpp);
            kt[IDX(i, j, k)] = fmx(kt[IDX(i, j, k)], kv_kpp);
            if (destab) {
              double gf = cfg->cs_nonlocal * g_shape;
              gamma_t[IDX(i, j, k)] = gf * qT;
              gamma_s[IDX(i, j, k)] = gf * qS;
            }
          }
        }
      }
    }
}

/* A9 split (ddiff off): ks := kt on all interfaces, rdb_ocean_vmix.F90:1921-1934. */
void orc_split(int nx, int ny, int nz, const double *kt, double *ks) {
  memcpy(ks, kt, sizeof(double) * (size_t)nx * ny * (nz + 1));
}

/* A10 vmix_assemble (scalar floors, no smoothing): guard then clip,
 * rdb_ocean_vmix.F90:2265-2277, 2368-2385, 2750-2769. */
int orc_assemble(const rdb_colphys_config *cfg, int nx, int ny, int nz,
                 double *kv, double *kt, double *ks) {
  int status = 0;
  if (cfg->vmix_guard) {
    size_t n3 = (size_t)nx * ny * (nz + 1);
    for (size_t p = 0; p < n3; ++p) {
      double a = kv[p], b = kt[p], c = ks[p];
      if (isnan(a) || isnan(b) || isnan(c) || a < 0.0 || b < 0.0 || c < 0.0) {
        status = RDB_COLPHYS_GUARD_TRIPPED;
        break;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int k = 1; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t p = IDX(i, j, k);
        kv[p] = fmn(fmx(kv[p], cfg->kv_bg), cfg->kv_max);
        kt[p] = fmn(fmx(kt[p], cfg->kt_bg), cfg->kd_max);
        ks[p] = fmn(fmx(ks[p], cfg->ks_bg), cfg->kd_max);
      }
  return status;
}

/* rdb_conc_add_source: rdb_tracer_conc.inc:33-50. */
static inline double conc_add_source(double c, double h, double src) {
  if (h > ORC_H_VANISHED) return c + src / (h + ORC_H_DIV_EPS);
  return c;
}

/* A2 surface deposit: apply_surface_src_2d_impl,
 * rdb_ocean_surface_flux.F90:1093-1142, driven as in :959-1070. */
void orc_surface_flux(const rdb_colphys_config *cfg, int nx, int ny, int nz,
                      double dt, double *T, double *S, const double *h,
                      const double *wet, const double *Q_heat,
                      const double *Q_salt, const int *k_top) {
  if (!cfg->has_heat && !cfg->has_salt) return;
  double inv_heat = dt / (cfg->rho0 * cfg->cp);
  double inv_salt = dt / cfg->rho0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int ks = k_top ? k_top[IDX2(i, j)] - 1 : nz - 1;
      int kt = first_live(h, nx, ny, i, j, ks);
      double hk = h[IDX(i, j, kt)];
      if (cfg->has_heat) {
        double cell = inv_heat * Q_heat[IDX2(i, j)] * wet[IDX2(i, j)];
        T[IDX(i, j, kt)] = conc_add_source(T[IDX(i, j, kt)], hk, cell);
      }
      if (cfg->has_salt) {
        double cell = inv_salt * Q_salt[IDX2(i, j)] * wet[IDX2(i, j)];
        S[IDX(i, j, kt)] = conc_add_source(S[IDX(i, j, kt)], hk, cell);
      }
    }
}

/* face_thick: rdb_ocean_vdiff.F90:1928-1962. */
static inline double face_thick(double ha, double hb, int use_harmonic) {
  if (use_harmonic) {
    double sum_h = ha + hb;
    if (sum_h > 1.0e-30) return 2.0 * ha * hb / sum_h;
    return 0.5 * sum_h;
  }
  return 0.5 * (ha + hb);
}

/* A12 build_factorize_tracer_matrix: rdb_ocean_vdiff.F90:1248-1349. */
void orc_build_factorize(int nx, int ny, int nz, int use_harmonic, double dt,
                         const double *K, const double *h, double *a,
                         double *b, double *c) {
  const double HV = ORC_H_VANISHED;
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double hc = fmx(h[IDX(i, j, 0)], HV);
      double alpha = 0.0;
      if (nz > 1) {
        double hp = fmx(h[IDX(i, j, 1)], HV);
        double dz_face = face_thick(hc, hp, use_harmonic);
        alpha = dt * K[IDX(i, j, 1)] / (hc * dz_face);
        if (h[IDX(i, j, 0)] <= HV || h[IDX(i, j, 1)] <= HV) alpha = 0.0;
      }
      a[IDX(i, j, 0)] = 0.0;
      c[IDX(i, j, 0)] = -alpha;
      b[IDX(i, j, 0)] = 1.0 + alpha;
      for (int k = 1; k < nz - 1; ++k) {
        double hm = fmx(h[IDX(i, j, k - 1)], HV);
        hc = fmx(h[IDX(i, j, k)], HV);
        double hp = fmx(h[IDX(i, j, k + 1)], HV);
        double beta = dt * K[IDX(i, j, k)] / (hc * face_thick(hm, hc, use_harmonic));
        alpha = dt * K[IDX(i, j, k + 1)] / (hc * face_thick(hc, hp, use_harmonic));
        if (h[IDX(i, j, k)] <= HV || h[IDX(i, j, k - 1)] <= HV) beta = 0.0;
        if (h[IDX(i, j, k)] <= HV || h[IDX(i, j, k + 1)] <= HV) alpha = 0.0;
        a[IDX(i, j, k)] = -beta;
        c[IDX(i, j, k)] = -alpha;
        b[IDX(i, j, k)] = 1.0 + alpha + beta;
      }
      if (nz > 1) {
        int k = nz - 1;
        double hm = fmx(h[IDX(i, j, k - 1)], HV);
        hc = fmx(h[IDX(i, j, k)], HV);
        double dz_face = face_thick(hm, hc, use_harmonic);
        double beta = dt * K[IDX(i, j, k)] / (hc * dz_face);
        if (h[IDX(i, j, k)] <= HV || h[IDX(i, j, k - 1)] <= HV) beta = 0.0;
        a[IDX(i, j, k)] = -beta;
        c[IDX(i, j, k)] = 0.0;
        b[IDX(i, j, k)] = 1.0 + beta;
      }
      c[IDX(i, j, 0)] = c[IDX(i, j, 0)] / b[IDX(i, j, 0)];
      for (int k = 1; k < nz; ++k) {
        double denom = b[IDX(i, j, k)] - a[IDX(i, j, k)] * c[IDX(i, j, k - 1)];
        c[IDX(i, j, k)] = c[IDX(i, j, k)] / denom;
        b[IDX(i, j, k)] = denom;
      }
    }
}

/* A13 apply_factored_tracer (no budget): rdb_ocean_vdiff.F90:1351-1414. */
void orc_apply_factored(int nx, int ny, int nz, double *x, const double *a,
                        const double *b, const double *c, double *rhs) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      for (int k = 0; k < nz; ++k) rhs[IDX(i, j, k)] = x[IDX(i, j, k)];
      rhs[IDX(i, j, 0)] = rhs[IDX(i, j, 0)] / b[IDX(i, j, 0)];
      for (int k = 1; k < nz; ++k)
        rhs[IDX(i, j, k)] = (rhs[IDX(i, j, k)] - a[IDX(i, j, k)] * rhs[IDX(i, j, k - 1)]) / b[IDX(i, j, k)];
      for (int k = nz - 2; k >= 0; --k)
        rhs[IDX(i, j, k)] = rhs[IDX(i, j, k)] - c[IDX(i, j, k)] * rhs[IDX(i, j, k + 1)];
      for (int k = 0; k < nz; ++k) x[IDX(i, j, k)] = rhs[IDX(i, j, k)];
    }
}

/* A11 vdiff_apply_tracers, two-pass dispatch (kt -> T; ks -> S + passives):
 * rdb_ocean_vdiff.F90:1107-1226. */
/* The budget of apply_factored_tracer (rdb_ocean_vdiff.F90:1397-1405):
 * budget += h*c_new - h*c_old per cell, the two products standalone. */
static void add_budget(size_t n, const double *h, const double *c_old, const double *c_new, double *budget) {
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n; ++q) {
    double p_new = h[q] * c_new[q];
    double p_old = h[q] * c_old[q];
    budget[q] = budget[q] + (p_new - p_old);
  }
}

/* A11 with the heat/salt budget slots (NULL: not kept). */
void orc_vdiff_tracers_budget(int nx, int ny, int nz, int use_harmonic, double dt, const double *kt,
                              const double *ks, const double *h, double *T, double *S, int n_passive,
                              double **passive, double *heat_budget, double *salt_budget) {
  size_t n = (size_t)nx * ny * nz;
  double *old = (heat_budget || salt_budget) ? (double *)malloc(sizeof(double) * n) : NULL;
  double *a = (double *)malloc(sizeof(double) * n);
  double *b = (double *)malloc(sizeof(double) * n);
  double *c = (double *)malloc(sizeof(double) * n);
  double *rhs = (double *)malloc(sizeof(double) * n);
  orc_build_factorize(nx, ny, nz, use_harmonic, dt, kt, h, a, b, c);
  if (heat_budget) memcpy(old, T, sizeof(double) * n);
  orc_apply_factored(nx, ny, nz, T, a, b, c, rhs);
  if (heat_budget) add_budget(n, h, old, T, heat_budget);
  orc_build_factorize(nx, ny, nz, use_harmonic, dt, ks, h, a, b, c);
  if (salt_budget) memcpy(old, S, sizeof(double) * n);
  orc_apply_factored(nx, ny, nz, S, a, b, c, rhs);
  if (salt_budget) add_budget(n, h, old, S, salt_budget);
  for (int t = 0; t < n_passive; ++t) orc_apply_factored(nx, ny, nz, passive[t], a, b, c, rhs);
  free(old);
  free(a);
  free(b);
  free(c);
  free(rhs);
}

void orc_vdiff_tracers(int nx, int ny, int nz, int use_harmonic, double dt,
                       const double *kt, const double *ks, const double *h,
                       double *T, double *S, int n_passive, double **passive) {
  orc_vdiff_tracers_budget(nx, ny, nz, use_harmonic, dt, kt, ks, h, T, S, n_passive, passive, NULL, NULL);
}

/* nonlocal_face_flux: rdb_ocean_vmix.F90:1694-1714. */
static inline double nl_face(double gamma, double h_below, double h_above) {
  return (h_below > ORC_H_VANISHED && h_above > ORC_H_VANISHED) ? gamma : 0.0;
}

/* A14 vmix_apply_nonlocal_tendencies (T with gamma_t, S with gamma_s; no
 * pseudo-salt registered): rdb_ocean_vmix.F90:1605-1692. */
void orc_nonlocal(int nx, int ny, int nz, double dt, const double *gamma_t,
                  const double *gamma_s, const double *h, double *T,
                  double *S) {
  for (int pass = 0; pass < 2; ++pass) {
    const double *g = pass == 0 ? gamma_t : gamma_s;
    double *x = pass == 0 ? T : S;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
          double f_up = 0.0, f_dn = 0.0;
          if (k < nz - 1) f_up = nl_face(g[IDX(i, j, k + 1)], h[IDX(i, j, k)], h[IDX(i, j, k + 1)]);
          if (k > 0) f_dn = nl_face(g[IDX(i, j, k)], h[IDX(i, j, k - 1)], h[IDX(i, j, k)]);
          x[IDX(i, j, k)] = conc_add_source(x[IDX(i, j, k)], h[IDX(i, j, k)], dt * (f_up - f_dn));
        }
  }
}

/* One column step: EOS -> vmix_build_coefficients -> surface flux ->
 * vdiff_apply_tracers -> non-local, the order of apply_deferred_thermo
 * (rdb_ocean_dyn.F90:3836-3856) with rho refreshed first.  The coefficient
 * build follows vmix_build_coefficients (rdb_ocean_dyn.F90:1473-1634): PP81,
 * then the KPP overlay or the EPBL solve and merge (EPBL turns KPP off,
 * rdb_ocean_setup_vmix.F90:552-719), KV_ML_INVZ2, convective adjustment
 * against the live boundary-layer depth of whichever surface scheme ran,
 * the heat/salt split (double diffusion when on), the assembly gate. */
int orc_column_step(const rdb_colphys_config *cfg, orc_state *st, double dt) {
  int nx = st->nx, ny = st->ny, nz = st->nz;
  int epbl = cfg->epbl_enable != 0;
  int kpp = cfg->use_kpp && !epbl;
  orc_eos_wright(nx, ny, nz, st->h, st->S, st->T, st->rho, cfg->rho0, cfg->p_ref);
  orc_pp81(cfg, nx, ny, nz, st->u, st->v, st->h, st->rho, st->kv, st->kt);
  if (kpp)
    orc_kpp_overlay(cfg, nx, ny, nz, st->u, st->v, st->h, st->rho, st->stress_mag,
                    st->stress_shelf, st->Q_heat, st->Q_salt, st->bl_depth, st->kv,
                    st->kt, st->gamma_t, st->gamma_s, st->b0, st->kbl);
  if (epbl)
    orc_epbl_compute_merge(cfg, nx, ny, nz, dt, st->h, st->T, st->S, st->stress_mag, st->stress_shelf,
                           st->Q_heat, st->Q_salt, st->wet, st->coriolis, st->kd_epbl, st->mld, st->b0_epbl,
                           st->la, cfg->epbl_tke_diags ? st->tke : NULL, st->kv, st->kt);
  orc_vmix_invz2(cfg, nx, ny, nz, st->h, st->kv);
  if (cfg->conv_enable)
    orc_vmix_convection(cfg, nx, ny, nz, st->h, st->rho, epbl ? st->mld : st->bl_depth, st->kv, st->kt);
  if (cfg->ddiff_enable)
    orc_vmix_split_ddiff(cfg, nx, ny, nz, st->h, st->T, st->S, st->kt, st->ks);
  else
    orc_split(nx, ny, nz, st->kt, st->ks);
  int status = orc_assemble(cfg, nx, ny, nz, st->kv, st->kt, st->ks);
  orc_surface_flux(cfg, nx, ny, nz, dt, st->T, st->S, st->h, st->wet, st->Q_heat,
                   st->Q_salt, st->k_top);
  orc_vdiff_tracers_budget(nx, ny, nz, cfg->use_harmonic, dt, st->kt, st->ks, st->h, st->T,
                           st->S, st->n_passive, st->passive, st->heat_budget, st->salt_budget);
  if (kpp) orc_nonlocal(nx, ny, nz, dt, st->gamma_t, st->gamma_s, st->h, st->T, st->S);
  return status;
}

void orc_totals(int nx, int ny, int nz, const double *h, const double *c,
                const double *wet, double *total) {
  long double acc = 0.0L;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        if (!wet || wet[IDX2(i, j)] > 0.5) acc += (long double)h[IDX(i, j, k)] * (long double)c[IDX(i, j, k)];
  *total = (double)acc;
}
