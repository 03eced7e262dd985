/*
 * epbl_oracle.c -- CPU FP64 restatement of the vertical-mixing contributors
 * that are off at the reference defaults: the energetics-based planetary
 * boundary layer (EPBL), KV_ML_INVZ2, convective adjustment and double
 * diffusion, with the position-independent exp/log/pow/erfc they call.
 *
 * TEST INFRASTRUCTURE.  The checker the HIP library is compared against;
 * imported only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg.  Built like colphys_oracle.c (gcc -O2
 * -ffp-contract=off): every expression keeps the reference's left-to-right
 * Fortran evaluation order, integer powers expand as gfortran's powi does
 * (x**3 = (x*x)*x, x**5 = ((x*x)*x)*(x*x)).
 *
 * Parity pins (tests/test_oracle_epbl.py, tests/test_oracle_contrib.py):
 *   orc_sm_exp / orc_sm_log / orc_sm_pow / orc_sm_erfc
 *                         - the reference's golden bit rows G_EXP, G_LOG,
 *                           G_POW, G_POWBIG, G_ERFC
 *                           (tests/test_safe_math_golden.inc), 48 each
 *   orc_specvol_derivs    - central finite differences of the Wright
 *                           specific volume (tests/test_ocean_epbl.F90)
 *   orc_epbl_find_mstar   - the OM4 / RH18 closed forms, the cap, the
 *                           convective reduction
 *   orc_epbl_lf17_*       - the reference's LF17 spot values (U10, La)
 *   orc_epbl_lt_enhance   - the hand value 1 + 0.447 * 0.3^-1.33, the cap
 *   orc_epbl_column       - the closed TKE ledger, the Kraus-Turner wind
 *                           deepening law, stable shoaling, convective
 *                           deepening, collapsed top layers
 *                           (tests/test_ocean_epbl.F90); the energy of the
 *                           final diffusivities against direct implicit
 *                           solves of the column (Wright EOS, both the
 *                           mixing and the convective branch); the closed
 *                           form of an unstratified column
 *                           Parity unpinned inside it (restated from the
 *                           reference's text, no check can tell a variant
 *                           apart): the rotational reduction of nstar
 *                           (:1374-1378), the restart of the mixing length
 *                           and the disconnect after a partly mixed layer
 *                           (:1566, :1606-1611), the fall-back to kd_g0 when
 *                           the convective diffusivity would cost energy
 *                           (:1535-1538) and the MLD root-finder's step rule
 *                           (:1625-1641).
 *   orc_vmix_invz2 / orc_vmix_convection / orc_vmix_split_ddiff
 *                         - the reference's closed forms
 *                           (tests/test_ocean_kv_ml_invz2.F90,
 *                           tests/test_ocean_convection.F90,
 *                           tests/test_ocean_double_diffusion.F90)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "colphys_oracle.h"

#define IDX(i, j, k) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * (size_t)(k)))
#define IDX2(i, j) ((size_t)(i) + (size_t)nx * (size_t)(j))

static inline double fmx(double a, double b) { return (a > b) ? a : b; }
static inline double fmn(double a, double b) { return (a < b) ? a : b; }

/* ------------------------------------------------------------------------
 * Position-independent exp / log / pow / erfc.  The polynomial tables and
 * the Cody-Waite split of ln 2 are the reference's
 * (src/shared_module_utilities/rdb_safe_math_consts.inc:29-128): the bits of
 * every result depend on them.
 * ---------------------------------------------------------------------- */
#define SM_TINY 2.2250738585072014e-308
#define SM_LN2_HI 0.69314718060195446014
#define SM_LN2_LO -4.2009150726810845979e-11
#define SM_INV_LN2 1.442695040888963387
#define SM_EXP_MAX 709.78271289338399679
#define SM_EXP_MIN -745.1332191019412
#define SM_SQRT_HALF 0.70710678118654757274
#define SM_TWO54 18014398509481984.0
#define SM_SPLIT 134217729.0
#define SM_INV_SQRTPI 0.56418958354775627928

static const double SM_EXP_C[14] = {1.0,        1.0,          1.0 / 2.0,        1.0 / 6.0,        1.0 / 24.0,
                                    1.0 / 120.0, 1.0 / 720.0, 1.0 / 5040.0,     1.0 / 40320.0,    1.0 / 362880.0,
                                    1.0 / 3628800.0, 1.0 / 39916800.0, 1.0 / 479001600.0, 1.0 / 6227020800.0};
static const double SM_LOG_C[10] = {1.0,       1.0 / 3.0,  1.0 / 5.0,  1.0 / 7.0,  1.0 / 9.0,
                                    1.0 / 11.0, 1.0 / 13.0, 1.0 / 15.0, 1.0 / 17.0, 1.0 / 19.0};
static const double SM_ERF_C[12] = {1.128379167095512559,    -0.376126389031837538,    0.1128379167095512614,
                                    -0.02686617064513124181, 0.005223977625441806293,  -0.0008548327023363942604,
                                    0.0001205533296923066544, -0.00001492564916290355896, 1.64620383639705829e-6,
                                    -1.636264690374603956e-7, 1.472152117638980096e-8, -1.095865453884735107e-9};
static const double SM_ERFC_A_P[7] = {0.9999999993850857694, 1.248626834329117141,  0.7867094724888269441,
                                      0.2819428661917610102, 0.05719027224971907214, 0.005247572786535052193,
                                      4.820036176802313081e-8};
static const double SM_ERFC_A_Q[7] = {1.0,                  2.377005991883625935, 2.468873583018134976,
                                      1.443014856236056964, 0.5046949490889137691, 0.1013327241858480304,
                                      0.009303534974169982985};
static const double SM_ERFC_B_P[7] = {0.9999921569124016196, 1.494269944137282868,  1.086387126830294259,
                                      0.4598879010194165118, 0.1135948493973390344, 0.01390038004034872994,
                                      1.000103316758091609e-9};
static const double SM_ERFC_B_Q[7] = {1.0,                  2.622591678871410537, 3.045871297831376889,
                                      2.025960043609839722, 0.8274770365185684096, 0.2013396690397278865,
                                      0.02463786793791475013};
static const double SM_ERFC_C_P[6] = {-0.4999999999999982792, -12.20993252402114315, -95.49719132910308872,
                                      -273.5933562329464053,  -231.8021338423288,    -12.38208675658035851};
static const double SM_ERFC_C_Q[6] = {1.0, 25.91986504803799107, 226.1241802321382579, 802.2984884770494318,
                                      1100.222105092567517, 428.3053107970796418};

/* rdb_sm_two_sum / rdb_sm_two_prod (rdb_safe_math_core.inc): Knuth's exact
 * sum and Dekker's exact product with a Veltkamp split. */
static inline void two_sum(double a, double b, double *s, double *e) {
  *s = a + b;
  double bb = *s - a;
  *e = (a - (*s - bb)) + (b - bb);
}

static inline void two_prod(double a, double b, double *p, double *e) {
  double c = SM_SPLIT * a;
  double ah = c - (c - a);
  double al = a - ah;
  c = SM_SPLIT * b;
  double bh = c - (c - b);
  double bl = b - bh;
  *p = a * b;
  *e = (((ah * bh - *p) + ah * bl) + al * bh) + al * bl;
}

/* rdb_sm_exp_core: exp(x + c) for a small correction c.  k = anint(x/ln2)
 * (round half away), r = x - k ln2_hi - k ln2_lo + c, degree-13 Taylor
 * polynomial by Horner, scale by 2^k, results below tiny flushed to 0. */
double orc_sm_exp_core(double x, double c) {
  double kr = round(x * SM_INV_LN2);
  int k = (int)kr;
  double r = x - kr * SM_LN2_HI;
  r = r - kr * SM_LN2_LO;
  r = r + c;
  double p = SM_EXP_C[13];
  for (int i = 12; i >= 0; --i) p = p * r + SM_EXP_C[i];
  double y = (k > 1000) ? 2.0 * ldexp(p, k - 1) : ldexp(p, k);
  if (y < SM_TINY) y = 0.0;
  return y;
}

double orc_sm_exp(double x) {
  if (isnan(x)) return x;
  if (x > SM_EXP_MAX) return INFINITY;
  if (x < SM_EXP_MIN) return 0.0;
  return orc_sm_exp_core(x, 0.0);
}

/* rdb_sm_reduce_log: x = m 2^k with m in [sqrt(1/2), sqrt(2)). */
static inline void reduce_log(double x, double *m, int *k) {
  double xs = x;
  int kadj = 0;
  if (x < SM_TINY) {
    xs = x * SM_TWO54;
    kadj = -54;
  }
  int e;
  double f = frexp(xs, &e); /* Fortran fraction / exponent */
  *k = e + kadj;
  *m = f;
  if (*m < SM_SQRT_HALF) {
    *m = 2.0 * *m;
    *k = *k - 1;
  }
}

double orc_sm_log(double x) {
  if (isnan(x)) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x < SM_TINY && x * SM_TWO54 == 0.0) return -INFINITY;
  if (!isfinite(x)) return x;
  double m;
  int k;
  reduce_log(x, &m, &k);
  double f = (m - 1.0) / (m + 1.0);
  double u = f * f;
  double p = SM_LOG_C[9];
  for (int i = 8; i >= 1; --i) p = p * u + SM_LOG_C[i];
  double kr = (double)k;
  double s = kr * SM_LN2_HI + 2.0 * f;
  double w = ((2.0 * f) * u) * p + kr * SM_LN2_LO;
  return s + w;
}

/* rdb_sm_log_dd: log(x) = lh + ll in double-double. */
void orc_sm_log_dd(double x, double *lh, double *ll) {
  if (x < SM_TINY && x * SM_TWO54 == 0.0) {
    *lh = -INFINITY;
    *ll = 0.0;
    return;
  }
  double m, bh, bl, p1, p2, s, e;
  int k;
  reduce_log(x, &m, &k);
  double a = m - 1.0;
  two_sum(m, 1.0, &bh, &bl);
  double q1 = a / bh;
  two_prod(q1, bh, &p1, &p2);
  double rr = ((a - p1) - p2) - q1 * bl;
  double q2 = rr / bh;
  double fh = q1 + q2;
  double fl = q2 - (fh - q1);
  double u = fh * fh;
  double p = SM_LOG_C[9];
  for (int i = 8; i >= 1; --i) p = p * u + SM_LOG_C[i];
  double tail = ((2.0 * fh) * u) * p;
  double kr = (double)k;
  two_sum(kr * SM_LN2_HI, 2.0 * fh, &s, &e);
  *lh = s;
  *ll = e + ((2.0 * fl + tail) + kr * SM_LN2_LO);
}

/* rdb_sm_pow: x^y = exp(y log x) with log in double-double and the product
 * split exactly, the IEEE special cases first. */
double orc_sm_pow(double x, double y) {
  if (isnan(x) || isnan(y)) return NAN;
  if (y == 0.0 || x == 1.0) return 1.0;
  double ax = fabs(x), ay = fabs(y);
  int yint, yodd;
  if (isfinite(y)) {
    if (ay >= 9007199254740992.0) {
      yint = 1;
      yodd = 0;
    } else {
      yint = (ay == round(ay));
      double h = 0.5 * ay;
      yodd = yint && (h != round(h));
    }
  } else {
    yint = 0;
    yodd = 0;
  }
  int xneg = (x < 0.0) || (x == 0.0 && signbit(x));
  double r;
  if (!isfinite(y)) {
    if (ax == 1.0) return 1.0;
    if ((ax < 1.0) == (y > 0.0)) return 0.0;
    return INFINITY;
  }
  if (ax == 0.0) {
    r = (y > 0.0) ? 0.0 : INFINITY;
    return (xneg && yodd) ? -r : r;
  }
  if (!isfinite(x)) {
    r = (y > 0.0) ? INFINITY : 0.0;
    return (xneg && yodd) ? -r : r;
  }
  if (xneg && !yint) return NAN;
  double lh, ll;
  orc_sm_log_dd(ax, &lh, &ll);
  double t = y * lh;
  if (t > 800.0) {
    r = INFINITY;
  } else if (t < -800.0) {
    r = 0.0;
  } else {
    double ph, pl, e, c2;
    two_prod(y, lh, &ph, &pl);
    double c = pl + y * ll;
    two_sum(ph, c, &e, &c2);
    if (e > SM_EXP_MAX)
      r = INFINITY;
    else if (e < SM_EXP_MIN)
      r = 0.0;
    else
      r = orc_sm_exp_core(e, c2);
  }
  return (xneg && yodd) ? -r : r;
}

/* rdb_sm_erfc: Taylor series of erf below 1/2; above, exp(-x^2) from the
 * exact split x^2 = xh^2 + xl (x + xh) times a rational fit on [1/2, 2),
 * [2, 4) or an asymptotic series in 1/x^2 beyond 4. */
double orc_sm_erfc(double x) {
  if (isnan(x)) return x;
  double ax = fabs(x), p, q, ex;
  if (ax < 0.5) {
    double v = x * x;
    p = SM_ERF_C[11];
    for (int i = 10; i >= 0; --i) p = p * v + SM_ERF_C[i];
    return 1.0 - x * p;
  }
  if (ax > 27.3) {
    ex = 0.0;
  } else {
    double c = SM_SPLIT * ax;
    double xh = c - (c - ax);
    double xl = ax - xh;
    ex = orc_sm_exp_core(-(xh * xh), -(xl * (ax + xh)));
  }
  if (ax < 2.0) {
    p = SM_ERFC_A_P[6];
    q = SM_ERFC_A_Q[6];
    for (int i = 5; i >= 0; --i) {
      p = p * ax + SM_ERFC_A_P[i];
      q = q * ax + SM_ERFC_A_Q[i];
    }
    ex = ex * (p / q);
  } else if (ax < 4.0) {
    p = SM_ERFC_B_P[6];
    q = SM_ERFC_B_Q[6];
    for (int i = 5; i >= 0; --i) {
      p = p * ax + SM_ERFC_B_P[i];
      q = q * ax + SM_ERFC_B_Q[i];
    }
    ex = ex * (p / q);
  } else {
    double rax = 1.0 / ax;
    double w = rax * rax;
    p = SM_ERFC_C_P[5];
    q = SM_ERFC_C_Q[5];
    for (int i = 4; i >= 0; --i) {
      p = p * w + SM_ERFC_C_P[i];
      q = q * w + SM_ERFC_C_Q[i];
    }
    ex = ((ex * SM_INV_SQRTPI) * rax) * (1.0 + w * (p / q));
  }
  if (ex < SM_TINY) ex = 0.0;
  return (x < 0.0) ? 2.0 - ex : ex;
}

/* ------------------------------------------------------------------------
 * eos_specvol_derivs, Wright (1997) branch (rdb_eos.F90:1551-1622):
 * SV = alpha0 + lambda / (p + p0), differentiated in closed form.
 * ---------------------------------------------------------------------- */
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

void orc_specvol_derivs(double T, double S, double p, double *dsv_dt, double *dsv_ds) {
  double T_sq = T * T;
  double p_0 = WB0 + WB1 * T + WB2 * T_sq + WB3 * T_sq * T + WB4 * S + WB5 * S * T;
  double lambda = WC0 + WC1 * T + WC2 * T_sq + WC3 * T_sq * T + WC4 * S + WC5 * S * T;
  double dp0_dt = WB1 + 2.0 * WB2 * T + 3.0 * WB3 * T_sq + WB5 * S;
  double dlam_dt = WC1 + 2.0 * WC2 * T + 3.0 * WC3 * T_sq + WC5 * S;
  double dp0_ds = WB4 + WB5 * T;
  double dlam_ds = WC4 + WC5 * T;
  double big_p = p + p_0;
  double inv_p = 1.0 / big_p;
  double inv_p2 = inv_p * inv_p;
  *dsv_dt = WA1 + dlam_dt * inv_p - lambda * dp0_dt * inv_p2;
  *dsv_ds = WA2 + dlam_ds * inv_p - lambda * dp0_ds * inv_p2;
}

/* ------------------------------------------------------------------------
 * EPBL helpers (rdb_ocean_epbl.F90:630-858).
 * ---------------------------------------------------------------------- */
static inline double cube(double x) { return (x * x) * x; }

/* epbl_find_mstar (:630-688). */
double orc_epbl_find_mstar(const rdb_colphys_config *c, double b0, double ustar, double bld, double absf) {
  double mstar;
  double absf_floor = fmx(absf, 1.0e-20);
  if (c->epbl_mstar_scheme == RDB_EPBL_MSTAR_OM4) {
    double mstar_s = c->epbl_mstar_coef1 * sqrt(fmx(0.0, b0) / (ustar * ustar * absf_floor));
    double mstar_n = 0.0;
    if (ustar > absf_floor * bld) mstar_n = c->epbl_c_ek * orc_sm_log(ustar / (absf_floor * fmx(bld, 1.0e-10)));
    mstar = fmx(mstar_s, fmn(1.25, mstar_n));
    if (c->epbl_mstar_cap > 0.0) mstar = fmn(c->epbl_mstar_cap, mstar);
  } else if (c->epbl_mstar_scheme == RDB_EPBL_MSTAR_RH18) {
    double msn_term = c->epbl_rh18_cn2 * orc_sm_exp(c->epbl_rh18_cn3 * bld * absf / ustar);
    double mstar_n = c->epbl_rh18_cn1 * msn_term / (1.0 + msn_term);
    double bp = fmx(0.0, b0);
    double u5 = cube(ustar) * (ustar * ustar);
    double mstar_s = c->epbl_rh18_cs1 * orc_sm_pow(bp * bp * bld / (u5 * absf_floor), c->epbl_rh18_cs2);
    mstar = mstar_n + mstar_s;
    if (c->epbl_mstar_cap > 0.0) mstar = fmn(c->epbl_mstar_cap, mstar);
  } else {
    mstar = c->epbl_mstar;
  }
  if (c->epbl_mstar_conv_adj > 0.0) {
    double t1 = -bld * fmn(0.0, b0);
    double t2 = 2.0 * mstar * cube(ustar);
    if (t2 > 0.0)
      mstar = mstar * ((1.0 - c->epbl_mstar_conv_adj) * t1 + t2) / (t1 + t2);
    else
      mstar = mstar * (1.0 - c->epbl_mstar_conv_adj);
  }
  return mstar;
}

/* epbl_mixlen_shape (:690-714). */
double orc_epbl_mixlen_shape(double z_depth, double mld_guess, double translay, double expo, int shaped) {
  if (!shaped || translay >= 1.0 || translay < 0.0 || mld_guess <= 0.0) return 1.0;
  double fr = fmx(0.0, (mld_guess - z_depth) / mld_guess);
  if (expo == 2.0) return translay + (1.0 - translay) * fr * fr;
  return translay + (1.0 - translay) * orc_sm_pow(fr, expo);
}

#define LT_VONKAR_WAVES 0.40
#define LT_NU_AIR 1.0e-6
#define LT_RHO_AIR 1.225
#define LT_CHARNOCK_MIN 0.028
#define LT_CHARNOCK_SLOPE 0.0017
#define LT_CHARNOCK_ICPT (-0.005)
#define LT_US_TO_U10 0.0162
#define LT_SWH_FROM_U10SQ 0.0246
#define LT_U19P5_TO_U10 1.075
#define LT_FM_INTO_FP 1.296
#define LT_R_LOSS 0.667
#define LT_PI 3.141592653589793116

/* epbl_lf17_wave_state (:716-765): COARE 3.5 u* -> U10 fixed point, then the
 * surface Stokes drift and Phillips peak wavenumber. */
void orc_epbl_lf17_wave_state(double ustar_w, double rho_ocn, double *u10, double *ustokes, double *kphil) {
  double ustar_air = ustar_w * sqrt(rho_ocn / LT_RHO_AIR);
  double z0_smooth = 0.11 * LT_NU_AIR / ustar_air;
  double u = ustar_air * sqrt(1000.0);
  int converged = 0;
  for (int itt = 1; itt <= 20; ++itt) {
    double alpha_ch = fmn(LT_CHARNOCK_MIN, LT_CHARNOCK_SLOPE * u + LT_CHARNOCK_ICPT);
    double z0w = fmx(z0_smooth + alpha_ch * (ustar_air * ustar_air) / ORC_GRAVITY, 1.0e-10);
    double u_new = ustar_air * orc_sm_log(10.0 / z0w) / LT_VONKAR_WAVES;
    if (fabs(u_new - u) <= 1.0e-3 * u) {
      u = u_new;
      converged = 1;
      break;
    }
    u = u_new;
  }
  if (!converged) u = 25.82 * ustar_air;
  *u10 = u;
  *ustokes = LT_US_TO_U10 * u;
  double hm0 = LT_SWH_FROM_U10SQ * u * u;
  double f_mean = LT_FM_INTO_FP * 0.877 * ORC_GRAVITY / (2.0 * LT_PI * LT_U19P5_TO_U10 * u);
  double vstokes = 0.125 * LT_PI * LT_R_LOSS * f_mean * hm0 * hm0;
  *kphil = 0.176 * *ustokes / fmx(vstokes, 1.0e-30);
}

/* one_m_exp_x (:806-816). */
static inline double one_m_exp_x(double x) {
  if (x < 1.0e-4) return 1.0 - x * (0.5 - x / 6.0);
  return (1.0 - orc_sm_exp(-x)) / x;
}

/* epbl_lf17_la (:767-804). */
double orc_epbl_lf17_la(double ustar_w, double zsl, double ustokes, double kphil) {
  double xkz = kphil * zsl;
  double root2kz = sqrt(2.0 * xkz);
  double r1 = (0.302 - 1.68 * xkz) * one_m_exp_x(2.0 * xkz);
  double r3 = (0.1264 + 0.64 * xkz) * one_m_exp_x(5.12 * xkz);
  double r5;
  if (root2kz > 1.0e-3) {
    double ec1 = orc_sm_erfc(root2kz);
    double ec2 = orc_sm_erfc(1.6 * root2kz);
    r5 = sqrt(LT_PI) * (root2kz * (-(0.84 * ec1) + 0.2 * ec2) + 0.1182 * (ec2 - ec1) / root2kz);
  } else {
    r5 = -(0.64 * sqrt(LT_PI) * root2kz) + (-0.14184 + 1.0839648 * root2kz * root2kz);
  }
  double ustokes_sl = fmx(ustokes * (0.715 + r1 + r3 + r5), 1.0e-10);
  return sqrt(ustar_w / ustokes_sl);
}

/* epbl_lt_enhance (:818-858). */
double orc_epbl_lt_enhance(const rdb_colphys_config *c, double la, double b0, double ustar, double bld, double absf,
                           double mstar) {
  double vk = c->epbl_von_karman;
  double mld_ek = bld * absf / ustar;
  double ek_ob = fabs(b0 * vk) / (fmx(absf, 1.0e-20) * ustar * ustar);
  double mld_ob = fabs(bld * b0 * vk) / cube(ustar);
  double la_mod;
  if (b0 >= 0.0)
    la_mod = la * ((1.0 + fmx(-0.5, c->epbl_lt_lac1 * mld_ek)) + c->epbl_lt_lac4 * ek_ob + c->epbl_lt_lac2 * mld_ob);
  else
    la_mod = la * ((1.0 + fmx(-0.5, c->epbl_lt_lac1 * mld_ek)) + c->epbl_lt_lac5 * ek_ob + c->epbl_lt_lac3 * mld_ob);
  la_mod = fmx(la_mod, 1.0e-10);
  if (c->epbl_lt_scheme == RDB_EPBL_LT_ADDITIVE)
    return mstar + c->epbl_lt_enhance_coef * orc_sm_pow(la_mod, c->epbl_lt_enhance_exp);
  double enh = fmn(c->epbl_lt_max_enhance, 1.0 + c->epbl_lt_enhance_coef * orc_sm_pow(la_mod, c->epbl_lt_enhance_exp));
  return mstar * enh;
}

/* ------------------------------------------------------------------------
 * EPBL column solve, epbl_column_kernel (rdb_ocean_epbl.F90:931-1659) for a
 * free surface (stack from 0 Pa) and no penetrating shortwave: the surface
 * layer is the first live layer counting down from nz; prep sweep; MLD
 * root-find (false position / bisection); downward interface sweep with the
 * closed-form energy solve and the embedded forward elimination.
 *
 * Column arrays are 0-based, index 0 the bed: h[nz], T[nz], S[nz],
 * kd_int[nz+1].  *mld holds the previous MLD on entry (the first guess when
 * mld_use_prev_guess) and the converged one on exit.  tke[6] (may be NULL)
 * receives wind, conv, forcing, mixing, mech_decay, conv_decay in W/m^2.
 * eos_linear selects the linear EOS's constant sensitivities (the variant
 * the reference's EPBL tests run) instead of the Wright EOS the column
 * engine uses.
 * ---------------------------------------------------------------------- */
#define H_NEGLECT ORC_H_DIV_EPS

static double vstar_of(const rdb_colphys_config *c, double tke_here, double conv_perel, double htot,
                       double mld_guess, double ustar, double dt, double rho0) {
  if (c->epbl_vstar_scheme == RDB_EPBL_VSTAR_RH18) {
    double surf_scale = fmx(0.05, 1.0 - htot / mld_guess);
    return c->epbl_vstar_scale_fac * surf_scale *
           (c->epbl_vstar_surf_fac * ustar + orc_cbrt(c->epbl_wstar_ustar_coef * conv_perel / (dt * rho0)));
  }
  return c->epbl_vstar_scale_fac * orc_cbrt(tke_here / (dt * rho0));
}

static double kd_of(const rdb_colphys_config *c, double vstar, double htot, double hbs, double absf) {
  if (c->epbl_mld_iteration) {
    double mixlen = fmx(c->epbl_min_mix_len, (htot * hbs * vstar) / (c->epbl_ekman_scale_coef * absf * htot * hbs + vstar));
    return vstar * c->epbl_von_karman * mixlen;
  }
  return vstar * c->epbl_von_karman * (htot * hbs * vstar) / (c->epbl_ekman_scale_coef * absf * htot * hbs + vstar);
}

void orc_epbl_column(const rdb_colphys_config *c, int nz, double dt, const double *h, const double *T,
                     const double *S, double tau_tot, double Q_heat, double Q_salt, double wet, double f_centre,
                     int eos_linear, double *kd_int, double *mld, double *b0_out, double *la_out, double *tke) {
  const double rho0 = c->rho0, G = ORC_GRAVITY, nstar = c->epbl_nstar;
  double *t0 = malloc(sizeof(double) * 6 * (size_t)(nz > 0 ? nz : 1));
  double *s0 = t0 + nz, *dpe_t = s0 + nz, *dpe_s = dpe_t + nz, *dch_t = dpe_s + nz, *dch_s = dch_t + nz;
  /* prep (:1087-1183) */
  int k_sfc = nz - 1;
  for (int k = nz - 1; k >= 0; --k)
    if (h[k] > ORC_H_VANISHED) {
      k_sfc = k;
      break;
    }
  double pres = 0.0, h_sum = 0.0, dsv_dt_sfc = 0.0, dsv_ds_sfc = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    double hk = h[k];
    double t0k = 0.0, s0k = 0.0, dsv_dt_k, dsv_ds_k;
    if (hk > 0.0) {
      t0k = T[k];
      s0k = S[k];
    }
    double dmass = rho0 * hk;
    double dpres = G * dmass;
    double p_mid = pres + 0.5 * dpres;
    if (eos_linear) {
      /* the linear variant of eos_specvol_derivs (rdb_eos.F90:1617-1620) */
      dsv_dt_k = c->alpha_T / (rho0 * rho0);
      dsv_ds_k = -c->beta_S / (rho0 * rho0);
    } else {
      orc_specvol_derivs(t0k, s0k, p_mid, &dsv_dt_k, &dsv_ds_k);
    }
    t0[k] = t0k;
    s0[k] = s0k;
    dpe_t[k] = dmass * p_mid * dsv_dt_k;
    dpe_s[k] = dmass * p_mid * dsv_ds_k;
    dch_t[k] = dmass * dsv_dt_k;
    dch_s[k] = dmass * dsv_ds_k;
    if (k == k_sfc) {
      dsv_dt_sfc = dsv_dt_k;
      dsv_ds_sfc = dsv_ds_k;
    }
    if (k <= k_sfc) {
      pres = pres + dpres;
      h_sum = h_sum + hk;
    }
  }
  h_sum = h_sum + H_NEGLECT;
  /* forcing (:1185-1226) */
  double q_t_kin = (1.0 / (rho0 * c->cp)) * Q_heat;
  double q_s_kin = (1.0 / rho0) * Q_salt;
  double b0 = G * rho0 * (dsv_dt_sfc * q_t_kin + dsv_ds_sfc * q_s_kin);
  *b0_out = b0;
  double ctke_sfc = -(0.5 * G * (rho0 * rho0) * h[k_sfc] * (q_t_kin * dt * dsv_dt_sfc + q_s_kin * dt * dsv_ds_sfc));
  double ustar = fmx(sqrt(tau_tot / rho0), c->epbl_ustar_min);
  double absf = sqrt((1.0 - c->epbl_omega_frac) * (f_centre * f_centre) +
                     c->epbl_omega_frac * 4.0 * (c->epbl_omega * c->epbl_omega));
  double idecay = c->epbl_tke_decay * absf / ustar;
  double mech_in = dt * rho0 * cube(ustar);
  double la_val = 0.0, lt_u10 = 0.0, lt_ustokes = 0.0, lt_kphil = 0.0;
  if (c->epbl_use_lt) orc_epbl_lf17_wave_state(ustar, rho0, &lt_u10, &lt_ustokes, &lt_kphil);
  double d_wind = 0.0, d_conv = 0.0, d_forcing = 0.0, d_mixing = 0.0, d_mdecay = 0.0, d_cdecay = 0.0;
  double mld_found = 0.0;

  if (wet <= 0.0 || h_sum <= 2.0 * H_NEGLECT) {
    for (int k = 0; k <= nz; ++k) kd_int[k] = 0.0;
  } else {
    double min_mld = 0.0, max_mld = h_sum, dmld_min = 0.0, dmld_max = 0.0;
    int have_min = 0, have_max = 0;
    double mld_guess = 0.5 * (min_mld + max_mld);
    if (c->epbl_mld_use_prev_guess && *mld > 0.0) mld_guess = fmn(*mld, max_mld);
    int n_its = c->epbl_mld_iteration ? c->epbl_mld_max_its : 1;
    for (int obl_it = 1; obl_it <= n_its; ++obl_it) {
      d_wind = d_conv = d_forcing = d_mixing = d_mdecay = d_cdecay = 0.0;
      double mstar_val = orc_epbl_find_mstar(c, b0, ustar, mld_guess, absf);
      if (c->epbl_use_lt) {
        la_val = orc_epbl_lf17_la(ustar, c->epbl_la_frac_hbl * mld_guess, lt_ustokes, lt_kphil);
        mstar_val = orc_epbl_lt_enhance(c, la_val, b0, ustar, mld_guess, absf, mstar_val);
      }
      double mech_tke = mstar_val * mech_in;
      d_wind = mech_tke;
      double conv_perel;
      if (ctke_sfc <= 0.0) {
        double forcing_clip = fmx(ctke_sfc, -mech_tke);
        mech_tke = mech_tke + forcing_clip;
        conv_perel = 0.0;
        d_forcing = forcing_clip;
      } else {
        conv_perel = ctke_sfc;
        d_conv = nstar * ctke_sfc;
      }
      /* sweep initialisation (:1330-1350) */
      double h_ka = h[k_sfc] + H_NEGLECT;
      double hp_a = h_ka;
      double dpe_t_a = dpe_t[k_sfc], dpe_s_a = dpe_s[k_sfc];
      double dch_t_a = dch_t[k_sfc], dch_s_a = dch_s[k_sfc];
      double th_a = h_ka * t0[k_sfc], sh_a = h_ka * s0[k_sfc];
      double te_lag = 0.0, se_lag = 0.0, kddt_prev = 0.0;
      double htot = h[k_sfc], z_int = h[k_sfc], pres_int = G * rho0 * h[k_sfc];
      double mld_output = h[k_sfc];
      int sfc_connected = 1;
      for (int ki = k_sfc + 1; ki <= nz; ++ki) kd_int[ki] = 0.0;
      /* interfaces ki = k_sfc .. 1 (0-based; the top of layer kb = ki - 1) */
      for (int ki = k_sfc; ki >= 1; --ki) {
        int ka = ki, kb = ki - 1;
        h_ka = h[ka] + H_NEGLECT;
        double h_kb = h[kb] + H_NEGLECT;
        int sfc_disconnect = 0;
        double exp_kh = orc_sm_exp(-h_ka * idecay);
        d_mdecay = d_mdecay + (1.0 - exp_kh) * mech_tke;
        mech_tke = mech_tke * exp_kh;
        double nstar_fc = nstar;
        if (conv_perel > 0.0 && absf > 0.0) {
          double ah = absf * htot;
          nstar_fc = nstar * conv_perel / (conv_perel + 0.2 * sqrt(0.5 * dt * rho0 * cube(ah) * conv_perel));
        }
        double tot_tke = mech_tke + nstar_fc * conv_perel;
        double kd_val, dt_h;
        if (tot_tke <= 0.0 &&
            0.0 <= (dch_t[kb] + dch_t[ka]) * (t0[ka] - t0[kb]) + (dch_s[kb] + dch_s[ka]) * (s0[ka] - s0[kb])) {
          kd_val = 0.0;
          sfc_disconnect = 1;
          /* the reference leaves dt_h at the previous interface's finite
           * value here; it only enters kddt_cur = 0 * dt_h = 0 */
          dt_h = dt / fmx(0.5 * (h_ka + h_kb), 1.0e-15 * h_sum);
        } else {
          dt_h = dt / fmx(0.5 * (h_ka + h_kb), 1.0e-15 * h_sum);
          double tke_here = mech_tke + c->epbl_wstar_ustar_coef * conv_perel;
          double hb_hs = (h_sum - z_int) / h_sum;
          double shape_fn = orc_epbl_mixlen_shape(z_int, mld_guess, c->epbl_translay_scale, c->epbl_mixlen_exponent,
                                                  c->epbl_mld_iteration);
          double hbs = fmn(hb_hs, shape_fn);
          double vstar, kd_g0;
          if (tke_here > 0.0) {
            vstar = vstar_of(c, tke_here, conv_perel, htot, mld_guess, ustar, dt, rho0);
            kd_g0 = kd_of(c, vstar, htot, hbs, absf);
          } else {
            vstar = 0.0;
            kd_g0 = 0.0;
          }
          double hp_b = h_kb;
          double th_b = h_kb * t0[kb], sh_b = h_kb * s0[kb];
          double hps = hp_a + hp_b;
          double bdt1 = hp_a * hp_b;
          double dt_c = hp_a * th_b - hp_b * th_a;
          double ds_c = hp_a * sh_b - hp_b * sh_a;
          double pec_core = hp_b * (dpe_t_a * dt_c + dpe_s_a * ds_c) - hp_a * (dpe_t[kb] * dt_c + dpe_s[kb] * ds_c);
          double colht_core = hp_b * (dch_t_a * dt_c + dch_s_a * ds_c) - hp_a * (dch_t[kb] * dt_c + dch_s[kb] * ds_c);
          if (colht_core < 0.0) pec_core = pec_core - pres_int * colht_core;
          double dkddt = kd_g0 * dt_h;
          double pe_max = pec_core / (bdt1 * hps);
          if (pe_max < 0.0) {
            tke_here = mech_tke + c->epbl_wstar_ustar_coef * (conv_perel - pe_max);
            if (tke_here > 0.0) {
              vstar = vstar_of(c, tke_here, conv_perel, htot, mld_guess, ustar, dt, rho0);
              kd_val = kd_of(c, vstar, htot, hbs, absf);
            } else {
              vstar = 0.0;
              kd_val = 0.0;
            }
            double pe_g0 = pec_core * dkddt / (bdt1 * (bdt1 + dkddt * hps));
            double dpe_conv = pec_core * (kd_val * dt_h) / (bdt1 * (bdt1 + (kd_val * dt_h) * hps));
            if (dpe_conv > 0.0) {
              kd_val = kd_g0;
              dpe_conv = pe_g0;
            }
            conv_perel = conv_perel - dpe_conv;
            d_conv = d_conv - nstar * dpe_conv;
            if (sfc_connected) mld_output = mld_output + h[kb];
          } else {
            double tke_used, frac_bl;
            if ((pec_core * dkddt <= tot_tke * (bdt1 * (bdt1 + dkddt * hps))) || (pec_core <= 0.0)) {
              kd_val = kd_g0;
              tke_used = pec_core * dkddt / (bdt1 * (bdt1 + dkddt * hps));
              frac_bl = 1.0;
            } else {
              kd_val = ((bdt1 * bdt1) * tot_tke) / (dt_h * (pec_core - bdt1 * hps * tot_tke));
              tke_used = tot_tke;
              frac_bl = tot_tke * (bdt1 * (bdt1 + dkddt * hps)) / (pec_core * dkddt);
            }
            if (sfc_connected) mld_output = mld_output + frac_bl * h[kb];
            if (frac_bl < 1.0) sfc_disconnect = 1;
            double r_reduc = 0.0;
            if (tot_tke > 0.0 && tot_tke > tke_used) r_reduc = (tot_tke - tke_used) / tot_tke;
            d_mixing = d_mixing + tke_used;
            d_cdecay = d_cdecay + (1.0 - r_reduc) * (nstar - nstar_fc) * conv_perel;
            mech_tke = r_reduc * mech_tke;
            conv_perel = r_reduc * conv_perel;
          }
        }
        kd_int[ki] = kd_val;
        double kddt_cur = kd_val * dt_h;
        double b1 = 1.0 / (hp_a + kddt_cur);
        double c1 = kddt_cur * b1;
        if (ki == k_sfc) {
          te_lag = b1 * (h_ka * t0[ka]);
          se_lag = b1 * (h_ka * s0[ka]);
        } else {
          double te_new = b1 * (h_ka * t0[ka] + kddt_prev * te_lag);
          double se_new = b1 * (h_ka * s0[ka] + kddt_prev * se_lag);
          te_lag = te_new;
          se_lag = se_new;
        }
        hp_a = h_kb + (hp_a * b1) * kddt_cur;
        dpe_t_a = dpe_t[kb] + c1 * dpe_t_a;
        dpe_s_a = dpe_s[kb] + c1 * dpe_s_a;
        dch_t_a = dch_t[kb] + c1 * dch_t_a;
        dch_s_a = dch_s[kb] + c1 * dch_s_a;
        th_a = h_kb * t0[kb] + kddt_cur * te_lag;
        sh_a = h_kb * s0[kb] + kddt_cur * se_lag;
        kddt_prev = kddt_cur;
        if (sfc_disconnect) {
          htot = h[kb];
          sfc_connected = 0;
        } else {
          htot = htot + h[kb];
        }
        z_int = z_int + h[kb];
        pres_int = pres_int + G * rho0 * h[kb];
      }
      kd_int[0] = 0.0;
      d_mdecay = d_mdecay + mech_tke;
      d_cdecay = d_cdecay + nstar * conv_perel;
      mld_found = mld_output;
      if (!c->epbl_mld_iteration) break;
      if (fabs(mld_found - mld_guess) < c->epbl_mld_tol) break;
      if (obl_it == n_its) break;
      if (mld_found > mld_guess) {
        min_mld = mld_guess;
        dmld_min = mld_found - mld_guess;
        have_min = 1;
      } else {
        max_mld = mld_guess;
        dmld_max = mld_found - mld_guess;
        have_max = 1;
      }
      if (c->epbl_mld_bisection)
        mld_guess = 0.5 * (min_mld + max_mld);
      else if (have_min && have_max && obl_it > 2 && ((obl_it - 1) % 4) > 0)
        mld_guess = min_mld + dmld_min * (max_mld - min_mld) / (dmld_min - dmld_max);
      else if (mld_found > min_mld && mld_found < max_mld)
        mld_guess = mld_found;
      else
        mld_guess = 0.5 * (min_mld + max_mld);
    }
  }
  *mld = mld_found;
  *la_out = la_val;
  if (tke) {
    tke[0] = d_wind / dt;
    tke[1] = d_conv / dt;
    tke[2] = d_forcing / dt;
    tke[3] = d_mixing / dt;
    tke[4] = d_mdecay / dt;
    tke[5] = d_cdecay / dt;
  }
  free(t0);
}

/* epbl_compute over a grid (rdb_ocean_epbl.F90:864-929) followed by the merge
 * into kv/kt (epbl_merge_into_kv_kt :1661-1688; "max_shear" without
 * kappa-shear is the additive merge).  tke: 6 planes of nx*ny or NULL. */
void orc_epbl_compute_merge(const rdb_colphys_config *c, int nx, int ny, int nz, double dt, const double *h,
                            const double *T, const double *S, const double *stress_mag, const double *stress_shelf,
                            const double *Q_heat, const double *Q_salt, const double *wet, const double *f_centre,
                            double *kd_int, double *mld, double *b0, double *la, double *tke, double *kv,
                            double *kt) {
  size_t n2 = (size_t)nx * ny;
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double hc[nz], tc[nz], sc[nz], kdc[nz + 1], tk[6];
      for (int k = 0; k < nz; ++k) {
        hc[k] = h[IDX(i, j, k)];
        tc[k] = T[IDX(i, j, k)];
        sc[k] = S[IDX(i, j, k)];
      }
      size_t c2 = IDX2(i, j);
      orc_epbl_column(c, nz, dt, hc, tc, sc, stress_mag[c2] + stress_shelf[c2], Q_heat[c2], Q_salt[c2], wet[c2],
                      f_centre ? f_centre[c2] : 0.0, 0, kdc, &mld[c2], &b0[c2], &la[c2], tke ? tk : NULL);
      for (int k = 0; k <= nz; ++k) kd_int[IDX(i, j, k)] = kdc[k];
      if (tke)
        for (int q = 0; q < 6; ++q) tke[q * n2 + c2] = tk[q];
      for (int k = 1; k < nz; ++k) {
        size_t p = IDX(i, j, k);
        if (c->epbl_combine == RDB_EPBL_COMBINE_MAX) {
          kt[p] = fmx(kt[p], kdc[k]);
          kv[p] = fmx(kv[p], c->epbl_prandtl * kdc[k]);
        } else {
          kt[p] = kt[p] + kdc[k];
          kv[p] = kv[p] + c->epbl_prandtl * kdc[k];
        }
      }
    }
}

/* ------------------------------------------------------------------------
 * vmix_add_kv_ml_invz2 (rdb_ocean_vmix.F90:1716-1748): kv += K hmix^2 /
 * max(z, dz_min)^2 at the interfaces shallower than hmix.
 * ---------------------------------------------------------------------- */
void orc_vmix_invz2(const rdb_colphys_config *c, int nx, int ny, int nz, const double *h, double *kv) {
  if (c->kv_ml_invz2 <= 0.0 || c->hmix_fixed <= 0.0) return;
  double hmix = c->hmix_fixed, hmix_sq = hmix * hmix;
  double dz_min_safe = 0.5 * hmix / (double)(nz > 1 ? nz : 1);
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double z = 0.0;
      for (int k = nz - 1; k >= 1; --k) {
        z = z + h[IDX(i, j, k)];
        if (z >= hmix) break;
        double m = fmx(z, dz_min_safe);
        kv[IDX(i, j, k)] = kv[IDX(i, j, k)] + c->kv_ml_invz2 * hmix_sq / (m * m);
      }
    }
}

/* vmix_convection_impl (rdb_ocean_vmix.F90:1768-1799): below the boundary
 * layer, N^2 < thresh floors kt at kd_conv and kv at prandtl * kd_conv. */
void orc_vmix_convection(const rdb_colphys_config *c, int nx, int ny, int nz, const double *h, const double *rho,
                         const double *bl_depth, double *kv, double *kt) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double z = 0.0;
      for (int k = nz - 1; k >= 1; --k) {
        z = z + h[IDX(i, j, k)];
        double dz_face = 0.5 * (h[IDX(i, j, k - 1)] + h[IDX(i, j, k)]);
        if (dz_face <= ORC_H_VANISHED) continue;
        if (z < bl_depth[IDX2(i, j)]) continue;
        double n2 = -(ORC_GRAVITY * (rho[IDX(i, j, k)] - rho[IDX(i, j, k - 1)]) / (c->rho0 * dz_face));
        if (n2 < c->conv_n2_thresh) {
          kt[IDX(i, j, k)] = fmx(kt[IDX(i, j, k)], c->conv_kd);
          kv[IDX(i, j, k)] = fmx(kv[IDX(i, j, k)], c->conv_prandtl * c->conv_kd);
        }
      }
    }
}

/* One interface of vmix_split_ddiff_impl (rdb_ocean_vmix.F90:1936-2028) with
 * the scalar alpha_T / beta_S: salt fingering (Large et al. 1994) and
 * diffusive convection (Marmorino-Caldwell 1976 or Kelley 1990).  T/S/h of
 * the layer above (u) and below (l). */
void orc_ddiff_iface(const rdb_colphys_config *c, double hu, double hl, double Tu, double Tl, double Su, double Sl,
                     double *kd_t, double *kd_s) {
  *kd_t = 0.0;
  *kd_s = 0.0;
  if (!(hu > ORC_H_VANISHED && hl > ORC_H_VANISHED)) return;
  double adT = c->alpha_T * (Tu - Tl);
  double bdS = c->beta_S * (Su - Sl);
  if (adT >= bdS && bdS > 0.0) {
    double rrho = adT / bdS;
    if (rrho < c->ddiff_strat_param_max) {
      double fing0 = (rrho - 1.0) / (c->ddiff_strat_param_max - 1.0);
      double fing = (c->ddiff_exp1 == 1.0) ? fing0 : orc_sm_pow(fing0, c->ddiff_exp1);
      double dd = orc_sm_pow(1.0 - fing, c->ddiff_exp2);
      *kd_s = c->ddiff_kappa_s * dd;
    }
    *kd_t = 0.7 * *kd_s;
  } else if (adT >= bdS && adT < 0.0) {
    double rrho = adT / bdS;
    double dd;
    if (c->ddiff_use_k90)
      dd = c->ddiff_mol_diff * 8.7 * orc_sm_pow(rrho, 1.1);
    else
      dd = c->ddiff_mol_diff * c->ddiff_param1 *
           orc_sm_exp(c->ddiff_param2 * orc_sm_exp(c->ddiff_param3 * (1.0 / rrho - 1.0)));
    *kd_t = dd;
    if (rrho < 0.5)
      *kd_s = 0.15 * rrho * dd;
    else
      *kd_s = (1.85 * rrho - 0.85) * dd;
  }
}

void orc_vmix_split_ddiff(const rdb_colphys_config *c, int nx, int ny, int nz, const double *h, const double *T,
                          const double *S, double *kt, double *ks) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (int k = 0; k <= nz; ++k) {
        size_t p = IDX(i, j, k);
        double kt_pre = kt[p], kd_t = 0.0, kd_s = 0.0;
        if (k >= 1 && k <= nz - 1)
          orc_ddiff_iface(c, h[IDX(i, j, k)], h[IDX(i, j, k - 1)], T[IDX(i, j, k)], T[IDX(i, j, k - 1)],
                          S[IDX(i, j, k)], S[IDX(i, j, k - 1)], &kd_t, &kd_s);
        ks[p] = kt_pre + kd_s;
        kt[p] = kt_pre + kd_t;
      }
}
