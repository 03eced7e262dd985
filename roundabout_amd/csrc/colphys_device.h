// colphys_device.h -- point formulas of the vertical-mixing path, shared by
// the staged kernels and the fused column-step kernel.
//
// Every function keeps the reference's Fortran left-to-right evaluation
// order and the library is compiled with -ffp-contract=off, so in FP64 each
// result is bitwise the value the reference's contraction-off CPU build
// computes.  The assembly functions are templates on the arithmetic type R
// so that the split-precision mode can run them in FP32; the EOS and the
// density/buoyancy differences that feed N^2 and the bulk Richardson number
// stay FP64 in every mode (single-precision densities of ~1027 kg/m^3 would
// cancel in those differences).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "safe_math.h"

namespace colphys {

constexpr double GRAVITY = 9.80665;     // rdb_constants.F90:37
constexpr double H_VANISHED = 1.5e-4;   // rdb_constants.F90:67
constexpr double H_DIV_EPS = 1.0e-20;   // rdb_constants.F90:85

// Scalars of the path, copied from rdb_colphys_config at create.
struct Params {
  double nu0, nu_bg, kappa_bg, pp_alpha, rho0, shear2_floor;
  double ri_crit, c_vt2, cs_nonlocal, alpha_T, beta_S, p_ref, cp;
  double kv_bg, kt_bg, ks_bg, kv_max, kd_max;
  int use_kpp, has_heat, has_salt, use_harmonic, guard;
  // double diffusion inside the heat/salt split (off: ks := kt)
  int ddiff, ddiff_k90;
  double dd_rmax, dd_kappa_s, dd_exp1, dd_exp2, dd_p1, dd_p2, dd_p3, dd_mol;
};

template <typename R>
__device__ __forceinline__ R fmx(R a, R b) { return (a > b) ? a : b; }
template <typename R>
__device__ __forceinline__ R fmn(R a, R b) { return (a < b) ? a : b; }

// ---------------------------------------------------------------------------
// Cube root (Kahan 1991; Muller 2016 ch. 5): exponent split modulo 3, degree-4
// seed on m in [1/2, 4), two Halley steps, following rdb_sm_cbrt
// (rdb_safe_math_pow.inc:107-142).  The seed coefficients must equal the
// reference's table (rdb_safe_math_consts.inc:104-106) for bitwise results.
// The exponent/fraction split is done on the bit pattern: exact.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double cbrt_kahan(double x) {
  uint64_t bits = (uint64_t)__double_as_longlong(x);
  uint32_t bexp = (uint32_t)((bits >> 52) & 0x7ffu);
  if (bexp == 0x7ffu) return x;  // NaN, +-Inf
  if (bexp == 0u) return __longlong_as_double((long long)(bits & 0x8000000000000000ull));  // +-0, subnormal
  int e = (int)bexp - 1022;  // ax = f * 2^e with f in [1/2, 1)
  double f = __longlong_as_double((long long)((bits & 0x000fffffffffffffull) | (1022ull << 52)));
  int rem = ((e % 3) + 3) % 3;
  int q = (e - rem) / 3;
  double m = f * (double)(1 << rem);  // exact: multiply by 1, 2 or 4
  double y = -0.003660165019409192949;
  y = y * m + 0.04437845996234061896;
  y = y * m + -0.2158001895775201462;
  y = y * m + 0.6542973117201534849;
  y = y * m + 0.5189978361037301813;
  for (int it = 0; it < 2; ++it) {
    double y3 = (y * y) * y;
    y = y + y * ((m - y3) / ((y3 + y3) + m));
  }
  // scale(y, q): y is in [0.79, 1.59]; 2^q stays in the normal range for
  // every normal input, so one exact multiply by a built power of two.
  double s = __longlong_as_double((long long)((uint64_t)(q + 1023) << 52));
  y = y * s;
  return (bits >> 63) ? -y : y;
}

// FP32 cube root for the split-precision assembly: same method and seed, two
// Halley steps.
__device__ __forceinline__ float cbrt_kahan(float x) {
  uint32_t bits = (uint32_t)__float_as_int(x);
  uint32_t bexp = (bits >> 23) & 0xffu;
  if (bexp == 0xffu) return x;
  if (bexp == 0u) return __int_as_float((int)(bits & 0x80000000u));
  int e = (int)bexp - 126;
  float f = __int_as_float((int)((bits & 0x007fffffu) | (126u << 23)));
  int rem = ((e % 3) + 3) % 3;
  int q = (e - rem) / 3;
  float m = f * (float)(1 << rem);
  float y = -0.003660165019409192949f;
  y = y * m + 0.04437845996234061896f;
  y = y * m + -0.2158001895775201462f;
  y = y * m + 0.6542973117201534849f;
  y = y * m + 0.5189978361037301813f;
  for (int it = 0; it < 2; ++it) {
    float y3 = (y * y) * y;
    y = y + y * ((m - y3) / ((y3 + y3) + m));
  }
  float s = __int_as_float((int)((uint32_t)(q + 127) << 23));
  y = y * s;
  return (bits >> 31) ? -y : y;
}

// ---------------------------------------------------------------------------
// A1 Wright (1997) reduced-range EOS (rdb_eos.F90:909-959, coeffs :203-217).
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R eos_wright(R T, R S, double h, R rho0, R p_ref) {
  if (h > H_VANISHED) {
    R Tsq = T * T;
    R Tcu = Tsq * T;
    R alpha0 = R(7.057924e-4) + R(3.480336e-7) * T + R(-1.112733e-7) * S;
    R p0 = R(5.790749e8) + R(3.516535e6) * T + R(-4.002714e4) * Tsq + R(2.084372e2) * Tcu +
           R(5.944068e5) * S + R(-9.643486e3) * S * T;
    R lambda = R(1.704853e5) + R(7.904722e2) * T + R(-7.984422e0) * Tsq + R(5.140652e-2) * Tcu +
               R(-2.302158e2) * S + R(-3.079464e0) * S * T;
    R ppp0 = p_ref + p0;
    R denom = lambda + alpha0 * ppp0;
    return ppp0 / denom;
  }
  return rho0;
}

// ---------------------------------------------------------------------------
// A3 PP81 at one interior interface between layers km1 (below) and k (above),
// from centred velocities (rdb_ocean_vmix.F90:1043-1072).
// ---------------------------------------------------------------------------
// The density difference is formed in FP64 and handed over as drho =
// rho_k - rho_km1: the first operation of the reference's N^2 expression, so
// the FP64 instantiation is bitwise the reference; the FP32 instantiation
// avoids cancelling two ~1027 kg/m^3 single-precision densities.
template <typename R>
__device__ __forceinline__ void pp81_iface(const Params& pr, R u_km1, R v_km1, R u_k, R v_k,
                                           R h_km1, R h_k, double drho, R& kv, R& kt) {
  R dz_face = R(0.5) * (h_km1 + h_k);
  if (dz_face <= R(0)) {
    kv = R(pr.nu_bg);
    kt = R(pr.kappa_bg);
    return;
  }
  R du_dz = (u_k - u_km1) / dz_face;
  R dv_dz = (v_k - v_km1) / dz_face;
  R shear2 = fmx(du_dz * du_dz + dv_dz * dv_dz, R(pr.shear2_floor));
  R n2 = -(R(GRAVITY) * R(drho) / (R(pr.rho0) * dz_face));
  R ri = n2 / shear2;
  R denom = R(1) + R(pr.pp_alpha) * fmx(ri, R(0));
  R ri_factor = R(1) / (denom * denom);
  kv = R(pr.nu_bg) + R(pr.nu0) * ri_factor;
  kt = R(pr.kappa_bg) + R(pr.nu0) * ri_factor / denom;
}

// A6 kpp_surface_buoyancy_flux (rdb_ocean_vmix.F90:1518-1559).
template <typename R>
__device__ __forceinline__ R kpp_b0(const Params& pr, R qT, R qS) {
  return (R(GRAVITY) / R(pr.rho0)) * (R(pr.alpha_T) * qT - R(pr.beta_S) * qS);
}

// face_thick (rdb_ocean_vdiff.F90:1928-1962).
__device__ __forceinline__ double face_thick(double ha, double hb, int use_harmonic) {
  if (use_harmonic) {
    double sum_h = ha + hb;
    if (sum_h > 1.0e-30) return 2.0 * ha * hb / sum_h;
    return 0.5 * sum_h;
  }
  return 0.5 * (ha + hb);
}

// rdb_conc_add_source (rdb_tracer_conc.inc:33-50).
__device__ __forceinline__ double conc_add_source(double c, double h, double src) {
  if (h > H_VANISHED) return c + src / (h + H_DIV_EPS);
  return c;
}

// One interface of vmix_split_ddiff_impl (rdb_ocean_vmix.F90:1936-2028),
// scalar alpha_T / beta_S: salt fingering (Large et al. 1994) and diffusive
// convection (Marmorino-Caldwell 1976, or Kelley 1990).  u is the layer
// above the interface, l the one below.
__device__ __forceinline__ void ddiff_iface(const Params& pr, double hu, double hl, double Tu, double Tl, double Su,
                                            double Sl, double& kd_t, double& kd_s) {
  kd_t = 0.0;
  kd_s = 0.0;
  if (!(hu > H_VANISHED && hl > H_VANISHED)) return;
  const double adT = pr.alpha_T * (Tu - Tl);
  const double bdS = pr.beta_S * (Su - Sl);
  if (adT >= bdS && bdS > 0.0) {
    const double rrho = adT / bdS;
    if (rrho < pr.dd_rmax) {
      const double fing0 = (rrho - 1.0) / (pr.dd_rmax - 1.0);
      const double fing = (pr.dd_exp1 == 1.0) ? fing0 : sm::pow(fing0, pr.dd_exp1);
      const double dd = sm::pow(1.0 - fing, pr.dd_exp2);
      kd_s = pr.dd_kappa_s * dd;
    }
    kd_t = 0.7 * kd_s;
  } else if (adT >= bdS && adT < 0.0) {
    const double rrho = adT / bdS;
    double dd;
    if (pr.ddiff_k90)
      dd = pr.dd_mol * 8.7 * sm::pow(rrho, 1.1);
    else
      dd = pr.dd_mol * pr.dd_p1 * sm::exp(pr.dd_p2 * sm::exp(pr.dd_p3 * (1.0 / rrho - 1.0)));
    kd_t = dd;
    if (rrho < 0.5)
      kd_s = 0.15 * rrho * dd;
    else
      kd_s = (1.85 * rrho - 0.85) * dd;
  }
}

// nonlocal_face_flux (rdb_ocean_vmix.F90:1694-1714).
__device__ __forceinline__ double nl_face(double g, double h_below, double h_above) {
  return (h_below > H_VANISHED && h_above > H_VANISHED) ? g : 0.0;
}

// One row of build_factorize_tracer_matrix (rdb_ocean_vdiff.F90:1288-1333):
// layer k of nz with thicknesses (h_km1, h_k, h_kp1) and diffusivities at its
// bottom (K_k) and top (K_kp1) interfaces.  Returns a (sub), b (diag),
// c (super) of the raw row.
__device__ __forceinline__ void tracer_row(int k, int nz, double dt, int use_harmonic, double h_km1,
                                           double h_k, double h_kp1, double K_k, double K_kp1,
                                           double& a, double& b, double& c) {
  const double HV = H_VANISHED;
  double hc = fmx(h_k, HV);
  if (k == 0) {
    double alpha = 0.0;
    if (nz > 1) {
      double hp = fmx(h_kp1, HV);
      double dz_face = face_thick(hc, hp, use_harmonic);
      alpha = dt * K_kp1 / (hc * dz_face);
      if (h_k <= HV || h_kp1 <= HV) alpha = 0.0;
    }
    a = 0.0;
    c = -alpha;
    b = 1.0 + alpha;
  } else if (k < nz - 1) {
    double hm = fmx(h_km1, HV);
    double hp = fmx(h_kp1, HV);
    double beta = dt * K_k / (hc * face_thick(hm, hc, use_harmonic));
    double alpha = dt * K_kp1 / (hc * face_thick(hc, hp, use_harmonic));
    if (h_k <= HV || h_km1 <= HV) beta = 0.0;
    if (h_k <= HV || h_kp1 <= HV) alpha = 0.0;
    a = -beta;
    c = -alpha;
    b = 1.0 + alpha + beta;
  } else {
    double hm = fmx(h_km1, HV);
    double dz_face = face_thick(hm, hc, use_harmonic);
    double beta = dt * K_k / (hc * dz_face);
    if (h_k <= HV || h_km1 <= HV) beta = 0.0;
    a = -beta;
    c = 0.0;
    b = 1.0 + beta;
  }
}

}  // namespace colphys
