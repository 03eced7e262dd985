// contrib.hip -- the vertical-mixing contributors that are off at the
// reference defaults, on the packed level-major layout (X[k * P + p], k = 0
// the bed, one lane per column):
//
//   k_epbl     energetics-based planetary boundary layer (Reichl & Hallberg
//              2018), epbl_column_kernel (rdb_ocean_epbl.F90:931-1659): MLD
//              root-find around a downward interface sweep with the
//              closed-form energy solve and the embedded forward elimination;
//   k_coeff_tail  the rest of the coefficient build in one top-down pass:
//              PP81 (when EPBL runs), the EPBL merge into kv/kt (:1661-1688),
//              KV_ML_INVZ2 (rdb_ocean_vmix.F90:1716-1748), convective
//              adjustment (:1768-1799), the heat/salt split with double
//              diffusion (ddiff_iface, :1936-2028) and vmix_assemble.
//
// The EPBL sweep needs, per level, the pre-mixing T/S and the pressure-
// weighted PE and column-height sensitivities.  The reference fills six
// column work arrays once and re-reads them in every MLD iteration; here
// each iteration recomputes them on the fly while it walks down (the same
// pressure stack, summed in the same order, so the same bits): every
// iteration streams h, T and S only, and no work array is written.
#include <type_traits>

#include "colphys_kernels.h"
#include "safe_math.h"

namespace colphys {

#define LV(arr, k) (arr)[(size_t)(k) * (size_t)P + (size_t)p]

namespace {

constexpr double H_NEGLECT = H_DIV_EPS;

// eos_specvol_derivs, Wright (1997) branch (rdb_eos.F90:1551-1622).
__device__ __forceinline__ void specvol_derivs(double T, double S, double p, double& dsv_dt, double& dsv_ds) {
  const double T_sq = T * T;
  const double p_0 = 5.790749e8 + 3.516535e6 * T + -4.002714e4 * T_sq + 2.084372e2 * T_sq * T + 5.944068e5 * S +
                     -9.643486e3 * S * T;
  const double lambda = 1.704853e5 + 7.904722e2 * T + -7.984422e0 * T_sq + 5.140652e-2 * T_sq * T +
                        -2.302158e2 * S + -3.079464e0 * S * T;
  const double dp0_dt = 3.516535e6 + 2.0 * -4.002714e4 * T + 3.0 * 2.084372e2 * T_sq + -9.643486e3 * S;
  const double dlam_dt = 7.904722e2 + 2.0 * -7.984422e0 * T + 3.0 * 5.140652e-2 * T_sq + -3.079464e0 * S;
  const double dp0_ds = 5.944068e5 + -9.643486e3 * T;
  const double dlam_ds = -2.302158e2 + -3.079464e0 * T;
  const double big_p = p + p_0;
  const double inv_p = 1.0 / big_p;
  const double inv_p2 = inv_p * inv_p;
  dsv_dt = 3.480336e-7 + dlam_dt * inv_p - lambda * dp0_dt * inv_p2;
  dsv_ds = -1.112733e-7 + dlam_ds * inv_p - lambda * dp0_ds * inv_p2;
}

__device__ __forceinline__ double cube(double x) { return (x * x) * x; }

// epbl_find_mstar (rdb_ocean_epbl.F90:630-688).
__device__ __forceinline__ double find_mstar(const EpblParams& c, double b0, double ustar, double bld, double absf) {
  double mstar;
  const double absf_floor = fmx(absf, 1.0e-20);
  if (c.mstar_scheme == RDB_EPBL_MSTAR_OM4) {
    const double mstar_s = c.mstar_coef1 * sqrt(fmx(0.0, b0) / (ustar * ustar * absf_floor));
    double mstar_n = 0.0;
    if (ustar > absf_floor * bld) mstar_n = c.c_ek * sm::log(ustar / (absf_floor * fmx(bld, 1.0e-10)));
    mstar = fmx(mstar_s, fmn(1.25, mstar_n));
    if (c.mstar_cap > 0.0) mstar = fmn(c.mstar_cap, mstar);
  } else if (c.mstar_scheme == RDB_EPBL_MSTAR_RH18) {
    const double msn_term = c.cn2 * sm::exp(c.cn3 * bld * absf / ustar);
    const double mstar_n = c.cn1 * msn_term / (1.0 + msn_term);
    const double bp = fmx(0.0, b0);
    const double u5 = cube(ustar) * (ustar * ustar);
    const double mstar_s = c.cs1 * sm::pow(bp * bp * bld / (u5 * absf_floor), c.cs2);
    mstar = mstar_n + mstar_s;
    if (c.mstar_cap > 0.0) mstar = fmn(c.mstar_cap, mstar);
  } else {
    mstar = c.mstar_const;
  }
  if (c.mstar_conv_adj > 0.0) {
    const double t1 = -bld * fmn(0.0, b0);
    const double t2 = 2.0 * mstar * cube(ustar);
    if (t2 > 0.0)
      mstar = mstar * ((1.0 - c.mstar_conv_adj) * t1 + t2) / (t1 + t2);
    else
      mstar = mstar * (1.0 - c.mstar_conv_adj);
  }
  return mstar;
}

// epbl_mixlen_shape (:690-714).
__device__ __forceinline__ double mixlen_shape(const EpblParams& c, double z, double mld_guess) {
  if (!c.mld_iteration || c.translay >= 1.0 || c.translay < 0.0 || mld_guess <= 0.0) return 1.0;
  const double fr = fmx(0.0, (mld_guess - z) / mld_guess);
  if (c.mixlen_exponent == 2.0) return c.translay + (1.0 - c.translay) * fr * fr;
  return c.translay + (1.0 - c.translay) * sm::pow(fr, c.mixlen_exponent);
}

constexpr double LT_PI = 3.141592653589793116;

// epbl_lf17_wave_state (:716-765).
__device__ __noinline__ void lf17_wave_state(double ustar_w, double rho_ocn, double& ustokes, double& kphil) {
  const double ustar_air = ustar_w * sqrt(rho_ocn / 1.225);
  const double z0_smooth = 0.11 * 1.0e-6 / ustar_air;
  double u = ustar_air * sqrt(1000.0);
  bool converged = false;
  for (int itt = 1; itt <= 20; ++itt) {
    const double alpha_ch = fmn(0.028, 0.0017 * u + -0.005);
    const double z0w = fmx(z0_smooth + alpha_ch * (ustar_air * ustar_air) / GRAVITY, 1.0e-10);
    const double u_new = ustar_air * sm::log(10.0 / z0w) / 0.40;
    if (fabs(u_new - u) <= 1.0e-3 * u) {
      u = u_new;
      converged = true;
      break;
    }
    u = u_new;
  }
  if (!converged) u = 25.82 * ustar_air;
  ustokes = 0.0162 * u;
  const double hm0 = 0.0246 * u * u;
  const double f_mean = 1.296 * 0.877 * GRAVITY / (2.0 * LT_PI * 1.075 * u);
  const double vstokes = 0.125 * LT_PI * 0.667 * f_mean * hm0 * hm0;
  kphil = 0.176 * ustokes / fmx(vstokes, 1.0e-30);
}

__device__ __forceinline__ double one_m_exp_x(double x) {
  if (x < 1.0e-4) return 1.0 - x * (0.5 - x / 6.0);
  return (1.0 - sm::exp(-x)) / x;
}

// epbl_lf17_la (:767-804).
__device__ __noinline__ double lf17_la(double ustar_w, double zsl, double ustokes, double kphil) {
  const double xkz = kphil * zsl;
  const double root2kz = sqrt(2.0 * xkz);
  const double r1 = (0.302 - 1.68 * xkz) * one_m_exp_x(2.0 * xkz);
  const double r3 = (0.1264 + 0.64 * xkz) * one_m_exp_x(5.12 * xkz);
  double r5;
  if (root2kz > 1.0e-3) {
    const double ec1 = sm::erfc(root2kz);
    const double ec2 = sm::erfc(1.6 * root2kz);
    r5 = sqrt(LT_PI) * (root2kz * (-(0.84 * ec1) + 0.2 * ec2) + 0.1182 * (ec2 - ec1) / root2kz);
  } else {
    r5 = -(0.64 * sqrt(LT_PI) * root2kz) + (-0.14184 + 1.0839648 * root2kz * root2kz);
  }
  const double ustokes_sl = fmx(ustokes * (0.715 + r1 + r3 + r5), 1.0e-10);
  return sqrt(ustar_w / ustokes_sl);
}

// epbl_lt_enhance (:818-858).
__device__ __noinline__ double lt_enhance(const EpblParams& c, double la, double b0, double ustar, double bld,
                                          double absf, double mstar) {
  const double vk = c.von_karman;
  const double mld_ek = bld * absf / ustar;
  const double ek_ob = fabs(b0 * vk) / (fmx(absf, 1.0e-20) * ustar * ustar);
  const double mld_ob = fabs(bld * b0 * vk) / cube(ustar);
  double la_mod;
  if (b0 >= 0.0)
    la_mod = la * ((1.0 + fmx(-0.5, c.lac1 * mld_ek)) + c.lac4 * ek_ob + c.lac2 * mld_ob);
  else
    la_mod = la * ((1.0 + fmx(-0.5, c.lac1 * mld_ek)) + c.lac5 * ek_ob + c.lac3 * mld_ob);
  la_mod = fmx(la_mod, 1.0e-10);
  if (c.lt_scheme == RDB_EPBL_LT_ADDITIVE) return mstar + c.lt_coef * sm::pow(la_mod, c.lt_exp);
  const double enh = fmn(c.lt_max, 1.0 + c.lt_coef * sm::pow(la_mod, c.lt_exp));
  return mstar * enh;
}

__device__ __forceinline__ double vstar_of(const EpblParams& c, double tke_here, double conv_perel, double htot,
                                           double mld_guess, double ustar, double dt, double rho0) {
  if (c.vstar_scheme == RDB_EPBL_VSTAR_RH18) {
    const double surf_scale = fmx(0.05, 1.0 - htot / mld_guess);
    return c.vstar_scale_fac * surf_scale *
           (c.vstar_surf_fac * ustar + cbrt_kahan(c.wstar_ustar_coef * conv_perel / (dt * rho0)));
  }
  return c.vstar_scale_fac * cbrt_kahan(tke_here / (dt * rho0));
}

__device__ __forceinline__ double kd_of(const EpblParams& c, double vstar, double htot, double hbs, double absf) {
  if (c.mld_iteration) {
    const double mixlen = fmx(c.min_mix_len, (htot * hbs * vstar) / (c.ekman_scale_coef * absf * htot * hbs + vstar));
    return vstar * c.von_karman * mixlen;
  }
  return vstar * c.von_karman * (htot * hbs * vstar) / (c.ekman_scale_coef * absf * htot * hbs + vstar);
}

// Pre-mixing T/S of a layer and its PE / column-height sensitivities at the
// mid-layer pressure of the stack (the prep sweep, :1087-1183).
struct LevelPrep {
  double t0, s0, dpe_t, dpe_s, dch_t, dch_s;
};

__device__ __forceinline__ LevelPrep level_prep(double hk, double T, double S, double rho0, double& pres) {
  LevelPrep q;
  q.t0 = 0.0;
  q.s0 = 0.0;
  if (hk > 0.0) {
    q.t0 = T;
    q.s0 = S;
  }
  const double dmass = rho0 * hk;
  const double dpres = GRAVITY * dmass;
  const double p_mid = pres + 0.5 * dpres;
  double dsv_dt, dsv_ds;
  specvol_derivs(q.t0, q.s0, p_mid, dsv_dt, dsv_ds);
  q.dpe_t = dmass * p_mid * dsv_dt;
  q.dpe_s = dmass * p_mid * dsv_ds;
  q.dch_t = dmass * dsv_dt;
  q.dch_s = dmass * dsv_ds;
  pres = pres + dpres;
  return q;
}

}  // namespace

// The EPBL solve for the packed columns [p0, p1): kd_int, MLD, B0, La and
// the TKE ledger of each (free surface, no penetrating shortwave).
__global__ void __launch_bounds__(64, 2) k_epbl(StateView s, EpblParams c, EpblView e, const double* __restrict__ fcor,
                                               double rho0, double cp, double dt, int p0, int p1) {
  const int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const int P = s.P, nz = s.nz;
  const double G = GRAVITY, nstar = c.nstar;
  // the surface layer: the first live layer counting down from the top
  int k_sfc = nz - 1;
  for (int k = nz - 1; k >= 0; --k)
    if (LV(s.h, k) > H_VANISHED) {
      k_sfc = k;
      break;
    }
  double h_sum = 0.0;
  for (int k = k_sfc; k >= 0; --k) h_sum = h_sum + LV(s.h, k);
  h_sum = h_sum + H_NEGLECT;
  const double h_sfc = LV(s.h, k_sfc);
  double pres0 = 0.0;
  const LevelPrep top = level_prep(h_sfc, LV(s.T, k_sfc), LV(s.S, k_sfc), rho0, pres0);
  const double dmass_s = rho0 * h_sfc;
  const double pmid_s = 0.0 + 0.5 * (G * dmass_s);
  double dsv_dt_sfc, dsv_ds_sfc;
  specvol_derivs(top.t0, top.s0, pmid_s, dsv_dt_sfc, dsv_ds_sfc);
  // forcing (:1185-1226)
  const double q_t_kin = (1.0 / (rho0 * cp)) * s.Qh[p];
  const double q_s_kin = (1.0 / rho0) * s.Qs[p];
  const double b0 = G * rho0 * (dsv_dt_sfc * q_t_kin + dsv_ds_sfc * q_s_kin);
  e.b0[p] = b0;
  const double ctke_sfc = -(0.5 * G * (rho0 * rho0) * h_sfc * (q_t_kin * dt * dsv_dt_sfc + q_s_kin * dt * dsv_ds_sfc));
  const double ustar = fmx(sqrt(s.tau_tot[p] / rho0), c.ustar_min);
  const double f = fcor[p];
  const double absf = sqrt((1.0 - c.omega_frac) * (f * f) + c.omega_frac * 4.0 * (c.omega * c.omega));
  const double idecay = c.tke_decay * absf / ustar;
  const double mech_in = dt * rho0 * cube(ustar);
  double la_val = 0.0, lt_ustokes = 0.0, lt_kphil = 0.0;
  if (c.use_lt) lf17_wave_state(ustar, rho0, lt_ustokes, lt_kphil);
  double d_wind = 0.0, d_conv = 0.0, d_forcing = 0.0, d_mixing = 0.0, d_mdecay = 0.0, d_cdecay = 0.0;
  double mld_found = 0.0;

  if (s.wet[p] <= 0.0 || h_sum <= 2.0 * H_NEGLECT) {
    for (int k = 0; k <= nz; ++k) LV(e.kd, k) = 0.0;
  } else {
    double min_mld = 0.0, max_mld = h_sum, dmld_min = 0.0, dmld_max = 0.0;
    bool have_min = false, have_max = false;
    double mld_guess = 0.5 * (min_mld + max_mld);
    if (c.mld_use_prev_guess && e.mld[p] > 0.0) mld_guess = fmn(e.mld[p], max_mld);
    const int n_its = c.mld_iteration ? c.mld_max_its : 1;
    for (int k = k_sfc + 1; k <= nz; ++k) LV(e.kd, k) = 0.0;
    LV(e.kd, 0) = 0.0;
    // One sweep of the column for the guess mld_guess.  A guess that is not
    // the last one only needs the MLD, which stops growing once the sweep
    // disconnects from the surface: those sweeps (FULL = false) stop there
    // and write nothing.  The last guess is swept in full (kd_int and the
    // TKE ledger); both kinds run the same arithmetic down to the
    // disconnect, so the MLD they find is the same to the bit.
    auto sweep = [&](auto full_tag, double mld_guess) -> double {
      constexpr bool FULL = decltype(full_tag)::value;
      d_wind = d_conv = d_forcing = d_mixing = d_mdecay = d_cdecay = 0.0;
      double mstar_val = find_mstar(c, b0, ustar, mld_guess, absf);
      if (c.use_lt) {
        la_val = lf17_la(ustar, c.la_frac_hbl * mld_guess, lt_ustokes, lt_kphil);
        mstar_val = lt_enhance(c, la_val, b0, ustar, mld_guess, absf, mstar_val);
      }
      double mech_tke = mstar_val * mech_in;
      d_wind = mech_tke;
      double conv_perel;
      if (ctke_sfc <= 0.0) {
        const double forcing_clip = fmx(ctke_sfc, -mech_tke);
        mech_tke = mech_tke + forcing_clip;
        conv_perel = 0.0;
        d_forcing = forcing_clip;
      } else {
        conv_perel = ctke_sfc;
        d_conv = nstar * ctke_sfc;
      }
      // sweep initialisation (:1330-1350); A is the layer above the
      // interface, B the one below
      double hA = h_sfc;
      LevelPrep A = top;
      double pres = pres0;
      double h_ka = hA + H_NEGLECT;
      double hp_a = h_ka;
      double dpe_t_a = A.dpe_t, dpe_s_a = A.dpe_s, dch_t_a = A.dch_t, dch_s_a = A.dch_s;
      double th_a = h_ka * A.t0, sh_a = h_ka * A.s0;
      double te_lag = 0.0, se_lag = 0.0, kddt_prev = 0.0;
      double htot = hA, z_int = hA, pres_int = G * rho0 * hA;
      double mld_output = hA;
      bool sfc_connected = true;
      for (int ki = k_sfc; ki >= 1; --ki) {
        const int kb = ki - 1;
        const double hB = LV(s.h, kb);
        const LevelPrep B = level_prep(hB, LV(s.T, kb), LV(s.S, kb), rho0, pres);
        h_ka = hA + H_NEGLECT;
        const double h_kb = hB + H_NEGLECT;
        bool sfc_disconnect = false;
        // with no mechanical TKE left the decay changes nothing:
        // (1 - e) * 0 and 0 * e are 0 for the finite e in (0, 1]
        if (mech_tke != 0.0) {
          const double exp_kh = sm::exp(-h_ka * idecay);
          d_mdecay = d_mdecay + (1.0 - exp_kh) * mech_tke;
          mech_tke = mech_tke * exp_kh;
        }
        double nstar_fc = nstar;
        if (conv_perel > 0.0 && absf > 0.0) {
          const double ah = absf * htot;
          nstar_fc = nstar * conv_perel / (conv_perel + 0.2 * sqrt(0.5 * dt * rho0 * cube(ah) * conv_perel));
        }
        const double tot_tke = mech_tke + nstar_fc * conv_perel;
        const double dt_h = dt / fmx(0.5 * (h_ka + h_kb), 1.0e-15 * h_sum);
        double kd_val;
        if (tot_tke <= 0.0 && 0.0 <= (B.dch_t + A.dch_t) * (A.t0 - B.t0) + (B.dch_s + A.dch_s) * (A.s0 - B.s0)) {
          kd_val = 0.0;
          sfc_disconnect = true;
        } else {
          double tke_here = mech_tke + c.wstar_ustar_coef * conv_perel;
          const double hb_hs = (h_sum - z_int) / h_sum;
          const double hbs = fmn(hb_hs, mixlen_shape(c, z_int, mld_guess));
          double kd_g0 = 0.0;
          if (tke_here > 0.0) {
            const double vstar = vstar_of(c, tke_here, conv_perel, htot, mld_guess, ustar, dt, rho0);
            kd_g0 = kd_of(c, vstar, htot, hbs, absf);
          }
          const double hp_b = h_kb;
          const double th_b = h_kb * B.t0, sh_b = h_kb * B.s0;
          const double hps = hp_a + hp_b;
          const double bdt1 = hp_a * hp_b;
          const double dt_c = hp_a * th_b - hp_b * th_a;
          const double ds_c = hp_a * sh_b - hp_b * sh_a;
          double pec_core = hp_b * (dpe_t_a * dt_c + dpe_s_a * ds_c) - hp_a * (B.dpe_t * dt_c + B.dpe_s * ds_c);
          const double colht_core = hp_b * (dch_t_a * dt_c + dch_s_a * ds_c) - hp_a * (B.dch_t * dt_c + B.dch_s * ds_c);
          if (colht_core < 0.0) pec_core = pec_core - pres_int * colht_core;
          const double dkddt = kd_g0 * dt_h;
          const double pe_max = pec_core / (bdt1 * hps);
          if (pe_max < 0.0) {
            tke_here = mech_tke + c.wstar_ustar_coef * (conv_perel - pe_max);
            kd_val = 0.0;
            if (tke_here > 0.0) {
              const double vstar = vstar_of(c, tke_here, conv_perel, htot, mld_guess, ustar, dt, rho0);
              kd_val = kd_of(c, vstar, htot, hbs, absf);
            }
            const double pe_g0 = pec_core * dkddt / (bdt1 * (bdt1 + dkddt * hps));
            double dpe_conv = pec_core * (kd_val * dt_h) / (bdt1 * (bdt1 + (kd_val * dt_h) * hps));
            if (dpe_conv > 0.0) {
              kd_val = kd_g0;
              dpe_conv = pe_g0;
            }
            conv_perel = conv_perel - dpe_conv;
            d_conv = d_conv - nstar * dpe_conv;
            if (sfc_connected) mld_output = mld_output + hB;
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
            if (sfc_connected) mld_output = mld_output + frac_bl * hB;
            if (frac_bl < 1.0) sfc_disconnect = true;
            double r_reduc = 0.0;
            if (tot_tke > 0.0 && tot_tke > tke_used) r_reduc = (tot_tke - tke_used) / tot_tke;
            d_mixing = d_mixing + tke_used;
            d_cdecay = d_cdecay + (1.0 - r_reduc) * (nstar - nstar_fc) * conv_perel;
            mech_tke = r_reduc * mech_tke;
            conv_perel = r_reduc * conv_perel;
          }
        }
        if (!FULL && sfc_disconnect) return mld_output;  // the MLD is final
        if (FULL) LV(e.kd, ki) = kd_val;
        const double kddt_cur = kd_val * dt_h;
        const double b1 = 1.0 / (hp_a + kddt_cur);
        const double c1 = kddt_cur * b1;
        if (ki == k_sfc) {
          te_lag = b1 * (h_ka * A.t0);
          se_lag = b1 * (h_ka * A.s0);
        } else {
          const double te_new = b1 * (h_ka * A.t0 + kddt_prev * te_lag);
          const double se_new = b1 * (h_ka * A.s0 + kddt_prev * se_lag);
          te_lag = te_new;
          se_lag = se_new;
        }
        hp_a = h_kb + (hp_a * b1) * kddt_cur;
        dpe_t_a = B.dpe_t + c1 * dpe_t_a;
        dpe_s_a = B.dpe_s + c1 * dpe_s_a;
        dch_t_a = B.dch_t + c1 * dch_t_a;
        dch_s_a = B.dch_s + c1 * dch_s_a;
        th_a = h_kb * B.t0 + kddt_cur * te_lag;
        sh_a = h_kb * B.s0 + kddt_cur * se_lag;
        kddt_prev = kddt_cur;
        if (sfc_disconnect) {
          htot = hB;
          sfc_connected = false;
        } else {
          htot = htot + hB;
        }
        z_int = z_int + hB;
        pres_int = pres_int + G * rho0 * hB;
        hA = hB;
        A = B;
      }
      d_mdecay = d_mdecay + mech_tke;
      d_cdecay = d_cdecay + nstar * conv_perel;
      return mld_output;
    };
    for (int obl_it = 1; obl_it <= n_its; ++obl_it) {
      // a guess known to be the last is swept in full at once
      if (!c.mld_iteration || obl_it == n_its) {
        mld_found = sweep(std::true_type{}, mld_guess);
        break;
      }
      mld_found = sweep(std::false_type{}, mld_guess);
      if (fabs(mld_found - mld_guess) < c.mld_tol) {
        sweep(std::true_type{}, mld_guess);
        break;
      }
      if (mld_found > mld_guess) {
        min_mld = mld_guess;
        dmld_min = mld_found - mld_guess;
        have_min = true;
      } else {
        max_mld = mld_guess;
        dmld_max = mld_found - mld_guess;
        have_max = true;
      }
      if (c.mld_bisection)
        mld_guess = 0.5 * (min_mld + max_mld);
      else if (have_min && have_max && obl_it > 2 && ((obl_it - 1) % 4) > 0)
        mld_guess = min_mld + dmld_min * (max_mld - min_mld) / (dmld_min - dmld_max);
      else if (mld_found > min_mld && mld_found < max_mld)
        mld_guess = mld_found;
      else
        mld_guess = 0.5 * (min_mld + max_mld);
    }
  }
  e.mld[p] = mld_found;
  e.la[p] = la_val;
  if (c.tke_diags) {
    LV(e.tke, 0) = d_wind / dt;
    LV(e.tke, 1) = d_conv / dt;
    LV(e.tke, 2) = d_forcing / dt;
    LV(e.tke, 3) = d_mixing / dt;
    LV(e.tke, 4) = d_mdecay / dt;
    LV(e.tke, 5) = d_cdecay / dt;
  }
}

// The coefficient build after the surface scheme, one pass per column from
// the top: PP81 at each interface (when the surface scheme did not need it
// first, i.e. with EPBL), the EPBL merge, KV_ML_INVZ2, convective
// adjustment, then the heat/salt split with double diffusion, the guard and
// the floors/ceilings of vmix_assemble (rdb_ocean_vmix.F90:2265-2385).  Each
// is a per-interface update applied in the reference's order, so one pass
// gives the bits of the separate sweeps; kv, kt and ks are written once.
__global__ void k_coeff_tail(StateView s, Params pr, ContribParams c, EpblView e, int do_pp81, int p0, int p1,
                             unsigned int* guard_count) {
  const int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const int P = s.P, nz = s.nz;
  const double bld = c.epbl ? e.mld[p] : s.bl[p];
  const bool invz2 = c.kv_ml_invz2 > 0.0 && c.hmix > 0.0;
  const double hmix_sq = c.hmix * c.hmix;
  const double dz_min_safe = 0.5 * c.hmix / (double)(nz > 1 ? nz : 1);
  bool inv_done = !invz2;
  unsigned int bad = 0;
  auto finish = [&](int k, double kv, double kt, double kd_t, double kd_s) {
    double ks = kt;
    if (pr.ddiff) {
      ks = kt + kd_s;
      kt = kt + kd_t;
    }
    if (pr.guard && (isnan(kv) || isnan(kt) || isnan(ks) || kv < 0.0 || kt < 0.0 || ks < 0.0)) ++bad;
    if (k >= 1 && k < nz) {
      kv = fmn(fmx(kv, pr.kv_bg), pr.kv_max);
      kt = fmn(fmx(kt, pr.kt_bg), pr.kd_max);
      ks = fmn(fmx(ks, pr.ks_bg), pr.kd_max);
    }
    LV(s.kv, k) = kv;
    LV(s.kt, k) = kt;
    LV(s.ks, k) = ks;
  };
  if (do_pp81) {
    finish(nz, 0.0, 0.0, 0.0, 0.0);
  } else {
    finish(nz, LV(s.kv, nz), LV(s.kt, nz), 0.0, 0.0);
  }
  double z = 0.0;
  double h_up = LV(s.h, nz - 1), rho_up = LV(s.rho, nz - 1);
  double ub_up = do_pp81 ? LV(s.ub, nz - 1) : 0.0, vb_up = do_pp81 ? LV(s.vb, nz - 1) : 0.0;
  double T_up = LV(s.T, nz - 1), S_up = LV(s.S, nz - 1);
  for (int k = nz - 1; k >= 1; --k) {
    const double h_dn = LV(s.h, k - 1), rho_dn = LV(s.rho, k - 1);
    const double T_dn = LV(s.T, k - 1), S_dn = LV(s.S, k - 1);
    double ub_dn = 0.0, vb_dn = 0.0;
    double kv, kt;
    if (do_pp81) {
      ub_dn = LV(s.ub, k - 1);
      vb_dn = LV(s.vb, k - 1);
      pp81_iface<double>(pr, ub_dn, vb_dn, ub_up, vb_up, h_dn, h_up, rho_up - rho_dn, kv, kt);
    } else {
      kv = LV(s.kv, k);
      kt = LV(s.kt, k);
    }
    if (c.epbl) {
      const double kd = LV(e.kd, k);
      if (c.combine == RDB_EPBL_COMBINE_MAX) {
        kt = fmx(kt, kd);
        kv = fmx(kv, c.prandtl * kd);
      } else {
        kt = kt + kd;
        kv = kv + c.prandtl * kd;
      }
    }
    z = z + h_up;
    if (!inv_done) {
      if (z >= c.hmix) {
        inv_done = true;
      } else {
        const double m = fmx(z, dz_min_safe);
        kv = kv + c.kv_ml_invz2 * hmix_sq / (m * m);
      }
    }
    if (c.conv) {
      const double dz_face = 0.5 * (h_dn + h_up);
      if (dz_face > H_VANISHED && !(z < bld)) {
        const double n2 = -(GRAVITY * (rho_up - rho_dn) / (c.rho0 * dz_face));
        if (n2 < c.n2_thresh) {
          kt = fmx(kt, c.conv_kd);
          kv = fmx(kv, c.conv_prandtl * c.conv_kd);
        }
      }
    }
    double kd_t = 0.0, kd_s = 0.0;
    if (pr.ddiff) ddiff_iface(pr, h_up, h_dn, T_up, T_dn, S_up, S_dn, kd_t, kd_s);
    finish(k, kv, kt, kd_t, kd_s);
    h_up = h_dn;
    rho_up = rho_dn;
    ub_up = ub_dn;
    vb_up = vb_dn;
    T_up = T_dn;
    S_up = S_dn;
  }
  if (do_pp81)
    finish(0, 0.0, 0.0, 0.0, 0.0);
  else
    finish(0, LV(s.kv, 0), LV(s.kt, 0), 0.0, 0.0);
  if (bad) atomicAdd(guard_count, bad);
}

// The device safe math over n arguments (fn: 0 exp, 1 log, 2 pow, 3 erfc),
// checked against the reference's golden rows.
__global__ void k_safe_math_eval(int fn, const double* x, const double* y, double* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r;
  switch (fn) {
    case 0: r = sm::exp(x[i]); break;
    case 1: r = sm::log(x[i]); break;
    case 2: r = sm::pow(x[i], y[i]); break;
    default: r = sm::erfc(x[i]); break;
  }
  out[i] = r;
}

}  // namespace colphys
