// colphys_kernels.h -- device state view and kernel declarations.
#pragma once

#include "colphys_device.h"
#include "../../include/rdb_colphys.h"

namespace colphys {

// Device pointers of one handle's packed state (pitch P, level-major).
struct StateView {
  int P, nz, n_passive, pad;
  double* h;        // [nz][P]   layer thickness
  double* ub;       // [nz][P]   centred x velocity
  double* vb;       // [nz][P]   centred y velocity
  double* T;        // [nz][P]
  double* S;        // [nz][P]
  double* tr;       // [n_passive][nz][P]
  double* rho;      // [nz][P]   (staged path)
  double* kv;       // [nz+1][P]
  double* kt;       // [nz+1][P]
  double* ks;       // [nz+1][P]
  double* gamma_t;  // [nz+1][P] (staged path)
  double* gamma_s;  // [nz+1][P] (staged path)
  double* fa;       // [nz][P]   tridiagonal scratch
  double* fb;       // [nz][P]
  double* fc;       // [nz][P]
  double* tau_tot;  // [P] stress_mag + stress_shelf
  double* Qh;       // [P]
  double* Qs;       // [P]
  double* wet;      // [P]
  double* bl;       // [P] bl_depth
  double* b0;       // [P]
  int* kbl;         // [P] KPP crossing layer (1-based, 0 none)
  int* ktop;        // [P] 0-based first-live scan start
  int* kb;          // [P] first live layer from the bed (nz: none)
  int* fdirty;      // [P] lowest bed-filler interface holding a KPP value
  // vdiff budget slots heat/salt_budget_vdiff (null: not kept): the content
  // change h*c_new - h*c_old of the tracer solve, accumulated per cell
  double* budT;     // [nz][P]
  double* budS;     // [nz][P]
  double* pold;     // [2][nz][P] h*c_old of T and S between the two sweeps
};

__global__ void k_pack_layer(const double* src, double* dst, const int* col, int ncol, int P, int nlev,
                             long long plane);
__global__ void k_permute_d(const double* src, double* dst, const int* newpos, int n, int P, int nlev);
__global__ void k_permute_i(const int* src, int* dst, const int* newpos, int n, int P, int nlev);
__global__ void k_unpack_layer(const double* src, double* dst, const int* col, int ncol, int P, int nlev,
                               long long plane);
__global__ void k_unpack_int(const int* src, double* dst, const int* col, int ncol);
__global__ void k_pack_centre_x(const double* uf, double* dst, const int* col, int ncol, int P, int nlev, int nx,
                                int ny);
__global__ void k_pack_centre_y(const double* vf, double* dst, const int* col, int ncol, int P, int nlev, int nx,
                                int ny);
__global__ void k_stress_mag(const double* txc, const double* tyc, double* smag, const double* shelf,
                             double* tau_tot, int ncol);
__global__ void k_cbrt_eval(const double* x, double* y, int n);
__global__ void k_eos(StateView s, Params pr, int p0, int p1);
__global__ void k_pp81(StateView s, Params pr, int p0, int p1);
__global__ void k_kpp(StateView s, Params pr, int p0, int p1);
__global__ void k_split_assemble(StateView s, Params pr, int p0, int p1, unsigned int* guard_count);
__global__ void k_surface_flux(StateView s, Params pr, double dt, int p0, int p1);
__global__ void k_vdiff(StateView s, Params pr, double dt, int p0, int p1, int from_kb);
__global__ void k_nonlocal(StateView s, double dt, int p0, int p1);
// Learned-closure inputs of the fused step: the network's outputs stored
// output-major (Yt[n * ldy + p], p the packed wet column), their
// normalisation, and a per-column guard flag (non-zero: plain path).
// Learned-closure predictions for the fused step: Yt is output-major
// (output n of packed column p at Yt[n * ldy + p], normalised by mean/stdv),
// guard[p] != 0 sends column p down the plain path.  The network is laid out
// on nz + lev_off levels with the surfaces aligned: interface kf of this
// grid is the network's interface kf + lev_off.
struct SurrView {
  const float* Yt;
  long long ldy;
  const float* mean;
  const float* stdv;
  const unsigned char* guard;
  int lev_off;
  int pad;
  const float* tmin;  // the range of each target over the training set: a
  const float* tmax;  // prediction outside it is clamped to it
};

__global__ void k_prepare(StateView s, Params pr, int p0, int p1);
template <typename R, int NTR, bool TWO_FACTOR, bool SURR>
__global__ void k_column_step(StateView s, Params pr, double dt, int p0, int p1, unsigned int* guard_count, int M,
                              SurrView sv);
template <typename R, int NTR, bool TWO_FACTOR, bool SURR>
__global__ void k_column_step_ens(const StateView* views, const SurrView* svs, const int* n_wet, Params pr, double dt,
                                  const int2* work, int n_work, unsigned int* guard_count, int M);
__global__ void k_totals(StateView s, int p0, int p1, unsigned long long* acc);

// EPBL knobs (ocean_epbl_t, rdb_ocean_epbl.F90:148-330), copied from the
// config at create.
struct EpblParams {
  int mstar_scheme, vstar_scheme, mld_iteration, mld_max_its, mld_bisection, mld_use_prev_guess, use_lt, lt_scheme;
  int tke_diags, pad;
  double mstar_const, mstar_cap, mstar_coef1, c_ek, mstar_conv_adj, cn1, cn2, cn3, cs1, cs2;
  double nstar, tke_decay, wstar_ustar_coef, vstar_scale_fac, vstar_surf_fac, von_karman, ekman_scale_coef;
  double min_mix_len, mixlen_exponent, translay, mld_tol, omega, omega_frac, ustar_min;
  double lt_coef, lt_exp, lt_max, la_frac_hbl, lac1, lac2, lac3, lac4, lac5;
};
// EPBL state, packed: kd [nz+1][P], tke [6][P] (wind, conv, forcing,
// mixing, mech_decay, conv_decay), the rest [P].
struct EpblView {
  double *kd, *mld, *b0, *la, *tke;
};
// The per-interface contributors after the surface scheme.
struct ContribParams {
  int epbl, conv, combine, pad;
  double prandtl, kv_ml_invz2, hmix, conv_kd, conv_prandtl, n2_thresh, rho0;
};
__global__ void k_epbl(StateView s, EpblParams c, EpblView e, const double* fcor, double rho0, double cp, double dt,
                       int p0, int p1);
__global__ void k_coeff_tail(StateView s, Params pr, ContribParams c, EpblView e, int do_pp81, int p0, int p1,
                             unsigned int* guard_count);
__global__ void k_safe_math_eval(int fn, const double* x, const double* y, double* out, int n);
// Allow every fused-step kernel lds_bytes of dynamic LDS (needed above 64 KiB).
hipError_t allow_full_lds(int lds_bytes);

}  // namespace colphys
