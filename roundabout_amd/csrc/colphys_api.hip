// colphys_api.hip -- the C ABI (include/rdb_colphys.h): handle lifecycle,
// state transfer, launches, hipGraph capture of the step loop, launch-shape
// autotuning, reproducible totals and the error ring.
//
// Conventions follow the reference ABI (include/rdb_ocean.h): int status on
// every entry point, never abort, a ring of error messages, an opaque handle
// that is NULL on failure, one live handle per process
// (src/api/rdb_ocean_api.F90:9-45).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/rdb_colphys.h"
#include "colphys_handle.h"
#include "colphys_kernels.h"

using colphys::Params;
using colphys::StateView;

namespace {

std::deque<std::string> g_ring;  // index 0 = most recent
bool g_live = false;

int push_error(int status, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_ring.push_front(buf);
  while (g_ring.size() > 32) g_ring.pop_back();
  return status;
}

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess)                                                                      \
      return push_error(RDB_COLPHYS_ERR_SETUP, "%s failed: %s (%s:%d)", #expr,                 \
                        hipGetErrorString(_e), __FILE__, __LINE__);                            \
  } while (0)

enum FieldKind { F_LAYER, F_IFACE, F_FACE_X, F_FACE_Y, F_TAU_X, F_TAU_Y, F_COL, F_COL_INT };

struct FieldSpec {
  FieldKind kind;
  bool settable;
  bool gettable;
};

}  // namespace

// error ring shared with the MLP entry points
int rdb_colphys_push_error(int status, const char* msg) { return push_error(status, "%s", msg); }

// the handle (struct rdb_colphys) is defined in colphys_handle.h

namespace {

Params make_params(const rdb_colphys_config& c) {
  Params p;
  p.nu0 = c.pp81_nu0;
  p.nu_bg = c.pp81_nu_bg;
  p.kappa_bg = c.pp81_kappa_bg;
  p.pp_alpha = c.pp81_alpha;
  p.rho0 = c.rho0;
  p.shear2_floor = c.shear2_floor;
  p.ri_crit = c.ri_crit;
  p.c_vt2 = c.c_vt2;
  p.cs_nonlocal = c.cs_nonlocal;
  p.alpha_T = c.alpha_T;
  p.beta_S = c.beta_S;
  p.p_ref = c.p_ref;
  p.cp = c.cp;
  p.kv_bg = c.kv_bg;
  p.kt_bg = c.kt_bg;
  p.ks_bg = c.ks_bg;
  p.kv_max = c.kv_max;
  p.kd_max = c.kd_max;
  p.use_kpp = c.use_kpp;
  p.has_heat = c.has_heat;
  p.has_salt = c.has_salt;
  p.use_harmonic = c.use_harmonic;
  p.guard = c.vmix_guard;
  p.ddiff = c.ddiff_enable ? 1 : 0;
  p.ddiff_k90 = c.ddiff_use_k90 ? 1 : 0;
  p.dd_rmax = c.ddiff_strat_param_max;
  p.dd_kappa_s = c.ddiff_kappa_s;
  p.dd_exp1 = c.ddiff_exp1;
  p.dd_exp2 = c.ddiff_exp2;
  p.dd_p1 = c.ddiff_param1;
  p.dd_p2 = c.ddiff_param2;
  p.dd_p3 = c.ddiff_param3;
  p.dd_mol = c.ddiff_mol_diff;
  return p;
}

colphys::EpblParams make_epbl_params(const rdb_colphys_config& c) {
  colphys::EpblParams e{};
  e.mstar_scheme = c.epbl_mstar_scheme;
  e.vstar_scheme = c.epbl_vstar_scheme;
  e.mld_iteration = c.epbl_mld_iteration ? 1 : 0;
  e.mld_max_its = c.epbl_mld_max_its;
  e.mld_bisection = c.epbl_mld_bisection ? 1 : 0;
  e.mld_use_prev_guess = c.epbl_mld_use_prev_guess ? 1 : 0;
  e.use_lt = c.epbl_use_lt ? 1 : 0;
  e.lt_scheme = c.epbl_lt_scheme;
  e.tke_diags = c.epbl_tke_diags ? 1 : 0;
  e.mstar_const = c.epbl_mstar;
  e.mstar_cap = c.epbl_mstar_cap;
  e.mstar_coef1 = c.epbl_mstar_coef1;
  e.c_ek = c.epbl_c_ek;
  e.mstar_conv_adj = c.epbl_mstar_conv_adj;
  e.cn1 = c.epbl_rh18_cn1;
  e.cn2 = c.epbl_rh18_cn2;
  e.cn3 = c.epbl_rh18_cn3;
  e.cs1 = c.epbl_rh18_cs1;
  e.cs2 = c.epbl_rh18_cs2;
  e.nstar = c.epbl_nstar;
  e.tke_decay = c.epbl_tke_decay;
  e.wstar_ustar_coef = c.epbl_wstar_ustar_coef;
  e.vstar_scale_fac = c.epbl_vstar_scale_fac;
  e.vstar_surf_fac = c.epbl_vstar_surf_fac;
  e.von_karman = c.epbl_von_karman;
  e.ekman_scale_coef = c.epbl_ekman_scale_coef;
  e.min_mix_len = c.epbl_min_mix_len;
  e.mixlen_exponent = c.epbl_mixlen_exponent;
  e.translay = c.epbl_translay_scale;
  e.mld_tol = c.epbl_mld_tol;
  e.omega = c.epbl_omega;
  e.omega_frac = c.epbl_omega_frac;
  e.ustar_min = c.epbl_ustar_min;
  e.lt_coef = c.epbl_lt_enhance_coef;
  e.lt_exp = c.epbl_lt_enhance_exp;
  e.lt_max = c.epbl_lt_max_enhance;
  e.la_frac_hbl = c.epbl_la_frac_hbl;
  e.lac1 = c.epbl_lt_lac1;
  e.lac2 = c.epbl_lt_lac2;
  e.lac3 = c.epbl_lt_lac3;
  e.lac4 = c.epbl_lt_lac4;
  e.lac5 = c.epbl_lt_lac5;
  return e;
}

// Whether a config turns on a contributor that the fused step does not hold.
bool has_extras(const rdb_colphys_config& c) {
  return c.epbl_enable || c.conv_enable || c.ddiff_enable || (c.kv_ml_invz2 > 0.0 && c.hmix_fixed > 0.0);
}

// Configure-time checks of the contributors off at the defaults
// (configure_ocean_epbl rdb_ocean_setup_vmix.F90:552-719,
// configure_ocean_conv :1200-1260, configure_ocean_ddiff :1262-1330).
int validate_extras(const rdb_colphys_config& c) {
  if (c.conv_enable) {
    if (c.conv_kd < 0.0) return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_conv_nml: kd_conv must be >= 0");
    if (c.conv_prandtl <= 0.0)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_conv_nml: prandtl_conv must be > 0");
  }
  if (c.ddiff_enable) {
    if (c.ddiff_strat_param_max <= 1.0)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                        "ocean_ddiff_nml: strat_param_max must be > 1 (the fingering form divides by "
                        "strat_param_max - 1)");
    if (c.ddiff_kappa_s < 0.0 || c.ddiff_mol_diff < 0.0)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_ddiff_nml: kappa_ddiff_s and mol_diff must be >= 0");
  }
  if (c.epbl_enable) {
    if (c.epbl_mstar_scheme < RDB_EPBL_MSTAR_CONSTANT || c.epbl_mstar_scheme > RDB_EPBL_MSTAR_RH18)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_epbl_nml: unknown mstar_scheme %d (constant/om4/rh18)",
                        c.epbl_mstar_scheme);
    if (c.epbl_vstar_scheme != RDB_EPBL_VSTAR_CUBE_ROOT && c.epbl_vstar_scheme != RDB_EPBL_VSTAR_RH18)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_epbl_nml: unknown vel_scale_scheme %d (cube_root/rh18)",
                        c.epbl_vstar_scheme);
    if (c.epbl_combine < RDB_EPBL_COMBINE_ADD || c.epbl_combine > RDB_EPBL_COMBINE_MAX_SHEAR)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_epbl_nml: unknown combine %d (add/max/max_shear)",
                        c.epbl_combine);
    if (c.epbl_use_lt && c.epbl_lt_scheme != RDB_EPBL_LT_RESCALE && c.epbl_lt_scheme != RDB_EPBL_LT_ADDITIVE)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_epbl_nml: unknown lt_scheme %d (rescale/additive)",
                        c.epbl_lt_scheme);
    if (c.epbl_mld_iteration && (c.epbl_translay_scale < 0.0 || c.epbl_translay_scale >= 1.0))
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                        "ocean_epbl_nml: translay_scale must be in [0,1) when mld_iteration is on");
    if (c.epbl_combine == RDB_EPBL_COMBINE_MAX_SHEAR && c.use_kpp)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                        "ocean_epbl_nml: combine='max_shear' is refused with use_kpp=.true.; set use_kpp=.false.");
    if (c.epbl_mld_iteration && c.epbl_mld_max_its < 1)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_epbl_nml: mld_max_its must be >= 1");
  }
  if (has_extras(c) && c.precision != RDB_COLPHYS_PREC_FP64)
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                      "EPBL, KV_ML_INVZ2, convection and double diffusion run in FP64 only (precision 0)");
  return RDB_COLPHYS_OK;
}

bool lookup_field(const char* name, int n_passive, FieldSpec& spec, int& tracer) {
  tracer = -1;
  struct E {
    const char* n;
    FieldSpec s;
  };
  static const E table[] = {
      {"h_layer", {F_LAYER, true, true}},
      {"u_face_x_layer", {F_FACE_X, true, true}},
      {"v_face_y_layer", {F_FACE_Y, true, true}},
      {"temperature", {F_LAYER, true, true}},
      {"salinity", {F_LAYER, true, true}},
      {"tau_x", {F_TAU_X, true, true}},
      {"tau_y", {F_TAU_Y, true, true}},
      {"stress_shelf", {F_COL, true, true}},
      {"Q_heat", {F_COL, true, true}},
      {"Q_salt", {F_COL, true, true}},
      {"bl_depth", {F_COL, true, true}},
      {"kv", {F_IFACE, true, true}},
      {"kt", {F_IFACE, true, true}},
      {"ks", {F_IFACE, true, true}},
      {"gamma_t", {F_IFACE, false, true}},
      {"gamma_s", {F_IFACE, false, true}},
      {"rho_layer", {F_LAYER, true, true}},
      {"stress_mag", {F_COL, false, true}},
      {"b0", {F_COL, false, true}},
      {"kbl", {F_COL_INT, false, true}},
      {"coriolis", {F_COL, true, true}},
      {"wet_t", {F_COL, false, true}},
      {"heat_budget_vdiff", {F_LAYER, true, true}},
      {"salt_budget_vdiff", {F_LAYER, true, true}},
      // EPBL (ocean_epbl_t, rdb_ocean_epbl.F90:283-312), kept when it is on
      {"mld", {F_COL, true, true}},
      {"kd_epbl", {F_IFACE, false, true}},
      {"la", {F_COL, false, true}},
      {"b0_epbl", {F_COL, false, true}},
      {"tke_wind", {F_COL, false, true}},
      {"tke_conv", {F_COL, false, true}},
      {"tke_forcing", {F_COL, false, true}},
      {"tke_mixing", {F_COL, false, true}},
      {"tke_mech_decay", {F_COL, false, true}},
      {"tke_conv_decay", {F_COL, false, true}},
  };
  for (const E& e : table)
    if (std::strcmp(e.n, name) == 0) {
      spec = e.s;
      return true;
    }
  if (std::strncmp(name, "tracer_", 7) == 0) {
    int t = std::atoi(name + 7);
    if (t >= 1 && t <= n_passive) {
      spec = {F_LAYER, true, true};
      tracer = t - 1;
      return true;
    }
  }
  return false;
}

bool valid(rdb_colphys* h) { return h != nullptr && g_live; }

double* layer_ptr(rdb_colphys* h, const char* name, int tracer) {
  StateView& s = h->s;
  if (tracer >= 0) return s.tr + (size_t)tracer * (size_t)s.nz * (size_t)s.P;
  if (!std::strcmp(name, "h_layer")) return s.h;
  if (!std::strcmp(name, "temperature")) return s.T;
  if (!std::strcmp(name, "salinity")) return s.S;
  if (!std::strcmp(name, "rho_layer")) return s.rho;
  if (!std::strcmp(name, "kv")) return s.kv;
  if (!std::strcmp(name, "kt")) return s.kt;
  if (!std::strcmp(name, "ks")) return s.ks;
  if (!std::strcmp(name, "gamma_t")) return s.gamma_t;
  if (!std::strcmp(name, "gamma_s")) return s.gamma_s;
  if (!std::strcmp(name, "stress_shelf")) return h->d_shelf;
  if (!std::strcmp(name, "Q_heat")) return s.Qh;
  if (!std::strcmp(name, "Q_salt")) return s.Qs;
  if (!std::strcmp(name, "bl_depth")) return s.bl;
  if (!std::strcmp(name, "stress_mag")) return h->d_smag;
  if (!std::strcmp(name, "b0")) return s.b0;
  if (!std::strcmp(name, "coriolis")) return h->d_coriolis;
  if (!std::strcmp(name, "wet_t")) return s.wet;
  if (!std::strcmp(name, "heat_budget_vdiff")) return s.budT;
  if (!std::strcmp(name, "salt_budget_vdiff")) return s.budS;
  const colphys::EpblView& e = h->ev;
  if (!std::strcmp(name, "mld")) return e.mld;
  if (!std::strcmp(name, "kd_epbl")) return e.kd;
  if (!std::strcmp(name, "la")) return e.la;
  if (!std::strcmp(name, "b0_epbl")) return e.b0;
  static const char* const tke_names[6] = {"tke_wind",   "tke_conv", "tke_forcing",
                                           "tke_mixing", "tke_mech_decay", "tke_conv_decay"};
  for (int q = 0; q < 6; ++q)
    if (!std::strcmp(name, tke_names[q])) return e.tke ? e.tke + (size_t)q * (size_t)s.P : nullptr;
  return nullptr;
}

int dalloc(rdb_colphys* h, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess)
    return push_error(RDB_COLPHYS_ERR_SETUP, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  h->allocations.push_back(*p);
  e = hipMemsetAsync(*p, 0, bytes ? bytes : 8, h->stream);
  if (e != hipSuccess) return push_error(RDB_COLPHYS_ERR_SETUP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
  return RDB_COLPHYS_OK;
}

void release(rdb_colphys* h) {
  for (rdb_colphys* m : h->members) {
    release(m);
    delete m;
  }
  h->members.clear();
  colphys::surr_free(h->surr);
  h->surr = nullptr;
  colphys::mom_release(h);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  for (hipEvent_t e : h->events) (void)hipEventDestroy(e);
  h->events.clear();
  for (void* p : h->allocations) (void)hipFree(p);
  h->allocations.clear();
  if (h->d_work) (void)hipFree(h->d_work);
  h->d_work = nullptr;
  if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
}

inline dim3 grid1(int n, int block) { return dim3((unsigned)((n + block - 1) / block)); }

int launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return push_error(RDB_COLPHYS_ERR_SETUP, "launch of %s failed: %s", what, hipGetErrorString(e));
  return RDB_COLPHYS_OK;
}

}  // namespace

namespace colphys {

// Launch the fused step over packed columns [p0, p1).
int launch_step(rdb_colphys* h, double dt, int p0, int p1) {
  if (p1 <= p0) return RDB_COLPHYS_OK;
  const bool two = (h->cfg.kt_bg != h->cfg.ks_bg);
  const bool split = (h->precision == RDB_COLPHYS_PREC_SPLIT);
  const int ntr = h->n_passive;
  dim3 g = grid1(p1 - p0, h->block), b(h->block);
  const int M = h->lds_levels;
  const size_t shmem = (size_t)h->block * (size_t)M * (size_t)(3 + std::min(ntr, 4) + (two ? 1 : 0)) * 8;
  const colphys::SurrView nosv{};  // the plain path reads none of it
#define LAUNCH(R, N, TF)                                                                                        \
  hipLaunchKernelGGL((colphys::k_column_step<R, N, TF, false>), g, b, shmem, h->stream, h->s, h->pr, dt, p0, p1, \
                     h->d_guard, M, nosv)
#define BY_NTR(R, TF)               \
  switch (ntr) {                    \
    case 0: LAUNCH(R, 0, TF); break; \
    case 1: LAUNCH(R, 1, TF); break; \
    case 2: LAUNCH(R, 2, TF); break; \
    case 3: LAUNCH(R, 3, TF); break; \
    default: LAUNCH(R, 4, TF); break; \
  }
  if (split) {
    if (two) { BY_NTR(float, true) } else { BY_NTR(float, false) }
  } else {
    if (two) { BY_NTR(double, true) } else { BY_NTR(double, false) }
  }
#undef BY_NTR
#undef LAUNCH
  return launch_check("k_column_step");
}

// Every wet column of the handle: the single-grid launch, or for an ensemble
// one persistent launch over the (member, column chunk) work list.  With surr
// the members' learned-closure views (d_svs) must be current.
// (Re)build the ensemble's work list for the current block size; outside
// any stream capture (it allocates and copies).
int ensure_work(rdb_colphys* h) {
  const int B = h->block;
  if (h->members.empty() || (h->work_block == B && h->d_work)) return RDB_COLPHYS_OK;
  {
    // chunks of B packed columns; members interleaved so that the deep
    // columns (each member is sorted deepest first) spread over the grid
    std::vector<int2> work;
    int maxc = 0;
    for (rdb_colphys* m : h->members) maxc = std::max(maxc, (m->n_wet + B - 1) / B);
    for (int c = 0; c < maxc; ++c)
      for (size_t m = 0; m < h->members.size(); ++m)
        if (c * B < h->members[m]->n_wet) work.push_back(make_int2((int)m, c * B));
    if (h->d_work) {
      HIP_TRY(hipStreamSynchronize(h->stream));
      HIP_TRY(hipFree(h->d_work));
      h->d_work = nullptr;
    }
    HIP_TRY(hipMalloc((void**)&h->d_work, sizeof(int2) * std::max<size_t>(work.size(), 1)));
    HIP_TRY(hipMemcpy(h->d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice));
    h->n_work = (int)work.size();
    h->work_block = B;
  }
  int dev = 0, n_cu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  // persistent grid: up to 2048 lanes per CU, never more workgroups than items
  h->persist_grid = std::max(1, std::min(h->n_work, std::max(n_cu, 1) * std::max(1, 2048 / B)));
  return RDB_COLPHYS_OK;
}

int launch_step_all(rdb_colphys* h, double dt, bool surr) {
  if (h->members.empty()) return surr ? push_error(RDB_COLPHYS_ERR_SETUP, "internal: single-grid surrogate launch")
                                      : launch_step(h, dt, 0, h->n_wet);
  int st = ensure_work(h);
  if (st) return st;
  if (h->n_work == 0) return RDB_COLPHYS_OK;
  const int B = h->block, grid = h->persist_grid;
  const bool two = (h->cfg.kt_bg != h->cfg.ks_bg);
  const bool split = (h->precision == RDB_COLPHYS_PREC_SPLIT);
  const int ntr = h->n_passive;
  const int M = h->lds_levels;
  const size_t shmem = (size_t)B * (size_t)M * (size_t)(3 + std::min(ntr, 4) + (two ? 1 : 0)) * 8;
#define LAUNCH(R, N, TF, SU)                                                                                      \
  hipLaunchKernelGGL((colphys::k_column_step_ens<R, N, TF, SU>), dim3(grid), dim3(B), shmem, h->stream, h->d_views, \
                     h->d_svs, h->d_nwet, h->pr, dt, h->d_work, h->n_work, h->d_guard, M)
#define BY_NTR(R, TF, SU)                \
  switch (ntr) {                         \
    case 0: LAUNCH(R, 0, TF, SU); break; \
    case 1: LAUNCH(R, 1, TF, SU); break; \
    case 2: LAUNCH(R, 2, TF, SU); break; \
    case 3: LAUNCH(R, 3, TF, SU); break; \
    default: LAUNCH(R, 4, TF, SU); break; \
  }
  if (surr) {
    if (two) { BY_NTR(double, true, true) } else { BY_NTR(double, false, true) }
  } else if (split) {
    if (two) { BY_NTR(float, true, false) } else { BY_NTR(float, false, false) }
  } else {
    if (two) { BY_NTR(double, true, false) } else { BY_NTR(double, false, false) }
  }
#undef BY_NTR
#undef LAUNCH
  return launch_check("k_column_step_ens");
}

// Staged sequence (one kernel per reference routine) over [p0, p1), in the
// order of vmix_build_coefficients (rdb_ocean_dyn.F90:1473-1634): PP81, the
// KPP overlay or the EPBL solve, the EPBL merge + KV_ML_INVZ2 + convective
// adjustment, the heat/salt split (double diffusion) and the assembly gate.
int staged_coefficients(rdb_colphys* h, int p0, int p1, double dt) {
  if (p1 <= p0) return RDB_COLPHYS_OK;
  dim3 g = grid1(p1 - p0, h->block), b(h->block);
  if (h->extras) {
    // EPBL reads only h, T and S: it runs first, and PP81 folds into the
    // one-pass tail (merge, KV_ML_INVZ2, convection, split, assemble); the
    // KPP overlay needs PP81's values first, so PP81 then stays separate
    // (over every column: a dry column still books its B0)
    if (h->cfg.epbl_enable)
      hipLaunchKernelGGL(colphys::k_epbl, grid1(h->n_all - p0, 64), dim3(64), 0, h->stream, h->s, h->ep, h->ev,
                         h->d_coriolis, h->cfg.rho0, h->cfg.cp, dt, p0, h->n_all);
    const int fold_pp81 = h->cfg.use_kpp ? 0 : 1;
    if (!fold_pp81) {
      hipLaunchKernelGGL(colphys::k_pp81, g, b, 0, h->stream, h->s, h->pr, p0, p1);
      hipLaunchKernelGGL(colphys::k_kpp, g, b, 0, h->stream, h->s, h->pr, p0, p1);
    }
    hipLaunchKernelGGL(colphys::k_coeff_tail, g, b, 0, h->stream, h->s, h->pr, h->cpar, h->ev, fold_pp81, p0, p1,
                       h->d_guard);
    return launch_check("coefficients");
  }
  hipLaunchKernelGGL(colphys::k_pp81, g, b, 0, h->stream, h->s, h->pr, p0, p1);
  if (h->cfg.use_kpp) hipLaunchKernelGGL(colphys::k_kpp, g, b, 0, h->stream, h->s, h->pr, p0, p1);
  dim3 g2(g.x, (unsigned)(h->nz + 1));
  hipLaunchKernelGGL(colphys::k_split_assemble, g2, b, 0, h->stream, h->s, h->pr, p0, p1, h->d_guard);
  return launch_check("coefficients");
}

int read_guard(rdb_colphys* h, unsigned int* out) {
  HIP_TRY(hipMemcpyAsync(out, h->d_guard, sizeof(unsigned int), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemsetAsync(h->d_guard, 0, sizeof(unsigned int), h->stream));
  return RDB_COLPHYS_OK;
}

// Dry columns (no live layer) have outputs that depend only on h, u, v, tau
// and Q: refresh them with one launch whenever one of those changed.
int refresh_static(rdb_colphys* h, double dt) {
  for (rdb_colphys* m : h->members) {
    int st = refresh_static(m, dt);
    if (st) return st;
  }
  if (!h->static_dirty) return RDB_COLPHYS_OK;
  hipLaunchKernelGGL(colphys::k_prepare, grid1(h->n_all, 256), dim3(256), 0, h->stream, h->s, h->pr, 0, h->n_all);
  int st = launch_check("k_prepare");
  if (st) return st;
  st = launch_step(h, dt, h->n_wet, h->n_all);
  if (st) return st;
  h->static_dirty = false;
  return RDB_COLPHYS_OK;
}

}  // namespace colphys

namespace {
using colphys::launch_step;
using colphys::read_guard;
using colphys::refresh_static;
using colphys::staged_coefficients;

void drop_graph(rdb_colphys* h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
  h->graph_n = 0;
  colphys::surr_invalidate(h);
}

int ensure_events(rdb_colphys* h, int n) {
  while ((int)h->events.size() < n) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    h->events.push_back(e);
  }
  return RDB_COLPHYS_OK;
}

// Capture n step launches (with event records between them when timing)
// into a graph.
int build_graph(rdb_colphys* h, double dt, int n) {
  drop_graph(h);
  int st0 = colphys::ensure_work(h);
  if (st0) return st0;
  HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int st = RDB_COLPHYS_OK;
  for (int i = 0; i < n && st == RDB_COLPHYS_OK; ++i) st = colphys::launch_step_all(h, dt, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess)
    return push_error(RDB_COLPHYS_ERR_SETUP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  h->graph = g;
  e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(h);
    return push_error(RDB_COLPHYS_ERR_SETUP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  h->graph_dt = dt;
  h->graph_n = n;
  h->graph_block = h->block;
  h->graph_lds = h->lds_levels;
  h->graph_prec = h->precision;
  return RDB_COLPHYS_OK;
}

}  // namespace

extern "C" {

int rdb_colphys_config_default(rdb_colphys_config* c) {
  if (!c) return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "config_default: NULL config");
  std::memset(c, 0, sizeof *c);
  c->pp81_nu0 = 1.0e-2;
  c->pp81_nu_bg = 1.0e-4;
  c->pp81_kappa_bg = 1.0e-5;
  c->pp81_alpha = 5.0;
  c->rho0 = 1035.0;
  c->shear2_floor = 1.0e-10;
  c->use_kpp = 1;
  c->ri_crit = 0.3;
  c->c_vt2 = 1.8;
  c->cs_nonlocal = 6.3;
  c->alpha_T = 1.7e-4;
  c->beta_S = 7.6e-4;
  c->p_ref = 0.0;
  c->cp = 3992.0;
  c->has_heat = 1;
  c->has_salt = 1;
  c->kv_bg = 1.0e-4;
  c->kt_bg = 1.0e-5;
  c->ks_bg = 1.0e-5;
  c->kv_max = 1.7976931348623157e308;
  c->kd_max = 1.7976931348623157e308;
  c->precision = RDB_COLPHYS_PREC_FP64;
  /* contributors off at the defaults: rdb_config.F90:1768-1772, 2036-2100,
   * 2644-2677 */
  c->hmix_fixed = 20.0;
  c->conv_kd = 1.0;
  c->conv_prandtl = 1.0;
  c->conv_n2_thresh = 0.0;
  c->ddiff_strat_param_max = 2.55;
  c->ddiff_kappa_s = 1.0e-4;
  c->ddiff_exp1 = 1.0;
  c->ddiff_exp2 = 3.0;
  c->ddiff_param1 = 0.909;
  c->ddiff_param2 = 4.6;
  c->ddiff_param3 = -0.54;
  c->ddiff_mol_diff = 1.5e-6;
  c->ddiff_use_k90 = 0;
  c->epbl_enable = 0;
  c->epbl_mstar_scheme = RDB_EPBL_MSTAR_OM4;
  c->epbl_vstar_scheme = RDB_EPBL_VSTAR_CUBE_ROOT;
  c->epbl_combine = RDB_EPBL_COMBINE_ADD;
  c->epbl_mld_iteration = 1;
  c->epbl_mld_max_its = 20;
  c->epbl_mld_bisection = 0;
  c->epbl_mld_use_prev_guess = 0;
  c->epbl_use_lt = 0;
  c->epbl_lt_scheme = RDB_EPBL_LT_RESCALE;
  c->epbl_tke_diags = 0;
  c->epbl_mstar = 1.2;
  c->epbl_mstar_cap = -1.0;
  c->epbl_mstar_coef1 = 0.3;
  c->epbl_c_ek = 0.085;
  c->epbl_mstar_conv_adj = 0.0;
  c->epbl_rh18_cn1 = 0.275;
  c->epbl_rh18_cn2 = 8.0;
  c->epbl_rh18_cn3 = -5.0;
  c->epbl_rh18_cs1 = 0.2;
  c->epbl_rh18_cs2 = 0.4;
  c->epbl_nstar = 0.2;
  c->epbl_tke_decay = 2.5;
  c->epbl_wstar_ustar_coef = 1.0;
  c->epbl_vstar_scale_fac = 1.0;
  c->epbl_vstar_surf_fac = 1.2;
  c->epbl_von_karman = 0.41;
  c->epbl_ekman_scale_coef = 1.0;
  c->epbl_min_mix_len = 0.0;
  c->epbl_mixlen_exponent = 2.0;
  c->epbl_translay_scale = 0.1;
  c->epbl_mld_tol = 1.0;
  c->epbl_omega = 7.2921e-5;
  c->epbl_omega_frac = 0.0;
  c->epbl_ustar_min = 1.0e-8;
  c->epbl_prandtl = 1.0;
  c->epbl_lt_enhance_coef = 0.447;
  c->epbl_lt_enhance_exp = -1.33;
  c->epbl_lt_max_enhance = 5.0;
  c->epbl_la_frac_hbl = 0.04;
  c->epbl_lt_lac1 = -0.87;
  c->epbl_lt_lac2 = 0.0;
  c->epbl_lt_lac3 = 0.0;
  c->epbl_lt_lac4 = 0.95;
  c->epbl_lt_lac5 = 0.95;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_working_precision(void) { return 8; }

}  // extern "C"

namespace {

// One grid's plan: validation, column packing, device buffers.  With a
// non-null `stream` the plan borrows it (an ensemble member).
int create_plan(const rdb_colphys_config* cfg, int nx, int ny, int nz, int n_passive, const double* wet,
                const int* k_top, const int* sort_key, hipStream_t stream, rdb_colphys** out) {
  *out = nullptr;
  if (!cfg || !wet) return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "create: NULL config or wet mask");
  if (nx < 1 || ny < 1 || nz < 1 || n_passive < 0 || n_passive > 16)
    return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "create: bad extents nx=%d ny=%d nz=%d n_passive=%d", nx, ny, nz,
                      n_passive);
  if ((long long)nx * ny > 0x7fffffffLL / 2)
    return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "create: %d x %d columns exceed the index range", nx, ny);
  if (int st0 = validate_extras(*cfg)) return st0;
  if (cfg->kd_smooth_iterations != 0)
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "kd_smooth_iterations > 0 needs a horizontal halo; not supported");
  if (cfg->use_harmonic != 0 && cfg->use_harmonic != 1)
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "use_harmonic must be 0 or 1");
  if (cfg->precision != RDB_COLPHYS_PREC_FP64 && cfg->precision != RDB_COLPHYS_PREC_SPLIT)
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "precision must be 0 (fp64) or 1 (split)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return push_error(RDB_COLPHYS_ERR_SETUP, "create: no HIP device available");

  rdb_colphys* h = new rdb_colphys();
  h->cfg = *cfg;
  // EPBL and the KPP overlay are mutually exclusive surface schemes: EPBL
  // turns KPP off (rdb_ocean_setup_vmix.F90:697-703)
  if (cfg->epbl_enable) h->cfg.use_kpp = 0;
  h->pr = make_params(h->cfg);
  h->extras = has_extras(h->cfg);
  h->ep = make_epbl_params(h->cfg);
  h->cpar.epbl = h->cfg.epbl_enable ? 1 : 0;
  h->cpar.conv = h->cfg.conv_enable ? 1 : 0;
  h->cpar.combine = h->cfg.epbl_combine;
  h->cpar.prandtl = h->cfg.epbl_prandtl;
  h->cpar.kv_ml_invz2 = h->cfg.kv_ml_invz2;
  h->cpar.hmix = h->cfg.hmix_fixed;
  h->cpar.conv_kd = h->cfg.conv_kd;
  h->cpar.conv_prandtl = h->cfg.conv_prandtl;
  h->cpar.n2_thresh = h->cfg.conv_n2_thresh;
  h->cpar.rho0 = h->cfg.rho0;
  h->nx = nx;
  h->ny = ny;
  h->nz = nz;
  h->n_passive = n_passive;
  h->precision = cfg->precision;
  {
    // the fused step may hold its Thomas state in all of a workgroup's LDS
    // (160 KiB on gfx950): allowed once per device for every instantiation
    static unsigned long long lds_allowed = 0;  // per device 0..63
    int dev = 0, mx = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || mx <= 0) {
      delete h;
      return push_error(RDB_COLPHYS_ERR_SETUP, "create: LDS size query failed");
    }
    if (dev >= 64 || !(lds_allowed >> dev & 1ull)) {
      const hipError_t e = colphys::allow_full_lds(mx);
      if (e != hipSuccess) {
        delete h;
        return push_error(RDB_COLPHYS_ERR_SETUP, "create: dynamic LDS of %d bytes refused: %s", mx,
                          hipGetErrorString(e));
      }
      if (dev < 64) lds_allowed |= 1ull << dev;
    }
    h->lds_max = mx;
  }
  // RDB_COLPHYS_BLOCK pins the launch shape (64, 128 or 256 threads) and
  // turns rdb_colphys_autotune into a no-op, so that a profiled run can
  // reproduce an autotuned configuration exactly.
  if (const char* env = std::getenv("RDB_COLPHYS_BLOCK")) {
    int b = std::atoi(env);
    if (b == 64 || b == 128 || b == 256) {
      h->block = b;
      h->block_pinned = true;
      const char* lv = std::getenv("RDB_COLPHYS_LDS_LEVELS");
      int m = lv ? std::atoi(lv) : 0;
      const size_t nq = (size_t)(3 + std::min(n_passive, 4) + (cfg->kt_bg != cfg->ks_bg ? 1 : 0));
      if (m > 0 && (size_t)b * (size_t)m * nq * 8 <= (size_t)h->lds_max) h->lds_levels = m;
    }
  }
  const int ncol = nx * ny;
  std::vector<int> wet_idx, dry_idx;
  h->dry_host.assign((size_t)ncol, 0);
  for (int c = 0; c < ncol; ++c) {
    if (wet[c] > 0.5)
      wet_idx.push_back(c);
    else {
      dry_idx.push_back(c);
      h->dry_host[(size_t)c] = 1;
    }
  }
  if (sort_key) {
    std::stable_sort(wet_idx.begin(), wet_idx.end(), [&](int a, int b) { return sort_key[a] > sort_key[b]; });
    h->order_mode = 1;
  }
  h->n_wet = (int)wet_idx.size();
  h->n_all = ncol;
  h->P = ((ncol + 63) / 64) * 64;
  h->col_host = wet_idx;
  h->col_host.insert(h->col_host.end(), dry_idx.begin(), dry_idx.end());

  if (stream) {
    h->stream = stream;
    h->own_stream = false;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete h;
      return push_error(RDB_COLPHYS_ERR_SETUP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
  }
  const size_t P = (size_t)h->P;
  const size_t L = (size_t)nz * P, I = (size_t)(nz + 1) * P;
  StateView& s = h->s;
  s.P = h->P;
  s.nz = nz;
  s.n_passive = n_passive;
  int st = RDB_COLPHYS_OK;
  auto A = [&](void* pp, size_t bytes) {
    if (st == RDB_COLPHYS_OK) st = dalloc(h, (void**)pp, bytes);
  };
  A(&h->d_col, sizeof(int) * P);
  A(&s.h, 8 * L);
  A(&s.ub, 8 * L);
  A(&s.vb, 8 * L);
  A(&s.T, 8 * L);
  A(&s.S, 8 * L);
  A(&s.tr, 8 * L * (size_t)std::max(n_passive, 1));
  A(&s.rho, 8 * L);
  A(&s.kv, 8 * I);
  A(&s.kt, 8 * I);
  A(&s.ks, 8 * I);
  A(&s.gamma_t, 8 * I);
  A(&s.gamma_s, 8 * I);
  A(&s.fa, 8 * L);
  A(&s.fb, 8 * L);
  A(&s.fc, 8 * L);
  A(&s.tau_tot, 8 * P);
  A(&s.Qh, 8 * P);
  A(&s.Qs, 8 * P);
  A(&s.wet, 8 * P);
  A(&s.bl, 8 * P);
  A(&s.b0, 8 * P);
  A(&s.kbl, 4 * P);
  A(&s.ktop, 4 * P);
  A(&s.kb, 4 * P);
  A(&s.fdirty, 4 * P);
  A(&h->d_txc, 8 * P);
  A(&h->d_tyc, 8 * P);
  A(&h->d_smag, 8 * P);
  A(&h->d_shelf, 8 * P);
  A(&h->d_coriolis, 8 * P);
  A(&h->d_guard, 64);
  A(&h->d_acc, 64);
  if (h->cfg.epbl_enable) {
    A(&h->ev.kd, 8 * I);
    A(&h->ev.mld, 8 * P);
    A(&h->ev.b0, 8 * P);
    A(&h->ev.la, 8 * P);
    if (h->cfg.epbl_tke_diags) A(&h->ev.tke, 8 * 6 * P);
  }
  h->stage_elems = (size_t)(nx + 1) * (size_t)(ny + 1) * (size_t)(nz + 1);
  A(&h->d_stage, 8 * h->stage_elems);
  if (st) {
    release(h);
    delete h;
    return st;
  }
  // column map, wet mask and k_top (0-based) in packed order
  std::vector<int> col_pad(P, 0), ktop(P, nz - 1);
  std::vector<double> wetp(P, 0.0);
  for (int p = 0; p < ncol; ++p) {
    int c = h->col_host[(size_t)p];
    col_pad[(size_t)p] = c;
    wetp[(size_t)p] = wet[c];
    ktop[(size_t)p] = k_top ? std::min(std::max(k_top[c], 1), nz) - 1 : nz - 1;
  }
  if (hipMemcpy(h->d_col, col_pad.data(), sizeof(int) * P, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s.wet, wetp.data(), 8 * P, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s.ktop, ktop.data(), 4 * P, hipMemcpyHostToDevice) != hipSuccess) {
    release(h);
    delete h;
    return push_error(RDB_COLPHYS_ERR_SETUP, "create: initial upload failed");
  }
  // a fresh state: every layer vanished (the land contract) until set
  std::vector<double> hv(L, colphys::H_VANISHED);
  if (hipMemcpy(s.h, hv.data(), 8 * L, hipMemcpyHostToDevice) != hipSuccess) {
    release(h);
    delete h;
    return push_error(RDB_COLPHYS_ERR_SETUP, "create: initial upload failed");
  }
  h->h_u.assign((size_t)(nx + 1) * ny * nz, 0.0);
  h->h_v.assign((size_t)nx * (ny + 1) * nz, 0.0);
  h->h_taux.assign((size_t)(nx + 1) * ny, 0.0);
  h->h_tauy.assign((size_t)nx * (ny + 1), 0.0);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    release(h);
    delete h;
    return push_error(RDB_COLPHYS_ERR_SETUP, "create: stream synchronize failed");
  }
  *out = h;
  return RDB_COLPHYS_OK;
}

}  // namespace

extern "C" {

int rdb_colphys_create(const rdb_colphys_config* cfg, int nx, int ny, int nz, int n_passive, const double* wet,
                       const int* k_top, const int* sort_key, rdb_colphys** out) {
  if (!out) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "create: NULL handle_out");
  *out = nullptr;
  if (g_live)
    return push_error(RDB_COLPHYS_ERR_ALREADY_EXISTS, "create: a column-physics handle is already live");
  int st = create_plan(cfg, nx, ny, nz, n_passive, wet, k_top, sort_key, nullptr, out);
  if (st) return st;
  g_live = true;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_ensemble_create(const rdb_colphys_config* cfg, int n_members, const int* nx, const int* ny,
                                const int* nz, int n_passive, const double* const* wet, const int* const* sort_key,
                                rdb_colphys** out) {
  if (!out) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "ensemble_create: NULL handle_out");
  *out = nullptr;
  if (g_live)
    return push_error(RDB_COLPHYS_ERR_ALREADY_EXISTS, "ensemble_create: a column-physics handle is already live");
  if (n_members < 1 || n_members > 256 || !nx || !ny || !nz || !wet)
    return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "ensemble_create: bad member count %d or NULL extents", n_members);
  if (n_passive > 4)
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ensemble_create: at most 4 passive tracers");
  if (cfg && has_extras(*cfg))
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                      "ensemble_create: EPBL, KV_ML_INVZ2, convection and double diffusion run on single-grid handles");
  // the parent: a 1 x 1 x 1 all-land grid that owns the stream, the graph,
  // the counters and the learned closure; its members hold the columns
  const double land = 0.0;
  rdb_colphys* h = nullptr;
  int st = create_plan(cfg, 1, 1, 1, n_passive, &land, nullptr, nullptr, nullptr, &h);
  if (st) return st;
  long long row = 0;
  for (int m = 0; m < n_members && !st; ++m) {
    rdb_colphys* mp = nullptr;
    st = create_plan(cfg, nx[m], ny[m], nz[m], n_passive, wet[m], nullptr, sort_key ? sort_key[m] : nullptr,
                     h->stream, &mp);
    if (st) {
      push_error(st, "ensemble_create: member %d (%d x %d x %d) failed", m, nx[m], ny[m], nz[m]);
      break;
    }
    mp->row_off = row;
    row += mp->n_wet;
    h->members.push_back(mp);
  }
  int nz_max = 1;
  for (rdb_colphys* mp : h->members) nz_max = std::max(nz_max, mp->nz);
  for (rdb_colphys* mp : h->members) mp->lev_off = nz_max - mp->nz;
  // the parent's own 1 x 1 x 1 grid keeps nz = 1 and no wet column, so any
  // single-grid code path that reaches it touches nothing
  h->nz_net = nz_max;
  h->ens_wet = row;
  if (!st && row > 0x7fffffffLL) st = push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "ensemble_create: too many wet columns");
  if (!st) {
    const size_t nm = h->members.size();
    std::vector<StateView> views(nm);
    std::vector<int> nw(nm);
    for (size_t m = 0; m < nm; ++m) {
      views[m] = h->members[m]->s;
      nw[m] = h->members[m]->n_wet;
      h->live_levels += h->members[m]->live_levels;
    }
    st = dalloc(h, (void**)&h->d_views, sizeof(StateView) * nm);
    if (!st) st = dalloc(h, (void**)&h->d_svs, sizeof(colphys::SurrView) * nm);
    if (!st) st = dalloc(h, (void**)&h->d_nwet, sizeof(int) * nm);
    if (!st && (hipMemcpy(h->d_views, views.data(), sizeof(StateView) * nm, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(h->d_nwet, nw.data(), sizeof(int) * nm, hipMemcpyHostToDevice) != hipSuccess))
      st = push_error(RDB_COLPHYS_ERR_SETUP, "ensemble_create: upload of the member table failed");
  }
  if (st) {
    release(h);
    delete h;
    return st;
  }
  g_live = true;
  *out = h;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_member_set_field(rdb_colphys* h, int member, const char* name, const double* data, int n1, int n2,
                                 int n3) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "member_set_field: bad handle");
  if (h->members.empty()) {
    if (member != 0) return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "member_set_field: member %d of a single grid", member);
    return rdb_colphys_set_field(h, name, data, n1, n2, n3);
  }
  if (member < 0 || member >= (int)h->members.size())
    return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "member_set_field: no member %d", member);
  int st = rdb_colphys_set_field(h->members[(size_t)member], name, data, n1, n2, n3);
  if (!st && name && !std::strcmp(name, "h_layer")) {
    h->live_levels = 0;
    for (rdb_colphys* m : h->members) h->live_levels += m->live_levels;
  }
  return st;
}

int rdb_colphys_member_get_field(rdb_colphys* h, int member, const char* name, double* out, int n1, int n2, int n3) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "member_get_field: bad handle");
  if (h->members.empty()) {
    if (member != 0) return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "member_get_field: member %d of a single grid", member);
    return rdb_colphys_get_field(h, name, out, n1, n2, n3);
  }
  if (member < 0 || member >= (int)h->members.size())
    return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "member_get_field: no member %d", member);
  return rdb_colphys_get_field(h->members[(size_t)member], name, out, n1, n2, n3);
}

int rdb_colphys_member_count(rdb_colphys* h, int* n) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "member_count: bad handle");
  if (!n) return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "member_count: NULL output");
  *n = h->members.empty() ? 1 : (int)h->members.size();
  return RDB_COLPHYS_OK;
}

int rdb_colphys_destroy(rdb_colphys** hp) {
  if (!hp || !*hp) return RDB_COLPHYS_OK;
  rdb_colphys* h = *hp;
  if (g_live) {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    release(h);
    delete h;
    g_live = false;
  }
  *hp = nullptr;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_set_field(rdb_colphys* h, const char* name, const double* data, int n1, int n2, int n3) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "set_field: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "set_field: an ensemble handle is reached per member (rdb_colphys_member_*) "
                      "and stepped with rdb_colphys_step");
  if (!name || !data) return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "set_field: NULL name or data");
  FieldSpec spec;
  int tracer;
  if (!lookup_field(name, h->n_passive, spec, tracer)) {
    int m1, m2, m3;
    bool settable;
    if (!colphys::mom_field_spec(h, name, &m1, &m2, &m3, &settable))
      return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "set_field: unknown field '%s'", name);
    if (!settable) return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "set_field: '%s' is read-only", name);
    if (n1 != m1 || n2 != m2 || n3 != m3)
      return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "set_field '%s': shape (%d,%d,%d), expected (%d,%d,%d)", name, n1,
                        n2, n3, m1, m2, m3);
    return colphys::mom_set_field(h, name, data, (size_t)m1 * m2 * m3);
  }
  if (!spec.settable) return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "set_field: '%s' is read-only", name);
  if (tracer < 0 && (spec.kind == F_LAYER || spec.kind == F_IFACE || spec.kind == F_COL) && !layer_ptr(h, name, tracer))
    return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "set_field: '%s' is not kept by this handle (budget slots off, EPBL off or its TKE diagnostics off)", name);
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  int e1 = nx, e2 = ny, e3 = 1;
  switch (spec.kind) {
    case F_LAYER: e3 = nz; break;
    case F_IFACE: e3 = nz + 1; break;
    case F_FACE_X: e1 = nx + 1; e3 = nz; break;
    case F_FACE_Y: e2 = ny + 1; e3 = nz; break;
    case F_TAU_X: e1 = nx + 1; break;
    case F_TAU_Y: e2 = ny + 1; break;
    default: break;
  }
  if (n1 != e1 || n2 != e2 || n3 != e3)
    return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "set_field '%s': shape (%d,%d,%d), expected (%d,%d,%d)", name, n1,
                      n2, n3, e1, e2, e3);
  const size_t n = (size_t)e1 * e2 * e3;
  if (!std::strcmp(name, "h_layer")) {
    // land contract (rdb_ocean_state_seed.F90:1842-1880): a dry column holds
    // no live layer; its outputs are refreshed only when inputs change.
    for (int k = 0; k < nz; ++k)
      for (int c = 0; c < nx * ny; ++c)
        if (h->dry_host[(size_t)c] && data[(size_t)k * nx * ny + c] > colphys::H_VANISHED)
          return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                            "set_field h_layer: dry column (i=%d, j=%d) has a live layer k=%d (h=%g > H_VANISHED)",
                            c % nx + 1, c / nx + 1, k + 1, data[(size_t)k * nx * ny + c]);
  }
  if (!std::strcmp(name, "h_layer")) {
    // live levels per wet column (from the lowest live layer up), for the
    // algorithmic-bytes figure of the step
    long long live = 0;
    h->col_live.assign((size_t)nx * ny, 0);
    for (int p = 0; p < h->n_wet; ++p) {
      int c = h->col_host[(size_t)p], kb = nz;
      for (int k = 0; k < nz; ++k)
        if (data[(size_t)k * nx * ny + c] > colphys::H_VANISHED) { kb = k; break; }
      live += nz - kb;
      h->col_live[(size_t)c] = nz - kb;
    }
    h->live_levels = live;
    ++h->h_gen;
  }
  HIP_TRY(hipMemcpyAsync(h->d_stage, data, 8 * n, hipMemcpyHostToDevice, h->stream));
  const int ncol = h->n_all, P = h->P, B = 256;
  dim3 b(B);
  switch (spec.kind) {
    case F_LAYER:
    case F_IFACE: {
      double* dst = layer_ptr(h, name, tracer);
      hipLaunchKernelGGL(colphys::k_pack_layer, dim3(grid1(ncol, B).x, e3), b, 0, h->stream, h->d_stage, dst, h->d_col,
                         ncol, P, e3, (long long)nx * ny);
      break;
    }
    case F_COL: {
      double* dst = layer_ptr(h, name, tracer);
      hipLaunchKernelGGL(colphys::k_pack_layer, dim3(grid1(ncol, B).x, 1), b, 0, h->stream, h->d_stage, dst, h->d_col,
                         ncol, P, 1, (long long)nx * ny);
      break;
    }
    case F_FACE_X:
      std::copy(data, data + n, h->h_u.begin());
      if (int e = colphys::mom_face_set(h, true)) return e;
      hipLaunchKernelGGL(colphys::k_pack_centre_x, dim3(grid1(ncol, B).x, nz), b, 0, h->stream, h->d_stage, h->s.ub,
                         h->d_col, ncol, P, nz, nx, ny);
      break;
    case F_FACE_Y:
      std::copy(data, data + n, h->h_v.begin());
      if (int e = colphys::mom_face_set(h, false)) return e;
      hipLaunchKernelGGL(colphys::k_pack_centre_y, dim3(grid1(ncol, B).x, nz), b, 0, h->stream, h->d_stage, h->s.vb,
                         h->d_col, ncol, P, nz, nx, ny);
      break;
    case F_TAU_X:
      std::copy(data, data + n, h->h_taux.begin());
      hipLaunchKernelGGL(colphys::k_pack_centre_x, dim3(grid1(ncol, B).x, 1), b, 0, h->stream, h->d_stage, h->d_txc,
                         h->d_col, ncol, P, 1, nx, ny);
      break;
    case F_TAU_Y:
      std::copy(data, data + n, h->h_tauy.begin());
      hipLaunchKernelGGL(colphys::k_pack_centre_y, dim3(grid1(ncol, B).x, 1), b, 0, h->stream, h->d_stage, h->d_tyc,
                         h->d_col, ncol, P, 1, nx, ny);
      break;
    default:
      break;
  }
  int st = launch_check("pack");
  if (st) return st;
  if (spec.kind == F_TAU_X || spec.kind == F_TAU_Y || !std::strcmp(name, "stress_shelf")) {
    hipLaunchKernelGGL(colphys::k_stress_mag, grid1(ncol, B), b, 0, h->stream, h->d_txc, h->d_tyc, h->d_smag,
                       h->d_shelf, h->s.tau_tot, ncol);
    st = launch_check("k_stress_mag");
    if (st) return st;
  }
  h->static_dirty = true;
  HIP_TRY(hipStreamSynchronize(h->stream));  // the caller may reuse its buffer
  ++h->state_ver;
  {
    // a host mirror of this field takes the new values directly (as the
    // reference's setters write the host array they push)
    auto it = h->mirrors.find(name);
    if (it != h->mirrors.end() && it->second.size() == n) {
      std::copy(data, data + n, it->second.begin());
      h->mirror_ver[name] = h->state_ver;
    }
  }
  return RDB_COLPHYS_OK;
}

int rdb_colphys_get_field(rdb_colphys* h, const char* name, double* out, int n1, int n2, int n3) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "get_field: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "get_field: an ensemble handle is reached per member (rdb_colphys_member_*) "
                      "and stepped with rdb_colphys_step");
  if (!name || !out) return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "get_field: NULL name or buffer");
  FieldSpec spec;
  int tracer;
  if (!lookup_field(name, h->n_passive, spec, tracer)) {
    int m1, m2, m3;
    bool settable;
    if (!colphys::mom_field_spec(h, name, &m1, &m2, &m3, &settable))
      return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "get_field: unknown field '%s'", name);
    if (n1 != m1 || n2 != m2 || n3 != m3)
      return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "get_field '%s': shape (%d,%d,%d), expected (%d,%d,%d)", name, n1,
                        n2, n3, m1, m2, m3);
    return colphys::mom_get_field(h, name, out, (size_t)m1 * m2 * m3);
  }
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  int e1 = nx, e2 = ny, e3 = 1;
  switch (spec.kind) {
    case F_LAYER: e3 = nz; break;
    case F_IFACE: e3 = nz + 1; break;
    case F_FACE_X: e1 = nx + 1; e3 = nz; break;
    case F_FACE_Y: e2 = ny + 1; e3 = nz; break;
    case F_TAU_X: e1 = nx + 1; break;
    case F_TAU_Y: e2 = ny + 1; break;
    default: break;
  }
  if (n1 != e1 || n2 != e2 || n3 != e3)
    return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "get_field '%s': shape (%d,%d,%d), expected (%d,%d,%d)", name, n1,
                      n2, n3, e1, e2, e3);
  const size_t n = (size_t)e1 * e2 * e3;
  switch (spec.kind) {
    case F_FACE_X:
      if (int e = colphys::mom_face_get(h, true)) return e;
      std::copy(h->h_u.begin(), h->h_u.end(), out);
      return RDB_COLPHYS_OK;
    case F_FACE_Y:
      if (int e = colphys::mom_face_get(h, false)) return e;
      std::copy(h->h_v.begin(), h->h_v.end(), out);
      return RDB_COLPHYS_OK;
    case F_TAU_X: std::copy(h->h_taux.begin(), h->h_taux.end(), out); return RDB_COLPHYS_OK;
    case F_TAU_Y: std::copy(h->h_tauy.begin(), h->h_tauy.end(), out); return RDB_COLPHYS_OK;
    default: break;
  }
  const int ncol = h->n_all, P = h->P, B = 256;
  if (spec.kind != F_COL_INT && !layer_ptr(h, name, tracer))
    return push_error(RDB_COLPHYS_ERR_NOT_FOUND, "get_field: '%s' is not kept by this handle (budget slots off, EPBL off or its TKE diagnostics off)", name);
  if (spec.kind == F_COL_INT) {
    hipLaunchKernelGGL(colphys::k_unpack_int, grid1(ncol, B), dim3(B), 0, h->stream, h->s.kbl, h->d_stage, h->d_col,
                       ncol);
  } else {
    double* src = layer_ptr(h, name, tracer);
    hipLaunchKernelGGL(colphys::k_unpack_layer, dim3(grid1(ncol, B).x, e3), dim3(B), 0, h->stream, src, h->d_stage,
                       h->d_col, ncol, P, e3, (long long)nx * ny);
  }
  int st = launch_check("unpack");
  if (st) return st;
  HIP_TRY(hipMemcpyAsync(out, h->d_stage, 8 * n, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return RDB_COLPHYS_OK;
}

int rdb_colphys_eos(rdb_colphys* h) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "eos: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "eos: an ensemble handle is reached per member (rdb_colphys_member_*) "
                      "and stepped with rdb_colphys_step");
  ++h->state_ver;
  dim3 g(grid1(h->n_all, h->block).x, (unsigned)h->nz);
  hipLaunchKernelGGL(colphys::k_eos, g, dim3(h->block), 0, h->stream, h->s, h->pr, 0, h->n_all);
  return launch_check("k_eos");
}

int rdb_colphys_coefficients(rdb_colphys* h, double dt) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "coefficients: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "coefficients: an ensemble handle is reached per member (rdb_colphys_member_*) "
                      "and stepped with rdb_colphys_step");
  ++h->state_ver;
  int st = staged_coefficients(h, 0, h->n_all, dt);
  if (st) return st;
  h->static_dirty = true;  // the staged pass rewrote every interface
  if (h->cfg.vmix_guard) {
    unsigned int bad = 0;
    st = read_guard(h, &bad);
    if (st) return st;
    if (bad) {
      push_error(RDB_COLPHYS_GUARD_TRIPPED, "vmix_assemble: negative or NaN diffusivity detected (%u values)", bad);
      return RDB_COLPHYS_GUARD_TRIPPED;
    }
  }
  return RDB_COLPHYS_OK;
}

int rdb_colphys_surface_flux(rdb_colphys* h, double dt) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "surface_flux: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "surface_flux: an ensemble handle is reached per member (rdb_colphys_member_*) "
                      "and stepped with rdb_colphys_step");
  ++h->state_ver;
  if (!h->cfg.has_heat && !h->cfg.has_salt) return RDB_COLPHYS_OK;
  hipLaunchKernelGGL(colphys::k_surface_flux, grid1(h->n_all, h->block), dim3(h->block), 0, h->stream, h->s, h->pr,
                     dt, 0, h->n_all);
  return launch_check("k_surface_flux");
}

int rdb_colphys_diffuse_tracers(rdb_colphys* h, double dt) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "diffuse_tracers: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "diffuse_tracers: an ensemble handle is reached per member (rdb_colphys_member_*) "
                      "and stepped with rdb_colphys_step");
  ++h->state_ver;
  hipLaunchKernelGGL(colphys::k_vdiff, grid1(h->n_all, h->block), dim3(h->block), 0, h->stream, h->s, h->pr, dt, 0,
                     h->n_all, 0);
  return launch_check("k_vdiff");
}

int rdb_colphys_nonlocal(rdb_colphys* h, double dt) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "nonlocal: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "nonlocal: an ensemble handle is reached per member (rdb_colphys_member_*) "
                      "and stepped with rdb_colphys_step");
  ++h->state_ver;
  if (!h->cfg.use_kpp) return RDB_COLPHYS_OK;
  dim3 g(grid1(h->n_all, h->block).x, (unsigned)h->nz);
  hipLaunchKernelGGL(colphys::k_nonlocal, g, dim3(h->block), 0, h->stream, h->s, dt, 0, h->n_all);
  return launch_check("k_nonlocal");
}

int rdb_colphys_step(rdb_colphys* h, double dt, int n_steps) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "step: bad handle");
  if (n_steps <= 0) return RDB_COLPHYS_OK;
  int st = refresh_static(h, dt);
  if (st) return st;
  if (h->closure == RDB_COLPHYS_CLOSURE_SURROGATE) {
    // features -> fused MLP forward -> column step with the predicted
    // coefficients, n steps captured in one hipGraph and replayed
    st = colphys::surr_run(h, dt, n_steps);
    if (st) return st;
  } else if (h->n_passive > 4 || h->extras) {
    // beyond the fused kernel's tracer slots, or a contributor it does not
    // hold: the staged sequence, per step (with a contributor on, over every
    // column: KV_ML_INVZ2 also reaches dry columns)
    // Dry columns hold no live layer (the land contract), so their tracer
    // solve is the identity and is skipped; their coefficients change only
    // through KV_ML_INVZ2.
    const bool invz2 = h->cpar.kv_ml_invz2 > 0.0 && h->cpar.hmix > 0.0;
    const int pc = (h->extras && invz2) ? h->n_all : h->n_wet;
    h->last_graph = 0;
    if (h->timing) {
      st = ensure_events(h, 2);
      if (st) return st;
      HIP_TRY(hipEventRecord(h->events[0], h->stream));
    }
    for (int i = 0; i < n_steps; ++i) {
      if ((st = rdb_colphys_eos(h)) || (st = staged_coefficients(h, 0, pc, dt)) ||
          (st = rdb_colphys_surface_flux(h, dt)))
        return st;
      if (h->n_wet > 0) {
        hipLaunchKernelGGL(colphys::k_vdiff, grid1(h->n_wet, h->block), dim3(h->block), 0, h->stream, h->s, h->pr,
                           dt, 0, h->n_wet, 1);
        if ((st = launch_check("k_vdiff"))) return st;
      }
      if ((st = rdb_colphys_nonlocal(h, dt))) return st;
    }
    if (h->timing) {
      HIP_TRY(hipEventRecord(h->events[1], h->stream));
      HIP_TRY(hipEventSynchronize(h->events[1]));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, h->events[0], h->events[1]));
      h->last_ms = ms / n_steps;
      h->sum_ms += ms;
      h->timed += n_steps;
    }
  } else {
    if (!h->graph_exec || h->graph_dt != dt || h->graph_n != n_steps || h->graph_block != h->block ||
        h->graph_lds != h->lds_levels || h->graph_prec != h->precision) {
      st = build_graph(h, dt, n_steps);
      if (st) return st;
    }
    // Timing: events on the handle's stream around the graph replay; the
    // bracketed region holds exactly the n_steps step kernels.
    if (h->timing) {
      st = ensure_events(h, 2);
      if (st) return st;
      HIP_TRY(hipEventRecord(h->events[0], h->stream));
    }
    HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
    h->last_graph = 1;
    if (h->timing) {
      HIP_TRY(hipEventRecord(h->events[1], h->stream));
      HIP_TRY(hipEventSynchronize(h->events[1]));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, h->events[0], h->events[1]));
      h->last_ms = ms / n_steps;
      h->sum_ms += ms;
      h->timed += n_steps;
    }
  }
  h->steps += n_steps;
  h->t_sim += dt * n_steps;
  ++h->state_ver;
  if (h->cfg.vmix_guard) {
    unsigned int bad = 0;
    st = read_guard(h, &bad);
    if (st) return st;
    if (bad) {
      push_error(RDB_COLPHYS_GUARD_TRIPPED, "vmix_assemble: negative or NaN diffusivity detected (%u values)", bad);
      return RDB_COLPHYS_GUARD_TRIPPED;
    }
  }
  return RDB_COLPHYS_OK;
}

int rdb_colphys_synchronize(rdb_colphys* h) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "synchronize: bad handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  return RDB_COLPHYS_OK;
}

}  // extern "C"

namespace {

// Re-pack the wet columns of a single-grid plan into `order` (a permutation
// of its wet grid columns): every packed array is permuted on the device, the
// column map follows.  Exact (a permutation), so the state is unchanged.
int repack(rdb_colphys* h, const std::vector<int>& order) {
  const int n = h->n_wet, P = h->P, nz = h->nz;
  if ((int)order.size() != n) return push_error(RDB_COLPHYS_ERR_SETUP, "repack: order has the wrong length");
  std::vector<int> where((size_t)h->n_all, -1), newpos((size_t)std::max(n, 1));
  for (int i = 0; i < n; ++i) where[(size_t)order[(size_t)i]] = i;
  for (int p = 0; p < n; ++p) {
    const int np = where[(size_t)h->col_host[(size_t)p]];
    if (np < 0) return push_error(RDB_COLPHYS_ERR_SETUP, "repack: order is not a permutation of the wet columns");
    newpos[(size_t)p] = np;
  }
  const int maxlev = std::max(nz + 1, nz * std::max(h->n_passive, 1));
  int* d_newpos = nullptr;
  void* tmp = nullptr;
  HIP_TRY(hipMalloc((void**)&d_newpos, sizeof(int) * newpos.size()));
  hipError_t e = hipMalloc(&tmp, (size_t)8 * (size_t)maxlev * (size_t)P);
  if (e != hipSuccess) {
    (void)hipFree(d_newpos);
    return push_error(RDB_COLPHYS_ERR_SETUP, "repack: hipMalloc failed: %s", hipGetErrorString(e));
  }
  e = hipMemcpyAsync(d_newpos, newpos.data(), sizeof(int) * newpos.size(), hipMemcpyHostToDevice, h->stream);
  StateView& sv = h->s;
  auto perm_d = [&](double* a, int nlev) {
    if (e != hipSuccess || !a || n == 0) return;
    hipLaunchKernelGGL(colphys::k_permute_d, dim3(grid1(n, 256).x, (unsigned)nlev), dim3(256), 0, h->stream, a,
                       (double*)tmp, d_newpos, n, P, nlev);
    e = hipMemcpy2DAsync(a, 8 * (size_t)P, tmp, 8 * (size_t)P, 8 * (size_t)n, (size_t)nlev, hipMemcpyDeviceToDevice,
                         h->stream);
  };
  auto perm_i = [&](int* a, int nlev) {
    if (e != hipSuccess || !a || n == 0) return;
    hipLaunchKernelGGL(colphys::k_permute_i, dim3(grid1(n, 256).x, (unsigned)nlev), dim3(256), 0, h->stream, a,
                       (int*)tmp, d_newpos, n, P, nlev);
    e = hipMemcpy2DAsync(a, 4 * (size_t)P, tmp, 4 * (size_t)P, 4 * (size_t)n, (size_t)nlev, hipMemcpyDeviceToDevice,
                         h->stream);
  };
  for (double* a : {sv.h, sv.ub, sv.vb, sv.T, sv.S, sv.rho}) perm_d(a, nz);
  perm_d(sv.tr, nz * std::max(h->n_passive, 1));
  for (double* a : {sv.kv, sv.kt, sv.ks, sv.gamma_t, sv.gamma_s}) perm_d(a, nz + 1);
  for (double* a : {sv.budT, sv.budS}) perm_d(a, nz);
  for (double* a : {sv.tau_tot, sv.Qh, sv.Qs, sv.wet, sv.bl, sv.b0, h->d_txc, h->d_tyc, h->d_smag, h->d_shelf,
                    h->d_coriolis})
    perm_d(a, 1);
  for (int* a : {sv.kbl, sv.ktop, sv.kb, sv.fdirty, h->d_col}) perm_i(a, 1);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  (void)hipFree(d_newpos);
  if (e != hipSuccess) return push_error(RDB_COLPHYS_ERR_SETUP, "repack failed: %s", hipGetErrorString(e));
  std::copy(order.begin(), order.end(), h->col_host.begin());
  return RDB_COLPHYS_OK;
}

}  // namespace

extern "C" {

int rdb_colphys_autotune(rdb_colphys* h, double dt) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "autotune: bad handle");
  // the staged path (a contributor the fused step does not hold) has no
  // launch shape to tune
  if (h->block_pinned || h->extras) return RDB_COLPHYS_OK;
  int st = refresh_static(h, dt);
  if (st) return st;
  // back up the evolving state (T, S, passives, bl_depth) of every grid
  const std::vector<rdb_colphys*> ps = colphys::plans(h);
  size_t nstate = 0;
  // (and the budget slots, which the candidate steps add to)
  for (rdb_colphys* m : ps)
    nstate += (size_t)m->nz * (size_t)m->P * (size_t)(2 + m->n_passive + (m->s.budT ? 2 : 0)) + (size_t)m->P;
  double* bak = nullptr;
  HIP_TRY(hipMalloc((void**)&bak, 8 * std::max<size_t>(nstate, 1)));
  auto save = [&](bool restore) {
    size_t off = 0;
    for (rdb_colphys* m : ps) {
      const size_t L = (size_t)m->nz * (size_t)m->P;
      double* parts[6] = {m->s.T, m->s.S, m->s.tr, m->s.bl, m->s.budT, m->s.budS};
      size_t sizes[6] = {L, L, L * (size_t)m->n_passive, (size_t)m->P, m->s.budT ? L : 0, m->s.budS ? L : 0};
      for (int q = 0; q < 6; ++q) {
        if (sizes[q])
          (void)hipMemcpyAsync(restore ? parts[q] : bak + off, restore ? bak + off : parts[q], 8 * sizes[q],
                               hipMemcpyDeviceToDevice, h->stream);
        off += sizes[q];
      }
    }
  };
  // candidates: launch block x live levels of Thomas state held in LDS
  // (up to the device's LDS per workgroup) x column packing (single grids: the
  // current order, grid order, and deepest-first by live levels, which
  // keeps a wavefront's columns of one depth together)
  const size_t nq = (size_t)(3 + std::min(h->n_passive, 4) + (h->cfg.kt_bg != h->cfg.ks_bg ? 1 : 0));
  std::vector<std::pair<int, int>> cands;
  for (int b : {64, 128, 256}) {
    int mmax = (int)((size_t)h->lds_max / ((size_t)b * nq * 8));
    for (int m : {0, 4, 8, 12, 16, 24, 32}) {
      if (m <= mmax) cands.push_back({b, m});
    }
  }
  std::vector<std::vector<int>> orders{std::vector<int>(h->col_host.begin(), h->col_host.begin() + h->n_wet)};
  std::vector<int> order_modes{h->order_mode};
  if (h->members.empty() && h->col_live.size() == (size_t)h->n_all) {
    std::vector<int> natural = orders[0];
    std::sort(natural.begin(), natural.end());
    std::vector<int> deep = natural;
    std::stable_sort(deep.begin(), deep.end(), [&](int a, int b) { return h->col_live[(size_t)a] > h->col_live[(size_t)b]; });
    for (auto& [o, mode] : {std::pair<std::vector<int>*, int>{&natural, 0}, {&deep, 1}})
      if (*o != orders[0]) {
        orders.push_back(*o);
        order_modes.push_back(mode);
      }
  }
  double best = 1e300;
  int best_b = h->block, best_m = h->lds_levels;
  size_t best_o = 0;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (size_t oi = 0; oi < orders.size() && !st; ++oi) {
    if (oi > 0) st = repack(h, orders[oi]);
    if (st) break;
    save(false);  // the state in this packing
    for (auto [b, mm] : cands) {
      h->block = b;
      h->lds_levels = mm;
      st = colphys::ensure_work(h);
      if (!st) st = colphys::launch_step_all(h, dt, false);  // warm-up
      if (st) break;
      (void)hipEventRecord(e0, h->stream);
      for (int r = 0; r < 3 && !st; ++r) st = colphys::launch_step_all(h, dt, false);
      (void)hipEventRecord(e1, h->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        best_b = b;
        best_m = mm;
        best_o = oi;
      }
      save(true);
    }
  }
  (void)hipStreamSynchronize(h->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(bak);
  h->block = best_b;
  h->lds_levels = best_m;
  if (!st && orders.size() > 1) {
    st = repack(h, orders[best_o]);  // the state itself is unchanged: a permutation
    if (!st) h->order_mode = order_modes[best_o];
  }
  if (st) return st;
  return colphys::ensure_work(h);
}

int rdb_colphys_set_precision(rdb_colphys* h, int mode) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "set_precision: bad handle");
  if (mode != RDB_COLPHYS_PREC_FP64 && mode != RDB_COLPHYS_PREC_SPLIT)
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "set_precision: unknown mode %d", mode);
  if (h->extras && mode != RDB_COLPHYS_PREC_FP64)
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                      "set_precision: EPBL, KV_ML_INVZ2, convection and double diffusion run in FP64 only");
  if (mode != h->precision) h->static_dirty = true;
  h->precision = mode;
  for (rdb_colphys* m : h->members) {
    if (mode != m->precision) m->static_dirty = true;
    m->precision = mode;
  }
  return RDB_COLPHYS_OK;
}

int rdb_colphys_set_closure(rdb_colphys* h, int mode) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "set_closure: bad handle");
  if (mode == RDB_COLPHYS_CLOSURE_PLAIN) {
    h->closure = mode;
    return RDB_COLPHYS_OK;
  }
  if (mode == RDB_COLPHYS_CLOSURE_SURROGATE) {
    if (h->extras)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                        "set_closure: the surrogate replaces PP81 + KPP; it is not defined with EPBL, KV_ML_INVZ2, "
                        "convection or double diffusion on");
    if (!colphys::surr_ready(h))
      return push_error(RDB_COLPHYS_ERR_NOT_INITIALISED,
                        "set_closure: no surrogate network is trained or loaded into this handle");
    if (h->precision != RDB_COLPHYS_PREC_FP64)
      return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "set_closure: the surrogate runs with the FP64 solve only");
    h->closure = mode;
    return RDB_COLPHYS_OK;
  }
  return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "set_closure: unknown mode %d", mode);
}

int rdb_colphys_heat_salt_totals(rdb_colphys* h, double* heat, double* salt) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "heat_salt_totals: bad handle");
  if (!heat || !salt) return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "heat_salt_totals: NULL output");
  HIP_TRY(hipMemsetAsync(h->d_acc, 0, 4 * sizeof(unsigned long long), h->stream));
  for (rdb_colphys* m : colphys::plans(h)) {
    // integer bins: the members' sums add exactly, in any order
    int nb = std::min(2048, std::max(1, (m->n_wet + 255) / 256));
    hipLaunchKernelGGL(colphys::k_totals, dim3(nb), dim3(256), 0, h->stream, m->s, 0, m->n_wet, h->d_acc);
    int st = launch_check("k_totals");
    if (st) return st;
  }
  unsigned long long acc[4];
  HIP_TRY(hipMemcpyAsync(acc, h->d_acc, sizeof acc, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  auto combine = [](unsigned long long hi, unsigned long long lo) {
    __int128 v = (__int128)(long long)hi * ((__int128)1 << 32) + (__int128)(long long)lo;
    return (double)((long double)v / 17592186044416.0L);
  };
  *heat = combine(acc[0], acc[1]);
  *salt = combine(acc[2], acc[3]);
  return RDB_COLPHYS_OK;
}

int rdb_colphys_set_timing(rdb_colphys* h, int enable) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "set_timing: bad handle");
  h->timing = enable != 0;
  if (!h->timing) {
    h->sum_ms = 0.0;
    h->timed = 0;
  }
  return RDB_COLPHYS_OK;
}

int rdb_colphys_stats(rdb_colphys* h, rdb_colphys_counters* out) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "stats: bad handle");
  if (!out) return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "stats: NULL output");
  std::memset(out, 0, sizeof *out);
  out->steps = h->steps;
  out->last_kernel_ms = h->last_ms;
  out->sum_kernel_ms = h->sum_ms;
  out->timed_launches = h->timed;
  // B_col = 8 (11 L + 3 (L + 1)) + 44 per wet column: reads h, u, v, T, S and
  // two passives, writes T, S, passives and kv/kt/ks, plus column scalars.
  // The step touches only live levels (bed fillers are static), so its
  // algorithmic figure uses each column's live-level count L_c.
  long long wet = 0, land = 0, live = 0;
  for (rdb_colphys* m : colphys::plans(h)) {
    const double L = (double)m->nz;
    out->reference_bytes_per_step += (8.0 * (11.0 * L + 3.0 * (L + 1.0)) + 44.0) * (double)m->n_wet;
    wet += m->n_wet;
    land += m->n_all - m->n_wet;
    live += m->live_levels;
  }
  out->wet_columns = wet;
  out->land_columns = land;
  out->algorithmic_bytes_per_step = 8.0 * (14.0 * (double)live + 3.0 * (double)wet) + 44.0 * (double)wet;
  out->live_levels = live;
  out->lds_levels = h->lds_levels;
  out->graph_replay = h->last_graph;
  if (h->closure != RDB_COLPHYS_CLOSURE_PLAIN) {
    const int st = colphys::surr_settle(h);
    if (st) return st;
  }
  out->fallback_fraction =
      h->closure == RDB_COLPHYS_CLOSURE_PLAIN ? 1.0 : (double)h->last_fallback / (double)std::max(wet, 1LL);
  out->block_threads = h->block;
  out->order_mode = h->order_mode;
  out->mlp_forward_ms = h->last_mlp_ms;
  out->mlp_flop_per_step = h->mlp_flop;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_mlp_config_default(rdb_colphys_mlp_config* c) {
  if (!c) return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp_config_default: NULL config");
  std::memset(c, 0, sizeof *c);
  c->hidden = 256;
  c->n_hidden = 4;
  c->batch = 8192;
  c->epochs = 8;
  c->lr = 1.0e-3;
  c->beta1 = 0.9;
  c->beta2 = 0.999;
  c->eps = 1.0e-8;
  c->seed = 1;
  c->guard_mode = RDB_COLPHYS_GUARD_STRICT;
  c->val_fraction = 0.05;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_mlp_collect(rdb_colphys* h, double dt, int n_steps, int stride, long long* n_samples) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "mlp_collect: bad handle");
  if (n_steps < 0) return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp_collect: negative step count");
  if (h->extras)
    return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                      "mlp_collect: the surrogate learns PP81 + KPP; it is not defined with EPBL, KV_ML_INVZ2, "
                      "convection or double diffusion on");
  return colphys::surr_collect(h, dt, n_steps, stride, n_samples);
}

int rdb_colphys_mlp_train(rdb_colphys* h, const rdb_colphys_mlp_config* cfg, float* train_loss, float* val_loss,
                          int n_hist, int* n_epochs_done) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "mlp_train: bad handle");
  return colphys::surr_train(h, cfg, train_loss, val_loss, n_hist, n_epochs_done);
}

int rdb_colphys_mlp_save(rdb_colphys* h, const char* path) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "mlp_save: bad handle");
  if (!path) return push_error(RDB_COLPHYS_ERR_IO, "mlp_save: NULL path");
  return colphys::surr_save(h, path);
}

int rdb_colphys_mlp_load(rdb_colphys* h, const char* path) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "mlp_load: bad handle");
  if (!path) return push_error(RDB_COLPHYS_ERR_IO, "mlp_load: NULL path");
  return colphys::surr_load(h, path);
}

int rdb_colphys_set_guard(rdb_colphys* h, int mode) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "set_guard: bad handle");
  return colphys::surr_set_guard(h, mode);
}

int rdb_colphys_set_budgets(rdb_colphys* h, int enable) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "set_budgets: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "set_budgets: the budget slots are kept for single grids only");
  if (enable && !h->s.budT) {
    const size_t L = (size_t)h->nz * (size_t)h->P;
    int st = dalloc(h, (void**)&h->s.budT, 8 * L);  // zeroed
    if (!st) st = dalloc(h, (void**)&h->s.budS, 8 * L);
    if (!st) st = dalloc(h, (void**)&h->s.pold, 16 * L);
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(h->stream));
  } else if (!enable) {
    h->s.budT = h->s.budS = h->s.pold = nullptr;  // the buffers stay allocated until destroy
  }
  drop_graph(h);  // the captured launches carry the state view
  return RDB_COLPHYS_OK;
}

int rdb_colphys_vdiff_momentum(rdb_colphys* h, double dt, const rdb_colphys_momentum_config* cfg) {
  if (!valid(h)) return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "vdiff_momentum: bad handle");
  if (!h->members.empty())
    return push_error(RDB_COLPHYS_ERR_BAD_HANDLE, "vdiff_momentum: an ensemble handle is reached per member "
                      "(rdb_colphys_member_*) and stepped with rdb_colphys_step");
  if (!cfg) return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "vdiff_momentum: NULL config");
  if (!(dt > 0.0) || !std::isfinite(dt)) return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "vdiff_momentum: dt = %g", dt);
  return colphys::mom_solve(h, dt, cfg);
}

int rdb_colphys_safe_math(int fn, const double* x, const double* y, double* out, int n) {
  if (fn < RDB_SM_EXP || fn > RDB_SM_ERFC) return push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "safe_math: unknown fn %d", fn);
  if (n < 0 || n > (1 << 26) || (n > 0 && (!x || !out || (fn == RDB_SM_POW && !y))))
    return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "safe_math: n = %d outside [0, 2^26] or NULL arrays", n);
  if (n == 0) return RDB_COLPHYS_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return push_error(RDB_COLPHYS_ERR_SETUP, "safe_math: no HIP device available");
  double* d = nullptr;
  HIP_TRY(hipMalloc((void**)&d, 24 * (size_t)n));
  hipError_t e = hipMemcpy(d, x, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess && fn == RDB_SM_POW) e = hipMemcpy(d + n, y, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(colphys::k_safe_math_eval, grid1(n, 256), dim3(256), 0, 0, fn, d, d + n, d + 2 * (size_t)n, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d + 2 * (size_t)n, 8 * (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return push_error(RDB_COLPHYS_ERR_SETUP, "safe_math: %s", hipGetErrorString(e));
  return RDB_COLPHYS_OK;
}

int rdb_colphys_cbrt(const double* x, double* y, int n) {
  if (n < 0 || n > (1 << 26) || (n > 0 && (!x || !y)))
    return push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "cbrt: n = %d outside [0, 2^26] or NULL arrays", n);
  if (n == 0) return RDB_COLPHYS_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return push_error(RDB_COLPHYS_ERR_SETUP, "cbrt: no HIP device available");
  double* d = nullptr;
  HIP_TRY(hipMalloc((void**)&d, 16 * (size_t)n));
  hipError_t e = hipMemcpy(d, x, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(colphys::k_cbrt_eval, grid1(n, 256), dim3(256), 0, 0, d, d + n, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(y, d + n, 8 * (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return push_error(RDB_COLPHYS_ERR_SETUP, "cbrt: %s", hipGetErrorString(e));
  return RDB_COLPHYS_OK;
}

int rdb_colphys_last_error(int idx, char* buf, int cap) {
  if (idx < 0 || idx >= (int)g_ring.size() || cap <= 0 || !buf) return 0;
  const std::string& m = g_ring[(size_t)idx];
  std::snprintf(buf, (size_t)cap, "%s", m.c_str());
  return (int)m.size();
}

}  // extern "C"

namespace colphys {
bool handle_valid(rdb_colphys* h) { return valid(h); }
}  // namespace colphys

namespace colphys {
int step_events(rdb_colphys* h, hipEvent_t* e0, hipEvent_t* e1) {
  int st = ensure_events(h, 2);
  if (st) return st;
  *e0 = h->events[0];
  *e1 = h->events[1];
  return RDB_COLPHYS_OK;
}
}  // namespace colphys
