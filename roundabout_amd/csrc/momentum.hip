// momentum.hip -- backward-Euler vertical viscosity of the face velocities:
// vdiff_apply_momentum -> diffuse_velocity_columns_impl
// (src/parameterizations/vertical/rdb_ocean_vdiff.F90:455-718, 1416-2115).
//
// One lane per face column.  The face velocities live on the device in the
// reference's face layout, level-major (u: (nx+1) x ny faces per level, v:
// nx x (ny+1)), so a wavefront's 64 faces of a level are one coalesced
// 512-byte access; the two abutting cells' layer thickness and viscosity are
// read from natural-order (i fastest) copies of h and kv, where the
// neighbouring faces of a wavefront share their cells in L1/L2.  Every
// expression keeps the reference's operation order (the library is built
// without FMA contraction), so the results are bitwise those of the
// reference's contraction-off CPU build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rdb_colphys.h"
#include "colphys_handle.h"
#include "colphys_kernels.h"

namespace colphys {

namespace {

struct MomFlags {
  int do_stress, do_drag, do_top, solve, remnant;
  int mom6, harm, upwind, glue, zlevel, corner, use_harmonic, per_face, scalar_kv;
  double rho0, hbbl_visc, kv_bbl_bg, prandtl, kv_bbl_c, bbl_thick_c, K;
};

struct MomView {
  int nu, nv, nz, nx, ny, x_face;
  double* u;           // [nz][nu * nv]
  const double* h;     // [nz][nx * ny]
  const double* kv;    // [nz + 1][nx * ny] (unused with scalar_kv)
  const double* wet;   // [nx * ny]
  const double* tau;   // [nu * nv]
  const double* lb;    // bottom drag rate
  const double* lt;    // top drag rate
  const double* cov;   // ice cover
  const double* kvb;   // per-face BBL viscosity
  const double* bth;   // per-face BBL thickness
  const double* kvc;   // [nz + 1][(nx + 1) * (ny + 1)] corner viscosity
  const double* ktop;  // 1-based layer index, or null: nz
  const double* kbot;  // 1-based layer index, or null: 1
  double* cd;          // Thomas c' [nz][nu * nv]
  double* rd;          // Thomas d'
  double* rem;         // visc_rem
};

__device__ __forceinline__ double mmax(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double mmin(double a, double b) { return (a < b) ? a : b; }

// face_thick (rdb_ocean_vdiff.F90:1928-1962)
__device__ __forceinline__ double m_face_thick(double ha, double hb, int use_harmonic) {
  if (use_harmonic) {
    double sum_h = ha + hb;
    if (sum_h > 1.0e-30) return 2.0 * ha * hb / sum_h;
    return 0.5 * sum_h;
  }
  return 0.5 * (ha + hb);
}

// vdiff_face_layer (rdb_ocean_vdiff.F90:1964-2065)
__device__ __forceinline__ void m_face_layer(const MomFlags& fl, double hl_c, double hr_c, double u_k, int k, int kb,
                                             double d_min, double i_hbbl, double& zacc, double& zcol_l,
                                             double& zcol_r, double& hvel, double& zint) {
  const double EPS_HVEL = 1.0e-30;
  zint = 0.0;
  if (fl.mom6 && !fl.harm) {
    double h_harm = 2.0 * hl_c * hr_c / (hl_c + hr_c + EPS_HVEL);
    double h_arith = 0.5 * (hl_c + hr_c);
    double h_delta = hr_c - hl_c;
    zcol_l = zcol_l + hl_c;
    zcol_r = zcol_r + hr_c;
    if (k >= kb) zacc = zacc + h_harm;
    double z_clear = mmax(zcol_l, zcol_r) + d_min;
    zint = mmax(zacc, z_clear) * i_hbbl;
    hvel = h_arith;
    if (u_k * h_delta > 0.0) {
      double z2 = zint;
      double botfn = 1.0 / (1.0 + 0.09 * z2 * z2 * z2 * z2 * z2 * z2);
      hvel = (1.0 - botfn) * h_arith + botfn * h_harm;
    }
  } else if (fl.mom6) {
    double h_harm = 2.0 * hl_c * hr_c / (hl_c + hr_c + EPS_HVEL);
    double h_arith = 0.5 * (hl_c + hr_c);
    double h_delta = hr_c - hl_c;
    hvel = h_harm;
    if (fl.upwind && u_k * h_delta < 0.0) {
      double z2 = zacc;
      double botfn = 1.0 / (1.0 + 0.09 * z2 * z2 * z2 * z2 * z2 * z2);
      hvel = (1.0 - botfn) * h_harm + botfn * h_arith;
    }
    if (k >= kb) zacc = zacc + h_harm * i_hbbl;
    zint = zacc;
  } else {
    hvel = 0.5 * (hl_c + hr_c);
  }
  if (fl.zlevel) hvel = mmin(hl_c, hr_c);
}

// vdiff_interface_coupling (rdb_ocean_vdiff.F90:2067-2115)
__device__ __forceinline__ void m_coupling(const MomFlags& fl, double hv_lo, double hv_hi, double zint_lo,
                                           double dt, double kv_bbl, double bbl_thick, double kv_cnr, double kv_l,
                                           double kv_r, double& num, double& dz, bool& cut) {
  const double HV = H_VANISHED;
  double hf_lo = mmax(hv_lo, HV);
  double hf_hi = mmax(hv_hi, HV);
  if (fl.mom6)
    dz = 0.5 * (hf_lo + hf_hi);
  else
    dz = m_face_thick(hf_lo, hf_hi, fl.use_harmonic);
  double nu_f = 0.5 * (kv_l + kv_r);
  if (fl.corner) nu_f = nu_f + kv_cnr;
  if (fl.glue) {
    double botfn_int = 1.0 / (1.0 + 0.09 * zint_lo * zint_lo * zint_lo * zint_lo * zint_lo * zint_lo);
    nu_f = mmax(0.0, nu_f + (kv_bbl - fl.kv_bbl_bg) * botfn_int);
    if (dz > bbl_thick) dz = (1.0 - botfn_int) * dz + botfn_int * bbl_thick;
  }
  num = dt * nu_f;
  cut = fl.zlevel && (hv_lo <= HV || hv_hi <= HV);
}

// diffuse_velocity_columns_impl (rdb_ocean_vdiff.F90:1416-1926): one face
// column per lane, the fused upward sweep (face layer, interface coupling,
// row, Thomas forward step) and one back substitution.
__global__ void __launch_bounds__(256) k_vdiff_momentum(MomView v, MomFlags fl, double dt) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t F = (size_t)v.nu * (size_t)v.nv;
  if ((size_t)f >= F) return;
  const double HV = H_VANISHED, EPS_HVEL = 1.0e-30, STRESS_H_MIN = 1.0e-3;
  const double inv_rho0 = 1.0 / fl.rho0;
  const int nz = v.nz, nx = v.nx, ny = v.ny;
  const int i = f % v.nu, j = f / v.nu;
  const size_t C = (size_t)nx * (size_t)ny, Ccr = (size_t)(nx + 1) * (size_t)(ny + 1);
  int i_left, i_right, j_below, j_above, i_c2, j_c2;
  if (v.x_face) {
    i_left = max(0, i - 1);
    i_right = min(nx - 1, i);
    j_below = j;
    j_above = j;
    i_c2 = i;
    j_c2 = j + 1;
  } else {
    i_left = i;
    i_right = i;
    j_below = max(0, j - 1);
    j_above = min(ny - 1, j);
    i_c2 = i + 1;
    j_c2 = j;
  }
  const size_t cl = (size_t)i_left + (size_t)nx * (size_t)j_below;
  const size_t cr = (size_t)i_right + (size_t)nx * (size_t)j_above;
  const size_t cn1 = (size_t)i + (size_t)(nx + 1) * (size_t)j;
  const size_t cn2 = (size_t)i_c2 + (size_t)(nx + 1) * (size_t)j_c2;
  const int kb = v.kbot ? (int)v.kbot[f] - 1 : 0;
  const int ktop = v.ktop ? (int)v.ktop[f] - 1 : nz - 1;
  double kv_bbl = 0.0, bbl_thick = fl.hbbl_visc;
  if (fl.glue) {
    kv_bbl = fl.per_face ? v.kvb[f] : fl.kv_bbl_c;
    bbl_thick = fl.per_face ? v.bth[f] : fl.bbl_thick_c;
  }
  const double i_hbbl = 1.0 / (bbl_thick + EPS_HVEL);
  double zacc = 0.0, zcol_l = 0.0, zcol_r = 0.0;
  if (fl.mom6 && !fl.harm) {
    for (int k = 0; k < nz; ++k) {
      zcol_l = zcol_l - v.h[cl + C * k];
      zcol_r = zcol_r - v.h[cr + C * k];
    }
  }
  const double d_min = mmin(-zcol_l, -zcol_r);
  int kb_live = kb;
  for (int k = kb; k < nz; ++k) {
    const bool live = fl.zlevel ? (mmin(v.h[cl + C * k], v.h[cr + C * k]) > HV) : true;
    if (live) {
      kb_live = k;
      break;
    }
  }
  // first row of the solve that is not the identity (a top row below the
  // first live one takes the top forcing: then every row is solved)
  const int k_lo = (ktop < kb_live) ? 0 : kb_live;
  auto kvat = [&](size_t c, int kf) -> double { return fl.scalar_kv ? fl.K : v.kv[c + C * kf]; };
  double hv_0, hv_p, zi_0, zi_p;
  m_face_layer(fl, v.h[cl], v.h[cr], v.u[f], 0, kb, d_min, i_hbbl, zacc, zcol_l, zcol_r, hv_p, zi_p);
  double num_t = 0.0, dz_t = 1.0, num_b, dz_b;
  bool cut_t = false, cut_b;
  double cp_prev = 0.0, dp_prev = 0.0, gp_prev = 0.0, cp = 0.0, dp = 0.0, gp = 0.0;
  for (int k = 0; k < nz; ++k) {
    hv_0 = hv_p;
    zi_0 = zi_p;
    num_b = num_t;
    dz_b = dz_t;
    cut_b = cut_t;
    if (k < nz - 1) {
      m_face_layer(fl, v.h[cl + C * (k + 1)], v.h[cr + C * (k + 1)], v.u[f + F * (k + 1)], k + 1, kb, d_min, i_hbbl,
                   zacc, zcol_l, zcol_r, hv_p, zi_p);
      if (k >= kb_live && k <= ktop) {
        double kv_cnr = 0.0;
        if (fl.corner) kv_cnr = fl.prandtl * (0.5 * (v.kvc[cn1 + Ccr * (k + 1)] + v.kvc[cn2 + Ccr * (k + 1)]));
        m_coupling(fl, hv_0, hv_p, zi_0, dt, kv_bbl, bbl_thick, kv_cnr, kvat(cl, k + 1), kvat(cr, k + 1), num_t,
                   dz_t, cut_t);
      }
    }
    double a_k = 0.0, b_k = 1.0, c_k = 0.0, alpha, beta;
    const double hf_k = mmax(hv_0, HV);
    if (k == kb_live && k <= ktop) {
      alpha = 0.0;
      if (kb_live < nz - 1) {
        alpha = num_t / (hf_k * dz_t);
        if (cut_t) alpha = 0.0;
      }
      c_k = -alpha;
      b_k = 1.0 + alpha;
      if (fl.glue) b_k = b_k + dt * kv_bbl / (hf_k * (mmin(0.5 * hv_0, bbl_thick) + EPS_HVEL));
      if (fl.do_drag && !fl.glue) b_k = b_k + dt * v.lb[f];
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
    double r_k = v.u[f + F * k];
    if (k == ktop) {
      if (fl.do_top) b_k = b_k + dt * v.lt[f];
      if (fl.do_stress) {
        double wind_open = 1.0;
        if (fl.do_top) wind_open = 1.0 - v.cov[f];
        r_k = r_k + dt * v.tau[f] * wind_open * mmin(v.wet[cl], v.wet[cr]) * inv_rho0 / mmax(hf_k, STRESS_H_MIN);
      }
    }
    if (k == 0) {
      cp = c_k / b_k;
      dp = r_k / b_k;
      gp = 1.0 / b_k;
    } else {
      double denom = b_k - a_k * cp_prev;
      cp = c_k / denom;
      dp = (r_k - a_k * dp_prev) / denom;
      gp = (1.0 - a_k * gp_prev) / denom;
    }
    // rows below k_lo are the identity (cp = 0, dp = u, gp = 1): nothing of
    // them is stored, and the back substitution stops at k_lo (u - 0 * u_next
    // leaves u as it is, the remnant is 1)
    if (k >= k_lo) {
      v.cd[f + F * k] = cp;
      if (fl.solve) v.rd[f + F * k] = dp;
      if (fl.remnant) v.rem[f + F * k] = gp;
    }
    cp_prev = cp;
    dp_prev = dp;
    gp_prev = gp;
  }
  double u_next = dp_prev;
  double g_next = mmin(gp_prev, 1.0);
  if (fl.solve) v.u[f + F * (nz - 1)] = u_next;
  if (fl.remnant) v.rem[f + F * (nz - 1)] = g_next;
  for (int k = nz - 2; k >= k_lo; --k) {
    double cpk = v.cd[f + F * k];
    if (fl.solve) {
      u_next = v.rd[f + F * k] - cpk * u_next;
      v.u[f + F * k] = u_next;
    }
    if (fl.remnant) {
      g_next = mmin(v.rem[f + F * k] - cpk * g_next, 1.0);
      v.rem[f + F * k] = g_next;
    }
  }
  if (fl.remnant)
    for (int k = min(k_lo, nz - 1) - 1; k >= 0; --k) v.rem[f + F * k] = 1.0;
}

inline dim3 mgrid(size_t n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

// face-located inputs of the folds, by name: extents (x face or y face, or
// the corner array) and whether they are settable
struct MomFieldSpec {
  const char* name;
  int kind;  // 0: u face 2-D, 1: v face 2-D, 2: corner 3-D, 3: u face 3-D (get only), 4: v face 3-D (get only)
};
const MomFieldSpec kMomFields[] = {
    {"lambda_bot_u", 0}, {"lambda_bot_v", 1}, {"lambda_top_u", 0}, {"lambda_top_v", 1}, {"cover_u", 0},
    {"cover_v", 1},      {"kv_bbl_u", 0},     {"kv_bbl_v", 1},     {"bbl_thick_u", 0}, {"bbl_thick_v", 1},
    {"k_top_u", 0},      {"k_bot_u", 0},      {"k_top_v", 1},      {"k_bot_v", 1},     {"kv_corner", 2},
    {"visc_rem_u", 3},   {"visc_rem_v", 4},
};

}  // namespace

struct MomState {
  size_t Fu = 0, Fv = 0, C = 0, Ccr = 0;
  double* uf = nullptr;  // [nz][Fu], reference face layout
  double* vf = nullptr;
  bool faces_on_device = false;  // uf/vf hold the current face velocities
  bool u_host_stale = false, v_host_stale = false;
  double* h_nat = nullptr;   // [nz][C]
  long long h_gen = -1;      // the handle's h_gen h_nat was copied at (-1: never)
  double* kv_nat = nullptr;  // [nz + 1][C]
  double* wet_nat = nullptr;
  double* tau_u = nullptr;
  double* tau_v = nullptr;
  double* cd = nullptr;  // [nz][max(Fu, Fv)]
  double* rd = nullptr;
  double* rem_u = nullptr;
  double* rem_v = nullptr;
  bool rem_valid = false;
  std::vector<std::pair<std::string, double*>> fields;  // set face fields (device copies)
  double* field(const char* n) const {
    for (auto& e : fields)
      if (e.first == n) return e.second;
    return nullptr;
  }
};

namespace {

int mom_alloc(rdb_colphys* h, double** p, size_t bytes) {
  hipError_t e = hipMalloc((void**)p, bytes ? bytes : 8);
  if (e != hipSuccess) {
    char buf[160];
    snprintf(buf, sizeof buf, "momentum: hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, buf);
  }
  h->allocations.push_back(*p);
  return 0;
}

int mom_fail(const char* what, hipError_t e) {
  char buf[200];
  snprintf(buf, sizeof buf, "momentum: %s: %s", what, hipGetErrorString(e));
  return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, buf);
}

#define MOM_TRY(what, expr)                  \
  do {                                       \
    hipError_t _e = (expr);                  \
    if (_e != hipSuccess) return mom_fail(what, _e); \
  } while (0)

// the momentum state of a single-grid handle, created on first use with
// the face velocities copied from the host mirrors
int mom_ensure(rdb_colphys* h, MomState** out) {
  if (!h->mom) {
    MomState* m = new MomState();
    h->mom = m;
    m->Fu = (size_t)(h->nx + 1) * h->ny;
    m->Fv = (size_t)h->nx * (h->ny + 1);
    m->C = (size_t)h->nx * h->ny;
    m->Ccr = (size_t)(h->nx + 1) * (h->ny + 1);
    const size_t nz = (size_t)h->nz;
    int st = mom_alloc(h, &m->uf, 8 * nz * m->Fu);
    if (!st) st = mom_alloc(h, &m->vf, 8 * nz * m->Fv);
    if (st) return st;
  }
  MomState* m = h->mom;
  if (!m->faces_on_device) {
    MOM_TRY("face upload", hipMemcpyAsync(m->uf, h->h_u.data(), 8 * h->h_u.size(), hipMemcpyHostToDevice, h->stream));
    MOM_TRY("face upload", hipMemcpyAsync(m->vf, h->h_v.data(), 8 * h->h_v.size(), hipMemcpyHostToDevice, h->stream));
    MOM_TRY("face upload", hipStreamSynchronize(h->stream));
    m->faces_on_device = true;
  }
  *out = m;
  return 0;
}

}  // namespace

void mom_release(rdb_colphys* h) {
  delete h->mom;  // device buffers are in h->allocations
  h->mom = nullptr;
}

int mom_face_set(rdb_colphys* h, bool x_face) {
  // a new host face field: the device copy follows it
  MomState* m = h->mom;
  if (!m || !m->faces_on_device) return 0;
  std::vector<double>& src = x_face ? h->h_u : h->h_v;
  MOM_TRY("face upload", hipMemcpyAsync(x_face ? m->uf : m->vf, src.data(), 8 * src.size(), hipMemcpyHostToDevice,
                                        h->stream));
  if (x_face)
    m->u_host_stale = false;
  else
    m->v_host_stale = false;
  return 0;
}

int mom_face_get(rdb_colphys* h, bool x_face) {
  // the host mirror of a face field after a momentum solve changed it
  MomState* m = h->mom;
  if (!m) return 0;
  bool& stale = x_face ? m->u_host_stale : m->v_host_stale;
  if (!stale) return 0;
  std::vector<double>& dst = x_face ? h->h_u : h->h_v;
  MOM_TRY("face download", hipMemcpyAsync(dst.data(), x_face ? m->uf : m->vf, 8 * dst.size(), hipMemcpyDeviceToHost,
                                          h->stream));
  MOM_TRY("face download", hipStreamSynchronize(h->stream));
  stale = false;
  return 0;
}

// 1: name is a momentum field (extents in e1..e3, settable); 0: not one
int mom_field_spec(rdb_colphys* h, const char* name, int* e1, int* e2, int* e3, bool* settable) {
  for (const MomFieldSpec& s : kMomFields)
    if (!std::strcmp(s.name, name)) {
      *e1 = h->nx + (s.kind == 0 || s.kind == 2 || s.kind == 3 ? 1 : 0);
      *e2 = h->ny + (s.kind == 1 || s.kind == 2 || s.kind == 4 ? 1 : 0);
      *e3 = s.kind == 2 ? h->nz + 1 : (s.kind >= 3 ? h->nz : 1);
      *settable = s.kind <= 2;
      return 1;
    }
  return 0;
}

int mom_set_field(rdb_colphys* h, const char* name, const double* data, size_t n) {
  MomState* m = nullptr;
  int st = mom_ensure(h, &m);
  if (st) return st;
  double* d = m->field(name);
  if (!d) {
    st = mom_alloc(h, &d, 8 * n);
    if (st) return st;
    m->fields.push_back({name, d});
  }
  MOM_TRY("field upload", hipMemcpyAsync(d, data, 8 * n, hipMemcpyHostToDevice, h->stream));
  MOM_TRY("field upload", hipStreamSynchronize(h->stream));
  return 0;
}

int mom_get_field(rdb_colphys* h, const char* name, double* out, size_t n) {
  MomState* m = h->mom;
  const double* d = nullptr;
  if (m) {
    if (!std::strcmp(name, "visc_rem_u"))
      d = m->rem_valid ? m->rem_u : nullptr;
    else if (!std::strcmp(name, "visc_rem_v"))
      d = m->rem_valid ? m->rem_v : nullptr;
    else
      d = m->field(name);
  }
  if (!d) {
    char buf[160];
    snprintf(buf, sizeof buf, "get_field '%s': not set (no momentum solve with it yet)", name);
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_NOT_FOUND, buf);
  }
  MOM_TRY("field download", hipMemcpyAsync(out, d, 8 * n, hipMemcpyDeviceToHost, h->stream));
  MOM_TRY("field download", hipStreamSynchronize(h->stream));
  return 0;
}

int mom_solve(rdb_colphys* h, double dt, const rdb_colphys_momentum_config* c) {
  // vdiff_apply_momentum's dispatch (rdb_ocean_vdiff.F90:596-717)
  MomState* m = nullptr;
  const bool do_stress = c->implicit_stress != 0;  // the handle always holds tau_x / tau_y
  const bool do_remnant = c->remnant != 0;
  const bool solve = !c->remnant_only;
  if (!solve && !do_remnant) return 0;  // nothing to produce
  const bool do_corner = c->use_corner != 0;
  if (!c->use_kv_source && c->K_v_momentum <= 0.0 && !do_stress && !c->implicit_drag && !c->implicit_top_drag &&
      !do_remnant && !do_corner && !c->bbl_glue)
    return 0;  // the pure-vdiff no-op
  int st = mom_ensure(h, &m);
  if (st) return st;
  const bool do_drag = c->implicit_drag && m->field("lambda_bot_u") && m->field("lambda_bot_v");
  const bool do_top = c->implicit_top_drag && m->field("lambda_top_u") && m->field("lambda_top_v") &&
                      m->field("cover_u") && m->field("cover_v");
  if (do_corner && !m->field("kv_corner"))
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "vdiff_momentum: use_corner without kv_corner");
  if (c->bbl_glue && !c->hvel_mom6)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "vdiff_momentum: bbl_glue requires hvel_mom6");
  if (c->bbl_glue && c->bbl_per_face &&
      !(m->field("kv_bbl_u") && m->field("kv_bbl_v") && m->field("bbl_thick_u") && m->field("bbl_thick_v")))
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE,
                                  "vdiff_momentum: bbl_per_face without kv_bbl_u/v and bbl_thick_u/v");
  const size_t nz = (size_t)h->nz, C = m->C, Fmax = std::max(m->Fu, m->Fv);
  if (!m->h_nat) {
    st = mom_alloc(h, &m->h_nat, 8 * nz * C);
    if (!st) st = mom_alloc(h, &m->kv_nat, 8 * (nz + 1) * C);
    if (!st) st = mom_alloc(h, &m->wet_nat, 8 * C);
    if (!st) st = mom_alloc(h, &m->tau_u, 8 * m->Fu);
    if (!st) st = mom_alloc(h, &m->tau_v, 8 * m->Fv);
    if (!st) st = mom_alloc(h, &m->cd, 8 * nz * Fmax);
    if (!st) st = mom_alloc(h, &m->rd, 8 * nz * Fmax);
    if (st) return st;
  }
  if (do_remnant && !m->rem_u) {
    st = mom_alloc(h, &m->rem_u, 8 * nz * m->Fu);
    if (!st) st = mom_alloc(h, &m->rem_v, 8 * nz * m->Fv);
    if (st) return st;
  }
  // the cell fields in natural order (the packed state is permuted by depth)
  const int ncol = h->n_all, P = h->P, B = 256;
  // (h and the wet mask only when h_layer was set since the last copy; the
  // wet mask is fixed at create)
  if (m->h_gen != h->h_gen) {
    hipLaunchKernelGGL(k_unpack_layer, dim3(mgrid(ncol, B).x, (unsigned)nz), dim3(B), 0, h->stream, h->s.h,
                       m->h_nat, h->d_col, ncol, P, (int)nz, (long long)C);
    hipLaunchKernelGGL(k_unpack_layer, dim3(mgrid(ncol, B).x, 1), dim3(B), 0, h->stream, h->s.wet, m->wet_nat,
                       h->d_col, ncol, P, 1, (long long)C);
    m->h_gen = h->h_gen;
  }
  if (c->use_kv_source)
    hipLaunchKernelGGL(k_unpack_layer, dim3(mgrid(ncol, B).x, (unsigned)(nz + 1)), dim3(B), 0, h->stream, h->s.kv,
                       m->kv_nat, h->d_col, ncol, P, (int)nz + 1, (long long)C);
  MOM_TRY("unpack", hipGetLastError());
  if (do_stress) {
    MOM_TRY("tau upload", hipMemcpyAsync(m->tau_u, h->h_taux.data(), 8 * m->Fu, hipMemcpyHostToDevice, h->stream));
    MOM_TRY("tau upload", hipMemcpyAsync(m->tau_v, h->h_tauy.data(), 8 * m->Fv, hipMemcpyHostToDevice, h->stream));
  }
  MomFlags fl{};
  fl.do_stress = do_stress;
  fl.do_drag = do_drag;
  fl.do_top = do_top;
  fl.solve = solve;
  fl.remnant = do_remnant;
  fl.mom6 = c->hvel_mom6 != 0;
  fl.harm = c->hvel_harmonic != 0;
  fl.upwind = c->hvel_upwind != 0;
  fl.glue = c->bbl_glue != 0;
  fl.zlevel = c->zlevel_faces != 0;
  fl.corner = do_corner;
  fl.use_harmonic = h->cfg.use_harmonic;
  fl.per_face = c->bbl_per_face != 0;
  fl.scalar_kv = !c->use_kv_source;
  fl.rho0 = h->cfg.rho0;
  fl.hbbl_visc = c->hbbl_visc;
  fl.kv_bbl_bg = c->kv_bbl_bg;
  fl.prandtl = c->kv_corner_prandtl;
  fl.kv_bbl_c = c->bbl_piston * c->hbbl_visc;  // fill_bbl_constants (rdb_ocean_vdiff.F90:1093-1105)
  fl.bbl_thick_c = c->hbbl_visc;
  fl.K = c->K_v_momentum;
  for (int xf = 1; xf >= 0; --xf) {
    MomView v{};
    v.nu = h->nx + xf;
    v.nv = h->ny + (1 - xf);
    v.nz = (int)nz;
    v.nx = h->nx;
    v.ny = h->ny;
    v.x_face = xf;
    v.u = xf ? m->uf : m->vf;
    v.h = m->h_nat;
    v.kv = m->kv_nat;
    v.wet = m->wet_nat;
    v.tau = xf ? m->tau_u : m->tau_v;
    v.lb = m->field(xf ? "lambda_bot_u" : "lambda_bot_v");
    v.lt = m->field(xf ? "lambda_top_u" : "lambda_top_v");
    v.cov = m->field(xf ? "cover_u" : "cover_v");
    v.kvb = m->field(xf ? "kv_bbl_u" : "kv_bbl_v");
    v.bth = m->field(xf ? "bbl_thick_u" : "bbl_thick_v");
    v.kvc = m->field("kv_corner");
    v.ktop = m->field(xf ? "k_top_u" : "k_top_v");
    v.kbot = m->field(xf ? "k_bot_u" : "k_bot_v");
    v.cd = m->cd;
    v.rd = m->rd;
    v.rem = xf ? m->rem_u : m->rem_v;
    const size_t F = xf ? m->Fu : m->Fv;
    hipLaunchKernelGGL(k_vdiff_momentum, mgrid(F, B), dim3(B), 0, h->stream, v, fl, dt);
    MOM_TRY("k_vdiff_momentum", hipGetLastError());
  }
  if (do_remnant) m->rem_valid = true;
  if (solve) {
    // the column step reads the centred velocities
    hipLaunchKernelGGL(k_pack_centre_x, dim3(mgrid(ncol, B).x, (unsigned)nz), dim3(B), 0, h->stream, m->uf, h->s.ub,
                       h->d_col, ncol, P, (int)nz, h->nx, h->ny);
    hipLaunchKernelGGL(k_pack_centre_y, dim3(mgrid(ncol, B).x, (unsigned)nz), dim3(B), 0, h->stream, m->vf, h->s.vb,
                       h->d_col, ncol, P, (int)nz, h->nx, h->ny);
    MOM_TRY("centre", hipGetLastError());
    m->u_host_stale = m->v_host_stale = true;
    h->static_dirty = true;  // land-column outputs depend on u and v
  }
  ++h->state_ver;
  return 0;
}

}  // namespace colphys

extern "C" {

int rdb_colphys_momentum_config_default(rdb_colphys_momentum_config* c) {
  if (!c) return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "momentum_config_default: NULL");
  std::memset(c, 0, sizeof *c);
  // ocean_vdiff_t defaults (rdb_ocean_vdiff.F90:77-280); the closure's kv
  // is the source, as vmix_apply_in_stage passes it
  c->use_kv_source = 1;
  c->K_v_momentum = 0.0;
  c->hvel_upwind = 1;
  c->hbbl_visc = 10.0;
  c->bbl_piston = 3.0e-4;
  c->kv_bbl_bg = 1.0e-4;
  c->kv_corner_prandtl = 1.0;
  return RDB_COLPHYS_OK;
}

}  // extern "C"
