// boundary.hip -- the rdb_ocean.h host surface of the column engine
// (include/rdb_ocean.h:108-119, 283-446): create from namelist text, the
// queries (grid, time, step count), pointers into host mirrors with the
// step-count generation and an explicit refresh, the tracer registry and
// the narrow setters.
//
// Host mirrors are full-extent host arrays in the reference layout, one per
// field a getter has been asked for, kept in the handle.  A getter never
// copies from the device except to fill a mirror it creates; every entry
// point that changes device state bumps the handle's state version, and
// rdb_colphys_refresh_host copies exactly the mirrors older than it.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/rdb_colphys.h"
#include "colphys_handle.h"
#include "colphys_kernels.h"

namespace {

int err(int status, const char* fmt, const char* a = "", const char* b = "") {
  char buf[400];
  snprintf(buf, sizeof buf, fmt, a, b);
  return rdb_colphys_push_error(status, buf);
}

bool single_grid(rdb_colphys* h, const char* who, int* st) {
  if (!colphys::handle_valid(h)) {
    *st = err(RDB_COLPHYS_ERR_BAD_HANDLE, "%s: bad handle", who);
    return false;
  }
  if (!h->members.empty()) {
    *st = err(RDB_COLPHYS_ERR_BAD_HANDLE, "%s: an ensemble handle is reached per member (rdb_colphys_member_*)", who);
    return false;
  }
  return true;
}

// full extents of a mirrored field (getter shape)
bool mirror_extents(rdb_colphys* h, const std::string& name, int* n1, int* n2, int* n3) {
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  *n1 = nx, *n2 = ny, *n3 = 1;
  if (name == "h_layer" || name == "rho_layer" || name == "temperature" || name == "salinity" ||
      name.compare(0, 7, "tracer_") == 0 || name == "heat_budget_vdiff" || name == "salt_budget_vdiff")
    *n3 = nz;
  else if (name == "kv" || name == "kt" || name == "ks")
    *n3 = nz + 1;
  else if (name == "u_face_x_layer")
    *n1 = nx + 1, *n3 = nz;
  else if (name == "v_face_y_layer")
    *n2 = ny + 1, *n3 = nz;
  else if (name == "tau_x")
    *n1 = nx + 1;
  else if (name == "tau_y")
    *n2 = ny + 1;
  else if (name != "wet_t" && name != "Q_heat" && name != "Q_salt")
    return false;
  return true;
}

// copy the device field into its mirror (rho_layer: the Wright EOS of the
// current T, S, h first, as the reference's step leaves rho_layer)
int fill_mirror(rdb_colphys* h, const std::string& name, std::vector<double>& m) {
  int n1, n2, n3;
  mirror_extents(h, name, &n1, &n2, &n3);
  m.resize((size_t)n1 * n2 * n3);
  if (name == "rho_layer") {
    dim3 g((unsigned)((h->n_all + h->block - 1) / h->block), (unsigned)h->nz);
    hipLaunchKernelGGL(colphys::k_eos, g, dim3(h->block), 0, h->stream, h->s, h->pr, 0, h->n_all);
    if (hipGetLastError() != hipSuccess) return err(RDB_COLPHYS_ERR_SETUP, "refresh_host: EOS launch failed");
  }
  const long long ver = h->state_ver;
  int st = rdb_colphys_get_field(h, name.c_str(), m.data(), n1, n2, n3);
  if (st) return st;
  h->mirror_ver[name] = ver;
  return 0;
}

int get_ptr(rdb_colphys* h, const char* who, const std::string& name, void** ptr, int* nx, int* ny, int* nz,
            int* gen) {
  int st = 0;
  if (!single_grid(h, who, &st)) return st;
  if (!ptr) return err(RDB_COLPHYS_ERR_BAD_SHAPE, "%s: NULL ptr", who);
  int n1, n2, n3;
  if (!mirror_extents(h, name, &n1, &n2, &n3)) return err(RDB_COLPHYS_ERR_NOT_FOUND, "%s: no field '%s'", who,
                                                          name.c_str());
  auto it = h->mirrors.find(name);
  if (it == h->mirrors.end()) {
    // the first request creates the mirror and fills it
    std::vector<double>& m = h->mirrors[name];
    st = fill_mirror(h, name, m);
    if (st) {
      h->mirrors.erase(name);
      return st;
    }
    it = h->mirrors.find(name);
  }
  *ptr = it->second.data();
  if (nx) *nx = n1;
  if (ny) *ny = n2;
  if (nz) *nz = n3;
  if (gen) *gen = (int)h->steps;
  return RDB_COLPHYS_OK;
}

std::string tracer_field(rdb_colphys* h, const char* name, int name_len, bool* ok) {
  *ok = false;
  if (!name || name_len <= 0) return "";
  std::string n(name, (size_t)strnlen(name, (size_t)name_len));
  if (n == "temperature" || n == "salinity") {
    *ok = true;
    return n;
  }
  if (n.compare(0, 7, "tracer_") == 0) {
    int t = std::atoi(n.c_str() + 7);
    if (t >= 1 && t <= h->n_passive && n == "tracer_" + std::to_string(t)) {
      *ok = true;
      return n;
    }
  }
  return n;
}

// ---- namelist text ---------------------------------------------------------
// Fortran namelist groups "&name key = value, ... /" (comments after '!',
// keys case-insensitive, logicals .true./.false./T/F, reals with e or d
// exponents, strings in quotes).  Only the groups of this path are read;
// other groups of a model namelist are skipped.

struct NmlEntry {
  std::string group, key, value;
};

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

int parse_nml(const char* text, int len, std::vector<NmlEntry>& out) {
  if (!text || len < 0) return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "namelist: NULL text or negative length");
  std::string t(text, (size_t)len);
  // drop comments (outside quotes)
  std::string s;
  bool in_q = false;
  char q = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    char c = t[i];
    if (in_q) {
      s += c;
      if (c == q) in_q = false;
    } else if (c == '\'' || c == '"') {
      in_q = true;
      q = c;
      s += c;
    } else if (c == '!') {
      while (i < t.size() && t[i] != '\n') ++i;
      s += '\n';
    } else {
      s += c;
    }
  }
  size_t i = 0;
  while (i < s.size()) {
    if (s[i] != '&') {
      ++i;
      continue;
    }
    size_t j = i + 1;
    while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
    std::string group = lower(s.substr(i + 1, j - i - 1));
    if (group == "end") {
      i = j;
      continue;
    }
    // body up to the terminating '/' (or &end) outside quotes
    size_t k = j;
    in_q = false;
    while (k < s.size()) {
      char c = s[k];
      if (in_q) {
        if (c == q) in_q = false;
      } else if (c == '\'' || c == '"') {
        in_q = true;
        q = c;
      } else if (c == '/' || c == '&') {
        break;
      }
      ++k;
    }
    if (k >= s.size()) return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "namelist: group &%s is not terminated by '/'",
                                  group.c_str());
    const std::string body = s.substr(j, k - j);
    // split "key = value" pairs: a new pair starts at an identifier followed by '='
    size_t p = 0;
    while (p < body.size()) {
      while (p < body.size() && (std::isspace((unsigned char)body[p]) || body[p] == ',')) ++p;
      if (p >= body.size()) break;
      size_t e = body.find('=', p);
      if (e == std::string::npos)
        return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "namelist &%s: expected 'key = value' near '%s'", group.c_str(),
                   trim(body.substr(p, 40)).c_str());
      std::string key = lower(trim(body.substr(p, e - p)));
      size_t v = e + 1, w = v;
      bool vq = false;
      char vqc = 0;
      while (w < body.size()) {
        char c = body[w];
        if (vq) {
          if (c == vqc) vq = false;
        } else if (c == '\'' || c == '"') {
          vq = true;
          vqc = c;
        } else if (c == ',' || c == '\n') {
          break;
        }
        ++w;
      }
      std::string val = trim(body.substr(v, w - v));
      if (key.empty() || val.empty())
        return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "namelist &%s: empty key or value for '%s'", group.c_str(),
                   key.c_str());
      out.push_back({group, key, val});
      p = w;
    }
    i = (s[k] == '/') ? k + 1 : k;
  }
  return 0;
}

int as_bool(const NmlEntry& e, int* out) {
  std::string v = lower(e.value);
  if (v == ".true." || v == "t" || v == ".t." || v == "true") {
    *out = 1;
    return 0;
  }
  if (v == ".false." || v == "f" || v == ".f." || v == "false") {
    *out = 0;
    return 0;
  }
  return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "namelist: '%s' expects a logical, got '%s'", e.key.c_str(),
             e.value.c_str());
}

int as_real(const NmlEntry& e, double* out) {
  std::string v = lower(e.value);
  for (char& c : v)
    if (c == 'd') c = 'e';  // Fortran double-precision exponent
  size_t us = v.find('_');   // a kind suffix (1.0_wp, 1.0_8)
  if (us != std::string::npos) v = v.substr(0, us);
  if (v == "huge" || v == "huge(1.0)") {
    *out = std::numeric_limits<double>::max();
    return 0;
  }
  char* end = nullptr;
  double x = std::strtod(v.c_str(), &end);
  if (v.empty() || !end || *end != '\0' || !std::isfinite(x))
    return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "namelist: '%s' expects a real, got '%s'", e.key.c_str(),
               e.value.c_str());
  *out = x;
  return 0;
}

int as_int(const NmlEntry& e, int* out) {
  char* end = nullptr;
  long x = std::strtol(e.value.c_str(), &end, 10);
  if (!end || *end != '\0')
    return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "namelist: '%s' expects an integer, got '%s'", e.key.c_str(),
               e.value.c_str());
  *out = (int)x;
  return 0;
}

std::string as_str(const NmlEntry& e) {
  std::string v = e.value;
  if (v.size() >= 2 && (v[0] == '\'' || v[0] == '"') && v.back() == v[0]) v = v.substr(1, v.size() - 2);
  return lower(trim(v));
}

bool group_is(const std::string& g, const char* base) {
  return g == base || g == std::string(base) + "_nml";
}

// Keys of the contributor groups: (name, kind, field).  kind 'b' logical,
// 'r' real, 'i' integer.
struct KeySpec {
  const char* key;
  char kind;
  size_t off;
};
#define RK(k, f) {k, 'r', offsetof(rdb_colphys_config, f)}
#define BK(k, f) {k, 'b', offsetof(rdb_colphys_config, f)}
#define IK(k, f) {k, 'i', offsetof(rdb_colphys_config, f)}
const KeySpec kConv[] = {BK("enable", conv_enable), RK("kd_conv", conv_kd), RK("prandtl_conv", conv_prandtl),
                         RK("n2_thresh", conv_n2_thresh)};
const KeySpec kDdiff[] = {BK("enable", ddiff_enable),          RK("strat_param_max", ddiff_strat_param_max),
                          RK("kappa_ddiff_s", ddiff_kappa_s),  RK("ddiff_exp1", ddiff_exp1),
                          RK("ddiff_exp2", ddiff_exp2),        RK("param1", ddiff_param1),
                          RK("param2", ddiff_param2),          RK("param3", ddiff_param3),
                          RK("mol_diff", ddiff_mol_diff),      BK("use_k90", ddiff_use_k90)};
const KeySpec kEpbl[] = {
    BK("enable", epbl_enable),
    RK("mstar", epbl_mstar),
    RK("mstar_cap", epbl_mstar_cap),
    RK("mstar_coef1", epbl_mstar_coef1),
    RK("c_ek", epbl_c_ek),
    RK("mstar_conv_adj", epbl_mstar_conv_adj),
    RK("rh18_cn1", epbl_rh18_cn1),
    RK("rh18_cn2", epbl_rh18_cn2),
    RK("rh18_cn3", epbl_rh18_cn3),
    RK("rh18_cs1", epbl_rh18_cs1),
    RK("rh18_cs2", epbl_rh18_cs2),
    RK("nstar", epbl_nstar),
    RK("tke_decay", epbl_tke_decay),
    RK("wstar_ustar_coef", epbl_wstar_ustar_coef),
    RK("vstar_scale_fac", epbl_vstar_scale_fac),
    RK("vstar_surf_fac", epbl_vstar_surf_fac),
    RK("von_karman", epbl_von_karman),
    RK("ekman_scale_coef", epbl_ekman_scale_coef),
    RK("min_mix_len", epbl_min_mix_len),
    RK("mixlen_exponent", epbl_mixlen_exponent),
    RK("translay_scale", epbl_translay_scale),
    BK("mld_iteration", epbl_mld_iteration),
    RK("mld_tol", epbl_mld_tol),
    IK("mld_max_its", epbl_mld_max_its),
    BK("mld_bisection", epbl_mld_bisection),
    BK("mld_use_prev_guess", epbl_mld_use_prev_guess),
    RK("omega", epbl_omega),
    RK("omega_frac", epbl_omega_frac),
    RK("prandtl", epbl_prandtl),
    BK("tke_diags", epbl_tke_diags),
    BK("use_lt", epbl_use_lt),
    RK("lt_enhance_coef", epbl_lt_enhance_coef),
    RK("lt_enhance_exp", epbl_lt_enhance_exp),
    RK("lt_max_enhance", epbl_lt_max_enhance),
    RK("la_frac_hbl", epbl_la_frac_hbl),
    RK("lt_lac1", epbl_lt_lac1),
    RK("lt_lac2", epbl_lt_lac2),
    RK("lt_lac3", epbl_lt_lac3),
    RK("lt_lac4", epbl_lt_lac4),
    RK("lt_lac5", epbl_lt_lac5),
};
#undef RK
#undef BK
#undef IK

int contrib_key(const NmlEntry& e, rdb_colphys_config* c, const char* group, const KeySpec* tab, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    if (e.key != tab[i].key) continue;
    char* base = reinterpret_cast<char*>(c) + tab[i].off;
    int st = 0;
    if (tab[i].kind == 'r') {
      double x = 0.0;
      if ((st = as_real(e, &x))) return st;
      *reinterpret_cast<double*>(base) = x;
    } else if (tab[i].kind == 'b') {
      int b = 0;
      if ((st = as_bool(e, &b))) return st;
      *reinterpret_cast<int*>(base) = b;
    } else {
      int v = 0;
      if ((st = as_int(e, &v))) return st;
      *reinterpret_cast<int*>(base) = v;
    }
    return 0;
  }
  return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "%s: unknown key '%s'", group, e.key.c_str());
}

// &ocean_vmix_nml (src/core/config/rdb_config_register_vmix.F90:738-870),
// &ocean_eos_nml (:594-650: c_p, p_ref, eos) and &ocean_ic_nml rho_0
// (src/core/config/rdb_config_register_core.F90:806) into a config that
// starts from the reference defaults.  Knobs of the model driver that do
// not act on the column path (time-step ratios, horizontal advection) are
// accepted and have no effect; switches of branches this engine does not
// run are refused with RDB_COLPHYS_ERR_CONFIG_VALIDATE.
int config_from_entries(const std::vector<NmlEntry>& ents, rdb_colphys_config* c) {
  int st = 0;
  for (const NmlEntry& e : ents) {
    const std::string& k = e.key;
    int b = 0;
    double x = 0.0;
    int n = 0;
    if (group_is(e.group, "ocean_vmix")) {
      if (k == "use_closure") {
        if ((st = as_bool(e, &b))) return st;
        if (!b) return err(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_vmix_nml: use_closure = .false. is not run by "
                                                            "the column engine (it replaces the closure)");
      } else if (k == "interior_closure") {
        if (as_str(e) != "pp81")
          return err(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_vmix_nml: interior_closure = '%s' (only pp81)",
                     as_str(e).c_str());
      } else if (k == "use_kpp") {
        if ((st = as_bool(e, &b))) return st;
        c->use_kpp = b;
      } else if (k == "direct_stress") {
        if ((st = as_bool(e, &b))) return st;
        if (b) return err(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_vmix_nml: direct_stress is not supported");
      } else if (k == "kv_ml_invz2") {
        if ((st = as_real(e, &x))) return st;
        c->kv_ml_invz2 = x;
      } else if (k == "harmonic_visc") {
        if ((st = as_bool(e, &b))) return st;
        c->use_harmonic = b;
      } else if (k == "kv_max") {
        if ((st = as_real(e, &x))) return st;
        c->kv_max = x;
      } else if (k == "kd_max") {
        if ((st = as_real(e, &x))) return st;
        c->kd_max = x;
      } else if (k == "kd_smooth_iterations") {
        if ((st = as_int(e, &n))) return st;
        c->kd_smooth_iterations = n;
      } else if (k == "vmix_guard") {
        if ((st = as_bool(e, &b))) return st;
        c->vmix_guard = b;
      } else if (k == "bkgnd_profile" || k == "bkgnd_henyey") {
        if ((st = as_bool(e, &b))) return st;
        if (b) return err(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_vmix_nml: %s is not supported", k.c_str());
      } else if (k == "bkgnd_prandtl") {
        if ((st = as_real(e, &x))) return st;
        if (x != -1.0) return err(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_vmix_nml: bkgnd_prandtl != -1 is not "
                                                                   "supported");
      } else if (k == "pp81_nu0") {
        if ((st = as_real(e, &x))) return st;
        c->pp81_nu0 = x;
      } else if (k == "pp81_nu_bg") {  // also seeds vmix%kv_bg
        if ((st = as_real(e, &x))) return st;
        c->pp81_nu_bg = x;
        c->kv_bg = x;
      } else if (k == "pp81_kappa_bg") {  // also seeds vmix%kt_bg and ks_bg
        if ((st = as_real(e, &x))) return st;
        c->pp81_kappa_bg = x;
        c->kt_bg = x;
        c->ks_bg = x;
      } else if (k == "pp81_alpha") {
        if ((st = as_real(e, &x))) return st;
        c->pp81_alpha = x;
      } else if (k == "shear2_floor") {
        if ((st = as_real(e, &x))) return st;
        c->shear2_floor = x;
      } else if (k == "kpp_ri_crit") {
        if ((st = as_real(e, &x))) return st;
        c->ri_crit = x;
      } else if (k == "kpp_cs_nonlocal") {
        if ((st = as_real(e, &x))) return st;
        c->cs_nonlocal = x;
      } else if (k == "kpp_c_vt2") {
        if ((st = as_real(e, &x))) return st;
        c->c_vt2 = x;
      } else if (k == "buoyancy_coeffs") {
        if (as_str(e) != "constant")
          return err(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_vmix_nml: buoyancy_coeffs = '%s' (only constant)",
                     as_str(e).c_str());
      } else if (k == "hmix_fixed") {
        if ((st = as_real(e, &x))) return st;
        c->hmix_fixed = x;
      } else if (k == "hmix_stress" || k == "bkgnd_kd_sfc" || k == "bkgnd_kd_deep" ||
                 k == "bkgnd_z0" || k == "bkgnd_delta" || k == "bkgnd_kd_min" || k == "bkgnd_henyey_n0_2omega" ||
                 k == "bkgnd_henyey_max_lat" || k == "dt_therm_ratio" || k == "diabatic_first" ||
                 k == "dt_tracer_advect_ratio" || k == "tracer_advection_scheme" || k == "tracer_recon") {
        // parameters of off branches, or of the model driver: no effect here
      } else {
        return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "ocean_vmix_nml: unknown key '%s'", k.c_str());
      }
    } else if (group_is(e.group, "ocean_eos")) {
      if (k == "eos") {
        if (as_str(e) != "wright")
          return err(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_eos_nml: eos = '%s' (only wright)", as_str(e).c_str());
      } else if (k == "c_p") {
        if ((st = as_real(e, &x))) return st;
        c->cp = x;
      } else if (k == "p_ref") {
        if ((st = as_real(e, &x))) return st;
        c->p_ref = x;
      }
      // freezing-point knobs and the rest of the group do not act on this path
    } else if (group_is(e.group, "ocean_ic")) {
      if (k == "rho_0") {
        if ((st = as_real(e, &x))) return st;
        c->rho0 = x;
      }
    } else if (group_is(e.group, "ocean_conv")) {
      // &ocean_conv_nml (rdb_config_register_vmix.F90:226-250)
      if ((st = contrib_key(e, c, "ocean_conv_nml", kConv, sizeof kConv / sizeof kConv[0]))) return st;
    } else if (group_is(e.group, "ocean_ddiff")) {
      // &ocean_ddiff_nml (:252-292)
      if ((st = contrib_key(e, c, "ocean_ddiff_nml", kDdiff, sizeof kDdiff / sizeof kDdiff[0]))) return st;
    } else if (group_is(e.group, "ocean_epbl")) {
      // &ocean_epbl_nml (:294-393)
      if (k == "mstar_scheme" || k == "vel_scale_scheme" || k == "combine" || k == "lt_scheme") {
        const std::string v = as_str(e);
        int tag = -1;
        if (k == "mstar_scheme") {
          tag = v == "constant" ? RDB_EPBL_MSTAR_CONSTANT
                : v == "om4"    ? RDB_EPBL_MSTAR_OM4
                : (v == "rh18" || v == "reichl_h18") ? RDB_EPBL_MSTAR_RH18 : -1;
          c->epbl_mstar_scheme = tag;
        } else if (k == "vel_scale_scheme") {
          tag = (v == "cube_root" || v == "cube_root_tke") ? RDB_EPBL_VSTAR_CUBE_ROOT
                : (v == "rh18" || v == "reichl_h18")      ? RDB_EPBL_VSTAR_RH18 : -1;
          c->epbl_vstar_scheme = tag;
        } else if (k == "combine") {
          tag = (v == "add" || v == "additive") ? RDB_EPBL_COMBINE_ADD
                : v == "max"       ? RDB_EPBL_COMBINE_MAX
                : v == "max_shear" ? RDB_EPBL_COMBINE_MAX_SHEAR : -1;
          c->epbl_combine = tag;
        } else {
          tag = v == "rescale" ? RDB_EPBL_LT_RESCALE : (v == "additive" || v == "add") ? RDB_EPBL_LT_ADDITIVE : -1;
          c->epbl_lt_scheme = tag;
        }
        if (tag < 0)
          return err(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "ocean_epbl_nml: unknown %s '%s'", k.c_str(), v.c_str());
      } else if ((st = contrib_key(e, c, "ocean_epbl_nml", kEpbl, sizeof kEpbl / sizeof kEpbl[0]))) {
        return st;
      }
    }
  }
  return 0;
}

// &ocean_vdiff_nml (rdb_config_register_vmix.F90:880-960): the momentum
// solve's switches
int momentum_from_entries(const std::vector<NmlEntry>& ents, rdb_colphys_momentum_config* m) {
  int st = 0;
  for (const NmlEntry& e : ents) {
    if (!group_is(e.group, "ocean_vdiff")) continue;
    const std::string& k = e.key;
    int b = 0;
    double x = 0.0;
    if (k == "implicit_stress" || k == "implicit_drag" || k == "implicit_top_drag" || k == "hvel_mom6" ||
        k == "hvel_harmonic" || k == "bbl_glue" || k == "hvel_upwind" || k == "accel_visc_rem") {
      if ((st = as_bool(e, &b))) return st;
      if (k == "implicit_stress") m->implicit_stress = b;
      if (k == "implicit_drag") m->implicit_drag = b;
      if (k == "implicit_top_drag") m->implicit_top_drag = b;
      if (k == "hvel_mom6") m->hvel_mom6 = b;
      if (k == "hvel_harmonic") m->hvel_harmonic = b;
      if (k == "bbl_glue") m->bbl_glue = b;
      if (k == "hvel_upwind") m->hvel_upwind = b;
      if (k == "accel_visc_rem") m->remnant = b;  // the remnant is what the acceleration weighting reads
    } else if (k == "hbbl_visc" || k == "bbl_piston") {
      if ((st = as_real(e, &x))) return st;
      if (k == "hbbl_visc") m->hbbl_visc = x;
      if (k == "bbl_piston") m->bbl_piston = x;
    } else if (k == "h_min" || k == "ppm_limit_pos" || k == "vol_cfl" || k == "positive_definite" ||
               k == "renorm_consistent_flux") {
      // knobs of the vertical advection in the same group: no effect here
    } else {
      return err(RDB_COLPHYS_ERR_CONFIG_PARSE, "ocean_vdiff_nml: unknown key '%s'", k.c_str());
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int rdb_colphys_config_from_string(const char* nml_text, int nml_len, rdb_colphys_config* cfg) {
  if (!cfg) return err(RDB_COLPHYS_ERR_BAD_SHAPE, "config_from_string: NULL config");
  std::vector<NmlEntry> ents;
  int st = parse_nml(nml_text, nml_len, ents);
  if (st) return st;
  rdb_colphys_config c;
  rdb_colphys_config_default(&c);
  st = config_from_entries(ents, &c);
  if (st) return st;
  *cfg = c;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_momentum_config_from_string(const char* nml_text, int nml_len, rdb_colphys_momentum_config* cfg) {
  if (!cfg) return err(RDB_COLPHYS_ERR_BAD_SHAPE, "momentum_config_from_string: NULL config");
  std::vector<NmlEntry> ents;
  int st = parse_nml(nml_text, nml_len, ents);
  if (st) return st;
  rdb_colphys_momentum_config m;
  rdb_colphys_momentum_config_default(&m);
  st = momentum_from_entries(ents, &m);
  if (st) return st;
  *cfg = m;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_create_from_string(const char* nml_text, int nml_len, int nx_tot, int ny_tot, int nz, int n_passive,
                                   const double* wet, const int* k_top, const int* sort_key, rdb_colphys** out) {
  if (out) *out = nullptr;
  rdb_colphys_config c;
  int st = rdb_colphys_config_from_string(nml_text, nml_len, &c);
  if (st) return st;
  return rdb_colphys_create(&c, nx_tot, ny_tot, nz, n_passive, wet, k_top, sort_key, out);
}

int rdb_colphys_get_grid_info(rdb_colphys* h, int* nx, int* ny, int* nz, int* nghost) {
  int st = 0;
  if (!single_grid(h, "get_grid_info", &st)) return st;
  if (nx) *nx = h->nx;
  if (ny) *ny = h->ny;
  if (nz) *nz = h->nz;
  if (nghost) *nghost = 0;  // the column engine's extents are the caller's full arrays
  return RDB_COLPHYS_OK;
}

int rdb_colphys_get_time(rdb_colphys* h, double* t_out) {
  int st = 0;
  if (!single_grid(h, "get_time", &st)) return st;
  if (!t_out) return err(RDB_COLPHYS_ERR_BAD_SHAPE, "get_time: NULL output");
  *t_out = h->t_sim;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_get_step_count(rdb_colphys* h, int* step_out) {
  int st = 0;
  if (!single_grid(h, "get_step_count", &st)) return st;
  if (!step_out) return err(RDB_COLPHYS_ERR_BAD_SHAPE, "get_step_count: NULL output");
  *step_out = (int)h->steps;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_refresh_host(rdb_colphys* h) {
  int st = 0;
  if (!single_grid(h, "refresh_host", &st)) return st;
  for (auto& kv : h->mirrors) {
    auto v = h->mirror_ver.find(kv.first);
    if (v != h->mirror_ver.end() && v->second == h->state_ver) continue;  // already current
    st = fill_mirror(h, kv.first, kv.second);
    if (st) return st;
  }
  return RDB_COLPHYS_OK;
}

#define RDB_GETTER3(fn, field)                                                                     \
  int fn(rdb_colphys* h, void** ptr, int* nx, int* ny, int* nz, int* gen) {                        \
    return get_ptr(h, #fn, field, ptr, nx, ny, nz, gen);                                          \
  }
#define RDB_GETTER2(fn, field)                                                                     \
  int fn(rdb_colphys* h, void** ptr, int* nx, int* ny, int* gen) {                                 \
    return get_ptr(h, #fn, field, ptr, nx, ny, nullptr, gen);                                     \
  }
RDB_GETTER3(rdb_colphys_get_h_layer_ptr, "h_layer")
RDB_GETTER3(rdb_colphys_get_u_face_x_layer_ptr, "u_face_x_layer")
RDB_GETTER3(rdb_colphys_get_v_face_y_layer_ptr, "v_face_y_layer")
RDB_GETTER3(rdb_colphys_get_rho_layer_ptr, "rho_layer")
RDB_GETTER3(rdb_colphys_get_kv_ptr, "kv")
RDB_GETTER3(rdb_colphys_get_kt_ptr, "kt")
RDB_GETTER3(rdb_colphys_get_ks_ptr, "ks")
RDB_GETTER2(rdb_colphys_get_wet_t_ptr, "wet_t")
RDB_GETTER2(rdb_colphys_get_tau_x_ptr, "tau_x")
RDB_GETTER2(rdb_colphys_get_tau_y_ptr, "tau_y")
RDB_GETTER2(rdb_colphys_get_q_heat_ptr, "Q_heat")
RDB_GETTER2(rdb_colphys_get_q_salt_ptr, "Q_salt")
#undef RDB_GETTER3
#undef RDB_GETTER2

int rdb_colphys_get_tracer_count(rdb_colphys* h, int* count_out) {
  int st = 0;
  if (!single_grid(h, "get_tracer_count", &st)) return st;
  if (!count_out) return err(RDB_COLPHYS_ERR_BAD_SHAPE, "get_tracer_count: NULL output");
  *count_out = 2 + h->n_passive;
  return RDB_COLPHYS_OK;
}

int rdb_colphys_list_tracers(rdb_colphys* h, int idx, char* buf, int cap) {
  int st = 0;
  if (!single_grid(h, "list_tracers", &st)) return st;
  if (idx < 0 || idx >= 2 + h->n_passive)
    return err(RDB_COLPHYS_ERR_NOT_FOUND, "list_tracers: index %s out of range", std::to_string(idx).c_str());
  const std::string name = idx == 0 ? "temperature" : idx == 1 ? "salinity" : "tracer_" + std::to_string(idx - 1);
  if (buf && cap > 0) snprintf(buf, (size_t)cap, "%s", name.c_str());
  return (int)name.size();  // snprintf semantics: the full length
}

int rdb_colphys_get_tracer_ptr(rdb_colphys* h, const char* name, int name_len, void** ptr, int* nx, int* ny, int* nz,
                               int* gen) {
  int st = 0;
  if (!single_grid(h, "get_tracer_ptr", &st)) return st;
  bool ok = false;
  std::string f = tracer_field(h, name, name_len, &ok);
  if (!ok) return err(RDB_COLPHYS_ERR_NOT_FOUND, "get_tracer_ptr: no tracer '%s'", f.c_str());
  return get_ptr(h, "get_tracer_ptr", f, ptr, nx, ny, nz, gen);
}

int rdb_colphys_set_h(rdb_colphys* h, const double* data, int nx, int ny, int nz) {
  return rdb_colphys_set_field(h, "h_layer", data, nx, ny, nz);
}

int rdb_colphys_set_u(rdb_colphys* h, const double* data, int nx, int ny, int nz) {
  return rdb_colphys_set_field(h, "u_face_x_layer", data, nx + 1, ny, nz);
}

int rdb_colphys_set_v(rdb_colphys* h, const double* data, int nx, int ny, int nz) {
  return rdb_colphys_set_field(h, "v_face_y_layer", data, nx, ny + 1, nz);
}

int rdb_colphys_set_wind(rdb_colphys* h, const double* taux, const double* tauy, int nx, int ny) {
  int st = rdb_colphys_set_field(h, "tau_x", taux, nx + 1, ny, 1);
  if (st) return st;
  return rdb_colphys_set_field(h, "tau_y", tauy, nx, ny + 1, 1);
}

int rdb_colphys_set_heat_flux(rdb_colphys* h, const double* q, int nx, int ny) {
  return rdb_colphys_set_field(h, "Q_heat", q, nx, ny, 1);
}

int rdb_colphys_set_salt_flux(rdb_colphys* h, const double* q, int nx, int ny) {
  return rdb_colphys_set_field(h, "Q_salt", q, nx, ny, 1);
}

int rdb_colphys_set_tracer(rdb_colphys* h, const char* name, int name_len, const double* data, int nx, int ny,
                           int nz) {
  int st = 0;
  if (!single_grid(h, "set_tracer", &st)) return st;
  bool ok = false;
  std::string f = tracer_field(h, name, name_len, &ok);
  if (!ok) return err(RDB_COLPHYS_ERR_NOT_FOUND, "set_tracer: no tracer '%s'", f.c_str());
  return rdb_colphys_set_field(h, f.c_str(), data, nx, ny, nz);
}

}  // extern "C"
