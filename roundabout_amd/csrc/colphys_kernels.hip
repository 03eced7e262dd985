// colphys_kernels.hip -- gfx950 kernels of the vertical-mixing column path.
//
// Data layout in HBM (all kernels): columns are packed, wet columns first
// (ordered by a depth key so that a wavefront holds columns of similar
// depth), then the dry columns.  Every per-level field is stored level-major
// with a pitch P that is a multiple of 64: X[k * P + p].  One lane owns one
// column, so the 64 lanes of a wavefront read 64 consecutive doubles (512 B)
// of one level: fully coalesced, and each column's serial k-sweeps (KPP
// bulk-Ri search, Thomas elimination) run down a register pipeline.
#include "colphys_kernels.h"

namespace colphys {

#define LV(arr, k) (arr)[(size_t)(k) * (size_t)P + (size_t)p]

// Levels of loads the fused step keeps in flight ahead of its serial
// sweeps (register ring depth).
constexpr int PF = 6;
// Sweep A usually stops a few levels below the surface: a shallower ring.
constexpr int PFA = 3;
// Cache policy of the fused step: the streamed inputs of sweep B and the
// final outputs (kv, kt, ks, T, S, tracers) are nontemporal loads and
// stores, so that L2 and the memory-side cache keep the Thomas state that
// sweep B spills and sweep C reads back.

// ---------------------------------------------------------------------------
// Pack / unpack between the reference's full-grid layout (i fastest) and the
// packed level-major layout.  col[p] = i + nx * j.
// ---------------------------------------------------------------------------
__global__ void k_pack_layer(const double* __restrict__ src, double* __restrict__ dst,
                             const int* __restrict__ col, int ncol, int P, int nlev, long long plane) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (p >= ncol || k >= nlev) return;
  LV(dst, k) = src[(size_t)k * (size_t)plane + (size_t)col[p]];
}

// Column repacking for the autotuner: dst[k][newpos[p]] = src[k][p] for the
// packed columns p < n (the wet ones), nlev levels of pitch P.
__global__ void k_permute_d(const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ newpos,
                            int n, int P, int nlev) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (p >= n || k >= nlev) return;
  dst[(size_t)k * (size_t)P + (size_t)newpos[p]] = src[(size_t)k * (size_t)P + (size_t)p];
}
__global__ void k_permute_i(const int* __restrict__ src, int* __restrict__ dst, const int* __restrict__ newpos, int n,
                            int P, int nlev) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (p >= n || k >= nlev) return;
  dst[(size_t)k * (size_t)P + (size_t)newpos[p]] = src[(size_t)k * (size_t)P + (size_t)p];
}

__global__ void k_unpack_layer(const double* __restrict__ src, double* __restrict__ dst,
                               const int* __restrict__ col, int ncol, int P, int nlev, long long plane) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (p >= ncol || k >= nlev) return;
  dst[(size_t)k * (size_t)plane + (size_t)col[p]] = LV(src, k);
}

__global__ void k_unpack_int(const int* __restrict__ src, double* __restrict__ dst,
                             const int* __restrict__ col, int ncol) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncol) return;
  dst[col[p]] = (double)src[p];
}

// Centred velocity from faces: 0.5 * (u(i) + u(i+1)) (x faces) or
// 0.5 * (v(j) + v(j+1)) (y faces), the face average PP81 and KPP form
// (rdb_ocean_vmix.F90:1045-1048, 1302-1303).
__global__ void k_pack_centre_x(const double* __restrict__ uf, double* __restrict__ dst,
                                const int* __restrict__ col, int ncol, int P, int nlev, int nx, int ny) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (p >= ncol || k >= nlev) return;
  int c = col[p];
  int i = c % nx, j = c / nx;
  size_t base = (size_t)k * (size_t)(nx + 1) * (size_t)ny + (size_t)j * (size_t)(nx + 1) + (size_t)i;
  LV(dst, k) = 0.5 * (uf[base] + uf[base + 1]);
}

__global__ void k_pack_centre_y(const double* __restrict__ vf, double* __restrict__ dst,
                                const int* __restrict__ col, int ncol, int P, int nlev, int nx, int ny) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (p >= ncol || k >= nlev) return;
  int c = col[p];
  int i = c % nx, j = c / nx;
  size_t base = (size_t)k * (size_t)nx * (size_t)(ny + 1) + (size_t)j * (size_t)nx + (size_t)i;
  LV(dst, k) = 0.5 * (vf[base] + vf[base + (size_t)nx]);
}

// A16 |tau| from centred stress (rdb_ocean_surface_stress.F90:867-874).
__global__ void k_stress_mag(const double* __restrict__ txc, const double* __restrict__ tyc,
                             double* __restrict__ smag, const double* __restrict__ shelf,
                             double* __restrict__ tau_tot, int ncol) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncol) return;
  double tx = txc[p], ty = tyc[p];
  double s = sqrt(tx * tx + ty * ty);
  smag[p] = s;
  tau_tot[p] = s + shelf[p];  // stress_mag + stress_shelf (rdb_ocean_vmix.F90:1318)
}

// The device cube root of the KPP path (cbrt_kahan) over n arguments.
__global__ void k_cbrt_eval(const double* __restrict__ x, double* __restrict__ y, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = cbrt_kahan(x[i]);
}

// ---------------------------------------------------------------------------
// Staged kernels: one per reference routine, over packed columns [p0, p1).
// ---------------------------------------------------------------------------
__global__ void k_eos(StateView s, Params pr, int p0, int p1) {
  int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (p >= p1 || k >= s.nz) return;
  const int P = s.P;
  LV(s.rho, k) = eos_wright<double>(LV(s.T, k), LV(s.S, k), LV(s.h, k), pr.rho0, pr.p_ref);
}

__global__ void k_pp81(StateView s, Params pr, int p0, int p1) {
  int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const int P = s.P, nz = s.nz;
  LV(s.kv, 0) = 0.0;
  LV(s.kt, 0) = 0.0;
  LV(s.kv, nz) = 0.0;
  LV(s.kt, nz) = 0.0;
  double ub0 = LV(s.ub, 0), vb0 = LV(s.vb, 0), h0 = LV(s.h, 0), r0 = LV(s.rho, 0);
  for (int k = 1; k < nz; ++k) {
    double ub1 = LV(s.ub, k), vb1 = LV(s.vb, k), h1 = LV(s.h, k), r1 = LV(s.rho, k);
    double kv, kt;
    pp81_iface<double>(pr, ub0, vb0, ub1, vb1, h0, h1, r1 - r0, kv, kt);
    LV(s.kv, k) = kv;
    LV(s.kt, k) = kt;
    ub0 = ub1; vb0 = vb1; h0 = h1; r0 = r1;
  }
}

// A4/A5 KPP overlay, both passes for one column (rdb_ocean_vmix.F90:1282-1515).
__global__ void k_kpp(StateView s, Params pr, int p0, int p1) {
  int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const int P = s.P, nz = s.nz;
  for (int k = 0; k <= nz; ++k) {
    LV(s.gamma_t, k) = 0.0;
    LV(s.gamma_s, k) = 0.0;
  }
  int k_sfc = nz - 1;
  for (int k = nz - 1; k >= 0; --k)
    if (LV(s.h, k) > H_VANISHED) { k_sfc = k; break; }
  const double rho0 = pr.rho0;
  double u_ref = LV(s.ub, k_sfc), v_ref = LV(s.vb, k_sfc);
  double b_ref = -(GRAVITY * LV(s.rho, k_sfc) / rho0);
  double d_centre_ref = 0.5 * LV(s.h, k_sfc);
  double tau_mag = s.tau_tot[p];
  double u_star = sqrt(tau_mag / rho0);
  double qT = s.Qh[p] / (rho0 * pr.cp);
  double qS = s.Qs[p] / rho0;
  double B0 = kpp_b0<double>(pr, qT, qS);
  double h_b_lagged = s.bl[p];
  double wstar3_lagged = fmx(0.0, -B0) * h_b_lagged;
  double cb = cbrt_kahan(wstar3_lagged);
  double w_s_col = sqrt(u_star * u_star + cb * cb);
  double ri_prev = 0.0, d_prev = d_centre_ref, h_b = 0.0, d_centre_k = d_centre_ref;
  int kcross = 0;
  double h_up = LV(s.h, k_sfc);
  for (int k = k_sfc - 1; k >= 0; --k) {
    double hk = LV(s.h, k);
    d_centre_k = d_centre_k + 0.5 * (h_up + hk);
    h_up = hk;
    double u_k = LV(s.ub, k), v_k = LV(s.vb, k);
    double b_k = -(GRAVITY * LV(s.rho, k) / rho0);
    double delta_d = d_centre_k - d_centre_ref;
    double delta_b = b_ref - b_k;
    double n_brunt = sqrt(fmx(0.0, delta_b / delta_d));
    double vt2 = (pr.c_vt2 / pr.ri_crit) * d_centre_k * n_brunt * w_s_col;
    double du = u_k - u_ref, dv = v_k - v_ref;
    double shear2 = fmx(du * du + dv * dv + vt2, pr.shear2_floor);
    double ri_bulk = delta_b * delta_d / shear2;
    if (ri_bulk >= pr.ri_crit) {
      double denom = ri_bulk - ri_prev;
      double frac = (fabs(denom) > 1.0e-12) ? (pr.ri_crit - ri_prev) / denom : 0.5;
      h_b = d_prev + frac * (d_centre_k - d_prev);
      kcross = k + 1;
      break;
    }
    ri_prev = ri_bulk;
    d_prev = d_centre_k;
  }
  if (kcross == 0) {
    h_b = 0.0;
    for (int k = 0; k <= k_sfc; ++k) h_b = h_b + LV(s.h, k);
  }
  s.bl[p] = h_b;
  s.kbl[p] = kcross;
  // pass 2
  s.b0[p] = B0;
  bool destab = (B0 < 0.0);
  double wstar3 = fmx(0.0, -B0) * h_b;
  double w_star = cbrt_kahan(wstar3);
  double w_s = sqrt(u_star * u_star + w_star * w_star);
  if (h_b > 0.0 && w_s > 0.0) {
    double d_running = 0.0;
    for (int k = k_sfc; k >= 1; --k) {
      d_running = d_running + LV(s.h, k);
      if (!(d_running < h_b)) break;  // d is non-decreasing: no later interface qualifies
      double sigma = d_running / h_b;
      double g_shape = sigma * (1.0 - sigma) * (1.0 - sigma);
      double kv_kpp = h_b * w_s * g_shape;
      LV(s.kv, k) = fmx(LV(s.kv, k), kv_kpp);
      LV(s.kt, k) = fmx(LV(s.kt, k), kv_kpp);
      if (destab) {
        double gf = pr.cs_nonlocal * g_shape;
        LV(s.gamma_t, k) = gf * qT;
        LV(s.gamma_s, k) = gf * qS;
      }
    }
  }
}

// A9 split + A10 assemble: guard count (before the clip), then floors and
// ceilings on interior interfaces (rdb_ocean_vmix.F90:1921-1934, 2265-2385).
__global__ void k_split_assemble(StateView s, Params pr, int p0, int p1, unsigned int* guard_count) {
  int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (p >= p1 || k > s.nz) return;
  const int P = s.P;
  double kt = LV(s.kt, k);
  double kv = LV(s.kv, k);
  double ks = kt;
  if (pr.ddiff) {
    double kd_t = 0.0, kd_s = 0.0;
    if (k >= 1 && k < s.nz)
      ddiff_iface(pr, LV(s.h, k), LV(s.h, k - 1), LV(s.T, k), LV(s.T, k - 1), LV(s.S, k), LV(s.S, k - 1), kd_t, kd_s);
    ks = kt + kd_s;
    kt = kt + kd_t;
  }
  if (pr.guard) {
    if (isnan(kv) || isnan(kt) || isnan(ks) || kv < 0.0 || kt < 0.0 || ks < 0.0) atomicAdd(guard_count, 1u);
  }
  if (k >= 1 && k < s.nz) {
    kv = fmn(fmx(kv, pr.kv_bg), pr.kv_max);
    kt = fmn(fmx(kt, pr.kt_bg), pr.kd_max);
    ks = fmn(fmx(ks, pr.ks_bg), pr.kd_max);
  }
  LV(s.kv, k) = kv;
  LV(s.kt, k) = kt;
  LV(s.ks, k) = ks;
}

// A2 surface deposit into the first live layer scanning down from k_top
// (apply_surface_src_2d_impl, rdb_ocean_surface_flux.F90:1093-1142).
__global__ void k_surface_flux(StateView s, Params pr, double dt, int p0, int p1) {
  int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const int P = s.P;
  int kstart = s.ktop[p];
  int kt = kstart;
  for (int k = kstart; k >= 0; --k)
    if (LV(s.h, k) > H_VANISHED) { kt = k; break; }
  double hk = LV(s.h, kt);
  double wet = s.wet[p];
  if (pr.has_heat) {
    double cell = dt / (pr.rho0 * pr.cp) * s.Qh[p] * wet;
    LV(s.T, kt) = conc_add_source(LV(s.T, kt), hk, cell);
  }
  if (pr.has_salt) {
    double cell = dt / pr.rho0 * s.Qs[p] * wet;
    LV(s.S, kt) = conc_add_source(LV(s.S, kt), hk, cell);
  }
}

// A11-A13 for a group of G tracers that share one diffusivity field K: the
// rows of build_factorize_tracer_matrix (rdb_ocean_vdiff.F90:1288-1333), the
// pivots and the forward sweep of apply_factored_tracer (:1351-1414) in one
// pass up the column, the back substitution in one pass down.  The
// pivot-scaled super-diagonal c' goes to the fc scratch and the forward
// residuals stay in place in the tracers, so every level is read and
// written once per pass; the arithmetic is the reference's, operation for
// operation (piv = b - a c'_{k-1}, c' = c / piv, r = (x - a r_{k-1}) / piv,
// x = r - c' x_{k+1}).  With a budget slot, budget += h*c_new - h*c_old.
// With kb > 0 the elimination starts at the first live layer: a bed filler
// row is the identity (alpha = beta = 0: a = -0, b = 1, c = 0), so the
// reference's elimination through the fillers leaves their values and the
// first live row's pivot, c' and residual as they are when it starts there
// (c'_{kb-1} = 0 and r_{kb-1} enter only through a * 0).
template <int G>
__device__ __forceinline__ void solve_group(const StateView& s, const double* __restrict__ K, double dt,
                                            int use_harmonic, int p, int kb, double* const (&xs)[G],
                                            double* const (&bud)[G], double* const (&pold)[G]) {
  const int P = s.P, nz = s.nz;
  double r[G];
#pragma unroll
  for (int t = 0; t < G; ++t) r[t] = 0.0;
  double h_km1 = kb > 0 ? LV(s.h, kb - 1) : 0.0, h_k = LV(s.h, kb), K_k = LV(K, kb);
  double cprev = 0.0;
  for (int k = kb; k < nz; ++k) {
    const double h_kp1 = (k + 1 < nz) ? LV(s.h, k + 1) : 0.0;
    const double K_kp1 = LV(K, k + 1);
    double a, b, c;
    tracer_row(k, nz, dt, use_harmonic, h_km1, h_k, h_kp1, K_k, K_kp1, a, b, c);
    double piv;
    if (k == 0) {
      cprev = c / b;
      piv = b;
    } else {
      piv = b - a * cprev;
      cprev = c / piv;
    }
    LV(s.fc, k) = cprev;
#pragma unroll
    for (int t = 0; t < G; ++t) {
      const double x = LV(xs[t], k);
      if (bud[t]) LV(pold[t], k) = h_k * x;
      r[t] = (k == 0) ? x / piv : (x - a * r[t]) / piv;
      LV(xs[t], k) = r[t];
    }
    h_km1 = h_k;
    h_k = h_kp1;
    K_k = K_kp1;
  }
  const double h_top = LV(s.h, nz - 1);
#pragma unroll
  for (int t = 0; t < G; ++t)
    if (bud[t]) LV(bud[t], nz - 1) = LV(bud[t], nz - 1) + (h_top * r[t] - LV(pold[t], nz - 1));
  for (int k = nz - 2; k >= kb; --k) {
    const double cp = LV(s.fc, k);
    const double hk = LV(s.h, k);
#pragma unroll
    for (int t = 0; t < G; ++t) {
      r[t] = LV(xs[t], k) - cp * r[t];
      LV(xs[t], k) = r[t];
      if (bud[t]) LV(bud[t], k) = LV(bud[t], k) + (hk * r[t] - LV(pold[t], k));
    }
  }
}

// A11 two-pass dispatch: kt -> T; ks -> S and passives (rdb_ocean_vdiff.F90:
// 1107-1226), the ks tracers in groups of up to four.  from_kb: start at the
// first live layer s.kb[p] (current only after the static refresh).
__global__ void k_vdiff(StateView s, Params pr, double dt, int p0, int p1, int from_kb) {
  int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const size_t L = (size_t)s.nz * (size_t)s.P;
  const int kb = from_kb ? min(s.kb[p], s.nz - 1) : 0;
  {
    double* const xs[1] = {s.T};
    double* const bud[1] = {s.budT};
    double* const pold[1] = {s.pold};
    solve_group<1>(s, s.kt, dt, pr.use_harmonic, p, kb, xs, bud, pold);
  }
  // S first, then the passives, four tracers a pass
  const int n = 1 + s.n_passive;
  for (int g0 = 0; g0 < n; g0 += 4) {
    auto tracer = [&](int t) { return t == 0 ? s.S : s.tr + (size_t)(t - 1) * L; };
    double* const b0 = g0 == 0 ? s.budS : nullptr;
    double* const q0 = s.pold ? s.pold + L : nullptr;
    switch (n - g0 >= 4 ? 4 : n - g0) {
      case 1: {
        double* const xs[1] = {tracer(g0)};
        double* const bud[1] = {b0};
        double* const pold[1] = {q0};
        solve_group<1>(s, s.ks, dt, pr.use_harmonic, p, kb, xs, bud, pold);
        break;
      }
      case 2: {
        double* const xs[2] = {tracer(g0), tracer(g0 + 1)};
        double* const bud[2] = {b0, nullptr};
        double* const pold[2] = {q0, nullptr};
        solve_group<2>(s, s.ks, dt, pr.use_harmonic, p, kb, xs, bud, pold);
        break;
      }
      case 3: {
        double* const xs[3] = {tracer(g0), tracer(g0 + 1), tracer(g0 + 2)};
        double* const bud[3] = {b0, nullptr, nullptr};
        double* const pold[3] = {q0, nullptr, nullptr};
        solve_group<3>(s, s.ks, dt, pr.use_harmonic, p, kb, xs, bud, pold);
        break;
      }
      default: {
        double* const xs[4] = {tracer(g0), tracer(g0 + 1), tracer(g0 + 2), tracer(g0 + 3)};
        double* const bud[4] = {b0, nullptr, nullptr, nullptr};
        double* const pold[4] = {q0, nullptr, nullptr, nullptr};
        solve_group<4>(s, s.ks, dt, pr.use_harmonic, p, kb, xs, bud, pold);
        break;
      }
    }
  }
}

// A14 non-local tendencies on T (gamma_t) and S (gamma_s)
// (rdb_ocean_vmix.F90:1652-1691).
__global__ void k_nonlocal(StateView s, double dt, int p0, int p1) {
  int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (p >= p1 || k >= s.nz) return;
  const int P = s.P, nz = s.nz;
  double hk = LV(s.h, k);
  double hkp1 = (k < nz - 1) ? LV(s.h, k + 1) : 0.0;
  double hkm1 = (k > 0) ? LV(s.h, k - 1) : 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    const double* g = pass == 0 ? s.gamma_t : s.gamma_s;
    double* x = pass == 0 ? s.T : s.S;
    double f_up = 0.0, f_dn = 0.0;
    if (k < nz - 1) f_up = nl_face(LV(g, k + 1), hk, hkp1);
    if (k > 0) f_dn = nl_face(LV(g, k), hkm1, hk);
    LV(x, k) = conc_add_source(LV(x, k), hk, dt * (f_up - f_dn));
  }
}

// ---------------------------------------------------------------------------
// Column preparation (run whenever h, u, v, tau or Q changed): per column the
// first live layer from the bed, kb (= nz when no layer is live), and the
// diffusivities of the bed-filler interfaces 1..kb-1.  A filler (h <=
// H_VANISHED) has rho = rho0 from the EOS, so N^2 = 0 across an interface
// between two fillers and PP81 gives the same kv/kt every step
// (rdb_ocean_vmix.F90:1050-1072 with Ri = 0): those interfaces are written
// here once; the step kernel only revisits them if the KPP boundary layer
// reaches them.  Interfaces 0 (bed) and nz (surface) are the closed-boundary
// zeros (rdb_ocean_vmix.F90:1031-1036).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void filler_iface(const StateView& s, const Params& pr, int p, int f, double& kv,
                                             double& kt) {
  const int P = s.P;
  pp81_iface<double>(pr, LV(s.ub, f - 1), LV(s.vb, f - 1), LV(s.ub, f), LV(s.vb, f), LV(s.h, f - 1), LV(s.h, f),
                     pr.rho0 - pr.rho0, kv, kt);
}

__global__ void k_prepare(StateView s, Params pr, int p0, int p1) {
  int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const int P = s.P, nz = s.nz;
  int kb = nz;
  for (int k = 0; k < nz; ++k)
    if (LV(s.h, k) > H_VANISHED) { kb = k; break; }
  s.kb[p] = kb;
  s.fdirty[p] = nz + 1;
  LV(s.kv, 0) = 0.0;
  LV(s.kt, 0) = 0.0;
  LV(s.ks, 0) = 0.0;
  LV(s.kv, nz) = 0.0;
  LV(s.kt, nz) = 0.0;
  LV(s.ks, nz) = 0.0;
  const int fmax = (kb < nz) ? kb - 1 : nz - 1;
  for (int f = 1; f <= fmax; ++f) {
    double kv, kt;
    filler_iface(s, pr, p, f, kv, kt);
    double ks = kt;
    LV(s.kv, f) = fmn(fmx(kv, pr.kv_bg), pr.kv_max);
    LV(s.kt, f) = fmn(fmx(kt, pr.kt_bg), pr.kd_max);
    LV(s.ks, f) = fmn(fmx(ks, pr.ks_bg), pr.kd_max);
  }
}

// ---------------------------------------------------------------------------
// Fused column step: A1-A14 for one column per lane in the order
// EOS -> PP81 -> KPP -> split -> assemble -> surface flux -> Thomas -> non-local
// (apply_deferred_thermo, rdb_ocean_dyn.F90:3836-3856).  Three sweeps:
//   A  (surface -> down, boundary layer only): KPP bulk-Ri search with the
//      EOS evaluated on the fly, then the shape-function values K_k of the
//      boundary-layer interfaces, parked in kv[] until sweep B reads them;
//   B  (live bed -> surface): EOS + PP81 + KPP max + split + clip at every
//      live interface (kv/kt/ks written once), the tracer row, the Thomas
//      elimination, the surface deposit folded into the right-hand side of
//      the first live layer.  The pivot-scaled super-diagonal c' and the
//      forward residuals r of the lowest M live levels stay in LDS; higher
//      levels spill (c' to scratch, r in place in T/S/tracers);
//   C  (surface -> live bed): back substitution and the non-local tendency,
//      gamma recomputed from the same top-down running depth as KPP pass 2.
// Bed fillers (k < kb) are identity rows of the reference's tridiagonal
// (alpha = beta = 0, b = 1), refuse deposits, and close their non-local
// faces; the reference's elimination through them leaves every filler value
// and the first live row's pivot, c' and r unchanged, so starting the
// elimination at kb gives the same bits.  Every other value comes from the
// same floating-point operations in the same order as the staged kernels and
// the reference.  TWO_FACTOR is needed only when kt_bg != ks_bg; otherwise
// one factorisation of kt serves every tracer (the reference repeats the
// identical arithmetic on its bitwise copy ks = kt).
// ---------------------------------------------------------------------------
template <typename R, int NTR, bool TWO_FACTOR, bool SURR>
__device__ __forceinline__ void column_step(const StateView& s, const Params& pr, double dt, int p,
                                            unsigned int* guard_count, int M, const SurrView& sv, double* lds) {
  const int P = s.P, nz = s.nz;
  const double HV = H_VANISHED;
  const R rho0 = R(pr.rho0);
  const int kb = s.kb[p];
  // learned closure: this column takes the network's kv/kt and h_b unless
  // its guard tripped (then the plain PP81 + KPP path below runs for it)
  const bool use_s = SURR && sv.guard[p] == 0;
  // (a network extrapolating far outside its training targets predicted
  // kt of 1e18 m^2/s on one column of a large ensemble, whose Thomas pivots
  // then cancelled to zero; the clamp keeps every prediction in the range
  // the network was trained on)
  auto pred = [&](int n) -> double {
    const float v = sv.Yt[(size_t)n * (size_t)sv.ldy + (size_t)p] * sv.stdv[n] + sv.mean[n];
    return (double)fmn(fmx(v, sv.tmin[n]), sv.tmax[n]);
  };
  const int B = blockDim.x, tid = threadIdx.x;
  // LDS slot of quantity q (c't, rT, rS, r of each passive, c's) at live
  // level m of this lane
#define LSLOT(q, m) lds[((size_t)(q) * (size_t)M + (size_t)(m)) * (size_t)B + (size_t)tid]

  // ---------------- sweep A: KPP pass 1 + pass 2 values -------------------
  int k_sfc = nz - 1;
  for (int k = nz - 1; k >= kb; --k)
    if (LV(s.h, k) > HV) { k_sfc = k; break; }

  const double tau_mag = s.tau_tot[p];
  const double qh = s.Qh[p], qsv = s.Qs[p];
  int k_lo = nz + 1;  // lowest interface holding a KPP value (none: > nz)
  R qT_r = R(0), qS_r = R(0), h_b_r = R(0);
  bool destab = false, kpp_on = false;
  if (use_s) {
    // predicted boundary-layer depth; B0, w_s and the gamma range follow the
    // KPP pass-2 formulas from it (rdb_ocean_vmix.F90:1428-1515)
    R u_star = sqrt(R(tau_mag) / rho0);
    qT_r = R(qh) / (rho0 * R(pr.cp));
    qS_r = R(qsv) / rho0;
    R B0 = kpp_b0<R>(pr, qT_r, qS_r);
    R h_b = fmx(R(pred(2 * (nz + sv.lev_off - 1))), R(0));
    s.b0[p] = (double)B0;
    destab = (B0 < R(0));
    R w_star = cbrt_kahan(fmx(R(0), -B0) * h_b);
    R w_s = sqrt(u_star * u_star + w_star * w_star);
    kpp_on = pr.use_kpp && (h_b > R(0) && w_s > R(0));
    h_b_r = h_b;
    if (kpp_on) {
      // the KPP shape-function profile of the predicted depth, parked for
      // sweep B's max with the predicted kv/kt exactly as on the plain path
      R d_running = R(0);
      for (int k = k_sfc; k >= 1; --k) {
        d_running = d_running + R(LV(s.h, k));
        if (!(d_running < h_b)) break;
        R sigma = d_running / h_b;
        R g_shape = sigma * (R(1) - sigma) * (R(1) - sigma);
        R kv_kpp = h_b * w_s * g_shape;
        k_lo = k;
        if (k >= kb) {
          LV(s.kv, k) = (double)kv_kpp;
        } else {
          double kv, kt;
          filler_iface(s, pr, p, k, kv, kt);
          double kk = (double)kv_kpp;
          kv = fmx(kv, kk);
          kt = fmx(kt, kk);
          double ks = kt;
          LV(s.kv, k) = fmn(fmx(kv, pr.kv_bg), pr.kv_max);
          LV(s.kt, k) = fmn(fmx(kt, pr.kt_bg), pr.kd_max);
          LV(s.ks, k) = fmn(fmx(ks, pr.ks_bg), pr.kd_max);
        }
      }
    }
    s.bl[p] = (double)h_b;
    s.kbl[p] = -1;  // no bulk-Richardson crossing was searched
  } else if (pr.use_kpp) {
    const R u_ref = R(LV(s.ub, k_sfc)), v_ref = R(LV(s.vb, k_sfc));
    double h_sfc = LV(s.h, k_sfc);
    // EOS and the buoyancy difference stay FP64 in every precision mode
    const double rho_sfc = eos_wright<double>(LV(s.T, k_sfc), LV(s.S, k_sfc), h_sfc, pr.rho0, pr.p_ref);
    const double b_ref = -(GRAVITY * rho_sfc / pr.rho0);
    const R d_centre_ref = R(0.5) * R(h_sfc);
    R u_star = sqrt(R(tau_mag) / rho0);
    qT_r = R(qh) / (rho0 * R(pr.cp));
    qS_r = R(qsv) / rho0;
    R B0 = kpp_b0<R>(pr, qT_r, qS_r);
    R wstar3_lagged = fmx(R(0), -B0) * R(s.bl[p]);
    R cb = cbrt_kahan(wstar3_lagged);
    R w_s_col = sqrt(u_star * u_star + cb * cb);
    const R vt2_coef = R(pr.c_vt2) / R(pr.ri_crit);
    R ri_prev = R(0), d_prev = d_centre_ref, h_b = R(0), d_centre_k = d_centre_ref;
    R h_up = R(h_sfc);
    int kcross = 0;
    // the levels below the surface layer come through a ring of PFA
    // register slots (slot j: level k_sfc - 1 - j - PFA n), loaded without a
    // branch on h so that the next levels are in flight while one is
    // tested; a level below the bed re-reads level 0, unused
    struct LevA {
      double h, ub, vb, T, S;
    };
    auto load_a = [&](int k, LevA& q) {
      const int kk = max(k, 0);
      q.h = LV(s.h, kk);
      q.ub = LV(s.ub, kk);
      q.vb = LV(s.vb, kk);
      q.T = LV(s.T, kk);
      q.S = LV(s.S, kk);
    };
    LevA aring[PFA];
#pragma unroll
    for (int j = 0; j < PFA; ++j) load_a(k_sfc - 1 - j, aring[j]);
    bool crossed = false;
    auto level_a = [&](int k, LevA& q) {
      double hk_d = q.h;
      R hk = R(hk_d);
      d_centre_k = d_centre_k + R(0.5) * (h_up + hk);
      h_up = hk;
      R u_k = R(q.ub), v_k = R(q.vb);
      double rho_k = (hk_d > HV) ? eos_wright<double>(q.T, q.S, hk_d, pr.rho0, pr.p_ref) : pr.rho0;
      load_a(k - PFA, q);
      double b_k = -(GRAVITY * rho_k / pr.rho0);
      R delta_d = d_centre_k - d_centre_ref;
      R delta_b = R(b_ref - b_k);
      R n_brunt = sqrt(fmx(R(0), delta_b / delta_d));
      R vt2 = vt2_coef * d_centre_k * n_brunt * w_s_col;
      R du = u_k - u_ref, dv = v_k - v_ref;
      R shear2 = fmx(du * du + dv * dv + vt2, R(pr.shear2_floor));
      R ri_bulk = delta_b * delta_d / shear2;
      if (ri_bulk >= R(pr.ri_crit)) {
        R denom = ri_bulk - ri_prev;
        R frac = (fabs(denom) > R(1.0e-12)) ? (R(pr.ri_crit) - ri_prev) / denom : R(0.5);
        h_b = d_prev + frac * (d_centre_k - d_prev);
        kcross = k + 1;
        crossed = true;
        return;
      }
      ri_prev = ri_bulk;
      d_prev = d_centre_k;
    };
    for (int k0 = k_sfc - 1; k0 >= 0 && !crossed; k0 -= PFA) {
#pragma unroll
      for (int j = 0; j < PFA; ++j)
        if (!crossed && k0 - j >= 0) level_a(k0 - j, aring[j]);
    }
    if (kcross == 0) {
      h_b = R(0);
      for (int k = 0; k <= k_sfc; ++k) h_b = h_b + R(LV(s.h, k));
    }
    s.b0[p] = (double)B0;
    destab = (B0 < R(0));
    R wstar3 = fmx(R(0), -B0) * h_b;
    R w_star = cbrt_kahan(wstar3);
    R w_s = sqrt(u_star * u_star + w_star * w_star);
    kpp_on = (h_b > R(0) && w_s > R(0));
    h_b_r = h_b;
    if (kpp_on) {
      R d_running = R(0);
      for (int k = k_sfc; k >= 1; --k) {
        d_running = d_running + R(LV(s.h, k));
        if (!(d_running < h_b)) break;
        R sigma = d_running / h_b;
        R g_shape = sigma * (R(1) - sigma) * (R(1) - sigma);
        R kv_kpp = h_b * w_s * g_shape;
        k_lo = k;
        if (k >= kb) {
          LV(s.kv, k) = (double)kv_kpp;  // parked until sweep B
        } else {
          // bed-filler interface inside the boundary layer (rare): final value
          double kv, kt;
          filler_iface(s, pr, p, k, kv, kt);
          double kk = (double)kv_kpp;
          kv = fmx(kv, kk);
          kt = fmx(kt, kk);
          double ks = kt;
          LV(s.kv, k) = fmn(fmx(kv, pr.kv_bg), pr.kv_max);
          LV(s.kt, k) = fmn(fmx(kt, pr.kt_bg), pr.kd_max);
          LV(s.ks, k) = fmn(fmx(ks, pr.ks_bg), pr.kd_max);
        }
      }
    }
    s.bl[p] = (double)h_b;
    s.kbl[p] = kcross;
  }
  {
    // filler interfaces the boundary layer reached last step but not now get
    // their plain value back
    const int fd = s.fdirty[p];
    const int k_lo_fill = k_lo;
    const int top = min(kb - 1, k_lo_fill - 1);
    for (int f = max(fd, 1); f <= top; ++f) {
      double kv, kt;
      filler_iface(s, pr, p, f, kv, kt);
      double ks = kt;
      LV(s.kv, f) = fmn(fmx(kv, pr.kv_bg), pr.kv_max);
      LV(s.kt, f) = fmn(fmx(kt, pr.kt_bg), pr.kd_max);
      LV(s.ks, f) = fmn(fmx(ks, pr.ks_bg), pr.kd_max);
    }
    int nd = (k_lo_fill < kb) ? k_lo_fill : nz + 1;
    if (nd != fd) s.fdirty[p] = nd;
  }
  if (kb >= nz) return;  // no live layer: nothing to mix or diffuse

  // ---------------- sweep B: coefficients + Thomas elimination ------------
  int k_fl = s.ktop[p];
  for (int k = k_fl; k >= 0; --k)
    if (LV(s.h, k) > HV) { k_fl = k; break; }
  const double wetv = s.wet[p];
  const double cellT = pr.has_heat ? dt / (pr.rho0 * pr.cp) * qh * wetv : 0.0;
  const double cellS = pr.has_salt ? dt / pr.rho0 * qsv * wetv : 0.0;

  double* const trb = s.tr;
  const size_t tstride = (size_t)nz * (size_t)P;
  unsigned int bad = 0;

  // coefficients of interface kf between layers kf-1 (below) and kf (above)
  auto iface = [&](int kf, R ub_lo, R vb_lo, R ub_hi, R vb_hi, double h_lo, double h_hi, double rho_lo,
                   double rho_hi, double& kt_out, double& ks_out) {
    double kvv, ktv;
    if (use_s) {
      // predicted log10 kt (outputs 0..nz-2) and log10 kv (nz-1..2nz-3)
      const int i_net = kf + sv.lev_off;  // the network's interface
      // 10^pred in single precision (one hardware exp2; a relative error of
      // a few 1e-8, far inside the network's own): the FP64 exp2 of two
      // predictions per interface had cost as much as the rest of the level
      ktv = (double)__builtin_amdgcn_exp2f((float)pred(i_net - 1) * 3.32192809f);
      kvv = (double)__builtin_amdgcn_exp2f((float)pred(nz + sv.lev_off - 1 + i_net - 1) * 3.32192809f);
    } else {
      R kv_r, kt_r;
      pp81_iface<R>(pr, ub_lo, vb_lo, ub_hi, vb_hi, R(h_lo), R(h_hi), rho_hi - rho_lo, kv_r, kt_r);
      kvv = (double)kv_r;
      ktv = (double)kt_r;
    }
    if (kf >= k_lo && kf <= k_sfc) {
      double kk = LV(s.kv, kf);  // KPP value parked by sweep A
      kvv = fmx(kvv, kk);
      ktv = fmx(ktv, kk);
    }
    if (pr.guard) {
      if (isnan(kvv) || isnan(ktv) || kvv < 0.0 || ktv < 0.0) ++bad;
    }
    double ksv = ktv;
    kvv = fmn(fmx(kvv, pr.kv_bg), pr.kv_max);
    ktv = fmn(fmx(ktv, pr.kt_bg), pr.kd_max);
    ksv = fmn(fmx(ksv, pr.ks_bg), pr.kd_max);
    __builtin_nontemporal_store(kvv, &LV(s.kv, kf));
    __builtin_nontemporal_store(ktv, &LV(s.kt, kf));
    __builtin_nontemporal_store(ksv, &LV(s.ks, kf));
    kt_out = ktv;
    ks_out = ksv;
  };

  // Level inputs of sweep B travel through a ring of PF register slots:
  // slot j holds level kb + j + PF n, and the loads of level k + PF are
  // issued as soon as level k has been consumed, so PF - 1 levels of loads
  // are in flight behind the serial elimination.  The loop is unrolled by
  // PF so that every slot index is a constant; a load past the surface
  // re-reads the top level (in bounds, unused).
  struct LevIn {
    double h, T, S, ub, vb;
    double tr[NTR > 0 ? NTR : 1];
  };
  auto load_lev = [&](int k, LevIn& q) {
    const int kk = min(k, nz - 1);
    q.h = __builtin_nontemporal_load(&LV(s.h, kk));
    q.T = __builtin_nontemporal_load(&LV(s.T, kk));
    q.S = __builtin_nontemporal_load(&LV(s.S, kk));
    q.ub = __builtin_nontemporal_load(&LV(s.ub, kk));
    q.vb = __builtin_nontemporal_load(&LV(s.vb, kk));
#pragma unroll
    for (int t = 0; t < NTR; ++t) q.tr[t] = __builtin_nontemporal_load(&LV(trb + (size_t)t * tstride, kk));
  };
  LevIn ring[PF];
#pragma unroll
  for (int j = 0; j < PF; ++j) load_lev(kb + j, ring[j]);

  double h0 = ring[0].h;
  double T0 = ring[0].T, S0 = ring[0].S;
  R ub0 = R(ring[0].ub), vb0 = R(ring[0].vb);
  double rho_0 = eos_wright<double>(T0, S0, h0, pr.rho0, pr.p_ref);
  double hm1 = 0.0;
  double Kt_below = 0.0, Ks_below = 0.0;  // interface kb
  if (kb >= 1) {
    hm1 = LV(s.h, kb - 1);
    iface(kb, R(LV(s.ub, kb - 1)), R(LV(s.vb, kb - 1)), ub0, vb0, hm1, h0, pr.rho0, rho_0, Kt_below, Ks_below);
  }
  double cp_t = 0.0, cp_s = 0.0;
  double rT = 0.0, rS = 0.0, rtr[NTR > 0 ? NTR : 1];
#pragma unroll
  for (int t = 0; t < NTR; ++t) rtr[t] = 0.0;

  // one level of sweep B: `cur` holds level k, `nxt` level k + 1
  auto level_b = [&](int k, LevIn& cur, const LevIn& nxt) {
    double h1 = 0.0, T1 = 0.0, S1 = 0.0;
    R ub1 = R(0), vb1 = R(0);
    double rho_1 = 0.0;
    double Kt_above = 0.0, Ks_above = 0.0;
    if (k + 1 < nz) {
      h1 = nxt.h;
      T1 = nxt.T;
      S1 = nxt.S;
      ub1 = R(nxt.ub);
      vb1 = R(nxt.vb);
      rho_1 = eos_wright<double>(T1, S1, h1, pr.rho0, pr.p_ref);
      iface(k + 1, ub0, vb0, ub1, vb1, h0, h1, rho_0, rho_1, Kt_above, Ks_above);
    }
    double a, b, c;
    tracer_row(k, nz, dt, pr.use_harmonic, hm1, h0, h1, Kt_below, Kt_above, a, b, c);
    double piv_t;
    if (k == kb) {
      cp_t = c / b;
      piv_t = b;
    } else {
      piv_t = b - a * cp_t;
      cp_t = c / piv_t;
    }
    double a_s = a, piv_s = piv_t;
    if (TWO_FACTOR) {
      double bs, cs;
      tracer_row(k, nz, dt, pr.use_harmonic, hm1, h0, h1, Ks_below, Ks_above, a_s, bs, cs);
      if (k == kb) {
        cp_s = cs / bs;
        piv_s = bs;
      } else {
        piv_s = bs - a_s * cp_s;
        cp_s = cs / piv_s;
      }
    }
    double xT = T0, xS = S0;
    if (k == k_fl) {
      if (pr.has_heat) xT = conc_add_source(xT, h0, cellT);
      if (pr.has_salt) xS = conc_add_source(xS, h0, cellS);
    }
    if (s.budT) {  // h*c_old of the tracer solve (after the surface deposit)
      LV(s.pold, k) = h0 * xT;
      LV(s.pold + tstride, k) = h0 * xS;
    }
    if (k == kb) {
      rT = xT / piv_t;
      rS = xS / piv_s;
    } else {
      rT = (xT - a * rT) / piv_t;
      rS = (xS - a_s * rS) / piv_s;
    }
#pragma unroll
    for (int t = 0; t < NTR; ++t) {
      double x = cur.tr[t];
      rtr[t] = (k == kb) ? x / piv_s : (x - a_s * rtr[t]) / piv_s;
    }
    // level k is consumed: its slot takes level k + PF
    load_lev(k + PF, cur);
    const int m = k - kb;
    if (m < M) {
      LSLOT(0, m) = cp_t;
      LSLOT(1, m) = rT;
      LSLOT(2, m) = rS;
#pragma unroll
      for (int t = 0; t < NTR; ++t) LSLOT(3 + t, m) = rtr[t];
      if (TWO_FACTOR) LSLOT(3 + NTR, m) = cp_s;
    } else {
      LV(s.fc, k) = cp_t;
      if (TWO_FACTOR) LV(s.fb, k) = cp_s;
      LV(s.T, k) = rT;
      LV(s.S, k) = rS;
#pragma unroll
      for (int t = 0; t < NTR; ++t) LV(trb + (size_t)t * tstride, k) = rtr[t];
    }
    hm1 = h0;
    h0 = h1;
    T0 = T1;
    S0 = S1;
    ub0 = ub1;
    vb0 = vb1;
    rho_0 = rho_1;
    Kt_below = Kt_above;
    Ks_below = Ks_above;
  };
  for (int k0 = kb; k0 < nz; k0 += PF) {
#pragma unroll
    for (int j = 0; j < PF; ++j)
      if (k0 + j < nz) level_b(k0 + j, ring[j], ring[(j + 1) % PF]);
  }

  // ---------------- sweep C: back substitution + non-local ----------------
  const bool nl_on = pr.use_kpp && kpp_on && destab;
  const int nl_lo = k_lo - 1;  // lowest layer touched by a non-zero gamma
  double xT_up = 0.0, xS_up = 0.0, xtr_up[NTR > 0 ? NTR : 1];
#pragma unroll
  for (int t = 0; t < NTR; ++t) xtr_up[t] = 0.0;
  double g_t_up = 0.0, g_s_up = 0.0;  // gamma at interface k+1
  R d_running = R(0);
  double h_up = 0.0;
  // the spilled Thomas state of the levels above the LDS tier comes back
  // through a ring of PF register slots, PF - 1 levels ahead of the back
  // substitution (slot j holds level nz - 1 - j - PF n; a level inside the
  // LDS tier, or below the bed, loads the lowest spilled level instead, in
  // bounds and unused)
  struct SpIn {
    double cpt, cps, T, S;
    double tr[NTR > 0 ? NTR : 1];
  };
  auto load_sp = [&](int k, SpIn& q) {
    const int kk = min(max(k, kb + M), nz - 1);
    q.cpt = LV(s.fc, kk);
    if (TWO_FACTOR) q.cps = LV(s.fb, kk);
    q.T = LV(s.T, kk);
    q.S = LV(s.S, kk);
#pragma unroll
    for (int t = 0; t < NTR; ++t) q.tr[t] = LV(trb + (size_t)t * tstride, kk);
  };
  SpIn sring[PF];
#pragma unroll
  for (int j = 0; j < PF; ++j) load_sp(nz - 1 - j, sring[j]);
  auto level_c = [&](int k, SpIn& q) {
    const int m = k - kb;
    double cpt, cps, xT, xS;
    double xt[NTR > 0 ? NTR : 1];
    if (m < M) {
      cpt = LSLOT(0, m);
      xT = LSLOT(1, m);
      xS = LSLOT(2, m);
#pragma unroll
      for (int t = 0; t < NTR; ++t) xt[t] = LSLOT(3 + t, m);
      cps = TWO_FACTOR ? LSLOT(3 + NTR, m) : cpt;
    } else {
      cpt = q.cpt;
      cps = TWO_FACTOR ? q.cps : cpt;
      xT = q.T;
      xS = q.S;
#pragma unroll
      for (int t = 0; t < NTR; ++t) xt[t] = q.tr[t];
    }
    // level k is consumed: its slot takes level k - PF
    load_sp(k - PF, q);
    if (k < nz - 1) {
      xT = xT - cpt * xT_up;
      xS = xS - cps * xS_up;
#pragma unroll
      for (int t = 0; t < NTR; ++t) xt[t] = xt[t] - cps * xtr_up[t];
    }
#pragma unroll
    for (int t = 0; t < NTR; ++t) {
      xtr_up[t] = xt[t];
      __builtin_nontemporal_store(xt[t], &LV(trb + (size_t)t * tstride, k));
    }
    xT_up = xT;
    xS_up = xS;
    if (s.budT) {  // budget += h*c_new - h*c_old, before the non-local tendency
      const double hb = LV(s.h, k);
      LV(s.budT, k) = LV(s.budT, k) + (hb * xT - LV(s.pold, k));
      LV(s.budS, k) = LV(s.budS, k) + (hb * xS - LV(s.pold + tstride, k));
    }
    if (nl_on && k <= k_sfc && k >= nl_lo) {
      // gamma at interface k (bottom of layer k), KPP pass 2 arithmetic
      double hk = LV(s.h, k);
      double g_t_dn = 0.0, g_s_dn = 0.0;
      d_running = d_running + R(hk);
      if (k >= 1 && d_running < h_b_r) {
        R sigma = d_running / h_b_r;
        R g_shape = sigma * (R(1) - sigma) * (R(1) - sigma);
        R gf = R(pr.cs_nonlocal) * g_shape;
        g_t_dn = (double)(gf * qT_r);
        g_s_dn = (double)(gf * qS_r);
      }
      double f_up_t = 0.0, f_dn_t = 0.0, f_up_s = 0.0, f_dn_s = 0.0;
      if (k < nz - 1) {
        f_up_t = nl_face(g_t_up, hk, h_up);
        f_up_s = nl_face(g_s_up, hk, h_up);
      }
      if (k > 0) {
        double hdn = LV(s.h, k - 1);
        f_dn_t = nl_face(g_t_dn, hdn, hk);
        f_dn_s = nl_face(g_s_dn, hdn, hk);
      }
      xT = conc_add_source(xT, hk, dt * (f_up_t - f_dn_t));
      xS = conc_add_source(xS, hk, dt * (f_up_s - f_dn_s));
      g_t_up = g_t_dn;
      g_s_up = g_s_dn;
      h_up = hk;
    }
    __builtin_nontemporal_store(xT, &LV(s.T, k));
    __builtin_nontemporal_store(xS, &LV(s.S, k));
  };
  for (int k0 = nz - 1; k0 >= kb; k0 -= PF) {
#pragma unroll
    for (int j = 0; j < PF; ++j)
      if (k0 - j >= kb) level_c(k0 - j, sring[j]);
  }
  if (pr.guard && bad) atomicAdd(guard_count, bad);
#undef LSLOT
}

template <typename R, int NTR, bool TWO_FACTOR, bool SURR>
__global__ void __launch_bounds__(256) k_column_step(StateView s, Params pr, double dt, int p0, int p1,
                                                     unsigned int* guard_count, int M, SurrView sv) {
  extern __shared__ double lds[];
  const int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p < p1) column_step<R, NTR, TWO_FACTOR, SURR>(s, pr, dt, p, guard_count, M, sv, lds);
}

// The ensemble form: one persistent launch over the wet columns of several
// members (each with its own grid, nz and packed layout).  Work item w is the
// column chunk [ck.y, ck.y + blockDim.x) of member ck.x; workgroups take
// items w = blockIdx.x, blockIdx.x + gridDim.x, ... so every workgroup
// leaves after a bounded number of items.
template <typename R, int NTR, bool TWO_FACTOR, bool SURR>
__global__ void __launch_bounds__(256) k_column_step_ens(const StateView* __restrict__ views,
                                                         const SurrView* __restrict__ svs, const int* __restrict__ n_wet,
                                                         Params pr, double dt, const int2* __restrict__ work,
                                                         int n_work, unsigned int* guard_count, int M) {
  extern __shared__ double lds[];
  for (int w = blockIdx.x; w < n_work; w += gridDim.x) {
    const int2 ck = work[w];
    const int p = ck.y + threadIdx.x;
    if (p < n_wet[ck.x]) {
      const StateView s = views[ck.x];
      const SurrView sv = SURR ? svs[ck.x] : SurrView{};
      column_step<R, NTR, TWO_FACTOR, SURR>(s, pr, dt, p, guard_count, M, sv, lds);
    }
  }
}

// ---------------------------------------------------------------------------
// A17 order-invariant integrals: each h*c product is rounded to a multiple of
// 2^-44 and split into a signed high word and an unsigned low 32-bit word;
// both are summed as 64-bit integers, which is associative, so the total does
// not depend on the reduction order (the idea of Hallberg & Adcroft 2014).
// ---------------------------------------------------------------------------
__global__ void k_totals(StateView s, int p0, int p1, unsigned long long* acc) {
  __shared__ long long sh[4][256];
  int tid = threadIdx.x;
  long long hiT = 0, loT = 0, hiS = 0, loS = 0;
  const int P = s.P;
  for (int p = p0 + blockIdx.x * blockDim.x + tid; p < p1; p += gridDim.x * blockDim.x) {
    for (int k = 0; k < s.nz; ++k) {
      double hk = LV(s.h, k);
      double vt = hk * LV(s.T, k);
      double vs = hk * LV(s.S, k);
      long long it = __double2ll_rn(vt * 17592186044416.0);  // 2^44
      long long is = __double2ll_rn(vs * 17592186044416.0);
      hiT += it >> 32;
      loT += (long long)(it & 0xffffffffll);
      hiS += is >> 32;
      loS += (long long)(is & 0xffffffffll);
    }
  }
  sh[0][tid] = hiT;
  sh[1][tid] = loT;
  sh[2][tid] = hiS;
  sh[3][tid] = loS;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (tid < w)
      for (int q = 0; q < 4; ++q) sh[q][tid] += sh[q][tid + w];
    __syncthreads();
  }
  if (tid == 0)
    for (int q = 0; q < 4; ++q) atomicAdd(&acc[q], (unsigned long long)sh[q][0]);
}

// Explicit instantiations of the fused step.
#define INST(R, N, TF, SU)                                                                                      \
  template __global__ void k_column_step<R, N, TF, SU>(StateView, Params, double, int, int, unsigned int*, int, \
                                                       SurrView);                                               \
  template __global__ void k_column_step_ens<R, N, TF, SU>(const StateView*, const SurrView*, const int*, Params, \
                                                           double, const int2*, int, unsigned int*, int);
INST(double, 0, false, false) INST(double, 1, false, false) INST(double, 2, false, false)
INST(double, 3, false, false) INST(double, 4, false, false)
INST(double, 0, true, false) INST(double, 1, true, false) INST(double, 2, true, false)
INST(double, 3, true, false) INST(double, 4, true, false)
INST(float, 0, false, false) INST(float, 1, false, false) INST(float, 2, false, false)
INST(float, 3, false, false) INST(float, 4, false, false)
INST(float, 0, true, false) INST(float, 1, true, false) INST(float, 2, true, false)
INST(float, 3, true, false) INST(float, 4, true, false)
INST(double, 0, false, true) INST(double, 1, false, true) INST(double, 2, false, true)
INST(double, 3, false, true) INST(double, 4, false, true)
INST(double, 0, true, true) INST(double, 1, true, true) INST(double, 2, true, true)
INST(double, 3, true, true) INST(double, 4, true, true)
#undef INST

// Dynamic LDS above 64 KiB has to be allowed per kernel: allow every
// fused-step instantiation the device's whole per-workgroup LDS.
hipError_t allow_full_lds(int lds_bytes) {
  hipError_t e = hipSuccess;
#define ALLOW(R, N, TF, SU)                                                                                  \
  if (e == hipSuccess)                                                                                       \
    e = hipFuncSetAttribute((const void*)k_column_step<R, N, TF, SU>,                                        \
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);                          \
  if (e == hipSuccess)                                                                                       \
    e = hipFuncSetAttribute((const void*)k_column_step_ens<R, N, TF, SU>,                                    \
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
#define ALLOW_N(R, TF, SU) ALLOW(R, 0, TF, SU) ALLOW(R, 1, TF, SU) ALLOW(R, 2, TF, SU) ALLOW(R, 3, TF, SU) \
  ALLOW(R, 4, TF, SU)
  ALLOW_N(double, false, false) ALLOW_N(double, true, false) ALLOW_N(float, false, false)
  ALLOW_N(float, true, false) ALLOW_N(double, false, true) ALLOW_N(double, true, true)
#undef ALLOW_N
#undef ALLOW
  return e;
}

}  // namespace colphys
