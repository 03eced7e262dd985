// surrogate.hip -- the learned closure of the column engine: training data
// from the plain path, feature and guard kernels, normalisation, the Adam
// training loop, guarded inference feeding the FP64 solve, save/load.
//
// Features per wet column (5 nz + 6): five per level, then |tau| (+ shelf
// stress), Q_heat, Q_salt, Coriolis f, the convective velocity scale of the
// lagged KPP boundary layer, w* = (max(0, -B0) h_lagged)^(1/3) (B0 the
// surface buoyancy flux), and the live column depth.  The lagged depth
// enters KPP only through w*, and only under cooling; given to the network
// as a raw input it became a persistence shortcut (h_b ~ h_lagged): in a
// surrogate run the network's own deep guesses fed back step after step
// while the plain boundary layer shoaled, and the worst 0.1 % of the
// columns -- all with h_b two to five times the plain one -- held a third
// of the 24-step temperature error.  The five level slots hold the column state
// (T, S, centred u, centred v, h) recoded against the live level above:
//   buoyancy jump  slog(g (rho_k - rho_above) / rho0, 1e-7 m/s^2)  [T slot]
//   S_k - S_above                                                  [S slot]
//   slog(u_k - u_above, 1e-3 m/s), slog(v_k - v_above, 1e-3 m/s)   [u, v slots]
//   log10 h_k                                                      [h slot]
// with slog(x, x0) = sign(x) log10(1 + |x| / x0) and rho from the FP64
// Wright EOS; a live level with no live level above (the surface) holds its
// T, S, u, v; a vanished level holds 0.  Given the surface values and the
// EOS the recoding is invertible, so it carries the same information as the
// raw profile, but (a) level-to-level contrasts survive the BF16 rounding of
// the normalised inputs (raw T would bury a 0.1 K step under the 8-bit
// mantissa of a 25 K profile) and (b) log Ri = log N^2 - log S^2 becomes
// nearly a linear combination of the inputs, leaving PP81 a function of one
// variable for the network to learn.
// Targets (2 (nz-1) + 1): log10 kt and log10 kv at the interior interfaces
// 1..nz-1 after the plain coefficient build, and the new boundary-layer
// depth.  Interfaces below the first live layer carry weight 0 (they keep
// their plain filler value at inference).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "colphys_handle.h"
#include "mlp_net.h"

namespace colphys {

using mlp::bf16;

struct SurrState {
  int nz = 0, din = 0, dout = 0;
  mlp::Net net;
  bool have_net = false;
  std::vector<int> widths;
  // raw dataset on the device
  float* Xraw = nullptr;
  float* Traw = nullptr;
  float* Wraw = nullptr;
  long long n = 0, cap = 0;
  // normalised training set
  bf16* Xn = nullptr;
  float* Tn = nullptr;
  float* Wn = nullptr;
  unsigned* idx = nullptr;
  // statistics
  std::vector<float> fmean, fstd, fmin, fmax, tmean, tstd, tmin, tmax;
  float* d_stats = nullptr;  // fmean | fstd | fmin | fmax | tmean | tstd | tmin | tmax
  // inference
  unsigned char* guard = nullptr;
  unsigned int* nfall = nullptr;
  int guard_mode = RDB_COLPHYS_GUARD_STRICT;
  int inf_rows = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};  // around the forward GEMMs (timing on)
  // a surrogate step leaves its fallback count (and, timing on, its event
  // pair) on the device; surr_settle reads them when they are asked for
  bool pending = false;
  bool timed = false;
  std::vector<SurrView> svs_up;  // the member views last uploaded (ensembles)
  // ensembles: the feature kernel's member table and work list (one launch)
  struct FeatMemberHost {
    const double* cor;
    long long row_off;
    int n, lev_off;
  };
  void* d_fmem = nullptr;
  int2* d_fwork = nullptr;
  int n_fwork = 0;
  // the captured surrogate step: rebuilt when dt, the step count, the launch
  // shape or any buffer its launches name (prep_gen) changes
  long long prep_gen = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double g_dt = 0.0;
  int g_n = 0, g_block = 0, g_lds = -1;
  long long g_prep = -1;
  ~SurrState();
  void free_dataset();
};

namespace {

#define LV(arr, k) (arr)[(size_t)(k) * (size_t)P + (size_t)p]

int perr(int st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return rdb_colphys_push_error(st, buf);
}

#define SG_TRY(expr)                                                                                \
  do {                                                                                              \
    hipError_t _e = (expr);                                                                         \
    if (_e != hipSuccess) return perr(RDB_COLPHYS_ERR_SETUP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

__device__ __forceinline__ bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

// One level of a column, as the features see it.
struct LevelVals {
  double h, T, S, U, V, rho;
  bool live;
};

// level k of packed column p; levels outside [0, nz) do not exist in this
// grid and read as a bed filler (h = 1e-4 m, not live).
// (split in the memory part, load_raw, and the arithmetic, finish_level,
// so that a walk can have the next levels' loads in flight while it
// computes the current one; load_raw always loads -- from the nearest level
// that exists -- and finish_level substitutes the filler, so that the loads
// never sit behind a branch, which made the compiler wait for every load in
// flight; the loads and the EOS do not depend on liveness, and a vanished
// level's values are never used)
// kfirst: the column's first live level; a bed filler below it is never
// live, so its T, S, u, v are never used and are read from level kfirst
// instead (lines the walk reads anyway: cache hits, not HBM traffic)
__device__ __forceinline__ LevelVals load_raw(const StateView& s, int p, int k, int kfirst = 0) {
  const int P = s.P;
  const int kc = min(max(k, 0), s.nz - 1);
  const int kv = max(kc, min(kfirst, s.nz - 1));
  LevelVals v;
  v.h = LV(s.h, kc);
  v.T = LV(s.T, kv);
  v.S = LV(s.S, kv);
  v.U = LV(s.ub, kv);
  v.V = LV(s.vb, kv);
  v.rho = 0.0;
  v.live = k >= 0 && k < s.nz;  // "exists"; finish_level turns it into "live"
  return v;
}
__device__ __forceinline__ LevelVals finish_level(LevelVals v, double rho0, double p_ref) {
  if (!v.live) return LevelVals{1.0e-4, 0.0, 0.0, 0.0, 0.0, 0.0, false};  // a level this grid does not have
  v.live = v.h > H_VANISHED;
  v.rho = eos_wright<double>(v.T, v.S, v.h, rho0, p_ref);
  return v;
}
__device__ __forceinline__ LevelVals load_level(const StateView& s, int p, int k, double rho0, double p_ref,
                                                int kfirst = 0) {
  return finish_level(load_raw(s, p, k, kfirst), rho0, p_ref);
}

// sign(x) log10(1 + |x| / x0), in single precision: the argument is formed
// in FP64 (a difference that would cancel in FP32) and the result is a
// network input, rounded to BF16 anyway
__device__ __forceinline__ float slog(double x, float inv_x0) {
  return copysignf(log10f(1.0f + fabsf((float)x) * inv_x0), (float)x);
}

// the five level slots of `cur` under the level `above`
__device__ __forceinline__ void level_slots(const LevelVals& cur, const LevelVals& above, double rho0, float out[5]) {
  out[4] = log10f(fmaxf((float)cur.h, 1e-30f));
  if (!cur.live) {
    out[0] = out[1] = out[2] = out[3] = 0.f;
  } else if (!above.live) {
    out[0] = (float)cur.T;
    out[1] = (float)cur.S;
    out[2] = (float)cur.U;
    out[3] = (float)cur.V;
  } else {
    out[0] = slog(GRAVITY * (cur.rho - above.rho) / rho0, 1.0e7f);
    out[1] = (float)(cur.S - above.S);
    out[2] = slog(cur.U - above.U, 1.0e3f);
    out[3] = slog(cur.V - above.V, 1.0e3f);
  }
}

// forcing feature i (0..5) of packed column p; the surface buoyancy flux is
// B0 = b0T Q_heat - b0S Q_salt (kpp_b0 of the fluxes)
__device__ __forceinline__ float forcing_feature(const StateView& s, const double* cor, int p, int i, double depth,
                                                 double b0T, double b0S) {
  switch (i) {
    case 0: return (float)s.tau_tot[p];
    case 1: return (float)s.Qh[p];
    case 2: return (float)s.Qs[p];
    case 3: return (float)cor[p];
    case 4: return (float)cbrt(fmax(0.0, b0S * s.Qs[p] - b0T * s.Qh[p]) * s.bl[p]);
    default: return (float)depth;
  }
}


// The network's levels are walked in FW groups of per = ceil(nzn / FW)
// (one wavefront each in k_features_infer); the live column depth is the
// sum, over the groups in order, of each group's live thicknesses summed
// from its top level down -- the same order in both feature kernels, so
// that training and inference features agree bit for bit.
constexpr int FC = 64, FW = 8, FB = 64 * FW;

// B0 = b0_T Q_heat - b0_S Q_salt: kpp_b0 with qT = Q_heat / (rho0 cp) and
// qS = Q_salt / rho0
inline double b0_T(const Params& pr) { return GRAVITY / pr.rho0 * pr.alpha_T / (pr.rho0 * pr.cp); }
inline double b0_S(const Params& pr) { return GRAVITY / pr.rho0 * pr.beta_S / pr.rho0; }

// Raw features of packed columns c = i * stride (i < ncols) into rows row0 + i.
// The network has nzn >= nz levels with the surfaces aligned: network level
// j is this grid's level j - lev_off.
__global__ void k_features_raw(StateView s, const double* cor, int ncols, int stride, int din, float* X, long long row0,
                               int nzn, int lev_off, double rho0, double p_ref, double b0T, double b0S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncols) return;
  int p = i * stride;
  float* row = X + (size_t)(row0 + i) * (size_t)din;
  LevelVals above = load_level(s, p, nzn - lev_off, rho0, p_ref);  // above the top: none
  const int per = (nzn + FW - 1) / FW;
  double dpart[FW];
  for (int w = 0; w < FW; ++w) dpart[w] = 0.0;
  for (int j = nzn - 1; j >= 0; --j) {
    const LevelVals cur = load_level(s, p, j - lev_off, rho0, p_ref);
    float o[5];
    level_slots(cur, above, rho0, o);
    for (int b = 0; b < 5; ++b) row[b * nzn + j] = o[b];
    if (cur.live) dpart[j / per] += cur.h;
    above = cur;
  }
  double depth = 0.0;
  for (int w = 0; w < FW; ++w) depth += dpart[w];
  for (int f = 5 * nzn; f < din; ++f) row[f] = forcing_feature(s, cor, p, f - 5 * nzn, depth, b0T, b0S);
}

// Targets of the same columns after the plain step: log10 kt, log10 kv at
// the network's interfaces 1..nzn-1, bl_depth.  Loss weight 1, except 0
// below this grid's first live layer kb and on interfaces the grid does not
// have.  (Weighting the kt targets by the diffusive flux they carry was
// tried and left the 24-step temperature error no better: the network's
// limit there is resolving Ri, not the weighting.)
__global__ void k_targets_raw(StateView s, int ncols, int stride, int dout, float* T, float* W, long long row0,
                              int nzn, int lev_off) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncols) return;
  const int P = s.P;
  int p = i * stride;
  const int kb = s.kb[p];
  float* t = T + (size_t)(row0 + i) * (size_t)dout;
  float* w = W + (size_t)(row0 + i) * (size_t)dout;
  for (int j = 1; j < nzn; ++j) {
    const int k = j - lev_off;  // this grid's interface (between layers k - 1 and k)
    const bool on = k >= 1 && k >= kb;
    const float wt = on ? 1.f : 0.f, wv = wt;
    t[j - 1] = k >= 1 ? (float)log10(fmax(LV(s.kt, k), 1e-30)) : 0.f;
    t[nzn - 1 + j - 1] = k >= 1 ? (float)log10(fmax(LV(s.kv, k), 1e-30)) : 0.f;
    w[j - 1] = wt;
    w[nzn - 1 + j - 1] = wv;
  }
  t[2 * (nzn - 1)] = (float)s.bl[p];
  w[2 * (nzn - 1)] = 1.f;
}

// Per-column mean/std (weighted when W is given) and min/max of an n x d
// row-major matrix; one block per column.
__global__ void k_col_stats(const float* X, const float* W, long long n, int d, float* mean, float* stdv, float* mn,
                            float* mx) {
  __shared__ double s1[256], s2[256], sw[256];
  __shared__ float smn[256], smx[256];
  const int j = blockIdx.x;
  double a = 0, b = 0, c = 0;
  float lo = 3.4e38f, hi = -3.4e38f;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    float w = W ? W[(size_t)i * d + j] : 1.f;
    if (w > 0.f) {
      double v = X[(size_t)i * d + j];
      a += w * v;
      b += w * v * v;
      c += w;
      lo = fminf(lo, (float)v);
      hi = fmaxf(hi, (float)v);
    }
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  sw[threadIdx.x] = c;
  smn[threadIdx.x] = lo;
  smx[threadIdx.x] = hi;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      s1[threadIdx.x] += s1[threadIdx.x + st];
      s2[threadIdx.x] += s2[threadIdx.x + st];
      sw[threadIdx.x] += sw[threadIdx.x + st];
      smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + st]);
      smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double cnt = sw[0] > 0 ? sw[0] : 1.0;
    double mu = s1[0] / cnt;
    double var = s2[0] / cnt - mu * mu;
    mean[j] = (float)mu;
    stdv[j] = (float)sqrt(fmax(var, 1e-12));
    if (mn) mn[j] = smn[0];
    if (mx) mx[j] = smx[0];
  }
}

// The normalisation kernels walk n x ld elements with a grid-stride loop:
// a dispatch counts its work-items in 32 bits, and a training set of 13 M
// columns x 384 inputs is 5e9 elements, so the grid is capped (gs_capped).
__global__ void k_norm_features(const float* X, long long n, int din, const float* mean, const float* stdv, bf16* Xn,
                                int ld) {
  const long long total = n * ld, stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    long long r = i / ld;
    int c = (int)(i % ld);
    Xn[i] = f2bf(c < din ? (X[(size_t)r * din + c] - mean[c]) / stdv[c] : 0.f);
  }
}

__global__ void k_norm_targets(const float* T, const float* W, long long n, int dout, const float* mean,
                               const float* stdv, float* Tn, float* Wn, int ld) {
  const long long total = n * ld, stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    long long r = i / ld;
    int c = (int)(i % ld);
    bool in = c < dout;
    Tn[i] = in ? (T[(size_t)r * dout + c] - mean[c]) / stdv[c] : 0.f;
    Wn[i] = in ? W[(size_t)r * dout + c] : 0.f;
  }
}

// Inference features of wet columns [0, n): normalised bf16 rows written
// through an LDS tile so that each row is stored with coalesced writes, and
// the per-column guard (non-zero: plain path).  Workgroup of FW waves over
// 64 columns: wave w takes an FW-th of the network levels for all 64
// columns (coalesced level-major reads), walking them top-down with the
// level above carried in registers; it also checks N^2 < 0 (strict guard)
// on the pairs whose lower level it owns.  Wave 0 adds the forcing features,
// the last wave (the surface levels) the surface-pair check; the flags meet
// in LDS.  Every value is the
// same arithmetic as feature(), so training and inference features agree
// bit for bit.
// The block of FC packed columns [p0, p0 + FC) of one grid (single grids:
// the launch's block index times FC; ensembles: one work item of a member).
__device__ __forceinline__ void features_block(int p0, const StateView& s, const double* cor, int n, int din,
                                               int ld, const float* st, bf16* X0, unsigned char* guard,
                                               unsigned int* nfall, int guard_mode, double rho0, double p_ref,
                                               int nzn, int lev_off, double b0T, double b0S) {
  __shared__ bf16 tile[384][FC + 1];
  __shared__ unsigned char flags[FW][FC];
  __shared__ double sdepth[FW][FC];
  // the normalisation statistics (mean, std, min, max per feature), staged
  // in LDS once: read from global memory inside the level walk, each of
  // their loads waited with vmcnt(0) -- which also drained the next levels'
  // loads in flight, leaving the walk latency-bound
  __shared__ float sst[4 * 384];
  const int tid = threadIdx.x, c = tid & (FC - 1), part = tid >> 6;
  for (int i = tid; i < 4 * din; i += FB) sst[i] = st[i];
  __syncthreads();
  const float* mean = sst;
  const float* stdv = sst + din;
  const float* mn = sst + 2 * din;
  const float* mx = sst + 3 * din;
  const int p = p0 + c;
  const int P = s.P, nz = s.nz;
  unsigned char fl = 0;
  auto put = [&](int f, float v) {
    if (v < mn[f] || v > mx[f]) fl |= 1;
    tile[f][c] = f2bf((v - mean[f]) / stdv[f]);
  };
  if (p < n) {
    const int per = (nzn + FW - 1) / FW;
    const int j0 = part * per, j1 = min(nzn, j0 + per);
    const int kfirst = s.kb[p];
    LevelVals above = load_level(s, p, j1 - lev_off, rho0, p_ref, kfirst);  // the level above this part's highest
    bool unstable = false;
    double dpart = 0.0;  // this group's live thickness, summed from its top down
    // the next two levels' loads stay in flight while a level is computed:
    // three named slots, the walk unrolled by three, so that a level's
    // values stay in their slot's registers (rotating them through copies
    // made the compiler wait for every load in flight at each step)
    // (the loads are unconditional -- past the group's lowest level they
    // read the next group's levels, which its own wave loads as well)
    auto step = [&](int j, const LevelVals& raw, LevelVals& ahead) {
      ahead = load_raw(s, p, j - 2 - lev_off, kfirst);
      __builtin_amdgcn_sched_barrier(0);
      if (j < j0) return;
      const LevelVals cur = finish_level(raw, rho0, p_ref);
      float o[5];
      level_slots(cur, above, rho0, o);
#pragma unroll
      for (int b = 0; b < 5; ++b) put(b * nzn + j, o[b]);
      // static stability of the pair (this level, the one above): N^2 < 0
      // where the upper of two live layers is denser (A3's sign)
      if (guard_mode == RDB_COLPHYS_GUARD_STRICT && cur.live && above.live && above.rho > cur.rho) unstable = true;
      if (cur.live) dpart += cur.h;
      above = cur;
    };
    LevelVals sa = load_raw(s, p, j1 - 1 - lev_off, kfirst);
    LevelVals sb = load_raw(s, p, j1 - 2 - lev_off, kfirst);
    LevelVals sc;
    for (int j = j1 - 1; j >= j0; j -= 3) {
      step(j, sa, sc);
      step(j - 1, sb, sa);
      step(j - 2, sc, sb);
    }
    sdepth[part][c] = dpart;
    if (part == FW - 1 && guard_mode == RDB_COLPHYS_GUARD_SURFACE) {
      const int kb = s.kb[p];
      int k_sfc = nz - 1;
      for (int k = nz - 1; k >= kb; --k)
        if (LV(s.h, k) > H_VANISHED) {
          k_sfc = k;
          break;
        }
      if (k_sfc > kb) {
        const double h1 = LV(s.h, k_sfc), h0 = LV(s.h, k_sfc - 1);
        if (h1 > H_VANISHED && h0 > H_VANISHED) {
          const double r1 = eos_wright<double>(LV(s.T, k_sfc), LV(s.S, k_sfc), h1, rho0, p_ref);
          const double r0 = eos_wright<double>(LV(s.T, k_sfc - 1), LV(s.S, k_sfc - 1), h0, rho0, p_ref);
          unstable = unstable || r1 > r0;
        }
      }
    }
    if (unstable) fl |= 2;
  }
  flags[part][c] = fl;
  __syncthreads();
  if (part == 0 && p < n) {
    // the forcing features, with the column depth from the groups' sums
    double depth = 0.0;
    for (int w = 0; w < FW; ++w) depth += sdepth[w][c];
    for (int f = 5 * nzn; f < din; ++f) put(f, forcing_feature(s, cor, p, f - 5 * nzn, depth, b0T, b0S));
    for (int f = din; f < ld; ++f) tile[f][c] = 0;
    unsigned char g = fl;  // this wave's flags, now with the forcing features' range checks
    for (int w = 1; w < FW; ++w) g |= flags[w][c];
    guard[p] = g;
    if (g) atomicAdd(nfall, 1u);
  }
  __syncthreads();  // the forcing features are in the tile
  // cooperative row stores: the FC rows of this block, 2 bf16 per lane,
  // consecutive lanes along a row
  const int rows = min(FC, n - p0);
  const int half = ld / 2;
  for (int idx = tid; idx < rows * half; idx += FB) {
    const int r = idx / half, c2 = idx - r * half;
    const uint32_t lo = tile[2 * c2][r], hi = tile[2 * c2 + 1][r];
    ((uint32_t*)(X0 + (size_t)(p0 + r) * (size_t)ld))[c2] = lo | (hi << 16);
  }
}

__global__ void __launch_bounds__(FB) k_features_infer(StateView s, const double* cor, int n, int din, int ld,
                                                     const float* st, bf16* X0, unsigned char* guard,
                                                     unsigned int* nfall, int guard_mode, double rho0, double p_ref,
                                                     int nzn, int lev_off, double b0T, double b0S) {
  features_block(blockIdx.x * FC, s, cor, n, din, ld, st, X0, guard, nfall, guard_mode, rho0, p_ref, nzn, lev_off,
                 b0T, b0S);
}

// Every member of an ensemble in one launch: block w takes the FC columns
// of work item (member, first packed column) work[w].
struct FeatMember {
  const double* cor;  // the member's Coriolis field (packed)
  long long row_off;  // the member's first row in the network batch
  int n;              // wet columns
  int lev_off;        // network levels below the member's bed
};

__global__ void __launch_bounds__(FB) k_features_infer_ens(const StateView* __restrict__ views,
                                                         const FeatMember* __restrict__ mem,
                                                         const int2* __restrict__ work, int din, int ld,
                                                         const float* st, bf16* X0, unsigned char* guard,
                                                         unsigned int* nfall, int guard_mode, double rho0,
                                                         double p_ref, int nzn, double b0T, double b0S) {
  const int2 w = work[blockIdx.x];
  const FeatMember m = mem[w.x];
  features_block(w.y, views[w.x], m.cor, m.n, din, ld, st, X0 + (size_t)m.row_off * (size_t)ld, guard + m.row_off,
                 nfall, guard_mode, rho0, p_ref, nzn, m.lev_off, b0T, b0S);
}

// the fallback counter is cleared by a kernel, so that the step can be
// captured and replayed (no memset node in the graph)
__global__ void k_clear_u32(unsigned int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

inline dim3 gs(long long n, int b) { return dim3((unsigned)((n + b - 1) / b)); }
// for grid-stride kernels: at most 2^20 workgroups (2^28 work-items at 256)
inline dim3 gs_capped(long long n, int b) { return dim3((unsigned)std::min<long long>((n + b - 1) / b, 1LL << 20)); }

}  // namespace

SurrState::~SurrState() {
  free_dataset();
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (d_fmem) (void)hipFree(d_fmem);
  if (d_fwork) (void)hipFree(d_fwork);
  if (d_stats) (void)hipFree(d_stats);
  if (guard) (void)hipFree(guard);
  if (nfall) (void)hipFree(nfall);
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
}

void SurrState::free_dataset() {
  for (void* q : {(void*)Xraw, (void*)Traw, (void*)Wraw, (void*)Xn, (void*)Tn, (void*)Wn, (void*)idx})
    if (q) (void)hipFree(q);
  Xraw = Traw = Wraw = Tn = Wn = nullptr;
  Xn = nullptr;
  idx = nullptr;
  n = cap = 0;
}

void surr_free(SurrState* s) { delete s; }

// levels of the network: the grid's nz, or the largest member nz
static int net_levels(rdb_colphys* h) { return h->members.empty() ? h->nz : h->nz_net; }
static long long wet_rows(rdb_colphys* h) { return h->members.empty() ? (long long)h->n_wet : h->ens_wet; }

static SurrState* surr_get(rdb_colphys* h) {
  if (!h->surr) {
    const int nzn = net_levels(h);
    h->surr = new SurrState();
    h->surr->nz = nzn;
    h->surr->din = 5 * nzn + 6;
    h->surr->dout = 2 * (nzn - 1) + 1;
  }
  return h->surr;
}

int surr_collect(rdb_colphys* h, double dt, int n_steps, int stride, long long* n_out) {
  const int nzn = net_levels(h);
  if (nzn < 2) return perr(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp_collect: needs nz >= 2");
  if (5 * nzn + 6 > 384) return perr(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp_collect: nz > 75 exceeds the feature tile");
  if (stride < 1) stride = 1;
  SurrState* S = surr_get(h);
  const std::vector<rdb_colphys*> ps = plans(h);
  long long ncols = 0;
  for (rdb_colphys* m : ps) ncols += (m->n_wet + stride - 1) / stride;
  const long long need = S->n + ncols * n_steps;
  if (need > S->cap) {
    // grow (copying what is already collected)
    const long long cap = std::max(need, S->cap * 2);
    float *X = nullptr, *T = nullptr, *W = nullptr;
    hipError_t ae = hipMalloc((void**)&X, 4 * (size_t)cap * S->din);
    if (ae == hipSuccess) ae = hipMalloc((void**)&T, 4 * (size_t)cap * S->dout);
    if (ae == hipSuccess) ae = hipMalloc((void**)&W, 4 * (size_t)cap * S->dout);
    if (ae != hipSuccess) {
      for (float* q : {X, T, W})
        if (q) (void)hipFree(q);
      return perr(RDB_COLPHYS_ERR_SETUP, "mlp_collect: no device memory for %lld training columns: %s", cap,
                  hipGetErrorString(ae));
    }
    if (S->n) {
      SG_TRY(hipMemcpyAsync(X, S->Xraw, 4 * (size_t)S->n * S->din, hipMemcpyDeviceToDevice, h->stream));
      SG_TRY(hipMemcpyAsync(T, S->Traw, 4 * (size_t)S->n * S->dout, hipMemcpyDeviceToDevice, h->stream));
      SG_TRY(hipMemcpyAsync(W, S->Wraw, 4 * (size_t)S->n * S->dout, hipMemcpyDeviceToDevice, h->stream));
      SG_TRY(hipStreamSynchronize(h->stream));
    }
    long long n_keep = S->n;
    S->free_dataset();
    S->Xraw = X;
    S->Traw = T;
    S->Wraw = W;
    S->cap = cap;
    S->n = n_keep;
  }
  int st = refresh_static(h, dt);
  if (st) return st;
  for (int i = 0; i < n_steps; ++i) {
    long long row = S->n;
    for (rdb_colphys* m : ps) {
      const int nc = (m->n_wet + stride - 1) / stride;
      if (nc > 0)
        hipLaunchKernelGGL(k_features_raw, gs(nc, 128), dim3(128), 0, h->stream, m->s, m->d_coriolis, nc, stride,
                           S->din, S->Xraw, row, nzn, nzn - m->nz, h->pr.rho0, h->pr.p_ref, b0_T(h->pr),
                           b0_S(h->pr));
      row += nc;
    }
    st = launch_step_all(h, dt, false);
    if (st) return st;
    row = S->n;
    for (rdb_colphys* m : ps) {
      const int nc = (m->n_wet + stride - 1) / stride;
      if (nc > 0)
        hipLaunchKernelGGL(k_targets_raw, gs(nc, 128), dim3(128), 0, h->stream, m->s, nc, stride, S->dout, S->Traw,
                           S->Wraw, row, nzn, nzn - m->nz);
      row += nc;
    }
    SG_TRY(hipGetLastError());
    S->n += ncols;
    h->steps += 1;
    h->t_sim += dt;
    ++h->state_ver;
  }
  SG_TRY(hipStreamSynchronize(h->stream));
  if (n_out) *n_out = S->n;
  return RDB_COLPHYS_OK;
}

int surr_train(rdb_colphys* h, const rdb_colphys_mlp_config* cfg, float* train_loss, float* val_loss, int n_hist,
               int* n_done) {
  SurrState* S = h->surr;
  if (!S || S->n < 64) return perr(RDB_COLPHYS_ERR_NOT_INITIALISED, "mlp_train: collect training columns first");
  if (!cfg || cfg->batch < 64 || cfg->batch > (1 << 20) || cfg->hidden < 1 || cfg->n_hidden < 1 || cfg->epochs < 0)
    return perr(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp_train: bad configuration");
  const long long N = S->n;
  const int din = S->din, dout = S->dout;
  if (N > 0xffffffffLL) return perr(RDB_COLPHYS_ERR_BAD_SHAPE, "mlp_train: more than 2^32 training columns");
  // statistics over the whole collected set
  std::vector<float> stats(4 * (size_t)din + 4 * (size_t)dout);
  float* dst = nullptr;
  SG_TRY(hipMalloc((void**)&dst, 4 * stats.size()));
  hipLaunchKernelGGL(k_col_stats, dim3(din), dim3(256), 0, h->stream, S->Xraw, (const float*)nullptr, N, din, dst,
                     dst + din, dst + 2 * din, dst + 3 * din);
  hipLaunchKernelGGL(k_col_stats, dim3(dout), dim3(256), 0, h->stream, S->Traw, S->Wraw, N, dout, dst + 4 * din,
                     dst + 4 * din + dout, dst + 4 * din + 2 * dout, dst + 4 * din + 3 * dout);
  SG_TRY(hipGetLastError());
  SG_TRY(hipMemcpyAsync(stats.data(), dst, 4 * stats.size(), hipMemcpyDeviceToHost, h->stream));
  SG_TRY(hipStreamSynchronize(h->stream));
  if (S->d_stats) (void)hipFree(S->d_stats);
  S->d_stats = dst;
  S->fmean.assign(stats.begin(), stats.begin() + din);
  S->fstd.assign(stats.begin() + din, stats.begin() + 2 * din);
  S->fmin.assign(stats.begin() + 2 * din, stats.begin() + 3 * din);
  S->fmax.assign(stats.begin() + 3 * din, stats.begin() + 4 * din);
  S->tmean.assign(stats.begin() + 4 * din, stats.begin() + 4 * din + dout);
  S->tstd.assign(stats.begin() + 4 * din + dout, stats.begin() + 4 * din + 2 * dout);
  S->tmin.assign(stats.begin() + 4 * din + 2 * dout, stats.begin() + 4 * din + 3 * dout);
  S->tmax.assign(stats.begin() + 4 * din + 3 * dout, stats.end());
  // a target no training row weights (an interface no collected column
  // has) has an empty range: its predictions are left unclamped
  bool patched = false;
  for (int n = 0; n < dout; ++n)
    if (!(S->tmin[(size_t)n] <= S->tmax[(size_t)n])) {
      S->tmin[(size_t)n] = -3.4e38f;
      S->tmax[(size_t)n] = 3.4e38f;
      patched = true;
    }
  if (patched) {
    SG_TRY(hipMemcpy(dst + 4 * din + 2 * dout, S->tmin.data(), 4 * (size_t)dout, hipMemcpyHostToDevice));
    SG_TRY(hipMemcpy(dst + 4 * din + 3 * dout, S->tmax.data(), 4 * (size_t)dout, hipMemcpyHostToDevice));
  }
  // network
  std::vector<int> widths{din};
  for (int l = 0; l < cfg->n_hidden; ++l) widths.push_back(cfg->hidden);
  widths.push_back(dout);
  if (!S->have_net || S->widths != widths) {
    int e = S->net.init(widths, h->stream);
    if (e) return perr(e, "mlp_train: %s", S->net.err.c_str());
    std::mt19937_64 rng((uint64_t)cfg->seed);
    std::vector<float> p(S->net.n_public());
    size_t q = 0;
    for (size_t l = 0; l + 1 < widths.size(); ++l) {
      const double lim = std::sqrt(6.0 / (widths[l] + widths[l + 1]));
      std::uniform_real_distribution<double> u(-lim, lim);
      for (long long i = 0; i < (long long)widths[l] * widths[l + 1]; ++i) p[q++] = (float)u(rng);
      for (int i = 0; i < widths[l + 1]; ++i) p[q++] = 0.f;
    }
    e = S->net.set_params(p.data());
    if (e) return perr(e, "mlp_train: %s", S->net.err.c_str());
    S->widths = widths;
    S->have_net = true;
  }
  mlp::Net& net = S->net;
  const int ldx = net.pin[0], ldt = net.pout[net.nl - 1];
  // normalised set
  if (S->Xn) (void)hipFree(S->Xn);
  if (S->Tn) (void)hipFree(S->Tn);
  if (S->Wn) (void)hipFree(S->Wn);
  if (S->idx) (void)hipFree(S->idx);
  SG_TRY(hipMalloc((void**)&S->Xn, 2 * (size_t)N * ldx));
  SG_TRY(hipMalloc((void**)&S->Tn, 4 * (size_t)N * ldt));
  SG_TRY(hipMalloc((void**)&S->Wn, 4 * (size_t)N * ldt));
  SG_TRY(hipMalloc((void**)&S->idx, 4 * (size_t)N));
  hipLaunchKernelGGL(k_norm_features, gs_capped(N * ldx, 256), dim3(256), 0, h->stream, S->Xraw, N, din, dst,
                     dst + din, S->Xn, ldx);
  hipLaunchKernelGGL(k_norm_targets, gs_capped(N * ldt, 256), dim3(256), 0, h->stream, S->Traw, S->Wraw, N, dout,
                     dst + 4 * din, dst + 4 * din + dout, S->Tn, S->Wn, ldt);
  SG_TRY(hipGetLastError());
  SG_TRY(hipStreamSynchronize(h->stream));
  // split: a held-out share for validation
  std::mt19937_64 rng((uint64_t)cfg->seed * 7919u + 17u);
  std::vector<unsigned> perm((size_t)N);
  for (long long i = 0; i < N; ++i) perm[(size_t)i] = (unsigned)i;
  std::shuffle(perm.begin(), perm.end(), rng);
  const long long n_val = std::min<long long>(N / 2, (long long)std::llround(cfg->val_fraction * (double)N));
  std::vector<unsigned> val(perm.begin(), perm.begin() + n_val), train(perm.begin() + n_val, perm.end());
  const int B = cfg->batch;
  const int Mpad = mlp::round_up(B, 64);
  int e = net.ensure_workspace(Mpad, true);
  if (e) return perr(e, "mlp_train: %s", net.err.c_str());
  const size_t nb_train = train.size() / (size_t)B, nb_val = (val.size() + (size_t)B - 1) / (size_t)B;
  if (cfg->epochs > 0 && nb_train == 0)
    return perr(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp_train: %zu training rows do not fill one batch of %d",
                train.size(), B);
  // per-batch loss and weight-sum slots on the device, read back once per
  // epoch into pinned memory; the training order lives in pinned memory
  // too, so that the next epoch's shuffle and upload overlap this epoch's
  // batches
  const size_t nslot = nb_train + nb_val, nt = train.size();
  struct Bufs {
    float* dl = nullptr;
    float* lh = nullptr;
    unsigned* order = nullptr;
    hipStream_t s;
    ~Bufs() {
      (void)hipStreamSynchronize(s);
      if (dl) (void)hipFree(dl);
      if (lh) (void)hipHostFree(lh);
      if (order) (void)hipHostFree(order);
    }
  } bufs{nullptr, nullptr, nullptr, h->stream};
  SG_TRY(hipMalloc((void**)&bufs.dl, 4 * (2 * nslot + 1)));
  SG_TRY(hipHostMalloc((void**)&bufs.lh, 4 * (2 * nslot + 1)));
  SG_TRY(hipHostMalloc((void**)&bufs.order, 4 * (nt + 1)));
  float* dl = bufs.dl;
  const float* lh = bufs.lh;
  unsigned* order = bufs.order;
  std::memcpy(order, train.data(), 4 * nt);
  // the validation rows sit after the training rows in S->idx
  SG_TRY(hipMemcpyAsync(S->idx + nt, val.data(), 4 * val.size(), hipMemcpyHostToDevice, h->stream));
  // this epoch's order: shuffled in place (the same sequence of shuffles
  // every run), uploaded behind the batches already queued
  auto next_order = [&]() -> hipError_t {
    std::shuffle(order, order + nt, rng);
    return hipMemcpyAsync(S->idx, order, 4 * nt, hipMemcpyHostToDevice, h->stream);
  };
  // one batch: forward (reading the set through the index), then either the
  // backward pass and an Adam step (training) or the loss alone (validation)
  auto run_batch = [&](size_t off, int nb, bool update, size_t q) -> int {
    net.src.X = S->Xn;
    net.src.T = S->Tn;
    net.src.W = S->Wn;
    net.src.idx = S->idx + off;
    net.src.nvalid = nb;
    int r = net.train_forward(Mpad, nb);
    if (!r && update) r = net.backward(Mpad);
    if (!r && update)
      r = net.update(true, (float)cfg->lr, (float)cfg->beta1, (float)cfg->beta2, (float)cfg->eps, dl + q,
                     dl + nslot + q);
    if (!r && !update) r = net.add_loss(dl + q, dl + nslot + q);
    if (r) return perr(r, "mlp_train: %s", net.err.c_str());
    return 0;
  };
  // the batches read the set through the index; inference after training
  // reads the network's own input buffer again
  struct SrcReset {
    mlp::Net& n;
    ~SrcReset() { n.src = mlp::Net::Source{}; }
  } src_reset{net};
  int done = 0;
  if (cfg->epochs > 0 && next_order() != hipSuccess)
    return perr(RDB_COLPHYS_ERR_SETUP, "mlp_train: batch index upload failed");
  for (int ep = 0; ep < cfg->epochs; ++ep) {
    e = 0;
    if (hipMemsetAsync(dl, 0, 4 * 2 * nslot, h->stream) != hipSuccess)
      e = perr(RDB_COLPHYS_ERR_SETUP, "mlp_train: loss slot clear failed");
    for (size_t q = 0; !e && q < nb_train; ++q) e = run_batch(q * B, B, true, q);
    for (size_t q = 0; !e && q < nb_val; ++q) {
      const size_t b0 = q * B;
      const int nb = (int)std::min<size_t>((size_t)B, val.size() - b0);
      e = run_batch(nt + b0, nb, false, nb_train + q);
    }
    if (!e && hipMemcpyAsync(bufs.lh, dl, 4 * 2 * nslot, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
      e = perr(RDB_COLPHYS_ERR_SETUP, "mlp_train: loss read-back failed");
    // the upload is queued behind this epoch's batches, and the previous
    // upload from `order` has completed (the stream was synchronised)
    if (!e && ep + 1 < cfg->epochs && next_order() != hipSuccess)
      e = perr(RDB_COLPHYS_ERR_SETUP, "mlp_train: batch index upload failed");
    if (!e && hipStreamSynchronize(h->stream) != hipSuccess)
      e = perr(RDB_COLPHYS_ERR_SETUP, "mlp_train: training batches failed");
    if (e) return e;
    // epoch losses: the batch means weighted by their sum(w)
    double ls = 0, cs = 0, vs = 0, vc = 0;
    const float* cnt = lh + nslot;
    for (size_t q = 0; q < nb_train; ++q) ls += (double)lh[q] * cnt[q], cs += cnt[q];
    for (size_t q = nb_train; q < nslot; ++q) vs += (double)lh[q] * cnt[q], vc += cnt[q];
    if (done < n_hist) {
      if (train_loss) train_loss[done] = cs > 0 ? (float)(ls / cs) : 0.f;
      if (val_loss) val_loss[done] = vc > 0 ? (float)(vs / vc) : 0.f;
    }
    ++done;
  }
  SG_TRY(hipStreamSynchronize(h->stream));
  if (n_done) *n_done = done;
  S->guard_mode = cfg->guard_mode;
  // the training batch may have resized the network's workspace: the next
  // surrogate step sets up its inference rows (and the member views) again
  S->inf_rows = 0;
  S->svs_up.clear();
  return RDB_COLPHYS_OK;
}

// Everything a surrogate step needs that is not a kernel launch: checks,
// the inference workspace, the guard and fallback buffers, the member views
// and the ensemble feature tables.  Run before a step is launched or
// captured; a change of any buffer the launches name bumps prep_gen.
int surr_prepare(rdb_colphys* h) {
  SurrState* S = h->surr;
  if (!S || !S->have_net || !S->d_stats) return perr(RDB_COLPHYS_ERR_NOT_INITIALISED, "surrogate: no trained network");
  mlp::Net& net = S->net;
  if (net.pin[0] > 384) return perr(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "surrogate: input width exceeds the feature tile");
  if (h->n_passive > 4) return perr(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "surrogate: at most 4 passive tracers");
  const int nzn = net_levels(h);
  if (S->din != 5 * nzn + 6) return perr(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "surrogate: network built for another nz");
  const long long rows = wet_rows(h);
  if (rows > 0x7fffff00LL) return perr(RDB_COLPHYS_ERR_BAD_SHAPE, "surrogate: too many wet columns");
  const int Mpad = mlp::round_up((int)std::max<long long>(rows, 1), 128);  // the fused inference kernel's row block
  // the inference workspace must hold every wet row: training in between
  // may have resized the network's buffers to its batch (Mcap)
  if (S->inf_rows != Mpad || net.Mcap < Mpad) {
    int e = net.ensure_workspace(Mpad, false);
    if (e) return perr(e, "surrogate: %s", net.err.c_str());
    SG_TRY(hipMemsetAsync(net.X0, 0, 2 * (size_t)Mpad * net.pin[0], h->stream));
    if (S->guard) (void)hipFree(S->guard);
    SG_TRY(hipMalloc((void**)&S->guard, (size_t)Mpad));
    if (!S->nfall) SG_TRY(hipMalloc((void**)&S->nfall, 64));
    S->inf_rows = Mpad;
    ++S->prep_gen;
  }
  if (net.Mcap < Mpad || !S->guard || !S->nfall)
    return perr(RDB_COLPHYS_ERR_SETUP, "surrogate: inference workspace of %d rows for %d wet rows", net.Mcap, Mpad);
  SG_TRY(mlp::prepare_forward());
  if (!h->members.empty()) {
    const std::vector<rdb_colphys*> ps = plans(h);
    const float* tmean = S->d_stats + 4 * S->din;
    const float* tstd = tmean + S->dout;
    // the member views change only with the network's workspace: upload
    // them (synchronously) when they do, never from a buffer a later step
    // could overwrite while the copy is pending
    std::vector<SurrView> svs;
    for (rdb_colphys* m : ps)
      svs.push_back(SurrView{net.Y + m->row_off, (long long)Mpad, tmean, tstd, S->guard + m->row_off, nzn - m->nz, 0,
                             tstd + S->dout, tstd + 2 * S->dout});
    if (svs.size() != S->svs_up.size() ||
        std::memcmp(svs.data(), S->svs_up.data(), sizeof(SurrView) * svs.size()) != 0) {
      SG_TRY(hipMemcpyAsync(h->d_svs, svs.data(), sizeof(SurrView) * svs.size(), hipMemcpyHostToDevice, h->stream));
      SG_TRY(hipStreamSynchronize(h->stream));
      S->svs_up = svs;
      ++S->prep_gen;
    }
    if (!S->d_fwork) {
      std::vector<SurrState::FeatMemberHost> fm;
      std::vector<int2> work;
      for (size_t mi = 0; mi < ps.size(); ++mi) {
        rdb_colphys* m = ps[mi];
        fm.push_back({m->d_coriolis, m->row_off, m->n_wet, nzn - m->nz});
        for (int p0 = 0; p0 < m->n_wet; p0 += FC) work.push_back(make_int2((int)mi, p0));
      }
      static_assert(sizeof(SurrState::FeatMemberHost) == sizeof(FeatMember), "feature member table layout");
      SG_TRY(hipMalloc(&S->d_fmem, sizeof(FeatMember) * std::max<size_t>(fm.size(), 1)));
      SG_TRY(hipMalloc((void**)&S->d_fwork, sizeof(int2) * std::max<size_t>(work.size(), 1)));
      SG_TRY(hipMemcpy(S->d_fmem, fm.data(), sizeof(FeatMember) * fm.size(), hipMemcpyHostToDevice));
      SG_TRY(hipMemcpy(S->d_fwork, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice));
      S->n_fwork = (int)work.size();
      ++S->prep_gen;
    }
    int st = ensure_work(h);  // the persistent step's work list, outside any capture
    if (st) return st;
  }
  h->mlp_flop = 0.0;
  for (int l = 0; l < net.nl; ++l) h->mlp_flop += 2.0 * (double)net.din[l] * (double)net.dout[l];
  h->mlp_flop *= (double)rows;
  return RDB_COLPHYS_OK;
}

// One surrogate step as kernel launches only (capturable): clear the
// fallback counter, features and guard flags of every wet column, the fused
// network forward, the column step with the predicted coefficients.
int surr_launch(rdb_colphys* h, double dt) {
  SurrState* S = h->surr;
  mlp::Net& net = S->net;
  const int nzn = net_levels(h);
  const int Mpad = S->inf_rows;
  const int ld = net.pin[0];
  hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, h->stream, S->nfall, 1);
  if (h->members.empty()) {
    if (h->n_wet > 0)
      hipLaunchKernelGGL(k_features_infer, gs(h->n_wet, FC), dim3(FB), 0, h->stream, h->s, h->d_coriolis, h->n_wet,
                         S->din, ld, S->d_stats, net.X0, S->guard, S->nfall, S->guard_mode, h->pr.rho0, h->pr.p_ref,
                         nzn, 0, b0_T(h->pr), b0_S(h->pr));
  } else if (S->n_fwork > 0) {
    hipLaunchKernelGGL(k_features_infer_ens, dim3((unsigned)S->n_fwork), dim3(FB), 0, h->stream, h->d_views,
                       (const FeatMember*)S->d_fmem, S->d_fwork, S->din, ld, S->d_stats, net.X0, S->guard, S->nfall,
                       S->guard_mode, h->pr.rho0, h->pr.p_ref, nzn, b0_T(h->pr), b0_S(h->pr));
  }
  SG_TRY(hipGetLastError());
  SG_TRY(mlp::forward_view(net.view(), net.X0, net.Y, Mpad, true, h->stream));
  if (!h->members.empty()) return launch_step_all(h, dt, true);
  const float* tmean = S->d_stats + 4 * S->din;
  const float* tstd = tmean + S->dout;
  const SurrView sv{net.Y, (long long)Mpad, tmean, tstd, S->guard, 0, 0, tstd + S->dout, tstd + 2 * S->dout};
  const bool two = (h->cfg.kt_bg != h->cfg.ks_bg);
  const int ntr = h->n_passive;
  dim3 g((unsigned)((h->n_wet + h->block - 1) / h->block)), b((unsigned)h->block);
  const int M = h->lds_levels;
  const size_t shmem = (size_t)h->block * (size_t)M * (size_t)(3 + std::min(ntr, 4) + (two ? 1 : 0)) * 8;
#define LAUNCH(N, TF)                                                                                              \
  hipLaunchKernelGGL((k_column_step<double, N, TF, true>), g, b, shmem, h->stream, h->s, h->pr, dt, 0, h->n_wet, \
                     h->d_guard, M, sv)
#define BY_NTR(TF)                  \
  switch (ntr) {                    \
    case 0: LAUNCH(0, TF); break;   \
    case 1: LAUNCH(1, TF); break;   \
    case 2: LAUNCH(2, TF); break;   \
    case 3: LAUNCH(3, TF); break;   \
    default: LAUNCH(4, TF); break;  \
  }
  if (h->n_wet > 0) {
    if (two) {
      BY_NTR(true)
    } else {
      BY_NTR(false)
    }
  }
#undef BY_NTR
#undef LAUNCH
  SG_TRY(hipGetLastError());
  return RDB_COLPHYS_OK;
}

// n surrogate steps: prepared, captured once into a hipGraph (rebuilt when
// dt, n, the launch shape or a buffer changes) and replayed.  With timing on
// the replay is bracketed by events on the handle's stream, and the network
// forward is timed on its own right after (one more forward over the last
// step's features: it rewrites the same outputs).  Nothing here waits for
// the device: the fallback count is read by surr_settle.
int surr_run(rdb_colphys* h, double dt, int n) {
  int st = surr_prepare(h);
  if (st) return st;
  SurrState* S = h->surr;
  if (!S->gexec || S->g_dt != dt || S->g_n != n || S->g_prep != S->prep_gen || S->g_block != h->block ||
      S->g_lds != h->lds_levels) {
    if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
    if (S->graph) (void)hipGraphDestroy(S->graph);
    S->gexec = nullptr;
    S->graph = nullptr;
    SG_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < n && !st; ++i) st = surr_launch(h, dt);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    SG_TRY(e);
    S->graph = g;
    SG_TRY(hipGraphInstantiate(&S->gexec, g, nullptr, nullptr, 0));
    S->g_dt = dt;
    S->g_n = n;
    S->g_prep = S->prep_gen;
    S->g_block = h->block;
    S->g_lds = h->lds_levels;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->timing) {
    st = step_events(h, &e0, &e1);
    if (st) return st;
    SG_TRY(hipEventRecord(e0, h->stream));
  }
  SG_TRY(hipGraphLaunch(S->gexec, h->stream));
  h->last_graph = 1;
  if (h->timing) {
    SG_TRY(hipEventRecord(e1, h->stream));
    SG_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    SG_TRY(hipEventElapsedTime(&ms, e0, e1));
    h->last_ms = ms / n;
    h->sum_ms += ms;
    h->timed += n;
    if (!S->ev[0]) SG_TRY(hipEventCreate(&S->ev[0]));
    if (!S->ev[1]) SG_TRY(hipEventCreate(&S->ev[1]));
    SG_TRY(hipEventRecord(S->ev[0], h->stream));
    SG_TRY(mlp::forward_view(S->net.view(), S->net.X0, S->net.Y, S->inf_rows, true, h->stream));
    SG_TRY(hipEventRecord(S->ev[1], h->stream));
  }
  S->pending = true;
  S->timed = h->timing;
  return RDB_COLPHYS_OK;
}

int surr_settle(rdb_colphys* h) {
  SurrState* S = h->surr;
  if (!S || !S->pending) return RDB_COLPHYS_OK;
  unsigned int nf = 0;
  SG_TRY(hipMemcpyAsync(&nf, S->nfall, 4, hipMemcpyDeviceToHost, h->stream));
  SG_TRY(hipStreamSynchronize(h->stream));
  h->last_fallback = nf;
  if (S->timed) {
    float ms = 0.f;
    SG_TRY(hipEventElapsedTime(&ms, S->ev[0], S->ev[1]));
    h->last_mlp_ms = ms;
  }
  S->pending = false;
  return RDB_COLPHYS_OK;
}

int surr_save(rdb_colphys* h, const char* path) {
  SurrState* S = h->surr;
  if (!S || !S->have_net || S->fmean.empty()) return perr(RDB_COLPHYS_ERR_NOT_INITIALISED, "mlp_save: no trained network");
  std::vector<float> p(S->net.n_public());
  if (S->net.get_params(p.data())) return perr(RDB_COLPHYS_ERR_SETUP, "mlp_save: %s", S->net.err.c_str());
  FILE* f = std::fopen(path, "wb");
  if (!f) return perr(RDB_COLPHYS_ERR_IO, "mlp_save: cannot open %s", path);
  // version 4: buoyancy-jump / signed-log level encoding of the features,
  // logarithms in single precision; version 5 adds the targets' range
  const char magic[8] = {'R', 'D', 'B', 'S', 'U', 'R', 'R', '5'};
  int hdr[4] = {S->nz, S->din, S->dout, (int)S->widths.size()};
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 4, f) == 4 &&
            std::fwrite(S->widths.data(), 4, S->widths.size(), f) == S->widths.size() &&
            std::fwrite(p.data(), 4, p.size(), f) == p.size();
  for (const std::vector<float>* v : {&S->fmean, &S->fstd, &S->fmin, &S->fmax, &S->tmean, &S->tstd, &S->tmin, &S->tmax})
    ok = ok && std::fwrite(v->data(), 4, v->size(), f) == v->size();
  ok = std::fclose(f) == 0 && ok;
  return ok ? RDB_COLPHYS_OK : perr(RDB_COLPHYS_ERR_IO, "mlp_save: write to %s failed", path);
}

int surr_load(rdb_colphys* h, const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return perr(RDB_COLPHYS_ERR_IO, "mlp_load: cannot open %s", path);
  char magic[8];
  int hdr[4];
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "RDBSURR", 7) == 0 &&
            (magic[7] == '4' || magic[7] == '5') && std::fread(hdr, 4, 4, f) == 4;
  const bool has_trange = ok && magic[7] == '5';
  if (!ok || hdr[0] != net_levels(h) || hdr[3] < 2 || hdr[3] > 16) {
    std::fclose(f);
    return perr(RDB_COLPHYS_ERR_BAD_SHAPE, "mlp_load: %s is not a network for nz = %d", path, net_levels(h));
  }
  SurrState* S = surr_get(h);
  std::vector<int> widths((size_t)hdr[3]);
  ok = std::fread(widths.data(), 4, widths.size(), f) == widths.size() && widths.front() == S->din &&
       widths.back() == S->dout;
  int e = ok ? S->net.init(widths, h->stream) : 2;
  std::vector<float> p(ok && !e ? S->net.n_public() : 0);
  ok = ok && !e && std::fread(p.data(), 4, p.size(), f) == p.size();
  S->fmean.resize(S->din);
  S->fstd.resize(S->din);
  S->fmin.resize(S->din);
  S->fmax.resize(S->din);
  S->tmean.resize(S->dout);
  S->tstd.resize(S->dout);
  for (std::vector<float>* v : {&S->fmean, &S->fstd, &S->fmin, &S->fmax, &S->tmean, &S->tstd})
    ok = ok && std::fread(v->data(), 4, v->size(), f) == v->size();
  // a version-4 file has no targets' range: predictions are not clamped
  S->tmin.assign(S->dout, -3.4e38f);
  S->tmax.assign(S->dout, 3.4e38f);
  if (has_trange)
    for (std::vector<float>* v : {&S->tmin, &S->tmax}) ok = ok && std::fread(v->data(), 4, v->size(), f) == v->size();
  std::fclose(f);
  if (!ok) return perr(RDB_COLPHYS_ERR_IO, "mlp_load: %s is truncated or malformed", path);
  e = S->net.set_params(p.data());
  if (e) return perr(e, "mlp_load: %s", S->net.err.c_str());
  S->widths = widths;
  S->have_net = true;
  std::vector<float> stats;
  for (const std::vector<float>* v : {&S->fmean, &S->fstd, &S->fmin, &S->fmax, &S->tmean, &S->tstd, &S->tmin, &S->tmax})
    stats.insert(stats.end(), v->begin(), v->end());
  if (S->d_stats) (void)hipFree(S->d_stats);
  S->d_stats = nullptr;
  SG_TRY(hipMalloc((void**)&S->d_stats, 4 * stats.size()));
  SG_TRY(hipMemcpy(S->d_stats, stats.data(), 4 * stats.size(), hipMemcpyHostToDevice));
  S->inf_rows = 0;
  S->svs_up.clear();
  return RDB_COLPHYS_OK;
}

int surr_set_guard(rdb_colphys* h, int mode) {
  if (mode < RDB_COLPHYS_GUARD_STRICT || mode > RDB_COLPHYS_GUARD_RANGE)
    return perr(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "set_guard: unknown mode %d", mode);
  SurrState* S = surr_get(h);
  if (S->guard_mode != mode) ++S->prep_gen;  // a launch argument of the captured step
  S->guard_mode = mode;
  return RDB_COLPHYS_OK;
}

void surr_invalidate(rdb_colphys* h) {
  if (h->surr) ++h->surr->prep_gen;
}

bool surr_ready(rdb_colphys* h) { return h->surr && h->surr->have_net && h->surr->d_stats; }

}  // namespace colphys
