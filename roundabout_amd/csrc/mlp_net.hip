// mlp_net.hip -- forward, backward and Adam of the BF16 MFMA MLP.
#include "mlp_net.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>

namespace mlp {

namespace {
inline dim3 g1(long long n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

// M, N multiples of 64, K of 32; splits > 1 (EPI_F32 only) writes the
// partial products of k-chunks of length kc to Cf + s * split_stride.
inline int gemm_blocks(const GemmArgs& g) { return (g.M / 64) * (g.N / 64) * g.splits; }
inline GemmArgs gargs(const bf16* A, int lda, const bf16* B, int ldb, int M, int N, int K, const GemmEpi& ep,
                      int splits = 1) {
  GemmArgs g{};
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.M = M;
  g.N = N;
  g.K = K;
  g.splits = splits;
  g.ep = ep;
  return g;
}
template <int EPI>
void gemm(hipStream_t s, const GemmArgs& g) {
  hipLaunchKernelGGL(k_gemm_bf16<EPI>, dim3((unsigned)gemm_blocks(g)), dim3(256), 0, s, g);
}
inline GemmEpi epi() { return GemmEpi{}; }

// k-splits of a weight-gradient product over Mpad batch rows: at most 16
// chunks of at least 128 rows, chunk length a multiple of 32 (32 chunks
// doubled the partials k_grad_step reads for no faster product)
inline void batch_splits(int Mpad, int* S, int* kc) {
  *S = std::max(1, std::min(16, Mpad / 128));
  *kc = round_up((Mpad + *S - 1) / *S, 32);
}

void transpose(hipStream_t s, const bf16* in, int ldi, bf16* out, int ldo, int R, int C) {
  dim3 grid((unsigned)((C + 31) / 32), (unsigned)((R + 31) / 32));
  hipLaunchKernelGGL(k_transpose_bf16, grid, dim3(32, 8), 0, s, in, ldi, out, ldo, R, C);
}
}  // namespace

int Net::fail(const char* what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  err = buf;
  return 3;
}

#define NT_TRY(expr)                      \
  do {                                    \
    hipError_t _e = (expr);               \
    if (_e != hipSuccess) return fail(#expr, _e); \
  } while (0)

int Net::init(const std::vector<int>& dims, hipStream_t s) {
  release();
  stream = s;
  if (dims.size() < 2) {
    err = "network needs at least an input and an output width";
    return 2;
  }
  nl = (int)dims.size() - 1;
  din.assign(dims.begin(), dims.end() - 1);
  dout.assign(dims.begin() + 1, dims.end());
  pin.resize(nl);
  pout.resize(nl);
  for (int l = 0; l < nl; ++l) {
    if (din[l] < 1 || dout[l] < 1) {
      err = "layer widths must be positive";
      return 2;
    }
    pin[l] = round_up(din[l], 64);
    pout[l] = round_up(dout[l], 64);
  }
  np = 0;
  nwb = 0;
  offW.resize(nl);
  offB.resize(nl);
  offWb.resize(nl);
  for (int l = 0; l < nl; ++l) {
    offW[l] = np;
    np += (size_t)pout[l] * pin[l];
    offB[l] = np;
    np += (size_t)pout[l];
    offWb[l] = nwb;
    nwb += (size_t)pout[l] * pin[l];
  }
  NT_TRY(hipMalloc((void**)&p, 4 * np));
  NT_TRY(hipMalloc((void**)&g, 4 * np));
  NT_TRY(hipMalloc((void**)&m1, 4 * np));
  NT_TRY(hipMalloc((void**)&m2, 4 * np));
  NT_TRY(hipMalloc((void**)&Wb, 2 * nwb));
  NT_TRY(hipMalloc((void**)&WTb, 2 * nwb));
  NT_TRY(hipMalloc((void**)&dloss, 64));
  NT_TRY(hipMemsetAsync(p, 0, 4 * np, stream));
  NT_TRY(hipMemsetAsync(g, 0, 4 * np, stream));
  NT_TRY(hipMemsetAsync(m1, 0, 4 * np, stream));
  NT_TRY(hipMemsetAsync(m2, 0, 4 * np, stream));
  adam_t = 0;
  return 0;
}

size_t Net::n_public() const {
  size_t n = 0;
  for (int l = 0; l < nl; ++l) n += (size_t)dout[l] * din[l] + dout[l];
  return n;
}

int Net::set_params(const float* host) {
  std::vector<float> pad(np, 0.f);
  size_t q = 0;
  for (int l = 0; l < nl; ++l) {
    for (int r = 0; r < dout[l]; ++r)
      for (int c = 0; c < din[l]; ++c) pad[offW[l] + (size_t)r * pin[l] + c] = host[q++];
    for (int r = 0; r < dout[l]; ++r) pad[offB[l] + r] = host[q++];
  }
  NT_TRY(hipMemcpyAsync(p, pad.data(), 4 * np, hipMemcpyHostToDevice, stream));
  NT_TRY(hipMemsetAsync(m1, 0, 4 * np, stream));
  NT_TRY(hipMemsetAsync(m2, 0, 4 * np, stream));
  adam_t = 0;
  for (int l = 0; l < nl; ++l) {
    dim3 grid((unsigned)((pin[l] + 255) / 256), (unsigned)pout[l]);
    hipLaunchKernelGGL(k_refresh_weights, grid, dim3(256), 0, stream, p + offW[l], pout[l], pin[l], Wb + offWb[l],
                       pin[l], WTb + offWb[l], pout[l], pout[l], pin[l]);
  }
  NT_TRY(hipGetLastError());
  NT_TRY(hipStreamSynchronize(stream));
  return 0;
}

static int copy_public(const Net& n, const float* dev, float* host, hipStream_t s, std::string& err) {
  std::vector<float> pad(n.np);
  if (hipMemcpyAsync(pad.data(), dev, 4 * n.np, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    err = "device to host copy of parameters failed";
    return 3;
  }
  size_t q = 0;
  for (int l = 0; l < n.nl; ++l) {
    for (int r = 0; r < n.dout[l]; ++r)
      for (int c = 0; c < n.din[l]; ++c) host[q++] = pad[n.offW[l] + (size_t)r * n.pin[l] + c];
    for (int r = 0; r < n.dout[l]; ++r) host[q++] = pad[n.offB[l] + r];
  }
  return 0;
}

int Net::get_params(float* host) const { return copy_public(*this, p, host, stream, const_cast<Net*>(this)->err); }
int Net::get_grads(float* host) const { return copy_public(*this, g, host, stream, const_cast<Net*>(this)->err); }

void Net::free_workspace() {
  auto F = [](void* q) {
    if (q) (void)hipFree(q);
  };
  F(X0);
  F(X0T);
  for (auto q : pre) F(q);
  for (auto q : act) F(q);
  for (auto q : actT) F(q);
  F(Y);
  F(T);
  F(Wm);
  F(dY);
  F(dYT);
  F(dpre);
  F(dpreT);
  F(wsw);
  F(wsb);
  F(lpart);
  F(cpart);
  wsw = wsb = lpart = cpart = nullptr;
  X0 = X0T = nullptr;
  pre.clear();
  act.clear();
  actT.clear();
  Y = T = Wm = nullptr;
  dY = dYT = dpre = dpreT = nullptr;
  Mcap = 0;
}

void Net::release() {
  free_workspace();
  auto F = [](void* q) {
    if (q) (void)hipFree(q);
  };
  F(p);
  F(g);
  F(m1);
  F(m2);
  F(Wb);
  F(WTb);
  F(dloss);
  p = g = m1 = m2 = dloss = nullptr;
  Wb = WTb = nullptr;
  nl = 0;
}

int Net::ensure_workspace(int Mpad, bool training) {
  if (Mpad <= Mcap && (!training || T)) return 0;
  free_workspace();
  const size_t M = (size_t)Mpad;
  int wmax = 0;
  for (int l = 0; l < nl; ++l) wmax = std::max(wmax, std::max(pin[l], pout[l]));
  NT_TRY(hipMalloc((void**)&X0, 2 * M * pin[0]));
  pre.assign(nl - 1, nullptr);
  act.assign(nl - 1, nullptr);
  actT.assign(nl - 1, nullptr);
  for (int l = 0; l < nl - 1; ++l) {
    NT_TRY(hipMalloc((void**)&act[l], 2 * M * pout[l]));
    if (training) {
      NT_TRY(hipMalloc((void**)&pre[l], 4 * M * pout[l]));
      NT_TRY(hipMalloc((void**)&actT[l], 2 * M * pout[l]));
    }
  }
  const int po = pout[nl - 1];
  NT_TRY(hipMalloc((void**)&Y, 4 * M * po));
  if (training) {
    NT_TRY(hipMalloc((void**)&X0T, 2 * M * pin[0]));
    NT_TRY(hipMalloc((void**)&T, 4 * M * po));
    NT_TRY(hipMalloc((void**)&Wm, 4 * M * po));
    NT_TRY(hipMalloc((void**)&dY, 2 * M * wmax));
    NT_TRY(hipMalloc((void**)&dYT, 2 * M * wmax));
    NT_TRY(hipMalloc((void**)&dpre, 2 * M * wmax));
    NT_TRY(hipMalloc((void**)&dpreT, 2 * M * wmax));
    int S, kc;
    batch_splits(Mpad, &S, &kc);
    const size_t SB = M / 64;
    offPW.assign(nl, 0);
    offPB.assign(nl, 0);
    size_t nw = 0, nbias = 0;
    for (int l = 0; l < nl; ++l) {
      offPW[l] = nw;
      nw += (size_t)S * pout[l] * pin[l];
      offPB[l] = nbias;
      nbias += SB * pout[l];
    }
    NT_TRY(hipMalloc((void**)&wsw, 4 * nw));
    NT_TRY(hipMalloc((void**)&wsb, 4 * nbias));
    NT_TRY(hipMalloc((void**)&lpart, 4 * SB * (size_t)(po / 64)));
    NT_TRY(hipMalloc((void**)&cpart, 4 * SB * (size_t)(po / 64)));
  }
  Mcap = Mpad;
  return 0;
}

static bool unfused_requested() {
  // RDB_MLP_UNFUSED=1 keeps inference on the per-layer GEMMs (A/B timing)
  static const bool unfused = [] {
    const char* e = std::getenv("RDB_MLP_UNFUSED");
    return e && e[0] == '1';
  }();
  return unfused;
}

bool fused_shape_ok(const LayerView& v, int Mpad) {
  if (unfused_requested() || Mpad % 128 != 0 || v.nl > FP_MAXL) return false;
  for (int l = 0; l < v.nl; ++l)
    if (v.pin[l] > (l == 0 ? FP_MAXK : FP_MAXN) || v.pout[l] > FP_MAXN || v.pin[l] % 64 || v.pout[l] % 64)
      return false;
  return true;
}

hipError_t prepare_forward() {
  // the fused kernel's dynamic LDS, allowed once per device (0..63)
  static unsigned long long lds_set = 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev >= 64 || !(lds_set >> dev & 1ull)) {
    e = hipFuncSetAttribute((const void*)k_mlp_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)FP_LDS_BYTES);
    if (e != hipSuccess) return e;
    if (dev < 64) lds_set |= 1ull << dev;
  }
  return hipSuccess;
}

hipError_t forward_view(const LayerView& v, const bf16* X0, float* Y, int Mpad, bool out_major, hipStream_t s) {
  if (v.nl < 1 || v.nl > LV_MAXL) return hipErrorInvalidValue;
  if (fused_shape_ok(v, Mpad)) {
    // inference: one launch over all layers, activations in LDS
    hipError_t e = prepare_forward();
    if (e != hipSuccess) return e;
    const size_t shmem = FP_LDS_BYTES;
    FusedNet fn{};
    fn.nl = v.nl;
    for (int l = 0; l < v.nl; ++l) {
      fn.W[l] = v.W[l];
      fn.b[l] = v.b[l];
      fn.pin[l] = v.pin[l];
      fn.pout[l] = v.pout[l];
    }
    const long long ldy = out_major ? (long long)Mpad : (long long)v.pout[v.nl - 1];
    hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)(Mpad / 128)), dim3(64 * FP_WAVES), shmem, s, fn, X0, v.pin[0], Y,
                       ldy, out_major ? 1 : 0);
    return hipGetLastError();
  }
  // per-layer GEMMs (batches the fused kernel does not take)
  const bf16* x = X0;
  int ldx = v.pin[0];
  for (int l = 0; l < v.nl; ++l) {
    GemmEpi ep = epi();
    ep.bias = v.b[l];
    if (l < v.nl - 1) {
      ep.Cb = v.act[l];
      ep.ldcb = v.pout[l];
      gemm<EPI_BIAS_GELU>(s, gargs(x, ldx, v.W[l], v.pin[l], Mpad, v.pout[l], v.pin[l], ep));
      x = v.act[l];
      ldx = v.pout[l];
    } else if (out_major) {
      ep.Cf = Y;
      ep.ldcf = Mpad;
      gemm<EPI_BIAS_F32_T>(s, gargs(x, ldx, v.W[l], v.pin[l], Mpad, v.pout[l], v.pin[l], ep));
    } else {
      ep.Cf = Y;
      ep.ldcf = v.pout[l];
      gemm<EPI_BIAS_F32>(s, gargs(x, ldx, v.W[l], v.pin[l], Mpad, v.pout[l], v.pin[l], ep));
    }
  }
  return hipGetLastError();
}

LayerView Net::view() const {
  LayerView v{};
  v.nl = nl;
  for (int l = 0; l < nl && l < LV_MAXL; ++l) {
    v.pin[l] = pin[l];
    v.pout[l] = pout[l];
    v.W[l] = Wb + offWb[l];
    v.b[l] = p + offB[l];
    v.act[l] = l < (int)act.size() ? act[l] : nullptr;
  }
  return v;
}

bool Net::fused_ok(int Mpad) const { return fused_shape_ok(view(), Mpad); }

int Net::forward(int Mpad, bool out_major) {
  if (nl > LV_MAXL) {
    err = "forward: more layers than the inference chain takes";
    return 2;
  }
  NT_TRY(forward_view(view(), X0, Y, Mpad, out_major, stream));
  return 0;
}

int Net::transpose_input(int Mpad) {
  transpose(stream, X0, pin[0], X0T, Mpad, Mpad, pin[0]);
  NT_TRY(hipGetLastError());
  return 0;
}

int Net::train_forward(int Mpad, int Mvalid) {
  const int L = nl - 1;
  const int po = pout[L];
  int kc;
  batch_splits(Mpad, &bw_S, &kc);
  bw_SB = Mpad / 64;
  bw_nlp = (Mpad / 64) * (po / 64);
  const bool gathered = src.X != nullptr;
  const bf16* x = gathered ? src.X : X0;
  int ldx = pin[0];
  for (int l = 0; l < nl; ++l) {
    GemmEpi ep = epi();
    ep.bias = p + offB[l];
    const bf16* W = Wb + offWb[l];
    if (l == 0 && gathered) {
      // the batch rows straight from the set, and their feature-major copy
      // (the first layer's weight-gradient operand) as they stream in
      ep.aidx = src.idx;
      ep.anv = src.nvalid;
      ep.AT = X0T;
      ep.ldat = Mpad;
    }
    if (l < L) {
      // hidden layer: the FP32 pre-activation (for GELU') and the activation
      // row-major (next layer's input) and feature-major (next layer's
      // weight-gradient operand)
      ep.Cf = pre[l];
      ep.ldcf = pout[l];
      ep.Cb = act[l];
      ep.ldcb = pout[l];
      ep.CbT = actT[l];
      ep.ldct = Mpad;
      gemm<EPI_BIAS_GELU>(stream, gargs(x, ldx, W, pin[l], Mpad, pout[l], pin[l], ep));
      x = act[l];
      ldx = pout[l];
    } else {
      // output layer: the loss partials, the output gradient in both
      // layouts and its bias-gradient partials straight from the tile
      ep.T = src.T ? src.T : T;
      ep.W = src.T ? src.W : Wm;
      ep.tidx = src.T ? src.idx : nullptr;
      ep.ldt = po;
      ep.mvalid = Mvalid;
      ep.nvalid = dout[L];
      ep.lpart = lpart;
      ep.cpart = cpart;
      ep.Cb = dY;
      ep.ldcb = po;
      ep.CbT = dYT;
      ep.ldct = Mpad;
      ep.colsum = wsb + offPB[L];
      gemm<EPI_MSE>(stream, gargs(x, ldx, W, pin[l], Mpad, po, pin[l], ep));
    }
  }
  NT_TRY(hipGetLastError());
  return 0;
}

int Net::backward(int Mpad) {
  const int L = nl - 1;
  int S, kc;
  batch_splits(Mpad, &S, &kc);
  for (int l = L; l >= 0; --l) {
    const int o = pout[l], i = pin[l];
    const bf16* xinT = (l == 0) ? X0T : actT[l - 1];
    // dW_l = d^T x_in : (o x Mpad) (i x Mpad)^T, split over the batch (K =
    // Mpad is long, the output small) into partials summed in k_grad_step
    GemmEpi ew = epi();
    ew.Cf = wsw + offPW[l];
    ew.ldcf = i;
    ew.kc = kc;
    ew.split_stride = (long long)o * i;
    const GemmArgs gw = gargs(dYT, Mpad, xinT, Mpad, o, i, Mpad, ew, S);
    if (l == 0) {
      gemm<EPI_F32>(stream, gw);
      break;
    }
    // with it, in the same launch: d_{l-1} = (d W_l) * GELU'(pre_{l-1}) :
    // (Mpad x o) (i x o)^T, its feature-major copy and the bias-gradient
    // partials of layer l - 1
    GemmEpi ex = epi();
    ex.pre = pre[l - 1];
    ex.ldp = i;
    ex.Cb = dpre;
    ex.ldcb = i;
    ex.CbT = dpreT;
    ex.ldct = Mpad;
    ex.colsum = wsb + offPB[l - 1];
    const GemmArgs gx = gargs(dY, o, WTb + offWb[l], o, Mpad, i, o, ex);
    hipLaunchKernelGGL((k_gemm_pair<EPI_F32, EPI_GELU_BWD>), dim3((unsigned)(gemm_blocks(gw) + gemm_blocks(gx))),
                       dim3(256), 0, stream, gw, gx);
    std::swap(dY, dpre);
    std::swap(dYT, dpreT);
  }
  NT_TRY(hipGetLastError());
  return 0;
}

int Net::update(bool adam_step, float lr, float b1, float b2, float eps, float* loss_dev, float* count_dev) {
  GradPlan gp{};
  gp.nl = nl;
  for (int l = 0; l < nl; ++l) {
    gp.pin[l] = pin[l];
    gp.pout[l] = pout[l];
    gp.offW[l] = (long long)offW[l];
    gp.offWb[l] = (long long)offWb[l];
    gp.pw[l] = wsw + offPW[l];
    gp.pb[l] = wsb + offPB[l];
  }
  float c1 = 1.f, c2 = 1.f;
  if (adam_step) {
    ++adam_t;
    // bias corrections in FP64: 1 - beta2^t cancels (beta2 = 0.999 leaves
    // 0.002 at t = 2), and in FP32 that costs ~1e-5 relative in the step
    c1 = (float)(1.0 - std::pow((double)b1, (double)adam_t));
    c2 = (float)(1.0 - std::pow((double)b2, (double)adam_t));
  }
  hipLaunchKernelGGL(k_grad_step, g1((long long)np, 256), dim3(256), 0, stream, gp, bw_S, bw_SB, (long long)np, p, g,
                     m1, m2, Wb, WTb, adam_step ? 1 : 0, lr, b1, b2, eps, c1, c2, (const float*)lpart,
                     (const float*)cpart, bw_nlp, loss_dev, count_dev);
  NT_TRY(hipGetLastError());
  return 0;
}

int Net::add_loss(float* loss_dev, float* count_dev) {
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, stream, (const float*)lpart, (const float*)cpart, bw_nlp,
                     loss_dev, count_dev);
  NT_TRY(hipGetLastError());
  return 0;
}

}  // namespace mlp
