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
template <int EPI>
void gemm(hipStream_t s, const bf16* A, int lda, const bf16* B, int ldb, int M, int N, int K, const float* bias,
          float* Cf, int ldcf, bf16* Cb, int ldcb, int splits = 1, int kc = 0, long long split_stride = 0) {
  dim3 grid((unsigned)((long long)(M / 64) * (N / 64)), (unsigned)splits);
  hipLaunchKernelGGL(k_gemm_bf16<EPI>, grid, dim3(64), 0, s, A, lda, B, ldb, M, N, K, bias, Cf, ldcf, Cb, ldcb,
                     splits > 1 ? kc : K, split_stride);
}

// k-splits of a weight-gradient product over Mpad batch rows: at most 32
// chunks of at least 64 rows, chunk length a multiple of 32
inline void batch_splits(int Mpad, int* S, int* kc) {
  *S = std::max(1, std::min(32, Mpad / 64));
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
  NT_TRY(hipMalloc((void**)&lpart, sizeof(float) * MSE_BLOCKS));
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
  F(dXf);
  F(dpre);
  F(dpreT);
  F(ws);
  ws = nullptr;
  X0 = X0T = nullptr;
  pre.clear();
  act.clear();
  actT.clear();
  Y = T = Wm = dXf = nullptr;
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
  F(lpart);
  p = g = m1 = m2 = dloss = lpart = nullptr;
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
    NT_TRY(hipMalloc((void**)&dXf, 4 * M * wmax));
    NT_TRY(hipMalloc((void**)&dpre, 2 * M * wmax));
    NT_TRY(hipMalloc((void**)&dpreT, 2 * M * wmax));
    int S, kc;
    batch_splits(Mpad, &S, &kc);
    size_t per = 0;
    for (int l = 0; l < nl; ++l) per = std::max(per, (size_t)pout[l] * pin[l] + pout[l]);
    NT_TRY(hipMalloc((void**)&ws, 4 * (size_t)S * per));
  }
  Mcap = Mpad;
  return 0;
}

bool Net::fused_ok(int Mpad) const {
  // RDB_MLP_UNFUSED=1 keeps inference on the per-layer GEMMs (A/B timing)
  static const bool unfused = [] {
    const char* e = std::getenv("RDB_MLP_UNFUSED");
    return e && e[0] == '1';
  }();
  if (unfused || Mpad % 128 != 0 || nl > FP_MAXL) return false;
  for (int l = 0; l < nl; ++l)
    if (pin[l] > FP_MAXK || pout[l] > FP_MAXN) return false;
  return true;
}

int Net::forward(int Mpad, bool keep_pre, bool out_major) {
  if (!keep_pre && fused_ok(Mpad)) {
    // inference: one launch over all layers, activations in LDS
    static bool lds_set = false;
    const size_t shmem = FP_LDS_BYTES;
    if (!lds_set) {
      NT_TRY(hipFuncSetAttribute((const void*)k_mlp_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
      lds_set = true;
    }
    FusedNet fn{};
    fn.nl = nl;
    for (int l = 0; l < nl; ++l) {
      fn.W[l] = Wb + offWb[l];
      fn.b[l] = p + offB[l];
      fn.pin[l] = pin[l];
      fn.pout[l] = pout[l];
    }
    const long long ldy = out_major ? (long long)Mpad : (long long)pout[nl - 1];
    hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)(Mpad / 128)), dim3(64 * FP_WAVES), shmem, stream, fn,
                       (const bf16*)X0,
                       pin[0], Y, ldy, out_major ? 1 : 0);
    NT_TRY(hipGetLastError());
    return 0;
  }
  const bf16* x = X0;
  int ldx = pin[0];
  for (int l = 0; l < nl; ++l) {
    const bf16* W = Wb + offWb[l];
    const float* b = p + offB[l];
    if (l < nl - 1) {
      gemm<EPI_BIAS_GELU>(stream, x, ldx, W, pin[l], Mpad, pout[l], pin[l], b, keep_pre ? pre[l] : nullptr, pout[l],
                          act[l], pout[l]);
      x = act[l];
      ldx = pout[l];
    } else if (out_major) {
      gemm<EPI_BIAS_F32_T>(stream, x, ldx, W, pin[l], Mpad, pout[l], pin[l], b, Y, Mpad, nullptr, 0);
    } else {
      gemm<EPI_BIAS_F32>(stream, x, ldx, W, pin[l], Mpad, pout[l], pin[l], b, Y, pout[l], nullptr, 0);
    }
  }
  NT_TRY(hipGetLastError());
  return 0;
}

int Net::backward(int Mpad, int Mvalid, double count, float* loss_out, float* loss_dev) {
  const int L = nl - 1;
  const int po = pout[L];
  float* lacc = loss_dev ? loss_dev : dloss;
  if (!loss_dev) NT_TRY(hipMemsetAsync(dloss, 0, 4, stream));
  const float inv = count > 0 ? (float)(1.0 / count) : 0.f;
  const int nb = (int)std::min<long long>(MSE_BLOCKS, ((long long)Mpad * po + 255) / 256);
  hipLaunchKernelGGL(k_mse, dim3((unsigned)nb), dim3(256), 0, stream, Y, T, Wm, Mpad, Mvalid, dout[L], po, inv, dY,
                     lpart);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, stream, lpart, nb, inv, lacc);
  int S, kc;
  batch_splits(Mpad, &S, &kc);
  // d (bf16, Mpad x pout[l]) is the gradient w.r.t. layer l's output (pre-activation)
  bf16* d = dY;
  for (int l = L; l >= 0; --l) {
    const int o = pout[l], i = pin[l];
    const bf16* xin = (l == 0) ? X0 : act[l - 1];
    bf16* xinT = (l == 0) ? X0T : actT[l - 1];
    // dW_l = d^T x_in : (o x Mpad) (i x Mpad)^T
    transpose(stream, d, o, dYT, Mpad, Mpad, o);
    transpose(stream, xin, i, xinT, Mpad, Mpad, i);
    // split over the batch (K = Mpad is long, the output small), then a
    // fixed-order sum of the partials
    const long long oi = (long long)o * i;
    gemm<EPI_F32>(stream, dYT, Mpad, xinT, Mpad, o, i, Mpad, nullptr, ws, i, nullptr, 0, S, kc, oi);
    hipLaunchKernelGGL(k_reduce_splits, g1(oi, 256), dim3(256), 0, stream, ws, S, oi, oi, g + offW[l]);
    float* bpart = ws + (size_t)S * (size_t)oi;
    hipLaunchKernelGGL(k_bias_partial, dim3((unsigned)(o / 64), (unsigned)S), dim3(64, BP_PHASES), 0, stream, d, Mpad, o, o,
                       kc, bpart);
    hipLaunchKernelGGL(k_reduce_splits, g1(o, 256), dim3(256), 0, stream, bpart, S, (long long)o, (long long)o,
                       g + offB[l]);
    if (l > 0) {
      // dX = d W_l : (Mpad x o) (i x o)^T, then through GELU of layer l-1
      gemm<EPI_F32>(stream, d, o, WTb + offWb[l], o, Mpad, i, o, nullptr, dXf, i, nullptr, 0);
      hipLaunchKernelGGL(k_gelu_backward, g1((long long)Mpad * i, 256), dim3(256), 0, stream, dXf, pre[l - 1], dpre,
                         (long long)Mpad * i);
      // the next layer's d lives in dpre; swap the role of the two buffers
      std::swap(dY, dpre);
      d = dY;
    }
  }
  NT_TRY(hipGetLastError());
  if (loss_out && !loss_dev) {
    NT_TRY(hipMemcpyAsync(loss_out, dloss, 4, hipMemcpyDeviceToHost, stream));
    NT_TRY(hipStreamSynchronize(stream));
  }
  return 0;
}

int Net::adam(float lr, float b1, float b2, float eps) {
  ++adam_t;
  float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
  hipLaunchKernelGGL(k_adam, g1((long long)np, 256), dim3(256), 0, stream, p, g, m1, m2, (long long)np, lr, b1, b2,
                     eps, c1, c2);
  for (int l = 0; l < nl; ++l) {
    dim3 grid((unsigned)((pin[l] + 255) / 256), (unsigned)pout[l]);
    hipLaunchKernelGGL(k_refresh_weights, grid, dim3(256), 0, stream, p + offW[l], pout[l], pin[l], Wb + offWb[l],
                       pin[l], WTb + offWb[l], pout[l], pout[l], pin[l]);
  }
  NT_TRY(hipGetLastError());
  return 0;
}

}  // namespace mlp
