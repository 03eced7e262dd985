// mlp_net.h -- host driver of the BF16 MFMA multilayer perceptron.
//
// Parameters live on the device as one padded FP32 master buffer: for each
// linear layer l, W_l (out_pad x in_pad, row-major: output-major like the
// public layout) followed by b_l (out_pad).  Padding rows/columns are zero
// and stay zero (their gradients vanish).  BF16 copies W_l and W_l^T are
// refreshed after every update.  The public (host) layout is unpadded:
// W_0, b_0, W_1, b_1, ... with W_l as (out_l x in_l).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mlp_kernels.h"

namespace mlp {

struct Net {
  int nl = 0;                      // linear layers
  std::vector<int> din, dout;      // unpadded in/out per layer
  std::vector<int> pin, pout;      // padded (multiples of 64)
  std::vector<size_t> offW, offB;  // offsets into the FP32 master buffer
  std::vector<size_t> offWb;       // offsets into the bf16 copies
  size_t np = 0, nwb = 0;
  float* p = nullptr;
  float* g = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  bf16* Wb = nullptr;
  bf16* WTb = nullptr;
  int adam_t = 0;
  // workspace for a batch of up to Mcap rows (multiple of 64)
  int Mcap = 0;
  bf16* X0 = nullptr;
  bf16* X0T = nullptr;
  std::vector<float*> pre;
  std::vector<bf16*> act, actT;
  float* Y = nullptr;
  float* T = nullptr;
  float* Wm = nullptr;
  bf16* dY = nullptr;
  bf16* dYT = nullptr;
  float* dXf = nullptr;
  bf16* dpre = nullptr;
  bf16* dpreT = nullptr;
  float* dloss = nullptr;
  float* lpart = nullptr;  // per-workgroup partial losses of k_mse
  float* ws = nullptr;  // split-K partials of the weight and bias gradients
  hipStream_t stream = nullptr;
  std::string err;

  ~Net() { release(); }
  int init(const std::vector<int>& dims, hipStream_t s);
  int ensure_workspace(int Mpad, bool training);
  int set_params(const float* host);   // unpadded public layout
  int get_params(float* host) const;
  int get_grads(float* host) const;
  size_t n_public() const;
  // X0 (Mpad x pin[0] bf16) must hold the inputs; Y receives the outputs,
  // row-major (Mpad x pout) or, with out_major, output-major (pout x Mpad)
  // Inference (keep_pre false) with Mpad a multiple of 128 and widths within
  // the fused kernel's limits runs as one fused launch (k_mlp_fused).
  int forward(int Mpad, bool keep_pre, bool out_major = false);
  bool fused_ok(int Mpad) const;
  // T, Wm (Mpad x pout[last]) hold targets and weights; rows >= Mvalid
  // ignored.  The loss is added to *loss_dev (device, no synchronisation)
  // when given, else returned in *loss_out (synchronises) when that is given.
  int backward(int Mpad, int Mvalid, double count, float* loss_out, float* loss_dev = nullptr);
  int adam(float lr, float b1, float b2, float eps);
  void release();

 private:
  void free_workspace();
  int fail(const char* what, hipError_t e);
};

inline int round_up(int x, int m) { return ((x + m - 1) / m) * m; }

}  // namespace mlp
