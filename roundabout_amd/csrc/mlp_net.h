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

// The inference chain on buffers the caller owns: padded bf16 weights
// (pout x pin, k contiguous), padded FP32 biases, and for the per-layer path
// one activation buffer per hidden layer (Mpad x pout[l] bf16; layer l only
// reads act[l - 1], so two buffers used alternately suffice).
constexpr int LV_MAXL = 16;
struct LayerView {
  int nl;
  int pin[LV_MAXL], pout[LV_MAXL];
  const bf16* W[LV_MAXL];
  const float* b[LV_MAXL];
  bf16* act[LV_MAXL];
};
// Whether forward_view runs the fused single-launch kernel for this shape.
bool fused_shape_ok(const LayerView& v, int Mpad);
// X0 (Mpad x pin[0] bf16) -> Y, row-major (Mpad x pout[nl-1]) or, with
// out_major, output-major (pout[nl-1] x Mpad); launches on s, allocates
// nothing.
hipError_t forward_view(const LayerView& v, const bf16* X0, float* Y, int Mpad, bool out_major, hipStream_t s);
// What forward_view sets up on first use on a device (the fused kernel's
// dynamic LDS allowance); call before capturing a forward into a graph.
hipError_t prepare_forward();

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
  bf16* dY = nullptr;      // gradient at the current layer's output, row-major
  bf16* dYT = nullptr;     // the same, feature-major (dW operand)
  bf16* dpre = nullptr;    // next (lower) layer's gradient, row-major
  bf16* dpreT = nullptr;   // and feature-major
  float* dloss = nullptr;
  float* lpart = nullptr;  // per-tile sums of w (y - t)^2 of the output layer
  float* cpart = nullptr;  // per-tile sums of w
  float* wsw = nullptr;    // weight-gradient partials: S x (pout * pin) per layer
  float* wsb = nullptr;    // bias-gradient partials: SB x pout per layer
  std::vector<size_t> offPW, offPB;  // layer offsets into wsw / wsb
  int bw_S = 0, bw_SB = 0, bw_nlp = 0;  // partials of the last training forward / backward pass
  hipStream_t stream = nullptr;
  std::string err;
  // Batch source of train_forward(): by default X0, T and Wm.
  // A caller that holds the whole training set on the device points the
  // first layer and the loss at it instead, with batch row m taken from set
  // row idx[m] (m < nvalid); the first layer then writes X0T itself.
  struct Source {
    const bf16* X = nullptr;  // rows of pin[0] bf16
    const float* T = nullptr;  // rows of pout[last] floats
    const float* W = nullptr;
    const unsigned* idx = nullptr;
    int nvalid = 0;
  } src;

  ~Net() { release(); }
  int init(const std::vector<int>& dims, hipStream_t s);
  int ensure_workspace(int Mpad, bool training);
  int set_params(const float* host);   // unpadded public layout
  int get_params(float* host) const;
  int get_grads(float* host) const;
  size_t n_public() const;
  // Inference: X0 (Mpad x pin[0] bf16) must hold the inputs; Y receives the
  // outputs, row-major (Mpad x pout) or, with out_major, output-major (pout
  // x Mpad).  Mpad a multiple of 128 and widths within the fused kernel's
  // limits run as one fused launch (k_mlp_fused), else per-layer GEMMs.
  int forward(int Mpad, bool out_major = false);
  bool fused_ok(int Mpad) const;
  LayerView view() const;  // the network's own weights and workspace
  // X0 -> X0T (feature-major copy of the inputs) for callers that fill X0
  // themselves.
  int transpose_input(int Mpad);
  // Training.  train_forward() runs the layers keeping what the backward
  // pass needs; the output layer's epilogue forms the masked MSE (targets
  // and 0/1 weights from T, Wm or the source set; rows >= Mvalid ignored;
  // the mean over sum(w) of the batch) as per-tile partials together with
  // its gradient.  backward() leaves the parameter gradient as split
  // partials; update() reduces and normalises them (into g), takes an Adam
  // step when adam_step is set, and adds the batch loss to *loss_dev and
  // sum(w) to *count_dev (device; count_dev may be null); add_loss() adds
  // the loss alone (validation).
  int train_forward(int Mpad, int Mvalid);
  int backward(int Mpad);
  int update(bool adam_step, float lr, float b1, float b2, float eps, float* loss_dev, float* count_dev);
  int add_loss(float* loss_dev, float* count_dev);
  void release();

 private:
  void free_workspace();
  int fail(const char* what, hipError_t e);
};

inline int round_up(int x, int m) { return ((x + m - 1) / m) * m; }

}  // namespace mlp
