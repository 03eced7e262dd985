// mlp_api.hip -- stateless C entry points of the BF16 MFMA MLP (forward and
// loss gradient on host arrays).  They have no counterpart in the reference;
// they exist so that the network kernels can be checked against an FP64 CPU
// oracle on small batches.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/rdb_colphys.h"
#include "mlp_net.h"

extern int rdb_colphys_push_error(int status, const char* msg);

namespace {

int build_net(mlp::Net& net, const int* widths, int n_widths, const float* params, hipStream_t s) {
  if (!widths || n_widths < 2 || n_widths > 16 || !params)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp: bad widths or NULL params");
  std::vector<int> dims(widths, widths + n_widths);
  int st = net.init(dims, s);
  if (st) return rdb_colphys_push_error(st, net.err.c_str());
  st = net.set_params(params);
  if (st) return rdb_colphys_push_error(st, net.err.c_str());
  return 0;
}

int upload_x(mlp::Net& net, int batch, int Mpad, const float* x) {
  const int in = net.din[0], ld = net.pin[0];
  std::vector<uint16_t> xb((size_t)Mpad * ld, 0);
  for (int m = 0; m < batch; ++m)
    for (int c = 0; c < in; ++c) {
      float f = x[(size_t)m * in + c];
      uint32_t u;
      std::memcpy(&u, &f, 4);
      u += 0x7fffu + ((u >> 16) & 1u);
      xb[(size_t)m * ld + c] = (uint16_t)(u >> 16);
    }
  // synchronised: the staging vector goes out of scope on return
  if (hipMemcpyAsync(net.X0, xb.data(), 2 * xb.size(), hipMemcpyHostToDevice, net.stream) != hipSuccess ||
      hipStreamSynchronize(net.stream) != hipSuccess)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp: input upload failed");
  return 0;
}

// targets and 0/1 weights (batch x out, row-major) into the padded T, Wm;
// synchronises, so the host staging may go out of scope
int upload_targets(mlp::Net& net, int batch, int Mpad, const float* t, const float* w) {
  const int out = net.dout[net.nl - 1], po = net.pout[net.nl - 1];
  std::vector<float> tp((size_t)Mpad * po, 0.f), wp((size_t)Mpad * po, 0.f);
  for (int m = 0; m < batch; ++m)
    for (int c = 0; c < out; ++c) {
      tp[(size_t)m * po + c] = t[(size_t)m * out + c];
      wp[(size_t)m * po + c] = w[(size_t)m * out + c];
    }
  if (hipMemcpyAsync(net.T, tp.data(), 4 * tp.size(), hipMemcpyHostToDevice, net.stream) != hipSuccess ||
      hipMemcpyAsync(net.Wm, wp.data(), 4 * wp.size(), hipMemcpyHostToDevice, net.stream) != hipSuccess ||
      hipStreamSynchronize(net.stream) != hipSuccess)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp: target upload failed");
  return 0;
}

}  // namespace

extern "C" {

int rdb_mlp_eval(const int* widths, int n_widths, const float* params, int batch, const float* x, float* y) {
  if (batch < 1 || batch > (1 << 22) || !x || !y)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "mlp_eval: batch outside [1, 2^22] or NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_eval: no HIP device available");
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_eval: stream create failed");
  int st;
  {
    mlp::Net net;
    st = build_net(net, widths, n_widths, params, s);
    const int Mpad = mlp::round_up(batch, 128);  // the fused inference kernel's row block
    if (!st && net.ensure_workspace(Mpad, false)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    if (!st) st = upload_x(net, batch, Mpad, x);
    if (!st && net.forward(Mpad)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    if (!st) {
      const int out = net.dout[net.nl - 1], po = net.pout[net.nl - 1];
      std::vector<float> yp((size_t)Mpad * po);
      if (hipMemcpyAsync(yp.data(), net.Y, 4 * yp.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_eval: output download failed");
      else
        for (int m = 0; m < batch; ++m) std::memcpy(y + (size_t)m * out, yp.data() + (size_t)m * po, 4 * (size_t)out);
    }
    (void)hipStreamSynchronize(s);
  }
  (void)hipStreamDestroy(s);
  return st;
}

int rdb_mlp_eval_t(const int* widths, int n_widths, const float* params, int batch, const float* x, float* yt) {
  if (batch < 1 || batch > (1 << 22) || !x || !yt)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "mlp_eval_t: batch outside [1, 2^22] or NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_eval_t: no HIP device available");
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_eval_t: stream create failed");
  int st;
  {
    mlp::Net net;
    st = build_net(net, widths, n_widths, params, s);
    const int Mpad = mlp::round_up(batch, 128);
    if (!st && net.ensure_workspace(Mpad, false)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    if (!st) st = upload_x(net, batch, Mpad, x);
    // the surrogate's layout: output n of row m at Y[n * Mpad + m]
    if (!st && net.forward(Mpad, true)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    if (!st) {
      const int out = net.dout[net.nl - 1], po = net.pout[net.nl - 1];
      std::vector<float> yp((size_t)Mpad * po);
      if (hipMemcpyAsync(yp.data(), net.Y, 4 * yp.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_eval_t: output download failed");
      else
        for (int n = 0; n < out; ++n)
          std::memcpy(yt + (size_t)n * batch, yp.data() + (size_t)n * Mpad, 4 * (size_t)batch);
    }
    (void)hipStreamSynchronize(s);
  }
  (void)hipStreamDestroy(s);
  return st;
}

int rdb_mlp_adam_steps(const int* widths, int n_widths, const float* params, int batch, const float* x,
                       const float* t, const float* w, int n_steps, double lr, double beta1, double beta2, double eps,
                       float* params_out) {
  if (batch < 1 || batch > (1 << 22) || !x || !t || !w || !params_out || n_steps < 0)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "mlp_adam_steps: bad batch or step count, or NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_adam_steps: no HIP device available");
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_adam_steps: stream create failed");
  int st;
  {
    mlp::Net net;
    st = build_net(net, widths, n_widths, params, s);
    const int Mpad = mlp::round_up(batch, 64);
    if (!st && net.ensure_workspace(Mpad, true)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    if (!st) st = upload_x(net, batch, Mpad, x);
    if (!st) st = upload_targets(net, batch, Mpad, t, w);
    if (!st && net.transpose_input(Mpad)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    // the training step of the learned closure, repeated on one batch:
    // forward keeping the activations, backward, gradient reduction + Adam +
    // BF16 refresh of W and W^T
    for (int i = 0; !st && i < n_steps; ++i)
      if (net.train_forward(Mpad, batch) || net.backward(Mpad) ||
          net.update(true, (float)lr, (float)beta1, (float)beta2, (float)eps, net.dloss, nullptr))
        st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    if (!st && net.get_params(params_out)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    (void)hipStreamSynchronize(s);
  }
  (void)hipStreamDestroy(s);
  return st;
}

int rdb_mlp_loss_grad(const int* widths, int n_widths, const float* params, int batch, const float* x,
                      const float* t, const float* w, float* loss, float* grad) {
  if (batch < 1 || batch > (1 << 22) || !x || !t || !w || !loss || !grad)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "mlp_loss_grad: batch outside [1, 2^22] or NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_loss_grad: no HIP device available");
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_loss_grad: stream create failed");
  int st;
  {
    mlp::Net net;
    st = build_net(net, widths, n_widths, params, s);
    const int Mpad = mlp::round_up(batch, 64);
    if (!st && net.ensure_workspace(Mpad, true)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    if (!st) st = upload_x(net, batch, Mpad, x);
    if (!st) st = upload_targets(net, batch, Mpad, t, w);
    if (!st) {
      if (hipMemsetAsync(net.dloss, 0, 4, s) != hipSuccess)
        st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_loss_grad: loss clear failed");
      if (!st && (net.transpose_input(Mpad) || net.train_forward(Mpad, batch) || net.backward(Mpad) ||
                  net.update(false, 0.f, 0.f, 0.f, 0.f, net.dloss, nullptr)))
        st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
      if (!st && (hipMemcpyAsync(loss, net.dloss, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
                  hipStreamSynchronize(s) != hipSuccess))
        st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_loss_grad: loss download failed");
      if (!st && net.get_grads(grad)) st = rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, net.err.c_str());
    }
    (void)hipStreamSynchronize(s);
  }
  (void)hipStreamDestroy(s);
  return st;
}

}  // extern "C"
