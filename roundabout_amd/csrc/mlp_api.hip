// mlp_api.hip -- stateless C entry points of the BF16 MFMA MLP (forward and
// loss gradient on host arrays).  They have no counterpart in the reference;
// they exist so that the network kernels can be checked against an FP64 CPU
// oracle on small batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
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

// ---- the allocation-free inference entry point --------------------------
// Its device memory is one file-scope array per GPU, held from library load
// to process end.  Host code never passes the array's name to a kernel or a
// copy (on the host it is only a stand-in symbol): it uses the device
// address hipGetSymbolAddress returns, looked up once per device with that
// device current and kept in g_scratch_ptr.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)80 << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES];

int scratch_of_current_device(unsigned char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_eval_noalloc: no current HIP device");
  if (dev >= NOALLOC_MAX_DEVICES)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, "mlp_eval_noalloc: device index above 63");
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      char buf[160];
      snprintf(buf, sizeof buf, "mlp_eval_noalloc: scratch address lookup failed: %s", hipGetErrorString(e));
      return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, buf);
    }
    g_scratch_ptr[dev] = (unsigned char*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

__device__ __forceinline__ mlp::bf16 to_bf16(float f) { return __builtin_bit_cast(mlp::bf16, (__bf16)f); }

// Layer l of the public parameter layout (W_l out x in row-major, then b_l)
// into the padded bf16 weights (out_pad x in_pad) and FP32 bias (out_pad)
// the inference chain reads; padding is written as zeros.
__global__ void k_pack_layer(const float* __restrict__ w, const float* __restrict__ b, int out, int in,
                             mlp::bf16* __restrict__ Wb, float* __restrict__ bp, int out_pad, int in_pad) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (r >= out_pad) return;
  if (c < in_pad) Wb[(size_t)r * in_pad + c] = to_bf16((r < out && c < in) ? w[(size_t)r * in + c] : 0.f);
  if (c == 0) bp[r] = r < out ? b[r] : 0.f;
}

// rows x cols of a row-major matrix of pitch ldy into a dense one
__global__ void k_unpad_rows(const float* __restrict__ Y, int ldy, int rows, int cols, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  out[i] = Y[(size_t)r * ldy + c];
}

// Byte offsets of everything the entry point keeps in the scratch array.
struct NoAllocPlan {
  int nl, Mpad;
  int din[mlp::LV_MAXL], dout[mlp::LV_MAXL], pin[mlp::LV_MAXL], pout[mlp::LV_MAXL];
  size_t pub_w[mlp::LV_MAXL], pub_b[mlp::LV_MAXL];  // float offsets into the public parameters
  size_t o_pub, o_wb[mlp::LV_MAXL], o_b[mlp::LV_MAXL], o_x0, o_act0, o_act1, o_y, o_io, total;
};

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

int plan_noalloc(const int* widths, int n_widths, int batch, NoAllocPlan* P) {
  if (!widths || n_widths < 2 || n_widths > mlp::LV_MAXL + 1)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp_eval_noalloc: 2..17 widths expected");
  P->nl = n_widths - 1;
  P->Mpad = mlp::round_up(batch, 128);  // the fused inference kernel's row block, as rdb_mlp_eval
  size_t q = 0, off = 0;
  for (int l = 0; l < P->nl; ++l) {
    // (a padded output width is a launch's grid y extent: at most 65535)
    if (widths[l] < 1 || widths[l + 1] < 1 || widths[l] > 16384 || widths[l + 1] > 16384)
      return rdb_colphys_push_error(RDB_COLPHYS_ERR_CONFIG_VALIDATE, "mlp_eval_noalloc: widths outside [1, 16384]");
    P->din[l] = widths[l];
    P->dout[l] = widths[l + 1];
    P->pin[l] = mlp::round_up(widths[l], 64);
    P->pout[l] = mlp::round_up(widths[l + 1], 64);
    P->pub_w[l] = q;
    q += (size_t)P->dout[l] * P->din[l];
    P->pub_b[l] = q;
    q += (size_t)P->dout[l];
  }
  const size_t M = (size_t)P->Mpad;
  P->o_pub = off;
  off = a256(off + 4 * q);
  for (int l = 0; l < P->nl; ++l) {
    P->o_wb[l] = off;
    off = a256(off + 2 * (size_t)P->pout[l] * P->pin[l]);
    P->o_b[l] = off;
    off = a256(off + 4 * (size_t)P->pout[l]);
  }
  size_t amax = 0;
  for (int l = 0; l + 1 < P->nl; ++l) amax = std::max(amax, (size_t)P->pout[l]);
  P->o_x0 = off;
  off = a256(off + 2 * M * P->pin[0]);
  P->o_act0 = off;
  off = a256(off + 2 * M * amax);
  P->o_act1 = off;
  off = a256(off + 2 * M * amax);
  P->o_y = off;
  off = a256(off + 4 * M * P->pout[P->nl - 1]);
  P->o_io = off;  // host input copy, later the dense output
  off = a256(off + 4 * (size_t)batch * std::max(P->din[0], P->dout[P->nl - 1]));
  P->total = off;
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

int rdb_mlp_eval_noalloc(const int* widths, int n_widths, const float* params, int batch, const float* x,
                         float* y) {
  if (batch < 1 || batch > (1 << 22) || !params || !x || !y)
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, "mlp_eval_noalloc: batch outside [1, 2^22] or NULL");
  NoAllocPlan P;
  int st = plan_noalloc(widths, n_widths, batch, &P);
  if (st) return st;
  if (P.total > NOALLOC_SCRATCH_BYTES) {
    char buf[200];
    snprintf(buf, sizeof buf, "mlp_eval_noalloc: batch %d of this network needs %zu bytes of scratch, above the %zu-byte maximum",
             batch, P.total, NOALLOC_SCRATCH_BYTES);
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_BAD_SHAPE, buf);
  }
  unsigned char* base = nullptr;
  st = scratch_of_current_device(&base);
  if (st) return st;
  // everything on the null stream: the last copy waits for the kernels
  const hipStream_t s = 0;
  const int nl = P.nl, in = P.din[0], out = P.dout[nl - 1];
  const size_t npub = P.pub_b[nl - 1] + (size_t)P.dout[nl - 1];
  float* pub = (float*)(base + P.o_pub);
  float* io = (float*)(base + P.o_io);
  mlp::bf16* X0 = (mlp::bf16*)(base + P.o_x0);
  float* Y = (float*)(base + P.o_y);
  mlp::LayerView v{};
  v.nl = nl;
  for (int l = 0; l < nl; ++l) {
    v.pin[l] = P.pin[l];
    v.pout[l] = P.pout[l];
    v.W[l] = (const mlp::bf16*)(base + P.o_wb[l]);
    v.b[l] = (const float*)(base + P.o_b[l]);
    v.act[l] = (mlp::bf16*)(base + ((l & 1) ? P.o_act1 : P.o_act0));
  }
  hipError_t e = hipMemcpy(pub, params, 4 * npub, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(io, x, 4 * (size_t)batch * in, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    for (int l = 0; l < nl; ++l)
      hipLaunchKernelGGL(k_pack_layer, dim3((unsigned)((P.pin[l] + 255) / 256), (unsigned)P.pout[l]), dim3(256), 0, s,
                         (const float*)(pub + P.pub_w[l]), (const float*)(pub + P.pub_b[l]), P.dout[l], P.din[l],
                         (mlp::bf16*)(base + P.o_wb[l]), (float*)(base + P.o_b[l]), P.pout[l], P.pin[l]);
    const long long nx0 = (long long)P.Mpad * P.pin[0];
    hipLaunchKernelGGL(mlp::k_to_bf16_padded, dim3((unsigned)((nx0 + 255) / 256)), dim3(256), 0, s, (const float*)io,
                       batch, in, X0, P.Mpad, P.pin[0]);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = mlp::forward_view(v, X0, Y, P.Mpad, false, s);
  if (e == hipSuccess) {
    const long long ny = (long long)batch * out;
    hipLaunchKernelGGL(k_unpad_rows, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, s, (const float*)Y,
                       P.pout[nl - 1], batch, out, io);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(y, io, 4 * (size_t)batch * out, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    char buf[160];
    snprintf(buf, sizeof buf, "mlp_eval_noalloc: %s", hipGetErrorString(e));
    return rdb_colphys_push_error(RDB_COLPHYS_ERR_SETUP, buf);
  }
  return RDB_COLPHYS_OK;
}

}  // extern "C"
