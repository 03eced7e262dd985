// mlp_kernels.h -- kernels of the learned closure (BF16 MFMA MLP).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlp {

typedef uint16_t bf16;  // raw bfloat16 bits in memory

enum { EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_GELU = 2, EPI_BIAS_F32_T = 3 };

// Fused inference over all layers (k_mlp_fused): per-layer weights (bf16,
// pout x pin, k contiguous) and FP32 biases, padded widths.
constexpr int FP_MAXL = 8;
constexpr int FP_MAXK = 384;
constexpr int FP_PITCH = FP_MAXK + 8;  // LDS row pitch (bf16): rows start 4 banks apart
constexpr int FP_MAXN = 256;           // widest layer output
constexpr int FP_WAVES = 8;            // wavefronts per workgroup
// dynamic LDS of k_mlp_fused (the activation tile); three weight slices of
// FP_MAXN x 32 bf16 (48 KB) are static LDS on top
constexpr size_t FP_LDS_BYTES = (size_t)128 * FP_PITCH * 2;
struct FusedNet {
  const bf16* W[FP_MAXL];
  const float* b[FP_MAXL];
  int pin[FP_MAXL];
  int pout[FP_MAXL];
  int nl;
};
__global__ void k_mlp_fused(FusedNet fn, const bf16* X0, int ldx, float* Y, long long ldy, int out_major);

template <int EPI>
__global__ void k_gemm_bf16(const bf16* A, int lda, const bf16* B, int ldb, int M, int N, int K, const float* bias,
                            float* Cf, int ldcf, bf16* Cb, int ldcb, int kc, long long split_stride);
__global__ void k_transpose_bf16(const bf16* in, int ldi, bf16* out, int ldo, int R, int C);
__global__ void k_gelu_backward(const float* g, const float* pre, bf16* d, long long n);
constexpr int MSE_BLOCKS = 512;  // workgroups of k_mse (grid-stride), one loss partial each
constexpr int BP_PHASES = 16;    // row phases of k_bias_partial
__global__ void k_mse(const float* Y, const float* T, const float* W, int M, int Mvalid, int N, int ld,
                      float inv_count, bf16* dY, float* part);
__global__ void k_loss_final(const float* part, int n, float inv_count, float* loss);
__global__ void k_bias_partial(const bf16* d, int M, int N, int ld, int rows, float* part);
__global__ void k_reduce_splits(const float* part, int S, long long n, long long stride, float* out);
__global__ void k_adam(float* p, const float* g, float* m1, float* m2, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2);
__global__ void k_refresh_weights(const float* w, int out, int in, bf16* Wb, int ldw, bf16* WTb, int ldwt,
                                  int out_pad, int in_pad);
__global__ void k_to_bf16_padded(const float* x, int rows, int cols, bf16* y, int rows_pad, int ld);

}  // namespace mlp
