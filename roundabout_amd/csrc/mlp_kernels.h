// mlp_kernels.h -- kernels of the learned closure (BF16 MFMA MLP).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlp {

typedef uint16_t bf16;  // raw bfloat16 bits in memory

enum { EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_GELU = 2, EPI_BIAS_F32_T = 3, EPI_GELU_BWD = 4, EPI_MSE = 5 };

// Fused inference over all layers (k_mlp_fused): per-layer weights (bf16,
// pout x pin, k contiguous) and FP32 biases, padded widths.
constexpr int FP_MAXL = 8;
constexpr int FP_MAXK = 384;
constexpr int FP_MAXN = 256;  // widest layer output
constexpr int FP_WAVES = 8;   // wavefronts per workgroup
// LDS row pitches (bf16) of the activation buffers: rows start 4 banks
// apart.  FA_PITCH: a layer's input or output (up to 256 columns); FB_PITCH:
// the first layer's input columns 256..383, staged into the second buffer
constexpr int FA_PITCH = 256 + 8;
constexpr int FB_PITCH = 128 + 8;
static_assert(FB_PITCH <= FA_PITCH, "the first layer's columns 256.. fit the second buffer");
// dynamic LDS of k_mlp_fused: two activation buffers of 128 rows
constexpr size_t FP_LDS_BYTES = (size_t)2 * 128 * FA_PITCH * 2;
struct FusedNet {
  const bf16* W[FP_MAXL];
  const float* b[FP_MAXL];
  int pin[FP_MAXL];
  int pout[FP_MAXL];
  int nl;
};
__global__ void k_mlp_fused(FusedNet fn, const bf16* X0, int ldx, float* Y, long long ldy, int out_major);

// Epilogue operands of k_gemm_bf16 (the ones an EPI does not name are unused).
struct GemmEpi {
  const float* bias;       // per-column bias (EPI_BIAS_*, EPI_MSE)
  const float* pre;        // EPI_GELU_BWD: FP32 pre-activations (M x N, pitch ldp)
  int ldp;
  float* Cf;               // FP32 output: product, partial, pre-activation
  int ldcf;
  bf16* Cb;                // bf16 output, row-major (M x N, pitch ldcb)
  int ldcb;
  bf16* CbT;               // optional bf16 copy, transposed (N x M, pitch ldct)
  int ldct;
  float* colsum;           // EPI_GELU_BWD, EPI_MSE: column sums of the bf16
                           // output per 64-row block, (M / 64) x N
  int kc;                  // split-K chunk (EPI_F32)
  long long split_stride;  // distance between the split partials of Cf
  const unsigned* aidx;    // optional row gather of A: row m is A row aidx[m]
  int anv;                 // (m < anv; rows past it read A row 0)
  bf16* AT;                // optional copy of the (gathered) A, feature-major
  int ldat;                // (K x M, pitch ldat), written as it streams in
  // EPI_MSE: targets T and 0/1 weights W (rows tidx[m] if tidx is given,
  // pitch ldt), valid rows / columns; the tile's sums of w (y - t)^2 and of
  // w go to lpart[tile] and cpart[tile]
  const float* T;
  const float* W;
  const unsigned* tidx;
  int ldt, mvalid, nvalid;
  float* lpart;
  float* cpart;
};

// One product C = A B^T of k_gemm_bf16: M, N multiples of 64, K of 32;
// splits > 1 (EPI_F32 only) sums k-chunks of kc into partials split_stride
// apart.  Workgroups: (M / 64) (N / 64) splits, n-tiles fastest.
struct GemmArgs {
  const bf16* A;
  int lda;
  const bf16* B;
  int ldb;
  int M, N, K, splits;
  GemmEpi ep;
};

template <int EPI>
__global__ void k_gemm_bf16(GemmArgs g);
// two independent products in one launch: the first g1 blocks, then g2's
template <int E1, int E2>
__global__ void k_gemm_pair(GemmArgs g1, GemmArgs g2);
__global__ void k_transpose_bf16(const bf16* in, int ldi, bf16* out, int ldo, int R, int C);
__global__ void k_loss_final(const float* lpart, const float* cpart, int n, float* loss, float* count);

// Gradient reduction + Adam + BF16 refresh over every parameter in one launch.
struct GradPlan {
  int nl;
  int pin[FP_MAXL], pout[FP_MAXL];
  long long offW[FP_MAXL];   // W_l in the FP32 buffers; b_l follows at offW + pout * pin
  long long offWb[FP_MAXL];  // W_l in the bf16 copies
  const float* pw[FP_MAXL];  // weight-gradient partials, S x (pout * pin)
  const float* pb[FP_MAXL];  // bias-gradient partials, SB x pout
};
__global__ void k_grad_step(GradPlan gp, int S, int SB, long long np, float* p, float* g, float* m1, float* m2,
                            bf16* Wb, bf16* WTb, int update, float lr, float b1, float b2, float eps, float c1,
                            float c2, const float* lpart, const float* cpart, int nlp, float* loss, float* count);
__global__ void k_refresh_weights(const float* w, int out, int in, bf16* Wb, int ldw, bf16* WTb, int ldwt,
                                  int out_pad, int in_pad);
__global__ void k_to_bf16_padded(const float* x, int rows, int cols, bf16* y, int rows_pad, int ld);

}  // namespace mlp
