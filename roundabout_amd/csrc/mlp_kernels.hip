// mlp_kernels.hip -- BF16 MFMA kernels of the learned vertical-mixing closure:
// GEMM with fused epilogues, GELU backward, bias gradients, masked MSE loss,
// Adam, transposes and weight refresh.
//
// The GEMM computes C[m][n] = sum_k A[m][k] * B[n][k] with both operands
// row-major and k contiguous (A: M x K, B: N x K), FP32 accumulation on
// v_mfma_f32_16x16x32_bf16.  Operand layout of that instruction (checked on
// the device): lane l holds A[l % 16][(l / 16) * 8 + i] and
// B[(l / 16) * 8 + i][l % 16] for i = 0..7, and D[(l / 16) * 4 + j][l % 16]
// for j = 0..3.  One wavefront owns a 64 x 64 tile of C (4 x 4 MFMA tiles);
// the forward pass Y = X W^T, the input gradient dX = dY W and the weight
// gradient dW = dY^T X are all this product with the right transposed copy.
#include "mlp_kernels.h"

namespace mlp {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float bf2f(bf16 h) {
  uint32_t u = (uint32_t)h << 16;
  return __uint_as_float(u);
}
// round to nearest even (no NaN inputs reach it in this path)
__device__ __forceinline__ bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
__device__ __forceinline__ float gelu(float x) {
  const float c = 0.7978845608028654f;
  float u = c * (x + 0.044715f * x * x * x);
  return 0.5f * x * (1.0f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float c = 0.7978845608028654f;
  float x2 = x * x;
  float u = c * (x + 0.044715f * x2 * x);
  float t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * x2);
}

template <int EPI>
__global__ void __launch_bounds__(64) k_gemm_bf16(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B,
                                                  int ldb, int M, int N, int K, const float* __restrict__ bias,
                                                  float* __restrict__ Cf, int ldcf, bf16* __restrict__ Cb, int ldcb,
                                                  int kc, long long split_stride) {
  const int lane = threadIdx.x;
  // n-tiles fastest, so that the workgroups sharing a row block of A run
  // together and find it in L2
  const int nt = N / 64;
  const int m0 = (int)(blockIdx.x / nt) * 64, n0 = (int)(blockIdx.x % nt) * 64;
  if (m0 >= M || n0 >= N) return;
  // split-K: this workgroup sums k in [kbeg, kend) into its own partial
  const int kbeg = blockIdx.y * kc;
  const int kend = min(K, kbeg + kc);
  A += kbeg;
  B += kbeg;
  K = max(kend - kbeg, 0);
  if (EPI == EPI_F32) Cf += (size_t)blockIdx.y * (size_t)split_stride;
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16* Ap[4];
  const bf16* Bp[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    Ap[i] = A + (size_t)(m0 + 16 * i + r) * (size_t)lda + q * 8;
    Bp[i] = B + (size_t)(n0 + 16 * i + r) * (size_t)ldb + q * 8;
  }
  bf16x8 a_cur[4], b_cur[4];
  if (K > 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a_cur[i] = *(const bf16x8*)(Ap[i]);
      b_cur[i] = *(const bf16x8*)(Bp[i]);
    }
  }
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a_nxt[4], b_nxt[4];
    const bool more = (k0 + 32 < K);
    if (more) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a_nxt[i] = *(const bf16x8*)(Ap[i] + k0 + 32);
        b_nxt[i] = *(const bf16x8*)(Bp[i] + k0 + 32);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_cur[i], b_cur[j], acc[i][j], 0, 0, 0);
    if (more) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a_cur[i] = a_nxt[i];
        b_cur[i] = b_nxt[i];
      }
    }
  }
  // epilogue: tile (i, j) element (row 16 i + 4 q + e, col 16 j + r)
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + 16 * j + r;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = m0 + 16 * i + 4 * q + e;
        float v = acc[i][j][e];
        if (EPI == EPI_F32) {
          Cf[(size_t)m * ldcf + n] = v;
        } else if (EPI == EPI_BIAS_F32) {
          Cf[(size_t)m * ldcf + n] = v + bias[n];
        } else if (EPI == EPI_BIAS_F32_T) {  // output-major: Cf[n][m]
          Cf[(size_t)n * ldcf + m] = v + bias[n];
        } else {  // EPI_BIAS_GELU: bf16 activation, optional fp32 pre-activation
          float pre = v + bias[n];
          if (Cf) Cf[(size_t)m * ldcf + n] = pre;
          Cb[(size_t)m * ldcb + n] = f2bf(gelu(pre));
        }
      }
    }
}

template __global__ void k_gemm_bf16<EPI_F32>(const bf16*, int, const bf16*, int, int, int, int, const float*,
                                             float*, int, bf16*, int, int, long long);
template __global__ void k_gemm_bf16<EPI_BIAS_F32>(const bf16*, int, const bf16*, int, int, int, int, const float*,
                                                  float*, int, bf16*, int, int, long long);
template __global__ void k_gemm_bf16<EPI_BIAS_F32_T>(const bf16*, int, const bf16*, int, int, int, int, const float*,
                                                    float*, int, bf16*, int, int, long long);
template __global__ void k_gemm_bf16<EPI_BIAS_GELU>(const bf16*, int, const bf16*, int, int, int, int, const float*,
                                                   float*, int, bf16*, int, int, long long);

// out[c][r] = in[r][c] for an R x C bf16 matrix (row pitches ldi, ldo).
__global__ void k_transpose_bf16(const bf16* __restrict__ in, int ldi, bf16* __restrict__ out, int ldo, int R, int C) {
  __shared__ bf16 tile[32][33];
  int c = blockIdx.x * 32 + threadIdx.x;
  int r0 = blockIdx.y * 32;
  for (int dy = threadIdx.y; dy < 32; dy += blockDim.y) {
    int r = r0 + dy;
    if (r < R && c < C) tile[dy][threadIdx.x] = in[(size_t)r * ldi + c];
  }
  __syncthreads();
  int rr = r0 + threadIdx.x;
  for (int dy = threadIdx.y; dy < 32; dy += blockDim.y) {
    int cc = blockIdx.x * 32 + dy;
    if (cc < C && rr < R) out[(size_t)cc * ldo + rr] = tile[threadIdx.x][dy];
  }
}

// d = g * gelu'(pre) (FP32 in), written as bf16 (row-major, pitch ld).
__global__ void k_gelu_backward(const float* __restrict__ g, const float* __restrict__ pre, bf16* __restrict__ d,
                                long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d[i] = f2bf(g[i] * gelu_grad(pre[i]));
}

// Masked mean-squared error of the output layer and its gradient (bf16).
// Y, T, W (weights 0/1) are M x N (pitch ld); rows >= Mvalid contribute 0.
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ W, int M,
                      int Mvalid, int N, int ld, float inv_count, bf16* __restrict__ dY, float* __restrict__ loss) {
  __shared__ float red[256];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (long long)M * ld) {
    int m = (int)(i / ld), n = (int)(i % ld);
    float g = 0.f;
    if (m < Mvalid && n < N) {
      float w = W[(size_t)m * ld + n];
      float e = Y[(size_t)m * ld + n] - T[(size_t)m * ld + n];
      l = w * e * e;
      g = 2.f * w * e * inv_count;
    }
    dY[i] = f2bf(g);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv_count);
}

// Partial column sums of d (bf16, M x N, pitch ld) over the row chunk
// [blockIdx.y * rows, (blockIdx.y + 1) * rows): part[blockIdx.y * N + n].
// Block (64, 4): 64 columns, 4 row phases combined in a fixed order.
__global__ void k_bias_partial(const bf16* __restrict__ d, int M, int N, int ld, int rows, float* __restrict__ part) {
  __shared__ float red[4][64];
  const int n = blockIdx.x * 64 + threadIdx.x;
  const int r0 = blockIdx.y * rows, r1 = min(M, r0 + rows);
  float s = 0.f;
  if (n < N)
    for (int m = r0 + threadIdx.y; m < r1; m += 4) s += bf2f(d[(size_t)m * ld + n]);
  red[threadIdx.y][threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.y == 0 && n < N)
    part[(size_t)blockIdx.y * N + n] = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                                       (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// out[i] = sum over s < S of part[s * stride + i], in order of s
// (deterministic split reductions).
__global__ void k_reduce_splits(const float* __restrict__ part, int S, long long n, long long stride,
                                float* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a = 0.f;
  for (int s = 0; s < S; ++s) a += part[(size_t)s * (size_t)stride + (size_t)i];
  out[i] = a;
}

// Adam on FP32 master parameters (Kingma & Ba 2015, bias-corrected).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float a = b1 * m1[i] + (1.f - b1) * gi;
  float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 copies of one layer: W (out x in, padded rows/cols zero) and its
// transpose, from the FP32 master (out x in, unpadded).
__global__ void k_refresh_weights(const float* __restrict__ w, int out, int in, bf16* __restrict__ Wb, int ldw,
                                  bf16* __restrict__ WTb, int ldwt, int out_pad, int in_pad) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;  // input index
  int r = blockIdx.y;                             // output index
  if (c >= in_pad || r >= out_pad) return;
  float v = (r < out && c < in) ? w[(size_t)r * in + c] : 0.f;
  bf16 b = f2bf(v);
  Wb[(size_t)r * ldw + c] = b;
  WTb[(size_t)c * ldwt + r] = b;
}

// Pack FP32 (rows x cols) into a padded bf16 matrix (pitch ld), zero padding.
__global__ void k_to_bf16_padded(const float* __restrict__ x, int rows, int cols, bf16* __restrict__ y, int rows_pad,
                                 int ld) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows_pad * ld) return;
  int r = (int)(i / ld), c = (int)(i % ld);
  y[i] = f2bf((r < rows && c < cols) ? x[(size_t)r * cols + c] : 0.f);
}

}  // namespace mlp
