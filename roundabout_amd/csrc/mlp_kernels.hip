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

// GELU, tanh form: 0.5 x (1 + tanh(u)), u = sqrt(2/pi) (x + 0.044715 x^3),
// evaluated as x * sigma(2 u) with sigma(z) = 1 / (1 + e^-z) (the identity
// 0.5 (1 + tanh(u)) = sigma(2 u)): one hardware exp and one divide.  Large
// |u| saturates correctly (e^-2u -> inf gives -0, -> 0 gives x).
__device__ __forceinline__ float gelu_sigma(float x, float* u_prime) {
  const float c = 0.7978845608028654f;
  const float x2 = x * x;
  const float u = c * (x + 0.044715f * x2 * x);
  if (u_prime) *u_prime = c * (1.0f + 3.0f * 0.044715f * x2);
  return 1.0f / (1.0f + __expf(-2.0f * u));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sigma(x, nullptr); }
// d/dx [x sigma(2u)] = sigma + 2 x sigma (1 - sigma) u'
__device__ __forceinline__ float gelu_grad(float x) {
  float up;
  const float sg = gelu_sigma(x, &up);
  return sg + 2.0f * x * sg * (1.0f - sg) * up;
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

// s_waitcnt vmcnt(n) (n = 0, 1, 2, 3) without waiting on the LDS or export
// counters; the gfx9 immediate holds vmcnt in bits 3:0, expcnt in 6:4 and
// lgkmcnt in 11:8.
__device__ __forceinline__ void wait_vm_except(int n) {
  switch (n) {
    case 0: __builtin_amdgcn_s_waitcnt(0xF70); break;
    case 1: __builtin_amdgcn_s_waitcnt(0xF71); break;
    case 2: __builtin_amdgcn_s_waitcnt(0xF72); break;
    default: __builtin_amdgcn_s_waitcnt(0xF73); break;
  }
}

// One layer of the fused kernel for a wave owning rows [row0, row0 + 64) of
// the block and NTW 16-wide output tiles from column nw0:
// acc[i][j] (rows row0 + 16 i.., columns nw0 + 16 j..) = act * W^T.
//
// W streams through LDS in k-slices of 32 (N rows x 64 B) in 3 buffers, with the direct global-to-LDS load (global_load_lds_dwordx4:
// lane l of a wave writes its 16 bytes at the instruction's LDS base + 16 l),
// two slices ahead of the MFMAs that use them.  Within a
// slice, the 16-byte chunk q of row n sits at slot 4 n + ((q + n / 4) mod 4):
// the 16 rows of one MFMA operand then start in 16 different bank groups.
// NTW is a compile-time constant, so the loop body has no branches; the
// scheduling barriers keep each k-step as: fragment reads, MFMAs, wait.
template <int NTW>
__device__ __forceinline__ void fused_layer(const bf16* __restrict__ W, int K, int N, int wave, int lane, int row0,
                                            int nw0, const bf16* act, bf16* w0, bf16* w1, bf16* w2,
                                            f32x4 (&acc)[4][4]) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NTW; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // a slice is N / 16 row groups of 16 rows (1 KB of LDS, one instruction
  // per wave each); lane l loads row g * 16 + l / 4 into LDS slot l % 4,
  // which holds global chunk (l % 4 - l / 16) mod 4
  const int lrow = lane >> 2;
  const int lchunk = ((lane & 3) - (lane >> 4) + 4) & 3;
  const int groups = N / 16;
  auto load_slice = [&](int k0, bf16* buf) {
    for (int g = wave; g < groups; g += FP_WAVES) {
      const bf16* src = W + (size_t)(g * 16 + lrow) * (size_t)K + k0 + lchunk * 8;
      __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)(buf + g * 512), 16,
                                       0, 0);
    }
  };
  // this wave's loads per slice (the vmcnt a slice still in flight holds)
  const int per_wave = wave < groups ? (groups - wave + FP_WAVES - 1) / FP_WAVES : 0;
  const int nsteps = K / 32;
  const bf16* arow = act + (row0 + r) * FP_PITCH + q * 8;
  // k-step t: slice t from `cur`; slice t + 2 starts into `ahead_buf`
  auto step = [&](int t, const bf16* cur, bf16* ahead_buf) {
    const int k0 = 32 * t;
    const bool ahead = t + 2 < nsteps;
    if (ahead) load_slice(k0 + 64, ahead_buf);
    bf16x8 a[4], b[NTW];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *(const bf16x8*)(arow + 16 * i * FP_PITCH + k0);
#pragma unroll
    for (int j = 0; j < NTW; ++j) {
      const int n = nw0 + 16 * j + r;
      b[j] = *(const bf16x8*)(cur + (4 * n + ((q + (n >> 2)) & 3)) * 8);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NTW; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    // slice t + 1 has landed (slice t + 2 may still be in flight); the
    // barrier also frees `cur` for slice t + 3.  A bare s_barrier: the loop
    // writes no LDS itself, the reads of `cur` are complete (their values
    // fed the MFMAs) and the direct loads are covered by the wait above, so
    // the full fence of __syncthreads -- which would drain slice t + 2 as
    // well -- is not needed.
    wait_vm_except(ahead ? per_wave : 0);
    __builtin_amdgcn_s_barrier();
  };
  load_slice(0, w0);
  if (nsteps > 1) load_slice(32, w1);
  wait_vm_except(nsteps > 1 ? per_wave : 0);  // slice 0 has landed
  __syncthreads();
  // unrolled by the buffer rotation, so that every LDS access names its
  // buffer and the fragment reads need not wait for the loads in flight
  for (int t = 0; t < nsteps; t += 3) {
    step(t, w0, w2);
    if (t + 1 < nsteps) step(t + 1, w1, w0);
    if (t + 2 < nsteps) step(t + 2, w2, w1);
  }
}

// Epilogue of one layer: GELU(acc + b) to the block's LDS input (hidden
// layers), or acc + b to Y (the last layer).
template <int NTW>
__device__ __forceinline__ void fused_epilogue(const f32x4 (&acc)[4][4], const float* __restrict__ bias, int row0,
                                               int nw0, int r, int q, bf16* act, bool last, float* __restrict__ Y,
                                               long long ldy, long long m0, int out_major) {
#pragma unroll
  for (int j = 0; j < NTW; ++j) {
    const int n = nw0 + 16 * j + r;
    const float bn = bias[n];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int mr = row0 + 16 * i + 4 * q;  // rows mr .. mr + 3 of the block
      if (!last) {
#pragma unroll
        for (int e = 0; e < 4; ++e) act[(mr + e) * FP_PITCH + n] = f2bf(gelu(acc[i][j][e] + bn));
      } else if (out_major) {
        f32x4 v = acc[i][j];
        v[0] += bn;
        v[1] += bn;
        v[2] += bn;
        v[3] += bn;
        *(f32x4*)(Y + (size_t)n * (size_t)ldy + (size_t)(m0 + mr)) = v;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) Y[(size_t)(m0 + mr + e) * (size_t)ldy + n] = acc[i][j][e] + bn;
      }
    }
  }
}

// The whole network on a block of 128 rows, activations resident in LDS.
//
// Workgroup of 8 wavefronts: the block's layer input (128 x pin bf16, row
// pitch FP_PITCH) sits in LDS; wave w owns rows 64 (w / 4) .. + 64 and
// output columns [(w % 4) pout / 4, (w % 4 + 1) pout / 4), two waves per
// SIMD.  The inputs are staged with direct-to-LDS loads, all in flight at
// once; W streams through LDS (fused_layer).  After a layer's k-loop the
// block overwrites its LDS input with GELU(acc + b) in bf16; the last layer
// adds the bias and stores FP32 to Y, row-major (Y[m ldy + n]) or
// output-major (Y[n ldy + m]).  HBM sees only the input features and the
// outputs: no activation round trips between layers.
// Limits (checked by the host): pin <= FP_MAXK, pout <= FP_MAXN, all
// multiples of 64; rows a multiple of 128.
__global__ void __launch_bounds__(64 * FP_WAVES) k_mlp_fused(FusedNet fn, const bf16* __restrict__ X0, int ldx,
                                                             float* __restrict__ Y, long long ldy, int out_major) {
  extern __shared__ __align__(16) bf16 act[];  // 128 x FP_PITCH (dynamic LDS)
  // weight slices: separate objects from `act`, so that reading one never
  // has to wait for the direct loads into another
  __shared__ __align__(16) bf16 w0[FP_MAXN * 32], w1[FP_MAXN * 32], w2[FP_MAXN * 32];
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;  // wave index in a scalar register
  const int r = lane & 15, q = lane >> 4;
  const long long m0 = (long long)blockIdx.x * 128;
  {
    // the block's LDS image (128 rows x FP_PITCH) in 1 KB pieces of 64
    // 16-byte chunks; the pad chunks at the end of each row re-read the
    // row's first chunk (never used)
    constexpr int CPR = FP_PITCH / 8;  // chunks per LDS row
    constexpr int NCH = 128 * CPR;     // a multiple of 64
    for (int c0 = wave * 64; c0 < NCH; c0 += 64 * FP_WAVES) {
      const int c = c0 + lane;
      const int row = c / CPR, col = (c % CPR) * 8;
      const bf16* src = X0 + (size_t)(m0 + row) * (size_t)ldx + (col < ldx ? col : 0);
      __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)(act + c0 * 8), 16,
                                       0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
  }
  __syncthreads();
  const int row0 = 64 * (wave >> 2);
  for (int l = 0; l < fn.nl; ++l) {
    const int K = fn.pin[l], N = fn.pout[l];
    const int ntw = N / 64;                // 16-wide n-tiles per wave (1..4)
    const int nw0 = (wave & 3) * (N / 4);  // this wave's first output column
    const bool last = (l == fn.nl - 1);
    f32x4 acc[4][4];
    switch (ntw) {
      case 4: fused_layer<4>(fn.W[l], K, N, wave, lane, row0, nw0, act, w0, w1, w2, acc); break;
      case 3: fused_layer<3>(fn.W[l], K, N, wave, lane, row0, nw0, act, w0, w1, w2, acc); break;
      case 2: fused_layer<2>(fn.W[l], K, N, wave, lane, row0, nw0, act, w0, w1, w2, acc); break;
      default: fused_layer<1>(fn.W[l], K, N, wave, lane, row0, nw0, act, w0, w1, w2, acc); break;
    }
    // (the layer ends with a barrier: every wave has read this layer's input)
    switch (ntw) {
      case 4: fused_epilogue<4>(acc, fn.b[l], row0, nw0, r, q, act, last, Y, ldy, m0, out_major); break;
      case 3: fused_epilogue<3>(acc, fn.b[l], row0, nw0, r, q, act, last, Y, ldy, m0, out_major); break;
      case 2: fused_epilogue<2>(acc, fn.b[l], row0, nw0, r, q, act, last, Y, ldy, m0, out_major); break;
      default: fused_epilogue<1>(acc, fn.b[l], row0, nw0, r, q, act, last, Y, ldy, m0, out_major); break;
    }
    __syncthreads();  // the next layer's input is complete
  }
}

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
// Masked mean-squared error of the output layer and its gradient (bf16):
// grid-stride over the M x N elements (pitch ld; rows >= Mvalid and
// columns >= N give 0); each workgroup writes its partial loss sum to
// part[blockIdx.x], which k_loss_final adds up in a fixed order (no
// atomics: the loss is reproducible).
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ W, int M,
                      int Mvalid, int N, int ld, float inv_count, bf16* __restrict__ dY, float* __restrict__ part) {
  __shared__ float red[256];
  const long long total = (long long)M * ld, stride = (long long)gridDim.x * blockDim.x;
  float l = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const int m = (int)(i / ld), n = (int)(i % ld);
    float g = 0.f;
    if (m < Mvalid && n < N) {
      const float w = W[i];
      const float e = Y[i] - T[i];
      l += w * e * e;
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
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// *loss += inv_count * sum of part[0..n) (one workgroup of 256, fixed order)
__global__ void k_loss_final(const float* __restrict__ part, int n, float inv_count, float* __restrict__ loss) {
  __shared__ float red[256];
  float a = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a += part[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss += red[0] * inv_count;
}

// Partial column sums of d (bf16, M x N, pitch ld) over the row chunk
// [blockIdx.y * rows, (blockIdx.y + 1) * rows): part[blockIdx.y * N + n].
// Block (64, BP_PHASES): 64 columns, BP_PHASES row phases combined in a
// fixed order.
__global__ void k_bias_partial(const bf16* __restrict__ d, int M, int N, int ld, int rows, float* __restrict__ part) {
  __shared__ float red[BP_PHASES][64];
  const int n = blockIdx.x * 64 + threadIdx.x;
  const int r0 = blockIdx.y * rows, r1 = min(M, r0 + rows);
  float s = 0.f;
  if (n < N)
    for (int m = r0 + threadIdx.y; m < r1; m += BP_PHASES) s += bf2f(d[(size_t)m * ld + n]);
  red[threadIdx.y][threadIdx.x] = s;
  __syncthreads();
  for (int w = BP_PHASES / 2; w > 0; w >>= 1) {
    if (threadIdx.y < w) red[threadIdx.y][threadIdx.x] += red[threadIdx.y + w][threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.y == 0 && n < N) part[(size_t)blockIdx.y * N + n] = red[0][threadIdx.x];
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
