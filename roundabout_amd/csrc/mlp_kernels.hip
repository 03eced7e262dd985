// mlp_kernels.hip -- BF16 MFMA kernels of the learned vertical-mixing closure:
// GEMM with fused epilogues (bias, GELU, GELU backward with bias-gradient
// partials, transposed copies), masked MSE loss, the fused gradient
// reduction + Adam + weight refresh, and the fused inference kernel.
//
// The GEMM computes C[m][n] = sum_k A[m][k] * B[n][k] with both operands
// row-major and k contiguous (A: M x K, B: N x K), FP32 accumulation on
// v_mfma_f32_16x16x32_bf16.  Operand layout of that instruction (checked on
// the device): lane l holds A[l % 16][(l / 16) * 8 + i] and
// B[(l / 16) * 8 + i][l % 16] for i = 0..7, and D[(l / 16) * 4 + j][l % 16]
// for j = 0..3.  One wavefront owns a 64 x 64 tile of C (4 x 4 MFMA tiles);
// the forward pass Y = X W^T, the input gradient dX = dY W and the weight
// gradient dW = dY^T X are all this product with the right transposed copy.
// The transposed copies the weight gradient needs (activations and output
// gradients, feature-major) are written by the epilogues that produce them.
#include "mlp_kernels.h"

namespace mlp {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float bf2f(bf16 h) {
  uint32_t u = (uint32_t)h << 16;
  return __uint_as_float(u);
}
// round to nearest even: gfx950's v_cvt_pk_bf16_f32 (one instruction; the
// integer form u += 0x7fff + ((u >> 16) & 1) gives the same bits for the
// finite values of this path and took four)
__device__ __forceinline__ bf16 f2bf(float f) { return __builtin_bit_cast(bf16, (__bf16)f); }

// GELU, tanh form: 0.5 x (1 + tanh(u)), u = sqrt(2/pi) (x + 0.044715 x^3),
// evaluated as x * sigma(2 u) with sigma(z) = 1 / (1 + e^-z) (the identity
// 0.5 (1 + tanh(u)) = sigma(2 u)).  e^-2u = 2^(x (k1 + k2 x^2)) with the
// constants -2 log2(e) sqrt(2/pi) and its 0.044715 multiple folded: one
// multiply, one fma and one multiply form the exponent (the unfolded form
// took six operations, every one a separate rounding without contraction),
// then one hardware exp2 and one hardware reciprocal (1 ulp; the result is
// rounded to BF16).  Large |u| saturates correctly (2^+big -> inf gives
// rcp(inf) = 0, 2^-big -> 0 gives x).
__device__ __forceinline__ float gelu_sigma(float x, float* u_prime) {
  constexpr double c = 0.7978845608028654;
  constexpr float k1 = (float)(-2.0 * 1.4426950408889634 * c);
  constexpr float k2 = (float)(-2.0 * 1.4426950408889634 * c * 0.044715);
  const float x2 = x * x;
  if (u_prime) *u_prime = (float)c * (1.0f + 3.0f * 0.044715f * x2);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(x * __builtin_fmaf(k2, x2, k1)));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sigma(x, nullptr); }
// d/dx [x sigma(2u)] = sigma + 2 x sigma (1 - sigma) u'
__device__ __forceinline__ float gelu_grad(float x) {
  float up;
  const float sg = gelu_sigma(x, &up);
  return sg + 2.0f * x * sg * (1.0f - sg) * up;
}

// s_waitcnt vmcnt(n) without waiting on the LDS or export counters; the gfx9
// immediate holds vmcnt[3:0] in bits 3:0, expcnt in 6:4, lgkmcnt in 11:8
// and vmcnt[5:4] in bits 15:14.
__device__ __forceinline__ void wait_vm_except(int n) {
  switch (n) {
    case 0: __builtin_amdgcn_s_waitcnt(0xF70); break;
    case 1: __builtin_amdgcn_s_waitcnt(0xF71); break;
    case 2: __builtin_amdgcn_s_waitcnt(0xF72); break;
    case 3: __builtin_amdgcn_s_waitcnt(0xF73); break;
    case 4: __builtin_amdgcn_s_waitcnt(0xF74); break;
    case 5: __builtin_amdgcn_s_waitcnt(0xF75); break;
    default: __builtin_amdgcn_s_waitcnt(0xF76); break;  // 6
  }
}

// Four 16-byte LDS reads (MFMA fragments) and the wait for them, as one asm
// statement.  Written as plain loads, the compiler's wait analysis could not
// tell these reads from the direct loads in flight into the other stages
// and waited for every outstanding load (a full memory round trip) before
// most k-steps; here the reads are ordered by the explicit vmcnt waits and
// barriers of the k-loop, and their own completion by lgkmcnt(0).
__device__ __forceinline__ void lds_read4(const bf16* p0, const bf16* p1, const bf16* p2, const bf16* p3,
                                          bf16x8& x0, bf16x8& x1, bf16x8& x2, bf16x8& x3) {
  typedef __attribute__((ext_vector_type(4))) unsigned u32x4;
  auto lds = [](const bf16* p) {
    return (unsigned)(uintptr_t)(const __attribute__((address_space(3))) bf16*)p;
  };
  u32x4 v0, v1, v2, v3;
  asm volatile(
      "ds_read_b128 %0, %4\n\t"
      "ds_read_b128 %1, %5\n\t"
      "ds_read_b128 %2, %6\n\t"
      "ds_read_b128 %3, %7\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3)
      : "v"(lds(p0)), "v"(lds(p1)), "v"(lds(p2)), "v"(lds(p3))
      : "memory");
  x0 = __builtin_bit_cast(bf16x8, v0);
  x1 = __builtin_bit_cast(bf16x8, v1);
  x2 = __builtin_bit_cast(bf16x8, v2);
  x3 = __builtin_bit_cast(bf16x8, v3);
}

// Eight 2-byte LDS reads and the wait for them, as one asm statement (for
// the same reason as lds_read4); returns the values packed low half first.
__device__ __forceinline__ uint4 lds_read8_u16(const unsigned (&ad)[8]) {
  unsigned v[8];
  asm volatile(
      "ds_read_u16 %0, %8\n\t"
      "ds_read_u16 %1, %9\n\t"
      "ds_read_u16 %2, %10\n\t"
      "ds_read_u16 %3, %11\n\t"
      "ds_read_u16 %4, %12\n\t"
      "ds_read_u16 %5, %13\n\t"
      "ds_read_u16 %6, %14\n\t"
      "ds_read_u16 %7, %15\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]), "=&v"(v[7])
      : "v"(ad[0]), "v"(ad[1]), "v"(ad[2]), "v"(ad[3]), "v"(ad[4]), "v"(ad[5]), "v"(ad[6]), "v"(ad[7])
      : "memory");
  return make_uint4(v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16), v[6] | (v[7] << 16));
}

// The k-loop of a workgroup's 64 x 64 tile over nsteps steps of 32, four
// waves each owning a 32 x 32 quarter (2 x 2 MFMA tiles).  Operands stream
// through LDS with direct global-to-LDS loads (global_load_lds_dwordx4: lane
// l writes its 16 bytes at the instruction's LDS base + 16 l) in four
// stages, three steps ahead of the MFMAs; a stage holds A rows 0..63 then B
// rows 0..63, 64 bytes each, with the 16-byte chunk q of row n at slot
// (q + n / 4) mod 4 so that the 16 rows of an MFMA operand start in 16 bank
// groups.  Wave w loads row group w of A and of B.  The four stages are
// separate LDS objects and the loop is unrolled by four, so that every
// access names its stage.  (One wave per 64 x 64 tile left half the SIMDs
// idle and serialised each wave's loads, MFMAs and epilogue; register
// prefetching was undone by the compiler, which moved the loads next to
// their MFMAs.)
__device__ __forceinline__ void gemm_steps(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb,
                                           const unsigned* __restrict__ aidx, int anv, bf16* __restrict__ AT,
                                           int ldat, int m0, int n0, int nsteps, int wave, int lane, bf16* st0,
                                           bf16* st1, bf16* st2, bf16* st3, f32x4 (&acc)[2][2]) {
  if (nsteps <= 0) return;
  const int r = lane & 15, q = lane >> 4;
  const int lrow = lane >> 2;
  const int lchunk = ((lane & 3) - (lane >> 4) + 4) & 3;
  // this lane's rows of group `wave` (A gathered through aidx if given)
  const int ma = m0 + 16 * wave + lrow;
  const size_t arow = aidx ? (ma < anv ? (size_t)aidx[ma] : 0) : (size_t)ma;
  const bf16* Ag = A + arow * (size_t)lda + lchunk * 8;
  const bf16* Bg = B + (size_t)(n0 + 16 * wave + lrow) * (size_t)ldb + lchunk * 8;
  auto load_stage = [&](int t, bf16* buf) {
    const int k0 = 32 * t;
    __builtin_amdgcn_global_load_lds((const void*)(Ag + k0),
                                     (__attribute__((address_space(3))) void*)(buf + wave * 512), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((const void*)(Bg + k0),
                                     (__attribute__((address_space(3))) void*)(buf + 2048 + wave * 512), 16, 0, 0);
  };
  const int ra = 32 * (wave >> 1), rb = 32 * (wave & 1);  // this wave's quarter
  // step t: stage t from `cur`; stage t + 3 starts into `ahead` (the stage
  // every wave finished reading before the barrier of step t - 1).  Past
  // the last stage the loads repeat it into the free buffer: every step
  // issues the same loads and waits the same count.
  const int last = nsteps - 1;
  auto step = [&](int t, const bf16* cur, bf16* ahead) {
    load_stage(min(t + 3, last), ahead);
    bf16x8 a[2], b[2];
    const bf16* pf[4];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int na = ra + 16 * i + r, nb = rb + 16 * i + r;
      pf[i] = cur + na * 32 + ((q + (na >> 2)) & 3) * 8;
      pf[2 + i] = cur + 2048 + nb * 32 + ((q + (nb >> 2)) & 3) * 8;
    }
    lds_read4(pf[0], pf[1], pf[2], pf[3], a[0], a[1], b[0], b[1]);
    if (AT) {
      // the stage's A slice, feature-major: thread -> feature k = tid / 8 of
      // the slice, rows 8 (tid % 8) .. + 8 of the tile (16 bytes; eight
      // threads write a 128-byte row piece)
      const int tid = threadIdx.x, k = tid >> 3, mr = (tid & 7) * 8;
      unsigned ad[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int m = mr + e;
        ad[e] = (unsigned)(uintptr_t)(const __attribute__((address_space(3))) bf16*)(
            cur + m * 32 + (((k >> 3) + (m >> 2)) & 3) * 8 + (k & 7));
      }
      *(uint4*)(AT + (size_t)(32 * t + k) * (size_t)ldat + mr) = lds_read8_u16(ad);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    // this wave's share of stage t + 1 has landed (its loads of steps t + 2
    // and t + 3, two each, may be in flight); the barrier makes every
    // wave's share visible.  A bare s_barrier: the loop writes no LDS
    // itself and the reads of `cur` are complete (they fed the MFMAs).
    wait_vm_except(4);
    __builtin_amdgcn_s_barrier();
  };
  load_stage(0, st0);
  load_stage(min(1, last), st1);
  load_stage(min(2, last), st2);
  wait_vm_except(4);  // stage 0 has landed
  __builtin_amdgcn_s_barrier();
  for (int t = 0; t < nsteps; t += 4) {
    step(t, st0, st3);
    if (t + 1 < nsteps) step(t + 1, st1, st0);
    if (t + 2 < nsteps) step(t + 2, st2, st1);
    if (t + 3 < nsteps) step(t + 3, st3, st2);
  }
  wait_vm_except(0);  // no load into LDS outlives the loop
}

// LDS of one GEMM workgroup: the four operand stages and the cross-wave
// column and loss sums of the epilogue
struct GemmSmem {
  bf16 st[4][4096];
  float csp[2][64];
  float lsp[4], wsp[4];
};

// Workgroup `id` of product g: tile (id % tiles), k-split (id / tiles).
template <int EPI>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, int id, GemmSmem& sm) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const GemmEpi& ep = g.ep;
  const int M = g.M, N = g.N;
  // n-tiles fastest, so that the workgroups sharing a row block of A run
  // together and find it in L2
  const int nt = N / 64, tiles = (M / 64) * nt;
  const int tile = id % tiles, split = id / tiles;
  const int m0 = (tile / nt) * 64, n0 = (tile % nt) * 64;
  // split-K: this workgroup sums k in [kbeg, kend) into its own partial
  const int kc = g.splits > 1 ? ep.kc : g.K;
  const int kbeg = split * kc;
  const int kend = min(g.K, kbeg + kc);
  const int K = max(kend - kbeg, 0);
  float* Cf = ep.Cf;
  if (EPI == EPI_F32) Cf += (size_t)split * (size_t)ep.split_stride;
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // the feature-major copy of A, from the workgroups of the first n-tile
  bf16* AT = (ep.AT && n0 == 0) ? ep.AT + (size_t)kbeg * (size_t)ep.ldat + m0 : nullptr;
  gemm_steps(g.A + kbeg, g.lda, g.B + kbeg, g.ldb, ep.aidx, ep.anv, AT, ep.ldat, m0, n0, K / 32, wave, lane,
             sm.st[0], sm.st[1], sm.st[2], sm.st[3], acc);
  // epilogue: this wave's tile (i, j) element (row ra + 16 i + 4 q + e,
  // column rb + 16 j + r of the workgroup's tile)
  const int ra = 32 * (wave >> 1), rb = 32 * (wave & 1);
  constexpr bool SUMS = (EPI == EPI_GELU_BWD || EPI == EPI_MSE);
  float lsum = 0.f, wsum = 0.f;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = n0 + rb + 16 * j + r;
    const float bn = (EPI == EPI_F32 || EPI == EPI_GELU_BWD) ? 0.f : ep.bias[n];
    float csum = 0.f;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int mb = m0 + ra + 16 * i + 4 * q;
      bf16 tv[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = mb + e;
        const float v = acc[i][j][e];
        if (EPI == EPI_F32) {
          Cf[(size_t)m * ep.ldcf + n] = v;
        } else if (EPI == EPI_BIAS_F32) {
          Cf[(size_t)m * ep.ldcf + n] = v + bn;
        } else if (EPI == EPI_BIAS_F32_T) {  // output-major: Cf[n][m]
          Cf[(size_t)n * ep.ldcf + m] = v + bn;
        } else if (EPI == EPI_BIAS_GELU) {  // bf16 activation, optional fp32 pre-activation
          const float pre = v + bn;
          if (Cf) Cf[(size_t)m * ep.ldcf + n] = pre;
          tv[e] = f2bf(gelu(pre));
          ep.Cb[(size_t)m * ep.ldcb + n] = tv[e];
        } else if (EPI == EPI_GELU_BWD) {  // gradient at the pre-activation of the layer below
          tv[e] = f2bf(v * gelu_grad(ep.pre[(size_t)m * ep.ldp + n]));
          ep.Cb[(size_t)m * ep.ldcb + n] = tv[e];
          csum += bf2f(tv[e]);
        } else {  // EPI_MSE: output y = v + b, loss w (y - t)^2 and its gradient
          // 2 w (y - t); the 1 / sum(w) of the mean is applied to the summed
          // gradient in k_grad_step, where the batch's sum(w) is known
          float gr = 0.f;
          if (m < ep.mvalid && n < ep.nvalid) {
            const size_t tt = (ep.tidx ? (size_t)ep.tidx[m] : (size_t)m) * (size_t)ep.ldt + n;
            const float w = ep.W[tt];
            const float d = (v + bn) - ep.T[tt];
            lsum += w * d * d;
            wsum += w;
            gr = 2.f * w * d;
          }
          tv[e] = f2bf(gr);
          ep.Cb[(size_t)m * ep.ldcb + n] = tv[e];
          csum += bf2f(tv[e]);
        }
      }
      if ((EPI == EPI_BIAS_GELU || SUMS) && ep.CbT) {
        // rows mb .. mb + 3 of column n are adjacent in the transposed copy
        uint2 w;
        w.x = (uint32_t)tv[0] | ((uint32_t)tv[1] << 16);
        w.y = (uint32_t)tv[2] | ((uint32_t)tv[3] << 16);
        *(uint2*)(ep.CbT + (size_t)n * ep.ldct + mb) = w;
      }
    }
    if (SUMS) {
      // the column's 32 rows of this wave: lanes r, r + 16, r + 32, r + 48
      // hold 8 rows each; pairwise sums are the same on every lane
      csum += __shfl_xor(csum, 16);
      csum += __shfl_xor(csum, 32);
      if (q == 0) sm.csp[wave >> 1][rb + 16 * j + r] = csum;
    }
  }
  if (EPI == EPI_MSE) {
    // the wave's loss and weight sums: butterfly over the 64 lanes (the
    // same value on each)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lsum += __shfl_xor(lsum, o);
      wsum += __shfl_xor(wsum, o);
    }
    if (lane == 0) {
      sm.lsp[wave] = lsum;
      sm.wsp[wave] = wsum;
    }
  }
  if (SUMS) {
    // the tile's 64 rows: upper half + lower half, in that order
    __syncthreads();
    if (threadIdx.x < 64)
      ep.colsum[(size_t)(m0 / 64) * (size_t)N + n0 + threadIdx.x] = sm.csp[0][threadIdx.x] + sm.csp[1][threadIdx.x];
    if (EPI == EPI_MSE && threadIdx.x == 0) {
      ep.lpart[tile] = (sm.lsp[0] + sm.lsp[1]) + (sm.lsp[2] + sm.lsp[3]);
      ep.cpart[tile] = (sm.wsp[0] + sm.wsp[1]) + (sm.wsp[2] + sm.wsp[3]);
    }
  }
}

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ __align__(16) GemmSmem sm;
  gemm_tile<EPI>(g, (int)blockIdx.x, sm);
}

template <int E1, int E2>
__global__ void __launch_bounds__(256) k_gemm_pair(GemmArgs g1, GemmArgs g2) {
  __shared__ __align__(16) GemmSmem sm;
  const int n1 = (g1.M / 64) * (g1.N / 64) * g1.splits;
  if ((int)blockIdx.x < n1)
    gemm_tile<E1>(g1, (int)blockIdx.x, sm);
  else
    gemm_tile<E2>(g2, (int)blockIdx.x - n1, sm);
}

template __global__ void k_gemm_bf16<EPI_BIAS_F32>(GemmArgs);
template __global__ void k_gemm_bf16<EPI_BIAS_F32_T>(GemmArgs);
template __global__ void k_gemm_bf16<EPI_BIAS_GELU>(GemmArgs);
template __global__ void k_gemm_bf16<EPI_MSE>(GemmArgs);
template __global__ void k_gemm_bf16<EPI_F32>(GemmArgs);
template __global__ void k_gemm_pair<EPI_F32, EPI_GELU_BWD>(GemmArgs, GemmArgs);

// ---------------------------------------------------------------------------
// Fused inference: the whole network on a block of 128 rows per workgroup.
//
// The block's activations stay in LDS, in two buffers of 128 x FA_PITCH
// bf16 (buf0, then buf1).  The first layer's input is staged into buf0
// (columns 0..255) and, for columns 256.., into buf1 at row pitch FB_PITCH.
//
// Each layer computes out^T = W act^T: W is the MFMA's A operand (16 output
// features x 32 k, one 16-byte global load per lane) and the activations its
// B operand (16 rows x 32 k, one 16-byte LDS read per lane), so a result
// tile holds, per lane, four consecutive output features of one row -- one
// 8-byte LDS write in the epilogue.  The 16-wide output tiles of a layer go
// to the waves round robin (wave w takes tiles w and w + 8) and every wave
// covers all 128 rows: no weight row is used by two waves, so each wave
// streams its own weight fragments from global memory (the 0.7 MB of
// weights stay in L2, shared by every workgroup) through a ring of D
// register slots, D k-steps ahead of the MFMAs, and the k-loop has no
// barrier.  The refills of a layer's last D steps load the next layer's
// first D steps.  Layer 0 reads buf0/buf1 and writes buf1 after a barrier
// (its input covers part of buf1); layer l >= 1 reads one buffer and writes
// the other, so one barrier per layer separates an epilogue's writes from
// the next layer's reads.
// ---------------------------------------------------------------------------

constexpr int FP_TILES = 2;  // 16-wide output tiles per wave and layer (FP_MAXN / 16 / FP_WAVES)
static_assert(FP_TILES * 16 * FP_WAVES == FP_MAXN, "two output tiles per wave cover the widest layer");

// LDS writes before the barrier are visible to the LDS reads after it; the
// fences name LDS only, so loads from global memory stay in flight across it
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// This lane's weight-fragment pointers for one layer: tile j of the wave is
// output tile wave + 8 j, W row 16 (wave + 8 j) + r, k chunk q.  A tile the
// layer does not have reads a tile it has (loaded, never used).
struct WPtrs {
  const bf16* w[FP_TILES];
};
__device__ __forceinline__ WPtrs wptrs(const bf16* W, int K, int N, int wave, int r, int q) {
  WPtrs p;
  const int tiles = N / 16;
#pragma unroll
  for (int j = 0; j < FP_TILES; ++j) {
    const int t = wave + FP_WAVES * j < tiles ? wave + FP_WAVES * j : wave % tiles;
    p.w[j] = W + (size_t)(16 * t + r) * (size_t)K + 8 * q;
  }
  return p;
}

// The k-loop of one layer for a wave with NT output tiles: acc[j][i] (output
// features of tile j, rows 16 i .. 16 i + 15) = W_tile act^T.  The ring
// holds the weight fragments of steps t .. t + D - 1 on entry (slot s % D);
// each step refills its slot with step s + D, of this layer or (for the
// last D steps) of the next.  The activation fragments are double-buffered:
// step s issues the LDS reads of step s + 1 (the last step re-reads its
// own), then its MFMAs, then the refill, and a scheduling barrier keeps that
// order (left to itself, the compiler sank the refills to the end of the
// unrolled block, where the next block's first MFMAs waited for them, and
// moved the LDS reads of step s + 1 down to their use: an empty asm
// statement with a memory clobber holds them in place).
// HASB: the input reaches past column 256 (the first layer), whose columns
// 256.. sit in `in1` at pitch FB_PITCH.
template <bool HASB>
__device__ __forceinline__ void fused_read_act(bf16x8 (&a)[8], const bf16* a0, const bf16* a1, int k0) {
  if (!HASB || k0 < 256) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = *(const bf16x8*)(a0 + k0 + 16 * i * FA_PITCH);
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = *(const bf16x8*)(a1 + (k0 - 256) + 16 * i * FB_PITCH);
  }
}

template <int NT, int D, bool HASB>
__device__ __forceinline__ void fused_kloop(bf16x8 (&ring)[4][FP_TILES], const WPtrs& cur, const WPtrs& nxt,
                                            int nsteps, const bf16* in0, const bf16* in1, int r, int q,
                                            const float* sb, int wave, f32x4 (&acc)[FP_TILES][8]) {
  // the accumulators start at the bias (a result tile holds features
  // n0 .. n0 + 3 of one row per lane; sb: the layer's biases in LDS): the
  // epilogue adds nothing
#pragma unroll
  for (int j = 0; j < FP_TILES; ++j) {
    const f32x4 b = j < NT ? *(const f32x4*)(sb + 16 * (wave + FP_WAVES * j) + 4 * q) : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[j][i] = b;
  }
  const bf16* a0 = in0 + r * FA_PITCH + 8 * q;
  const bf16* a1 = in1 + r * FB_PITCH + 8 * q;
  bf16x8 a[2][8];
  if (NT > 0) fused_read_act<HASB>(a[0], a0, a1, 0);
  for (int t = 0; t < nsteps; t += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int s = t + u;
      if (NT > 0) {
        const int s1 = s + 1 < nsteps ? s + 1 : s;
        fused_read_act<HASB>(a[(u + 1) & 1], a0, a1, 32 * s1);
        __asm__ volatile("" ::: "memory");  // the reads stay ahead of this step's MFMAs
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
          for (int i = 0; i < 8; ++i)
            acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ring[u][j], a[u & 1][i], acc[j][i], 0, 0, 0);
      }
      const int sr = s + D;
#pragma unroll
      for (int j = 0; j < FP_TILES; ++j)
        ring[u][j] = *(const bf16x8*)(sr < nsteps ? cur.w[j] + 32 * sr : nxt.w[j] + 32 * (sr - nsteps));
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int D, bool HASB>
__device__ __forceinline__ void fused_kloop_n(int nt, bf16x8 (&ring)[4][FP_TILES], const WPtrs& cur, const WPtrs& nxt,
                                              int nsteps, const bf16* in0, const bf16* in1, int r, int q,
                                              const float* sb, int wave, f32x4 (&acc)[FP_TILES][8]) {
  if (nt >= 2)
    fused_kloop<2, D, HASB>(ring, cur, nxt, nsteps, in0, in1, r, q, sb, wave, acc);
  else if (nt == 1)
    fused_kloop<1, D, HASB>(ring, cur, nxt, nsteps, in0, in1, r, q, sb, wave, acc);
  else
    fused_kloop<0, D, HASB>(ring, cur, nxt, nsteps, in0, in1, r, q, sb, wave, acc);
}

// two floats to two bf16 (round to nearest even) in one word: one
// v_cvt_pk_bf16_f32 (two one-lane conversions and the shift and or took four)
__device__ __forceinline__ unsigned pack_bf16x2(float lo, float hi) {
  typedef __attribute__((ext_vector_type(2))) float f32x2;
  typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2_t;
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{lo, hi}, bf16x2_t));
}

// Hidden-layer epilogue: GELU(acc) (acc started at the bias) in bf16 to the
// next layer's LDS input, four features of one row per 8-byte write.
template <int NT>
__device__ __forceinline__ void fused_store_act(const f32x4 (&acc)[FP_TILES][8], int wave, int r, int q,
                                                bf16* out) {
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int n0 = 16 * (wave + FP_WAVES * j) + 4 * q;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      uint2 w;
      w.x = pack_bf16x2(gelu(acc[j][i][0]), gelu(acc[j][i][1]));
      w.y = pack_bf16x2(gelu(acc[j][i][2]), gelu(acc[j][i][3]));
      *(uint2*)(out + (16 * i + r) * FA_PITCH + n0) = w;
    }
  }
}

// Last-layer epilogue: acc (bias included) in FP32 to Y, row-major
// (Y[m ldy + n]) or output-major (Y[n ldy + m]).
template <int NT>
__device__ __forceinline__ void fused_store_out(const f32x4 (&acc)[FP_TILES][8], int wave, int r, int q,
                                                float* __restrict__ Y, long long ldy, long long m0, int out_major) {
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int n0 = 16 * (wave + FP_WAVES * j) + 4 * q;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const long long m = m0 + 16 * i + r;
      const f32x4 v = acc[j][i];
      if (out_major) {
#pragma unroll
        for (int e = 0; e < 4; ++e) Y[(size_t)(n0 + e) * (size_t)ldy + (size_t)m] = v[e];
      } else {
        *(f32x4*)(Y + (size_t)m * (size_t)ldy + n0) = v;
      }
    }
  }
}

template <int D>
__device__ __forceinline__ void fused_body(const FusedNet& fn, const bf16* __restrict__ X0, int ldx,
                                           float* __restrict__ Y, long long ldy, int out_major, bf16* buf0,
                                           bf16* buf1, float* sbias) {
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;  // wave index in a scalar register
  const int r = lane & 15, q = lane >> 4;
  const long long m0 = (long long)blockIdx.x * 128;
  {
    // the block's input, region by region, with direct global-to-LDS loads
    // in 1 KB pieces of 64 16-byte chunks (lane l writes the piece's chunk
    // l); the pad chunk at the end of each LDS row re-reads the row's first
    // chunk of that region (never used)
    auto stage = [&](bf16* dst, int pitch, int col0, int ncol) {
      const int cpr = pitch / 8;  // chunks per LDS row
      const int nch = 128 * cpr;  // a multiple of 64
      for (int c0 = wave * 64; c0 < nch; c0 += 64 * FP_WAVES) {
        const int c = c0 + lane;
        const int row = c / cpr, col = (c % cpr) * 8;
        const bf16* src = X0 + (size_t)(m0 + row) * (size_t)ldx + col0 + (col < ncol ? col : 0);
        __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)(dst + c0 * 8),
                                         16, 0, 0);
      }
    };
    stage(buf0, FA_PITCH, 0, min(ldx, 256));
    if (ldx > 256) stage(buf1, FB_PITCH, 256, ldx - 256);
  }
  bf16x8 ring[4][FP_TILES];
  WPtrs cur = wptrs(fn.W[0], fn.pin[0], fn.pout[0], wave, r, q);
#pragma unroll
  for (int u = 0; u < D; ++u)
#pragma unroll
    for (int j = 0; j < FP_TILES; ++j) ring[u][j] = *(const bf16x8*)(cur.w[j] + 32 * u);
  for (int l = 0; l < fn.nl; ++l)
    for (int n = tid; n < fn.pout[l]; n += 64 * FP_WAVES) sbias[l * FP_MAXN + n] = fn.b[l][n];
  __builtin_amdgcn_s_waitcnt(0);  // the staged input has landed (and the first weight fragments)
  __syncthreads();
  const bf16* in0 = buf0;
  bf16* out = buf1;
  for (int l = 0; l < fn.nl; ++l) {
    const int K = fn.pin[l], N = fn.pout[l];
    const int tiles = N / 16;
    const int nt = (wave < tiles ? 1 : 0) + (wave + FP_WAVES < tiles ? 1 : 0);
    const bool last = l == fn.nl - 1;
    const WPtrs nxt = last ? cur : wptrs(fn.W[l + 1], fn.pin[l + 1], fn.pout[l + 1], wave, r, q);
    const float* sb = sbias + l * FP_MAXN;
    f32x4 acc[FP_TILES][8];
    if (K > 256)
      fused_kloop_n<D, true>(nt, ring, cur, nxt, K / 32, in0, buf1, r, q, sb, wave, acc);
    else
      fused_kloop_n<D, false>(nt, ring, cur, nxt, K / 32, in0, buf1, r, q, sb, wave, acc);
    if (last) {
      if (nt >= 2)
        fused_store_out<2>(acc, wave, r, q, Y, ldy, m0, out_major);
      else if (nt == 1)
        fused_store_out<1>(acc, wave, r, q, Y, ldy, m0, out_major);
      break;
    }
    if (l == 0) lds_barrier();  // the first layer's input covers part of buf1: all reads done first
    if (nt >= 2)
      fused_store_act<2>(acc, wave, r, q, out);
    else if (nt == 1)
      fused_store_act<1>(acc, wave, r, q, out);
    lds_barrier();  // the next layer's input is complete
    in0 = out;
    out = (out == buf1) ? buf0 : buf1;
    cur = nxt;
  }
}

// Limits (checked by the host): pin[0] <= FP_MAXK, every other pin and every
// pout <= FP_MAXN, all multiples of 64; rows a multiple of 128.  Weight
// fragments run 4 k-steps ahead when every pin is a multiple of 128, else 2.
__global__ void __launch_bounds__(64 * FP_WAVES) k_mlp_fused(FusedNet fn, const bf16* __restrict__ X0, int ldx,
                                                             float* __restrict__ Y, long long ldy, int out_major) {
  extern __shared__ __align__(16) bf16 act[];  // buf0 then buf1 (dynamic LDS)
  __shared__ __align__(16) float sbias[FP_MAXL * FP_MAXN];
  bool d4 = true;
  for (int l = 0; l < fn.nl; ++l) d4 = d4 && (fn.pin[l] % 128 == 0);
  if (d4)
    fused_body<4>(fn, X0, ldx, Y, ldy, out_major, act, act + 128 * FA_PITCH, sbias);
  else
    fused_body<2>(fn, X0, ldx, Y, ldy, out_major, act, act + 128 * FA_PITCH, sbias);
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

// Fixed-order sum of part[0..n) by one workgroup of 256 (the same value in
// every thread; red is overwritten).
__device__ __forceinline__ float block_sum(const float* __restrict__ part, int n, float* red) {
  float a = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a += part[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const float v = red[0];
  __syncthreads();
  return v;
}

// The batch loss from the output layer's per-tile partials: *loss +=
// sum(lpart) / sum(cpart), *count += sum(cpart).
__device__ __forceinline__ void batch_loss(const float* __restrict__ lpart, const float* __restrict__ cpart, int n,
                                           float* __restrict__ loss, float* __restrict__ count, float* red) {
  const float c = block_sum(cpart, n, red);
  const float l = block_sum(lpart, n, red);
  if (threadIdx.x == 0) {
    *loss += c > 0.f ? l * (1.f / c) : 0.f;
    if (count) *count += c;
  }
}

__global__ void __launch_bounds__(256) k_loss_final(const float* __restrict__ lpart, const float* __restrict__ cpart,
                                                    int n, float* __restrict__ loss, float* __restrict__ count) {
  __shared__ float red[256];
  batch_loss(lpart, cpart, n, loss, count, red);
}

// One launch after the backward pass, thread per parameter of the padded
// FP32 buffer: the gradient is the fixed-order sum of its split partials
// (S k-chunks of the batch for weights, SB 64-row blocks for biases) times
// 1 / sum(w) of the batch (every workgroup forms that sum from the output
// layer's per-tile partials, in the same order), kept in g; with update,
// Adam (Kingma & Ba 2015, bias-corrected by c1, c2) moves the FP32 master
// and the BF16 copies W and W^T follow.  The last workgroup also adds the
// batch loss to *loss and sum(w) to *count.
__global__ void __launch_bounds__(256) k_grad_step(GradPlan gp, int S, int SB, long long np, float* __restrict__ p,
                                                   float* __restrict__ g, float* __restrict__ m1,
                                                   float* __restrict__ m2, bf16* __restrict__ Wb,
                                                   bf16* __restrict__ WTb, int update, float lr, float b1, float b2,
                                                   float eps, float c1, float c2, const float* __restrict__ lpart,
                                                   const float* __restrict__ cpart, int nlp, float* __restrict__ loss,
                                                   float* __restrict__ count) {
  __shared__ float red[256];
  const float cnt = block_sum(cpart, nlp, red);
  const float inv = cnt > 0.f ? 1.f / cnt : 0.f;
  if (loss && blockIdx.x == gridDim.x - 1) batch_loss(lpart, cpart, nlp, loss, count, red);
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  int l = 0;
  while (l + 1 < gp.nl && j >= gp.offW[l + 1]) ++l;
  const long long e = j - gp.offW[l];
  const long long oi = (long long)gp.pout[l] * gp.pin[l];
  // the partials of a group of 16 are loaded together, then added in order
  const bool wt = e < oi;
  const float* q = wt ? gp.pw[l] + e : gp.pb[l] + (e - oi);
  const size_t st = wt ? (size_t)oi : (size_t)gp.pout[l];
  const int ns = wt ? S : SB;
  float gv = 0.f;
  for (int s0 = 0; s0 < ns; s0 += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = s0 + u < ns ? q[(size_t)(s0 + u) * st] : 0.f;
#pragma unroll
    for (int u = 0; u < 16; ++u) gv += v[u];
  }
  gv *= inv;
  g[j] = gv;
  if (!update) return;
  const float a = b1 * m1[j] + (1.f - b1) * gv;
  const float b = b2 * m2[j] + (1.f - b2) * gv * gv;
  m1[j] = a;
  m2[j] = b;
  const float pn = p[j] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  p[j] = pn;
  if (e < oi) {
    const int rr = (int)(e / gp.pin[l]), cc = (int)(e % gp.pin[l]);
    const bf16 v = f2bf(pn);
    Wb[gp.offWb[l] + e] = v;
    WTb[gp.offWb[l] + (long long)cc * gp.pout[l] + rr] = v;
  }
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
