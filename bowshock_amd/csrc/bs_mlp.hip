// Learned warm start: a multilayer perceptron trained and evaluated with
// hand-written BF16 MFMA GEMM kernels (FP32 accumulation), FP32 master
// weights and Adam (bs_mlp_* in include/bowshock.h). No counterpart in the
// reference, which contains no learned components.
//
// Network: inputs -> hidden (SiLU) x depth -> outputs (linear). Layer l maps
// in_l to out_l with Z_l = A_{l-1} W_l + b_l, A_l = SiLU(Z_l); W_l is stored
// in-major [in_l][out_l]. Every dimension the GEMMs see is a multiple of 64
// (the input width is zero-padded to 64), so the kernels need no edge
// handling.
//
// GEMM: C[M x N] = A[M x K] B[K x N], A and B BF16 with element strides
// (row-major or transposed views), a 64 x 64 block tile in four wavefronts
// of 32 x 32, K steps of 32 staged in LDS, v_mfma_f32_16x16x32_bf16 with
// FP32 accumulators, and a fused epilogue (bias, SiLU, SiLU' for the
// backward pass, FP32 and/or BF16 stores).

#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "bs_internal.hpp"

using namespace bs;

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int GM = 64, GN = 64, GK = 32, GPAD = 8;

enum Epilogue : int { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2, EPI_DSILU = 3 };

struct GemmArgs {
  const bf16* A;
  long lda_m, lda_k;
  const bf16* B;
  long ldb_k, ldb_n;
  float* C;          // FP32 result (EPI_STORE, EPI_BIAS: the output; EPI_BIAS_SILU: pre-activation; EPI_DSILU: gradient)
  bf16* Cb;          // BF16 copy (EPI_BIAS_SILU: activation; EPI_DSILU: gradient), or null
  long ldc;
  const float* bias; // per column
  const float* pre;  // EPI_DSILU: the forward pre-activation, ldc layout
  int M, N, K;
};

__device__ __forceinline__ float sigmoid_f(float z) { return 1.f / (1.f + __expf(-z)); }

// The eight BF16 values of a 16-byte vector to eight LDS places `stride`
// elements apart (a transposed tile), by bit fields: taking the vector's
// address would put it in scratch memory.
__device__ __forceinline__ void scatter8(bf16* dst, int stride, const uint4& v) {
  unsigned short* d = (unsigned short*)dst;
  d[0] = (unsigned short)(v.x & 0xFFFFu);
  d[stride] = (unsigned short)(v.x >> 16);
  d[2 * stride] = (unsigned short)(v.y & 0xFFFFu);
  d[3 * stride] = (unsigned short)(v.y >> 16);
  d[4 * stride] = (unsigned short)(v.z & 0xFFFFu);
  d[5 * stride] = (unsigned short)(v.z >> 16);
  d[6 * stride] = (unsigned short)(v.w & 0xFFFFu);
  d[7 * stride] = (unsigned short)(v.w >> 16);
}

// A: K contiguous (A_KC) or M contiguous; B: N contiguous (B_NC) or K
// contiguous. LDS holds A as [m][k] and B as [n][k], so both MFMA operands
// are 16-byte LDS reads of 8 consecutive k.
template <int EPI, bool A_KC, bool B_NC>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[GM][GK + GPAD];
  __shared__ __attribute__((aligned(16))) bf16 Bs[GN][GK + GPAD];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const long m0 = (long)blockIdx.y * GM, n0 = (long)blockIdx.x * GN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += GK) {
    if (A_KC) {
      const int m = t >> 2, kc = (t & 3) * 8;
      *(uint4*)&As[m][kc] = *(const uint4*)(g.A + (m0 + m) * g.lda_m + (k0 + kc));
    } else {
      const int k = t >> 3, mc = (t & 7) * 8;
      const uint4 v = *(const uint4*)(g.A + (m0 + mc) + (long)(k0 + k) * g.lda_k);
      scatter8(&As[mc][k], GK + GPAD, v);
    }
    if (B_NC) {
      const int k = t >> 3, nc = (t & 7) * 8;
      const uint4 v = *(const uint4*)(g.B + (long)(k0 + k) * g.ldb_k + (n0 + nc));
      scatter8(&Bs[nc][k], GK + GPAD, v);
    } else {
      const int n = t >> 2, kc = (t & 3) * 8;
      *(uint4*)&Bs[n][kc] = *(const uint4*)(g.B + (n0 + n) * g.ldb_n + (k0 + kc));
    }
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) a[i] = *(const bf16x8*)&As[wm + i * 16 + (lane & 15)][8 * (lane >> 4)];
#pragma unroll
    for (int j = 0; j < 2; j++) b[j] = *(const bf16x8*)&Bs[wn + j * 16 + (lane & 15)][8 * (lane >> 4)];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // Accumulator r of lane l holds row 4 (l / 16) + r, column l % 16 of its
  // 16 x 16 block.
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long m = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
        const long n = n0 + wn + j * 16 + (lane & 15);
        const long o = m * g.ldc + n;
        float v = acc[i][j][r];
        if (EPI == EPI_STORE) {
          g.C[o] = v;
        } else if (EPI == EPI_BIAS) {
          g.C[o] = v + g.bias[n];
        } else if (EPI == EPI_BIAS_SILU) {
          const float z = v + g.bias[n];
          g.C[o] = z;
          g.Cb[o] = (bf16)(z * sigmoid_f(z));
        } else {
          const float z = g.pre[o];
          const float s = sigmoid_f(z);
          const float d = v * (s * (1.f + z * (1.f - s)));
          g.C[o] = d;
          g.Cb[o] = (bf16)d;
        }
      }
}

// The epilogue of one output element (shared by the GEMM kernels and the
// split-K reduction).
template <int EPI>
__device__ __forceinline__ void gemm_store(const GemmArgs& g, long m, long n, float v) {
  const long o = m * g.ldc + n;
  if (EPI == EPI_STORE) {
    g.C[o] = v;
  } else if (EPI == EPI_BIAS) {
    g.C[o] = v + g.bias[n];
  } else if (EPI == EPI_BIAS_SILU) {
    const float z = v + g.bias[n];
    g.C[o] = z;
    g.Cb[o] = (bf16)(z * sigmoid_f(z));
  } else {
    const float z = g.pre[o];
    const float s = sigmoid_f(z);
    const float d = v * (s * (1.f + z * (1.f - s)));
    g.C[o] = d;
    g.Cb[o] = (bf16)d;
  }
}

// 128 x 128 block tile: four wavefronts of 64 x 64, each 2 x 2 blocks of
// v_mfma_f32_32x32x16_bf16 (16 FP32 accumulators per lane per block),
// K steps of 32 staged in LDS, the next step's operands loaded into
// registers while the current one is multiplied. With `split` > 1 the
// z-th block of the grid covers K range [z K / split, (z + 1) K / split)
// and stores raw sums to ws[z] for gemm_split_reduce.
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int BM = 128, BN = 128, BK = 32, BPAD = 8;

template <int EPI, bool A_KC, bool B_NC>
__global__ __launch_bounds__(256) void gemm128_kernel(GemmArgs g, int split, float* __restrict__ ws) {
  __shared__ __attribute__((aligned(16))) bf16 As[BM][BK + BPAD];
  __shared__ __attribute__((aligned(16))) bf16 Bs[BN][BK + BPAD];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const long m0 = (long)blockIdx.y * BM, n0 = (long)blockIdx.x * BN;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int kc = g.K / split;
  const int kbeg = blockIdx.z * kc, kend = kbeg + kc;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;
  // Each thread moves two 16-byte vectors of A and two of B per K step:
  // rows (or k) `t / 4` and `t / 4 + 64` (K-contiguous sources), or k
  // `t / 16` and `t / 16 + 16` (transposed sources).
  const int rq = t >> 2, kq = (t & 3) * 8;   // K-contiguous: row, k offset
  const int kt = t >> 4, xq = (t & 15) * 8;  // transposed: k, row offset
  const bf16* a0;
  const bf16* a1;
  long a_step;
  if (A_KC) {
    a0 = g.A + (m0 + rq) * g.lda_m + kq;
    a1 = a0 + 64 * g.lda_m;
    a_step = BK;
  } else {
    a0 = g.A + (m0 + xq) + (long)kt * g.lda_k;
    a1 = a0 + 16 * g.lda_k;
    a_step = (long)BK * g.lda_k;
  }
  const bf16* b0;
  const bf16* b1;
  long b_step;
  if (B_NC) {
    b0 = g.B + (long)kt * g.ldb_k + (n0 + xq);
    b1 = b0 + 16 * g.ldb_k;
    b_step = (long)BK * g.ldb_k;
  } else {
    b0 = g.B + (n0 + rq) * g.ldb_n + kq;
    b1 = b0 + 64 * g.ldb_n;
    b_step = BK;
  }
  const long first = kbeg / BK;
  a0 += first * a_step;
  a1 += first * a_step;
  b0 += first * b_step;
  b1 += first * b_step;
  uint4 ra0 = *(const uint4*)a0, ra1 = *(const uint4*)a1, rb0 = *(const uint4*)b0, rb1 = *(const uint4*)b1;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    if (A_KC) {
      *(uint4*)&As[rq][kq] = ra0;
      *(uint4*)&As[rq + 64][kq] = ra1;
    } else {
      scatter8(&As[xq][kt], BK + BPAD, ra0);
      scatter8(&As[xq][kt + 16], BK + BPAD, ra1);
    }
    if (B_NC) {
      scatter8(&Bs[xq][kt], BK + BPAD, rb0);
      scatter8(&Bs[xq][kt + 16], BK + BPAD, rb1);
    } else {
      *(uint4*)&Bs[rq][kq] = rb0;
      *(uint4*)&Bs[rq + 64][kq] = rb1;
    }
    __syncthreads();
    if (k0 + BK < kend) {
      a0 += a_step;
      a1 += a_step;
      b0 += b_step;
      b1 += b_step;
      ra0 = *(const uint4*)a0;
      ra1 = *(const uint4*)a1;
      rb0 = *(const uint4*)b0;
      rb1 = *(const uint4*)b1;
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const bf16x8 fa0 = *(const bf16x8*)&As[wm + (lane & 31)][kk + 8 * (lane >> 5)];
      const bf16x8 fa1 = *(const bf16x8*)&As[wm + 32 + (lane & 31)][kk + 8 * (lane >> 5)];
      const bf16x8 fb0 = *(const bf16x8*)&Bs[wn + (lane & 31)][kk + 8 * (lane >> 5)];
      const bf16x8 fb1 = *(const bf16x8*)&Bs[wn + 32 + (lane & 31)][kk + 8 * (lane >> 5)];
      acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa0, fb0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa0, fb1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa1, fb0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa1, fb1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  // Accumulator r of lane l holds row 8 (r / 4) + 4 (l / 32) + r % 4,
  // column l % 32 of its 32 x 32 block.
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const long m = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long n = n0 + wn + 32 * j + (lane & 31);
        if (split > 1) ws[(long)blockIdx.z * g.M * g.N + m * g.N + n] = acc[i][j][r];
        else gemm_store<EPI>(g, m, n, acc[i][j][r]);
      }
}

// 64 x 64 block tile for operands that are both K-contiguous (A [m][k],
// B [n][k]): four wavefronts of 32 x 32 as gemm_bf16_kernel, K steps of 64
// through two LDS buffers, and the next step's operands loaded into
// registers while the current step is multiplied, so that one barrier per
// step suffices and the global loads' latency overlaps the MFMAs. Each
// thread moves two 16-byte vectors of A and two of B per step.
constexpr int PK = 64, PPAD = 8;

//
// A one-dimensional grid of (M / 64) (N / 64) blocks placed XCD-aware
// (xcd_tile): the tiles of one XCD are a contiguous run with the M tiles of
// an N column adjacent, so the blocks that share a column of B run on one
// XCD together and all but the first read it from that XCD's L2.
template <int EPI>
__global__ __launch_bounds__(256) void gemm_pipe_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[2][GM][PK + PPAD];
  __shared__ __attribute__((aligned(16))) bf16 Bs[2][GN][PK + PPAD];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int tile = xcd_tile(blockIdx.x, gridDim.x), mt = g.M / GM;
  const long m0 = (long)(tile % mt) * GM, n0 = (long)(tile / mt) * GN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  // Rows t / 8 and t / 8 + 32, k offset 8 (t % 8).
  const int r = t >> 3, kq = (t & 7) * 8;
  const bf16* a0 = g.A + (m0 + r) * g.lda_m + kq;
  const bf16* a1 = a0 + 32 * g.lda_m;
  const bf16* b0 = g.B + (n0 + r) * g.ldb_n + kq;
  const bf16* b1 = b0 + 32 * g.ldb_n;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  uint4 ra0 = *(const uint4*)a0, ra1 = *(const uint4*)a1, rb0 = *(const uint4*)b0, rb1 = *(const uint4*)b1;
  const int steps = g.K / PK;
  for (int s = 0; s < steps; s++) {
    const int buf = s & 1;
    *(uint4*)&As[buf][r][kq] = ra0;
    *(uint4*)&As[buf][r + 32][kq] = ra1;
    *(uint4*)&Bs[buf][r][kq] = rb0;
    *(uint4*)&Bs[buf][r + 32][kq] = rb1;
    __syncthreads();
    if (s + 1 < steps) {
      const long o = (long)(s + 1) * PK;
      ra0 = *(const uint4*)(a0 + o);
      ra1 = *(const uint4*)(a1 + o);
      rb0 = *(const uint4*)(b0 + o);
      rb1 = *(const uint4*)(b1 + o);
    }
#pragma unroll
    for (int kk = 0; kk < PK; kk += 32) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++) a[i] = *(const bf16x8*)&As[buf][wm + i * 16 + (lane & 15)][kk + 8 * (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; j++) b[j] = *(const bf16x8*)&Bs[buf][wn + j * 16 + (lane & 15)][kk + 8 * (lane >> 4)];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        gemm_store<EPI>(g, m0 + wm + i * 16 + 4 * (lane >> 4) + q, n0 + wn + j * 16 + (lane & 15), acc[i][j][q]);
}

// The same tile for K = 64 STEPS with STEPS small: every step's operands
// are loaded into registers at the start, so the block waits for global
// memory once instead of once per step (a 256 x 512 x 12800 GEMM spends
// most of a step waiting on its loads otherwise).
template <int EPI, int STEPS>
__global__ __launch_bounds__(256) void gemm_allk_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[2][GM][PK + PPAD];
  __shared__ __attribute__((aligned(16))) bf16 Bs[2][GN][PK + PPAD];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int tile = xcd_tile(blockIdx.x, gridDim.x), mt = g.M / GM;
  const long m0 = (long)(tile % mt) * GM, n0 = (long)(tile / mt) * GN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int r = t >> 3, kq = (t & 7) * 8;
  const bf16* a0 = g.A + (m0 + r) * g.lda_m + kq;
  const bf16* a1 = a0 + 32 * g.lda_m;
  const bf16* b0 = g.B + (n0 + r) * g.ldb_n + kq;
  const bf16* b1 = b0 + 32 * g.ldb_n;
  // Native vector types: arrays of them stay in registers.
  u32x4 ra0[STEPS], ra1[STEPS], rb0[STEPS], rb1[STEPS];
#pragma unroll
  for (int s = 0; s < STEPS; s++) {
    ra0[s] = *(const u32x4*)(a0 + s * PK);
    ra1[s] = *(const u32x4*)(a1 + s * PK);
    rb0[s] = *(const u32x4*)(b0 + s * PK);
    rb1[s] = *(const u32x4*)(b1 + s * PK);
  }
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < STEPS; s++) {
    const int buf = s & 1;
    *(u32x4*)&As[buf][r][kq] = ra0[s];
    *(u32x4*)&As[buf][r + 32][kq] = ra1[s];
    *(u32x4*)&Bs[buf][r][kq] = rb0[s];
    *(u32x4*)&Bs[buf][r + 32][kq] = rb1[s];
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < PK; kk += 32) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++) a[i] = *(const bf16x8*)&As[buf][wm + i * 16 + (lane & 15)][kk + 8 * (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; j++) b[j] = *(const bf16x8*)&Bs[buf][wn + j * 16 + (lane & 15)][kk + 8 * (lane >> 4)];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        gemm_store<EPI>(g, m0 + wm + i * 16 + 4 * (lane >> 4) + q, n0 + wn + j * 16 + (lane & 15), acc[i][j][q]);
}

// Sum of the split partial products in a fixed order, then the epilogue.
template <int EPI>
__global__ __launch_bounds__(256) void gemm_split_reduce(GemmArgs g, int split, const float* __restrict__ ws) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  const long mn = (long)g.M * g.N;
  if (e >= mn) return;
  float v = 0.f;
  for (int z = 0; z < split; z++) v += ws[z * mn + e];
  gemm_store<EPI>(g, e / g.N, e % g.N, v);
}

// Rows idx[b] of the data set (or rows b of a staged block) into the padded
// BF16 input; rows at or beyond `rows` are zero.
__global__ void gather_inputs(const float* __restrict__ x, const int* __restrict__ idx, int rows, int inputs,
                              int in_pad, int batch_pad, bf16* __restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)batch_pad * in_pad) return;
  const int b = (int)(e / in_pad), k = (int)(e % in_pad);
  float v = 0.f;
  if (b < rows && k < inputs) {
    const long row = idx ? idx[b] : b;
    v = x[row * inputs + k];
  }
  out[e] = (bf16)v;
}

// dY = 2 (Y - T) / (rows * outputs) for the batch's rows (FP32 and BF16),
// and per block the FP64 sum of squared errors.
__global__ __launch_bounds__(256) void mse_grad(const float* __restrict__ y, const float* __restrict__ data_y,
                                                const int* __restrict__ idx, int rows, int outputs, float scale,
                                                float* __restrict__ dyf, bf16* __restrict__ dyb,
                                                double* __restrict__ partials) {
  __shared__ double red[4];
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  double sq = 0.0;
  if (e < (long)rows * outputs) {
    const int b = (int)(e / outputs), n = (int)(e % outputs);
    const float d = y[e] - data_y[(long)idx[b] * outputs + n];
    const float g = scale * d;
    dyf[e] = g;
    dyb[e] = (bf16)g;
    sq = (double)d * (double)d;
  }
  double s = bs::block_sum<256>(sq, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Column sums of a [rows x cols] FP32 matrix (bias gradients).
__global__ void column_sums(const float* __restrict__ g, int rows, int cols, float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= cols) return;
  float s = 0.f;
  for (int b = 0; b < rows; b++) s += g[(long)b * cols + n];
  out[n] = s;
}

// Adam over the flat parameter vector; the BF16 mirror follows the FP32
// master weights.
__global__ void adam_step(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, bf16* __restrict__ wb, long n, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[e] = wi;
  wb[e] = (bf16)wi;
}

__global__ void to_bf16(const float* __restrict__ w, bf16* __restrict__ wb, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) wb[e] = (bf16)w[e];
}

// The transposed BF16 mirror of one layer's weights: wt[n][k] = w[k][n]
// for w of rows x cols (both multiples of 64), through a 64 x 64 LDS tile
// so that the loads and the stores are both coalesced.
__global__ __launch_bounds__(256) void to_bf16_transposed(const float* __restrict__ w, bf16* __restrict__ wt,
                                                          int rows, int cols) {
  __shared__ float tile[64][65];
  const long r0 = (long)blockIdx.y * 64, c0 = (long)blockIdx.x * 64;
  for (int k = threadIdx.x; k < 64 * 64; k += 256) {
    const int r = k >> 6, c = k & 63;
    tile[r][c] = w[(r0 + r) * cols + c0 + c];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 64 * 64; k += 256) {
    const int c = k >> 6, r = k & 63;
    wt[(c0 + c) * rows + r0 + r] = (bf16)tile[r][c];
  }
}

int round64(int v) { return (v + 63) / 64 * 64; }

}  // namespace

struct bs_mlp {
  int device;
  hipStream_t stream = nullptr;
  int inputs, in_pad, hidden, depth, outputs, max_batch;
  int layers;                     // depth + 1
  std::vector<int> lin, lout;     // padded widths per layer
  std::vector<long> w_off, b_off; // offsets into the flat parameter vector
  long n_params = 0;              // padded flat count
  float *w = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  bf16* wb = nullptr;   // BF16 mirror of w: per layer [in][out]
  bf16* wbt = nullptr;  // per layer [out][in], the forward GEMMs' K-contiguous operand
  bf16* x = nullptr;                      // [max_batch x in_pad]
  std::vector<float*> z;                  // per hidden layer [max_batch x hidden], and the output
  std::vector<bf16*> act;                 // per hidden layer
  float *dzf[2] = {nullptr, nullptr};     // ping-pong gradients [max_batch x max(outputs, hidden)]
  bf16* dzb[2] = {nullptr, nullptr};
  double* partials = nullptr;
  int* idx = nullptr;
  float *data_x = nullptr, *data_y = nullptr;
  int data_rows = 0;
  float* stage_x = nullptr;               // predict: staged host block
  float* ws = nullptr;                    // split-K partial products
  long ws_floats = 0;
  long step = 0;
  std::atomic<bool> busy{false};
};

namespace {

struct MlpGuard {
  bs_mlp* p;
  bool ok;
  explicit MlpGuard(bs_mlp* m) : p(m) {
    bool expected = false;
    ok = p->busy.compare_exchange_strong(expected, true);
  }
  ~MlpGuard() {
    if (ok) p->busy.store(false);
  }
};

#define MLP_GUARD(p)                                                                         \
  if ((p) == nullptr) return fail(BS_ERR_RANGE, "no network");                               \
  MlpGuard _mguard(p);                                                                       \
  if (!_mguard.ok) return fail(BS_ERR_BUSY, "bs_mlp: await the previous operation first"); \
  (void)hipSetDevice((p)->device)

void mlp_free(bs_mlp* p) {
  std::vector<void*> all = {p->w, p->grad, p->m, p->v, p->wb, p->wbt, p->x, p->dzf[0], p->dzf[1], p->dzb[0], p->dzb[1],
                            p->partials, p->idx, p->data_x, p->data_y, p->stage_x, p->ws};
  for (float* q : p->z) all.push_back(q);
  for (bf16* q : p->act) all.push_back(q);
  for (void* q : all)
    if (q) (void)hipFree(q);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  p->stream = nullptr;
}

constexpr int MAX_SPLIT = 16;

template <int EPI, bool AK, bool BN_>
void launch_gemm128(bs_mlp* p, const GemmArgs& g) {
  // Split K while the tile grid leaves CUs idle and each part keeps at
  // least 256 of K.
  const int tiles = (g.M / BM) * (g.N / BN);
  int split = 1;
  while (split < MAX_SPLIT && tiles * split < 512 && g.K % (2 * split * BK) == 0 && g.K / (2 * split) >= 256 &&
         (long)2 * split * g.M * g.N <= p->ws_floats)
    split *= 2;
  dim3 grid(g.N / BN, g.M / BM, split);
  hipLaunchKernelGGL((gemm128_kernel<EPI, AK, BN_>), grid, dim3(256), 0, p->stream, g, split, p->ws);
  if (split > 1) {
    const long mn = (long)g.M * g.N;
    hipLaunchKernelGGL(gemm_split_reduce<EPI>, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, p->stream, g, split,
                       (const float*)p->ws);
  }
}

template <int EPI>
void launch_gemm(bs_mlp* p, const GemmArgs& g, bool a_kc, bool b_nc) {
  // Both operands K-contiguous (the forward GEMMs on the transposed weight
  // mirror, dZ W^T): the pipelined 64 x 64 tile. Otherwise the 128 x 128
  // tile where it fills the chip (or its K is long enough to split), else
  // the 64 x 64 tile, which gives four times the blocks.
  if (a_kc && !b_nc && g.K % PK == 0 && g.K < 4096) {
    const dim3 grid((g.N / GN) * (g.M / GM));
    if (g.K == 8 * PK) hipLaunchKernelGGL((gemm_allk_kernel<EPI, 8>), grid, dim3(256), 0, p->stream, g);
    else if (g.K == PK) hipLaunchKernelGGL((gemm_allk_kernel<EPI, 1>), grid, dim3(256), 0, p->stream, g);
    else hipLaunchKernelGGL(gemm_pipe_kernel<EPI>, grid, dim3(256), 0, p->stream, g);
    return;
  }
  const long tiles128 = (long)(g.M / BM) * (g.N / BN);
  if (g.M % BM == 0 && g.N % BN == 0 && g.K % BK == 0 && (tiles128 >= 256 || g.K >= 4096)) {
    if (a_kc && b_nc) launch_gemm128<EPI, true, true>(p, g);
    else if (a_kc) launch_gemm128<EPI, true, false>(p, g);
    else if (b_nc) launch_gemm128<EPI, false, true>(p, g);
    else launch_gemm128<EPI, false, false>(p, g);
    return;
  }
  dim3 grid(g.N / GN, g.M / GM);
  if (a_kc && b_nc) hipLaunchKernelGGL((gemm_bf16_kernel<EPI, true, true>), grid, dim3(256), 0, p->stream, g);
  else if (a_kc) hipLaunchKernelGGL((gemm_bf16_kernel<EPI, true, false>), grid, dim3(256), 0, p->stream, g);
  else if (b_nc) hipLaunchKernelGGL((gemm_bf16_kernel<EPI, false, true>), grid, dim3(256), 0, p->stream, g);
  else hipLaunchKernelGGL((gemm_bf16_kernel<EPI, false, false>), grid, dim3(256), 0, p->stream, g);
}

// The BF16 mirrors of the FP32 master weights (wb is also written by the
// Adam kernel; the transposed mirror is refreshed here).
void mlp_mirror(bs_mlp* p, bool plain) {
  if (plain)
    hipLaunchKernelGGL(to_bf16, dim3((unsigned)((p->n_params + 255) / 256)), dim3(256), 0, p->stream, p->w, p->wb,
                       p->n_params);
  for (int l = 0; l < p->layers; l++)
    hipLaunchKernelGGL(to_bf16_transposed, dim3(p->lout[l] / 64, p->lin[l] / 64), dim3(256), 0, p->stream,
                       p->w + p->w_off[l], p->wbt + p->w_off[l], p->lin[l], p->lout[l]);
}

// Forward pass of `rows` staged input rows (p->x, padded to batch_pad).
void mlp_forward(bs_mlp* p, int batch_pad) {
  const bf16* a = p->x;
  for (int l = 0; l < p->layers; l++) {
    GemmArgs g{};
    g.A = a;
    g.lda_m = p->lin[l];
    g.lda_k = 1;
    g.B = p->wbt + p->w_off[l];  // B[k][n] = W[k][n], K contiguous
    g.ldb_k = 1;
    g.ldb_n = p->lin[l];
    g.C = p->z[l];
    g.ldc = p->lout[l];
    g.bias = p->w + p->b_off[l];
    g.M = batch_pad;
    g.N = p->lout[l];
    g.K = p->lin[l];
    if (l + 1 < p->layers) {
      g.Cb = p->act[l];
      launch_gemm<EPI_BIAS_SILU>(p, g, true, false);
      a = p->act[l];
    } else {
      launch_gemm<EPI_BIAS>(p, g, true, false);
    }
  }
}

// Loss and gradients of the rows idx[0..batch) (batch a multiple of 64).
bs_status mlp_backward(bs_mlp* p, int batch, double* loss) {
  const int L = p->layers;
  const long ne = (long)batch * p->outputs;
  const int nb = (int)((ne + 255) / 256);
  hipLaunchKernelGGL(mse_grad, dim3(nb), dim3(256), 0, p->stream, p->z[L - 1], p->data_y, p->idx, batch, p->outputs,
                     2.f / (float)ne, p->dzf[0], p->dzb[0], p->partials);
  hipLaunchKernelGGL(finalize_sum, dim3(1), dim3(256), 0, p->stream, p->partials, nb, 1.0, 0, p->partials + nb);
  int cur = 0;
  for (int l = L - 1; l >= 0; l--) {
    const bf16* a_prev = l == 0 ? p->x : p->act[l - 1];
    GemmArgs g{};  // dW_l = A_{l-1}^T dZ_l
    g.A = a_prev;
    g.lda_m = 1;
    g.lda_k = p->lin[l];
    g.B = p->dzb[cur];
    g.ldb_k = p->lout[l];
    g.ldb_n = 1;
    g.C = p->grad + p->w_off[l];
    g.ldc = p->lout[l];
    g.M = p->lin[l];
    g.N = p->lout[l];
    g.K = batch;
    launch_gemm<EPI_STORE>(p, g, false, true);
    hipLaunchKernelGGL(column_sums, dim3((p->lout[l] + 255) / 256), dim3(256), 0, p->stream, p->dzf[cur], batch,
                       p->lout[l], p->grad + p->b_off[l]);
    if (l > 0) {
      GemmArgs h{};  // dZ_{l-1} = (dZ_l W_l^T) * SiLU'(Z_{l-1})
      h.A = p->dzb[cur];
      h.lda_m = p->lout[l];
      h.lda_k = 1;
      h.B = p->wb + p->w_off[l];
      h.ldb_k = 1;
      h.ldb_n = p->lout[l];
      h.C = p->dzf[1 - cur];
      h.Cb = p->dzb[1 - cur];
      h.ldc = p->lin[l];
      h.pre = p->z[l - 1];
      h.M = batch;
      h.N = p->lin[l];
      h.K = p->lout[l];
      launch_gemm<EPI_DSILU>(p, h, true, false);
      cur = 1 - cur;
    }
  }
  HIP_TRY(hipGetLastError(), "network launch");
  double out = 0;
  HIP_TRY(hipMemcpyAsync(&out, p->partials + nb, sizeof out, hipMemcpyDeviceToHost, p->stream), "loss readback");
  HIP_TRY(hipStreamSynchronize(p->stream), "network step");
  if (loss) *loss = out / (double)ne;  // mean squared error
  return BS_OK;
}

bs_status mlp_stage_indices(bs_mlp* p, const int32_t* idx, int batch) {
  if (batch < 64 || batch > p->max_batch || batch % 64)
    return fail(BS_ERR_RANGE, "the batch must be a multiple of 64 between 64 and %d", p->max_batch);
  if (!p->data_x) return fail(BS_ERR_RANGE, "no data set loaded");
  for (int b = 0; b < batch; b++)
    if (idx[b] < 0 || idx[b] >= p->data_rows) return fail(BS_ERR_RANGE, "row index %d out of range", idx[b]);
  HIP_TRY(hipMemcpyAsync(p->idx, idx, batch * sizeof(int), hipMemcpyHostToDevice, p->stream), "index upload");
  const long ne = (long)batch * p->in_pad;
  hipLaunchKernelGGL(gather_inputs, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, p->stream, p->data_x, p->idx,
                     batch, p->inputs, p->in_pad, batch, p->x);
  return BS_OK;
}

// External parameter layout: per layer W [in][out] (in = inputs for the
// first layer, unpadded) then b [out].
long external_count(const bs_mlp* p) {
  long c = 0;
  for (int l = 0; l < p->layers; l++) c += (long)(l == 0 ? p->inputs : p->lin[l]) * p->lout[l] + p->lout[l];
  return c;
}

}  // namespace

extern "C" {

bs_status bs_mlp_create(bs_ctx* ctx, int32_t inputs, int32_t hidden, int32_t depth, int32_t outputs,
                        int32_t max_batch, bs_mlp** out) {
  if (!ctx || !out) return fail(BS_ERR_RANGE, "null argument");
  *out = nullptr;
  if (inputs < 1 || hidden < 64 || hidden % 64 || depth < 1 || outputs < 64 || outputs % 64 || max_batch < 64 ||
      max_batch % 64)
    return fail(BS_ERR_RANGE, "hidden width, outputs and batch must be positive multiples of 64");
  bs_mlp* p = new bs_mlp();
  p->device = ctx->device;
  p->inputs = inputs;
  p->in_pad = round64(inputs);
  p->hidden = hidden;
  p->depth = depth;
  p->outputs = outputs;
  p->max_batch = max_batch;
  p->layers = depth + 1;
  for (int l = 0; l < p->layers; l++) {
    p->lin.push_back(l == 0 ? p->in_pad : hidden);
    p->lout.push_back(l == depth ? outputs : hidden);
    p->w_off.push_back(p->n_params);
    p->n_params += (long)p->lin[l] * p->lout[l];
    p->b_off.push_back(p->n_params);
    p->n_params += p->lout[l];
  }
  (void)hipSetDevice(p->device);
  auto bail = [&](bs_status code) {
    mlp_free(p);
    delete p;
    return code;
  };
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(BS_ERR_EXECUTION, "hipStreamCreate failed"));
  const long wide = outputs > hidden ? outputs : hidden;
  bs_status st;
  if ((st = dev_alloc(&p->w, p->n_params, "weights")) || (st = dev_alloc(&p->grad, p->n_params, "gradients")) ||
      (st = dev_alloc(&p->m, p->n_params, "moments")) || (st = dev_alloc(&p->v, p->n_params, "moments")) ||
      (st = dev_alloc(&p->wb, p->n_params, "weights")) || (st = dev_alloc(&p->wbt, p->n_params, "weights")) ||
      (st = dev_alloc(&p->x, (size_t)max_batch * p->in_pad, "inputs")) ||
      (st = dev_alloc(&p->dzf[0], (size_t)max_batch * wide, "gradients")) ||
      (st = dev_alloc(&p->dzf[1], (size_t)max_batch * wide, "gradients")) ||
      (st = dev_alloc(&p->dzb[0], (size_t)max_batch * wide, "gradients")) ||
      (st = dev_alloc(&p->dzb[1], (size_t)max_batch * wide, "gradients")) ||
      (st = dev_alloc(&p->partials, (size_t)((long)max_batch * outputs / 256 + 2), "partials")) ||
      (st = dev_alloc(&p->idx, (size_t)max_batch, "indices")) ||
      (st = dev_alloc(&p->stage_x, (size_t)max_batch * inputs, "inputs")))
    return bail(st);
  // Split-K products: up to MAX_SPLIT parts of the batch x hidden GEMMs.
  p->ws_floats = (long)MAX_SPLIT * max_batch * hidden;
  if ((st = dev_alloc(&p->ws, (size_t)p->ws_floats, "split-K workspace"))) return bail(st);
  for (int l = 0; l < p->layers; l++) {
    float* zl = nullptr;
    if ((st = dev_alloc(&zl, (size_t)max_batch * p->lout[l], "activations"))) return bail(st);
    p->z.push_back(zl);
    if (l + 1 < p->layers) {
      bf16* al = nullptr;
      if ((st = dev_alloc(&al, (size_t)max_batch * p->lout[l], "activations"))) return bail(st);
      p->act.push_back(al);
    }
  }
  const size_t bytes = p->n_params * sizeof(float);
  if (hipMemsetAsync(p->w, 0, bytes, p->stream) != hipSuccess || hipMemsetAsync(p->m, 0, bytes, p->stream) != hipSuccess ||
      hipMemsetAsync(p->v, 0, bytes, p->stream) != hipSuccess ||
      hipMemsetAsync(p->wb, 0, p->n_params * sizeof(bf16), p->stream) != hipSuccess ||
      hipMemsetAsync(p->wbt, 0, p->n_params * sizeof(bf16), p->stream) != hipSuccess ||
      hipStreamSynchronize(p->stream) != hipSuccess)
    return bail(fail(BS_ERR_EXECUTION, "network initialisation failed"));
  *out = p;
  return BS_OK;
}

bs_status bs_mlp_destroy(bs_mlp* p) {
  if (!p) return BS_OK;
  bool idle = false;
  if (!p->busy.compare_exchange_strong(idle, true))
    return fail(BS_ERR_BUSY, "bs_mlp: await the previous operation first (destroy while a call runs)");
  (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  mlp_free(p);
  delete p;
  return BS_OK;
}

int64_t bs_mlp_param_count(const bs_mlp* p) { return p ? external_count(p) : 0; }

bs_status bs_mlp_set_params(bs_mlp* p, const float* params, int64_t count) {
  MLP_GUARD(p);
  if (!params || count != external_count(p)) return fail(BS_ERR_RANGE, "expected %ld parameters", external_count(p));
  std::vector<float> flat(p->n_params, 0.f);
  const float* src = params;
  for (int l = 0; l < p->layers; l++) {
    const int in = l == 0 ? p->inputs : p->lin[l], out_w = p->lout[l];
    for (int k = 0; k < in; k++)
      for (int n = 0; n < out_w; n++) flat[p->w_off[l] + (long)k * out_w + n] = src[(long)k * out_w + n];
    src += (long)in * out_w;
    for (int n = 0; n < out_w; n++) flat[p->b_off[l] + n] = src[n];
    src += out_w;
  }
  HIP_TRY(hipMemcpyAsync(p->w, flat.data(), flat.size() * sizeof(float), hipMemcpyHostToDevice, p->stream), "upload");
  mlp_mirror(p, true);
  HIP_TRY(hipMemsetAsync(p->m, 0, p->n_params * sizeof(float), p->stream), "reset");
  HIP_TRY(hipMemsetAsync(p->v, 0, p->n_params * sizeof(float), p->stream), "reset");
  HIP_TRY(hipStreamSynchronize(p->stream), "upload");
  p->step = 0;
  return BS_OK;
}

static bs_status mlp_read_flat(bs_mlp* p, const float* dev, float* params) {
  std::vector<float> flat(p->n_params);
  HIP_TRY(hipMemcpyAsync(flat.data(), dev, flat.size() * sizeof(float), hipMemcpyDeviceToHost, p->stream), "readback");
  HIP_TRY(hipStreamSynchronize(p->stream), "readback");
  float* dst = params;
  for (int l = 0; l < p->layers; l++) {
    const int in = l == 0 ? p->inputs : p->lin[l], out_w = p->lout[l];
    for (int k = 0; k < in; k++)
      for (int n = 0; n < out_w; n++) dst[(long)k * out_w + n] = flat[p->w_off[l] + (long)k * out_w + n];
    dst += (long)in * out_w;
    for (int n = 0; n < out_w; n++) dst[n] = flat[p->b_off[l] + n];
    dst += out_w;
  }
  return BS_OK;
}

bs_status bs_mlp_get_params(bs_mlp* p, float* params, int64_t count) {
  MLP_GUARD(p);
  if (!params || count != external_count(p)) return fail(BS_ERR_RANGE, "expected %ld parameters", external_count(p));
  return mlp_read_flat(p, p->w, params);
}

bs_status bs_mlp_load_data(bs_mlp* p, const float* x, const float* y, int32_t rows) {
  MLP_GUARD(p);
  if (!x || !y || rows < 1) return fail(BS_ERR_RANGE, "empty data set");
  if (p->data_x) (void)hipFree(p->data_x);
  if (p->data_y) (void)hipFree(p->data_y);
  p->data_x = p->data_y = nullptr;
  p->data_rows = 0;
  bs_status st;
  if ((st = dev_alloc(&p->data_x, (size_t)rows * p->inputs, "data set")) ||
      (st = dev_alloc(&p->data_y, (size_t)rows * p->outputs, "data set")))
    return st;
  HIP_TRY(copy_on(p->stream, p->data_x, x, (size_t)rows * p->inputs * sizeof(float), hipMemcpyHostToDevice), "data upload");
  HIP_TRY(copy_on(p->stream, p->data_y, y, (size_t)rows * p->outputs * sizeof(float), hipMemcpyHostToDevice), "data upload");
  p->data_rows = rows;
  return BS_OK;
}

bs_status bs_mlp_gradients(bs_mlp* p, const int32_t* idx, int32_t batch, float* grads, double* loss) {
  MLP_GUARD(p);
  if (!idx) return fail(BS_ERR_RANGE, "no indices");
  bs_status st = mlp_stage_indices(p, idx, batch);
  if (st) return st;
  mlp_forward(p, batch);
  if ((st = mlp_backward(p, batch, loss))) return st;
  return grads ? mlp_read_flat(p, p->grad, grads) : BS_OK;
}

bs_status bs_mlp_train_batch(bs_mlp* p, const int32_t* idx, int32_t batch, double lr, double* loss) {
  MLP_GUARD(p);
  if (!idx || !(lr > 0)) return fail(BS_ERR_RANGE, "no indices or learning rate");
  bs_status st = mlp_stage_indices(p, idx, batch);
  if (st) return st;
  mlp_forward(p, batch);
  if ((st = mlp_backward(p, batch, loss))) return st;
  p->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)p->step), c2 = 1.f - powf(b2, (float)p->step);
  hipLaunchKernelGGL(adam_step, dim3((unsigned)((p->n_params + 255) / 256)), dim3(256), 0, p->stream, p->w, p->grad,
                     p->m, p->v, p->wb, p->n_params, (float)lr, b1, b2, 1e-8f, c1, c2);
  mlp_mirror(p, false);
  HIP_TRY(hipGetLastError(), "adam launch");
  HIP_TRY(hipStreamSynchronize(p->stream), "adam step");
  return BS_OK;
}

bs_status bs_mlp_predict(bs_mlp* p, const float* x, int32_t rows, float* y) {
  MLP_GUARD(p);
  if (!x || !y || rows < 0) return fail(BS_ERR_RANGE, "null argument");
  for (int r0 = 0; r0 < rows; r0 += p->max_batch) {
    const int n = rows - r0 < p->max_batch ? rows - r0 : p->max_batch;
    const int pad = round64(n);
    HIP_TRY(hipMemcpyAsync(p->stage_x, x + (long)r0 * p->inputs, (size_t)n * p->inputs * sizeof(float),
                           hipMemcpyHostToDevice, p->stream),
            "input upload");
    const long ne = (long)pad * p->in_pad;
    hipLaunchKernelGGL(gather_inputs, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, p->stream, p->stage_x,
                       (const int*)nullptr, n, p->inputs, p->in_pad, pad, p->x);
    mlp_forward(p, pad);
    HIP_TRY(hipMemcpyAsync(y + (long)r0 * p->outputs, p->z[p->layers - 1], (size_t)n * p->outputs * sizeof(float),
                           hipMemcpyDeviceToHost, p->stream),
            "output readback");
  }
  HIP_TRY(hipGetLastError(), "network launch");
  HIP_TRY(hipStreamSynchronize(p->stream), "predict");
  return BS_OK;
}

bs_status bs_mlp_time_gemm(bs_mlp* p, int32_t layer, int32_t reps, double* ms_per_gemm, double* flops_per_gemm) {
  MLP_GUARD(p);
  if (layer < 0 || layer >= p->layers || reps < 1) return fail(BS_ERR_RANGE, "no such layer");
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0), "event");
  HIP_TRY(hipEventCreate(&e1), "event");
  GemmArgs g{};
  g.A = layer == 0 ? p->x : p->act[layer - 1];
  g.lda_m = p->lin[layer];
  g.lda_k = 1;
  g.B = p->wbt + p->w_off[layer];
  g.ldb_k = 1;
  g.ldb_n = p->lin[layer];
  g.C = p->z[layer];
  g.ldc = p->lout[layer];
  g.bias = p->w + p->b_off[layer];
  g.M = p->max_batch;
  g.N = p->lout[layer];
  g.K = p->lin[layer];
  launch_gemm<EPI_BIAS>(p, g, true, false);
  (void)hipEventRecord(e0, p->stream);
  for (int r = 0; r < reps; r++) launch_gemm<EPI_BIAS>(p, g, true, false);
  (void)hipEventRecord(e1, p->stream);
  HIP_TRY(hipEventSynchronize(e1), "gemm timing");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_gemm) *ms_per_gemm = ms / reps;
  if (flops_per_gemm) *flops_per_gemm = 2.0 * g.M * (double)g.N * g.K;
  return BS_OK;
}

}  // extern "C"
