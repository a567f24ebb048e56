// mh_nn_kernels.h -- device kernels shared by the surrogate networks
// (mlp.hip, cnn.hip): a matrix-core GEMM (bf16 operands, FP32 accumulate)
// with bias/GELU epilogues, the squared-error loss gradient, fixed-order column sums (bias
// gradients), Adam, and small helpers.  Every reduction runs in a fixed
// order, so training is deterministic.
//
// gelu(x) = 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).  Adam follows
// Kingma & Ba (2015), Algorithm 1 (bias-corrected moments).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "mercury_hip.h"
#include "mh_internal.h"

namespace {

constexpr int kTile = 64;
constexpr float kGeluC = 0.7978845608028654f;

__device__ __forceinline__ float gelu_f(float x) {
  const float u = kGeluC * (x + 0.044715f * x * x * x);
  return 0.5f * x * (1.0f + tanhf(u));
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float u = kGeluC * (x + 0.044715f * x * x * x);
  const float th = tanhf(u);
  return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * kGeluC * (1.0f + 3.0f * 0.044715f * x * x);
}

template <typename T>
__device__ __forceinline__ float ld(const T* p, size_t i);
template <>
__device__ __forceinline__ float ld<float>(const float* p, size_t i) { return p[i]; }
template <>
__device__ __forceinline__ float ld<__hip_bfloat16>(const __hip_bfloat16* p, size_t i) {
  return __bfloat162float(p[i]);
}

// bf16 operands for the matrix cores: round to nearest even, as
// __float2bfloat16 does (finite inputs).
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  const uint32_t u = __float_as_uint(x);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
template <typename T>
__device__ __forceinline__ uint16_t ld_bf16(const T* p, size_t i);
template <>
__device__ __forceinline__ uint16_t ld_bf16<float>(const float* p, size_t i) { return bf16_bits(p[i]); }
template <>
__device__ __forceinline__ uint16_t ld_bf16<__hip_bfloat16>(const __hip_bfloat16* p, size_t i) {
  return __builtin_bit_cast(uint16_t, p[i]);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// One 64 x 64 tile of op(A) op(B) over k in [k_begin, k_end) on the matrix
// cores (v_mfma_f32_16x16x32_bf16: bf16 operands, fp32 accumulate), 256
// threads.  op(A)[m][k] = TRA ? A[k*lda + m] : A[m*lda + k], op(B)[k][n] =
// TRB ? B[n*ldb + k] : B[k*ldb + n]; elements outside M, N or the k range
// are zeros.  Each step stages a 64 x 32 slice of both operands in LDS,
// rounded to bf16 and k-contiguous per row of op(A) and per column of
// op(B); wave w owns the 32 x 32 quarter (32 (w / 2), 32 (w % 2)) as 2 x 2
// blocks of 16 x 16.  Lane l feeds row (or column) l % 16 of a block with
// the 8 k values 8 (l / 16) .. 8 (l / 16) + 7 and receives rows
// 4 (l / 16) .. 4 (l / 16) + 3 of column l % 16.  Every output sums its k
// in the same order on every run.
constexpr int kMk = 32;                  // k per step
constexpr int kMRow = kMk + 8;           // LDS row pitch in bf16 (80 B, 16 B aligned)
template <typename TA, typename TB, bool TRA, bool TRB>
__device__ __forceinline__ void mfma_tile(const TA* __restrict__ A, int lda, const TB* __restrict__ B, int ldb,
                                          int M, int N, int k_begin, int k_end, int m0, int n0,
                                          uint16_t (*sa)[kMRow], uint16_t (*sb)[kMRow], f32x4_t (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = k_begin; k0 < k_end; k0 += kMk) {
    // 64 x 32 elements of each operand, 8 per thread, read along the
    // operand's contiguous dimension
    for (int e = threadIdx.x; e < kTile * kMk; e += 256) {
      int mm, kk;
      if (TRA) { mm = e % kTile; kk = e / kTile; } else { kk = e % kMk; mm = e / kMk; }
      const int m = m0 + mm, k = k0 + kk;
      uint16_t va = 0;
      if (m < M && k < k_end)
        va = TRA ? ld_bf16(A, static_cast<size_t>(k) * lda + m) : ld_bf16(A, static_cast<size_t>(m) * lda + k);
      sa[mm][kk] = va;
      int nn, kb;
      if (TRB) { kb = e % kMk; nn = e / kMk; } else { nn = e % kTile; kb = e / kTile; }
      const int n = n0 + nn, kq = k0 + kb;
      uint16_t vb = 0;
      if (n < N && kq < k_end)
        vb = TRB ? ld_bf16(B, static_cast<size_t>(n) * ldb + kq) : ld_bf16(B, static_cast<size_t>(kq) * ldb + n);
      sb[nn][kb] = vb;
    }
    __syncthreads();
    bf16x8_t fa[2], fb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      fa[i] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u16x8_t*>(&sa[wm + 16 * i + (lane & 15)][8 * (lane >> 4)]));
      fb[i] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u16x8_t*>(&sb[wn + 16 * i + (lane & 15)][8 * (lane >> 4)]));
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
}

// Output position of accumulator (i, j, r) of this lane in the tile.
__device__ __forceinline__ int mfma_row(int i, int r) {
  return ((threadIdx.x >> 6) >> 1) * 32 + 16 * i + 4 * ((threadIdx.x & 63) >> 4) + r;
}
__device__ __forceinline__ int mfma_col(int j) { return ((threadIdx.x >> 6) & 1) * 32 + 16 * j + (threadIdx.x & 15); }

// C[M x N] = op(A) op(B) (mfma_tile's operands, bf16 products, fp32 sums).
// Epilogue EPI: 0 = store; 1 = + bias[n], store Z and gelu(Z) (hidden layer
// forward); 2 = + bias[n], store (output layer forward); 3 = multiply by
// gelu'(Zprev[m][n]) (hidden layer backward); 4 = + bias[n], store gelu
// (hidden layer forward without keeping Z: the same values as 1 stores in C).
template <typename TA, typename TB, bool TRA, bool TRB, int EPI>
__global__ __launch_bounds__(256) void gemm(const TA* __restrict__ A, int lda, const TB* __restrict__ B,
                                            int ldb, float* __restrict__ C, int ldc, int M, int N, int K,
                                            const float* __restrict__ bias, float* __restrict__ Z,
                                            const float* __restrict__ Zprev) {
  __shared__ __align__(16) uint16_t sa[kTile][kMRow];
  __shared__ __align__(16) uint16_t sb[kTile][kMRow];
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  f32x4_t acc[2][2];
  mfma_tile<TA, TB, TRA, TRB>(A, lda, B, ldb, M, N, 0, K, m0, n0, sa, sb, acc);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + mfma_row(i, r), n = n0 + mfma_col(j);
        if (m >= M || n >= N) continue;
        const size_t o = static_cast<size_t>(m) * ldc + n;
        float v = acc[i][j][r];
        if (EPI == 1) {
          v += bias[n];
          Z[o] = v;
          v = gelu_f(v);
        } else if (EPI == 2) {
          v += bias[n];
        } else if (EPI == 3) {
          v *= gelu_grad(Zprev[o]);
        } else if (EPI == 4) {
          v = gelu_f(v + bias[n]);
        }
        C[o] = v;
      }
}

// dY = (Y - T) / B, per-row loss terms 0.5 (Y - T)^2
__global__ void loss_grad(const float* Y, const float* T, int n, float inv_b, float* dY, float* terms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float e = Y[i] - T[i];
  dY[i] = e * inv_b;
  terms[i] = 0.5f * e * e;
}

// out[n] = sum over rows of X[m][n] (row order fixed)
__global__ void col_sum(const float* X, int M, int N, float* out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += X[static_cast<size_t>(m) * N + n];
  out[n] = s;
}

__global__ void adam_step(float* p, const float* g, float* m, float* v, __hip_bfloat16* w16, int n,
                          float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  p[i] = pi;
  if (w16) w16[i] = __float2bfloat16(pi);
}

__global__ void to_bf16(const float* src, __hip_bfloat16* dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = __float2bfloat16(src[i]);
}

__global__ void gather_rows(const float* src, const int* idx, int n_rows, int width, float* dst) {
  const int r = blockIdx.x, c = threadIdx.x;
  if (r >= n_rows) return;
  for (int j = c; j < width; j += blockDim.x)
    dst[static_cast<size_t>(r) * width + j] = src[static_cast<size_t>(idx[r]) * width + j];
}

void hipcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw mh::Error(MH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  void ensure(size_t count) {
    if (count <= n) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    hipcheck(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
    n = count;
  }
};

dim3 grid_for(int M, int N) { return dim3((N + kTile - 1) / kTile, (M + kTile - 1) / kTile); }

}  // namespace
