// mh_nn_kernels.h -- device kernels shared by the surrogate networks
// (mlp.hip, cnn.hip): a tiled FP32-accumulate GEMM with bias/GELU
// epilogues, the squared-error loss gradient, fixed-order column sums (bias
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
constexpr int kTk = 16;
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

// C[M x N] = op(A) op(B) with op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k]
// and op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n].  Epilogue EPI:
// 0 = store; 1 = + bias[n], store Z and gelu(Z) (hidden layer forward);
// 2 = + bias[n], store (output layer forward); 3 = multiply by
// gelu'(Zprev[m][n]) (hidden layer backward); 4 = + bias[n], store gelu
// (hidden layer forward without keeping Z: the same values as 1 stores in C).
template <typename TA, typename TB, bool TRA, bool TRB, int EPI>
__global__ __launch_bounds__(256) void gemm(const TA* __restrict__ A, int lda, const TB* __restrict__ B,
                                            int ldb, float* __restrict__ C, int ldc, int M, int N, int K,
                                            const float* __restrict__ bias, float* __restrict__ Z,
                                            const float* __restrict__ Zprev) {
  __shared__ float sa[kTk][kTile + 1];
  __shared__ float sb[kTk][kTile + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += kTk) {
    for (int e = threadIdx.x; e < kTile * kTk; e += 256) {
      const int mm = e % kTile, kk = e / kTile;
      const int m = m0 + mm, k = k0 + kk;
      float va = 0.f;
      if (m < M && k < K) va = TRA ? ld(A, static_cast<size_t>(k) * lda + m) : ld(A, static_cast<size_t>(m) * lda + k);
      sa[kk][mm] = va;
      const int n = n0 + mm;
      float vb = 0.f;
      if (n < N && k < K) vb = TRB ? ld(B, static_cast<size_t>(n) * ldb + k) : ld(B, static_cast<size_t>(k) * ldb + n);
      sb[kk][mm] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTk; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = sa[kk][ty * 4 + i];
        b[i] = sb[kk][tx * 4 + i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
      if (m >= M || n >= N) continue;
      const size_t o = static_cast<size_t>(m) * ldc + n;
      float v = acc[i][j];
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
