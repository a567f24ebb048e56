// mh_nn_kernels.h -- device kernels shared by the surrogate networks
// (mlp.hip, cnn.hip): a matrix-core GEMM (bf16 operands, FP32 accumulate)
// with bias/GELU epilogues, the squared-error loss gradient, fixed-order column sums (bias
// gradients), split-K weight gradients, Adam, and small helpers.  Every reduction runs in a fixed
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

// bf16 bits of 8 consecutive elements at p: 16-byte loads where p is
// 16-byte aligned, element loads otherwise.
__device__ __forceinline__ u16x8_t ld8_bf16_a16(const float* p) {
  const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
  return u16x8_t{bf16_bits(a.x), bf16_bits(a.y), bf16_bits(a.z), bf16_bits(a.w),
                 bf16_bits(b.x), bf16_bits(b.y), bf16_bits(b.z), bf16_bits(b.w)};
}
__device__ __forceinline__ u16x8_t ld8_bf16(const float* p) {
  if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0) return ld8_bf16_a16(p);
  u16x8_t r;
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = bf16_bits(p[j]);
  return r;
}
__device__ __forceinline__ u16x8_t ld8_bf16(const __hip_bfloat16* p) {
  if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0) return *reinterpret_cast<const u16x8_t*>(p);
  u16x8_t r;
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = __builtin_bit_cast(uint16_t, p[j]);
  return r;
}

// 8 consecutive elements as loaded, converted to bf16 bits only by
// bits8(): a tile step issues the loads of all its slices before the first
// conversion waits for one.  p must be 16-byte aligned.
template <typename T>
struct Raw8;
template <>
struct Raw8<float> {
  float4 a, b;
};
template <>
struct Raw8<__hip_bfloat16> {
  u16x8_t v;
};
__device__ __forceinline__ Raw8<float> ld_raw8(const float* p) {
  return Raw8<float>{reinterpret_cast<const float4*>(p)[0], reinterpret_cast<const float4*>(p)[1]};
}
__device__ __forceinline__ Raw8<__hip_bfloat16> ld_raw8(const __hip_bfloat16* p) {
  return Raw8<__hip_bfloat16>{*reinterpret_cast<const u16x8_t*>(p)};
}
__device__ __forceinline__ u16x8_t bits8(const Raw8<float>& r) {
  return u16x8_t{bf16_bits(r.a.x), bf16_bits(r.a.y), bf16_bits(r.a.z), bf16_bits(r.a.w),
                 bf16_bits(r.b.x), bf16_bits(r.b.y), bf16_bits(r.b.z), bf16_bits(r.b.w)};
}
__device__ __forceinline__ u16x8_t bits8(const Raw8<__hip_bfloat16>& r) { return r.v; }

// Operand loaders of the matrix-core tile: A(m, k) and B(k, n) return the
// bf16 bits of op(A)[m][k] and op(B)[k][n] (callers stay within M, N and
// the k range).  kContigK says which index is contiguous in memory (k, or
// m / n), so a tile's elements are read along it.  A loader with kVec8
// also reads 8 consecutive elements along that index in one call,
// vec8(a, b), with (a, b) the first element's indices in the order of
// operator(); callers use it only when all 8 lie inside the operand.
//   StridedA: op(A)[m][k] = TR ? p[k*ld + m] : p[m*ld + k]
//   StridedB: op(B)[k][n] = TR ? p[n*ld + k] : p[k*ld + n]
template <typename T, bool TR>
struct StridedA {
  const T* p;
  int ld;
  static constexpr bool kContigK = !TR;
  static constexpr bool kVec8 = true;
  __device__ uint16_t operator()(int m, int k) const {
    return TR ? ld_bf16(p, static_cast<size_t>(k) * ld + m) : ld_bf16(p, static_cast<size_t>(m) * ld + k);
  }
  __device__ u16x8_t vec8(int m, int k) const {
    return ld8_bf16(p + (TR ? static_cast<size_t>(k) * ld + m : static_cast<size_t>(m) * ld + k));
  }
  // every 8-element run a tile step reads starts 16-byte aligned
  __device__ bool aligned16() const { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0 && ld % 8 == 0; }
  __device__ Raw8<T> raw8(int m, int k) const {
    return ld_raw8(p + (TR ? static_cast<size_t>(k) * ld + m : static_cast<size_t>(m) * ld + k));
  }
};
template <typename T, bool TR>
struct StridedB {
  const T* p;
  int ld;
  static constexpr bool kContigK = TR;
  static constexpr bool kVec8 = true;
  __device__ uint16_t operator()(int k, int n) const {
    return TR ? ld_bf16(p, static_cast<size_t>(n) * ld + k) : ld_bf16(p, static_cast<size_t>(k) * ld + n);
  }
  __device__ u16x8_t vec8(int k, int n) const {
    return ld8_bf16(p + (TR ? static_cast<size_t>(n) * ld + k : static_cast<size_t>(k) * ld + n));
  }
  __device__ bool aligned16() const { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0 && ld % 8 == 0; }
  __device__ Raw8<T> raw8(int k, int n) const {
    return ld_raw8(p + (TR ? static_cast<size_t>(n) * ld + k : static_cast<size_t>(k) * ld + n));
  }
};

// One 64 x 32 slice of an operand (rows along m or n, k contiguous), 8
// elements per thread of 256: fetch_slice reads this thread's 8 as bf16
// bits (zeros outside ROWS and the k range), raw_at reads them unconverted
// with no checks, put_slice writes them to LDS.  ROWS = M (A) or N (B); A
// loaders take (row, k), B loaders (k, row).
template <bool IS_A, typename L>
__device__ __forceinline__ uint16_t load_at(const L& l, int row, int k) {
  if constexpr (IS_A) return l(row, k);
  else return l(k, row);
}
template <bool IS_A, typename L>
__device__ __forceinline__ u16x8_t vec8_at(const L& l, int row, int k) {
  if constexpr (IS_A) return l.vec8(row, k);
  else return l.vec8(k, row);
}
template <bool IS_A, typename L>
__device__ __forceinline__ auto raw_at(const L& l, int r0, int k0) {
  if constexpr (L::kContigK) {
    const int r = r0 + (threadIdx.x >> 2), k = k0 + ((threadIdx.x & 3) << 3);
    if constexpr (IS_A) return l.raw8(r, k);
    else return l.raw8(k, r);
  } else {
    const int r = r0 + ((threadIdx.x & 7) << 3), k = k0 + (threadIdx.x >> 3);
    if constexpr (IS_A) return l.raw8(r, k);
    else return l.raw8(k, r);
  }
}
template <bool IS_A, typename L>
__device__ __forceinline__ u16x8_t fetch_slice(const L& l, int ROWS, int r0, int k0, int k_end) {
  u16x8_t v;
  if constexpr (L::kVec8) {
    if constexpr (L::kContigK) {
      const int r = r0 + (threadIdx.x >> 2), k = k0 + ((threadIdx.x & 3) << 3);
      if (r < ROWS && k + 8 <= k_end) v = vec8_at<IS_A>(l, r, k);
      else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (r < ROWS && k + j < k_end) ? load_at<IS_A>(l, r, k + j) : uint16_t(0);
      }
    } else {
      const int r = r0 + ((threadIdx.x & 7) << 3), k = k0 + (threadIdx.x >> 3);
      if (r + 8 <= ROWS && k < k_end) v = vec8_at<IS_A>(l, r, k);
      else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (r + j < ROWS && k < k_end) ? load_at<IS_A>(l, r + j, k) : uint16_t(0);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = threadIdx.x + 256 * j;
      const int rr = L::kContigK ? e / 32 : e % 64, kk = L::kContigK ? e % 32 : e / 64;
      const int r = r0 + rr, k = k0 + kk;
      v[j] = (r < ROWS && k < k_end) ? load_at<IS_A>(l, r, k) : uint16_t(0);
    }
  }
  return v;
}
template <typename L>
__device__ __forceinline__ void put_slice(u16x8_t v, uint16_t (*s)[40]) {
  if constexpr (L::kVec8) {
    if constexpr (L::kContigK) {
      *reinterpret_cast<u16x8_t*>(&s[threadIdx.x >> 2][(threadIdx.x & 3) << 3]) = v;
    } else {
      const int kk = threadIdx.x >> 3, rr = (threadIdx.x & 7) << 3;
#pragma unroll
      for (int j = 0; j < 8; ++j) s[rr + j][kk] = v[j];
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = threadIdx.x + 256 * j;
      if constexpr (L::kContigK) s[e / 32][e % 32] = v[j];
      else s[e % 64][e / 64] = v[j];
    }
  }
}

// One 64 x 64 tile of op(A) op(B) over k in [k_begin, k_end) on the matrix
// cores (v_mfma_f32_16x16x32_bf16: bf16 operands, fp32 accumulate), 256
// threads.  Elements outside M, N or the k range are zeros.  Each step
// stages S slices of 64 x 32 of both operands in LDS as bf16, k-contiguous
// per row of op(A) and per column of op(B), between two barriers.  With
// FAST (strided operands), a tile whose rows lie inside M and N, whose k
// range is whole slices and whose operand runs are 16-byte aligned takes a
// branch-free path that issues the loads of all S slices before the first
// conversion to bf16 waits for one; otherwise each thread's 8 elements are
// read with per-element bounds checks.  Wave w owns the 32 x 32 quarter
// (32 (w / 2), 32 (w % 2)) as 2 x 2 blocks of 16 x 16.  Lane l feeds row
// (or column) l % 16 of a block with the 8 k values 8 (l / 16) ..
// 8 (l / 16) + 7 and receives rows 4 (l / 16) .. 4 (l / 16) + 3 of column
// l % 16.  Every output sums its k in ascending slices of 32, the same
// order on every path and every run.
constexpr int kMk = 32;                  // k per slice
constexpr int kMRow = kMk + 8;           // LDS row pitch in bf16 (80 B, 16 B aligned)
// slices per step: 2 for the dense layers' GEMMs (k of 64 or 256); the
// convolution kernels' gathered operands take one (a deeper stage lowers
// their occupancy without shortening them)
constexpr int kStageDense = 2, kStageConv = 1, kStageWgrad = 1;
template <int S, bool FAST, typename LA, typename LB>
__device__ __forceinline__ void mfma_tile(const LA& la, const LB& lb, int M, int N, int k_begin, int k_end, int m0,
                                          int n0, uint16_t (*sa)[kTile][kMRow], uint16_t (*sb)[kTile][kMRow],
                                          f32x4_t (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  bool fast = false;
  if constexpr (FAST)
    fast = m0 + kTile <= M && n0 + kTile <= N && (k_end - k_begin) % kMk == 0 && k_begin % 8 == 0 &&
           la.aligned16() && lb.aligned16();
  for (int k0 = k_begin; k0 < k_end; k0 += S * kMk) {
    u16x8_t va[S], vb[S];
    bool staged = false;
    if constexpr (FAST) {
      if (fast) {
        // whole slices in range and aligned: every load of the step first,
        // then the conversions
        decltype(raw_at<true>(la, m0, k0)) ra[S];
        decltype(raw_at<false>(lb, n0, k0)) rb[S];
#pragma unroll
        for (int s = 0; s < S; ++s)
          if (k0 + s * kMk < k_end) {
            ra[s] = raw_at<true>(la, m0, k0 + s * kMk);
            rb[s] = raw_at<false>(lb, n0, k0 + s * kMk);
          }
#pragma unroll
        for (int s = 0; s < S; ++s) {
          va[s] = k0 + s * kMk < k_end ? bits8(ra[s]) : u16x8_t{};
          vb[s] = k0 + s * kMk < k_end ? bits8(rb[s]) : u16x8_t{};
        }
        staged = true;
      }
    }
    if (staged || S > 1) {
      if (!staged) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
          va[s] = fetch_slice<true>(la, M, m0, k0 + s * kMk, k_end);
          vb[s] = fetch_slice<false>(lb, N, n0, k0 + s * kMk, k_end);
        }
      }
#pragma unroll
      for (int s = 0; s < S; ++s) {
        put_slice<LA>(va[s], sa[s]);
        put_slice<LB>(vb[s], sb[s]);
      }
    } else {
      // one slice with checks: each operand written as soon as it is read
      // (measured 10 % faster for the convolution weight gradients than
      // reading both first)
      put_slice<LA>(fetch_slice<true>(la, M, m0, k0, k_end), sa[0]);
      put_slice<LB>(fetch_slice<false>(lb, N, n0, k0, k_end), sb[0]);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (k0 + s * kMk < k_end) {
        bf16x8_t fa[2], fb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          fa[i] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u16x8_t*>(&sa[s][wm + 16 * i + (lane & 15)][8 * (lane >> 4)]));
          fb[i] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u16x8_t*>(&sb[s][wn + 16 * i + (lane & 15)][8 * (lane >> 4)]));
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
}

// Output position of accumulator (i, j, r) of this lane in the tile.
__device__ __forceinline__ int mfma_row(int i, int r) {
  return ((threadIdx.x >> 6) >> 1) * 32 + 16 * i + 4 * ((threadIdx.x & 63) >> 4) + r;
}
__device__ __forceinline__ int mfma_col(int j) { return ((threadIdx.x >> 6) & 1) * 32 + 16 * j + (threadIdx.x & 15); }

// Writes a tile's accumulators to C with epilogue EPI: 0 = store; 1 =
// + bias[n], store Z and gelu(Z) (hidden layer forward); 2 = + bias[n],
// store (output layer forward); 3 = multiply by gelu'(Zprev[m][n]) (hidden
// layer backward); 4 = + bias[n], store gelu (hidden layer forward without
// keeping Z: the same values as 1 stores in C).
template <int EPI>
__device__ __forceinline__ void mfma_store(const f32x4_t (&acc)[2][2], float* __restrict__ C, int ldc, int M, int N,
                                           int m0, int n0, const float* __restrict__ bias, float* __restrict__ Z,
                                           const float* __restrict__ Zprev) {
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

// C[M x N] = op(A) op(B) (bf16 products, fp32 sums) with loaders la, lb,
// then epilogue EPI.
template <typename LA, typename LB, int EPI>
__global__ __launch_bounds__(256) void gemm_l(LA la, LB lb, float* __restrict__ C, int ldc, int M, int N, int K,
                                              const float* __restrict__ bias, float* __restrict__ Z,
                                              const float* __restrict__ Zprev) {
  __shared__ __align__(16) uint16_t sa[kStageConv][kTile][kMRow];
  __shared__ __align__(16) uint16_t sb[kStageConv][kTile][kMRow];
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  f32x4_t acc[2][2];
  mfma_tile<kStageConv, false>(la, lb, M, N, 0, K, m0, n0, sa, sb, acc);
  mfma_store<EPI>(acc, C, ldc, M, N, m0, n0, bias, Z, Zprev);
}

// The same for strided operands: op(A)[m][k] = TRA ? A[k*lda + m] :
// A[m*lda + k], op(B)[k][n] = TRB ? B[n*ldb + k] : B[k*ldb + n].
template <typename TA, typename TB, bool TRA, bool TRB, int EPI>
__global__ __launch_bounds__(256) void gemm(const TA* __restrict__ A, int lda, const TB* __restrict__ B,
                                            int ldb, float* __restrict__ C, int ldc, int M, int N, int K,
                                            const float* __restrict__ bias, float* __restrict__ Z,
                                            const float* __restrict__ Zprev) {
  __shared__ __align__(16) uint16_t sa[kStageDense][kTile][kMRow];
  __shared__ __align__(16) uint16_t sb[kStageDense][kTile][kMRow];
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  f32x4_t acc[2][2];
  mfma_tile<kStageDense, true>(StridedA<TA, TRA>{A, lda}, StridedB<TB, TRB>{B, ldb}, M, N, 0, K, m0, n0, sa, sb, acc);
  mfma_store<EPI>(acc, C, ldc, M, N, m0, n0, bias, Z, Zprev);
}

// dY = (Y - T) / B, per-row loss terms 0.5 (Y - T)^2
__global__ void loss_grad(const float* Y, const float* T, int n, float inv_b, float* dY, float* terms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float e = Y[i] - T[i];
  dY[i] = e * inv_b;
  terms[i] = 0.5f * e * e;
}

// Column sums in a fixed order: part[s split_stride + n] = sum of X[r][n]
// over the rows r of split s ([s rows, (s + 1) rows) within M).  A block of kColThreads
// takes kColTile columns: thread (g, c) adds rows g, g + kColGroups, ... of
// its split in ascending order, then one thread per column adds the
// kColGroups partials in ascending g.
constexpr int kColTile = 64, kColGroups = 16, kColThreads = kColTile * kColGroups;
__global__ __launch_bounds__(kColThreads) void col_partial(const float* __restrict__ X, int M, int N, int rows,
                                                           float* __restrict__ part, size_t split_stride) {
  __shared__ float acc[kColGroups][kColTile];
  const int c = threadIdx.x % kColTile, g = threadIdx.x / kColTile;
  const int n = blockIdx.x * kColTile + c;
  const int r0 = blockIdx.y * rows, r1 = min(M, r0 + rows);
  float sum = 0.f;
  if (n < N)
    for (int r = r0 + g; r < r1; r += kColGroups) sum += X[static_cast<size_t>(r) * N + n];
  acc[g][c] = sum;
  __syncthreads();
  if (g == 0 && n < N) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < kColGroups; ++k) t += acc[k][c];
    part[blockIdx.y * split_stride + n] = t;
  }
}

// Weight gradient with split-K: part[s split_stride + i Nj + j] = sum over
// rows r of split s of D[r][i] * X(r, j) on the matrix cores (bf16 operands,
// fp32 sums in a fixed order); D is [M x Ni] row-major, X a B-operand loader
// giving row r, column j of an [M x Nj] matrix.
template <typename LX>
__global__ __launch_bounds__(256) void wgrad_l(const float* __restrict__ Dm, int Ni, LX lx, int Nj, int M, int rows,
                                               float* __restrict__ part, size_t split_stride) {
  __shared__ __align__(16) uint16_t sd[kStageWgrad][kTile][kMRow];
  __shared__ __align__(16) uint16_t sx[kStageWgrad][kTile][kMRow];
  const int i0 = blockIdx.y * kTile, j0 = blockIdx.x * kTile;
  const int r_begin = blockIdx.z * rows, r_end = min(M, r_begin + rows);
  f32x4_t acc[2][2];
  // op(A)[i][r] = D[r][i]
  mfma_tile<kStageWgrad, false>(StridedA<float, true>{Dm, Ni}, lx, Ni, Nj, r_begin, r_end, i0, j0, sd, sx, acc);
  float* out = part + blockIdx.z * split_stride;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + mfma_row(a, q), j = j0 + mfma_col(b);
        if (i < Ni && j < Nj) out[static_cast<size_t>(i) * Nj + j] = acc[a][b][q];
      }
}

// out[e] = sum over splits s (ascending) of part[s][e]
__global__ void sum_splits(const float* part, int n, int splits, float* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.f;
  for (int k = 0; k < splits; ++k) s += part[static_cast<size_t>(k) * n + e];
  out[e] = s;
}

__device__ __forceinline__ void adam_update(int i, float gi, float* p, float* m, float* v, __hip_bfloat16* w16,
                                            float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  p[i] = pi;
  if (w16) w16[i] = __float2bfloat16(pi);
}

__global__ void adam_step(float* p, const float* g, float* m, float* v, __hip_bfloat16* w16, int n,
                          float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  adam_update(i, g[i], p, m, v, w16, lr, b1, b2, eps, bc1, bc2);
}

// Adam with each gradient summed from its split-K partials, part[k stride +
// i] for k = 0 .. splits - 1, in sum_splits' order: the same gradient
// without a sum_splits launch per weight and bias tensor.
__global__ void adam_step_splits(float* p, const float* part, int splits, size_t stride, float* m, float* v,
                                 __hip_bfloat16* w16, int n, float lr, float b1, float b2, float eps, float bc1,
                                 float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = 0.f;
  for (int k = 0; k < splits; ++k) g += part[k * stride + i];
  adam_update(i, g, p, m, v, w16, lr, b1, b2, eps, bc1, bc2);
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
