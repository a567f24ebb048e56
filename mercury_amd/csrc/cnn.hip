// cnn.hip -- convolutional surrogate for voxel genomes: a learned 64-d
// feature extractor in front of the surrogate MLP (64 -> 256 -> 256 -> 1).
//
// Input: a genome, one material index (0 = empty, 1..4) per voxel of an 8^3
// grid, voxel (x, y, z) at x + 8 (y + 8 z).  Network:
//   conv1  3x3x3, 5 (one-hot materials) -> 32, zero padding, GELU
//   conv2  3x3x3, 32 -> 32, GELU
//   pool   2x2x2 average, 8^3 -> 4^3
//   conv3  3x3x3, 32 -> 64, GELU
//   mean over the 4^3 positions -> 64 features
//   dense  64 -> 256 GELU -> 256 GELU -> 1
// Convolutions run as GEMMs over im2col rows (one row per output position,
// column k*C + c for kernel offset k = (dz+1)*9 + (dy+1)*3 + (dx+1)) whose
// operands the matrix-core tile gathers as it loads them, so no im2col
// matrix is written; the input gradient is a GEMM over (k, output channel)
// of the shifted output gradients; weight gradients reduce over the batch
// with split-K partials summed in a fixed order, so training is
// deterministic.  Weights
// are bf16 with an fp32 master copy for Adam; activations are fp32 and
// enter every GEMM rounded to bf16 (weights, im2col columns, gradients),
// with fp32 accumulation on the matrix cores.  Loss = mean 0.5 (y - target)^2.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "mercury_hip.h"
#include "mh_internal.h"
#include "mh_nn_kernels.h"

namespace {

constexpr int kG = 8, kP1 = kG * kG * kG;        // genome grid and positions
constexpr int kG2 = kG / 2, kP2 = kG2 * kG2 * kG2;
constexpr int kC0 = 5, kC1 = 32, kC2 = 32, kC3 = 64, kH = 256;
constexpr int kK1 = 27 * kC0, kK2 = 27 * kC1, kK3 = 27 * kC2;
constexpr int kSplit = 64;                        // split-K partials for weight gradients

__device__ __forceinline__ bool inside(int z, int y, int x, int D) {
  return z >= 0 && z < D && y >= 0 && y < D && x >= 0 && x < D;
}

// Convolutions run as GEMMs whose operands are gathered as the tile loads
// them (no im2col matrix in memory).  Row r = n * D^3 + pos of an
// activation [B * D^3][C]; column k*C + c of its im2col matrix is channel c
// at pos + offset(k), offset(k) = (k / 9 - 1, (k / 3) % 3 - 1, k % 3 - 1) in
// (z, y, x), zero outside the grid.
template <int D, int C>
__device__ __forceinline__ uint16_t conv_col(const float* act, int r, int kk) {
  constexpr int P = D * D * D;
  const int k = kk / C, c = kk - k * C;
  const int n = r / P, pos = r - n * P;
  const int x = pos % D + k % 3 - 1, y = (pos / D) % D + (k / 3) % 3 - 1, z = pos / (D * D) + k / 9 - 1;
  if (!inside(z, y, x, D)) return 0;
  return bf16_bits(act[(static_cast<size_t>(n) * P + (z * D + y) * D + x) * C + c]);
}
// The same for the genome itself: channel c of a voxel is 1 where it holds
// material c (0 = empty, 1..4).
__device__ __forceinline__ uint16_t onehot_col(const uint8_t* g, int r, int kk) {
  const int k = kk / kC0, c = kk - k * kC0;
  const int n = r / kP1, pos = r - n * kP1;
  const int x = pos % kG + k % 3 - 1, y = (pos / kG) % kG + (k / 3) % 3 - 1, z = pos / (kG * kG) + k / 9 - 1;
  if (!inside(z, y, x, kG)) return 0;
  return g[static_cast<size_t>(n) * kP1 + (z * kG + y) * kG + x] == c ? uint16_t(0x3f80) : uint16_t(0);
}
// Columns kk .. kk + 7 of row r (kk a multiple of 8, C a multiple of 8: one
// kernel offset, 8 consecutive channels, 32-byte aligned in the activation
// buffers).
template <int D, int C>
__device__ __forceinline__ u16x8_t conv_col8(const float* act, int r, int kk) {
  static_assert(C % 8 == 0, "8 channels of one kernel offset");
  constexpr int P = D * D * D;
  const int k = kk / C, c = kk - k * C;
  const int n = r / P, pos = r - n * P;
  const int x = pos % D + k % 3 - 1, y = (pos / D) % D + (k / 3) % 3 - 1, z = pos / (D * D) + k / 9 - 1;
  if (!inside(z, y, x, D)) return u16x8_t{0, 0, 0, 0, 0, 0, 0, 0};
  return ld8_bf16_a16(act + (static_cast<size_t>(n) * P + (z * D + y) * D + x) * C + c);
}
// im2col matrices as the forward GEMM's A operand (rows r, columns k*C + c)
// and as the weight gradient's X operand (the same matrix)
template <int D, int C>
struct ConvColsA {
  const float* act;
  static constexpr bool kContigK = true;
  static constexpr bool kVec8 = true;
  __device__ uint16_t operator()(int m, int k) const { return conv_col<D, C>(act, m, k); }
  __device__ u16x8_t vec8(int m, int k) const { return conv_col8<D, C>(act, m, k); }
};
template <int D, int C>
struct ConvColsB {
  const float* act;
  static constexpr bool kContigK = false;
  static constexpr bool kVec8 = true;
  __device__ uint16_t operator()(int k, int n) const { return conv_col<D, C>(act, k, n); }
  __device__ u16x8_t vec8(int k, int n) const { return conv_col8<D, C>(act, k, n); }
};
// Columns kk .. kk + 7 of row r of the one-hot im2col matrix: 8 columns
// cross kernel offsets (5 channels each), so the row is decoded once and
// the offset and channel advance column by column.
__device__ __forceinline__ u16x8_t onehot_col8(const uint8_t* g, int r, int kk) {
  const int n = r / kP1, pos = r - n * kP1;
  const int px = pos % kG, py = (pos / kG) % kG, pz = pos / (kG * kG);
  const uint8_t* gn = g + static_cast<size_t>(n) * kP1;
  int k = kk / kC0, c = kk - k * kC0;
  u16x8_t v;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    uint16_t e = 0;
    if (k < 27) {
      const int x = px + k % 3 - 1, y = py + (k / 3) % 3 - 1, z = pz + k / 9 - 1;
      if (inside(z, y, x, kG) && gn[(z * kG + y) * kG + x] == c) e = 0x3f80;
    }
    v[j] = e;
    if (++c == kC0) {
      c = 0;
      ++k;
    }
  }
  return v;
}
struct OneHotA {
  const uint8_t* g;
  static constexpr bool kContigK = true;
  static constexpr bool kVec8 = true;
  __device__ uint16_t operator()(int m, int k) const { return onehot_col(g, m, k); }
  __device__ u16x8_t vec8(int m, int k) const { return onehot_col8(g, m, k); }
};
struct OneHotB {
  const uint8_t* g;
  static constexpr bool kContigK = false;
  static constexpr bool kVec8 = true;
  __device__ uint16_t operator()(int k, int n) const { return onehot_col(g, k, n); }
  __device__ u16x8_t vec8(int k, int n) const { return onehot_col8(g, k, n); }
};
// Input gradient of a convolution, d[n, pos, c] = sum over k, o of
// dZ[n, pos - offset(k), o] W[o][k*C + c] (positions inside the grid), as a
// GEMM over k*O + o: A gathers the shifted output gradients, B reads the
// weights [O x 27 C] re-indexed.
template <int D, int O>
struct ShiftedGrad {
  const float* dz;
  static constexpr bool kContigK = true;
  static constexpr bool kVec8 = O % 8 == 0;
  __device__ int row(int m, int k) const {  // -1 outside the grid
    constexpr int P = D * D * D;
    const int n = m / P, pos = m - n * P;
    const int x = pos % D - (k % 3 - 1), y = (pos / D) % D - ((k / 3) % 3 - 1), z = pos / (D * D) - (k / 9 - 1);
    return inside(z, y, x, D) ? n * P + (z * D + y) * D + x : -1;
  }
  __device__ uint16_t operator()(int m, int kk) const {
    const int k = kk / O, o = kk - k * O;
    const int r = row(m, k);
    return r < 0 ? uint16_t(0) : bf16_bits(dz[static_cast<size_t>(r) * O + o]);
  }
  // kk a multiple of 8: 8 consecutive output channels of one offset
  __device__ u16x8_t vec8(int m, int kk) const {
    const int k = kk / O, o = kk - k * O;
    const int r = row(m, k);
    return r < 0 ? u16x8_t{0, 0, 0, 0, 0, 0, 0, 0} : ld8_bf16_a16(dz + static_cast<size_t>(r) * O + o);
  }
};
template <int C, int O>
struct ConvWeightT {
  const __hip_bfloat16* w;
  static constexpr bool kContigK = false;
  static constexpr bool kVec8 = C % 8 == 0;
  __device__ uint16_t operator()(int kk, int c) const {
    const int k = kk / O, o = kk - k * O;
    return ld_bf16(w, static_cast<size_t>(o) * (27 * C) + k * C + c);
  }
  // c a multiple of 8: 8 consecutive input channels
  __device__ u16x8_t vec8(int kk, int c) const {
    const int k = kk / O, o = kk - k * O;
    return ld8_bf16(w + static_cast<size_t>(o) * (27 * C) + k * C + c);
  }
};

// 2x2x2 average pooling 8^3 -> 4^3: the 8 inputs summed z, y, x ascending
__global__ void pool2_fwd(const float* a, int B, float* p) {
  const size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<size_t>(B) * kP2 * kC2) return;
  const int c = static_cast<int>(t % kC2);
  const size_t r = t / kC2;
  const int n = static_cast<int>(r / kP2), q = static_cast<int>(r % kP2);
  const int X = q % kG2, Y = (q / kG2) % kG2, Z = q / (kG2 * kG2);
  float s = 0.f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx)
        s += a[(static_cast<size_t>(n) * kP1 + ((2 * Z + dz) * kG + 2 * Y + dy) * kG + 2 * X + dx) * kC2 + c];
  p[t] = 0.125f * s;
}

// pooling backward fused with conv2's activation: dZ2 = dp / 8 * gelu'(Z2)
__global__ void pool2_bwd_gelu(const float* dp, const float* Z, int B, float* dZ) {
  const size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<size_t>(B) * kP1 * kC2) return;
  const int c = static_cast<int>(t % kC2);
  const size_t r = t / kC2;
  const int n = static_cast<int>(r / kP1), pos = static_cast<int>(r % kP1);
  const int x = pos % kG, y = (pos / kG) % kG, z = pos / (kG * kG);
  const int q = ((z / 2) * kG2 + y / 2) * kG2 + x / 2;
  dZ[t] = 0.125f * dp[(static_cast<size_t>(n) * kP2 + q) * kC2 + c] * gelu_grad(Z[t]);
}

// mean over the 4^3 positions (ascending) -> features [B][64]
__global__ void gap_fwd(const float* a, int B, float* f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kC3) return;
  const int n = t / kC3, c = t % kC3;
  float s = 0.f;
  for (int q = 0; q < kP2; ++q) s += a[(static_cast<size_t>(n) * kP2 + q) * kC3 + c];
  f[t] = s * (1.0f / kP2);
}

// mean backward fused with conv3's activation: dZ3 = df / 64 * gelu'(Z3)
__global__ void gap_bwd_gelu(const float* df, const float* Z, int B, float* dZ) {
  const size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<size_t>(B) * kP2 * kC3) return;
  const int c = static_cast<int>(t % kC3);
  const int n = static_cast<int>(t / (static_cast<size_t>(kP2) * kC3));
  dZ[t] = df[n * kC3 + c] * (1.0f / kP2) * gelu_grad(Z[t]);
}

__global__ void gather_genomes(const uint8_t* src, const int* idx, int n, uint8_t* dst) {
  const int r = blockIdx.x;
  if (r >= n) return;
  for (int j = threadIdx.x; j < kP1; j += blockDim.x)
    dst[static_cast<size_t>(r) * kP1 + j] = src[static_cast<size_t>(idx[r]) * kP1 + j];
}

__global__ void gather_scalar(const float* src, const int* idx, int n, float* dst) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[r] = src[idx[r]];
}

unsigned blocks_for(size_t n) { return static_cast<unsigned>((n + 255) / 256); }

struct LayerShape {
  int out, in;  // weight [out x in], bias [out]
};
constexpr LayerShape kLayers[] = {{kC1, kK1}, {kC2, kK2}, {kC3, kK3}, {kH, kC3}, {kH, kH}, {1, kH}};
constexpr int kNL = 6;

}  // namespace

struct mh_cnn {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t w_off[kNL] = {}, b_off[kNL] = {};
  size_t n_params = 0;
  Dev<float> params, grads, m, v, part;
  Dev<__hip_bfloat16> w16;
  int cap = 0;
  // activations (B rows of genomes)
  Dev<uint8_t> g;
  Dev<float> z1, a1, z2, a2, p2, z3, a3, feat, zh1, ah1, zh2, ah2, y, t;
  // gradients
  Dev<float> epoch_terms;  // loss terms of every batch of a training run's last epoch
  Dev<float> dy, terms, dzh2, dzh1, dfeat, dz3, dp2, dz2, dz1;
  Dev<uint8_t> data_g;
  Dev<float> data_t;
  Dev<int> idx;
  int adam_t = 0;
  ~mh_cnn() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  const __hip_bfloat16* W(int l) const { return w16.p + w_off[l]; }
  const float* b(int l) const { return params.p + b_off[l]; }
};

namespace {

void ensure(mh_cnn* net, int B) {
  if (B <= net->cap) return;
  const size_t R1 = static_cast<size_t>(B) * kP1, R2 = static_cast<size_t>(B) * kP2;
  net->g.ensure(static_cast<size_t>(B) * kP1);
  net->z1.ensure(R1 * kC1);
  net->a1.ensure(R1 * kC1);
  net->z2.ensure(R1 * kC2);
  net->a2.ensure(R1 * kC2);
  net->p2.ensure(R2 * kC2);
  net->z3.ensure(R2 * kC3);
  net->a3.ensure(R2 * kC3);
  net->feat.ensure(static_cast<size_t>(B) * kC3);
  net->zh1.ensure(static_cast<size_t>(B) * kH);
  net->ah1.ensure(static_cast<size_t>(B) * kH);
  net->zh2.ensure(static_cast<size_t>(B) * kH);
  net->ah2.ensure(static_cast<size_t>(B) * kH);
  net->y.ensure(B);
  net->t.ensure(B);
  net->dy.ensure(B);
  net->terms.ensure(B);
  net->dzh2.ensure(static_cast<size_t>(B) * kH);
  net->dzh1.ensure(static_cast<size_t>(B) * kH);
  net->dfeat.ensure(static_cast<size_t>(B) * kC3);
  net->dz3.ensure(R2 * kC3);
  net->dp2.ensure(R2 * kC2);
  net->dz2.ensure(R1 * kC2);
  net->dz1.ensure(R1 * kC1);
  net->cap = B;
}

// out[M x N] = act(in[M x K] W^T + b): EPI 1 stores Z and gelu, EPI 2 plain
template <int EPI>
void dense(mh_cnn* net, int l, const float* in, int M, float* out, float* Z) {
  const int N = kLayers[l].out, K = kLayers[l].in;
  hipLaunchKernelGGL((gemm<float, __hip_bfloat16, false, true, EPI>), grid_for(M, N), dim3(256), 0, net->stream,
                     in, K, net->W(l), K, out, N, M, N, K, net->b(l), Z, static_cast<const float*>(nullptr));
}

// convolution l as a GEMM over its gathered operand la: EPI 1 stores Z and gelu
template <typename LA>
void conv(mh_cnn* net, int l, LA la, int M, float* out, float* Z) {
  const int N = kLayers[l].out, K = kLayers[l].in;
  hipLaunchKernelGGL((gemm_l<LA, StridedB<__hip_bfloat16, true>, 1>), grid_for(M, N), dim3(256), 0, net->stream, la,
                     StridedB<__hip_bfloat16, true>{net->W(l), K}, out, N, M, N, K, net->b(l), Z,
                     static_cast<const float*>(nullptr));
}

void forward(mh_cnn* net, int B) {
  const size_t R1 = static_cast<size_t>(B) * kP1, R2 = static_cast<size_t>(B) * kP2;
  hipStream_t s = net->stream;
  conv(net, 0, OneHotA{net->g.p}, static_cast<int>(R1), net->a1.p, net->z1.p);
  conv(net, 1, ConvColsA<kG, kC1>{net->a1.p}, static_cast<int>(R1), net->a2.p, net->z2.p);
  hipLaunchKernelGGL(pool2_fwd, dim3(blocks_for(R2 * kC2)), dim3(256), 0, s, net->a2.p, B, net->p2.p);
  conv(net, 2, ConvColsA<kG2, kC2>{net->p2.p}, static_cast<int>(R2), net->a3.p, net->z3.p);
  hipLaunchKernelGGL(gap_fwd, dim3(blocks_for(static_cast<size_t>(B) * kC3)), dim3(256), 0, s, net->a3.p, B,
                     net->feat.p);
  dense<1>(net, 3, net->feat.p, B, net->ah1.p, net->zh1.p);
  dense<1>(net, 4, net->ah1.p, B, net->ah2.p, net->zh2.p);
  dense<2>(net, 5, net->ah2.p, B, net->y.p, nullptr);
  hipcheck(hipGetLastError(), "cnn forward");
}

// gradient of layer l's weights and bias from dZ [M x out] and its input
// (loader lx of the [M x in] matrix), reduced over rows in kSplit
// fixed-order partials: summed into net->grads, or with to_adam left in
// net->part at the parameters' offsets, n_params apart, for
// adam_step_splits to sum
template <typename LX>
void layer_grads(mh_cnn* net, int l, const float* dZ, LX lx, int M, bool to_adam) {
  const int No = kLayers[l].out, Ni = kLayers[l].in;
  const int rows = (M + kSplit - 1) / kSplit;
  hipStream_t s = net->stream;
  float* pW = to_adam ? net->part.p + net->w_off[l] : net->part.p;
  float* pb = to_adam ? net->part.p + net->b_off[l] : net->part.p;
  const size_t sW = to_adam ? net->n_params : static_cast<size_t>(No) * Ni;
  const size_t sb = to_adam ? net->n_params : static_cast<size_t>(No);
  hipLaunchKernelGGL(wgrad_l<LX>, dim3((Ni + kTile - 1) / kTile, (No + kTile - 1) / kTile, kSplit), dim3(256), 0, s,
                     dZ, No, lx, Ni, M, rows, pW, sW);
  if (!to_adam)
    hipLaunchKernelGGL(sum_splits, dim3(blocks_for(static_cast<size_t>(No) * Ni)), dim3(256), 0, s, net->part.p,
                       No * Ni, kSplit, net->grads.p + net->w_off[l]);
  hipLaunchKernelGGL(col_partial, dim3((No + kColTile - 1) / kColTile, kSplit), dim3(kColThreads), 0, s, dZ, M, No,
                     rows, pb, sb);
  if (!to_adam)
    hipLaunchKernelGGL(sum_splits, dim3(blocks_for(No)), dim3(256), 0, s, net->part.p, No, kSplit,
                       net->grads.p + net->b_off[l]);
}

// input gradient of convolution l (input grid D, C channels, O outputs)
// from dZ [M x O]: EPI 3 multiplies by gelu'(Zprev), EPI 0 plain
template <int D, int C, int O, int EPI>
void conv_back(mh_cnn* net, int l, const float* dZ, int M, float* dIn, const float* Zprev) {
  hipLaunchKernelGGL((gemm_l<ShiftedGrad<D, O>, ConvWeightT<C, O>, EPI>), grid_for(M, C), dim3(256), 0, net->stream,
                     ShiftedGrad<D, O>{dZ}, ConvWeightT<C, O>{net->W(l)}, dIn, C, M, C, 27 * O,
                     static_cast<const float*>(nullptr), static_cast<float*>(nullptr), Zprev);
}

// dIn [M x in] = dZ [M x out] W (EPI 3: times gelu'(Zprev), EPI 0 plain)
template <int EPI>
void dense_back(mh_cnn* net, int l, const float* dZ, int M, float* dIn, const float* Zprev) {
  const int No = kLayers[l].out, Ni = kLayers[l].in;
  hipLaunchKernelGGL((gemm<float, __hip_bfloat16, false, false, EPI>), grid_for(M, Ni), dim3(256), 0, net->stream,
                     dZ, No, net->W(l), Ni, dIn, Ni, M, Ni, No, static_cast<const float*>(nullptr),
                     static_cast<float*>(nullptr), Zprev);
}

void backward(mh_cnn* net, int B, bool to_adam = false) {
  const size_t R1 = static_cast<size_t>(B) * kP1, R2 = static_cast<size_t>(B) * kP2;
  hipStream_t s = net->stream;
  using Rows = StridedB<float, false>;
  layer_grads(net, 5, net->dy.p, Rows{net->ah2.p, kH}, B, to_adam);
  dense_back<3>(net, 5, net->dy.p, B, net->dzh2.p, net->zh2.p);
  layer_grads(net, 4, net->dzh2.p, Rows{net->ah1.p, kH}, B, to_adam);
  dense_back<3>(net, 4, net->dzh2.p, B, net->dzh1.p, net->zh1.p);
  layer_grads(net, 3, net->dzh1.p, Rows{net->feat.p, kC3}, B, to_adam);
  dense_back<0>(net, 3, net->dzh1.p, B, net->dfeat.p, nullptr);
  hipLaunchKernelGGL(gap_bwd_gelu, dim3(blocks_for(R2 * kC3)), dim3(256), 0, s, net->dfeat.p, net->z3.p, B,
                     net->dz3.p);
  layer_grads(net, 2, net->dz3.p, ConvColsB<kG2, kC2>{net->p2.p}, static_cast<int>(R2), to_adam);
  conv_back<kG2, kC2, kC3, 0>(net, 2, net->dz3.p, static_cast<int>(R2), net->dp2.p, nullptr);
  hipLaunchKernelGGL(pool2_bwd_gelu, dim3(blocks_for(R1 * kC2)), dim3(256), 0, s, net->dp2.p, net->z2.p, B,
                     net->dz2.p);
  layer_grads(net, 1, net->dz2.p, ConvColsB<kG, kC1>{net->a1.p}, static_cast<int>(R1), to_adam);
  conv_back<kG, kC1, kC2, 3>(net, 1, net->dz2.p, static_cast<int>(R1), net->dz1.p, net->z1.p);
  layer_grads(net, 0, net->dz1.p, OneHotB{net->g.p}, static_cast<int>(R1), to_adam);
  hipcheck(hipGetLastError(), "cnn backward");
}

// Loss terms (into `terms`, B floats on the device) and dy = (y - t) / B for
// the batch in net->g / net->t, then the backward pass (to_adam as in
// layer_grads).
void loss_backward_to(mh_cnn* net, int B, float* terms, bool to_adam = false) {
  forward(net, B);
  hipLaunchKernelGGL(loss_grad, dim3(blocks_for(B)), dim3(256), 0, net->stream, net->y.p, net->t.p, B, 1.0f / B,
                     net->dy.p, terms);
  backward(net, B, to_adam);
}

// A batch's loss from its B terms on the host: their sum in order / B.
float batch_loss(const float* terms, int B) {
  double sum = 0.0;
  for (int i = 0; i < B; i++) sum += terms[i];
  return static_cast<float>(sum / B);
}

float loss_backward(mh_cnn* net, int B, bool want_loss) {
  loss_backward_to(net, B, net->terms.p);
  if (!want_loss) return 0.f;
  std::vector<float> terms(B);
  hipcheck(hipMemcpyAsync(terms.data(), net->terms.p, B * sizeof(float), hipMemcpyDeviceToHost, net->stream), "D2H");
  hipcheck(hipStreamSynchronize(net->stream), "sync");
  return batch_loss(terms.data(), B);
}

int fail(const mh::Error& e) {
  mh::set_error(e.what());
  return e.code;
}

}  // namespace

extern "C" {

int mh_cnn_create(int device, uint64_t seed, mh_cnn** out) {
  try {
    if (!out) throw mh::Error(MH_ERR_ARG, "null argument");
    auto net = std::make_unique<mh_cnn>();
    net->device = device;
    hipcheck(hipSetDevice(device), "hipSetDevice");
    hipcheck(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking), "stream");
    size_t off = 0, max_w = 0;
    for (int l = 0; l < kNL; l++) {
      net->w_off[l] = off;
      off += static_cast<size_t>(kLayers[l].out) * kLayers[l].in;
      net->b_off[l] = off;
      off += kLayers[l].out;
      max_w = std::max(max_w, static_cast<size_t>(kLayers[l].out) * kLayers[l].in);
    }
    net->n_params = off;
    // He-uniform weights (bound sqrt(6 / fan_in)), zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> p(off, 0.f);
    for (int l = 0; l < kNL; l++) {
      std::uniform_real_distribution<float> u(-std::sqrt(6.0f / kLayers[l].in), std::sqrt(6.0f / kLayers[l].in));
      for (size_t i = 0; i < static_cast<size_t>(kLayers[l].out) * kLayers[l].in; i++) p[net->w_off[l] + i] = u(rng);
    }
    net->params.ensure(off);
    net->grads.ensure(off);
    net->m.ensure(off);
    net->v.ensure(off);
    net->w16.ensure(off);
    // split partials of one tensor, or of every parameter for adam_step_splits
    net->part.ensure(static_cast<size_t>(kSplit) * std::max(max_w, off));
    std::vector<float> zeros(off, 0.f);
    hipcheck(hipMemcpy(net->params.p, p.data(), off * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipcheck(hipMemcpy(net->m.p, zeros.data(), off * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipcheck(hipMemcpy(net->v.p, zeros.data(), off * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipLaunchKernelGGL(to_bf16, dim3(blocks_for(off)), dim3(256), 0, net->stream, net->params.p, net->w16.p,
                       static_cast<int>(off));
    hipcheck(hipStreamSynchronize(net->stream), "init");
    *out = net.release();
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_cnn_destroy(mh_cnn* net) {
  delete net;
  return MH_OK;
}

int mh_cnn_param_count(const mh_cnn* net) { return net ? static_cast<int>(net->n_params) : MH_ERR_ARG; }

int mh_cnn_get_params(const mh_cnn* net, float* out) {
  try {
    if (!net || !out) throw mh::Error(MH_ERR_ARG, "null argument");
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    hipcheck(hipMemcpy(out, net->params.p, net->n_params * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_cnn_set_params(mh_cnn* net, const float* in) {
  try {
    if (!net || !in) throw mh::Error(MH_ERR_ARG, "null argument");
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    hipcheck(hipMemcpy(net->params.p, in, net->n_params * sizeof(float), hipMemcpyHostToDevice), "H2D");
    std::vector<float> zeros(net->n_params, 0.f);
    hipcheck(hipMemcpy(net->m.p, zeros.data(), net->n_params * sizeof(float), hipMemcpyHostToDevice), "H2D");
    hipcheck(hipMemcpy(net->v.p, zeros.data(), net->n_params * sizeof(float), hipMemcpyHostToDevice), "H2D");
    net->adam_t = 0;
    hipLaunchKernelGGL(to_bf16, dim3(blocks_for(net->n_params)), dim3(256), 0, net->stream, net->params.p,
                       net->w16.p, static_cast<int>(net->n_params));
    hipcheck(hipStreamSynchronize(net->stream), "set params");
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_cnn_forward(mh_cnn* net, const uint8_t* genomes, int n, float* y) {
  try {
    if (!net || n < 0 || (n > 0 && (!genomes || !y))) throw mh::Error(MH_ERR_ARG, "null argument");
    if (n == 0) return MH_OK;
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    const int chunk = 256;
    ensure(net, std::min(n, chunk));
    for (int b0 = 0; b0 < n; b0 += chunk) {
      const int B = std::min(chunk, n - b0);
      hipcheck(hipMemcpyAsync(net->g.p, genomes + static_cast<size_t>(b0) * kP1, static_cast<size_t>(B) * kP1,
                              hipMemcpyHostToDevice, net->stream), "H2D");
      forward(net, B);
      hipcheck(hipMemcpyAsync(y + b0, net->y.p, B * sizeof(float), hipMemcpyDeviceToHost, net->stream), "D2H");
    }
    hipcheck(hipStreamSynchronize(net->stream), "sync");
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_cnn_loss_grad(mh_cnn* net, const uint8_t* genomes, const float* target, int n, float* loss, float* grad) {
  try {
    if (!net || !genomes || !target || n < 1) throw mh::Error(MH_ERR_ARG, "null argument");
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    ensure(net, n);
    hipcheck(hipMemcpyAsync(net->g.p, genomes, static_cast<size_t>(n) * kP1, hipMemcpyHostToDevice, net->stream),
             "H2D");
    hipcheck(hipMemcpyAsync(net->t.p, target, n * sizeof(float), hipMemcpyHostToDevice, net->stream), "H2D");
    const float l = loss_backward(net, n, true);
    if (loss) *loss = l;
    if (grad) hipcheck(hipMemcpy(grad, net->grads.p, net->n_params * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

int mh_cnn_train(mh_cnn* net, const uint8_t* genomes, const float* target, int n, int batch, int n_epochs, float lr,
                 uint64_t seed, float* final_loss) {
  try {
    if (!net || !genomes || !target || n < 1 || batch < 1 || n_epochs < 0) throw mh::Error(MH_ERR_ARG, "bad arguments");
    hipcheck(hipSetDevice(net->device), "hipSetDevice");
    batch = std::min(batch, n);
    ensure(net, batch);
    net->data_g.ensure(static_cast<size_t>(n) * kP1);
    net->data_t.ensure(n);
    net->idx.ensure(n);
    hipcheck(hipMemcpyAsync(net->data_g.p, genomes, static_cast<size_t>(n) * kP1, hipMemcpyHostToDevice,
                            net->stream), "H2D");
    hipcheck(hipMemcpyAsync(net->data_t.p, target, n * sizeof(float), hipMemcpyHostToDevice, net->stream), "H2D");
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    // the last epoch keeps every batch's loss terms on the device, so no
    // batch waits for a copy to the host; they are summed after the loop
    const int n_batches = n / batch;
    net->epoch_terms.ensure(static_cast<size_t>(n_batches) * batch);
    int batches = 0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < n_epochs; ep++) {
      for (int i = 0; i < n; i++) perm[i] = i;
      std::shuffle(perm.begin(), perm.end(), rng);
      hipcheck(hipMemcpyAsync(net->idx.p, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, net->stream), "H2D");
      const bool last = ep == n_epochs - 1;
      for (int b0 = 0; b0 + batch <= n; b0 += batch) {
        hipLaunchKernelGGL(gather_genomes, dim3(batch), dim3(128), 0, net->stream, net->data_g.p, net->idx.p + b0,
                           batch, net->g.p);
        hipLaunchKernelGGL(gather_scalar, dim3(blocks_for(batch)), dim3(256), 0, net->stream, net->data_t.p,
                           net->idx.p + b0, batch, net->t.p);
        loss_backward_to(net, batch,
                         last ? net->epoch_terms.p + static_cast<size_t>(batches) * batch : net->terms.p, true);
        net->adam_t++;
        const float bc1 = 1.f - std::pow(b1, static_cast<float>(net->adam_t));
        const float bc2 = 1.f - std::pow(b2, static_cast<float>(net->adam_t));
        hipLaunchKernelGGL(adam_step_splits, dim3(blocks_for(net->n_params)), dim3(256), 0, net->stream,
                           net->params.p, net->part.p, kSplit, net->n_params, net->m.p, net->v.p, net->w16.p,
                           static_cast<int>(net->n_params), lr, b1, b2, eps, bc1, bc2);
        hipcheck(hipGetLastError(), "train step");
        if (last) batches++;
      }
    }
    // the last epoch's mean of the batch losses
    std::vector<float> terms(static_cast<size_t>(batches) * batch);
    if (batches)
      hipcheck(hipMemcpyAsync(terms.data(), net->epoch_terms.p, terms.size() * sizeof(float),
                              hipMemcpyDeviceToHost, net->stream), "D2H");
    hipcheck(hipStreamSynchronize(net->stream), "sync");
    double epoch_loss = 0.0;
    for (int bi = 0; bi < batches; bi++) epoch_loss += batch_loss(terms.data() + static_cast<size_t>(bi) * batch, batch);
    if (batches) epoch_loss /= batches;
    if (final_loss) *final_loss = static_cast<float>(epoch_loss);
    return MH_OK;
  } catch (const mh::Error& e) {
    return fail(e);
  }
}

}  // extern "C"
