// calibrate.hip -- measurement utilities for the roofline figures.
//
//   mh_calibrate_lds:  LDS read bandwidth of the device (ds_read_b128 from
//                      a 64 KB buffer, conflict-free lane pattern, 4 waves
//                      per SIMD), the peak the LDS-resident step kernel is
//                      measured against.
//   mh_calibrate_copy: a plain float4 copy of a known byte count, the
//                      reference that scales rocprof's FETCH_SIZE /
//                      WRITE_SIZE counters to bytes.
//   mh_calibrate_valu: FP32 VALU issue rate (wave64 v_fma_f32 or v_mul_f32
//                      per second, 8 independent chains per lane, 8 waves
//                      per SIMD),
//                      the peak the VALU-bound step kernels are measured
//                      against.
#include <hip/hip_runtime.h>

#include <vector>

#include "mercury_hip.h"
#include "mh_device.h"
#include "mh_internal.h"

namespace {

constexpr int kLdsFloat4 = 4096;  // 64 KB

__global__ __launch_bounds__(1024) void lds_read_bw(float* out, int iters) {
  __shared__ float4 buf[kLdsFloat4];
  const int tid = threadIdx.x;
  for (int i = tid; i < kLdsFloat4; i += 1024)
    buf[i] = make_float4(static_cast<float>(i), 1.f, 2.f, 3.f);
  __syncthreads();
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  int base = tid;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float4 v = buf[(base + j * 1024) & (kLdsFloat4 - 1)];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    base += 64;
  }
  if (acc.x == -1.f) out[blockIdx.x * 1024 + tid] = acc.x + acc.y + acc.z + acc.w;
}

// kind 0: v_fma_f32 chains; kind 1: v_mul_f32 chains (two sources);
// kind 2: v_fma_f32, twice the chains (more independent work per wave);
// kind 3: v_rcp_f32 chains; kind 4: v_rcp_f32 then v_fma_f32 in each chain
// step (counted as two instructions)
template <int KIND>
__global__ __launch_bounds__(256) void valu_rate(float* out, int iters, float a, float b) {
  constexpr int N = KIND == 2 ? 16 : 8;
  float x[N];
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = static_cast<float>(threadIdx.x) + static_cast<float>(i) + 1.0f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (KIND == 1)
        x[i] = x[i] * a;
      else if (KIND == 3)
        x[i] = __builtin_amdgcn_rcpf(x[i]);
      else if (KIND == 4)
        x[i] = fmaf(__builtin_amdgcn_rcpf(x[i]), a, b);
      else
        x[i] = fmaf(x[i], a, b);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) s = s + x[i];
  if (s == -1.f) out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void copy_f4(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

// Bit-exact comparison of the fast arithmetic sequences with the compiler's
// correctly rounded operations over a contiguous range of float bit
// patterns; one mismatch count per block.
//   which 0: hardware sqrt          vs sqrtf
//         1: mh::sqrt_rn_inrange    vs sqrtf
//         2: hardware reciprocal    vs 1.0f / x
//         3: x                      vs the next float (all differ: checks
//                                      the counting itself)
//         4: mh::rcp_rn_inrange     vs 1.0f / x
//         5: hardware sqrt above sqrtf (count of such inputs)
//         6: hardware sqrt below sqrtf
//         7: hardware reciprocal above 1.0f / x
//         8: hardware reciprocal below 1.0f / x
//   (5-8 count one direction of the error for positive x)
//         9: mh::div_rn_from_rcp(a, b, rcp_rn_inrange(b)) vs a / b for a
//            pseudo-random pair drawn from the pattern (|a| in [2^-79, 2^63)
//            or 0, b in [2^-48, 2^40), random signs of a; pairs whose
//            quotient is not a normal float are skipped)
//        10: mh::sqrt_rn_tuckerman_inrange (hardware sqrt with Tuckerman's
//            rounding test) vs sqrtf
__global__ __launch_bounds__(256) void check_arith(int which, uint32_t begin, uint32_t count,
                                                   uint32_t* out) {
  __shared__ uint32_t part[256];
  uint32_t bad = 0;
  const uint32_t stride = gridDim.x * 256u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += stride) {
    const float x = __uint_as_float(begin + i);
    float a, b;
    if (which == 0) {
      a = __builtin_amdgcn_sqrtf(x);
      b = sqrtf(x);
    } else if (which == 1) {
      a = mh::sqrt_rn_inrange(x);
      b = sqrtf(x);
    } else if (which == 2) {
      a = __builtin_amdgcn_rcpf(x);
      b = 1.0f / x;
    } else if (which == 3) {
      a = x;
      b = __uint_as_float(__float_as_uint(x) + 1u);
    } else if (which == 4) {
      a = mh::rcp_rn_inrange(x);
      b = 1.0f / x;
    } else if (which == 9) {
      // splitmix64 of the pattern: exponents and mantissas of a and b
      uint64_t z = (static_cast<uint64_t>(begin + i) + 0x9e3779b97f4a7c15ull) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 31)) * 0x94d049bb133111ebull;
      z ^= z >> 29;
      const int ea = -79 + static_cast<int>((z & 0xff) % 142);
      const int eb = -48 + static_cast<int>(((z >> 8) & 0xff) % 88);
      const uint32_t ma = static_cast<uint32_t>(z >> 16) & 0x7fffffu;
      const uint32_t mb = static_cast<uint32_t>(z >> 39) & 0x7fffffu;
      float x = __uint_as_float((static_cast<uint32_t>(ea + 127) << 23) | ma);
      if ((z >> 62) & 1u) x = -x;
      if (((z >> 8) & 0xff) == 0xff) x = 0.0f;
      const float y = __uint_as_float((static_cast<uint32_t>(eb + 127) << 23) | mb);
      b = x / y;
      const float ab = fabsf(b);
      a = (ab == 0.0f && x == 0.0f) || (ab >= 0x1p-126f && ab < 0x1p127f)
              ? mh::div_rn_from_rcp(x, y, mh::rcp_rn_inrange(y))
              : b;
    } else if (which == 10) {
      a = mh::sqrt_rn_tuckerman_inrange(x);
      b = sqrtf(x);
    } else if (which <= 6) {
      a = __builtin_amdgcn_sqrtf(x);
      b = sqrtf(x);
    } else {
      a = __builtin_amdgcn_rcpf(x);
      b = 1.0f / x;
    }
    if (which <= 4 || which == 10)
      bad += (__float_as_uint(a) != __float_as_uint(b)) ? 1u : 0u;
    else if (which == 9)  // a zero quotient may differ in sign
      bad += (__float_as_uint(a) != __float_as_uint(b) && !(a == 0.0f && b == 0.0f)) ? 1u : 0u;
    else if (which == 5 || which == 7)
      bad += (a > b) ? 1u : 0u;
    else
      bad += (a < b) ? 1u : 0u;
  }
  part[threadIdx.x] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

int fail(hipError_t e, const char* what) {
  mh::set_error(std::string(what) + ": " + hipGetErrorString(e));
  return MH_ERR_HIP;
}

}  // namespace

extern "C" {

int mh_calibrate_lds(int device, float* gbytes_per_s) {
  if (!gbytes_per_s) return MH_ERR_ARG;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(e, "hipSetDevice");
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  const int blocks = (cus > 0 ? cus : 256) * 8;
  const int iters = 4096;
  float* out = nullptr;
  if ((e = hipMalloc(&out, sizeof(float) * static_cast<size_t>(blocks) * 1024)) != hipSuccess)
    return fail(e, "hipMalloc");
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  hipLaunchKernelGGL(lds_read_bw, dim3(blocks), dim3(1024), 0, nullptr, out, 16);
  (void)hipEventRecord(a, nullptr);
  hipLaunchKernelGGL(lds_read_bw, dim3(blocks), dim3(1024), 0, nullptr, out, iters);
  (void)hipEventRecord(b, nullptr);
  e = hipEventSynchronize(b);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  if (e != hipSuccess) return fail(e, "lds calibration");
  const double bytes = static_cast<double>(blocks) * 1024.0 * iters * 4 * 16;
  *gbytes_per_s = static_cast<float>(bytes / (ms * 1e-3) / 1e9);
  return MH_OK;
}

int mh_calibrate_valu(int device, int kind, float* giga_instr_per_s) {
  if (!giga_instr_per_s || kind < 0 || kind > 4) return MH_ERR_ARG;
  auto kernel = kind == 0   ? valu_rate<0>
                : kind == 1 ? valu_rate<1>
                : kind == 2 ? valu_rate<2>
                : kind == 3 ? valu_rate<3>
                            : valu_rate<4>;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(e, "hipSetDevice");
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  // 8 blocks of 4 waves per CU: 8 waves per SIMD
  const int blocks = (cus > 0 ? cus : 256) * 8;
  const int iters = 1 << 14;
  float* out = nullptr;
  if ((e = hipMalloc(&out, sizeof(float) * static_cast<size_t>(blocks) * 256)) != hipSuccess)
    return fail(e, "hipMalloc");
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, nullptr, out, 64, 0.999f, 1e-3f);
  (void)hipEventRecord(a, nullptr);
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, nullptr, out, iters, 0.999f, 1e-3f);
  (void)hipEventRecord(b, nullptr);
  e = hipEventSynchronize(b);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  if (e != hipSuccess) return fail(e, "valu calibration");
  const double chains = kind == 2 ? 16.0 : kind == 4 ? 2.0 * 8.0 : 8.0;
  const double instr = static_cast<double>(blocks) * 4.0 * iters * chains;  // wave instructions
  *giga_instr_per_s = static_cast<float>(instr / (ms * 1e-3) / 1e9);
  return MH_OK;
}

int mh_calibrate_copy(int device, size_t bytes, int reps, float* avg_ms) {
  if (!avg_ms || bytes < 16 || reps < 1) return MH_ERR_ARG;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(e, "hipSetDevice");
  const size_t n = bytes / 16;
  float4 *src = nullptr, *dst = nullptr;
  if ((e = hipMalloc(&src, n * 16)) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&dst, n * 16)) != hipSuccess) {
    (void)hipFree(src);
    return fail(e, "hipMalloc");
  }
  (void)hipMemset(src, 0, n * 16);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const int blocks = 256 * 16;
  hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, nullptr, src, dst, n);
  (void)hipEventRecord(a, nullptr);
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, nullptr, src, dst, n);
  (void)hipEventRecord(b, nullptr);
  e = hipEventSynchronize(b);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(src);
  (void)hipFree(dst);
  if (e != hipSuccess) return fail(e, "copy calibration");
  *avg_ms = ms / reps;
  return MH_OK;
}

int mh_check_arith(int device, int which, uint32_t begin, uint32_t count, uint64_t* mismatches) {
  if (!mismatches || which < 0 || which > 10 || count == 0 ||
      static_cast<uint64_t>(begin) + count > 0x100000000ull)
    return MH_ERR_ARG;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(e, "hipSetDevice");
  const int blocks = 256 * 32;
  uint32_t* out = nullptr;
  if ((e = hipMalloc(&out, sizeof(uint32_t) * blocks)) != hipSuccess) return fail(e, "hipMalloc");
  hipLaunchKernelGGL(check_arith, dim3(blocks), dim3(256), 0, nullptr, which, begin, count, out);
  std::vector<uint32_t> h(blocks);
  e = hipMemcpy(h.data(), out, sizeof(uint32_t) * blocks, hipMemcpyDeviceToHost);
  (void)hipFree(out);
  if (e != hipSuccess) return fail(e, "arith check");
  uint64_t total = 0;
  for (uint32_t v : h) total += v;
  *mismatches = total;
  return MH_OK;
}

}  // extern "C"
