// calibrate.hip -- measurement utilities for the roofline figures.
//
//   mh_calibrate_lds:  LDS read bandwidth of the device (ds_read_b128 from
//                      a 64 KB buffer, conflict-free lane pattern, 4 waves
//                      per SIMD), the peak the LDS-resident step kernel is
//                      measured against.
//   mh_calibrate_copy: a plain float4 copy of a known byte count, the
//                      reference that scales rocprof's FETCH_SIZE /
//                      WRITE_SIZE counters to bytes.
#include <hip/hip_runtime.h>

#include "mercury_hip.h"
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

__global__ void copy_f4(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
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
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  const int blocks = (cus > 0 ? cus : 256) * 8;
  const int iters = 4096;
  float* out = nullptr;
  if ((e = hipMalloc(&out, sizeof(float) * static_cast<size_t>(blocks) * 1024)) != hipSuccess)
    return fail(e, "hipMalloc");
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipLaunchKernelGGL(lds_read_bw, dim3(blocks), dim3(1024), 0, nullptr, out, 16);
  hipEventRecord(a, nullptr);
  hipLaunchKernelGGL(lds_read_bw, dim3(blocks), dim3(1024), 0, nullptr, out, iters);
  hipEventRecord(b, nullptr);
  e = hipEventSynchronize(b);
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(out);
  if (e != hipSuccess) return fail(e, "lds calibration");
  const double bytes = static_cast<double>(blocks) * 1024.0 * iters * 4 * 16;
  *gbytes_per_s = static_cast<float>(bytes / (ms * 1e-3) / 1e9);
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
    hipFree(src);
    return fail(e, "hipMalloc");
  }
  hipMemset(src, 0, n * 16);
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  const int blocks = 256 * 16;
  hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, nullptr, src, dst, n);
  hipEventRecord(a, nullptr);
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(copy_f4, dim3(blocks), dim3(256), 0, nullptr, src, dst, n);
  hipEventRecord(b, nullptr);
  e = hipEventSynchronize(b);
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(src);
  hipFree(dst);
  if (e != hipSuccess) return fail(e, "copy calibration");
  *avg_ms = ms / reps;
  return MH_OK;
}

}  // extern "C"
