// knn.hip -- K nearest neighbours of every mass of every robot, on the GPU.
//
// Replaces Triangulation::KNN (common/evolvables/knn.cu:62-144: dense M x M
// distance matrix + segmented radix sort per row) with a per-robot tiled
// scan: one lane per query mass keeps its K best (distance, index) pairs
// sorted in registers while candidate positions stream through LDS in
// tiles.  Distances are sqrtf(dx*dx + (dy*dy + dz*dz)) (Eigen's
// Vector3f::norm order, as KNN_CPU computes them, knn.cu:147-158); a zero
// distance -- the mass itself or a coincident mass -- counts as +inf
// (knn.cu:166-167); equal distances keep the lower index first, the order
// the reference's stable radix sort of index-ordered rows produces.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "mercury_hip.h"
#include "mh_internal.h"

namespace {

constexpr int kQueryBlock = 256;
constexpr int kTile = 1024;
constexpr int kMaxK = 32;

struct KnnTile {
  uint32_t mass_begin;  // robot's first mass (global)
  uint32_t n_mass;      // robot's mass count
  uint32_t query0;      // first query mass of this workgroup (robot-local)
  uint32_t out_begin;   // first output row (global mass index)
};

__global__ __launch_bounds__(kQueryBlock) void knn_tiles(const float* __restrict__ pos,
                                                         const KnnTile* __restrict__ tiles,
                                                         int K, uint16_t* __restrict__ nbr,
                                                         float* __restrict__ dist) {
  __shared__ float sx[kTile], sy[kTile], sz[kTile];
  const KnnTile t = tiles[blockIdx.x];
  const int q = static_cast<int>(t.query0) + static_cast<int>(threadIdx.x);
  const bool active = q < static_cast<int>(t.n_mass);
  float qx = 0.f, qy = 0.f, qz = 0.f;
  if (active) {
    const size_t g = 3u * (static_cast<size_t>(t.mass_begin) + q);
    qx = pos[g];
    qy = pos[g + 1];
    qz = pos[g + 2];
  }
  float bd[kMaxK];
  int bi[kMaxK];
#pragma unroll
  for (int k = 0; k < kMaxK; ++k) {
    bd[k] = INFINITY;
    bi[k] = 0x7fffffff;
  }
  float worst = INFINITY;  // bd[K - 1], kept in its own register
  for (int base = 0; base < static_cast<int>(t.n_mass); base += kTile) {
    const int n = min(kTile, static_cast<int>(t.n_mass) - base);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kQueryBlock) {
      const size_t g = 3u * (static_cast<size_t>(t.mass_begin) + base + i);
      sx[i] = pos[g];
      sy[i] = pos[g + 1];
      sz[i] = pos[g + 2];
    }
    __syncthreads();
    for (int c = 0; c < n; ++c) {
      const float dx = qx - sx[c], dy = qy - sy[c], dz = qz - sz[c];
      const float yz = dy * dy + dz * dz;
      float d = sqrtf(dx * dx + yz);
      if (d == 0.0f) d = INFINITY;
      // rows are ranked by (distance, index) and candidates arrive in index
      // order, so a candidate enters only if strictly closer than the K-th
      if (active && d < worst) {
        float cd = d;
        int ci = base + c;
        bool moving = false;
#pragma unroll
        for (int k = 0; k < kMaxK; ++k) {
          if (k < K) {
            const bool take = moving || cd < bd[k];
            const float td = bd[k];
            const int ti = bi[k];
            bd[k] = take ? cd : td;
            bi[k] = take ? ci : ti;
            cd = take ? td : cd;
            ci = take ? ti : ci;
            moving = take;
            if (k == K - 1) worst = bd[k];
          }
        }
      }
    }
  }
  if (active) {
    const size_t o = (static_cast<size_t>(t.out_begin) + q) * K;
#pragma unroll
    for (int k = 0; k < kMaxK; ++k)
      if (k < K) {
        nbr[o + k] = static_cast<uint16_t>(bi[k] == 0x7fffffff ? 0xffff : bi[k]);
        dist[o + k] = bd[k];
      }
  }
}

int fail(int code, const std::string& m) {
  mh::set_error(m);
  return code;
}

}  // namespace

extern "C" {

int mh_knn_batch(int device, int n_robots, const uint32_t* mass_offset, const float* pos, int K,
                 uint16_t* nbr, float* dist) {
  if (n_robots < 0 || (n_robots > 0 && (!mass_offset || !pos || !nbr || !dist)))
    return fail(MH_ERR_ARG, "null argument");
  if (K < 1 || K > kMaxK) return fail(MH_ERR_RANGE, "K must be 1..32");
  if (n_robots == 0) return MH_OK;
  std::vector<KnnTile> tiles;
  for (int r = 0; r < n_robots; r++) {
    const uint32_t m = mass_offset[r + 1] - mass_offset[r];
    if (mass_offset[r + 1] < mass_offset[r]) return fail(MH_ERR_ARG, "offsets must be non-decreasing");
    if (m > 65535) return fail(MH_ERR_RANGE, "robot has more than 65535 masses");
    if (m > 0 && static_cast<int>(m) <= K) return fail(MH_ERR_RANGE, "robot has no more than K masses");
    for (uint32_t q = 0; q < m; q += kQueryBlock) tiles.push_back({mass_offset[r], m, q, mass_offset[r]});
  }
  const size_t M = mass_offset[n_robots];
  if (tiles.empty()) return MH_OK;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(MH_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  float* d_pos = nullptr;
  float* d_dist = nullptr;
  uint16_t* d_nbr = nullptr;
  KnnTile* d_tiles = nullptr;
  auto cleanup = [&]() {
    hipFree(d_pos);
    hipFree(d_dist);
    hipFree(d_nbr);
    hipFree(d_tiles);
  };
  if ((e = hipMalloc(&d_pos, 3 * M * sizeof(float))) != hipSuccess ||
      (e = hipMalloc(&d_dist, M * K * sizeof(float))) != hipSuccess ||
      (e = hipMalloc(&d_nbr, M * K * sizeof(uint16_t))) != hipSuccess ||
      (e = hipMalloc(&d_tiles, tiles.size() * sizeof(KnnTile))) != hipSuccess) {
    cleanup();
    return fail(MH_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  e = hipMemcpy(d_pos, pos, 3 * M * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d_tiles, tiles.data(), tiles.size() * sizeof(KnnTile), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(knn_tiles, dim3(static_cast<unsigned>(tiles.size())), dim3(kQueryBlock), 0,
                       nullptr, d_pos, d_tiles, K, d_nbr, d_dist);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(nbr, d_nbr, M * K * sizeof(uint16_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(dist, d_dist, M * K * sizeof(float), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return fail(MH_ERR_HIP, std::string("knn: ") + hipGetErrorString(e));
  return MH_OK;
}

}  // extern "C"
