// voxel_kernels.hip -- voxel-robot populations built on the device.
//
// One workgroup per robot turns the robot's voxel encodings into everything
// the mass-parallel step kernel (step_mass) reads, with the results of the
// host build (VoxelRobot::Build, common/evolvables/VoxelRobot.cpp:30-331,
// as mh_voxel_build restates it) followed by the host packing:
//   1. decode every lattice point's encoding (two lowest set bits, bit 0 =
//      air; material.h:275-322);
//   2. Strip (VoxelRobot.cpp:267-283): keep the largest face-connected
//      component of non-air voxels, the one with the lowest first voxel
//      among equal sizes.  Components are labelled by their lowest voxel
//      index with a min-label propagation (each round takes the minimum
//      over a voxel's face neighbours and jumps to its label's label);
//      labels only decrease and always name a voxel of the same component,
//      so the result does not depend on the order of the updates;
//   3. masses at the shape's positions, material = the voxel's;
//   4. spring materials: decode(OR of the non-air bits of the up to 4
//      voxels the shape template lists per spring);
//   5. the robot's (rest length, material) pair table: a presence bit per
//      (length index, material id), ranked by a prefix popcount;
//   6. lane order: masses by non-air springs, descending, ties by mass
//      index;
//   7. per lane one record per non-air spring in ascending spring index,
//      the mask of the records whose first endpoint the mass is, and each
//      record's spring index, as the host packing writes them.
// Integer atomics only (LDS counts, bit sets, the error flag), so every
// output is deterministic.
#include <hip/hip_runtime.h>

#include <climits>

#include "mercury_hip.h"
#include "mh_kernels.h"

namespace mh {
namespace {

constexpr int kVoxThreads = 1024;

__device__ __forceinline__ int encoded_index(uint32_t enc) {
  if (enc == 0u || (enc & 1u)) return 0;
  const int lo = __builtin_ctz(enc);
  const uint32_t rest = enc & (enc - 1u);
  if (rest == 0u) return 1 + lo * (lo - 1) / 2;
  const int hi = __builtin_ctz(rest);
  return 1 + hi * (hi - 1) / 2 + lo;
}

__device__ __forceinline__ void flag_error(int* error, int code) { atomicCAS(error, 0, code); }

__global__ __launch_bounds__(kVoxThreads) void voxel_population(VoxelBuildParams p) {
  __shared__ uint32_t nab[kVoxMaxPoints];     // non-air bits of each voxel after Strip
  __shared__ int label[kVoxMaxPoints];        // component label (lowest voxel index)
  __shared__ int cnt[kVoxMaxPoints];          // component sizes, by label
  __shared__ uint16_t key[kVoxMaxPoints];     // lane-order key per mass
  __shared__ uint16_t lane_of[kVoxMaxPoints];
  __shared__ uint8_t sid[kVoxMaxSprings];     // spring material ids
  __shared__ uint32_t bits[kVoxPairWords];    // (length index, material id) presence
  __shared__ uint32_t below[kVoxPairWords];   // set bits in the words before
  __shared__ int best_key, n_pairs, n_active;

  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const int n = p.n, S = p.n_springs;
  const int words = p.n_lengths * 8;  // 256 material ids per length: 8 words

  // 1. decode
  const uint32_t* enc = p.enc + static_cast<size_t>(r) * n;
  int code = 0;
  if (tid < n) {
    const uint32_t e = enc[tid];
    if (e & ~p.valid_bits) flag_error(p.error, MH_ERR_ARG);
    code = (e & ~p.valid_bits) ? 0 : encoded_index(e);
    label[tid] = p.code_solid[code] ? tid : INT_MAX;
    cnt[tid] = 0;
  }
  for (int w = tid; w < words; w += kVoxThreads) bits[w] = 0u;
  if (tid == 0) {
    best_key = 0;
    n_active = 0;
  }
  __syncthreads();

  // 2. Strip: min-label propagation over face neighbours
  int x = 0, y = 0, z = 0;
  if (tid < n) {
    x = tid % p.nx;
    y = (tid / p.nx) % p.ny;
    z = tid / (p.nx * p.ny);
  }
  for (;;) {
    int changed = 0;
    if (tid < n && label[tid] != INT_MAX) {
      int m = label[tid];
      const int nb[6] = {x + 1 < p.nx ? tid + 1 : -1, x > 0 ? tid - 1 : -1,
                         y + 1 < p.ny ? tid + p.nx : -1, y > 0 ? tid - p.nx : -1,
                         z + 1 < p.nz ? tid + p.nx * p.ny : -1, z > 0 ? tid - p.nx * p.ny : -1};
#pragma unroll
      for (int k = 0; k < 6; k++)
        if (nb[k] >= 0) m = min(m, label[nb[k]]);
      m = min(m, label[m]);
      if (m < label[tid]) {
        label[tid] = m;
        changed = 1;
      }
    }
    if (!__syncthreads_or(changed)) break;
  }
  if (tid < n && label[tid] != INT_MAX) atomicAdd(&cnt[label[tid]], 1);
  __syncthreads();
  // the largest component, the lowest first voxel among equal sizes
  if (tid < n && label[tid] == tid) atomicMax(&best_key, (cnt[tid] << 11) | (2047 - tid));
  __syncthreads();
  const int best = best_key > 0 ? 2047 - (best_key & 2047) : -1;

  // 3. masses
  if (tid < n) {
    if (label[tid] != best || best < 0) code = 0;
    const uint32_t ce = p.code_enc[code];
    nab[tid] = ce != 1u ? ce : 0u;
    const size_t gi = static_cast<size_t>(r) * n + tid;
    p.pos[gi] = make_float4(p.tpos[3 * tid], p.tpos[3 * tid + 1], p.tpos[3 * tid + 2], p.mass);
    p.vel[gi] = make_float4(0.f, 0.f, 0.f, 0.f);
    p.mass_id[gi] = p.code_id[code];
    p.stripped[gi] = ce;
  }
  __syncthreads();

  // 4. spring materials, 5. their (length, material) presence bits
  for (int s = tid; s < S; s += kVoxThreads) {
    uint32_t bitsor = 0u;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint16_t v = p.voxels[4 * s + k];
      if (v != 0xffffu) bitsor |= nab[v];
    }
    const uint8_t id = p.code_id[encoded_index(bitsor)];
    sid[s] = id;
    p.spring_id[static_cast<size_t>(r) * S + s] = id;
    if (id != 0) {
      const int kk = p.length_index[s] * 256 + id;
      atomicOr(&bits[kk >> 5], 1u << (kk & 31));
    }
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int w = 0; w < words; w++) {
      below[w] = static_cast<uint32_t>(acc);
      acc += __popc(bits[w]);
    }
    n_pairs = acc;
    if (acc > kMassPairNull) flag_error(p.error, MH_ERR_RANGE);
  }
  __syncthreads();
  const int np_r = n_pairs;
  float* tab = p.lbtab + static_cast<size_t>(r) * 2 * kMassPairMax;
  for (int w = tid; w < words; w += kVoxThreads) {
    uint32_t b = bits[w];
    uint32_t rank = below[w];
    while (b) {
      const int bit = __builtin_ctz(b);
      b &= b - 1u;
      const int kk = w * 32 + bit;
      if (rank < static_cast<uint32_t>(kMassPairNull)) {
        tab[2 * rank] = p.lengths[kk >> 8];
        tab[2 * rank + 1] = __uint_as_float(static_cast<uint32_t>(kk & 255));
      }
      rank++;
    }
  }
  auto pair_rank = [&](int s) -> uint32_t {
    const int kk = p.length_index[s] * 256 + sid[s];
    const uint32_t w = bits[kk >> 5];
    return below[kk >> 5] + __popc(w & ((1u << (kk & 31)) - 1u));
  };

  // 6. lane order: most springs first (ties by mass index)
  int nq = 0, npf = 0;
  if (tid < n) {
    for (int j = p.first_off[tid]; j < p.first_off[tid + 1]; j++) npf += sid[p.first_list[j]] != 0;
    for (int j = p.second_off[tid]; j < p.second_off[tid + 1]; j++) nq += sid[p.second_list[j]] != 0;
    if (npf > p.maxp || nq > p.maxp) flag_error(p.error, MH_ERR_RANGE);
    key[tid] = static_cast<uint16_t>(npf + nq);
    atomicAdd(&n_active, npf);
  }
  __syncthreads();
  if (tid < n) {
    const uint16_t mine = key[tid];
    int rank = 0;
    for (int j = 0; j < n; j++) {
      const uint16_t kj = key[j];
      rank += (kj > mine) | ((kj == mine) & (j < tid));
    }
    lane_of[tid] = static_cast<uint16_t>(rank);
    p.mperm[static_cast<size_t>(r) * n + rank] = static_cast<uint16_t>(tid);
  }
  __syncthreads();

  // 7. records: one per non-air spring of the mass in ascending spring
  // index (merging its first-endpoint and second-endpoint lists), the mask
  // of the first-endpoint ones and each record's spring index
  const int P = p.maxp;
  uint32_t* rec = p.mrec + static_cast<size_t>(r) * (4 * P + 1) * n;
  if (tid < n && np_r <= kMassPairNull) {
    const int lane = lane_of[tid];
    int jf = p.first_off[tid], jf_end = p.first_off[tid + 1];
    int js = p.second_off[tid], js_end = p.second_off[tid + 1];
    uint32_t mask = 0u;
    int k = 0;
    while ((jf < jf_end || js < js_end) && k < 2 * P) {
      // the lower spring index next; a spring from the mass to itself
      // (none in a lattice) would come first as its first endpoint
      const bool take_first = js >= js_end || (jf < jf_end && p.first_list[jf] <= p.second_list[js]);
      const int s = take_first ? p.first_list[jf++] : p.second_list[js++];
      if (!sid[s]) continue;
      const uint32_t other = take_first ? p.pairs[2 * s + 1] : p.pairs[2 * s];
      rec[static_cast<size_t>(k) * n + lane] =
          16u * static_cast<uint32_t>(mass_slot(static_cast<int>(other))) | (8u * pair_rank(s)) << 16;
      rec[static_cast<size_t>(2 * P + 1 + k) * n + lane] = static_cast<uint32_t>(s);
      if (take_first) mask |= 1u << k;
      k++;
    }
    rec[static_cast<size_t>(2 * P) * n + lane] = mask;
    uint8_t* c = p.mcnt + static_cast<size_t>(r) * 2 * n;
    c[2 * lane] = static_cast<uint8_t>(k);
    c[2 * lane + 1] = 0;
  }
  if (tid == 0) {
    PackDesc d;
    d.mass_begin = static_cast<uint32_t>(r) * n;
    d.n_mass = static_cast<uint32_t>(n);
    d.spring_begin = static_cast<uint32_t>(r) * S;
    d.n_spring = static_cast<uint32_t>(n_active);
    d.csr_begin = static_cast<uint32_t>(r) * (4 * P + 1) * n;
    d.row_begin = static_cast<uint32_t>(r) * 2 * n;
    d.pad0 = static_cast<uint32_t>(r) * 2 * kMassPairMax;
    d.pad1 = static_cast<uint32_t>(np_r);
    p.packs[r] = d;
    p.active[r] = static_cast<uint32_t>(n_active);
  }
}

}  // namespace

hipError_t launch_voxel_population(const VoxelBuildParams& p, int n_robots, hipStream_t stream) {
  if (n_robots <= 0) return hipSuccess;
  hipLaunchKernelGGL(voxel_population, dim3(n_robots), dim3(kVoxThreads), 0, stream, p);
  return hipGetLastError();
}

}  // namespace mh
