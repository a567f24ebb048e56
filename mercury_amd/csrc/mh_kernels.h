// mh_kernels.h -- launch interface between the host simulator (sim.cpp) and
// the HIP step kernels (step_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace mh {

// One workgroup's share of the population: a contiguous range of robots.
// All offsets are global element offsets into the device arrays.
struct PackDesc {
  uint32_t mass_begin;    // first mass (float4 pos/vel arrays)
  uint32_t n_mass;
  uint32_t spring_begin;  // first active spring
  uint32_t n_spring;
  uint32_t csr_begin;     // first gather entry (mass-parallel kernels: first record)
  uint32_t row_begin;     // first row pointer (mass-parallel kernels: first record count)
  // kernel-specific: chunked kernel, offset of the pack's chunk table in
  // ctab and its chunk count; mass-parallel kernels, offset of the pack's
  // (rest length, material) pairs in lbtab and the pair count
  uint32_t pad0, pad1;
};

struct StepParams {
  const PackDesc* packs;
  float4* pos;               // xyz + mass
  float4* vel;               // xyz + 0
  const uint32_t* spair;     // pack-local a | b << 16
  const float* slbar;
  const uint8_t* smat;       // compact material index
  float* sstress;            // per active spring (may be null)
  const uint32_t* csr;       // (pack-local spring << 1) | (mass is endpoint b);
                             // LDS plans pad rows to 4 with the zero slot
  const uint32_t* row;       // pack-local gather offsets, n_mass + 1 per pack
  const uint16_t* mperm;     // per pack: phase-2 visiting order of its masses
  const uint32_t* mrec;      // mass-parallel kernel: per-mass spring records
  const uint8_t* mcnt;       // mass-parallel kernel: record counts per mass (2 each)
  const float* lbtab;        // mass-parallel kernels: per-pack (rest length, material) pairs
  const uint16_t* ccsr;      // chunked kernel: per chunk, entries 3 * (spring - chunk start)
  const uint16_t* crow;      // chunked kernel: per chunk, 2 n_mass + 1 row bounds (padded to quads)
  const uint32_t* ctab;      // chunked kernel: per pack (at d.pad0), entry offset of each chunk + end
  float* scratch;            // streaming kernel: float4 per active spring
  const float* mat_k;        // spring constant
  const float* mat_kx;       // xpbd: 2 + ((1/k)/dt)/dt
  const float* mat_dL0;
  const float* mat_omega;
  const float* mat_phi;
  int n_mat;
  int mass_cap;              // LDS capacity in masses (layout stride)
  int spring_cap;            // LDS capacity in springs (layout stride); the
                             // chunked kernel's springs per chunk
  int csr_cap;               // LDS capacity in gather entries (multiple of 4)
  int n_steps;
  float dt;
  const float* clock;        // start time of each launch of the call (clock kernel)
  int launch;                // this launch's index into clock
  int track_stress;
  int fast_div;              // xpbd, gather kernel: every material stiffness, rest length
                             // and amplitude lies in the fast division's range
  float gravity, ground_k, mu_s, mu_k, drag, damping;
};

// A compiled step-kernel variant.
struct KernelVariant {
  int solver;     // MH_SOLVER_*
  int kind;       // MH_KERNEL_LDS / MH_KERNEL_STREAM / MH_KERNEL_MASS
  int chunked;    // STREAM kind: 1 = spring terms in LDS chunks, 0 = in global memory,
                  // 2 = mass-parallel with streamed records (step_mstream);
                  // MASS kind: 1 = terms computed once and gathered (step_massg)
  int block;      // threads per workgroup
  int nspt;       // springs per thread held in registers (LDS kind); records
                  // per endpoint side (MASS kind); record bytes, 8 or 6
                  // (mass-parallel streaming kind)
  int mpt;        // masses per thread
  const void* fn; // kernel symbol
  const void* fn_stress = nullptr;  // variant used when stress is tracked (null: fn)
};

// All compiled variants (table lives in step_kernels.hip).
const KernelVariant* kernel_variants(int* count);

// LDS bytes a pack of (masses, springs, materials) needs in the LDS kernel;
// the streaming kernel needs only masses and the material table.
// LDS position slot of pack-local mass i in the LDS kernel, and the slot
// count for n masses: one unused slot after every 8 skews consecutive
// masses across the LDS banks.
__host__ __device__ inline int pos_slot(int i) { return i + (i >> 3); }
__host__ __device__ inline int pos_slots(int n) { return n + (n >> 3) + 1; }
// The mass-parallel kernels' slots: one unused slot after every
// 2^MH_MASS_SKEW_SHIFT masses, and a last (ghost) slot.
#ifndef MH_MASS_SKEW_SHIFT
#define MH_MASS_SKEW_SHIFT 3
#endif
__host__ __device__ constexpr int mass_slot(int i) { return i + (i >> MH_MASS_SKEW_SHIFT); }
__host__ __device__ constexpr int mass_slots(int block) { return block + (block >> MH_MASS_SKEW_SHIFT) + 1; }
// Mass-parallel kernels: distinct (rest length, material) pairs per pack;
// the LDS pair tables hold kMassPairMax, the last one reserved for padding
// records.  Per pair: {L, k} (float2; L the actuated rest length), {1/k,
// rest length} (float2) and the constants {dL0, omega, phi, rest length}
// (float4), each in its own table so that a record's byte offset
// pair * 8 indexes the first two.
constexpr int kMassPairMax = 512;
constexpr int kMassPairNull = kMassPairMax - 1;
constexpr int kMassDynBytes = 8 * kMassPairMax;
constexpr int kMassCstBytes = 16 * kMassPairMax;
// Their LDS layout (byte offsets): a zero term and two flags (step_massg)
// below kMassPos | nbuf position buffers float4[mass_slots(block)] | {L, k}
// (step_mass: two buffers; step_massg: one) | {1/k, rest length} | the
// constants | step_massg's term slots.
constexpr int kMassZero = 0, kMassPos = 64;
__host__ __device__ constexpr int mass_pair_base(int block, int nbuf) {
  return kMassPos + 16 * nbuf * mass_slots(block);
}
// step_massg: byte offset of the term slots and of the term of (lane, k)
__host__ __device__ constexpr int massg_term_base(int block) {
  return mass_pair_base(block, 1) + 2 * kMassDynBytes + kMassCstBytes;
}
__host__ __device__ constexpr uint32_t massg_term_offset(int block, int maxp, uint32_t lane, uint32_t k) {
  return static_cast<uint32_t>(massg_term_base(block)) + 12u * (static_cast<uint32_t>(maxp) * lane + k);
}
// Words of a stored spring term by solver: xpbd (cx, cy, cz), force (Dx,
// Dy, Dz, f).
__host__ __device__ constexpr int solver_term_words(int solver) { return solver == MH_SOLVER_FORCE ? 4 : 3; }
// Bytes of the LDS spring-term array (tw floats per slot, spring_cap + 2
// slots), rounded up to 16 B so the float4 positions after it stay aligned.
__host__ __device__ inline size_t lds_term_bytes(int spring_cap, int tw) {
  return (4u * static_cast<size_t>(tw) * static_cast<size_t>(spring_cap + 2) + 15u) & ~static_cast<size_t>(15);
}
size_t lds_bytes_lds(int mass_cap, int spring_cap, int csr_cap, int n_mat, int tw);
size_t lds_bytes_stream(int mass_cap, int n_mat);
size_t lds_bytes_mass(int block, int n_mat);
// Chunked streaming kernel: per lane and chunk at most kChunkSpl springs,
// kChunkEpl quads of gather entries and kChunkRpl quads of row bounds are in
// flight; a robot has at most kChunkMax chunks.  Row bounds per chunk:
// n_mass + 1 entries padded to whole quads.
constexpr int kChunkSpl = 8, kChunkEpl = 8, kChunkRpl = 2, kChunkMax = 256;
__host__ __device__ constexpr int chunk_row_quads(int n_mass) { return (n_mass + 1 + 3) / 4; }
size_t lds_bytes_chunk(int mass_cap, int spring_cap, int csr_cap, int n_mat, int tw);
size_t lds_bytes_massg(int block, int maxp);
// Mass-parallel streaming kernel (step_mstream): its per-material table
// holds kMStreamMats entries, the compact materials and a null one.  The
// 6-byte record format holds the material in 4 bits (kMStream6Mats
// entries) and 16 * slot in the other 12 of a u16.
constexpr int kMStreamMats = 256;
constexpr int kMStream6Mats = 16;
size_t lds_bytes_mstream(int block, int mpt);

hipError_t launch_step(const KernelVariant& v, int n_packs, size_t lds_bytes,
                       hipStream_t stream, const StepParams& p);

// Start time of each of the ceil(n_steps / K) launches of a call into
// times[], by the reference's float recurrence t += dt from t0 (one lane;
// the step launches that follow read it).
hipError_t launch_clock(float* times, float t0, float dt, int n_steps, int K, hipStream_t stream);

// Voxel populations built on the device (voxel_kernels.hip): one
// workgroup per robot of an nx*ny*nz lattice (at most kVoxMaxPoints
// points) writes the state and the mass-parallel kernel's packing.
constexpr int kVoxMaxPoints = 1024;
constexpr int kVoxMaxSprings = 13 * kVoxMaxPoints;
constexpr int kVoxPairWords = 512;  // (length index, material id) bits: 64 lengths x 256 ids
struct VoxelBuildParams {
  int nx, ny, nz, n, n_springs;
  int n_lengths;                 // distinct spring lengths of the shape (at most kVoxPairWords / 8)
  int maxp;                      // record slots per side of the mass kernel
  uint32_t valid_bits;           // encoding bits of the material set
  float mass;                    // every mass
  const uint32_t* enc;           // robot-major voxel encodings, n per robot
  const float* tpos;             // the shape: 3n positions
  const uint16_t* pairs;         // 2 per spring, DFS order
  const uint16_t* voxels;        // 4 per spring: points deciding its material (0xffff none)
  const uint8_t* length_index;   // per spring: index into lengths
  const float* lengths;          // distinct lengths
  const uint16_t* first_off;     // n + 1: springs whose first endpoint is the point
  const uint16_t* first_list;    //   (ascending spring index)
  const uint16_t* second_off;    // n + 1: springs whose second endpoint is the point
  const uint16_t* second_list;
  const uint8_t* code_id;        // decode table by encoded index: composite id,
  const uint32_t* code_enc;      //   the material's encoding,
  const uint8_t* code_solid;     //   1 unless every parameter is 0 (Strip's air test)
  // outputs, robot r at r * n (masses), r * n_springs (springs)
  float4* pos;
  float4* vel;
  uint8_t* mass_id;
  uint32_t* stripped;            // voxel encodings after Strip
  uint8_t* spring_id;
  uint32_t* mrec;                // (4 maxp + 1) n per robot
  uint8_t* mcnt;                 // 2n per robot
  uint16_t* mperm;               // n per robot
  float* lbtab;                  // 2 kMassPairMax per robot
  PackDesc* packs;               // one per robot
  uint32_t* active;              // non-air springs per robot
  int* error;                    // first error code, 0 if none
};
hipError_t launch_voxel_population(const VoxelBuildParams& p, int n_robots, hipStream_t stream);

// Per-robot COM of non-air masses (SoftBody::updateCOM order).
hipError_t launch_com(const float4* pos, const uint8_t* mass_id,
                      const uint32_t* robot_mass_begin, int n_robots,
                      float* com_out, hipStream_t stream);

}  // namespace mh
