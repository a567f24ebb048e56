// step_kernels.hip -- fused persistent step kernels for gfx950.
//
// One workgroup owns one pack (a contiguous range of robots) for n_steps
// substeps; the state is read from HBM once per launch and written back
// once.  No atomics: every mass receives one sum from +0 over its springs
// in ascending spring index, each spring's contribution added where the
// mass is the first endpoint and subtracted where it is the second, so
// results are bit-reproducible, independent of the launch plan and equal
// to the CPU mirror of the same step.
//
// Step semantics (reference sim_kernel.cu:141-269 for the XPBD solver;
// force solver per the project spec written in the header of oracle/step.c):
//   XPBD  x~ = x + v dt;  lambda = -(d - L)/(2 + ((1/k)/dt)/dt);
//         dp = lambda * (x~a - x~b)/(d + 1e-12);  x' = x~ + sum(+-dp);
//         v' = (float)(0.99 * (double)(x' - x) / (double)dt)
//   force D = xa - xb;  f = (k (L - d)) (1 / (d + 1e-12));
//         F = sum fma(+-f, D, .) - m g y^ - c_d v + ground normal/friction
//         (each spring's f and D folded into the running sum with one
//         rounding per component);  u = v + (F (1/m)) dt;  x' = x + u dt;
//         v' = damping * u
//   L = fma(Lbar, A sin(fma(omega, t, phi)), Lbar)
//
// Kernels:
//   step_mass    one lane per mass, robot resident in LDS; each spring is
//                evaluated by both endpoints from records held in VGPRs;
//                one barrier per substep (voxel robots);
//   step_massg   the same lanes, XPBD only; each spring evaluated once by
//                its first endpoint, the term passed to the second through
//                LDS; two barriers per substep;
//   step_lds     one lane per spring (records in VGPRs), terms in LDS, then
//                one lane per mass gathering them through CSR lists;
//   step_mstream robots over the register budget, force solver: one lane
//                per mass walks its springs from its own side, records
//                streamed from HBM every substep; one barrier per substep;
//   step_chunk   robots too large for LDS: positions in LDS, springs
//                streamed from HBM in LDS-sized chunks, the next chunk's
//                inputs loaded while the current one is worked on;
//   step_stream  the fallback for robots too large for step_chunk: spring
//                terms in HBM.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mercury_hip.h"
#include "mh_device.h"
#include "mh_kernels.h"

namespace mh {

namespace {

constexpr float kEps = 1e-12f;

struct MassState {
  float x, y, z;     // position
  float vx, vy, vz;  // velocity
  float tx, ty, tz;  // xpbd: predicted position x~
  float m, inv_m, mg;
};

// Phase-2 update of one mass from its gathered sum (ax, ay, az).
template <int SOLVER>
__device__ __forceinline__ float4 integrate(MassState& s, float ax, float ay, float az,
                                            const StepParams& p) {
  if constexpr (SOLVER == MH_SOLVER_XPBD_REF) {
    const float nx = s.tx + ax, ny = s.ty + ay, nz = s.tz + az;
    const double ddt = static_cast<double>(p.dt);
    s.vx = static_cast<float>((0.99 * static_cast<double>(nx - s.x)) / ddt);
    s.vy = static_cast<float>((0.99 * static_cast<double>(ny - s.y)) / ddt);
    s.vz = static_cast<float>((0.99 * static_cast<double>(nz - s.z)) / ddt);
    s.x = nx;
    s.y = ny;
    s.z = nz;
    s.tx = s.x + s.vx * p.dt;
    s.ty = s.y + s.vy * p.dt;
    s.tz = s.z + s.vz * p.dt;
    return make_float4(s.tx, s.ty, s.tz, 0.f);
  } else {
    float fx = ax, fy = ay, fz = az;
    float vx = s.vx, vy = s.vy, vz = s.vz;
    fy = fy - s.mg;
    if (p.drag != 0.0f) {
      fx = fx - p.drag * vx;
      fy = fy - p.drag * vy;
      fz = fz - p.drag * vz;
    }
    if (p.ground_k != 0.0f && s.y < 0.0f) {
      const float fn = p.ground_k * (-s.y);
      fy = fy + fn;
      const float fh = sqrtf(fx * fx + fz * fz);
      if (fh <= p.mu_s * fn) {
        fx = 0.0f;
        fz = 0.0f;
        vx = 0.0f;
        vz = 0.0f;
      } else {
        const float sc = 1.0f - (p.mu_k * fn) / fh;
        fx = fx * sc;
        fz = fz * sc;
      }
    }
    const float ux = vx + (fx * s.inv_m) * p.dt;
    const float uy = vy + (fy * s.inv_m) * p.dt;
    const float uz = vz + (fz * s.inv_m) * p.dt;
    s.x = s.x + ux * p.dt;
    s.y = s.y + uy * p.dt;
    s.z = s.z + uz * p.dt;
    s.vx = p.damping * ux;
    s.vy = p.damping * uy;
    s.vz = p.damping * uz;
    return make_float4(s.x, s.y, s.z, 0.f);
  }
}

__device__ __forceinline__ void load_mass(MassState& s, const float4* pos, const float4* vel,
                                          uint32_t gi, const StepParams& p) {
  const float4 q = pos[gi];
  const float4 w = vel[gi];
  s.x = q.x; s.y = q.y; s.z = q.z; s.m = q.w;
  s.vx = w.x; s.vy = w.y; s.vz = w.z;
  s.inv_m = 1.0f / s.m;
  s.mg = s.m * p.gravity;
  s.tx = s.x + s.vx * p.dt;
  s.ty = s.y + s.vy * p.dt;
  s.tz = s.z + s.vz * p.dt;
}

template <int SOLVER>
__device__ __forceinline__ float4 lds_position(const MassState& s) {
  if constexpr (SOLVER == MH_SOLVER_XPBD_REF) return make_float4(s.tx, s.ty, s.tz, 0.f);
  return make_float4(s.x, s.y, s.z, 0.f);
}

// XPBD term of one spring with actuated rest length L: the correction
// (cx, cy, cz) = lambda n and lambda.
template <int SOLVER>
__device__ __forceinline__ void spring_term_L(float4 pa, float4 pb, float L, float kk, float& cx, float& cy,
                                              float& cz, float& lambda) {
  static_assert(SOLVER == MH_SOLVER_XPBD_REF, "the force solver's term is force_factor");
  const float dx = pa.x - pb.x;
  const float dy = pa.y - pb.y;
  const float dz = pa.z - pb.z;
  const float d = norm3(dx, dy, dz);
  const float den = d + kEps;
  const float nx = dx / den, ny = dy / den, nz = dz / den;
  const float C = d - L;
  lambda = -C / kk;
  cx = lambda * nx;
  cy = lambda * ny;
  cz = lambda * nz;
}

// The same from the rest length lbar and actuation factor R:
// L = fma(lbar, R, lbar).
template <int SOLVER>
__device__ __forceinline__ void spring_term(float4 pa, float4 pb, float lbar, float R,
                                            float kk, float& cx, float& cy, float& cz,
                                            float& lambda) {
  spring_term_L<SOLVER>(pa, pb, fmaf(lbar, R, lbar), kk, cx, cy, cz, lambda);
}

// Force solver, one spring with actuated rest length L and stiffness kk:
// Hooke's law with one correctly rounded reciprocal, f = (k (L - d)) *
// (1 / (d + eps)), and the difference D = pa - pb.  Its term f D enters a
// mass's sum as fmaf(f, D, sum) per component (one rounding), or
// fmaf(-f, D, sum) where the mass is the second endpoint; from the second
// endpoint's own side D comes out as -D exactly, so both are fmaf(f, D,
// sum) with that side's D.
__device__ __forceinline__ void force_factor(float4 pa, float4 pb, float L, float kk, float& f, float& dx,
                                             float& dy, float& dz) {
  dx = pa.x - pb.x;
  dy = pa.y - pb.y;
  dz = pa.z - pb.z;
  float d, inv;
  length_and_inverse(dx, dy, dz, kEps, d, inv);
  f = (kk * (L - d)) * inv;
}

// Words of a stored spring term: xpbd (cx, cy, cz); force (Dx, Dy, Dz, f).
template <int SOLVER>
constexpr int term_words() { return solver_term_words(SOLVER); }

// Marks the unused w of a position read from LDS as used, with no
// instruction.  Left alone, the compiler narrows such float4 reads to
// ds_read_b96, which on gfx950 costs more LDS cycles than ds_read_b128
// when a wave's addresses are scattered (a wave reading random 16-byte
// slots: 17 cycles against 12).  Placed after x, y and z have been used,
// so it adds no wait.
__device__ __forceinline__ void keep_b128(float w) { asm volatile("" ::"v"(w)); }

// Adds four gathered terms to a sum, in entry order.  An entry (16 bits) is
// twice the word offset of a stored term, + 1 where the mass is the
// spring's second endpoint; g = +-1 from that bit.  xpbd terms (cx, cy, cz):
// fma(term, g, sum) is sum +- term rounded once.  force terms (Dx, Dy, Dz,
// f): fma(g f, D, sum), one rounding per component.
template <int SOLVER>
__device__ __forceinline__ void gather4(const float* s_term, uint2 w, float& ax, float& ay, float& az) {
  const float* q[4] = {s_term + ((w.x & 0xffffu) >> 1), s_term + (w.x >> 17), s_term + ((w.y & 0xffffu) >> 1),
                       s_term + (w.y >> 17)};
  const float g[4] = {__uint_as_float(0x3f800000u | (w.x << 31)),
                      __uint_as_float(0x3f800000u | ((w.x & 0x10000u) << 15)),
                      __uint_as_float(0x3f800000u | (w.y << 31)),
                      __uint_as_float(0x3f800000u | ((w.y & 0x10000u) << 15))};
  if constexpr (SOLVER == MH_SOLVER_FORCE) {
    float4 t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = *reinterpret_cast<const float4*>(q[i]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float sf = t[i].w * g[i];
      ax = fmaf(sf, t[i].x, ax); ay = fmaf(sf, t[i].y, ay); az = fmaf(sf, t[i].z, az);
    }
  } else {
    float x[4], y[4], z[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) { x[i] = q[i][0]; y[i] = q[i][1]; z[i] = q[i][2]; }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ax = fmaf(x[i], g[i], ax); ay = fmaf(y[i], g[i], ay); az = fmaf(z[i], g[i], az);
    }
  }
}

// ----------------------------------------------------------------- LDS kernel
// LDS layout (from LDS address 0): spring terms, TW floats per slot
// (xpbd (cx, cy, cz), force (Dx, Dy, Dz, f)), spring_cap + 2 slots rounded
// up to 16 B | positions float4[pos_slots(mass_cap)] | gather entries
// u16[csr_cap] | row bounds u16[mass_cap + 1] (the layout reserves
// 2 mass_cap + 2) | actuation float[n_mat].  Slot spring_cap is the sink of
// lanes without a spring; slot spring_cap + 1 holds zeros.  A gather entry
// is 2 * (TW * slot) of a term, + 1 where the mass is the spring's second
// endpoint.  Mass i's row [row[i], row[i+1]) lists its
// springs in ascending spring index, padded to a multiple of 4 with the
// zero slot (adding +0 leaves a sum that started at +0 unchanged); the
// mass receives the canonical sum.
// Springs are dealt to waves in contiguous runs of whole 64-lane
// iterations, so every wave runs ceil(springs / BLOCK) iterations.
// Phase-2 lanes visit masses in the order of p.mperm (longest rows first).
// Mass i's position sits in slot i + i/8 (pos_slot): the skew spreads the
// 16-byte position reads of a wave over the LDS banks; spring records carry
// skewed endpoint indices.
template <int SOLVER, int BLOCK, int NSPT, int MPT>
__global__ __launch_bounds__(BLOCK) void step_lds(StepParams p) {
  constexpr int kWaves = BLOCK / 64;
  extern __shared__ __align__(16) unsigned char smem[];
  const PackDesc d = p.packs[blockIdx.x];
  const int nm = static_cast<int>(d.n_mass);
  const int ns = static_cast<int>(d.n_spring);
  constexpr int TW = term_words<SOLVER>();
  const int sink = p.spring_cap, zero = p.spring_cap + 1;
  float* s_term = reinterpret_cast<float*>(smem);
  float4* s_pos = reinterpret_cast<float4*>(smem + lds_term_bytes(p.spring_cap, TW));
  uint16_t* s_csr = reinterpret_cast<uint16_t*>(s_pos + pos_slots(p.mass_cap));
  uint16_t* s_row = s_csr + p.csr_cap;
  float* s_R = reinterpret_cast<float*>(s_row + 2 * p.mass_cap + 2);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int iters = (ns + 64 * kWaves - 1) / (64 * kWaves);  // <= NSPT
  const int wave_first = wave * iters * 64;

  const int n_csr = static_cast<int>(p.row[d.row_begin + nm]);
  for (int i = tid; i < n_csr; i += BLOCK) s_csr[i] = static_cast<uint16_t>(p.csr[d.csr_begin + i]);
  for (int i = tid; i <= nm; i += BLOCK) s_row[i] = static_cast<uint16_t>(p.row[d.row_begin + i]);
  if (tid < TW) s_term[TW * zero + tid] = 0.f;

  // spring records -> registers; lanes past the pack's springs compute a
  // dummy spring (0,0) whose term goes to the sink slot
  uint32_t sp[NSPT];
  float lb[NSPT], kk[NSPT], st[NSPT];
  int mi[NSPT], slot[NSPT];
#pragma unroll
  for (int k = 0; k < NSPT; ++k) {
    const int s = wave_first + k * 64 + lane;
    sp[k] = 0; lb[k] = 1.f; kk[k] = 1.f; st[k] = 0.f; mi[k] = 0; slot[k] = TW * sink;
    if (k < iters && s < ns) {
      const uint32_t g = d.spring_begin + s;
      sp[k] = p.spair[g];
      lb[k] = p.slbar[g];
      mi[k] = p.smat[g];
      kk[k] = (SOLVER == MH_SOLVER_XPBD_REF) ? p.mat_kx[mi[k]] : p.mat_k[mi[k]];
      if (p.track_stress) st[k] = p.sstress[g];
      slot[k] = TW * s;
    }
  }
  // masses -> registers, first positions -> LDS
  MassState ms[MPT];
  int mid[MPT], pslot[MPT];
#pragma unroll
  for (int j = 0; j < MPT; ++j) {
    const int i = tid + j * BLOCK;
    mid[j] = -1;
    pslot[j] = 0;
    if (i < nm) {
      mid[j] = p.mperm[d.mass_begin + i];
      pslot[j] = pos_slot(mid[j]);
      load_mass(ms[j], p.pos, p.vel, d.mass_begin + mid[j], p);
      s_pos[pslot[j]] = lds_position<SOLVER>(ms[j]);
    }
  }
  float t = p.clock[p.launch];
  if (tid < p.n_mat) s_R[tid] = actuation(p.mat_dL0[tid], p.mat_omega[tid], p.mat_phi[tid], t);
  __syncthreads();

  for (int step = 0; step < p.n_steps; ++step) {
#pragma unroll
    for (int k = 0; k < NSPT; ++k) {
      if (k < iters) {
        const float4 pa = s_pos[sp[k] & 0xffffu];
        const float4 pb = s_pos[sp[k] >> 16];
        if constexpr (SOLVER == MH_SOLVER_FORCE) {
          float f, dx, dy, dz;
          force_factor(pa, pb, fmaf(lb[k], s_R[mi[k]], lb[k]), kk[k], f, dx, dy, dz);
          *reinterpret_cast<float4*>(s_term + slot[k]) = make_float4(dx, dy, dz, f);
        } else {
          float cx, cy, cz, lambda;
          spring_term<SOLVER>(pa, pb, lb[k], s_R[mi[k]], kk[k], cx, cy, cz, lambda);
          if (p.track_stress) st[k] = st[k] + lambda / lb[k];
          s_term[slot[k]] = cx;
          s_term[slot[k] + 1] = cy;
          s_term[slot[k] + 2] = cz;
        }
      }
    }
    __syncthreads();
    const float tn = t + p.dt;
#pragma unroll
    for (int j = 0; j < MPT; ++j) {
      if (mid[j] >= 0) {
        const int i = mid[j];
        const int e2 = s_row[i + 1];
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        // the mass's springs in ascending spring order, four gathers in
        // flight; an entry is 2 * (TW * slot) + 1 where the mass is the
        // spring's second endpoint
        for (int e = s_row[i]; e < e2; e += 4) {
          const uint2 w = *reinterpret_cast<const uint2*>(s_csr + e);
          gather4<SOLVER>(s_term, w, ax, ay, az);
        }
        s_pos[pslot[j]] = integrate<SOLVER>(ms[j], ax, ay, az, p);
      }
    }
    if (tid < p.n_mat) s_R[tid] = actuation(p.mat_dL0[tid], p.mat_omega[tid], p.mat_phi[tid], tn);
    t = tn;
    __syncthreads();
  }

#pragma unroll
  for (int j = 0; j < MPT; ++j) {
    if (mid[j] >= 0) {
      const uint32_t gi = d.mass_begin + mid[j];
      p.pos[gi] = make_float4(ms[j].x, ms[j].y, ms[j].z, ms[j].m);
      p.vel[gi] = make_float4(ms[j].vx, ms[j].vy, ms[j].vz, 0.f);
    }
  }
  if (SOLVER == MH_SOLVER_XPBD_REF && p.track_stress) {
#pragma unroll
    for (int k = 0; k < NSPT; ++k) {
      const int s = wave_first + k * 64 + lane;
      if (k < iters && s < ns) p.sstress[d.spring_begin + s] = st[k];
    }
  }
}

// ----------------------------------------------------------- streaming kernel
template <int SOLVER, int BLOCK>
__global__ __launch_bounds__(BLOCK) void step_stream(StepParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PackDesc d = p.packs[blockIdx.x];
  const int nm = static_cast<int>(d.n_mass);
  const int ns = static_cast<int>(d.n_spring);
  float4* s_pos = reinterpret_cast<float4*>(smem);
  float* s_R = reinterpret_cast<float*>(s_pos + p.mass_cap);
  const int tid = threadIdx.x;
  const uint32_t* spair = p.spair + d.spring_begin;
  const float* slbar = p.slbar + d.spring_begin;
  const uint8_t* smat = p.smat + d.spring_begin;
  float4* sc = reinterpret_cast<float4*>(p.scratch) + d.spring_begin;
  const uint32_t* csr = p.csr + d.csr_begin;
  const uint32_t* row = p.row + d.row_begin;

  // masses stay in the registers of their owner across substeps; the
  // owner loop is strided so any mass count up to the LDS capacity works
  for (int i = tid; i < nm; i += BLOCK) {
    MassState s;
    load_mass(s, p.pos, p.vel, d.mass_begin + i, p);
    s_pos[i] = lds_position<SOLVER>(s);
  }
  float t = p.clock[p.launch];
  if (tid < p.n_mat) s_R[tid] = actuation(p.mat_dL0[tid], p.mat_omega[tid], p.mat_phi[tid], t);
  __syncthreads();

  for (int step = 0; step < p.n_steps; ++step) {
    for (int s = tid; s < ns; s += BLOCK) {
      const uint32_t pr = spair[s];
      const int m = smat[s];
      const float kk = (SOLVER == MH_SOLVER_XPBD_REF) ? p.mat_kx[m] : p.mat_k[m];
      if constexpr (SOLVER == MH_SOLVER_FORCE) {
        float f, dx, dy, dz;
        force_factor(s_pos[pr & 0xffffu], s_pos[pr >> 16], fmaf(slbar[s], s_R[m], slbar[s]), kk, f, dx, dy, dz);
        sc[s] = make_float4(dx, dy, dz, f);
      } else {
        float cx, cy, cz, lambda;
        spring_term<SOLVER>(s_pos[pr & 0xffffu], s_pos[pr >> 16], slbar[s], s_R[m], kk, cx, cy,
                            cz, lambda);
        if (p.track_stress) {
          float* stp = p.sstress + d.spring_begin + s;
          *stp = *stp + lambda / slbar[s];
        }
        sc[s] = make_float4(cx, cy, cz, 0.f);
      }
    }
    __syncthreads();
    const float tn = t + p.dt;
    for (int i = tid; i < nm; i += BLOCK) {
      MassState s;
      // state lives in global memory between substeps for this kernel
      const float4 q = p.pos[d.mass_begin + i];
      const float4 w = p.vel[d.mass_begin + i];
      s.x = q.x; s.y = q.y; s.z = q.z; s.m = q.w;
      s.vx = w.x; s.vy = w.y; s.vz = w.z;
      s.inv_m = 1.0f / s.m;
      s.mg = s.m * p.gravity;
      const float4 lp = s_pos[i];
      s.tx = lp.x; s.ty = lp.y; s.tz = lp.z;
      // the mass's springs in ascending spring order; an entry is
      // 2 * slot, + 1 where the mass is the spring's second endpoint
      const int e2 = static_cast<int>(row[i + 1]);
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
      for (int e = static_cast<int>(row[i]); e < e2; ++e) {
        const uint32_t c = csr[e];
        const float4 q = sc[c >> 1];
        const float g = __uint_as_float(0x3f800000u | (c << 31));
        if constexpr (SOLVER == MH_SOLVER_FORCE) {
          const float sf = q.w * g;  // (Dx, Dy, Dz, f): fma(+-f, D, sum)
          ax = fmaf(sf, q.x, ax);
          ay = fmaf(sf, q.y, ay);
          az = fmaf(sf, q.z, az);
        } else {
          ax = fmaf(q.x, g, ax);
          ay = fmaf(q.y, g, ay);
          az = fmaf(q.z, g, az);
        }
      }
      s_pos[i] = integrate<SOLVER>(s, ax, ay, az, p);
      p.pos[d.mass_begin + i] = make_float4(s.x, s.y, s.z, s.m);
      p.vel[d.mass_begin + i] = make_float4(s.vx, s.vy, s.vz, 0.f);
    }
    if (tid < p.n_mat) s_R[tid] = actuation(p.mat_dL0[tid], p.mat_omega[tid], p.mat_phi[tid], tn);
    t = tn;
    __syncthreads();
  }
}

// ------------------------------------------------------------- chunked kernel
// For robots whose springs do not fit in LDS at once.  Positions stay in LDS
// (skewed slots, as in step_lds) and each lane keeps MPT masses in
// registers; every substep walks the robot's springs in chunks of
// p.spring_cap.  A chunk's inputs -- its spring records, its u16 gather
// lists and its row bounds (both prepared on the host, 8-byte aligned) --
// are loaded into registers one chunk ahead, so the global-memory latency
// of chunk c + 1 overlaps chunk c's work.  Phase 1 copies the chunk's gather
// lists and row bounds into LDS and writes its spring terms to LDS; after a
// barrier, phase 2 adds the chunk's terms to the running sum of each
// lane's masses in ascending spring order (+ where the mass is the first
// endpoint, - where it is the second); a barrier ends the chunk.  Chunks
// run in ascending spring order, so after the last chunk each mass holds
// its canonical sum and integrates.  LDS: terms | positions | gather entries
// u16[csr_cap] | row bounds u16[4 chunk_row_quads(mass_cap)] | actuation
// float[n_mat] | stiffness float[n_mat] | chunk entry offsets
// u32[kChunkMax + 1] (from p.ctab at d.pad0).
template <int SOLVER, int BLOCK, int MPT>
__global__ __launch_bounds__(BLOCK) void step_chunk(StepParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PackDesc d = p.packs[blockIdx.x];
  const int nm = static_cast<int>(d.n_mass);
  const int ns = static_cast<int>(d.n_spring);
  const int C = p.spring_cap;
  const int n_chunks = ns > 0 ? (ns + C - 1) / C : 1;
  constexpr int TW = term_words<SOLVER>();
  const int zero = C + 1;
  const int rq = chunk_row_quads(nm);
  float* s_term = reinterpret_cast<float*>(smem);
  float4* s_pos = reinterpret_cast<float4*>(smem + lds_term_bytes(C, TW));
  uint16_t* s_csr = reinterpret_cast<uint16_t*>(s_pos + pos_slots(p.mass_cap));
  uint16_t* s_row = s_csr + p.csr_cap;
  float* s_R = reinterpret_cast<float*>(s_row + 4 * chunk_row_quads(p.mass_cap));
  float* s_K = s_R + p.n_mat;
  uint32_t* s_eoff = reinterpret_cast<uint32_t*>(s_K + p.n_mat);
  const int tid = threadIdx.x;
  const uint32_t* spair = p.spair + d.spring_begin;
  const float* slbar = p.slbar + d.spring_begin;
  const uint8_t* smat = p.smat + d.spring_begin;
  const uint16_t* ccsr = p.ccsr + d.csr_begin;
  const uint16_t* crow = p.crow + d.row_begin;
  if (tid < TW) s_term[TW * zero + tid] = 0.f;
  for (int i = tid; i <= n_chunks; i += BLOCK) s_eoff[i] = p.ctab[d.pad0 + i];
  for (int i = tid; i < p.n_mat; i += BLOCK) s_K[i] = (SOLVER == MH_SOLVER_XPBD_REF) ? p.mat_kx[i] : p.mat_k[i];

  MassState ms[MPT];
  int pslot[MPT];
#pragma unroll
  for (int j = 0; j < MPT; ++j) {
    const int i = tid + j * BLOCK;
    pslot[j] = pos_slot(i);
    if (i < nm) {
      load_mass(ms[j], p.pos, p.vel, d.mass_begin + i, p);
      s_pos[pslot[j]] = lds_position<SOLVER>(ms[j]);
    }
  }
  float t = p.clock[p.launch];
  if (tid < p.n_mat) s_R[tid] = actuation(p.mat_dL0[tid], p.mat_omega[tid], p.mat_phi[tid], t);
  __syncthreads();

  // the next chunk's inputs, in flight while the current one is worked on
  uint32_t pr[kChunkSpl], mt[kChunkSpl];
  float lbv[kChunkSpl];
  uint2 ce[kChunkEpl], rw[kChunkRpl];
  auto fetch = [&](int c) {
    const int s0 = c * C, s1 = min(ns, s0 + C);
#pragma unroll
    for (int j = 0; j < kChunkSpl; ++j) {
      const int s = s0 + tid + j * BLOCK;
      if (s < s1) {
        pr[j] = spair[s];
        lbv[j] = slbar[s];
        mt[j] = smat[s];
      }
    }
    const int e0 = static_cast<int>(s_eoff[c]);
    const int ne4 = (static_cast<int>(s_eoff[c + 1]) - e0) >> 2;
    const uint2* src = reinterpret_cast<const uint2*>(ccsr + e0);
#pragma unroll
    for (int j = 0; j < kChunkEpl; ++j) {
      const int i = tid + j * BLOCK;
      if (i < ne4) ce[j] = src[i];
    }
    const uint2* rsrc = reinterpret_cast<const uint2*>(crow + static_cast<size_t>(4 * rq) * c);
#pragma unroll
    for (int j = 0; j < kChunkRpl; ++j) {
      const int i = tid + j * BLOCK;
      if (i < rq) rw[j] = rsrc[i];
    }
  };
  fetch(0);

  for (int step = 0; step < p.n_steps; ++step) {
    float A[MPT][3];
#pragma unroll
    for (int j = 0; j < MPT; ++j) A[j][0] = A[j][1] = A[j][2] = 0.0f;
    const float tn = t + p.dt;
    for (int c = 0; c < n_chunks; ++c) {
      const int s0 = c * C;
      const int s1 = min(ns, s0 + C);
      const int ne4 = (static_cast<int>(s_eoff[c + 1]) - static_cast<int>(s_eoff[c])) >> 2;
#pragma unroll
      for (int j = 0; j < kChunkEpl; ++j) {
        const int i = tid + j * BLOCK;
        if (i < ne4) reinterpret_cast<uint2*>(s_csr)[i] = ce[j];
      }
#pragma unroll
      for (int j = 0; j < kChunkRpl; ++j) {
        const int i = tid + j * BLOCK;
        if (i < rq) reinterpret_cast<uint2*>(s_row)[i] = rw[j];
      }
#pragma unroll
      for (int j = 0; j < kChunkSpl; ++j) {
        const int s = s0 + tid + j * BLOCK;
        if (s < s1) {
          const int m = static_cast<int>(mt[j]);
          const int o = TW * (s - s0);
          if constexpr (SOLVER == MH_SOLVER_FORCE) {
            float f, dx, dy, dz;
            force_factor(s_pos[pr[j] & 0xffffu], s_pos[pr[j] >> 16], fmaf(lbv[j], s_R[m], lbv[j]), s_K[m], f,
                         dx, dy, dz);
            *reinterpret_cast<float4*>(s_term + o) = make_float4(dx, dy, dz, f);
          } else {
            float cx, cy, cz, lambda;
            spring_term<SOLVER>(s_pos[pr[j] & 0xffffu], s_pos[pr[j] >> 16], lbv[j], s_R[m], s_K[m], cx, cy,
                                cz, lambda);
            if (p.track_stress) {
              float* stp = p.sstress + d.spring_begin + s;
              *stp = *stp + lambda / lbv[j];
            }
            s_term[o] = cx;
            s_term[o + 1] = cy;
            s_term[o + 2] = cz;
          }
        }
      }
      // the following chunk (the first one again after the last: the next
      // substep starts with it)
      fetch(c + 1 < n_chunks ? c + 1 : 0);
      __syncthreads();
      const bool last = c == n_chunks - 1;
#pragma unroll
      for (int j = 0; j < MPT; ++j) {
        const int i = tid + j * BLOCK;
        if (i < nm) {
          // the chunk's springs of the mass in ascending spring order; an
          // entry is 2 * (TW * (spring - chunk start)), + 1 where the mass
          // is the second endpoint
          const int e2 = s_row[i + 1];
          float ax = A[j][0], ay = A[j][1], az = A[j][2];
          for (int e = s_row[i]; e < e2; e += 4) {
            const uint2 w = *reinterpret_cast<const uint2*>(s_csr + e);
            gather4<SOLVER>(s_term, w, ax, ay, az);
          }
          A[j][0] = ax; A[j][1] = ay; A[j][2] = az;
          if (last) s_pos[pslot[j]] = integrate<SOLVER>(ms[j], ax, ay, az, p);
        }
      }
      if (last && tid < p.n_mat)
        s_R[tid] = actuation(p.mat_dL0[tid], p.mat_omega[tid], p.mat_phi[tid], tn);
      __syncthreads();
    }
    t = tn;
  }

#pragma unroll
  for (int j = 0; j < MPT; ++j) {
    const int i = tid + j * BLOCK;
    if (i < nm) {
      const uint32_t gi = d.mass_begin + i;
      p.pos[gi] = make_float4(ms[j].x, ms[j].y, ms[j].z, ms[j].m);
      p.vel[gi] = make_float4(ms[j].vx, ms[j].vy, ms[j].vz, 0.f);
    }
  }
}

// ------------------------------------------- mass-parallel streaming kernel
// For robots too large for the register-resident mass kernel (NN robots:
// 25 springs per mass, up to 2000 masses).  Like step_mass, every spring is
// evaluated by both of its endpoints from their own sides and each mass adds
// its list in ascending spring index (the canonical sum, one accumulator),
// so there is no term buffer, no gather list and one barrier per substep;
// unlike it, the records stream from memory every substep.  Each thread
// serves MPT virtual lanes v = tid + j BLOCK; virtual lanes are the robot's
// masses by degree, most springs first (p.mperm).  A record holds the
// other endpoint's slot, the spring's material and its rest length; the
// actuated length is fma(rest, R[material], rest).  The records of the 64
// virtual lanes of group g start at p.row[d.pad0 + g] (in 8-byte units,
// from d.csr_begin).  Lists are padded to the group's longest with records
// of the null material (stiffness 0, a ghost far from every mass), whose
// terms are signed zeros.  Two record formats (REC):
//   8  one uint2 per record {16 * slot | 8 * material << 16, rest length
//      bits}, both LDS byte offsets, laid out [k][64];
//   6  a u16 code 16 * slot | material (at most 15 materials and the null
//      one) and the f32 rest length, in two arrays of quads: the codes
//      uint2[n4][64] (4 codes per lane), then the rest lengths
//      float4[n4][64], n4 = ceil(longest / 4).  12 bytes per spring-update
//      instead of 16, in two loads per 4 records.
// A lane loads its records BATCH at a time, the next BATCH in flight while
// the current ones are evaluated.  LDS, a static array so that every base
// is an instruction offset: per material {R(t), k} float2[2][kMStreamMats]
// (the null material last) | position buffers
// float4[2][mass_slots(BLOCK MPT)] (read one, write the other).  Substeps
// run in pairs, so the buffer a substep reads is a compile-time choice.
template <int BLOCK, int MPT>
constexpr int mstream_pos_bytes() { return 16 * mass_slots(BLOCK * MPT); }
constexpr int kMStreamMatBytes = 2 * 8 * kMStreamMats;
template <int BLOCK, int MPT>
constexpr int mstream_lds_bytes() { return kMStreamMatBytes + 2 * mstream_pos_bytes<BLOCK, MPT>(); }

// Adds one record's spring to (ax, ay, az): the other endpoint's position
// at LDS byte offset po, its material's {R(t), k} at mo.
__device__ __forceinline__ void mstream_record(const unsigned char* pos_r, const unsigned char* mat_r, float4 own,
                                               uint32_t po, uint32_t mo, float lbar, float& ax, float& ay,
                                               float& az) {
  const float2 Rk = *reinterpret_cast<const float2*>(mat_r + mo);
  const float4 other = *reinterpret_cast<const float4*>(pos_r + po);
  float f, dx, dy, dz;
  force_factor(own, other, fmaf(lbar, Rk.x, lbar), Rk.y, f, dx, dy, dz);
  ax = fmaf(f, dx, ax);
  ay = fmaf(f, dy, ay);
  az = fmaf(f, dz, az);
  keep_b128(other.w);
}

// WPE: waves per SIMD the register allocation must allow (8 at BLOCK 1024:
// at most 64 VGPRs, so two workgroups share a CU); BATCH: records per load
// batch (a multiple of 4 for REC 6); REC: record format, 8 or 6 bytes.
template <int SOLVER, int BLOCK, int MPT, int WPE, int BATCH, int REC>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) void step_mstream(StepParams p) {
  static_assert(SOLVER == MH_SOLVER_FORCE, "the mass-parallel streaming kernel serves the force solver");
  static_assert(REC == 8 || (REC == 6 && BATCH % 4 == 0), "record format");
  static_assert(BLOCK % 64 == 0, "a wave serves one group of 64 virtual lanes");
  constexpr int kMStreamBatch = BATCH;
  __shared__ __align__(16) unsigned char lds[mstream_lds_bytes<BLOCK, MPT>()];
  unsigned char* const smem = lds + kMStreamMatBytes;  // the position buffers
  const PackDesc d = p.packs[blockIdx.x];
  const int nm = static_cast<int>(d.n_mass);
  constexpr int kSlots = mass_slots(BLOCK * MPT);
  constexpr int kGhost = kSlots - 1;
  constexpr int kPos = mstream_pos_bytes<BLOCK, MPT>();
  const int tid = threadIdx.x;
  const int n_mat = p.n_mat;
  const uint2* recs = reinterpret_cast<const uint2*>(p.mrec) + d.csr_begin;

  MassState ms[MPT];
  int slot[MPT], mass[MPT], nr_w[MPT];
  uint32_t gbase[MPT];
#pragma unroll
  for (int j = 0; j < MPT; ++j) {
    const int v = tid + j * BLOCK;
    const bool live = v < nm;
    mass[j] = live ? p.mperm[d.mass_begin + v] : 0;
    slot[j] = live ? mass_slot(mass[j]) : kGhost;
    if (live) load_mass(ms[j], p.pos, p.vel, d.mass_begin + mass[j], p);
    int n = live ? p.mcnt[d.row_begin + v] : 0;
    for (int off = 32; off > 0; off >>= 1) n = max(n, __shfl_xor(n, off));
    // wave-uniform (a wave serves one group): scalar registers and branches
    nr_w[j] = __builtin_amdgcn_readfirstlane(n);
    gbase[j] = __builtin_amdgcn_readfirstlane((v < ((nm + 63) & ~63)) ? p.row[d.pad0 + (v >> 6)] : 0u);
    if (live) reinterpret_cast<float4*>(smem)[slot[j]] = lds_position<SOLVER>(ms[j]);
  }
  if (tid == 0) {
    reinterpret_cast<float4*>(smem)[kGhost] = make_float4(1.0e3f, 1.0e3f, 1.0e3f, 0.f);
    reinterpret_cast<float4*>(smem + kPos)[kGhost] = make_float4(1.0e3f, 1.0e3f, 1.0e3f, 0.f);
  }
  float t = p.clock[p.launch];
  auto mat_table = [&](int buf, float tm) {
    for (int i = tid; i <= n_mat; i += BLOCK) {
      float2 e = make_float2(0.f, 0.f);  // the null material: R 0, stiffness 0
      if (i < n_mat)
        e = make_float2(actuation(p.mat_dL0[i], p.mat_omega[i], p.mat_phi[i], tm),
                        (SOLVER == MH_SOLVER_XPBD_REF) ? p.mat_kx[i] : p.mat_k[i]);
      reinterpret_cast<float2*>(lds + buf * 8 * kMStreamMats)[i] = e;
    }
  };
  mat_table(0, t);
  __syncthreads();

  auto substep = [&](auto cur_c) {
    constexpr int cur = decltype(cur_c)::value;
    const unsigned char* pos_r = smem + cur * kPos;
    unsigned char* pos_w = smem + (1 - cur) * kPos;
    const unsigned char* mat_r = lds + cur * 8 * kMStreamMats;
#pragma unroll
    for (int j = 0; j < MPT; ++j) {
      const int v = tid + j * BLOCK;
      const bool live = v < nm;
      const float4 own = live ? *reinterpret_cast<const float4*>(pos_r + 16 * slot[j])
                              : make_float4(0.f, 0.f, 0.f, 0.f);
      const int n = nr_w[j];
      const int lane = tid & 63;  // v & 63: BLOCK is a multiple of 64
      float ax = 0.f, ay = 0.f, az = 0.f;
      // The loads inside the loops are unconditional (an index past the
      // list reads the list's last record or quad again, never evaluated),
      // so the compiler waits only for the batch in use, not for the one
      // in flight.  n is wave-uniform.
      if constexpr (REC == 8) {
        const uint2* rv = recs + gbase[j] + lane;
        uint2 q[kMStreamBatch];
#pragma unroll
        for (int i = 0; i < kMStreamBatch; ++i) q[i] = make_uint2(0u, 0u);
        if (n > 0) {
#pragma unroll
          for (int i = 0; i < kMStreamBatch; ++i) q[i] = rv[64 * min(i, n - 1)];
        }
        for (int k0 = 0; k0 < n; k0 += kMStreamBatch) {
          uint2 nq[kMStreamBatch];
#pragma unroll
          for (int i = 0; i < kMStreamBatch; ++i) nq[i] = rv[64 * min(k0 + kMStreamBatch + i, n - 1)];
#pragma unroll
          for (int i = 0; i < kMStreamBatch; ++i) {
            if (k0 + i < n)
              mstream_record(pos_r, mat_r, own, q[i].x & 0xffffu, q[i].x >> 16, __uint_as_float(q[i].y), ax, ay,
                             az);
            // one record at a time: keeps the registers of the unrolled batch bounded
            __builtin_amdgcn_sched_barrier(0);
          }
#pragma unroll
          for (int i = 0; i < kMStreamBatch; ++i) q[i] = nq[i];
        }
      } else {
        // quads: codes cv[64 b], rest lengths lv[64 b]
        constexpr int kQ = kMStreamBatch / 4;
        const int n4 = (n + 3) >> 2;
        const uint2* cv = recs + gbase[j] + lane;
        const float4* lv = reinterpret_cast<const float4*>(recs + gbase[j] + 64 * n4) + lane;
        uint2 qc[kQ];
        float4 ql[kQ];
#pragma unroll
        for (int i = 0; i < kQ; ++i) {
          qc[i] = make_uint2(0u, 0u);
          ql[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (n4 > 0) {
#pragma unroll
          for (int i = 0; i < kQ; ++i) {
            qc[i] = cv[64 * min(i, n4 - 1)];
            ql[i] = lv[64 * min(i, n4 - 1)];
          }
        }
        for (int b0 = 0; b0 < n4; b0 += kQ) {
          uint2 nc[kQ];
          float4 nl[kQ];
#pragma unroll
          for (int i = 0; i < kQ; ++i) {
            const int b = min(b0 + kQ + i, n4 - 1);
            nc[i] = cv[64 * b];
            nl[i] = lv[64 * b];
          }
#pragma unroll
          for (int i = 0; i < kQ; ++i) {
            const int k = 4 * (b0 + i);
            const uint32_t w0 = qc[i].x, w1 = qc[i].y;
            if (k < n) mstream_record(pos_r, mat_r, own, w0 & 0xfff0u, (w0 << 3) & 0x78u, ql[i].x, ax, ay, az);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 1 < n)
              mstream_record(pos_r, mat_r, own, (w0 >> 16) & 0xfff0u, (w0 >> 13) & 0x78u, ql[i].y, ax, ay, az);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 2 < n) mstream_record(pos_r, mat_r, own, w1 & 0xfff0u, (w1 << 3) & 0x78u, ql[i].z, ax, ay, az);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 3 < n)
              mstream_record(pos_r, mat_r, own, (w1 >> 16) & 0xfff0u, (w1 >> 13) & 0x78u, ql[i].w, ax, ay, az);
            __builtin_amdgcn_sched_barrier(0);
          }
#pragma unroll
          for (int i = 0; i < kQ; ++i) {
            qc[i] = nc[i];
            ql[i] = nl[i];
          }
        }
      }
      if (live) *reinterpret_cast<float4*>(pos_w + 16 * slot[j]) = integrate<SOLVER>(ms[j], ax, ay, az, p);
    }
    t = t + p.dt;
    mat_table(1 - cur, t);
    __syncthreads();
  };
  int step = 0;
  for (; step + 1 < p.n_steps; step += 2) {
    substep(std::integral_constant<int, 0>{});
    substep(std::integral_constant<int, 1>{});
  }
  if (step < p.n_steps) substep(std::integral_constant<int, 0>{});

  // the mass index and the mass itself are read again here rather than
  // kept in registers through the substep loop
#pragma unroll
  for (int j = 0; j < MPT; ++j) {
    if (tid + j * BLOCK < nm) {
      const uint32_t gi = d.mass_begin + p.mperm[d.mass_begin + tid + j * BLOCK];
      p.pos[gi] = make_float4(ms[j].x, ms[j].y, ms[j].z, p.pos[gi].w);
      p.vel[gi] = make_float4(ms[j].vx, ms[j].vy, ms[j].vz, 0.f);
    }
  }
}

// ------------------------------------------------------- mass-parallel kernel
// One lane per mass; every spring is evaluated by both of its endpoints
// (the same operands in the same order, so both get the same bits), which
// removes the spring-term buffer and gather lists from LDS and leaves one
// barrier per substep.  LDS (fixed layout, so table and buffer bases are
// instruction offsets): a small area below kMassPos | positions
// float4[2][mass_slots(BLOCK)] (read one buffer, write the other; skewed
// slots mass_slot(i); the last slot of each buffer is a far-away ghost that
// padding records point at) | the pack's pair table, float2[2][kMassPairMax]
// {L, k} per (rest length, material) pair, L the actuated rest length of the
// substep that reads the buffer | float2[kMassPairMax] {1/k, rest length} |
// the pairs' constants float4[kMassPairMax] {dL0, omega, phi, rest
// length}.  Each substep the
// lanes below the pair count compute the next substep's L of one pair each
// (the actuation once per pair instead of once per spring).  Lanes take
// masses in the order of p.mperm (most records first, so a wave's lists
// have similar lengths).  Each lane keeps its records in registers: one
// record per spring of its mass, in ascending spring index (a spring from
// the mass to itself twice), padded to the wave's longest list with
// records of the null pair (last entry: stiffness 0 for force, +inf for
// xpbd), whose terms are signed zeros; a record holds byte offsets: the
// other endpoint's slot * 16 | pair index * 8 << 16.  The lane computes
// every term from its own side, which for a spring whose second endpoint
// it is gives the exact negation of the first endpoint's term, and adds
// them in list order: the canonical sum.  Records live in p.mrec as
// [part][lane]: parts 0..2 MAXP - 1 the records, part 2 MAXP a bit mask of
// the records whose first endpoint the mass is (they carry the stress),
// parts 2 MAXP + 1 + k the pack-local spring index of record k; p.mcnt
// holds two bytes per lane, the record count first.
template <int BLOCK>
constexpr int mass_pos_bytes() { return 16 * mass_slots(BLOCK); }

// Fills a pack's pair tables (d.pad0: p.lbtab offset of its pairs, two
// words each: rest length, compact material; d.pad1: pair count) for the
// time t: the constants, {1/k, rest length} and {L(t), k} in both dynamic
// buffers (the second's L is overwritten before it is read).  null_k and
// null_ky are the null pair's stiffness and reciprocal entries.
template <int SOLVER, int BLOCK>
__device__ __forceinline__ void mass_pairs_init(unsigned char* smem, const StepParams& p, const PackDesc& d,
                                                int dyn0, int dyn1, int aux, int cst, float t, float null_k,
                                                float null_ky) {
  const int n = static_cast<int>(d.pad1);
  for (int i = static_cast<int>(threadIdx.x); i < n; i += BLOCK) {
    const float lb = p.lbtab[d.pad0 + 2 * i];
    const uint32_t c = __float_as_uint(p.lbtab[d.pad0 + 2 * i + 1]);
    const float k = (SOLVER == MH_SOLVER_XPBD_REF) ? p.mat_kx[c] : p.mat_k[c];
    const float4 cs = make_float4(p.mat_dL0[c], p.mat_omega[c], p.mat_phi[c], lb);
    const float L = fmaf(lb, actuation(cs.x, cs.y, cs.z, t), lb);
    const float ky = (SOLVER == MH_SOLVER_XPBD_REF) ? 1.0f / k : 0.0f;
    reinterpret_cast<float4*>(smem + cst)[i] = cs;
    reinterpret_cast<float2*>(smem + aux)[i] = make_float2(ky, lb);
    reinterpret_cast<float2*>(smem + dyn0)[i] = make_float2(L, k);
    reinterpret_cast<float2*>(smem + dyn1)[i] = make_float2(L, k);
  }
  if (threadIdx.x == 0) {
    reinterpret_cast<float4*>(smem + cst)[kMassPairNull] = make_float4(0.f, 0.f, 0.f, 1.f);
    reinterpret_cast<float2*>(smem + aux)[kMassPairNull] = make_float2(null_ky, 1.f);
    reinterpret_cast<float2*>(smem + dyn0)[kMassPairNull] = make_float2(1.f, null_k);
    reinterpret_cast<float2*>(smem + dyn1)[kMassPairNull] = make_float2(1.f, null_k);
  }
}

// The next substep's actuated rest lengths: L(t) of every pair into the
// first word of its dyn entry (STRIDE words per entry).
template <int BLOCK, int STRIDE = 2>
__device__ __forceinline__ void mass_pairs_advance(unsigned char* smem, int n_pairs, int dyn, int cst, float t) {
  for (int i = static_cast<int>(threadIdx.x); i < n_pairs; i += BLOCK) {
    const float4 cs = reinterpret_cast<const float4*>(smem + cst)[i];
    reinterpret_cast<float*>(smem + dyn)[STRIDE * i] = fmaf(cs.w, actuation(cs.x, cs.y, cs.z, t), cs.w);
  }
}

// step_massg's pair table: one float4 {L(t), k, 1/k, rest length} per pair
// at dyn (the null pair's stiffness 1 with reciprocal 0: a signed-zero
// term), so a record reads its pair with one ds_read_b128; the constants
// {dL0, omega, phi, rest length} at cst.
template <int BLOCK>
__device__ __forceinline__ void massg_pairs_init(unsigned char* smem, const StepParams& p, const PackDesc& d,
                                                 int dyn, int cst, float t) {
  const int n = static_cast<int>(d.pad1);
  for (int i = static_cast<int>(threadIdx.x); i < n; i += BLOCK) {
    const float lb = p.lbtab[d.pad0 + 2 * i];
    const uint32_t c = __float_as_uint(p.lbtab[d.pad0 + 2 * i + 1]);
    const float k = p.mat_kx[c];
    const float4 cs = make_float4(p.mat_dL0[c], p.mat_omega[c], p.mat_phi[c], lb);
    reinterpret_cast<float4*>(smem + cst)[i] = cs;
    reinterpret_cast<float4*>(smem + dyn)[i] = make_float4(fmaf(lb, actuation(cs.x, cs.y, cs.z, t), lb), k, 1.0f / k, lb);
  }
  if (threadIdx.x == 0) {
    reinterpret_cast<float4*>(smem + cst)[kMassPairNull] = make_float4(0.f, 0.f, 0.f, 1.f);
    reinterpret_cast<float4*>(smem + dyn)[kMassPairNull] = make_float4(1.f, 1.f, 0.f, 1.f);
  }
}

// One substep reading position and pair buffer CUR and writing the other.
template <int SOLVER, int BLOCK, int MAXP, bool STRESS, int CUR>
__device__ __forceinline__ void mass_substep(unsigned char* smem, const StepParams& p, MassState& ms,
                                             uint32_t (&rr)[2 * MAXP], float (&st)[STRESS ? 2 * MAXP : 1],
                                             uint32_t first, int nr, int nr_w, bool live, int pslot, int n_pairs,
                                             float t) {
  constexpr int kP0 = kMassPos + CUR * mass_pos_bytes<BLOCK>();
  constexpr int kP1 = kMassPos + (1 - CUR) * mass_pos_bytes<BLOCK>();
  constexpr int kPairs = mass_pair_base(BLOCK, 2);
  constexpr int kD0 = kPairs + CUR * kMassDynBytes;
  constexpr int kD1 = kPairs + (1 - CUR) * kMassDynBytes;
  constexpr int kAux = kPairs + 2 * kMassDynBytes;
  constexpr int kCst = kPairs + 3 * kMassDynBytes;
  // idle lanes sit at the origin, 1e3*sqrt(3) from the ghost, so their
  // padding terms stay in the fast arithmetic range
  const float4 own = live ? *reinterpret_cast<const float4*>(smem + kP0 + 16 * pslot)
                          : make_float4(0.f, 0.f, 0.f, 0.f);
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
  for (int k = 0; k < 2 * MAXP; ++k) {
    if (k < nr_w) {
      // opaque to the optimizer: decoding the record per substep costs two
      // VALU, hoisting every decoded field out of the step loop would cost
      // registers
      asm volatile("" : "+v"(rr[k]));
      const uint32_t r = rr[k];
      const float2 Lk = *reinterpret_cast<const float2*>(smem + kD0 + (r >> 16));
      const float4 other = *reinterpret_cast<const float4*>(smem + kP0 + (r & 0xffffu));
      // padding records add a signed zero, which leaves a sum that
      // started at +0 unchanged
      if constexpr (SOLVER == MH_SOLVER_FORCE) {
        float f, dx, dy, dz;
        force_factor(own, other, Lk.x, Lk.y, f, dx, dy, dz);
        ax = fmaf(f, dx, ax); ay = fmaf(f, dy, ay); az = fmaf(f, dz, az);
      } else {
        float cx, cy, cz, lambda;
        spring_term_L<SOLVER>(own, other, Lk.x, Lk.y, cx, cy, cz, lambda);
        ax = ax + cx; ay = ay + cy; az = az + cz;
        if constexpr (STRESS)
          if (k < nr && ((first >> k) & 1u))
            st[k] = st[k] + lambda / *reinterpret_cast<const float*>(smem + kAux + (r >> 16) + 4);
      }
      keep_b128(other.w);
    }
    // one record at a time: keeps the unrolled loop's registers bounded
    __builtin_amdgcn_sched_barrier(0);
  }
  if (live) *reinterpret_cast<float4*>(smem + kP1 + 16 * pslot) = integrate<SOLVER>(ms, ax, ay, az, p);
  mass_pairs_advance<BLOCK>(smem, n_pairs, kD1, kCst, t + p.dt);
  __syncthreads();
}

// STRESS (xpbd only) accumulates per-spring stress; the variant without it
// frees 2 MAXP registers per lane.
template <int SOLVER, int BLOCK, int MAXP, bool STRESS>
__global__ __launch_bounds__(BLOCK) void step_mass(StepParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PackDesc d = p.packs[blockIdx.x];
  const int nm = static_cast<int>(d.n_mass);
  constexpr int kSlots = mass_slots(BLOCK);
  constexpr int kGhost = kSlots - 1;
  constexpr int kPairs = mass_pair_base(BLOCK, 2);
  const int tid = threadIdx.x;
  const bool live = tid < nm;
  const uint32_t* rec = p.mrec + d.csr_begin;

  MassState ms;
  uint32_t rr[2 * MAXP];
  float st[STRESS ? 2 * MAXP : 1];
  uint32_t first = 0u;
  int nr = 0, mass = 0;
  // padding record: the ghost, with the null pair
  const uint32_t pad = 16u * kGhost | (8u * kMassPairNull) << 16;
#pragma unroll
  for (int k = 0; k < 2 * MAXP; ++k) rr[k] = pad;
#pragma unroll
  for (int k = 0; k < (STRESS ? 2 * MAXP : 1); ++k) st[k] = 0.f;
  if (live) {
    mass = p.mperm[d.mass_begin + tid];
    load_mass(ms, p.pos, p.vel, d.mass_begin + mass, p);
    nr = p.mcnt[d.row_begin + 2 * tid];
#pragma unroll
    for (int k = 0; k < 2 * MAXP; ++k)
      if (k < nr) rr[k] = rec[static_cast<size_t>(k) * nm + tid];
    if constexpr (STRESS) {
      first = rec[static_cast<size_t>(2 * MAXP) * nm + tid];
#pragma unroll
      for (int k = 0; k < 2 * MAXP; ++k)
        if (k < nr && ((first >> k) & 1u))
          st[k] = p.sstress[d.spring_begin + rec[static_cast<size_t>(2 * MAXP + 1 + k) * nm + tid]];
    }
  }
  const int pslot = live ? mass_slot(mass) : kGhost;
  float4* pos0 = reinterpret_cast<float4*>(smem + kMassPos);
  float4* pos1 = reinterpret_cast<float4*>(smem + kMassPos + mass_pos_bytes<BLOCK>());
  if (live) pos0[pslot] = lds_position<SOLVER>(ms);
  if (tid == 0) {
    pos0[kGhost] = make_float4(1.0e3f, 1.0e3f, 1.0e3f, 0.f);
    pos1[kGhost] = make_float4(1.0e3f, 1.0e3f, 1.0e3f, 0.f);
  }
  float t = p.clock[p.launch];
  // null pair: a term with stiffness 0 (force) or divisor +inf (xpbd) is a
  // signed zero
  mass_pairs_init<SOLVER, BLOCK>(smem, p, d, kPairs, kPairs + kMassDynBytes, kPairs + 2 * kMassDynBytes,
                                 kPairs + 3 * kMassDynBytes, t,
                                 (SOLVER == MH_SOLVER_XPBD_REF) ? __builtin_inff() : 0.0f, 0.0f);
  const int n_pairs = static_cast<int>(d.pad1);
  // wave-wide loop bound
  int nr_w = nr;
  for (int off = 32; off > 0; off >>= 1) nr_w = max(nr_w, __shfl_xor(nr_w, off));
  __syncthreads();

  // substeps in pairs, so each buffer's base is a constant offset
  int step = 0;
  for (; step + 1 < p.n_steps; step += 2) {
    mass_substep<SOLVER, BLOCK, MAXP, STRESS, 0>(smem, p, ms, rr, st, first, nr, nr_w, live, pslot, n_pairs, t);
    t = t + p.dt;
    mass_substep<SOLVER, BLOCK, MAXP, STRESS, 1>(smem, p, ms, rr, st, first, nr, nr_w, live, pslot, n_pairs, t);
    t = t + p.dt;
  }
  if (step < p.n_steps)
    mass_substep<SOLVER, BLOCK, MAXP, STRESS, 0>(smem, p, ms, rr, st, first, nr, nr_w, live, pslot, n_pairs, t);

  if (live) {
    const uint32_t gi = d.mass_begin + mass;
    p.pos[gi] = make_float4(ms.x, ms.y, ms.z, ms.m);
    p.vel[gi] = make_float4(ms.vx, ms.vy, ms.vz, 0.f);
    if constexpr (STRESS) {
#pragma unroll
      for (int k = 0; k < 2 * MAXP; ++k)
        if (k < nr && ((first >> k) & 1u))
          p.sstress[d.spring_begin + rec[static_cast<size_t>(2 * MAXP + 1 + k) * nm + tid]] = st[k];
    }
  }
}

// XPBD term of the gather kernel.  fast_ok says that the workgroup's
// positions and the pack's materials lie where div_rn_from_rcp is exact
// (coordinates 0 or of magnitude in [2^-56, 2^20]; stiffness divisor in
// [2^-40, 2^40], rest lengths in [2^-40, 2^20], amplitude at most 1/2); the
// wave then takes the short sequence if every lane's squared length is also
// in the fast sqrt range.  ky = RN(1/kk); 0 marks the null pair, whose
// term is a signed zero.  Same results as spring_term_L<XPBD> bit for bit.
__device__ __forceinline__ void xpbd_term_L(float4 pa, float4 pb, float L, float kk, float ky,
                                            bool fast_ok, float& cx, float& cy, float& cz, float& lambda) {
  const float dx = pa.x - pb.x;
  const float dy = pa.y - pb.y;
  const float dz = pa.z - pb.z;
  const float s = sum_sq3(dx, dy, dz);
  // in [2^-30, 2^126]: d + kEps == d (see kLenFastLo)
  const bool fast = fast_ok && __float_as_uint(s) - __float_as_uint(kLenFastLo) <=
                                   __float_as_uint(kSqrtFastHi) - __float_as_uint(kLenFastLo);
  float nx, ny, nz;
  if (__builtin_amdgcn_ballot_w64(fast) == __builtin_amdgcn_read_exec()) {
    const float d = sqrt_rn_inrange(s);
    const float y = rcp_rn_inrange(d);
    nx = div_rn_from_rcp(dx, d, y);
    ny = div_rn_from_rcp(dy, d, y);
    nz = div_rn_from_rcp(dz, d, y);
    lambda = div_rn_from_rcp(-(d - L), kk, ky);
  } else {
    const float d = sqrtf(s);
    const float den = d + kEps;
    nx = dx / den;
    ny = dy / den;
    nz = dz / den;
    lambda = ky == 0.0f ? 0.0f : -(d - L) / kk;
  }
  cx = lambda * nx;
  cy = lambda * ny;
  cz = lambda * nz;
}

// Position of a mass in the fast division's range: each coordinate 0 or of
// magnitude in [2^-56, 2^20] (NaN is not).
__device__ __forceinline__ bool position_in_div_range(float4 q) {
  auto ok = [](float x) {
    const float a = fabsf(x);
    return a == 0.0f || (a >= 0x1p-56f && a <= 0x1p20f);
  };
  return ok(q.x) && ok(q.y) && ok(q.z);
}

// ------------------------------------------------ mass-parallel gather kernel
// As step_mass, but each spring is evaluated once, by its first endpoint.
// Phase 1: a lane computes the terms of the springs whose first endpoint it
// is and stores each in its own LDS term slots (12 B per (lane, k),
// lane-major).  Barrier.  Phase 2: the lane walks one list of entries, one
// per spring of its mass in ascending spring index, each the LDS byte
// offset of the spring's term (its own slot or the first endpoint's) with
// bit 0 set where the mass is the second endpoint; it adds
// fma(term, +-1, sum) (the product by +-1 is exact, so this is the sum plus
// or minus the term, rounded once: the canonical sum), integrates and
// writes its position in place; the next substep's actuated rest lengths
// are written in place too.  Barrier.  Two barriers per substep, half the
// spring arithmetic of step_mass.  LDS: a zero term at kMassZero and two
// flags at kMassZero + 16 | positions float4[mass_slots(BLOCK)] | one pair
// table float4 {L, k, 1/k, rest length} and the constants | terms.  p.mrec
// parts: 0..MAXP - 1 the first-endpoint records (as step_mass's, with the
// pair's offset 16 * pair for the float4 table), MAXP..3 MAXP - 1 the
// entries (padding: the zero term), 3 MAXP + k the pack-local spring index
// of first-endpoint record k; p.mcnt: two counts per lane (records,
// entries).  XPBD: the
// flags (one per substep parity) record whether any position written for
// that substep left the fast division's range.
template <int SOLVER, int BLOCK, int MAXP, bool STRESS>
__global__ __launch_bounds__(BLOCK) void step_massg(StepParams p) {
  static_assert(SOLVER == MH_SOLVER_XPBD_REF,
                "step_massg stores 12-byte terms: the xpbd step only (the force step_mass is faster)");
  extern __shared__ __align__(16) unsigned char smem[];
  const PackDesc d = p.packs[blockIdx.x];
  const int nm = static_cast<int>(d.n_mass);
  constexpr int kSlots = mass_slots(BLOCK);
  constexpr int kGhost = kSlots - 1;
  constexpr int kDyn = mass_pair_base(BLOCK, 1);  // float4 per pair, over two float2 tables' room
  constexpr int kCst = kDyn + 2 * kMassDynBytes;
  constexpr int kTerm = massg_term_base(BLOCK);
  const int tid = threadIdx.x;
  const bool live = tid < nm;
  const uint32_t* rec = p.mrec + d.csr_begin;

  MassState ms;
  uint32_t rp[MAXP], re[2 * MAXP];
  float st[MAXP];
  int np = 0, ne = 0, mass = 0;
  const uint32_t pad = 16u * kGhost | (16u * kMassPairNull) << 16;
#pragma unroll
  for (int k = 0; k < MAXP; ++k) {
    rp[k] = pad;
    st[k] = 0.f;
  }
#pragma unroll
  for (int k = 0; k < 2 * MAXP; ++k) re[k] = kMassZero;
  if (live) {
    mass = p.mperm[d.mass_begin + tid];
    load_mass(ms, p.pos, p.vel, d.mass_begin + mass, p);
    np = p.mcnt[d.row_begin + 2 * tid];
    ne = p.mcnt[d.row_begin + 2 * tid + 1];
#pragma unroll
    for (int k = 0; k < MAXP; ++k) {
      if (k < np) rp[k] = rec[static_cast<size_t>(k) * nm + tid];
      if (STRESS && k < np) st[k] = p.sstress[d.spring_begin + rec[static_cast<size_t>(3 * MAXP + k) * nm + tid]];
    }
#pragma unroll
    for (int k = 0; k < 2 * MAXP; ++k)
      if (k < ne) re[k] = rec[static_cast<size_t>(MAXP + k) * nm + tid];
  }
  const int pslot = live ? mass_slot(mass) : kGhost;
  float4* pos = reinterpret_cast<float4*>(smem + kMassPos);
  if (live) pos[pslot] = lds_position<SOLVER>(ms);
  uint32_t* s_flag = reinterpret_cast<uint32_t*>(smem + kMassZero + 16);
  float t = p.clock[p.launch];
  // one pair table, its rest lengths updated in place in phase 2
  massg_pairs_init<BLOCK>(smem, p, d, kDyn, kCst, t);
  const int n_pairs = static_cast<int>(d.pad1);
  if (tid == 0) {
    pos[kGhost] = make_float4(1.0e3f, 1.0e3f, 1.0e3f, 0.f);
    float* z = reinterpret_cast<float*>(smem + kMassZero);
    z[0] = z[1] = z[2] = 0.0f;
    s_flag[0] = 0u;
    s_flag[1] = 0u;
  }
  int np_w = np, ne_w = ne;
  for (int off = 32; off > 0; off >>= 1) {
    np_w = max(np_w, __shfl_xor(np_w, off));
    ne_w = max(ne_w, __shfl_xor(ne_w, off));
  }
  unsigned char* mine = smem + kTerm + 12 * MAXP * tid;
  if (SOLVER == MH_SOLVER_XPBD_REF) {
    __syncthreads();  // flags cleared
    if (live && !position_in_div_range(lds_position<SOLVER>(ms))) atomicOr(&s_flag[0], 1u);
  }
  __syncthreads();

  for (int step = 0; step < p.n_steps; ++step) {
    const float4 own = live ? pos[pslot] : make_float4(0.f, 0.f, 0.f, 0.f);
    bool fast_ok = false;
    if (SOLVER == MH_SOLVER_XPBD_REF) {
      fast_ok = p.fast_div && s_flag[step & 1] == 0u;
      if (tid == 0) s_flag[(step + 1) & 1] = 0u;  // read last in the previous substep
    }
#pragma unroll
    for (int k = 0; k < MAXP; ++k) {
      if (k < np_w) {
        const uint32_t r = rp[k];
        const float4 other = *reinterpret_cast<const float4*>(smem + kMassPos + (r & 0xffffu));
        float cx, cy, cz, lambda;
        const float4 e = *reinterpret_cast<const float4*>(smem + kDyn + (r >> 16));  // {L, k, 1/k, rest}
        xpbd_term_L(own, other, e.x, e.y, e.z, fast_ok, cx, cy, cz, lambda);
        keep_b128(other.w);
        keep_b128(e.w);
        if (STRESS && k < np) st[k] = st[k] + lambda / e.w;
        float* slot = reinterpret_cast<float*>(mine + 12 * k);
        slot[0] = cx;
        slot[1] = cy;
        slot[2] = cz;
      }
    }
    __syncthreads();
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
    for (int k = 0; k < 2 * MAXP; ++k) {
      if (k < ne_w) {
        const uint32_t e = re[k];
        const float* q = reinterpret_cast<const float*>(smem + (e & ~1u));
        const float sg = __uint_as_float(0x3f800000u | (e << 31));
        ax = fmaf(q[0], sg, ax);
        ay = fmaf(q[1], sg, ay);
        az = fmaf(q[2], sg, az);
      }
    }
    if (live) {
      const float4 np4 = integrate<SOLVER>(ms, ax, ay, az, p);
      pos[pslot] = np4;
      if (SOLVER == MH_SOLVER_XPBD_REF && !position_in_div_range(np4)) atomicOr(&s_flag[(step + 1) & 1], 1u);
    }
    mass_pairs_advance<BLOCK, 4>(smem, n_pairs, kDyn, kCst, t + p.dt);
    t = t + p.dt;
    __syncthreads();
  }

  if (live) {
    const uint32_t gi = d.mass_begin + mass;
    p.pos[gi] = make_float4(ms.x, ms.y, ms.z, ms.m);
    p.vel[gi] = make_float4(ms.vx, ms.vy, ms.vz, 0.f);
    if (STRESS) {
#pragma unroll
      for (int k = 0; k < MAXP; ++k)
        if (k < np) p.sstress[d.spring_begin + rec[static_cast<size_t>(3 * MAXP + k) * nm + tid]] = st[k];
    }
  }
}

// ---------------------------------------------------------------- clock
// times[i] = start time of launch i: the float recurrence t += dt, K
// additions per launch, from t0.
__global__ void set_clock(float* times, float t0, float dt, int n_steps, int K) {
  if (threadIdx.x != 0) return;
  float t = t0;
  int launch = 0;
  for (int done = 0; done < n_steps; done += K) {
    times[launch++] = t;
    const int n = min(K, n_steps - done);
    for (int k = 0; k < n; ++k) t = t + dt;
  }
}

// ------------------------------------------------------------------ COM
__global__ void robot_com(const float4* pos, const uint8_t* mass_id, const uint32_t* begin,
                          int n_robots, float* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_robots) return;
  float cx = 0.f, cy = 0.f, cz = 0.f, count = 0.f;
  for (uint32_t i = begin[r]; i < begin[r + 1]; ++i) {
    if (mass_id[i] == 0) continue;
    const float4 q = pos[i];
    const float den = count + 1.f;
    cx = cx + (q.x - cx) / den;
    cy = cy + (q.y - cy) / den;
    cz = cz + (q.z - cz) / den;
    count = den;
  }
  out[3 * r] = cx;
  out[3 * r + 1] = cy;
  out[3 * r + 2] = cz;
}

#define MH_LDS(SOLVER, B, S, M) \
  {SOLVER, MH_KERNEL_LDS, 0, B, S, M, reinterpret_cast<const void*>(&step_lds<SOLVER, B, S, M>)}
#define MH_MASS(SOLVER, B, P)                                                         \
  {SOLVER, MH_KERNEL_MASS, 0, B, P, 1, reinterpret_cast<const void*>(&step_mass<SOLVER, B, P, false>), \
   reinterpret_cast<const void*>(&step_mass<SOLVER, B, P, (SOLVER == MH_SOLVER_XPBD_REF)>)}
#define MH_MASSG(SOLVER, B, P)                                                        \
  {SOLVER, MH_KERNEL_MASS, 1, B, P, 1, reinterpret_cast<const void*>(&step_massg<SOLVER, B, P, false>), \
   reinterpret_cast<const void*>(&step_massg<SOLVER, B, P, (SOLVER == MH_SOLVER_XPBD_REF)>)}
#define MH_CHUNK(SOLVER, B, M) \
  {SOLVER, MH_KERNEL_STREAM, 1, B, 0, M, reinterpret_cast<const void*>(&step_chunk<SOLVER, B, M>)}
#define MH_STREAM(SOLVER, B) \
  {SOLVER, MH_KERNEL_STREAM, 0, B, 0, 0, reinterpret_cast<const void*>(&step_stream<SOLVER, B>)}
#define MH_MSTREAM(B, M, W, BT, R)                                                             \
  {MH_SOLVER_FORCE, MH_KERNEL_STREAM, 2, B, R, M,                                           \
   reinterpret_cast<const void*>(&step_mstream<MH_SOLVER_FORCE, B, M, W, BT, R>)}
#define MH_LDS_SET(SOLVER)                                                         \
      MH_MASS(SOLVER, 768, 13), MH_MASS(SOLVER, 1024, 13), MH_MASS(SOLVER, 512, 16),   \
      MH_MASS(SOLVER, 256, 16), MH_LDS(SOLVER, 1024, 8, 1), MH_LDS(SOLVER, 1024, 8, 2), MH_LDS(SOLVER, 1024, 4, 1), \
      MH_LDS(SOLVER, 512, 8, 2), MH_LDS(SOLVER, 512, 16, 2), \
      MH_LDS(SOLVER, 512, 16, 4), MH_LDS(SOLVER, 256, 16, 4), MH_LDS(SOLVER, 256, 8, 2), \
      MH_CHUNK(SOLVER, 1024, 2), MH_CHUNK(SOLVER, 512, 4),                            \
      MH_STREAM(SOLVER, 1024), MH_STREAM(SOLVER, 512), MH_STREAM(SOLVER, 256)

// The mass-parallel streaming kernel serves the force solver (the XPBD
// step's stress output needs each spring's index); it precedes the other
// streaming variants, so plans that are not autotuned stream with it.
// The gather flavour of the mass-parallel kernel serves the xpbd step (its
// division-heavy term is computed once); the force solver's step_mass is
// faster than it.
const KernelVariant kVariants[] = {MH_MASSG(MH_SOLVER_XPBD_REF, 768, 13), MH_MASSG(MH_SOLVER_XPBD_REF, 512, 16),
                                   MH_MASSG(MH_SOLVER_XPBD_REF, 256, 16), MH_LDS_SET(MH_SOLVER_XPBD_REF),
                                   MH_MSTREAM(1024, 2, 8, 4, 6), MH_MSTREAM(512, 4, 4, 4, 6),
                                   MH_MSTREAM(512, 4, 4, 8, 6), MH_MSTREAM(1024, 2, 8, 4, 8),
                                   MH_MSTREAM(1024, 2, 8, 2, 8), MH_MSTREAM(512, 4, 4, 4, 8),
                                   MH_LDS_SET(MH_SOLVER_FORCE)};

}  // namespace

const KernelVariant* kernel_variants(int* count) {
  *count = static_cast<int>(sizeof(kVariants) / sizeof(kVariants[0]));
  return kVariants;
}

size_t lds_bytes_lds(int mass_cap, int spring_cap, int csr_cap, int n_mat, int tw) {
  size_t b = lds_term_bytes(spring_cap, tw);            // spring terms, sink, zero
  b += 16u * static_cast<size_t>(pos_slots(mass_cap));  // float4 positions (skewed)
  b += 2u * static_cast<size_t>(csr_cap);               // gather entries (u16)
  b += 2u * static_cast<size_t>(2 * mass_cap + 2);      // row bounds
  b += 4u * static_cast<size_t>(n_mat);                 // actuation table
  return b;
}

size_t lds_bytes_mass(int block, int n_mat) {
  // the small area, two position buffers, two {L, k} buffers, {1/k, rest
  // length} and the pair constants
  (void)n_mat;
  return static_cast<size_t>(mass_pair_base(block, 2)) + 3u * kMassDynBytes + kMassCstBytes;
}

size_t lds_bytes_massg(int block, int maxp) {
  // tables, one position buffer, 12 B of terms per (lane, record)
  return static_cast<size_t>(massg_term_base(block)) + 12u * static_cast<size_t>(maxp) * block;
}

size_t lds_bytes_chunk(int mass_cap, int spring_cap, int csr_cap, int n_mat, int tw) {
  size_t b = lds_term_bytes(spring_cap, tw);                      // spring terms, sink, zero
  b += 16u * static_cast<size_t>(pos_slots(mass_cap));            // float4 positions (skewed)
  b += 2u * static_cast<size_t>(csr_cap);                         // gather entries (u16)
  b += 8u * static_cast<size_t>(chunk_row_quads(mass_cap));       // row bounds
  b += 8u * static_cast<size_t>(n_mat);                           // actuation, stiffness
  b += 4u * static_cast<size_t>(kChunkMax + 1);                   // chunk entry offsets
  return b;
}

size_t lds_bytes_mstream(int block, int mpt) {
  return kMStreamMatBytes + 2u * 16u * static_cast<size_t>(mass_slots(block * mpt));
}

size_t lds_bytes_stream(int mass_cap, int n_mat) {
  return 16u * static_cast<size_t>(mass_cap) + 4u * static_cast<size_t>(n_mat);
}

hipError_t launch_step(const KernelVariant& v, int n_packs, size_t lds_bytes, hipStream_t stream,
                       const StepParams& p) {
  StepParams args = p;
  void* kargs[] = {&args};
  const void* fn = (p.track_stress && v.fn_stress) ? v.fn_stress : v.fn;
  // the mass-parallel streaming kernel's LDS is a static array
  const size_t dyn = (v.kind == MH_KERNEL_STREAM && v.chunked == 2) ? 0 : lds_bytes;
  return hipLaunchKernel(fn, dim3(n_packs), dim3(v.block), kargs, dyn, stream);
}

hipError_t launch_clock(float* times, float t0, float dt, int n_steps, int K, hipStream_t stream) {
  hipLaunchKernelGGL(set_clock, dim3(1), dim3(64), 0, stream, times, t0, dt, n_steps, K);
  return hipGetLastError();
}

hipError_t launch_com(const float4* pos, const uint8_t* mass_id, const uint32_t* robot_mass_begin,
                      int n_robots, float* com_out, hipStream_t stream) {
  const int threads = 256;
  const int blocks = (n_robots + threads - 1) / threads;
  hipLaunchKernelGGL(robot_com, dim3(blocks), dim3(threads), 0, stream, pos, mass_id,
                     robot_mass_begin, n_robots, com_out);
  return hipGetLastError();
}

}  // namespace mh
