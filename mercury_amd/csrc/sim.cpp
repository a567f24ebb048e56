// sim.cpp -- the simulator handle: population packing, launch planning,
// step loop (optionally captured in a hipGraph), collect and fitness.
//
// Reference: common/simulator/Simulator.cu (SetElements 206-367,
// Simulate 369-458, AllocateElement 460-529, Collect 537-595).  Differences
// by design: robots may have different mass and spring counts (per-robot
// offsets instead of a uniform stride), air springs are dropped from the
// device arrays (the reference skips them in the kernel, sim_kernel.cu:211),
// all state stays on the device between calls, errors are return codes.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <tuple>

#include "mh_internal.h"
#include "mh_kernels.h"

namespace mh {
namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(MH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// fn(i) for i in [0, n) on up to 16 host threads; items are independent.
// The first exception thrown by any item is rethrown after all threads end.
template <typename F>
void parallel_for(size_t n, F&& fn) {
  size_t nt = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
  nt = std::min(nt, n / 8 + 1);  // small batches stay on the calling thread
  if (nt <= 1) {
    for (size_t i = 0; i < n; i++) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::exception_ptr err;
  std::mutex mu;
  auto work = [&]() {
    for (size_t i = next++; i < n; i = next++) {
      try {
        fn(i);
      } catch (...) {
        std::lock_guard<std::mutex> lock(mu);
        if (!err) err = std::current_exception();
      }
    }
  };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; t++) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hip_check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)), "hipMalloc");
    n = count;
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    alloc(count);
    if (count)
      hip_check(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s), "H2D copy");
  }
};

// One kernel's share of the population.
struct Part {
  const KernelVariant* v = nullptr;
  std::vector<PackDesc> packs;      // launch order (largest work first)
  int mass_cap = 0, spring_cap = 0, csr_cap = 0;
  size_t lds = 0;
  uint32_t pack_offset = 0;         // first descriptor in the device array
  int n_robots = 0;
};

// Robots that fit the LDS kernel go to the LDS part, the rest to the
// streaming part; each part is one launch per group of substeps.
struct Plan {
  std::vector<Part> parts;
  int rpb = 0;  // robots-per-block cap of the LDS part (0 = as many as fit)
};

constexpr int kMaxLdsBytes = 160 * 1024;

}  // namespace
}  // namespace mh

struct mh_sim {
  mh_config cfg{};
  std::vector<mh::Material> mats_by_id;
  int composite_count = 0;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;         // second part of a mixed plan
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int lds_limit = mh::kMaxLdsBytes;

  // population (host copies used for packing and collect)
  int n_robots = 0;
  std::vector<uint32_t> mass_off, spring_off, active_off;  // per robot prefix offsets
  std::vector<uint32_t> gather_pad;  // per robot: sum over masses of the degree rounded up to 4
  std::vector<uint8_t> max_side;     // per robot: most springs one mass has on one side (capped 255)
  std::vector<uint16_t> max_deg;     // per robot: most springs of one mass (capped 65535)
  // per robot: distinct (rest length, material) pairs of its springs as
  // rest-length bits << 8 | compact material, sorted; empty when there are
  // more than the mass-parallel kernels' pair table holds
  std::vector<std::vector<uint64_t>> pair_set;
  std::vector<uint16_t> act_pairs;   // active springs, robot-local
  std::vector<uint32_t> active_src;  // active spring -> original spring index
  // stress held by air springs (original order; empty = all zero): the step
  // never touches them, so they keep what mh_sim_set_stress gave them
  std::vector<float> held_stress;
  std::vector<uint8_t> act_mat;      // compact material index
  std::vector<float> act_lbar;       // mean length per active spring
  std::vector<uint8_t> compact_to_id;
  std::vector<float> base_com, base_len;
  size_t n_masses = 0, n_springs = 0, n_active = 0;
  bool has_population = false;
  bool fast_div_mats = false;   // xpbd divisors and amplitudes in the fast division's range
  bool fast_div_lbar = false;   // active rest lengths in its range

  // device state
  mh::DevBuf<float4> pos, vel;
  mh::DevBuf<float> lbar, stress, scratch, com;
  mh::DevBuf<uint8_t> smat, mass_id;
  mh::DevBuf<uint16_t> mperm, ccsr, crow;
  mh::DevBuf<uint32_t> mrec, ctab;
  mh::DevBuf<uint8_t> mcnt;
  mh::DevBuf<float> lbtab;
  mh::DevBuf<uint32_t> spair, csr, row, robot_begin;
  mh::DevBuf<mh::PackDesc> packs;
  mh::DevBuf<float> mat_k, mat_kx, mat_dL0, mat_omega, mat_phi;

  // decode table by encoded index, for populations built on the device:
  // composite id, the material's encoding, 1 unless every parameter is 0
  std::vector<uint8_t> code_id, code_solid;
  std::vector<uint32_t> code_enc;
  uint32_t valid_bits = 0;  // encoding bits of the material set
  // a voxel population built on the device (mh_sim_set_voxel_population):
  // stress is kept per spring in the robots' own order, air springs
  // included; the shape template of the last lattice shape stays resident
  bool voxel_device = false;
  int vt_nx = 0, vt_ny = 0, vt_nz = 0, vt_n = 0, vt_springs = 0, vt_lengths = 0;
  uint32_t vt_res_bits = 0;
  std::vector<float> vt_pos;  // host copy of the shape's positions (baselines)
  mh::DevBuf<float> vt_dpos, vt_dlen;
  mh::DevBuf<uint16_t> vt_pairs, vt_vox, vt_foff, vt_flist, vt_soff, vt_slist;
  mh::DevBuf<uint8_t> vt_lidx, vt_code_id, vt_code_solid;
  mh::DevBuf<uint32_t> vt_code_enc, vox_enc, vox_stripped, vox_active;
  mh::DevBuf<uint8_t> vox_spring_id;
  mh::DevBuf<int> vox_error;

  mh::Plan plan;
  int substeps = 0;  // substeps per launch; 0 = every substep of a call in one launch
  float total_time = 0.f;
  // start time of every launch of the current call, written on the device
  // by the clock kernel ahead of the step launches (so a captured graph
  // replays for any starting time)
  mh::DevBuf<float> clock;

  // graph cache: one instantiated graph per (steps, stress, substeps, plan)
  hipGraphExec_t graph = nullptr;
  int g_steps = -1, g_track = -1, g_sub = -1;
  const void* g_key = nullptr;
  int graph_builds = 0;

  // timing
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int last_launches = 0;
  bool have_timing = false;

  ~mh_sim() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (aux) (void)hipStreamDestroy(aux);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace mh {
namespace {

void drop_graph(mh_sim* s) {
  if (s->graph) (void)hipGraphExecDestroy(s->graph);
  s->graph = nullptr;
  s->g_steps = -1;
}

// Material tables for the compact indices (xpbd stiffness term depends on dt).
void upload_materials(mh_sim* s) {
  const int n = static_cast<int>(s->compact_to_id.size());
  std::vector<float> k(n), kx(n), a(n), w(n), ph(n);
  const float dt = s->cfg.dt;
  for (int i = 0; i < n; i++) {
    const Material& m = s->mats_by_id[s->compact_to_id[i]];
    k[i] = m.k;
    float alpha = 1.0f / m.k;  // sim_kernel.cu:220, ((1/k)/dt)/dt
    alpha = alpha / dt;
    alpha = alpha / dt;
    kx[i] = 2.0f + alpha;      // sim_kernel.cu:225
    a[i] = m.dL0;
    w[i] = m.omega;
    ph[i] = m.phi;
  }
  // the gather kernel's fast xpbd division needs divisors in [2^-40, 2^40]
  // and amplitudes of at most 1/2
  s->fast_div_mats = true;
  for (int i = 0; i < n; i++)
    s->fast_div_mats = s->fast_div_mats && kx[i] >= 0x1p-40f && kx[i] <= 0x1p40f && std::fabs(a[i]) <= 0.5f;
  s->mat_k.upload(k.data(), n, s->stream);
  s->mat_kx.upload(kx.data(), n, s->stream);
  s->mat_dL0.upload(a.data(), n, s->stream);
  s->mat_omega.upload(w.data(), n, s->stream);
  s->mat_phi.upload(ph.data(), n, s->stream);
}

int robot_masses(const mh_sim* s, int r) { return static_cast<int>(s->mass_off[r + 1] - s->mass_off[r]); }
int robot_active(const mh_sim* s, int r) { return static_cast<int>(s->active_off[r + 1] - s->active_off[r]); }

bool fits_lds_alone(const mh_sim* s, const KernelVariant& v, int r, int n_mat) {
  const int m = robot_masses(s, r), sp = robot_active(s, r), g = static_cast<int>(s->gather_pad[r]);
  const int tw = solver_term_words(s->cfg.solver);
  return m <= v.block * v.mpt && sp <= v.block * v.nspt && 2 * tw * (sp + 1) + 1 <= 65535 && g <= 65535 &&
         lds_bytes_lds(std::max(m, 1), std::max(sp, 1), g, n_mat, tw) <= static_cast<size_t>(s->lds_limit);
}

bool fits_mass(const mh_sim* s, const KernelVariant& v, int r, int n_mat) {
  const int m = robot_masses(s, r);
  // a robot with springs needs its pairs in the table (an empty set marks
  // one with too many)
  const bool table_ok = robot_active(s, r) == 0 || !s->pair_set[r].empty();
  const size_t lds = v.chunked ? lds_bytes_massg(v.block, v.nspt) : lds_bytes_mass(v.block, n_mat);
  return m <= v.block && s->max_side[r] <= v.nspt && table_ok && lds <= static_cast<size_t>(s->lds_limit);
}

// MH_KERNEL_* family bit of a variant (the gather flavour of the
// mass-parallel kernel is a family of its own).
int family(const KernelVariant& v) {
  if (v.kind == MH_KERNEL_MASS && v.chunked) return MH_KERNEL_MASS_GATHER;
  if (v.kind == MH_KERNEL_STREAM && v.chunked == 2) return MH_KERNEL_MASS_STREAM;
  return v.kind;
}
constexpr int kResident = MH_KERNEL_LDS | MH_KERNEL_MASS | MH_KERNEL_MASS_GATHER;
constexpr int kStreaming = MH_KERNEL_STREAM | MH_KERNEL_MASS_STREAM;
constexpr int kAllFamilies = kResident | kStreaming;

bool fits_resident(const mh_sim* s, const KernelVariant& v, int r, int n_mat) {
  return v.kind == MH_KERNEL_MASS ? fits_mass(s, v, r, n_mat) : fits_lds_alone(s, v, r, n_mat);
}

// The mass-parallel streaming kernel: the robot's masses fill at most
// block * mpt virtual lanes, no mass has more than 255 springs (u8
// counts), the materials and the null one fit its table.
bool fits_mstream(const mh_sim* s, const KernelVariant& v, int r, int n_mat) {
  const bool compact = v.nspt == 6;
  if (compact && (n_mat + 1 > kMStream6Mats || 16 * mass_slots(v.block * v.mpt) > 65536)) return false;
  return robot_masses(s, r) <= v.block * v.mpt && s->max_deg[r] <= 255 && n_mat + 1 <= kMStreamMats &&
         lds_bytes_mstream(v.block, v.mpt) <= static_cast<size_t>(s->lds_limit);
}

// Gather-entry capacity of a chunk of C springs for a robot of m masses: 2C
// entries plus padding of up to 3 per row, in whole quads.
int chunk_csr_cap(int C, int m) { return (2 * C + 3 * m + 3) & ~3; }

// Springs per chunk for the chunked kernel and a robot of m masses: the
// largest candidate whose LDS layout (chunk terms, skewed positions, one
// chunk's gather entries -- at most 2C plus padding of 3 per row part --
// and row bounds) fits; 0 if none does.
int chunk_springs(const mh_sim* s, const KernelVariant& v, int m, int n_mat) {
  if (m > v.block * v.mpt || chunk_row_quads(m) > kChunkRpl * v.block) return 0;
  for (int C : {8192, 6144, 5120, 4096, 3072, 2048, 1536, 1024, 512}) {
    const int entries = chunk_csr_cap(C, m);
    if (entries > 65535 || 2 * solver_term_words(s->cfg.solver) * (C + 1) + 1 > 65535) continue;
    // what one chunk puts in flight per lane
    if (C > kChunkSpl * v.block || entries / 4 > kChunkEpl * v.block) continue;
    if (lds_bytes_chunk(std::max(m, 1), C, entries, n_mat, solver_term_words(s->cfg.solver)) <=
        static_cast<size_t>(s->lds_limit))
      return C;
  }
  return 0;
}

// Greedy packing of consecutive LDS-eligible robots into workgroups of the
// LDS variant (at most rpb robots each); every other robot gets a
// workgroup of the streaming variant.  Returns false if a robot fits
// neither kernel.
bool make_plan(const mh_sim* s, const KernelVariant* lv, const KernelVariant* sv, int rpb, Plan& plan) {
  plan = Plan{};
  plan.rpb = rpb;
  const int n_mat = static_cast<int>(s->compact_to_id.size());
  const int tw = solver_term_words(s->cfg.solver);
  const int cap = rpb > 0 ? rpb : std::numeric_limits<int>::max();
  Part lp, stp;
  lp.v = lv;
  stp.v = sv;
  int r = 0;
  std::vector<uint64_t> pair_union, merged;
  while (r < s->n_robots) {
    if (!lv || !fits_resident(s, *lv, r, n_mat)) {
      const int m = robot_masses(s, r);
      if (!sv) return false;
      if (sv->chunked == 2 ? !fits_mstream(s, *sv, r, n_mat)
          : sv->chunked    ? chunk_springs(s, *sv, m, n_mat) == 0
                           : lds_bytes_stream(std::max(m, 1), n_mat) > static_cast<size_t>(s->lds_limit))
        return false;
      PackDesc d{};
      d.mass_begin = s->mass_off[r];
      d.n_mass = static_cast<uint32_t>(m);
      d.spring_begin = s->active_off[r];
      d.n_spring = static_cast<uint32_t>(robot_active(s, r));
      stp.mass_cap = std::max(stp.mass_cap, m);
      stp.spring_cap = std::max(stp.spring_cap, static_cast<int>(d.n_spring));
      stp.packs.push_back(d);
      stp.n_robots++;
      r++;
      continue;
    }
    if (lv->kind == MH_KERNEL_MASS) {
      // consecutive robots while their masses fill one workgroup and their
      // (rest length, material) pairs share one table
      PackDesc d{};
      d.mass_begin = s->mass_off[r];
      d.spring_begin = s->active_off[r];
      pair_union.clear();
      int count = 0;
      while (r < s->n_robots && count < cap && fits_mass(s, *lv, r, n_mat)) {
        const int m = static_cast<int>(d.n_mass) + robot_masses(s, r);
        merged.clear();
        std::set_union(pair_union.begin(), pair_union.end(), s->pair_set[r].begin(), s->pair_set[r].end(),
                       std::back_inserter(merged));
        if (m > lv->block || static_cast<int>(merged.size()) > kMassPairNull) break;
        pair_union.swap(merged);
        d.n_mass = static_cast<uint32_t>(m);
        d.n_spring += static_cast<uint32_t>(robot_active(s, r));
        r++;
        count++;
      }
      lp.mass_cap = std::max(lp.mass_cap, static_cast<int>(d.n_mass));
      lp.spring_cap = std::max(lp.spring_cap, static_cast<int>(d.n_spring));
      lp.packs.push_back(d);
      lp.n_robots += count;
      continue;
    }
    PackDesc d{};
    d.mass_begin = s->mass_off[r];
    d.spring_begin = s->active_off[r];
    int count = 0, gathers = 0;
    const int mcap = lv->block * lv->mpt, scap = lv->block * lv->nspt;
    while (r < s->n_robots && count < cap && fits_lds_alone(s, *lv, r, n_mat)) {
      const int m = static_cast<int>(d.n_mass) + robot_masses(s, r);
      const int sp = static_cast<int>(d.n_spring) + robot_active(s, r);
      const int g = gathers + static_cast<int>(s->gather_pad[r]);
      // gather entries hold 2 * (tw * slot) + 1 in 16 bits: slots up to
      // spring_cap + 1
      const bool fits = m <= mcap && sp <= scap && 2 * tw * (sp + 1) + 1 <= 65535 && g <= 65535 &&
                        lds_bytes_lds(std::max(lp.mass_cap, m), std::max(lp.spring_cap, sp),
                                      std::max(lp.csr_cap, g), n_mat, tw) <= static_cast<size_t>(s->lds_limit);
      if (!fits) break;
      d.n_mass = static_cast<uint32_t>(m);
      d.n_spring = static_cast<uint32_t>(sp);
      gathers = g;
      r++;
      count++;
    }
    lp.mass_cap = std::max(lp.mass_cap, static_cast<int>(d.n_mass));
    lp.spring_cap = std::max(lp.spring_cap, static_cast<int>(d.n_spring));
    lp.csr_cap = std::max(lp.csr_cap, gathers);
    lp.packs.push_back(d);
    lp.n_robots += count;
  }
  if (!lp.packs.empty()) {
    lp.lds = lv->kind == MH_KERNEL_MASS
                 ? (lv->chunked ? lds_bytes_massg(lv->block, lv->nspt) : lds_bytes_mass(lv->block, n_mat))
                 : lds_bytes_lds(std::max(lp.mass_cap, 1), std::max(lp.spring_cap, 1), lp.csr_cap, n_mat, tw);
    plan.parts.push_back(std::move(lp));
  }
  if (!stp.packs.empty()) {
    if (sv->chunked == 2) {
      stp.lds = lds_bytes_mstream(sv->block, sv->mpt);
    } else if (sv->chunked) {
      // one chunk size for the part: the one that fits its largest robot
      stp.spring_cap = chunk_springs(s, *sv, std::max(stp.mass_cap, 1), n_mat);
      stp.csr_cap = chunk_csr_cap(stp.spring_cap, std::max(stp.mass_cap, 1));
      stp.lds = lds_bytes_chunk(std::max(stp.mass_cap, 1), stp.spring_cap, stp.csr_cap, n_mat, tw);
      for (const PackDesc& d : stp.packs)
        if ((d.n_spring + stp.spring_cap - 1) / stp.spring_cap > static_cast<uint32_t>(kChunkMax)) return false;
    } else {
      stp.lds = lds_bytes_stream(std::max(stp.mass_cap, 1), n_mat);
    }
    plan.parts.push_back(std::move(stp));
  }
  return true;
}

// Chunked part (one robot per pack): skewed spring pairs, and per chunk of
// part.spring_cap springs the u16 gather entries 3 * (spring - chunk start)
// and 2 n_mass + 1 chunk-relative row bounds (padded to whole quads), each
// row part in ascending spring index padded to 4 with the zero slot; per
// pack (d.pad0) the entry offset of every chunk and the end.
void upload_chunked(mh_sim* s, Part& part, std::vector<uint32_t>& spair, std::vector<uint16_t>& ccsr,
                    std::vector<uint16_t>& crow, std::vector<uint32_t>& ctab) {
  const uint32_t C = static_cast<uint32_t>(part.spring_cap);
  const uint32_t tw2 = 2u * static_cast<uint32_t>(solver_term_words(s->cfg.solver));
  const uint16_t zero_entry = static_cast<uint16_t>(tw2 * (C + 1));
  auto pad = [](uint32_t n) { return (n + 3u) & ~3u; };
  std::vector<uint32_t> lpair, count, fill_plus;
  for (PackDesc& d : part.packs) {
    d.csr_begin = static_cast<uint32_t>(ccsr.size());
    d.row_begin = static_cast<uint32_t>(crow.size());
    const uint32_t nm = d.n_mass, ns = d.n_spring;
    lpair.resize(2 * static_cast<size_t>(ns));
    for (uint32_t k = 0; k < ns; k++) {
      const uint32_t g = d.spring_begin + k;
      const uint32_t a = s->act_pairs[2 * g], b = s->act_pairs[2 * g + 1];
      lpair[2 * k] = a;
      lpair[2 * k + 1] = b;
      spair[g] = static_cast<uint32_t>(pos_slot(static_cast<int>(a))) |
                 (static_cast<uint32_t>(pos_slot(static_cast<int>(b))) << 16);
    }
    const uint32_t n_chunks = ns > 0 ? (ns + C - 1) / C : 1;
    d.pad0 = static_cast<uint32_t>(ctab.size());
    d.pad1 = n_chunks;
    for (uint32_t c = 0; c < n_chunks; c++) {
      ctab.push_back(static_cast<uint32_t>(ccsr.size()) - d.csr_begin);
      const uint32_t s0 = c * C, s1 = std::min(ns, s0 + C);
      // one row per mass: the chunk's springs of the mass in ascending
      // spring index, entry 2 * (3 * (spring - s0)), + 1 for the second
      // endpoint, padded to a quad with the zero slot
      count.assign(nm + 1, 0);
      for (uint32_t k = s0; k < s1; k++) {
        count[lpair[2 * k] + 1]++;
        count[lpair[2 * k + 1] + 1]++;
      }
      for (uint32_t i = 0; i < nm; i++) count[i + 1] = count[i] + pad(count[i + 1]);
      if (count[nm] > static_cast<uint32_t>(part.csr_cap))
        throw Error(MH_ERR_INTERNAL, "chunk gather list over its LDS capacity");
      const size_t base = ccsr.size();
      ccsr.resize(base + count[nm], zero_entry);
      fill_plus.assign(count.begin(), count.end() - 1);
      for (uint32_t k = s0; k < s1; k++) {
        const uint16_t e = static_cast<uint16_t>(tw2 * (k - s0));
        ccsr[base + fill_plus[lpair[2 * k]]++] = e;
        ccsr[base + fill_plus[lpair[2 * k + 1]]++] = static_cast<uint16_t>(e + 1u);
      }
      for (uint32_t i = 0; i <= nm; i++) crow.push_back(static_cast<uint16_t>(count[i]));
      for (uint32_t i = nm + 1; i < 4u * static_cast<uint32_t>(chunk_row_quads(static_cast<int>(nm))); i++)
        crow.push_back(static_cast<uint16_t>(count[nm]));
    }
    ctab.push_back(static_cast<uint32_t>(ccsr.size()) - d.csr_begin);
  }
  std::stable_sort(part.packs.begin(), part.packs.end(), [](const PackDesc& x, const PackDesc& y) {
    return static_cast<uint64_t>(x.n_spring) + x.n_mass > static_cast<uint64_t>(y.n_spring) + y.n_mass;
  });
}

// Mass-parallel streaming part (one robot per pack): its masses in
// virtual-lane order, most springs first (ties by index), into perm; per
// virtual lane its spring count (u8, mcnt at d.row_begin); per group of 64
// virtual lanes the offset of its records, in records from d.csr_begin
// (u32, row at d.pad0); the records of group g as [k][64] for k below the
// group's longest list: per spring of the mass in ascending spring index
// {16 * slot of the other endpoint | 8 * compact material << 16, rest
// length bits}, padded with {16 * ghost | 8 * n_mat << 16, 1.0f}.  mrec
// holds two words per record; d.csr_begin counts records.
void upload_mstream(mh_sim* s, Part& part, std::vector<uint16_t>& perm, std::vector<uint32_t>& mrec,
                    std::vector<uint8_t>& mcnt, std::vector<uint32_t>& row) {
  const int n_mat = static_cast<int>(s->compact_to_id.size());
  const bool compact = part.v->nspt == 6;
  const uint32_t ghost = static_cast<uint32_t>(mass_slots(part.v->block * part.v->mpt) - 1);
  const uint32_t pad_w0 = 16u * ghost | (8u * static_cast<uint32_t>(n_mat)) << 16;
  const uint16_t pad_code = static_cast<uint16_t>(16u * ghost | static_cast<uint32_t>(n_mat));
  std::vector<uint32_t> deg, order, lane, fill, start;
  // records are 8-byte aligned, the compact format's rest-length quads 16-byte
  while (mrec.size() % (compact ? 4u : 2u)) mrec.push_back(0u);
  for (PackDesc& d : part.packs) {
    const uint32_t nm = d.n_mass, ns = d.n_spring;  // one robot: pack-local = robot-local
    deg.assign(nm, 0);
    for (uint32_t k = 0; k < ns; k++) {
      const uint32_t g = d.spring_begin + k;
      deg[s->act_pairs[2 * g]]++;
      deg[s->act_pairs[2 * g + 1]]++;
    }
    order.resize(nm);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return deg[x] > deg[y]; });
    lane.assign(nm, 0);
    for (uint32_t v = 0; v < nm; v++) {
      lane[order[v]] = v;
      perm[d.mass_begin + v] = static_cast<uint16_t>(order[v]);
    }
    d.row_begin = static_cast<uint32_t>(mcnt.size());
    for (uint32_t v = 0; v < nm; v++) mcnt.push_back(static_cast<uint8_t>(deg[order[v]]));
    d.csr_begin = static_cast<uint32_t>(mrec.size() / 2);
    d.pad0 = static_cast<uint32_t>(row.size());
    const uint32_t groups = (nm + 63) / 64;
    // per group, in 8-byte units: 64 records per k ([k][64]), or in the
    // compact format 64 code quads then 64 rest-length quads (2 units
    // each) per quad of k
    start.assign(groups + 1, 0);
    std::vector<uint32_t> quads(groups);
    for (uint32_t gi = 0; gi < groups; gi++) {
      uint32_t longest = 0;
      for (uint32_t v = 64 * gi; v < std::min(nm, 64 * gi + 64); v++) longest = std::max(longest, deg[order[v]]);
      row.push_back(start[gi]);
      quads[gi] = (longest + 3) / 4;
      start[gi + 1] = start[gi] + (compact ? 192 * quads[gi] : 64 * longest);
    }
    const size_t base = mrec.size();
    mrec.resize(base + 2 * static_cast<size_t>(start[groups]));
    if (compact) {
      for (uint32_t gi = 0; gi < groups; gi++) {
        uint32_t* w = mrec.data() + base + 2 * static_cast<size_t>(start[gi]);
        const size_t nq = 64u * quads[gi];
        std::fill(w, w + 2 * nq, static_cast<uint32_t>(pad_code) * 0x10001u);
        std::fill(w + 2 * nq, w + 6 * nq, 0x3f800000u);
      }
    } else {
      for (size_t i = base; i < mrec.size(); i += 2) {
        mrec[i] = pad_w0;
        mrec[i + 1] = 0x3f800000u;
      }
    }
    fill.assign(nm, 0);
    for (uint32_t k = 0; k < ns; k++) {  // ascending spring index
      const uint32_t g = d.spring_begin + k;
      const uint32_t a = s->act_pairs[2 * g], b = s->act_pairs[2 * g + 1];
      uint32_t lb;
      std::memcpy(&lb, &s->act_lbar[g], 4);
      const uint32_t m8 = (8u * s->act_mat[g]) << 16;
      for (int e = 0; e < 2; e++) {
        const uint32_t own = e ? b : a, other = e ? a : b;
        const uint32_t v = lane[own];
        const uint32_t kk = fill[own]++;
        const uint32_t oslot = 16u * static_cast<uint32_t>(mass_slot(static_cast<int>(other)));
        if (compact) {
          // code quad (kk / 4, lane) holds 4 u16 codes, rest-length quad 4 floats
          uint32_t* w = mrec.data() + base + 2 * static_cast<size_t>(start[v / 64]);
          const size_t q = 64u * (kk / 4) + (v & 63u);
          uint32_t& cw = w[2 * q + ((kk & 3u) >> 1)];
          const uint32_t sh = 16u * (kk & 1u);
          cw = (cw & ~(0xffffu << sh)) | (oslot | s->act_mat[g]) << sh;
          w[2 * 64u * quads[v / 64] + 4 * q + (kk & 3u)] = lb;
        } else {
          const size_t at = base + 2 * (static_cast<size_t>(start[v / 64]) + 64u * kk + (v & 63u));
          mrec[at] = oslot | m8;
          mrec[at + 1] = lb;
        }
      }
    }
  }
  std::stable_sort(part.packs.begin(), part.packs.end(), [](const PackDesc& x, const PackDesc& y) {
    return static_cast<uint64_t>(x.n_spring) + x.n_mass > static_cast<uint64_t>(y.n_spring) + y.n_mass;
  });
}

// Mass-parallel part: per pack its (rest length, material) pairs, two
// words each (the rest length, the compact material as u32 bits; d.pad0 =
// offset, d.pad1 = pair count), the lane order of its masses (most records
// first, in perm), and per lane its mass's first-endpoint and
// second-endpoint records in ascending spring index, laid out
// [part][k][lane] (part 2: pack-local spring index of the part-0 records),
// with the two counts per lane.  A record holds LDS byte offsets: the other
// endpoint's skewed slot * 16 | pair index * 8 << 16.  For the gather
// kernel a second-endpoint record is instead the LDS offset of the term
// slot its first endpoint fills.
void upload_mass(mh_sim* s, Part& part, std::vector<uint16_t>& perm, std::vector<uint32_t>& mrec,
                 std::vector<uint8_t>& mcnt, std::vector<float>& lbtab) {
  const int P = part.v->nspt;
  const bool gather = part.v->chunked != 0;
  // record words per lane: step_mass 2P records, the first-endpoint mask
  // and 2P spring indices; step_massg P records, 2P entries and P spring
  // indices
  const int parts = gather ? 4 * P : 4 * P + 1;
  const size_t n_packs = part.packs.size();
  auto first_robot = [s](const PackDesc& d) {
    return static_cast<int>(std::upper_bound(s->mass_off.begin(), s->mass_off.end(), d.mass_begin) -
                            s->mass_off.begin()) - 1;
  };
  // pass 1 (packs in parallel): the pack's pairs, the union of its robots'
  // sets, sorted
  std::vector<std::vector<uint64_t>> tables(n_packs);
  parallel_for(n_packs, [&](size_t pi) {
    const PackDesc& d = part.packs[pi];
    std::vector<uint64_t>& table = tables[pi];
    std::vector<uint64_t> merged;
    for (int q = first_robot(d); q < s->n_robots && s->mass_off[q] < d.mass_begin + d.n_mass; q++) {
      merged.clear();
      std::set_union(table.begin(), table.end(), s->pair_set[q].begin(), s->pair_set[q].end(),
                     std::back_inserter(merged));
      table.swap(merged);
    }
    if (table.size() > static_cast<size_t>(kMassPairNull))
      throw Error(MH_ERR_INTERNAL, "mass pack pair table over the kernel's capacity");
  });
  // offsets of every pack's pairs, records and counts
  size_t lb_end = lbtab.size(), rec_end = mrec.size(), cnt_end = mcnt.size();
  for (size_t pi = 0; pi < n_packs; pi++) {
    PackDesc& d = part.packs[pi];
    d.pad0 = static_cast<uint32_t>(lb_end);
    d.pad1 = static_cast<uint32_t>(tables[pi].size());
    lb_end += 2 * tables[pi].size();
    d.csr_begin = static_cast<uint32_t>(rec_end);
    rec_end += static_cast<size_t>(parts) * d.n_mass;
    d.row_begin = static_cast<uint32_t>(cnt_end);
    cnt_end += 2u * d.n_mass;
  }
  lbtab.resize(lb_end);
  mrec.resize(rec_end, 0u);
  mcnt.resize(cnt_end);
  // pass 2 (packs in parallel): pairs, lane order and records
  parallel_for(n_packs, [&](size_t pi) {
    const PackDesc& d = part.packs[pi];
    const std::vector<uint64_t>& table = tables[pi];
    const uint32_t nm = d.n_mass, ns = d.n_spring;
    const int r0 = first_robot(d);
    float* tab = lbtab.data() + d.pad0;
    for (size_t t = 0; t < table.size(); t++) {
      const uint32_t lb_bits = static_cast<uint32_t>(table[t] >> 8), mat = static_cast<uint32_t>(table[t] & 0xffu);
      std::memcpy(&tab[2 * t], &lb_bits, 4);
      std::memcpy(&tab[2 * t + 1], &mat, 4);
    }
    uint32_t* out = mrec.data() + d.csr_begin;
    // record counts per mass, then the lane order: most records first
    // (step_massg: most first-endpoint records, then most entries; ties
    // keep order)
    std::vector<uint32_t> fill_p(nm, 0), fill_m(nm, 0), order(nm), lane(nm), first(nm, 0);
    int q = r0;
    for (uint32_t k = 0; k < ns; k++) {
      const uint32_t g = d.spring_begin + k;
      while (g >= s->active_off[q + 1]) q++;
      const uint32_t off = s->mass_off[q] - d.mass_begin;
      fill_p[s->act_pairs[2 * g] + off]++;
      fill_m[s->act_pairs[2 * g + 1] + off]++;
    }
    for (uint32_t i = 0; i < nm; i++) {
      if (fill_p[i] > static_cast<uint32_t>(P) || fill_m[i] > static_cast<uint32_t>(P))
        throw Error(MH_ERR_INTERNAL, "mass record lists over the kernel's capacity");
    }
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
      const uint32_t tx = fill_p[x] + fill_m[x], ty = fill_p[y] + fill_m[y];
      if (gather && fill_p[x] != fill_p[y]) return fill_p[x] > fill_p[y];
      return tx > ty;
    });
    for (uint32_t i = 0; i < nm; i++) {
      perm[d.mass_begin + i] = static_cast<uint16_t>(order[i]);
      lane[order[i]] = i;
    }
    // table position of a pair by value
    auto index_of = [&table](uint64_t v) {
      return static_cast<uint32_t>(std::lower_bound(table.begin(), table.end(), v) - table.begin());
    };
    std::vector<uint32_t> fill(nm, 0);  // entries (records, for step_mass) so far per mass
    std::fill(fill_p.begin(), fill_p.end(), 0u);
    q = r0;
    uint64_t last_key = ~0ull;
    uint32_t last_pi = 0;
    for (uint32_t k = 0; k < ns; k++) {  // ascending spring index
      const uint32_t g = d.spring_begin + k;
      while (g >= s->active_off[q + 1]) q++;
      const uint32_t off = s->mass_off[q] - d.mass_begin;
      const uint32_t a = s->act_pairs[2 * g] + off, b = s->act_pairs[2 * g + 1] + off;
      const uint32_t la = lane[a], lb = lane[b];
      uint32_t bits;
      std::memcpy(&bits, &s->act_lbar[g], 4);
      const uint64_t key = static_cast<uint64_t>(bits) << 8 | s->act_mat[g];
      if (key != last_key) {
        last_key = key;
        last_pi = index_of(key);
      }
      // the pair's byte offset in the kernel's table: float2 entries
      // (step_mass), float4 (step_massg)
      const uint32_t tail = ((gather ? 16u : 8u) * last_pi) << 16;
      const uint32_t rec_a = 16u * static_cast<uint32_t>(mass_slot(static_cast<int>(b))) | tail;
      if (!gather) {
        // both endpoints hold a record, in ascending spring index; the
        // first endpoint's carries the stress
        const uint32_t ka = fill[a]++;
        out[static_cast<size_t>(ka) * nm + la] = rec_a;
        out[static_cast<size_t>(2 * P + 1 + ka) * nm + la] = k;
        first[a] |= 1u << ka;
        const uint32_t kb = fill[b]++;
        out[static_cast<size_t>(kb) * nm + lb] = 16u * static_cast<uint32_t>(mass_slot(static_cast<int>(a))) | tail;
        out[static_cast<size_t>(2 * P + 1 + kb) * nm + lb] = k;
      } else {
        // the first endpoint computes the term into its slot ka; both
        // endpoints list the slot, the second with bit 0 set
        const uint32_t ka = fill_p[a]++;
        out[static_cast<size_t>(ka) * nm + la] = rec_a;
        out[static_cast<size_t>(3 * P + ka) * nm + la] = k;
        const uint32_t slot = massg_term_offset(part.v->block, P, la, ka);
        out[static_cast<size_t>(P + fill[a]++) * nm + la] = slot;
        out[static_cast<size_t>(P + fill[b]++) * nm + lb] = slot | 1u;
      }
    }
    if (!gather)
      for (uint32_t i = 0; i < nm; i++) out[static_cast<size_t>(2 * P) * nm + lane[i]] = first[i];
    uint8_t* cnt = mcnt.data() + d.row_begin;
    for (uint32_t i = 0; i < nm; i++) {
      cnt[2 * i] = static_cast<uint8_t>(gather ? fill_p[order[i]] : fill[order[i]]);
      cnt[2 * i + 1] = static_cast<uint8_t>(gather ? fill[order[i]] : 0u);
    }
  });
  std::stable_sort(part.packs.begin(), part.packs.end(), [](const PackDesc& x, const PackDesc& y) {
    return static_cast<uint64_t>(x.n_spring) + x.n_mass > static_cast<uint64_t>(y.n_spring) + y.n_mass;
  });
}

// Device arrays that depend on the packing: pack-local spring pairs, the
// gather lists, row bounds and phase-2 order.  Mass i's row holds its
// springs in ascending spring index, row[i] and row[i+1] bound it; an entry
// is twice the term's offset, + 1 where the mass is the spring's second
// endpoint.  LDS parts store the offset as tw * slot (tw = 3 words per xpbd
// term, 4 per force term) and pad each row to a multiple of 4 with the zero
// slot; streaming parts store the slot.  The mass-parallel and chunked
// parts have layouts of their own (upload_mstream, upload_mass,
// upload_chunked).
void upload_plan(mh_sim* s, Plan& plan) {
  const size_t na = s->n_active;
  std::vector<uint32_t> spair(na), csr, row;
  csr.reserve(2 * na + 6 * s->n_masses);
  std::vector<uint16_t> perm(s->n_masses);
  std::vector<uint32_t> order, count, nplus, lpair;
  std::vector<PackDesc> all;
  row.reserve(2 * s->n_masses + 64);
  bool need_scratch = false;
  std::vector<uint16_t> ccsr, crow;
  std::vector<uint32_t> ctab;
  std::vector<uint32_t> mrec;
  std::vector<uint8_t> mcnt;
  std::vector<float> lbtab;
  for (Part& part : plan.parts) {
    const bool padded = part.v->kind == MH_KERNEL_LDS;
    need_scratch |= part.v->kind == MH_KERNEL_STREAM && !part.v->chunked;
    if (part.v->kind == MH_KERNEL_STREAM && part.v->chunked == 2) {
      upload_mstream(s, part, perm, mrec, mcnt, row);
    } else if (part.v->kind == MH_KERNEL_MASS) {
      upload_mass(s, part, perm, mrec, mcnt, lbtab);
    } else if (part.v->chunked == 1) {
      upload_chunked(s, part, spair, ccsr, crow, ctab);
    } else {
      // entries: 2 * (tw * slot) (LDS parts) or 2 * slot (streaming parts),
      // + 1 where the mass is the spring's second endpoint
      const uint32_t tw2 = 2u * static_cast<uint32_t>(solver_term_words(s->cfg.solver));
      const uint32_t zero_entry = tw2 * static_cast<uint32_t>(part.spring_cap + 1);
      const uint32_t scale = padded ? tw2 : 2u;
      auto pad = [padded](uint32_t n) { return padded ? (n + 3u) & ~3u : n; };
      for (PackDesc& d : part.packs) {
        d.csr_begin = static_cast<uint32_t>(csr.size());
        d.row_begin = static_cast<uint32_t>(row.size());
        const uint32_t nm = d.n_mass;
        nplus.assign(nm, 0);
        lpair.resize(2 * static_cast<size_t>(d.n_spring));
        // pack-local pairs: robot-local index + robot's offset inside the pack
        int r = static_cast<int>(std::upper_bound(s->mass_off.begin(), s->mass_off.end(), d.mass_begin) -
                                 s->mass_off.begin()) - 1;
        for (uint32_t k = 0; k < d.n_spring; k++) {
          const uint32_t g = d.spring_begin + k;
          while (g >= s->active_off[r + 1]) r++;
          const uint32_t off = s->mass_off[r] - d.mass_begin;
          const uint32_t a = s->act_pairs[2 * g] + off, b = s->act_pairs[2 * g + 1] + off;
          // LDS parts address positions by skewed slot
          spair[g] = padded ? static_cast<uint32_t>(pos_slot(static_cast<int>(a))) |
                                  (static_cast<uint32_t>(pos_slot(static_cast<int>(b))) << 16)
                            : a | (b << 16);
          lpair[2 * k] = a;
          lpair[2 * k + 1] = b;
          nplus[a]++;
          nplus[b]++;
        }
        // row bounds: mass i's springs at [row[i], row[i+1])
        count.assign(nm + 1, 0);
        for (uint32_t i = 0; i < nm; i++) count[i + 1] = count[i] + pad(nplus[i]);
        // phase-2 visiting order: longest rows first (ties keep order)
        order.resize(nm);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
          return count[x + 1] - count[x] > count[y + 1] - count[y];
        });
        for (uint32_t i = 0; i < nm; i++) perm[d.mass_begin + i] = static_cast<uint16_t>(order[i]);
        for (uint32_t i = 0; i <= nm; i++) row.push_back(count[i]);
        csr.resize(d.csr_begin + count[nm], zero_entry);
        std::vector<uint32_t> fill(count.begin(), count.end() - 1);
        for (uint32_t k = 0; k < d.n_spring; k++) {
          const uint32_t a = lpair[2 * k], b = lpair[2 * k + 1];
          csr[d.csr_begin + fill[a]++] = scale * k;
          csr[d.csr_begin + fill[b]++] = scale * k + 1u;
        }
      }
      // launch order: most work first, so uneven packs balance across CUs
      std::stable_sort(part.packs.begin(), part.packs.end(), [](const PackDesc& x, const PackDesc& y) {
        return static_cast<uint64_t>(x.n_spring) + x.n_mass > static_cast<uint64_t>(y.n_spring) + y.n_mass;
      });
    }
    part.pack_offset = static_cast<uint32_t>(all.size());
    all.insert(all.end(), part.packs.begin(), part.packs.end());
    // dynamic LDS above 64 KB needs the attribute (the mass-parallel
    // streaming kernel's LDS is static)
    if (part.lds > 64 * 1024 && !(part.v->kind == MH_KERNEL_STREAM && part.v->chunked == 2)) {
      for (const void* fn : {part.v->fn, part.v->fn_stress})
        if (fn)
          (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(part.lds));
    }
  }
  s->spair.upload(spair.data(), na, s->stream);
  s->csr.upload(csr.data(), csr.size(), s->stream);
  s->row.upload(row.data(), row.size(), s->stream);
  s->packs.upload(all.data(), all.size(), s->stream);
  s->mperm.upload(perm.data(), perm.size(), s->stream);
  s->ccsr.upload(ccsr.data(), ccsr.size(), s->stream);
  s->crow.upload(crow.data(), crow.size(), s->stream);
  s->ctab.upload(ctab.data(), ctab.size(), s->stream);
  s->mrec.upload(mrec.data(), mrec.size(), s->stream);
  s->mcnt.upload(mcnt.data(), mcnt.size(), s->stream);
  s->lbtab.upload(lbtab.data(), lbtab.size(), s->stream);
  if (need_scratch) s->scratch.alloc(4 * std::max<size_t>(na, 1));
  hip_check(hipStreamSynchronize(s->stream), "plan upload");
  s->plan = std::move(plan);
  drop_graph(s);
}

// Compiled variants of the given families (bit set) allowed by the config,
// in table order.
std::vector<const KernelVariant*> candidates(const mh_sim* s, int kinds) {
  int nv = 0;
  const KernelVariant* all = kernel_variants(&nv);
  const int allowed = s->cfg.kernel == MH_KERNEL_AUTO ? kAllFamilies : s->cfg.kernel;
  std::vector<const KernelVariant*> out;
  for (int i = 0; i < nv; i++) {
    const KernelVariant& v = all[i];
    if (v.solver != s->cfg.solver || !(family(v) & kinds & allowed)) continue;
    if (s->cfg.block_size && v.block != s->cfg.block_size) continue;
    out.push_back(&v);
  }
  return out;
}

// Default plan: the LDS variant (table order, largest blocks first) that
// takes the most robots, and for the rest the first streaming variant that
// fits them (table order: for the force solver the mass-parallel streaming
// kernel, then the chunked and global-memory ones).
void choose_default_plan(mh_sim* s) {
  auto lc = candidates(s, kResident);
  auto sc = candidates(s, kStreaming);
  const KernelVariant* sv = sc.empty() ? nullptr : sc.front();
  const int n_mat = static_cast<int>(s->compact_to_id.size());
  const KernelVariant* best = nullptr;
  int best_fit = -1;
  for (const KernelVariant* v : lc) {
    int fit = 0;
    for (int r = 0; r < s->n_robots; r++) fit += fits_resident(s, *v, r, n_mat) ? 1 : 0;
    if (fit > best_fit) {
      best_fit = fit;
      best = v;
    }
  }
  Plan p;
  bool ok = make_plan(s, best, sv, s->cfg.robots_per_block, p);
  for (size_t i = 1; !ok && i < sc.size(); i++) ok = make_plan(s, best, sc[i], s->cfg.robots_per_block, p);
  if (!ok) throw Error(MH_ERR_RANGE, "no step kernel fits this population (robot too large for LDS)");
  upload_plan(s, p);
}

StepParams make_params(mh_sim* s, const Part& part, int launch, int n_steps, int track) {
  StepParams p{};
  p.packs = s->packs.p + part.pack_offset;
  p.pos = s->pos.p;
  p.vel = s->vel.p;
  p.spair = s->spair.p;
  p.slbar = s->lbar.p;
  p.smat = s->smat.p;
  p.sstress = s->stress.p;
  p.csr = s->csr.p;
  p.row = s->row.p;
  p.mperm = s->mperm.p;
  p.ccsr = s->ccsr.p;
  p.crow = s->crow.p;
  p.ctab = s->ctab.p;
  p.mrec = s->mrec.p;
  p.mcnt = s->mcnt.p;
  p.lbtab = s->lbtab.p;
  p.scratch = s->scratch.p;
  p.mat_k = s->mat_k.p;
  p.mat_kx = s->mat_kx.p;
  p.mat_dL0 = s->mat_dL0.p;
  p.mat_omega = s->mat_omega.p;
  p.mat_phi = s->mat_phi.p;
  p.n_mat = static_cast<int>(s->compact_to_id.size());
  p.mass_cap = part.mass_cap;
  p.spring_cap = part.spring_cap;
  p.csr_cap = part.csr_cap;
  p.n_steps = n_steps;
  p.dt = s->cfg.dt;
  p.clock = s->clock.p;
  p.launch = launch;
  p.track_stress = track;
  p.fast_div = s->fast_div_mats && s->fast_div_lbar;
  p.gravity = s->cfg.gravity;
  p.ground_k = s->cfg.ground_k;
  p.mu_s = s->cfg.mu_s;
  p.mu_k = s->cfg.mu_k;
  p.drag = s->cfg.drag;
  p.damping = s->cfg.damping;
  return p;
}

int launch_count(const mh_sim* s, int n_steps) {
  const int K = s->substeps > 0 ? s->substeps : n_steps;
  return (n_steps + K - 1) / K;
}

// Issue the step launches of n_steps steps.  Launch i starts at
// s->clock[i], which the clock kernel writes before them.  The parts own
// disjoint robots: the second one runs on the aux stream, forked from and
// joined back into the main stream.  Returns launches per part.
int issue(mh_sim* s, int n_steps, int track) {
  const int K = s->substeps > 0 ? s->substeps : n_steps;
  const size_t np = s->plan.parts.size();
  if (np > 1) {
    hip_check(hipEventRecord(s->ev_fork, s->stream), "fork");
    hip_check(hipStreamWaitEvent(s->aux, s->ev_fork, 0), "fork");
  }
  int launches = 0;
  for (size_t i = 0; i < np; i++) {
    const Part& part = s->plan.parts[i];
    hipStream_t st = i == 0 ? s->stream : s->aux;
    launches = 0;
    for (int done = 0; done < n_steps;) {
      const int n = std::min(K, n_steps - done);
      StepParams p = make_params(s, part, launches, n, track);
      hip_check(launch_step(*part.v, static_cast<int>(part.packs.size()), part.lds, st, p), "step launch");
      done += n;
      launches++;
    }
  }
  if (np > 1) {
    hip_check(hipEventRecord(s->ev_join, s->aux), "join");
    hip_check(hipStreamWaitEvent(s->stream, s->ev_join, 0), "join");
  }
  return launches;
}

void run_steps(mh_sim* s, int n_steps, int track) {
  if (n_steps <= 0 || s->n_robots == 0 || s->plan.parts.empty()) {
    for (int k = 0; k < n_steps; k++) s->total_time += s->cfg.dt;
    return;
  }
  if (track && !s->stress.p) throw Error(MH_ERR_STATE, "no stress buffer");
  const int launches = launch_count(s, n_steps);
  if (s->clock.n < static_cast<size_t>(launches)) {
    // the table's address is baked into captured graphs
    hip_check(hipStreamSynchronize(s->stream), "sync");
    drop_graph(s);
    s->clock.alloc(static_cast<size_t>(launches));
  }
  // launch start times from the reference's float time recurrence, on the
  // device: the step launches (captured or not) read them
  hip_check(launch_clock(s->clock.p, s->total_time, s->cfg.dt, n_steps,
                         s->substeps > 0 ? s->substeps : n_steps, s->stream),
            "clock launch");
  if (s->timing) hip_check(hipEventRecord(s->ev0, s->stream), "event");
  if (s->cfg.use_graph) {
    const bool hit = s->graph && s->g_steps == n_steps && s->g_track == track &&
                     s->g_sub == s->substeps && s->g_key == &s->plan;
    if (!hit) {
      drop_graph(s);
      hipGraph_t g = nullptr;
      hip_check(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      try {
        issue(s, n_steps, track);
      } catch (...) {
        (void)hipStreamEndCapture(s->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hip_check(hipStreamEndCapture(s->stream, &g), "end capture");
      hipError_t e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      hip_check(e, "graph instantiate");
      s->g_steps = n_steps;
      s->g_track = track;
      s->g_sub = s->substeps;
      s->g_key = &s->plan;
      s->graph_builds++;
    }
    hip_check(hipGraphLaunch(s->graph, s->stream), "graph launch");
  } else {
    issue(s, n_steps, track);
  }
  if (s->timing) {
    hip_check(hipEventRecord(s->ev1, s->stream), "event");
    s->last_launches = launches;
    s->have_timing = true;
  }
  for (int k = 0; k < n_steps; k++) s->total_time += s->cfg.dt;
}

// The population's tuned launch plan if the tuner has one cached for its
// shape and this config, else the default plan (defined with the tuner).
void choose_plan(mh_sim* s);

int step_count(float duration, float dt) {
  // Simulator.cu:428-449: while (remaining > 0) { ...; remaining -= dt; }
  float remaining = duration;
  long n = 0;
  while (remaining > 0.0f) {
    remaining -= dt;
    if (++n > 1000000000L) throw Error(MH_ERR_RANGE, "step count overflow (dt too small)");
  }
  return static_cast<int>(n);
}

}  // namespace
}  // namespace mh

#define MH_TRY try {
#define MH_CATCH                                   \
  }                                                \
  catch (const mh::Error& e) {                     \
    mh::set_error(e.what());                       \
    return e.code;                                 \
  }                                                \
  catch (const std::bad_alloc&) {                  \
    mh::set_error("host allocation failed");       \
    return MH_ERR_NOMEM;                           \
  }                                                \
  catch (const std::exception& e) {                \
    mh::set_error(e.what());                       \
    return MH_ERR_INTERNAL;                        \
  }

extern "C" {

void mh_config_default(mh_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->solver = MH_SOLVER_XPBD_REF;
  c->dt = 0.0001f;      // Simulator.h:55
  c->gravity = 9.81f;   // environment.h:19 (EnvironmentLand)
  c->ground_k = 1.0e5f;
  c->mu_s = 1.0f;
  c->mu_k = 0.8f;       // environment.h:20
  c->drag = 0.0f;
  c->damping = 0.99f;   // environment.h:13
  c->device = 0;
  c->kernel = MH_KERNEL_AUTO;
  c->block_size = 0;
  c->robots_per_block = 0;
  c->substeps_per_launch = 0;
  c->use_graph = 1;
}

int mh_sim_create(const mh_config* cfg, const mh_material_set* set, mh_sim** out) {
  MH_TRY
  if (!cfg || !set || !out) throw mh::Error(MH_ERR_ARG, "null argument");
  if (cfg->solver != MH_SOLVER_XPBD_REF && cfg->solver != MH_SOLVER_FORCE)
    throw mh::Error(MH_ERR_ARG, "unknown solver");
  if (!(cfg->dt > 0.f)) throw mh::Error(MH_ERR_ARG, "dt must be > 0");
  if (cfg->block_size && cfg->block_size != 256 && cfg->block_size != 512 && cfg->block_size != 768 &&
      cfg->block_size != 1024)
    throw mh::Error(MH_ERR_ARG, "block_size must be 0, 256, 512, 768 or 1024");
  if (cfg->kernel < MH_KERNEL_AUTO || cfg->kernel > mh::kAllFamilies)
    throw mh::Error(MH_ERR_ARG, "unknown kernel kind");
  int ndev = 0;
  mh::hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (cfg->device < 0 || cfg->device >= ndev) throw mh::Error(MH_ERR_HIP, "no such HIP device");
  mh::hip_check(hipSetDevice(cfg->device), "hipSetDevice");
  auto s = std::make_unique<mh_sim>();
  s->cfg = *cfg;
  s->mats_by_id = set->by_id;
  s->composite_count = set->composite_count;
  for (const mh::Material& m : set->by_enc) {
    s->code_id.push_back(static_cast<uint8_t>(m.id));
    s->code_enc.push_back(m.encoding);
    s->code_solid.push_back(!(m.k == 0.f && m.dL0 == 0.f && m.omega == 0.f && m.phi == 0.f));
  }
  s->valid_bits = set->base.size() >= 32 ? 0xffffffffu : ((1u << set->base.size()) - 1u);
  s->substeps = cfg->substeps_per_launch > 0 ? cfg->substeps_per_launch : 0;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, cfg->device) == hipSuccess &&
      lds > 0)
    s->lds_limit = std::min(lds, mh::kMaxLdsBytes);
  mh::hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "stream");
  mh::hip_check(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking), "stream");
  mh::hip_check(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming), "event");
  mh::hip_check(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming), "event");
  mh::hip_check(hipEventCreate(&s->ev0), "event");
  mh::hip_check(hipEventCreate(&s->ev1), "event");
  *out = s.release();
  return MH_OK;
  MH_CATCH
}

int mh_sim_destroy(mh_sim* sim) {
  if (sim) {
    (void)hipSetDevice(sim->cfg.device);
    if (sim->stream) (void)hipStreamSynchronize(sim->stream);
    delete sim;
  }
  return MH_OK;
}

int mh_sim_set_population(mh_sim* s, const mh_population* pop) {
  MH_TRY
  if (!s || !pop) throw mh::Error(MH_ERR_ARG, "null argument");
  if (pop->n_robots < 0) throw mh::Error(MH_ERR_ARG, "n_robots < 0");
  const int R = pop->n_robots;
  if (R > 0 && (!pop->mass_offset || !pop->spring_offset))
    throw mh::Error(MH_ERR_ARG, "offsets are null");
  mh::hip_check(hipSetDevice(s->cfg.device), "hipSetDevice");
  std::vector<uint32_t> mo(R + 1, 0), so(R + 1, 0);
  if (R > 0) {
    std::copy(pop->mass_offset, pop->mass_offset + R + 1, mo.begin());
    std::copy(pop->spring_offset, pop->spring_offset + R + 1, so.begin());
  }
  if (mo[0] != 0 || so[0] != 0) throw mh::Error(MH_ERR_ARG, "offsets must start at 0");
  for (int r = 0; r < R; r++) {
    if (mo[r + 1] < mo[r] || so[r + 1] < so[r]) throw mh::Error(MH_ERR_ARG, "offsets must be non-decreasing");
    if (mo[r + 1] - mo[r] > 65536) throw mh::Error(MH_ERR_RANGE, "robot has more than 65536 masses");
  }
  const size_t M = mo[R], S = so[R];
  if ((M && (!pop->pos || !pop->mass_material_id)) ||
      (S && (!pop->spring_pairs || !pop->spring_mean_length || !pop->spring_material_id)))
    throw mh::Error(MH_ERR_ARG, "population arrays are null");

  // validate and compact springs (air springs are no-ops, sim_kernel.cu:211).
  // Pass 1, robots in parallel: validation, active counts and each robot's
  // material ids in first-appearance order.  Compact ids then follow the
  // first appearance over the whole population, robot by robot.  Pass 2,
  // robots in parallel: the compacted arrays at each robot's offset.
  std::vector<int> id_to_compact(256, -1);
  std::vector<uint8_t> compact_to_id;
  std::vector<uint32_t> ao(R + 1, 0), src;
  std::vector<uint16_t> ap;
  std::vector<float> alb;
  std::vector<uint8_t> amat;
  std::vector<std::vector<uint8_t>> ids_seen(R);
  mh::parallel_for(static_cast<size_t>(R), [&](size_t ri) {
    const int r = static_cast<int>(ri);
    const uint32_t nm = mo[r + 1] - mo[r];
    bool seen[256] = {};
    uint32_t active = 0;
    for (uint32_t k = so[r]; k < so[r + 1]; k++) {
      const uint16_t a = pop->spring_pairs[2 * k], b = pop->spring_pairs[2 * k + 1];
      const uint8_t id = pop->spring_material_id[k];
      if (a >= nm || b >= nm) throw mh::Error(MH_ERR_ARG, "spring endpoint outside its robot");
      if (id >= s->composite_count) throw mh::Error(MH_ERR_ARG, "spring material id outside the material set");
      if (id == 0) continue;
      active++;
      if (!seen[id]) {
        seen[id] = true;
        ids_seen[r].push_back(id);
      }
    }
    ao[r + 1] = active;
  });
  for (int r = 0; r < R; r++) {
    ao[r + 1] += ao[r];
    for (uint8_t id : ids_seen[r]) {
      if (id_to_compact[id] < 0) {
        id_to_compact[id] = static_cast<int>(compact_to_id.size());
        compact_to_id.push_back(id);
      }
    }
  }
  src.resize(ao[R]);
  ap.resize(2 * static_cast<size_t>(ao[R]));
  alb.resize(ao[R]);
  amat.resize(ao[R]);
  mh::parallel_for(static_cast<size_t>(R), [&](size_t ri) {
    const int r = static_cast<int>(ri);
    uint32_t j = ao[r];
    for (uint32_t k = so[r]; k < so[r + 1]; k++) {
      const uint8_t id = pop->spring_material_id[k];
      if (id == 0) continue;
      src[j] = k;
      ap[2 * static_cast<size_t>(j)] = pop->spring_pairs[2 * k];
      ap[2 * static_cast<size_t>(j) + 1] = pop->spring_pairs[2 * k + 1];
      alb[j] = pop->spring_mean_length[k];
      amat[j] = static_cast<uint8_t>(id_to_compact[id]);
      j++;
    }
  });
  if (compact_to_id.empty()) compact_to_id.push_back(0);  // keeps tables non-empty

  // per robot: gather entries with every mass's row rounded up to 4;
  // the most springs one mass has on one side; distinct (rest length,
  // material) pairs (left empty when they exceed the mass-parallel kernels'
  // table)
  std::vector<uint32_t> gpad(R, 0);
  std::vector<uint8_t> mside(R, 0);
  std::vector<uint16_t> mdeg(R, 0);
  std::vector<std::vector<uint64_t>> psets(R);
  mh::parallel_for(static_cast<size_t>(R), [&](size_t ri) {
    const int r = static_cast<int>(ri);
    {
      std::vector<uint32_t> plus(mo[r + 1] - mo[r], 0), minus(mo[r + 1] - mo[r], 0);
      std::vector<uint64_t>& ps = psets[r];
      uint64_t last = ~0ull;
      for (uint32_t k = ao[r]; k < ao[r + 1]; k++) {
        plus[ap[2 * k]]++;
        minus[ap[2 * k + 1]]++;
        uint32_t bits;
        std::memcpy(&bits, &alb[k], 4);
        const uint64_t key = static_cast<uint64_t>(bits) << 8 | amat[k];
        if (key != last && ps.size() <= static_cast<size_t>(mh::kMassPairNull)) {
          last = key;
          auto it = std::lower_bound(ps.begin(), ps.end(), key);
          if (it == ps.end() || *it != key) ps.insert(it, key);
        }
      }
      if (ps.size() > static_cast<size_t>(mh::kMassPairNull)) ps.clear();
      uint32_t g = 0, side = 0, deg = 0;
      for (size_t i = 0; i < plus.size(); i++) {
        g += (plus[i] + minus[i] + 3u) & ~3u;
        side = std::max(side, std::max(plus[i], minus[i]));
        deg = std::max(deg, plus[i] + minus[i]);
      }
      gpad[r] = g;
      mside[r] = static_cast<uint8_t>(std::min<uint32_t>(side, 255));
      mdeg[r] = static_cast<uint16_t>(std::min<uint32_t>(deg, 65535));
    }
  });

  // masses: float4 (xyz, mass) and (v, 0); baseline COM and length per
  // robot (SoftBody::updateBaseline at Reset)
  std::vector<float4> p4(M), v4(M);
  std::vector<float> bc(3 * R), bl(R);
  mh::parallel_for(static_cast<size_t>(R), [&](size_t ri) {
    const int r = static_cast<int>(ri);
    for (size_t i = mo[r]; i < mo[r + 1]; i++) {
      const float m = pop->mass ? pop->mass[i] : 1.0f;
      if (s->cfg.solver == MH_SOLVER_FORCE && !(m > 0.f)) throw mh::Error(MH_ERR_ARG, "mass must be > 0");
      p4[i] = make_float4(pop->pos[3 * i], pop->pos[3 * i + 1], pop->pos[3 * i + 2], m);
      v4[i] = pop->vel ? make_float4(pop->vel[3 * i], pop->vel[3 * i + 1], pop->vel[3 * i + 2], 0.f)
                       : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    mh::baseline(pop->pos + 3 * mo[r], pop->mass_material_id + mo[r], static_cast<int>(mo[r + 1] - mo[r]),
                 &bc[3 * r], &bl[r]);
  });

  mh::drop_graph(s);
  s->has_population = false;
  s->n_robots = R;
  s->mass_off = mo;
  s->spring_off = so;
  s->active_off = ao;
  s->gather_pad = std::move(gpad);
  s->max_side = std::move(mside);
  s->max_deg = std::move(mdeg);
  s->pair_set = std::move(psets);
  s->act_pairs = std::move(ap);
  s->act_lbar = alb;
  s->fast_div_lbar = true;
  for (float L : alb) s->fast_div_lbar = s->fast_div_lbar && L >= 0x1p-40f && L <= 0x1p20f;
  s->active_src = std::move(src);
  s->act_mat = std::move(amat);
  s->compact_to_id = std::move(compact_to_id);
  s->base_com = std::move(bc);
  s->base_len = std::move(bl);
  s->n_masses = M;
  s->n_springs = S;
  s->n_active = s->active_src.size();
  s->held_stress.clear();
  s->voxel_device = false;
  s->total_time = 0.f;

  s->pos.upload(p4.data(), M, s->stream);
  s->vel.upload(v4.data(), M, s->stream);
  s->lbar.upload(alb.data(), s->n_active, s->stream);
  s->smat.upload(s->act_mat.data(), s->n_active, s->stream);
  s->mass_id.upload(pop->mass_material_id, M, s->stream);
  s->robot_begin.upload(mo.data(), R + 1, s->stream);
  s->stress.alloc(std::max<size_t>(s->n_active, 1));
  mh::hip_check(hipMemsetAsync(s->stress.p, 0, sizeof(float) * s->stress.n, s->stream), "memset");
  s->com.alloc(3 * std::max(R, 1));
  mh::upload_materials(s);
  if (R > 0) mh::choose_plan(s);
  mh::hip_check(hipStreamSynchronize(s->stream), "upload");
  s->has_population = true;
  return MH_OK;
  MH_CATCH
}

namespace mh {
namespace {

// The mass-parallel variant that takes one robot of n masses per workgroup:
// the smallest block that holds it, within what the config allows.
const KernelVariant* voxel_variant(const mh_sim* s, int n) {
  int count = 0;
  const KernelVariant* all = kernel_variants(&count);
  const KernelVariant* best = nullptr;
  for (int i = 0; i < count; i++) {
    const KernelVariant& v = all[i];
    if (v.solver != s->cfg.solver || v.kind != MH_KERNEL_MASS || v.chunked) continue;
    if (s->cfg.kernel && !(s->cfg.kernel & MH_KERNEL_MASS)) continue;
    if (s->cfg.block_size && v.block != s->cfg.block_size) continue;
    if (v.block < n || v.nspt < 13) continue;
    if (lds_bytes_mass(v.block, s->composite_count) > static_cast<size_t>(s->lds_limit)) continue;
    if (!best || v.block < best->block) best = &v;
  }
  return best;
}

// The shape template of an nx*ny*nz lattice on the device, kept while
// populations of the same shape follow.
void prepare_voxel_shape(mh_sim* s, int nx, int ny, int nz, float resolution) {
  uint32_t res_bits;
  std::memcpy(&res_bits, &resolution, 4);
  if (s->vt_n && s->vt_nx == nx && s->vt_ny == ny && s->vt_nz == nz && s->vt_res_bits == res_bits) return;
  s->vt_n = 0;
  VoxelTemplate t;
  voxel_template(nx, ny, nz, resolution, t);
  const int n = t.n, S = t.n_springs;
  // distinct lengths (bit patterns), each spring's index among them
  std::vector<float> lengths(t.length);
  std::sort(lengths.begin(), lengths.end(), [](float a, float b) {
    uint32_t x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    return x < y;
  });
  lengths.erase(std::unique(lengths.begin(), lengths.end(),
                            [](float a, float b) { return std::memcmp(&a, &b, 4) == 0; }),
                lengths.end());
  if (static_cast<int>(lengths.size()) > kVoxPairWords / 8)
    throw Error(MH_ERR_RANGE, "lattice has more distinct spring lengths than the device build's pair index holds");
  std::vector<uint8_t> lidx(S);
  for (int k = 0; k < S; k++) {
    for (size_t j = 0; j < lengths.size(); j++)
      if (std::memcmp(&lengths[j], &t.length[k], 4) == 0) lidx[k] = static_cast<uint8_t>(j);
  }
  // per point: springs whose first (second) endpoint it is, ascending
  std::vector<uint16_t> foff(n + 1, 0), soff(n + 1, 0), flist(S), slist(S);
  for (int k = 0; k < S; k++) {
    foff[t.pairs[2 * k] + 1]++;
    soff[t.pairs[2 * k + 1] + 1]++;
  }
  for (int i = 0; i < n; i++) {
    foff[i + 1] = static_cast<uint16_t>(foff[i + 1] + foff[i]);
    soff[i + 1] = static_cast<uint16_t>(soff[i + 1] + soff[i]);
  }
  std::vector<uint16_t> fill_f(foff.begin(), foff.end() - 1), fill_s(soff.begin(), soff.end() - 1);
  for (int k = 0; k < S; k++) {
    flist[fill_f[t.pairs[2 * k]]++] = static_cast<uint16_t>(k);
    slist[fill_s[t.pairs[2 * k + 1]]++] = static_cast<uint16_t>(k);
  }
  for (int i = 0; i < n; i++)
    if (foff[i + 1] - foff[i] > 13 || soff[i + 1] - soff[i] > 13)
      throw Error(MH_ERR_INTERNAL, "lattice point with more than 13 springs on one side");
  s->vt_dpos.upload(t.pos.data(), t.pos.size(), s->stream);
  s->vt_pairs.upload(t.pairs.data(), t.pairs.size(), s->stream);
  s->vt_vox.upload(t.voxels.data(), t.voxels.size(), s->stream);
  s->vt_lidx.upload(lidx.data(), lidx.size(), s->stream);
  s->vt_dlen.upload(lengths.data(), lengths.size(), s->stream);
  s->vt_foff.upload(foff.data(), foff.size(), s->stream);
  s->vt_flist.upload(flist.data(), flist.size(), s->stream);
  s->vt_soff.upload(soff.data(), soff.size(), s->stream);
  s->vt_slist.upload(slist.data(), slist.size(), s->stream);
  s->vt_code_id.upload(s->code_id.data(), s->code_id.size(), s->stream);
  s->vt_code_enc.upload(s->code_enc.data(), s->code_enc.size(), s->stream);
  s->vt_code_solid.upload(s->code_solid.data(), s->code_solid.size(), s->stream);
  hip_check(hipStreamSynchronize(s->stream), "voxel shape upload");
  s->vt_pos = t.pos;
  s->vt_nx = nx;
  s->vt_ny = ny;
  s->vt_nz = nz;
  s->vt_res_bits = res_bits;
  s->vt_springs = S;
  s->vt_lengths = static_cast<int>(lengths.size());
  s->vt_n = n;
}

}  // namespace
}  // namespace mh

int mh_sim_set_voxel_population(mh_sim* s, int n_robots, int nx, int ny, int nz, float resolution, float mass,
                                const uint32_t* voxel_encoding) {
  MH_TRY
  if (!s) throw mh::Error(MH_ERR_ARG, "null sim");
  if (n_robots < 0 || (n_robots > 0 && !voxel_encoding)) throw mh::Error(MH_ERR_ARG, "invalid robot list");
  if (!(resolution > 0.f)) throw mh::Error(MH_ERR_ARG, "resolution must be > 0");
  if (!(mass > 0.f) || !std::isfinite(mass)) throw mh::Error(MH_ERR_ARG, "mass must be > 0");
  if (nx < 1 || ny < 1 || nz < 1 || static_cast<long>(nx) * ny * nz > mh::kVoxMaxPoints)
    throw mh::Error(MH_ERR_RANGE, "device-built voxel robots have 1..1024 lattice points");
  mh::hip_check(hipSetDevice(s->cfg.device), "hipSetDevice");
  const int n = nx * ny * nz;
  const mh::KernelVariant* v = mh::voxel_variant(s, n);
  if (!v)
    throw mh::Error(MH_ERR_ARG, "no mass-parallel step kernel holds this lattice within the config "
                                "(kernel family, block size) and the LDS budget");
  mh::hip_check(hipStreamSynchronize(s->stream), "synchronize");
  mh::drop_graph(s);
  s->has_population = false;
  mh::prepare_voxel_shape(s, nx, ny, nz, resolution);
  const int R = n_robots, S = s->vt_springs, P = v->nspt;
  const size_t M = static_cast<size_t>(R) * n, ST = static_cast<size_t>(R) * S;

  s->pos.alloc(M);
  s->vel.alloc(M);
  s->mass_id.alloc(M);
  s->stress.alloc(std::max<size_t>(ST, 1));
  mh::hip_check(hipMemsetAsync(s->stress.p, 0, sizeof(float) * s->stress.n, s->stream), "memset");
  s->com.alloc(3 * std::max(R, 1));
  s->vox_enc.alloc(M);
  if (M)
    mh::hip_check(hipMemcpyAsync(s->vox_enc.p, voxel_encoding, M * sizeof(uint32_t), hipMemcpyDefault, s->stream),
                  "encodings copy");
  s->vox_stripped.alloc(M);
  s->vox_spring_id.alloc(ST);
  s->mrec.alloc(static_cast<size_t>(R) * (4 * P + 1) * n);
  s->mcnt.alloc(static_cast<size_t>(R) * 2 * n);
  s->mperm.alloc(M);
  s->lbtab.alloc(static_cast<size_t>(R) * 2 * mh::kMassPairMax);
  s->packs.alloc(R);
  s->vox_active.alloc(R);
  s->vox_error.alloc(1);
  mh::hip_check(hipMemsetAsync(s->vox_error.p, 0, sizeof(int), s->stream), "memset");
  mh::VoxelBuildParams p{};
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.n = n;
  p.n_springs = S;
  p.n_lengths = s->vt_lengths;
  p.maxp = P;
  p.valid_bits = s->valid_bits;
  p.mass = mass;
  p.enc = s->vox_enc.p;
  p.tpos = s->vt_dpos.p;
  p.pairs = s->vt_pairs.p;
  p.voxels = s->vt_vox.p;
  p.length_index = s->vt_lidx.p;
  p.lengths = s->vt_dlen.p;
  p.first_off = s->vt_foff.p;
  p.first_list = s->vt_flist.p;
  p.second_off = s->vt_soff.p;
  p.second_list = s->vt_slist.p;
  p.code_id = s->vt_code_id.p;
  p.code_enc = s->vt_code_enc.p;
  p.code_solid = s->vt_code_solid.p;
  p.pos = s->pos.p;
  p.vel = s->vel.p;
  p.mass_id = s->mass_id.p;
  p.stripped = s->vox_stripped.p;
  p.spring_id = s->vox_spring_id.p;
  p.mrec = s->mrec.p;
  p.mcnt = s->mcnt.p;
  p.mperm = s->mperm.p;
  p.lbtab = s->lbtab.p;
  p.packs = s->packs.p;
  p.active = s->vox_active.p;
  p.error = s->vox_error.p;
  mh::hip_check(mh::launch_voxel_population(p, R, s->stream), "voxel build launch");
  int err = 0;
  std::vector<uint32_t> active(R);
  std::vector<uint8_t> mid(M);
  mh::hip_check(hipMemcpyAsync(&err, s->vox_error.p, sizeof(int), hipMemcpyDeviceToHost, s->stream), "D2H");
  if (R)
    mh::hip_check(hipMemcpyAsync(active.data(), s->vox_active.p, R * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 s->stream), "D2H");
  if (M) mh::hip_check(hipMemcpyAsync(mid.data(), s->mass_id.p, M, hipMemcpyDeviceToHost, s->stream), "D2H");
  mh::hip_check(hipStreamSynchronize(s->stream), "voxel build");
  if (err == MH_ERR_ARG) throw mh::Error(MH_ERR_ARG, "voxel encoding uses a bit outside the material set");
  if (err)
    throw mh::Error(MH_ERR_RANGE, "a robot has more distinct (rest length, material) pairs than the "
                                  "mass-parallel kernel's table holds; use mh_sim_set_population");

  // host bookkeeping: offsets, baselines (SoftBody::updateBaseline at Reset)
  s->n_robots = R;
  s->mass_off.assign(R + 1, 0);
  s->spring_off.assign(R + 1, 0);
  s->active_off.assign(R + 1, 0);
  for (int r = 0; r < R; r++) {
    s->mass_off[r + 1] = static_cast<uint32_t>((r + 1) * n);
    s->spring_off[r + 1] = static_cast<uint32_t>((r + 1) * S);
    s->active_off[r + 1] = s->active_off[r] + active[r];
  }
  s->base_com.assign(3 * static_cast<size_t>(R), 0.f);
  s->base_len.assign(R, 0.f);
  mh::parallel_for(static_cast<size_t>(R), [&](size_t r) {
    mh::baseline(s->vt_pos.data(), mid.data() + r * n, n, &s->base_com[3 * r], &s->base_len[r]);
  });
  s->gather_pad.clear();
  s->max_side.clear();
  s->max_deg.clear();
  s->pair_set.clear();
  s->act_pairs.clear();
  s->active_src.clear();
  s->act_mat.clear();
  s->act_lbar.clear();
  s->held_stress.clear();
  s->compact_to_id.resize(s->composite_count);
  std::iota(s->compact_to_id.begin(), s->compact_to_id.end(), 0);
  s->n_masses = M;
  s->n_springs = ST;
  s->n_active = s->active_off[R];
  s->total_time = 0.f;
  s->fast_div_lbar = false;
  std::vector<uint32_t> rb(R + 1);
  for (int r = 0; r <= R; r++) rb[r] = static_cast<uint32_t>(r * n);
  s->robot_begin.upload(rb.data(), R + 1, s->stream);
  mh::upload_materials(s);

  // the plan: one part, one robot per workgroup
  mh::Plan plan;
  if (R > 0) {
    mh::Part part;
    part.v = v;
    part.packs.resize(R);
    for (int r = 0; r < R; r++) {
      mh::PackDesc& d = part.packs[r];
      d.mass_begin = static_cast<uint32_t>(r * n);
      d.n_mass = static_cast<uint32_t>(n);
      d.spring_begin = static_cast<uint32_t>(r) * S;
      d.n_spring = active[r];
      d.csr_begin = static_cast<uint32_t>(r) * (4 * P + 1) * n;
      d.row_begin = static_cast<uint32_t>(r) * 2 * n;
      d.pad0 = static_cast<uint32_t>(r) * 2 * mh::kMassPairMax;
    }
    part.mass_cap = n;
    part.spring_cap = static_cast<int>(*std::max_element(active.begin(), active.end()));
    part.lds = mh::lds_bytes_mass(v->block, s->composite_count);
    part.n_robots = R;
    if (part.lds > 64 * 1024)
      for (const void* fn : {v->fn, v->fn_stress})
        if (fn) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(part.lds));
    plan.parts.push_back(std::move(part));
  }
  s->plan = std::move(plan);
  if (s->cfg.substeps_per_launch > 0) s->substeps = s->cfg.substeps_per_launch;
  s->voxel_device = true;
  mh::hip_check(hipStreamSynchronize(s->stream), "upload");
  s->has_population = true;
  return MH_OK;
  MH_CATCH
}

int mh_sim_voxel_materials(mh_sim* s, uint8_t* mass_material_id, uint8_t* spring_material_id,
                           uint32_t* stripped_encoding) {
  MH_TRY
  if (!s) throw mh::Error(MH_ERR_ARG, "null sim");
  if (!s->has_population || !s->voxel_device) throw mh::Error(MH_ERR_STATE, "no device-built voxel population");
  mh::hip_check(hipSetDevice(s->cfg.device), "hipSetDevice");
  if (mass_material_id && s->n_masses)
    mh::hip_check(hipMemcpyAsync(mass_material_id, s->mass_id.p, s->n_masses, hipMemcpyDeviceToHost, s->stream),
                  "D2H");
  if (spring_material_id && s->n_springs)
    mh::hip_check(hipMemcpyAsync(spring_material_id, s->vox_spring_id.p, s->n_springs, hipMemcpyDeviceToHost,
                                 s->stream), "D2H");
  if (stripped_encoding && s->n_masses)
    mh::hip_check(hipMemcpyAsync(stripped_encoding, s->vox_stripped.p, s->n_masses * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s->stream), "D2H");
  mh::hip_check(hipStreamSynchronize(s->stream), "D2H");
  return MH_OK;
  MH_CATCH
}

int mh_sim_simulate(mh_sim* s, float duration, int track_stresses) {
  MH_TRY
  if (!s) throw mh::Error(MH_ERR_ARG, "null sim");
  if (!s->has_population) throw mh::Error(MH_ERR_STATE, "no population set");
  mh::hip_check(hipSetDevice(s->cfg.device), "hipSetDevice");
  mh::run_steps(s, mh::step_count(duration, s->cfg.dt), track_stresses ? 1 : 0);
  return MH_OK;
  MH_CATCH
}

int mh_sim_step(mh_sim* s, int n_steps, int track_stresses) {
  MH_TRY
  if (!s) throw mh::Error(MH_ERR_ARG, "null sim");
  if (!s->has_population) throw mh::Error(MH_ERR_STATE, "no population set");
  if (n_steps < 0) throw mh::Error(MH_ERR_ARG, "n_steps < 0");
  mh::hip_check(hipSetDevice(s->cfg.device), "hipSetDevice");
  mh::run_steps(s, n_steps, track_stresses ? 1 : 0);
  return MH_OK;
  MH_CATCH
}

int mh_sim_synchronize(mh_sim* s) {
  MH_TRY
  if (!s) throw mh::Error(MH_ERR_ARG, "null sim");
  mh::hip_check(hipStreamSynchronize(s->stream), "synchronize");
  return MH_OK;
  MH_CATCH
}

int mh_sim_collect(mh_sim* s, float* pos, float* vel, float* stress) {
  MH_TRY
  if (!s) throw mh::Error(MH_ERR_ARG, "null sim");
  if (!s->has_population) throw mh::Error(MH_ERR_STATE, "no population set");
  const size_t M = s->n_masses;
  std::vector<float4> tmp(M);
  if (pos && M) {
    mh::hip_check(hipMemcpyAsync(tmp.data(), s->pos.p, M * sizeof(float4), hipMemcpyDeviceToHost, s->stream), "D2H");
    mh::hip_check(hipStreamSynchronize(s->stream), "D2H");
    for (size_t i = 0; i < M; i++) {
      pos[3 * i] = tmp[i].x;
      pos[3 * i + 1] = tmp[i].y;
      pos[3 * i + 2] = tmp[i].z;
    }
  }
  if (vel && M) {
    mh::hip_check(hipMemcpyAsync(tmp.data(), s->vel.p, M * sizeof(float4), hipMemcpyDeviceToHost, s->stream), "D2H");
    mh::hip_check(hipStreamSynchronize(s->stream), "D2H");
    for (size_t i = 0; i < M; i++) {
      vel[3 * i] = tmp[i].x;
      vel[3 * i + 1] = tmp[i].y;
      vel[3 * i + 2] = tmp[i].z;
    }
  }
  if (stress && s->voxel_device) {
    if (s->n_springs)
      mh::hip_check(hipMemcpyAsync(stress, s->stress.p, s->n_springs * sizeof(float), hipMemcpyDeviceToHost,
                                   s->stream), "D2H");
    mh::hip_check(hipStreamSynchronize(s->stream), "D2H");
  } else if (stress) {
    if (s->held_stress.empty())
      std::fill(stress, stress + s->n_springs, 0.f);
    else
      std::copy(s->held_stress.begin(), s->held_stress.end(), stress);
    if (s->n_active) {
      std::vector<float> st(s->n_active);
      mh::hip_check(hipMemcpyAsync(st.data(), s->stress.p, s->n_active * sizeof(float), hipMemcpyDeviceToHost,
                                   s->stream), "D2H");
      mh::hip_check(hipStreamSynchronize(s->stream), "D2H");
      for (size_t k = 0; k < s->n_active; k++) stress[s->active_src[k]] = st[k];
    }
  }
  return MH_OK;
  MH_CATCH
}

int mh_sim_set_stress(mh_sim* s, const float* stress) {
  MH_TRY
  if (!s || !stress) throw mh::Error(MH_ERR_ARG, "null argument");
  if (!s->has_population) throw mh::Error(MH_ERR_STATE, "no population set");
  mh::hip_check(hipSetDevice(s->cfg.device), "hipSetDevice");
  if (s->voxel_device) {
    // one slot per spring in the robots' own order; air springs are never stepped
    if (s->n_springs) {
      mh::hip_check(hipMemcpyAsync(s->stress.p, stress, s->n_springs * sizeof(float), hipMemcpyHostToDevice,
                                   s->stream), "H2D");
      mh::hip_check(hipStreamSynchronize(s->stream), "H2D");
    }
    return MH_OK;
  }
  // air springs keep their value, as slots of the reference's buffer do
  // while solveDistance skips them (sim_kernel.cu:211)
  s->held_stress.assign(stress, stress + s->n_springs);
  for (uint32_t k : s->active_src) s->held_stress[k] = 0.f;
  if (s->n_active) {
    std::vector<float> st(s->n_active);
    for (size_t k = 0; k < s->n_active; k++) st[k] = stress[s->active_src[k]];
    // the stream's earlier steps may still be accumulating into the buffer
    mh::hip_check(hipMemcpyAsync(s->stress.p, st.data(), s->n_active * sizeof(float), hipMemcpyHostToDevice,
                                 s->stream), "H2D");
    mh::hip_check(hipStreamSynchronize(s->stream), "H2D");
  }
  return MH_OK;
  MH_CATCH
}

int mh_sim_com(mh_sim* s, float* out) {
  MH_TRY
  if (!s || !out) throw mh::Error(MH_ERR_ARG, "null argument");
  if (!s->has_population) throw mh::Error(MH_ERR_STATE, "no population set");
  if (s->n_robots == 0) return MH_OK;
  mh::hip_check(mh::launch_com(s->pos.p, s->mass_id.p, s->robot_begin.p, s->n_robots, s->com.p, s->stream),
                "com launch");
  mh::hip_check(hipMemcpyAsync(out, s->com.p, 3 * s->n_robots * sizeof(float), hipMemcpyDeviceToHost, s->stream),
                "D2H");
  mh::hip_check(hipStreamSynchronize(s->stream), "D2H");
  return MH_OK;
  MH_CATCH
}

int mh_sim_fitness(mh_sim* s, int mode, float* out) {
  MH_TRY
  if (!s || !out) throw mh::Error(MH_ERR_ARG, "null argument");
  if (mode != MH_FITNESS_REFERENCE && mode != MH_FITNESS_DISPLACEMENT) throw mh::Error(MH_ERR_ARG, "bad mode");
  std::vector<float> c(3 * std::max(s->n_robots, 1));
  int rc = mh_sim_com(s, c.data());
  if (rc != MH_OK) return rc;
  for (int r = 0; r < s->n_robots; r++) {
    const float dx = c[3 * r] - s->base_com[3 * r];
    if (mode == MH_FITNESS_DISPLACEMENT) {
      out[r] = dx;
    } else {
      const float f = dx / s->base_len[r];  // SoftBody.cpp:132
      out[r] = f > 0.0f ? f : 0.0f;
    }
  }
  return MH_OK;
  MH_CATCH
}

int mh_sim_reset(mh_sim* s) {
  if (!s) return MH_ERR_ARG;
  s->total_time = 0.f;
  return MH_OK;
}

float mh_sim_total_time(const mh_sim* s) { return s ? s->total_time : 0.f; }

int mh_sim_set_total_time(mh_sim* s, float t) {
  if (!s || !std::isfinite(t)) return MH_ERR_ARG;
  s->total_time = t;
  return MH_OK;
}
float mh_sim_get_time_step(const mh_sim* s) { return s ? s->cfg.dt : 0.f; }

int mh_sim_set_time_step(mh_sim* s, float dt) {
  MH_TRY
  if (!s || !(dt > 0.f)) throw mh::Error(MH_ERR_ARG, "dt must be > 0");
  s->cfg.dt = dt;
  if (s->has_population) {
    mh::upload_materials(s);
    mh::hip_check(hipStreamSynchronize(s->stream), "upload");
  }
  mh::drop_graph(s);
  return MH_OK;
  MH_CATCH
}

int mh_sim_get_info(const mh_sim* s, mh_sim_info* info) {
  MH_TRY
  if (!s || !info) throw mh::Error(MH_ERR_ARG, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->n_robots = s->n_robots;
  info->n_masses = static_cast<int>(s->n_masses);
  info->n_springs = static_cast<int>(s->n_springs);
  info->n_active_springs = static_cast<int>(s->n_active);
  for (const mh::Part& part : s->plan.parts) {
    info->kernel |= mh::family(*part.v);
    info->n_blocks += static_cast<int>(part.packs.size());
    if (part.v->kind != MH_KERNEL_STREAM) {
      info->block_size = part.v->block;
      info->lds_bytes = static_cast<int>(part.lds);
      info->max_masses_per_block = part.mass_cap;
      info->max_springs_per_block = part.spring_cap;
      info->n_lds_robots = part.n_robots;
    } else {
      info->stream_block_size = part.v->block;
      info->n_stream_robots = part.n_robots;
      for (const mh::PackDesc& d : part.packs) info->n_stream_springs += static_cast<int>(d.n_spring);
      info->stream_chunk = part.v->chunked == 1 ? part.spring_cap : 0;
      info->stream_record_bytes = part.v->chunked == 2 ? part.v->nspt : 0;
    }
  }
  info->robots_per_block = s->plan.rpb;
  info->substeps_per_launch = s->substeps;
  info->graph_builds = s->graph_builds;
  return MH_OK;
  MH_CATCH
}

namespace mh {
namespace {

// Tuning results by (solver, device, what the config allows, population
// shape); the chosen plan never changes results, only speed.
struct TuneKey {
  int solver, device, kernel, block_size, robots_per_block, substeps, use_graph, lds_limit;
  std::vector<uint64_t> shape;
  auto tie() const {
    return std::tie(solver, device, kernel, block_size, robots_per_block, substeps, use_graph, lds_limit, shape);
  }
  bool operator<(const TuneKey& o) const { return tie() < o.tie(); }
};
struct TuneChoice {
  const KernelVariant* lds;
  const KernelVariant* stream;
  int rpb;
  int substeps;  // 0 = every substep of a call in one launch
};
std::map<TuneKey, TuneChoice>& tune_cache() {
  static std::map<TuneKey, TuneChoice> c;
  return c;
}
std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}

TuneKey tune_key(const mh_sim* s) {
  // the population's shape, coarsely: robot count, mean and largest robot
  // (masses, active springs), each rounded down to a quarter of a power of
  // two, the most springs on one side of a mass and whether every robot's
  // (rest length, material) pairs fit the mass kernels' table -- so that
  // populations drawn from one distribution share a choice
  auto bucket = [](uint64_t v) -> uint64_t {
    if (v < 8) return v;
    const int e = 63 - __builtin_clzll(v);
    return (static_cast<uint64_t>(e) << 2) | ((v >> (e - 2)) & 3u);
  };
  uint64_t sm = 0, sa = 0, mm = 0, ma = 0, side = 0, tables = 1;
  for (int r = 0; r < s->n_robots; r++) {
    const uint64_t m = static_cast<uint64_t>(robot_masses(s, r)), a = static_cast<uint64_t>(robot_active(s, r));
    sm += m;
    sa += a;
    mm = std::max(mm, m);
    ma = std::max(ma, a);
    side = std::max<uint64_t>(side, s->max_side[r]);
    tables &= (a == 0 || !s->pair_set[r].empty()) ? 1u : 0u;
  }
  const uint64_t n = static_cast<uint64_t>(std::max(s->n_robots, 1));
  const mh_config& c = s->cfg;
  return TuneKey{c.solver, c.device, c.kernel, c.block_size, c.robots_per_block, c.substeps_per_launch,
                 c.use_graph, s->lds_limit,
                 {bucket(n), bucket(sm / n), bucket(sa / n), bucket(mm), bucket(ma), side, tables}};
}

// Applies a cached choice if its variants are allowed by this simulator's
// config and its plan fits the population; false otherwise.
bool apply_choice(mh_sim* s, const TuneChoice& c) {
  auto allowed = [s](const KernelVariant* v, int kinds) {
    if (!v) return true;
    auto cs = candidates(s, kinds);
    return std::find(cs.begin(), cs.end(), v) != cs.end();
  };
  if (!allowed(c.lds, kResident) || !allowed(c.stream, kStreaming)) return false;
  Plan p;
  if (!make_plan(s, c.lds, c.stream, c.rpb, p)) return false;
  upload_plan(s, p);
  s->substeps = s->cfg.substeps_per_launch > 0 ? s->cfg.substeps_per_launch : c.substeps;
  return true;
}

// The population's tuned plan, if one is cached; else the default plan.
void choose_plan(mh_sim* s) {
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    auto it = tune_cache().find(tune_key(s));
    if (it != tune_cache().end() && apply_choice(s, it->second)) return;
  }
  s->substeps = s->cfg.substeps_per_launch > 0 ? s->cfg.substeps_per_launch : 0;
  choose_default_plan(s);
}

// Seconds for n_steps substeps with the current plan: the best of three
// runs after a warm-up (plans a few percent apart must not swap places on
// noise).
double time_steps(mh_sim* s, int n_steps) {
  run_steps(s, std::max(1, n_steps / 4), 0);  // warm-up
  hip_check(hipStreamSynchronize(s->stream), "sync");
  double best = std::numeric_limits<double>::infinity();
  for (int rep = 0; rep < 3; rep++) {
    auto t0 = std::chrono::steady_clock::now();
    run_steps(s, n_steps, 0);
    hip_check(hipStreamSynchronize(s->stream), "sync");
    best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  return best;
}

double time_plan(mh_sim* s, Plan& p, int n_steps) {
  upload_plan(s, p);
  return time_steps(s, n_steps);
}

// Saves the simulated state and the settings the tuner changes; restores
// them when it goes out of scope, also when tuning throws.
struct TuneScope {
  mh_sim* s;
  std::vector<float4> p0, v0;
  std::vector<float> st0;
  float t_save;
  int use_graph, substeps;
  bool restored = false;
  explicit TuneScope(mh_sim* sim)
      : s(sim), p0(sim->n_masses), v0(sim->n_masses), st0(sim->stress.n), t_save(sim->total_time),
        use_graph(sim->cfg.use_graph), substeps(sim->substeps) {
    copy(hipMemcpyDeviceToHost);
  }
  void copy(hipMemcpyKind kind) {
    const bool d2h = kind == hipMemcpyDeviceToHost;
    auto one = [&](void* host, void* dev, size_t bytes) {
      if (!bytes) return;
      hip_check(hipMemcpyAsync(d2h ? host : dev, d2h ? dev : host, bytes, kind, s->stream), "tune state copy");
    };
    one(p0.data(), s->pos.p, p0.size() * sizeof(float4));
    one(v0.data(), s->vel.p, v0.size() * sizeof(float4));
    one(st0.data(), s->stress.p, st0.size() * sizeof(float));
    hip_check(hipStreamSynchronize(s->stream), "tune state copy");
  }
  void restore() {
    restored = true;
    s->cfg.use_graph = use_graph;
    s->total_time = t_save;
    (void)hipStreamSynchronize(s->stream);
    copy(hipMemcpyHostToDevice);
  }
  ~TuneScope() {
    if (restored) return;
    s->substeps = substeps;
    try {
      restore();
    } catch (...) {
    }
  }
};

}  // namespace
}  // namespace mh

int mh_sim_autotune(mh_sim* s, int n_steps) {
  MH_TRY
  if (!s) throw mh::Error(MH_ERR_ARG, "null sim");
  if (!s->has_population) throw mh::Error(MH_ERR_STATE, "no population set");
  if (s->n_robots == 0) return MH_OK;
  if (n_steps < 1) n_steps = 1;
  mh::hip_check(hipSetDevice(s->cfg.device), "hipSetDevice");
  // earlier asynchronous steps finish before the state is saved
  mh::hip_check(hipStreamSynchronize(s->stream), "synchronize");
  if (s->voxel_device) {
    // the device-built packing is fixed: only substeps per launch are timed
    if (s->cfg.substeps_per_launch > 0) return MH_OK;
    mh::TuneScope scope(s);
    double best_k = std::numeric_limits<double>::infinity();
    int choice = 0;
    for (int K : {0, 2500, 1000, 250}) {
      if (K > 0 && 2 * K > n_steps) continue;
      s->substeps = K;
      const double t = mh::time_steps(s, n_steps);
      if (t < best_k * (K == 0 ? 1.0 : 0.99)) {
        best_k = t;
        choice = K;
      }
    }
    s->substeps = choice;
    scope.substeps = choice;
    scope.restore();
    return MH_OK;
  }
  const mh::TuneKey key = mh::tune_key(s);
  {
    std::lock_guard<std::mutex> lk(mh::tune_mutex());
    auto it = mh::tune_cache().find(key);
    if (it != mh::tune_cache().end() && mh::apply_choice(s, it->second)) return MH_OK;
  }
  // candidates run on a copy of the state; the population is restored after
  mh::TuneScope scope(s);
  s->cfg.use_graph = 0;
  s->substeps = 0;
  auto lc = mh::candidates(s, mh::kResident);
  auto sc = mh::candidates(s, mh::kStreaming);
  lc.push_back(nullptr);  // everything streamed
  // streaming variant for the LDS search: the first that can take every robot
  const mh::KernelVariant* sv0 = sc.empty() ? nullptr : sc.front();
  for (const mh::KernelVariant* sv : sc) {
    mh::Plan p;
    if (mh::make_plan(s, nullptr, sv, 0, p)) {
      sv0 = sv;
      break;
    }
  }
  std::vector<int> rpbs = {s->cfg.robots_per_block};
  if (s->cfg.robots_per_block == 0) rpbs = {0, 1, 2, 4};
  double best = std::numeric_limits<double>::infinity();
  mh::TuneChoice choice{nullptr, nullptr, 0, 0};
  // coordinate search: LDS variant x robots per block, then streaming
  // variant, then substeps per launch
  for (const mh::KernelVariant* lv : lc) {
    for (int rpb : rpbs) {
      if (!lv && rpb != rpbs.front()) continue;
      mh::Plan p;
      if (!mh::make_plan(s, lv, sv0, rpb, p)) continue;
      // skip robots-per-block caps that produce the same packing
      if (rpb > 0 && !p.parts.empty() && p.parts[0].v == lv &&
          static_cast<int>(p.parts[0].packs.size()) == p.parts[0].n_robots && rpb > 1)
        continue;
      const double t = mh::time_plan(s, p, n_steps);
      if (t < best) {
        best = t;
        choice = {lv, sv0, rpb, 0};
      }
    }
  }
  for (const mh::KernelVariant* sv : sc) {
    if (sv == sv0) continue;
    mh::Plan p;
    if (!mh::make_plan(s, choice.lds, sv, choice.rpb, p)) continue;
    bool uses_stream = false;
    for (auto& part : p.parts) uses_stream |= part.v->kind == MH_KERNEL_STREAM;
    if (!uses_stream) break;
    const double t = mh::time_plan(s, p, n_steps);
    if (t < best) {
      best = t;
      choice.stream = sv;
    }
  }
  if (!std::isfinite(best)) throw mh::Error(MH_ERR_RANGE, "no step kernel fits this population");
  mh::Plan p;
  mh::make_plan(s, choice.lds, choice.stream, choice.rpb, p);
  mh::upload_plan(s, p);
  // substeps per launch, timed with the caller's graph setting: one launch
  // per call, or launches of K substeps (a caller's synchronize waits for
  // at most one launch)
  s->cfg.use_graph = scope.use_graph;
  if (s->cfg.substeps_per_launch > 0) {
    choice.substeps = s->cfg.substeps_per_launch;
  } else {
    double best_k = std::numeric_limits<double>::infinity();
    for (int K : {0, 2500, 1000, 250}) {
      if (K > 0 && 2 * K > n_steps) continue;
      s->substeps = K;
      const double t = mh::time_steps(s, n_steps);
      // shorter launches must win by more than noise
      if (t < best_k * (K == 0 ? 1.0 : 0.99)) {
        best_k = t;
        choice.substeps = K;
      }
    }
  }
  s->substeps = choice.substeps;
  scope.restore();
  {
    std::lock_guard<std::mutex> lk(mh::tune_mutex());
    mh::tune_cache()[key] = choice;
  }
  return MH_OK;
  MH_CATCH
}

int mh_sim_device_state(mh_sim* s, void** pos4, void** vel4, void** stream) {
  if (!s) return MH_ERR_ARG;
  if (pos4) *pos4 = s->pos.p;
  if (vel4) *vel4 = s->vel.p;
  if (stream) *stream = s->stream;
  return MH_OK;
}

int mh_sim_set_timing(mh_sim* s, int enable) {
  if (!s) return MH_ERR_ARG;
  s->timing = enable != 0;
  s->have_timing = false;
  return MH_OK;
}

int mh_sim_last_kernel_ms(const mh_sim* s, float* avg_ms, int* launches) {
  MH_TRY
  if (!s || !avg_ms) throw mh::Error(MH_ERR_ARG, "null argument");
  *avg_ms = 0.f;
  if (launches) *launches = 0;
  if (!s->have_timing || s->last_launches == 0) return MH_OK;
  mh::hip_check(hipEventSynchronize(s->ev1), "event sync");
  float ms = 0.f;
  mh::hip_check(hipEventElapsedTime(&ms, s->ev0, s->ev1), "elapsed");
  *avg_ms = ms / s->last_launches;
  if (launches) *launches = s->last_launches;
  return MH_OK;
  MH_CATCH
}

}  // extern "C"
