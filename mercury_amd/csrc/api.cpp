// api.cpp -- C entry points for errors, materials, voxel topology, baseline.
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

#include "mh_internal.h"

namespace mh {
namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }
void clear_error() { g_error.clear(); }
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

const char* mh_last_error(void) { return mh::g_error.c_str(); }
int mh_version(void) { return 100; }

int mh_materials_reference(mh_material_set** out) {
  MH_TRY
  if (!out) throw mh::Error(MH_ERR_ARG, "out is null");
  *out = mh::make_material_set(mh::reference_base()).release();
  return MH_OK;
  MH_CATCH
}

int mh_materials_create(int n_base, const mh_base_material* base, mh_material_set** out) {
  MH_TRY
  if (!out || (n_base > 0 && !base) || n_base < 0)
    throw mh::Error(MH_ERR_ARG, "invalid material list");
  std::vector<mh::Material> b(n_base + 1);
  b[0].encoding = 1u;
  for (int i = 0; i < n_base; i++) {
    b[i + 1].id = static_cast<uint32_t>(i + 1);
    b[i + 1].k = base[i].k;
    b[i + 1].dL0 = base[i].dL0;
    b[i + 1].omega = base[i].omega;
    b[i + 1].phi = base[i].phi;
    b[i + 1].encoding = 1u << (i + 1);
    if (!(base[i].k > 0.f)) throw mh::Error(MH_ERR_ARG, "base material stiffness must be > 0");
  }
  *out = mh::make_material_set(b).release();
  return MH_OK;
  MH_CATCH
}

int mh_materials_destroy(mh_material_set* set) {
  delete set;
  return MH_OK;
}

int mh_materials_count(const mh_material_set* set) {
  if (!set) {
    mh::set_error("set is null");
    return MH_ERR_ARG;
  }
  return set->composite_count;
}

int mh_materials_by_id(const mh_material_set* set, int id, float kdwp[4], uint32_t* encoding) {
  MH_TRY
  if (!set || id < 0 || id >= set->composite_count) throw mh::Error(MH_ERR_ARG, "bad material id");
  const mh::Material& m = set->by_id[id];
  if (kdwp) {
    kdwp[0] = m.k;
    kdwp[1] = m.dL0;
    kdwp[2] = m.omega;
    kdwp[3] = m.phi;
  }
  if (encoding) *encoding = m.encoding;
  return MH_OK;
  MH_CATCH
}

int mh_materials_decode(const mh_material_set* set, uint32_t encoding, uint32_t* id,
                        int32_t* encoded_index) {
  MH_TRY
  if (!set) throw mh::Error(MH_ERR_ARG, "set is null");
  // the reference scans composite_count bits of the encoding
  uint32_t e = encoding;
  if (set->composite_count < 32) e &= (1u << set->composite_count) - 1u;
  const int idx = set->encoded_index(e);
  if (idx >= set->composite_count) throw mh::Error(MH_ERR_ARG, "encoding outside the material set");
  if (id) *id = set->by_enc[idx].id;
  if (encoded_index) *encoded_index = idx;
  return MH_OK;
  MH_CATCH
}

int mh_voxel_max_springs(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return 0;
  return 13 * nx * ny * nz;
}

int mh_voxel_build(const mh_material_set* set, int nx, int ny, int nz, float resolution,
                   const uint32_t* voxel_encoding, float* mass_pos, uint8_t* mass_material_id,
                   uint16_t* spring_pairs, float* spring_length, uint8_t* spring_material_id,
                   uint32_t* spring_encoding, int max_springs, int* n_springs,
                   uint32_t* stripped_encoding) {
  MH_TRY
  if (!set || !voxel_encoding || !mass_pos || !mass_material_id || !spring_pairs ||
      !spring_length || !spring_material_id || !n_springs)
    throw mh::Error(MH_ERR_ARG, "null argument");
  if (!(resolution > 0.f)) throw mh::Error(MH_ERR_ARG, "resolution must be > 0");
  mh::VoxelOut out;
  mh::voxel_build(*set, nx, ny, nz, resolution, voxel_encoding, out);
  const int ns = static_cast<int>(out.length.size());
  *n_springs = ns;
  if (ns > max_springs) throw mh::Error(MH_ERR_RANGE, "max_springs too small");
  const int n = nx * ny * nz;
  std::memcpy(mass_pos, out.pos.data(), sizeof(float) * 3 * n);
  std::memcpy(mass_material_id, out.mass_id.data(), n);
  std::memcpy(spring_pairs, out.pairs.data(), sizeof(uint16_t) * 2 * ns);
  std::memcpy(spring_length, out.length.data(), sizeof(float) * ns);
  std::memcpy(spring_material_id, out.spring_id.data(), ns);
  if (spring_encoding) std::memcpy(spring_encoding, out.spring_enc.data(), sizeof(uint32_t) * ns);
  if (stripped_encoding) std::memcpy(stripped_encoding, out.stripped_enc.data(), sizeof(uint32_t) * n);
  return MH_OK;
  MH_CATCH
}

int mh_voxel_template(int nx, int ny, int nz, float resolution, float* mass_pos, uint16_t* spring_pairs,
                      float* spring_length, uint16_t* spring_voxels, int max_springs, int* n_springs) {
  MH_TRY
  if (!mass_pos || !spring_pairs || !spring_length || !spring_voxels || !n_springs)
    throw mh::Error(MH_ERR_ARG, "null argument");
  if (!(resolution > 0.f)) throw mh::Error(MH_ERR_ARG, "resolution must be > 0");
  mh::VoxelTemplate t;
  mh::voxel_template(nx, ny, nz, resolution, t);
  *n_springs = t.n_springs;
  if (t.n_springs > max_springs) throw mh::Error(MH_ERR_RANGE, "max_springs too small");
  std::memcpy(mass_pos, t.pos.data(), sizeof(float) * 3 * t.n);
  std::memcpy(spring_pairs, t.pairs.data(), sizeof(uint16_t) * 2 * t.n_springs);
  std::memcpy(spring_length, t.length.data(), sizeof(float) * t.n_springs);
  std::memcpy(spring_voxels, t.voxels.data(), sizeof(uint16_t) * 4 * t.n_springs);
  return MH_OK;
  MH_CATCH
}

int mh_voxel_build_batch(const mh_material_set* set, int n_robots, int nx, int ny, int nz,
                         float resolution, const uint32_t* voxel_encoding, int threads,
                         float* mass_pos, uint8_t* mass_material_id, uint32_t* spring_offset,
                         uint16_t* spring_pairs, float* spring_length, uint8_t* spring_material_id,
                         int64_t spring_capacity) {
  MH_TRY
  if (!set || n_robots < 0 || (n_robots > 0 && (!voxel_encoding || !mass_pos || !mass_material_id ||
                                                !spring_offset || !spring_pairs || !spring_length ||
                                                !spring_material_id)))
    throw mh::Error(MH_ERR_ARG, "null argument");
  if (!(resolution > 0.f)) throw mh::Error(MH_ERR_ARG, "resolution must be > 0");
  if (nx < 1 || ny < 1 || nz < 1 || static_cast<long>(nx) * ny * nz > 65536)
    throw mh::Error(MH_ERR_ARG, "lattice must have 1..65536 points");
  const size_t nv = static_cast<size_t>(nx) * ny * nz;
  std::vector<mh::VoxelOut> out(static_cast<size_t>(n_robots));
  int nt = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
  nt = std::max(1, std::min(nt, std::max(n_robots, 1)));
  std::vector<std::string> errors(static_cast<size_t>(nt));
  std::atomic<int> next{0};
  auto work = [&](int t) {
    try {
      for (int r = next++; r < n_robots; r = next++)
        mh::voxel_build(*set, nx, ny, nz, resolution, voxel_encoding + nv * r, out[r]);
    } catch (const std::exception& e) {
      errors[t] = e.what();
      next = n_robots;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  for (const std::string& e : errors)
    if (!e.empty()) throw mh::Error(MH_ERR_ARG, e);
  spring_offset[0] = 0;
  for (int r = 0; r < n_robots; r++) {
    const size_t ns = out[r].length.size();
    if (spring_offset[r] + ns > static_cast<uint64_t>(spring_capacity))
      throw mh::Error(MH_ERR_RANGE, "spring_capacity too small");
    spring_offset[r + 1] = spring_offset[r] + static_cast<uint32_t>(ns);
  }
  for (int r = 0; r < n_robots; r++) {
    const mh::VoxelOut& o = out[r];
    const size_t s0 = spring_offset[r], ns = o.length.size();
    std::memcpy(mass_pos + 3 * nv * r, o.pos.data(), sizeof(float) * 3 * nv);
    std::memcpy(mass_material_id + nv * r, o.mass_id.data(), nv);
    std::memcpy(spring_pairs + 2 * s0, o.pairs.data(), sizeof(uint16_t) * 2 * ns);
    std::memcpy(spring_length + s0, o.length.data(), sizeof(float) * ns);
    std::memcpy(spring_material_id + s0, o.spring_id.data(), ns);
  }
  return MH_OK;
  MH_CATCH
}

int mh_rest_lengths(const float* pos, const uint16_t* pairs, int n_springs, float* out) {
  MH_TRY
  if (n_springs < 0 || (n_springs > 0 && (!pos || !pairs || !out)))
    throw mh::Error(MH_ERR_ARG, "null argument");
  for (int s = 0; s < n_springs; s++) {
    const float* a = pos + 3 * pairs[2 * s];
    const float* b = pos + 3 * pairs[2 * s + 1];
    const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    const float yz = dy * dy + dz * dz;  // Eigen's 3-vector squaredNorm: x^2 + (y^2 + z^2)
    out[s] = std::sqrt(dx * dx + yz);
  }
  return MH_OK;
  MH_CATCH
}

int mh_baseline(const float* pos, const uint8_t* mass_material_id, int n_masses, float com[3],
                float* length) {
  MH_TRY
  if (!pos || !mass_material_id || !com || !length || n_masses < 0)
    throw mh::Error(MH_ERR_ARG, "null argument");
  mh::baseline(pos, mass_material_id, n_masses, com, length);
  return MH_OK;
  MH_CATCH
}

}  // extern "C"
