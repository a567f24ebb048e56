// mh_internal.h -- shared host-side types of libmercury_hip.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mercury_hip.h"

namespace mh {

// Error carrying an MH_ERR_* code; every C entry point converts it.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string& msg);
void clear_error();

// One material record (material.h:42-48 without the colour).
struct Material {
  uint32_t id = 0;
  float k = 0, dL0 = 0, omega = 0, phi = 0;
  uint32_t encoding = 0;
  bool same_params(const Material& o) const {  // Material::operator==
    return k == o.k && dL0 == o.dL0 && omega == o.omega && phi == o.phi;
  }
};

}  // namespace mh

// Composite material tables of one base set (material.h:192-322).
struct mh_material_set {
  std::vector<mh::Material> base;    // index 0 = air, index i has id i
  std::vector<mh::Material> by_id;   // id_lookup
  std::vector<mh::Material> by_enc;  // decode, indexed by encoded index
  int composite_count = 0;
  int encoded_index(uint32_t enc) const;
  const mh::Material& decode(uint32_t enc) const {
    return by_enc[encoded_index(enc)];
  }
};

namespace mh {
std::unique_ptr<mh_material_set> make_material_set(const std::vector<Material>& base);
std::vector<Material> reference_base();

// Host-side copy of one voxel robot topology.
struct VoxelOut {
  std::vector<float> pos;          // 3n
  std::vector<uint8_t> mass_id;    // n
  std::vector<uint16_t> pairs;     // 2s
  std::vector<float> length;       // s
  std::vector<uint8_t> spring_id;  // s
  std::vector<uint32_t> spring_enc;
  std::vector<uint32_t> stripped_enc;
};
// What every voxel robot of one lattice shape shares: mass positions
// (after ShiftX / ShiftY), springs in the DFS order with their lengths,
// and per spring the up to 4 voxels (lattice point indices, 0xffff = none)
// whose non-air encoding bits, OR-ed and decoded, give its material.
struct VoxelTemplate {
  int nx = 0, ny = 0, nz = 0, n = 0, n_springs = 0;
  std::vector<float> pos;          // 3n
  std::vector<uint16_t> pairs;     // 2 n_springs
  std::vector<float> length;       // n_springs
  std::vector<uint16_t> voxels;    // 4 n_springs
};
void voxel_template(int nx, int ny, int nz, float resolution, VoxelTemplate& t);
void voxel_build(const mh_material_set& set, int nx, int ny, int nz,
                 float resolution, const uint32_t* voxel_enc, VoxelOut& out);

// SoftBody fitness helpers (SoftBody.cpp:98-137).
void baseline(const float* pos, const uint8_t* mass_id, int n, float com[3],
              float* length);

}  // namespace mh
