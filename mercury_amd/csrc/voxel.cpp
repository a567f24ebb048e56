// voxel.cpp -- voxel-robot topology, bit-exact with VoxelRobot::Build.
//
// Reference: common/evolvables/VoxelRobot.cpp
//   Strip (267-283): keep the largest face-connected non-air component; among
//     equal sizes the one reached first in voxel-index order.  Done here with
//     an explicit-stack flood fill in index order.
//   Build (310-331): one mass per lattice point at (1/res)*index with the
//     material of the voxel at that index, springs by DFS from (0,0,0) over
//     13 forward offsets, then ShiftX (max x -> 0) and ShiftY (min y -> 0).
//   BuildSpringsRecurse (60-193): every (node, offset) with a valid target
//     emits a spring before the target's visited check; the DFS is run here
//     with an explicit stack of (node, next offset) frames, which visits in
//     the same order as the recursion.
//   Spring length: Eigen Vector3f::norm() = sqrtf(x*x + (y*y + z*z)).
//   Spring material: cardinal = decode(OR of the up to 4 non-air voxels
//     sharing the edge); face diagonal = decode(source-or-backtracked voxel
//     | its in-plane neighbour), or that voxel's material when the neighbour
//     is off-lattice; body diagonal = the source-or-backtracked voxel.
#include <cmath>
#include <cstdlib>

#include "mh_internal.h"

namespace mh {

namespace {

struct Lattice {
  int nx, ny, nz;
  bool valid(int x, int y, int z) const {
    return x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
  }
  int index(int x, int y, int z) const { return x + y * nx + z * nx * ny; }
  void coords(int i, int& x, int& y, int& z) const {
    x = i % nx;
    y = (i / nx) % ny;
    z = (i / nx / ny) % nz;
  }
};

// The 13 forward offsets in the order the reference recursion tries them.
constexpr int kOffsets[13][3] = {
    {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1},
    {1, -1, 0}, {0, 1, -1}, {-1, 0, 1},
    {-1, 1, 1}, {1, -1, 1}, {1, 1, -1}};

bool is_air(const Material& m) { return m.k == 0.f && m.dL0 == 0.f && m.omega == 0.f && m.phi == 0.f; }

// Material of the spring from voxel `s` (coords sx,sy,sz) along offset o.
Material spring_material(const mh_material_set& set, const Lattice& L,
                         const std::vector<Material>& vox, int sx, int sy, int sz,
                         const int* o) {
  const int l1 = std::abs(o[0]) + std::abs(o[1]) + std::abs(o[2]);
  auto non_air_bits = [&](const Material& m) -> uint32_t {
    return m.encoding != 1u ? m.encoding : 0u;
  };
  if (l1 == 1) {
    // the four voxels whose corner set contains this edge: the source and
    // its neighbours one step back along the two axes normal to the edge
    const int ax = o[0] ? 0 : (o[1] ? 1 : 2);
    const int u = ax == 0 ? 1 : 0, w = ax == 2 ? 1 : 2;
    uint32_t enc = 0;
    for (int du = 0; du <= 1; du++)
      for (int dw = 0; dw <= 1; dw++) {
        int c[3] = {sx, sy, sz};
        c[u] -= du;
        c[w] -= dw;
        if (!L.valid(c[0], c[1], c[2])) continue;
        enc |= non_air_bits(vox[L.index(c[0], c[1], c[2])]);
      }
    return set.decode(enc);
  }
  // voxel carrying the spring: the source, moved back one step along every
  // negative component of the offset
  int m[3] = {sx - (o[0] == -1), sy - (o[1] == -1), sz - (o[2] == -1)};
  const Material& carrier = vox[L.index(m[0], m[1], m[2])];
  if (l1 == 3) return carrier;
  // face diagonal: pair the carrier with its neighbour one step back along
  // the axis normal to the face
  int nb[3] = {m[0] - (o[0] == 0), m[1] - (o[1] == 0), m[2] - (o[2] == 0)};
  if (!L.valid(nb[0], nb[1], nb[2])) return carrier;
  uint32_t enc = non_air_bits(vox[L.index(nb[0], nb[1], nb[2])]) | non_air_bits(carrier);
  return set.decode(enc);
}

// The voxels whose non-air bits decide the material of the spring from
// (sx, sy, sz) along offset o (0xffff = none): as spring_material, the
// material is decode(OR of their non-air bits); decode(0) is air, and a
// single voxel's own non-air bits decode to its own material.
void spring_voxels(const Lattice& L, int sx, int sy, int sz, const int* o, uint16_t out[4]) {
  for (int k = 0; k < 4; k++) out[k] = 0xffff;
  const int l1 = std::abs(o[0]) + std::abs(o[1]) + std::abs(o[2]);
  if (l1 == 1) {
    const int ax = o[0] ? 0 : (o[1] ? 1 : 2);
    const int u = ax == 0 ? 1 : 0, w = ax == 2 ? 1 : 2;
    int k = 0;
    for (int du = 0; du <= 1; du++)
      for (int dw = 0; dw <= 1; dw++) {
        int c[3] = {sx, sy, sz};
        c[u] -= du;
        c[w] -= dw;
        if (L.valid(c[0], c[1], c[2])) out[k++] = static_cast<uint16_t>(L.index(c[0], c[1], c[2]));
      }
    return;
  }
  const int m[3] = {sx - (o[0] == -1), sy - (o[1] == -1), sz - (o[2] == -1)};
  out[0] = static_cast<uint16_t>(L.index(m[0], m[1], m[2]));
  if (l1 == 3) return;
  const int nb[3] = {m[0] - (o[0] == 0), m[1] - (o[1] == 0), m[2] - (o[2] == 0)};
  if (L.valid(nb[0], nb[1], nb[2])) out[1] = static_cast<uint16_t>(L.index(nb[0], nb[1], nb[2]));
}

// The reference's DFS from (0,0,0) over the 13 forward offsets, with an
// explicit stack of (node, next offset) frames: emit(src, dst, sx, sy, sz,
// offset) for every (node, offset) with a valid target, before the
// target's visited check (BuildSpringsRecurse, VoxelRobot.cpp:60-193).
template <typename Emit>
void dfs_springs(const Lattice& L, Emit&& emit) {
  const int n = L.nx * L.ny * L.nz;
  std::vector<uint8_t> visited(n, 0);
  std::vector<std::pair<int, int>> frames;
  visited[0] = 1;
  frames.emplace_back(0, 0);
  while (!frames.empty()) {
    auto& f = frames.back();
    if (f.second == 13) {
      frames.pop_back();
      continue;
    }
    const int src = f.first;
    const int* o = kOffsets[f.second++];
    int sx, sy, sz;
    L.coords(src, sx, sy, sz);
    const int tx = sx + o[0], ty = sy + o[1], tz = sz + o[2];
    if (!L.valid(tx, ty, tz)) continue;
    const int dst = L.index(tx, ty, tz);
    emit(src, dst, sx, sy, sz, o);
    if (!visited[dst]) {
      visited[dst] = 1;
      frames.emplace_back(dst, 0);  // invalidates f; not used below
    }
  }
}

// Lattice positions (1/res * index), then ShiftX (max x -> 0) and ShiftY
// (min y -> 0) over all masses, the translation added in float.
void lattice_positions(const Lattice& L, float resolution, std::vector<float>& base, std::vector<float>& pos) {
  const int n = L.nx * L.ny * L.nz;
  const float inv = 1.0f / resolution;
  base.assign(3 * static_cast<size_t>(n), 0.f);
  for (int i = 0; i < n; i++) {
    int x, y, z;
    L.coords(i, x, y, z);
    base[3 * i] = inv * static_cast<float>(x);
    base[3 * i + 1] = inv * static_cast<float>(y);
    base[3 * i + 2] = inv * static_cast<float>(z);
  }
  float max_x = base[0];
  for (int i = 1; i < n; i++) max_x = base[3 * i] > max_x ? base[3 * i] : max_x;
  const float tx = -max_x;
  float min_y = base[1];
  for (int i = 1; i < n; i++) min_y = base[3 * i + 1] < min_y ? base[3 * i + 1] : min_y;
  const float ty = -min_y;
  pos.assign(3 * static_cast<size_t>(n), 0.f);
  for (int i = 0; i < n; i++) {
    pos[3 * i] = base[3 * i] + tx;
    pos[3 * i + 1] = base[3 * i + 1] + ty;
    pos[3 * i + 2] = base[3 * i + 2];
  }
}

// Spring length as Eigen's Vector3f::norm(): sqrtf(x*x + (y*y + z*z)).
float spring_length(const std::vector<float>& base, int src, int dst) {
  const float dx = base[3 * src] - base[3 * dst];
  const float dy = base[3 * src + 1] - base[3 * dst + 1];
  const float dz = base[3 * src + 2] - base[3 * dst + 2];
  const float yz = dy * dy + dz * dz;
  return std::sqrt(dx * dx + yz);
}

}  // namespace

void voxel_template(int nx, int ny, int nz, float resolution, VoxelTemplate& t) {
  if (nx < 1 || ny < 1 || nz < 1 || static_cast<long>(nx) * ny * nz > 65536)
    throw Error(MH_ERR_RANGE, "voxel lattice must have 1..65536 points (u16 mass indices)");
  const Lattice L{nx, ny, nz};
  t.nx = nx;
  t.ny = ny;
  t.nz = nz;
  t.n = nx * ny * nz;
  std::vector<float> base;
  lattice_positions(L, resolution, base, t.pos);
  t.pairs.clear();
  t.length.clear();
  t.voxels.clear();
  dfs_springs(L, [&](int src, int dst, int sx, int sy, int sz, const int* o) {
    t.pairs.push_back(static_cast<uint16_t>(src));
    t.pairs.push_back(static_cast<uint16_t>(dst));
    t.length.push_back(spring_length(base, src, dst));
    uint16_t v[4];
    spring_voxels(L, sx, sy, sz, o, v);
    t.voxels.insert(t.voxels.end(), v, v + 4);
  });
  t.n_springs = static_cast<int>(t.length.size());
}

void voxel_build(const mh_material_set& set, int nx, int ny, int nz, float resolution,
                 const uint32_t* voxel_enc, VoxelOut& out) {
  if (nx < 1 || ny < 1 || nz < 1 || static_cast<long>(nx) * ny * nz > 65536)
    throw Error(MH_ERR_RANGE, "voxel lattice must have 1..65536 points (u16 mass indices)");
  const Lattice L{nx, ny, nz};
  const int n = nx * ny * nz;
  const uint32_t valid_bits =
      set.base.size() >= 32 ? 0xffffffffu : ((1u << set.base.size()) - 1u);
  std::vector<Material> vox(n);
  for (int i = 0; i < n; i++) {
    if (voxel_enc[i] & ~valid_bits)
      throw Error(MH_ERR_ARG, "voxel encoding uses a bit outside the material set");
    vox[i] = set.decode(voxel_enc[i]);
  }

  // ---- Strip: component sizes in index order, keep the first largest
  std::vector<int> comp(n, -1);
  std::vector<int> stack;
  int best = -1, best_size = 0, ncomp = 0;
  for (int i = 0; i < n; i++) {
    if (comp[i] != -1 || is_air(vox[i])) continue;
    int size = 0;
    stack.assign(1, i);
    comp[i] = ncomp;
    while (!stack.empty()) {
      int c = stack.back();
      stack.pop_back();
      size++;
      int x, y, z;
      L.coords(c, x, y, z);
      const int nbr[6][3] = {{x + 1, y, z}, {x - 1, y, z}, {x, y + 1, z},
                             {x, y - 1, z}, {x, y, z + 1}, {x, y, z - 1}};
      for (const auto& q : nbr) {
        if (!L.valid(q[0], q[1], q[2])) continue;
        int j = L.index(q[0], q[1], q[2]);
        if (comp[j] != -1 || is_air(vox[j])) continue;
        comp[j] = ncomp;
        stack.push_back(j);
      }
    }
    if (size > best_size) {
      best_size = size;
      best = ncomp;
    }
    ncomp++;
  }
  for (int i = 0; i < n; i++)
    if (best < 0 || comp[i] != best) vox[i] = set.by_enc[0];

  // ---- masses
  std::vector<float> base;
  lattice_positions(L, resolution, base, out.pos);
  out.mass_id.assign(n, 0);
  out.stripped_enc.assign(n, 0);
  for (int i = 0; i < n; i++) {
    out.mass_id[i] = static_cast<uint8_t>(vox[i].id);
    out.stripped_enc[i] = vox[i].encoding;
  }

  // ---- springs in the DFS emission order
  out.pairs.clear();
  out.length.clear();
  out.spring_id.clear();
  out.spring_enc.clear();
  dfs_springs(L, [&](int src, int dst, int sx, int sy, int sz, const int* o) {
    const Material m = spring_material(set, L, vox, sx, sy, sz, o);
    out.pairs.push_back(static_cast<uint16_t>(src));
    out.pairs.push_back(static_cast<uint16_t>(dst));
    out.length.push_back(spring_length(base, src, dst));
    out.spring_id.push_back(static_cast<uint8_t>(m.id));
    out.spring_enc.push_back(m.encoding);
  });
}

void baseline(const float* pos, const uint8_t* mass_id, int n, float com[3], float* length) {
  // SoftBody::updateCOM (112-122): incremental mean in float, divide by i+1
  float c[3] = {0.f, 0.f, 0.f};
  float count = 0.f;
  float lo = 0.f, hi = 0.f;  // updateLength (98-110) starts from the origin
  for (int i = 0; i < n; i++) {
    if (mass_id[i] == 0) continue;
    const float den = count + 1.f;
    for (int k = 0; k < 3; k++) c[k] = c[k] + (pos[3 * i + k] - c[k]) / den;
    count = den;
    const float x = pos[3 * i];
    if (x < lo) lo = x;
    if (x > hi) hi = x;
  }
  com[0] = c[0];
  com[1] = c[1];
  com[2] = c[2];
  const float l = std::fabs(hi - lo);
  *length = l > 1.0f ? l : 1.0f;
}

}  // namespace mh
