// Robot text formats of the C++ view (include/mercury_hip_io.hpp): builds
// one voxel robot with the reference material set, writes its soft-body
// and voxel-robot encodings, decodes both and checks that re-encoding gives
// the same bytes.
//   io_demo <softbody.txt> <voxel.txt> <n_lattice>
// Lattice point i holds base material (7 i) % 23 (0 = air), air on the last
// layer of every axis.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "mercury_hip_io.hpp"

using namespace mercury_hip;

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const int n = atoi(argv[3]), nv = n * n * n;
  MaterialSet ms;
  VoxelGrid grid;
  grid.nx = grid.ny = grid.nz = static_cast<unsigned>(n);
  grid.resolution = 10.0f;
  grid.xSize = grid.ySize = grid.zSize = static_cast<float>(n) / grid.resolution;
  grid.material.resize(nv);
  std::vector<uint32_t> enc(nv);
  for (int i = 0; i < nv; i++) {
    const int x = i % n, y = (i / n) % n, z = i / (n * n);
    const bool edge = x == n - 1 || y == n - 1 || z == n - 1;
    grid.material[i] = edge ? 0 : static_cast<uint8_t>((7 * i) % 23);
    enc[i] = 1u << grid.material[i];
  }
  std::vector<float> pos(3 * nv), len(mh_voxel_max_springs(n, n, n));
  std::vector<uint8_t> mid(nv), sid(len.size());
  std::vector<uint16_t> pairs(2 * len.size());
  int ns = 0;
  check(mh_voxel_build(ms.get(), n, n, n, grid.resolution, enc.data(), pos.data(), mid.data(), pairs.data(),
                       len.data(), sid.data(), nullptr, static_cast<int>(len.size()), &ns, nullptr));
  Element e;
  for (int i = 0; i < nv; i++) {
    Mass m;
    m.id = static_cast<uint32_t>(i);
    for (int a = 0; a < 3; a++) m.pos[a] = m.protoPos[a] = pos[3 * i + a];
    m.material = ms.by_id(mid[i]);
    m.mass = 0.001f;
    e.masses.push_back(m);
  }
  for (int k = 0; k < ns; k++) {
    Spring s;
    s.m0 = pairs[2 * k];
    s.m1 = pairs[2 * k + 1];
    s.rest_length = s.mean_length = len[k];
    s.material = ms.by_id(sid[k]);
    e.springs.push_back(s);
  }
  const std::string soft = EncodeSoftBody(e);
  const auto table = ReferenceVoxelMaterials();
  const std::string vox = EncodeVoxelRobot(grid, table);
  std::ofstream(argv[1], std::ios::binary) << soft;
  std::ofstream(argv[2], std::ios::binary) << vox;
  // round trips
  if (EncodeSoftBody(DecodeSoftBody(soft, ms)) != soft) {
    std::fprintf(stderr, "soft body round trip differs\n");
    return 1;
  }
  const VoxelGrid back = DecodeVoxelRobot(vox, table);
  if (back.material != grid.material || back.nx != grid.nx || EncodeVoxelRobot(back, table) != vox) {
    std::fprintf(stderr, "voxel robot round trip differs\n");
    return 1;
  }
  Element empty;
  const std::string es = EncodeSoftBody(empty);
  if (es != "type=NNRobot\nmasses=springs=faces=cells=boundary_count=0\n" ||
      !DecodeSoftBody(es, ms).masses.empty()) {
    std::fprintf(stderr, "empty soft body\n");
    return 1;
  }
  std::printf("ok %d masses %d springs\n", nv, ns);
  return 0;
}
