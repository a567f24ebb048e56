// ref_dump.cpp -- harness around the reference's own host code.
// TEST INFRASTRUCTURE ONLY: builds against the reference sources where they
// lie (see oracle/ref/Makefile) and writes golden vectors that pin the
// oracle and the product.  Nothing here is shipped.
//
// Modes (little-endian binary I/O, parsed by tests/golden/make_golden.py):
//   materials <out>        material.h id_lookup / decode tables and
//                          decode()/encodedCompositeIdx() of listed
//                          encodings (input list read from <in>)
//   voxel <in> <out>       VoxelRobot::Initialize + per-voxel decode(enc) +
//                          Build(); dumps masses, springs, baseline COM and
//                          length; SoftBody::Reset() mean lengths; and the
//                          fitness of given position sets via Update().
//   encode <in> <out>      the same robot inputs as `voxel`; writes the
//                          reference's VoxelRobot::Encode and
//                          SoftBody::Encode texts of the built robot, then
//                          each decoded by VoxelRobot::Decode /
//                          SoftBody::Decode and encoded again (a decoder
//                          that throws gives "exception: <what>").
//   decode <kind> <in> <out>  decodes the text file <in> with the
//                          reference's VoxelRobot::Decode (kind "voxel") or
//                          SoftBody::Decode (kind "soft") and writes its
//                          Encode text.
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <vector>
#include <stdexcept>
#include <string>
#include "VoxelRobot.h"

namespace {

struct Probe : public VoxelRobot {
  Eigen::Vector3f baseCOM() const { return mBaseCOM; }
  float length() const { return mLength; }
};

template <typename T>
void put(FILE* f, const T& v) { fwrite(&v, sizeof(T), 1, f); }

template <typename T>
T get(FILE* f) {
  T v;
  if (fread(&v, sizeof(T), 1, f) != 1) { fprintf(stderr, "short read\n"); exit(2); }
  return v;
}

void put_mat(FILE* f, const Material& m) {
  put<uint32_t>(f, m.id);
  put<float>(f, m.k);
  put<float>(f, m.dL0);
  put<float>(f, m.omega);
  put<float>(f, m.phi);
  put<uint32_t>(f, m.encoding);
}

int do_materials(const char* in, const char* out) {
  FILE* fi = fopen(in, "rb");
  FILE* fo = fopen(out, "wb");
  if (!fi || !fo) return 1;
  put<int32_t>(fo, COMPOSITE_COUNT);
  for (int i = 0; i < COMPOSITE_COUNT; i++) put_mat(fo, materials::id_lookup((uint8_t)i));
  for (int i = 0; i < MATERIAL_COUNT; i++) put_mat(fo, materials::matLookup(i));
  int32_t n = get<int32_t>(fi);
  put<int32_t>(fo, n);
  for (int i = 0; i < n; i++) {
    uint32_t e = get<uint32_t>(fi);
    Material m = materials::decode(e);
    put<uint32_t>(fo, e);
    put<int32_t>(fo, materials::encodedCompositeIdx(e));
    put_mat(fo, m);
  }
  // TestMatEncoding known answers (common/tests/test_simulator.cpp:65-81)
  put<int32_t>(fo, materials::decode(materials::air.encoding) != materials::air);
  put<int32_t>(fo, materials::decode(materials::bone.encoding) != materials::bone);
  put<int32_t>(fo, materials::decode(materials::adductor_muscle0.encoding) != materials::adductor_muscle0);
  put<int32_t>(fo, materials::decode(materials::air.encoding | materials::bone.encoding) != materials::air);
  fclose(fi);
  fclose(fo);
  return 0;
}

int do_voxel(const char* in, const char* out) {
  FILE* fi = fopen(in, "rb");
  FILE* fo = fopen(out, "wb");
  if (!fi || !fo) return 1;
  float size = get<float>(fi);
  float res = get<float>(fi);
  int32_t nvox = get<int32_t>(fi);
  std::vector<uint32_t> enc(nvox);
  for (int i = 0; i < nvox; i++) enc[i] = get<uint32_t>(fi);
  Probe R;
  R.setResolution(res);
  R.setUniformSize(size);
  R.Initialize();
  if ((int)R.getVoxels().size() != nvox) {
    fprintf(stderr, "voxel count %zu != %d\n", R.getVoxels().size(), nvox);
    return 3;
  }
  for (int i = 0; i < nvox; i++) R.getVoxels()[i].mat = materials::decode(enc[i]);
  R.Build();
  const auto& ms = R.getMasses();
  const auto& ss = R.getSprings();
  put<int32_t>(fo, (int32_t)ms.size());
  for (const Mass& m : ms) {
    put<float>(fo, m.pos.x()); put<float>(fo, m.pos.y()); put<float>(fo, m.pos.z());
    put<uint32_t>(fo, m.material.id);
    put<uint32_t>(fo, m.material.encoding);
  }
  put<int32_t>(fo, (int32_t)ss.size());
  for (const Spring& s : ss) {
    put<uint16_t>(fo, s.m0); put<uint16_t>(fo, s.m1);
    put<float>(fo, s.rest_length); put<float>(fo, s.mean_length);
    put<uint32_t>(fo, s.material.id);
    put<uint32_t>(fo, s.material.encoding);
  }
  Eigen::Vector3f b = R.baseCOM();
  put<float>(fo, b.x()); put<float>(fo, b.y()); put<float>(fo, b.z());
  put<float>(fo, R.length());
  // Reset(): mean lengths recomputed from the shifted proto positions
  R.Reset();
  for (const Spring& s : R.getSprings()) put<float>(fo, s.mean_length);
  b = R.baseCOM();
  put<float>(fo, b.x()); put<float>(fo, b.y()); put<float>(fo, b.z());
  // fitness of supplied position sets
  int32_t nsets = get<int32_t>(fi);
  put<int32_t>(fo, nsets);
  for (int k = 0; k < nsets; k++) {
    Element e;
    e.masses = R.getMasses();
    e.springs = R.getSprings();
    for (auto& m : e.masses) {
      float x = get<float>(fi), y = get<float>(fi), z = get<float>(fi);
      m.pos = Eigen::Vector3f(x, y, z);
    }
    Probe P = R;
    P.Update(e);
    Eigen::Vector3f c = P.getCOM();
    put<float>(fo, c.x()); put<float>(fo, c.y()); put<float>(fo, c.z());
    put<float>(fo, P.fitness());
  }
  fclose(fi);
  fclose(fo);
  return 0;
}

void put_text(FILE* f, const std::string& s) {
  put<int32_t>(f, (int32_t)s.size());
  fwrite(s.data(), 1, s.size(), f);
}

bool write_file(const std::string& path, const std::string& s) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  fwrite(s.data(), 1, s.size(), f);
  fclose(f);
  return true;
}

std::string reencode(const std::string& kind, const std::string& path) {
  try {
    if (kind == "voxel") {
      VoxelRobot V;
      V.Decode(path);
      return V.Encode();
    }
    SoftBody S;
    S.Decode(path);
    return S.Encode();
  } catch (const std::exception& e) {
    return std::string("exception: ") + e.what();
  }
}

int do_encode(const char* in, const char* out) {
  FILE* fi = fopen(in, "rb");
  FILE* fo = fopen(out, "wb");
  if (!fi || !fo) return 1;
  float size = get<float>(fi);
  float res = get<float>(fi);
  int32_t nvox = get<int32_t>(fi);
  std::vector<uint32_t> enc(nvox);
  for (int i = 0; i < nvox; i++) enc[i] = get<uint32_t>(fi);
  VoxelRobot R;
  R.setResolution(res);
  R.setUniformSize(size);
  R.Initialize();
  if ((int)R.getVoxels().size() != nvox) return 3;
  for (int i = 0; i < nvox; i++) R.getVoxels()[i].mat = materials::decode(enc[i]);
  R.Build();
  const std::string vox = R.Encode();
  const std::string soft = R.SoftBody::Encode();
  put_text(fo, vox);
  put_text(fo, soft);
  const std::string tmp = std::string(out) + ".txt";
  if (!write_file(tmp, vox)) return 4;
  put_text(fo, reencode("voxel", tmp));
  if (!write_file(tmp, soft)) return 4;
  put_text(fo, reencode("soft", tmp));
  remove(tmp.c_str());
  fclose(fi);
  fclose(fo);
  return 0;
}

int do_decode(const char* kind, const char* in, const char* out) {
  FILE* fo = fopen(out, "wb");
  if (!fo) return 1;
  const std::string s = reencode(kind, in);
  fwrite(s.data(), 1, s.size(), fo);
  fclose(fo);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 4 && std::string(argv[1]) == "materials") return do_materials(argv[2], argv[3]);
  if (argc >= 4 && std::string(argv[1]) == "voxel") return do_voxel(argv[2], argv[3]);
  if (argc >= 4 && std::string(argv[1]) == "encode") return do_encode(argv[2], argv[3]);
  if (argc >= 5 && std::string(argv[1]) == "decode") return do_decode(argv[2], argv[3], argv[4]);
  fprintf(stderr, "usage: ref_dump materials <in> <out> | voxel <in> <out> | encode <in> <out> | "
                  "decode voxel|soft <in> <out>\n");
  return 1;
}
