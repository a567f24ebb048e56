// mercury_hip_io.hpp -- header-only text formats of the reference's robots,
// so solutions written here load in the reference's decoder and visualizer
// and the reference's files load here.
//
// Mirrors (paths in the reference repository):
//   EncodeSoftBody / DecodeSoftBody   SoftBody::Encode / Decode
//                                     (common/evolvables/SoftBody.cpp:147-309) with the
//                                     stream formats of Mass, Spring, Face, Cell
//                                     (common/simulator/{mass,spring,face,cell}.h)
//   EncodeVoxelRobot / DecodeVoxelRobot  VoxelRobot::Encode / Decode
//                                     (common/evolvables/VoxelRobot.cpp:504-610), voxel
//                                     lines as Voxel::to_string (VoxelRobot.h:51-63)
//   SolutionFileName                  Optimizer::WriteSolutions (optimizer_impl.h:35-48)
//
// Formats, byte for byte:
//   soft body: "type=NNRobot\n", then "masses=", "springs=", "faces=", "cells="
//   each followed by its items joined with ';' and a '\n' after the last item
//   (no '\n' for an empty list, as in the reference), then
//   "boundary_count=<n>\n".  Numbers use 9 significant digits (%.9g):
//   mass "id,x,y,z,mass,material_id" (x, y, z = the proto position), spring
//   "m0,m1,rest_length,mean_length,material_id", face "m0,m1,m2", cell
//   "m0,m1,m2,m3,mean_volume,material_encoding".
//   voxel robot: "type=VoxelRobot\n", xSize, ySize, zSize and resolution on
//   lines of their own (%f), then per voxel
//   "v ID,ix,iy,iz,cx,cy,cz,bx,by,bz,k, dL0, omega, phi, r, g, b, a\n"
//   (%f for every float), voxel i at indices (i % nx, i / nx % ny,
//   i / nx / ny), base = index / resolution, center = base + 0.5 / resolution.
// Decoding accepts what the encoder writes, including a soft body whose
// empty lists leave several keys on one line (the reference's reader stops
// there).  Decoded voxel materials are matched to the base table by
// (k, dL0, omega, phi) at the 6 decimals the format keeps.
#ifndef MERCURY_HIP_IO_HPP
#define MERCURY_HIP_IO_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mercury_hip_sim.hpp"

namespace mercury_hip {

namespace io_detail {

inline std::string g9(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.9g", v);
  return buf;
}

inline std::string f6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%f", v);
  return buf;
}

// Splits s at every occurrence of sep.
inline std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  size_t b = 0;
  for (;;) {
    const size_t e = s.find(sep, b);
    out.push_back(s.substr(b, e == std::string::npos ? std::string::npos : e - b));
    if (e == std::string::npos) break;
    b = e + 1;
  }
  return out;
}

inline float to_f(const std::string& s) {
  char* end = nullptr;
  const float v = std::strtof(s.c_str(), &end);
  if (end == s.c_str()) throw Error(MH_ERR_ARG, "robot file: bad number '" + s + "'");
  return v;
}

inline long to_i(const std::string& s) {
  char* end = nullptr;
  const long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str()) throw Error(MH_ERR_ARG, "robot file: bad integer '" + s + "'");
  return v;
}

}  // namespace io_detail

// SoftBody::Encode.  Masses are written with their proto positions and ids.
inline std::string EncodeSoftBody(const Element& e) {
  using io_detail::g9;
  std::string s = "type=NNRobot\n";
  s += "masses=";
  for (size_t i = 0; i < e.masses.size(); i++) {
    const Mass& m = e.masses[i];
    s += std::to_string(m.id) + "," + g9(m.protoPos[0]) + "," + g9(m.protoPos[1]) + "," + g9(m.protoPos[2]) +
         "," + g9(m.mass) + "," + std::to_string(static_cast<unsigned>(m.material.id));
    s += i + 1 < e.masses.size() ? ";" : "\n";
  }
  s += "springs=";
  for (size_t i = 0; i < e.springs.size(); i++) {
    const Spring& p = e.springs[i];
    s += std::to_string(p.m0) + "," + std::to_string(p.m1) + "," + g9(p.rest_length) + "," + g9(p.mean_length) +
         "," + std::to_string(static_cast<unsigned>(p.material.id));
    s += i + 1 < e.springs.size() ? ";" : "\n";
  }
  s += "faces=";
  for (size_t i = 0; i < e.faces.size(); i++) {
    const Face& f = e.faces[i];
    s += std::to_string(f.m0) + "," + std::to_string(f.m1) + "," + std::to_string(f.m2);
    s += i + 1 < e.faces.size() ? ";" : "\n";
  }
  s += "cells=";
  for (size_t i = 0; i < e.cells.size(); i++) {
    const Cell& c = e.cells[i];
    s += std::to_string(c.m0) + "," + std::to_string(c.m1) + "," + std::to_string(c.m2) + "," +
         std::to_string(c.m3) + "," + g9(c.mean_volume) + "," +
         std::to_string(static_cast<uint16_t>(c.material.encoding));
    s += i + 1 < e.cells.size() ? ";" : "\n";
  }
  s += "boundary_count=" + std::to_string(e.boundaryCount) + "\n";
  return s;
}

// SoftBody::Decode from the file's text: materials by composite id (cells
// by encoding) from ms; pos = protoPos, velocity zero.
inline Element DecodeSoftBody(const std::string& text, const MaterialSet& ms) {
  using namespace io_detail;
  Element e;
  // key=value fields in order; an empty list leaves the next key on the
  // same line
  std::vector<std::pair<std::string, std::string>> fields;
  for (const std::string& line : split(text, '\n')) {
    std::string rest = line;
    while (!rest.empty()) {
      const size_t eq = rest.find('=');
      if (eq == std::string::npos) break;
      const std::string key = rest.substr(0, eq);
      std::string value = rest.substr(eq + 1);
      std::string tail;
      for (const char* k : {"springs=", "faces=", "cells=", "boundary_count="}) {
        const size_t at = value.find(k);
        if (at == 0) {  // empty list: the value is the next field
          tail = value;
          value.clear();
          break;
        }
      }
      fields.emplace_back(key, value);
      rest = tail;
    }
  }
  for (const auto& kv : fields) {
    const std::string& key = kv.first;
    const std::string& value = kv.second;
    if (key == "type" && value != "NNRobot") throw Error(MH_ERR_ARG, "robot file: not a soft body");
    if (value.empty() && key != "type") continue;
    if (key == "masses") {
      for (const std::string& item : split(value, ';')) {
        const auto f = split(item, ',');
        if (f.size() != 6) throw Error(MH_ERR_ARG, "robot file: bad mass '" + item + "'");
        Mass m;
        m.id = static_cast<uint32_t>(to_i(f[0]));
        for (int a = 0; a < 3; a++) m.pos[a] = m.protoPos[a] = to_f(f[1 + a]);
        m.mass = to_f(f[4]);
        m.material = ms.by_id(static_cast<int>(to_i(f[5])));
        e.masses.push_back(m);
      }
    } else if (key == "springs") {
      for (const std::string& item : split(value, ';')) {
        const auto f = split(item, ',');
        if (f.size() != 5) throw Error(MH_ERR_ARG, "robot file: bad spring '" + item + "'");
        Spring p;
        p.m0 = static_cast<uint16_t>(to_i(f[0]));
        p.m1 = static_cast<uint16_t>(to_i(f[1]));
        p.rest_length = to_f(f[2]);
        p.mean_length = to_f(f[3]);
        p.material = ms.by_id(static_cast<int>(to_i(f[4])));
        e.springs.push_back(p);
      }
    } else if (key == "faces") {
      for (const std::string& item : split(value, ';')) {
        const auto f = split(item, ',');
        if (f.size() != 3) throw Error(MH_ERR_ARG, "robot file: bad face '" + item + "'");
        e.faces.push_back({static_cast<uint16_t>(to_i(f[0])), static_cast<uint16_t>(to_i(f[1])),
                           static_cast<uint16_t>(to_i(f[2]))});
      }
    } else if (key == "cells") {
      for (const std::string& item : split(value, ';')) {
        const auto f = split(item, ',');
        if (f.size() != 6) throw Error(MH_ERR_ARG, "robot file: bad cell '" + item + "'");
        Cell c;
        c.m0 = static_cast<uint16_t>(to_i(f[0]));
        c.m1 = static_cast<uint16_t>(to_i(f[1]));
        c.m2 = static_cast<uint16_t>(to_i(f[2]));
        c.m3 = static_cast<uint16_t>(to_i(f[3]));
        c.mean_volume = to_f(f[4]);
        c.material = ms.decode(static_cast<uint32_t>(to_i(f[5])));
        e.cells.push_back(c);
      }
    } else if (key == "boundary_count") {
      e.boundaryCount = static_cast<unsigned>(to_i(value));
    }
  }
  return e;
}

// One base material as a voxel line carries it (Material::to_string).
struct VoxelMaterial {
  float k = 0, dL0 = 0, omega = 0, phi = 0;
  float r = 0, g = 0, b = 0, a = 0;
};

// The voxel lattice of a VoxelRobot: nx * ny * nz voxels (lattice points),
// voxel i holding base material material[i] (index into a VoxelMaterial
// table; 0 = air), as genome_to_lattice / mh_voxel_build take it.
struct VoxelGrid {
  float xSize = 0, ySize = 0, zSize = 0, resolution = 1;
  unsigned nx = 0, ny = 0, nz = 0;
  std::vector<uint8_t> material;
};

// The reference's 23 base materials (air, adductor muscles 0-9, abductor
// muscles 0-9, tissue, bone) with their display colours
// (common/simulator/material.h:111-133), indexed by base id.
inline std::vector<VoxelMaterial> ReferenceVoxelMaterials() {
  MaterialSet ms;
  std::vector<VoxelMaterial> t(23);
  const float add[3] = {106.0f / 255.0f, 211.0f / 255.0f, 250.0f / 255.0f};
  const float abd[3] = {89.0f / 255.0f, 217.0f / 255.0f, 78.0f / 255.0f};
  const float tis[3] = {217.0f / 255.0f, 102.0f / 255.0f, 78.0f / 255.0f};
  const float bon[3] = {240.0f / 255.0f, 209.0f / 255.0f, 98.0f / 255.0f};
  for (int i = 1; i <= 22; i++) {
    const Material m = ms.by_id(i);
    const float* c = i <= 10 ? add : i <= 20 ? abd : i == 21 ? tis : bon;
    t[i] = {m.k, m.dL0, m.omega, m.phi, c[0], c[1], c[2], 0.0f};
  }
  return t;
}

// VoxelRobot::Encode.
inline std::string EncodeVoxelRobot(const VoxelGrid& v, const std::vector<VoxelMaterial>& table) {
  using io_detail::f6;
  const size_t n = static_cast<size_t>(v.nx) * v.ny * v.nz;
  if (v.material.size() != n) throw Error(MH_ERR_ARG, "voxel grid size mismatch");
  std::string s = "type=VoxelRobot\n";
  s += std::to_string(v.xSize) + "\n" + std::to_string(v.ySize) + "\n" + std::to_string(v.zSize) + "\n" +
       std::to_string(v.resolution) + "\n";
  const float inv = 1.0f / v.resolution;
  const float half = inv * 0.5f;
  for (size_t i = 0; i < n; i++) {
    const unsigned ix = static_cast<unsigned>(i % v.nx), iy = static_cast<unsigned>(i / v.nx % v.ny),
                   iz = static_cast<unsigned>(i / v.nx / v.ny % v.nz);
    const float bx = inv * static_cast<float>(ix), by = inv * static_cast<float>(iy),
                bz = inv * static_cast<float>(iz);
    if (v.material[i] >= table.size()) throw Error(MH_ERR_ARG, "voxel material outside the table");
    const VoxelMaterial& m = table[v.material[i]];
    s += "v " + std::to_string(static_cast<unsigned>(i)) + "," + std::to_string(ix) + "," + std::to_string(iy) +
         "," + std::to_string(iz) + "," + f6(bx + half) + "," + f6(by + half) + "," + f6(bz + half) + "," +
         f6(bx) + "," + f6(by) + "," + f6(bz) + "," + f6(m.k) + ", " + f6(m.dL0) + ", " + f6(m.omega) + ", " +
         f6(m.phi) + ", " + f6(m.r) + ", " + f6(m.g) + ", " + f6(m.b) + ", " + f6(m.a) + "\n";
  }
  return s;
}

// VoxelRobot::Decode from the file's text.
inline VoxelGrid DecodeVoxelRobot(const std::string& text, const std::vector<VoxelMaterial>& table) {
  using namespace io_detail;
  const auto lines = split(text, '\n');
  if (lines.size() < 5 || lines[0] != "type=VoxelRobot") throw Error(MH_ERR_ARG, "robot file: not a voxel robot");
  VoxelGrid v;
  v.xSize = to_f(lines[1]);
  v.ySize = to_f(lines[2]);
  v.zSize = to_f(lines[3]);
  v.resolution = to_f(lines[4]);
  v.nx = static_cast<unsigned>(std::lround(v.xSize * v.resolution));
  v.ny = static_cast<unsigned>(std::lround(v.ySize * v.resolution));
  v.nz = static_cast<unsigned>(std::lround(v.zSize * v.resolution));
  const size_t n = static_cast<size_t>(v.nx) * v.ny * v.nz;
  v.material.assign(n, 0);
  auto r6 = [](float x) { return std::nearbyint(static_cast<double>(x) * 1e6); };
  size_t count = 0;
  for (size_t l = 5; l < lines.size(); l++) {
    if (lines[l].empty()) continue;
    if (lines[l][0] != 'v') throw Error(MH_ERR_ARG, "robot file: bad voxel line");
    const auto f = split(lines[l].substr(1), ',');
    if (f.size() != 18) throw Error(MH_ERR_ARG, "robot file: bad voxel line '" + lines[l] + "'");
    const long id = to_i(f[0]);
    if (id < 0 || static_cast<size_t>(id) >= n) throw Error(MH_ERR_ARG, "robot file: voxel id out of range");
    const float k = to_f(f[10]), dl = to_f(f[11]), om = to_f(f[12]), ph = to_f(f[13]);
    int found = -1;
    for (size_t t = 0; t < table.size() && found < 0; t++)
      if (r6(table[t].k) == r6(k) && r6(table[t].dL0) == r6(dl) && r6(table[t].omega) == r6(om) &&
          r6(table[t].phi) == r6(ph))
        found = static_cast<int>(t);
    if (found < 0) throw Error(MH_ERR_ARG, "robot file: voxel material not in the table");
    v.material[static_cast<size_t>(id)] = static_cast<uint8_t>(found);
    count++;
  }
  if (count != n) throw Error(MH_ERR_ARG, "robot file: voxel count does not match the sizes");
  return v;
}

// Optimizer::WriteSolutions file name of solution i.
inline std::string SolutionFileName(unsigned i, float fitness) {
  return "solution_" + std::to_string(i) + "_fitness_" + std::to_string(fitness);
}

}  // namespace mercury_hip

#endif  // MERCURY_HIP_IO_HPP
