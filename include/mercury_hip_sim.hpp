// mercury_hip_sim.hpp -- header-only C++ view of the C-ABI with the
// reference simulator's class interface, so evaluate-style callers port
// unchanged.
//
// Mirrors (paths in the reference repository):
//   Mass / Spring / Face / Cell / Element   common/simulator/{mass,spring,
//                                           face,cell,element}.h
//   ElementTracker, class Simulator         common/simulator/Simulator.h:11-149
//     Initialize, SetElement(s), Simulate, Collect, Reset, getDeltaT,
//     setTimeStep, getTotalTime
// Differences: positions are float[3] instead of Eigen::Vector3f; a
// Material carries the composite id that indexes the material set; errors
// throw mercury_hip::Error instead of calling exit(); several simulators
// may coexist (no process-global state).  Fitness (SoftBody::updateFitness,
// SoftBody.cpp:124-137) is computed on the device by Fitness().
#ifndef MERCURY_HIP_SIM_HPP
#define MERCURY_HIP_SIM_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mercury_hip.h"

namespace mercury_hip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc) {
  if (rc != MH_OK) throw Error(rc, mh_last_error());
}

struct Material {
  uint8_t id = 0;  // composite id in the simulator's material set; 0 = air
  float k = 0, dL0 = 0, omega = 0, phi = 0;
  uint32_t encoding = 1;
};

struct Mass {
  uint32_t id = 0;
  float pos[3] = {0, 0, 0};
  float protoPos[3] = {0, 0, 0};
  float vel[3] = {0, 0, 0};
  Material material;
  float mass = 1.0f;
};

struct Spring {
  uint16_t m0 = 0, m1 = 0;
  float rest_length = 0, mean_length = 0;
  Material material;
};

struct Face {
  uint16_t m0 = 0, m1 = 0, m2 = 0;
};

struct Cell {
  uint16_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  float mean_volume = 0;
  Material material;
};

struct Element {
  std::vector<Mass> masses;
  std::vector<Spring> springs;
  std::vector<Face> faces;
  std::vector<Cell> cells;
  unsigned int boundaryCount = 0;
  float sim_time = 0;
  float total_sim_time = 0;
};

struct ElementTracker {
  uint32_t ID = 0;
  size_t mass_begin = 0, mass_end = 0, spring_begin = 0, spring_end = 0;
};

// Owns an mh_material_set (reference set by default).
class MaterialSet {
 public:
  MaterialSet() { check(mh_materials_reference(&h_)); }
  explicit MaterialSet(const std::vector<mh_base_material>& base) {
    check(mh_materials_create(static_cast<int>(base.size()), base.data(), &h_));
  }
  ~MaterialSet() { mh_materials_destroy(h_); }
  MaterialSet(const MaterialSet&) = delete;
  MaterialSet& operator=(const MaterialSet&) = delete;
  const mh_material_set* get() const { return h_; }
  Material by_id(int id) const {
    float v[4];
    uint32_t enc = 0;
    check(mh_materials_by_id(h_, id, v, &enc));
    Material m;
    m.id = static_cast<uint8_t>(id);
    m.k = v[0]; m.dL0 = v[1]; m.omega = v[2]; m.phi = v[3];
    m.encoding = enc;
    return m;
  }
  Material decode(uint32_t encoding) const {
    uint32_t id = 0;
    check(mh_materials_decode(h_, encoding, &id, nullptr));
    return by_id(static_cast<int>(id));
  }

 private:
  mh_material_set* h_ = nullptr;
};

class Simulator {
 public:
  explicit Simulator(const MaterialSet& set) : set_(set) { mh_config_default(&cfg_); }
  ~Simulator() { mh_sim_destroy(sim_); }
  Simulator(const Simulator&) = delete;
  Simulator& operator=(const Simulator&) = delete;

  // Simulator::Initialize(Config::Simulator): solver, time step, environment.
  void Initialize(const mh_config& cfg) {
    cfg_ = cfg;
    mh_sim_destroy(sim_);
    sim_ = nullptr;
    check(mh_sim_create(&cfg_, set_.get(), &sim_));
  }
  void Initialize() { Initialize(cfg_); }

  ElementTracker SetElement(const Element& e) { return SetElements({e})[0]; }

  std::vector<ElementTracker> SetElements(const std::vector<Element>& elements) {
    if (!sim_) Initialize();
    elements_ = elements;
    const int R = static_cast<int>(elements.size());
    mo_.assign(R + 1, 0);
    so_.assign(R + 1, 0);
    for (int r = 0; r < R; r++) {
      mo_[r + 1] = mo_[r] + static_cast<uint32_t>(elements[r].masses.size());
      so_[r + 1] = so_[r] + static_cast<uint32_t>(elements[r].springs.size());
    }
    pos_.assign(3 * mo_[R], 0.f);
    vel_.assign(3 * mo_[R], 0.f);
    mass_.assign(mo_[R], 1.f);
    mid_.assign(mo_[R], 0);
    pairs_.assign(2 * so_[R], 0);
    lbar_.assign(so_[R], 0.f);
    smat_.assign(so_[R], 0);
    for (int r = 0; r < R; r++) {
      const Element& e = elements[r];
      for (size_t i = 0; i < e.masses.size(); i++) {
        const size_t g = mo_[r] + i;
        for (int c = 0; c < 3; c++) {
          pos_[3 * g + c] = e.masses[i].pos[c];
          vel_[3 * g + c] = e.masses[i].vel[c];
        }
        mass_[g] = e.masses[i].mass;
        mid_[g] = e.masses[i].material.id;
      }
      for (size_t k = 0; k < e.springs.size(); k++) {
        const size_t g = so_[r] + k;
        pairs_[2 * g] = e.springs[k].m0;
        pairs_[2 * g + 1] = e.springs[k].m1;
        lbar_[g] = e.springs[k].mean_length;
        smat_[g] = e.springs[k].material.id;
      }
    }
    mh_population p{R, mo_.data(), so_.data(), pos_.data(), vel_.data(), mass_.data(),
                    mid_.data(), pairs_.data(), lbar_.data(), smat_.data()};
    check(mh_sim_set_population(sim_, &p));
    std::vector<ElementTracker> t(R);
    for (int r = 0; r < R; r++)
      t[r] = ElementTracker{static_cast<uint32_t>(r), mo_[r], mo_[r + 1], so_[r], so_[r + 1]};
    return t;
  }

  void Simulate(float sim_duration, bool trackStresses = false) {
    check(mh_sim_simulate(sim_, sim_duration, trackStresses ? 1 : 0));
  }

  Element Collect(const ElementTracker& t) { return Collect(std::vector<ElementTracker>{t})[0]; }

  std::vector<Element> Collect(const std::vector<ElementTracker>& trackers) {
    check(mh_sim_collect(sim_, pos_.data(), vel_.data(), nullptr));
    std::vector<Element> out;
    out.reserve(trackers.size());
    for (const ElementTracker& t : trackers) {
      Element e;
      e.masses = elements_.at(t.ID).masses;
      e.springs = elements_.at(t.ID).springs;
      for (size_t i = 0; i < e.masses.size(); i++)
        for (int c = 0; c < 3; c++) {
          e.masses[i].pos[c] = pos_[3 * (t.mass_begin + i) + c];
          e.masses[i].vel[c] = vel_[3 * (t.mass_begin + i) + c];
        }
      out.push_back(std::move(e));
    }
    return out;
  }

  // Accumulated spring stresses (Simulate(..., trackStresses = true)).
  std::vector<float> Stresses() {
    std::vector<float> st(so_.empty() ? 0 : so_.back());
    check(mh_sim_collect(sim_, nullptr, nullptr, st.data()));
    return st;
  }

  // Per-robot fitness (MH_FITNESS_REFERENCE or MH_FITNESS_DISPLACEMENT).
  std::vector<float> Fitness(int mode = MH_FITNESS_REFERENCE) {
    std::vector<float> f(elements_.size());
    if (!f.empty()) check(mh_sim_fitness(sim_, mode, f.data()));
    return f;
  }

  float getDeltaT() const { return sim_ ? mh_sim_get_time_step(sim_) : cfg_.dt; }
  void setTimeStep(float dt) {
    cfg_.dt = dt;
    if (sim_) check(mh_sim_set_time_step(sim_, dt));
  }
  float getTotalTime() const { return sim_ ? mh_sim_total_time(sim_) : 0.f; }
  void Reset() {
    if (sim_) check(mh_sim_reset(sim_));
  }
  mh_sim* handle() { return sim_; }

 private:
  const MaterialSet& set_;
  mh_config cfg_{};
  mh_sim* sim_ = nullptr;
  std::vector<Element> elements_;
  std::vector<uint32_t> mo_, so_;
  std::vector<float> pos_, vel_, mass_, lbar_;
  std::vector<uint8_t> mid_, smat_;
  std::vector<uint16_t> pairs_;
};

}  // namespace mercury_hip

#endif  // MERCURY_HIP_SIM_HPP
