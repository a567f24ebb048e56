// Evaluate-style use of the C++ shim (include/mercury_hip_sim.hpp): build
// voxel robots with the library's VoxelRobot::Build equivalent, evaluate
// them like Evaluator<T>::BatchEvaluate (Reset -> SetElements -> Simulate ->
// Collect -> fitness), and write the final positions for the parity test.
//   shim_demo <out.bin> <n_robots> <voxels> <steps_duration_s> <seed>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "mercury_hip_sim.hpp"

using namespace mercury_hip;

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  const char* out = argv[1];
  const int R = atoi(argv[2]), vox = atoi(argv[3]);
  const float duration = static_cast<float>(atof(argv[4]));
  const unsigned seed = static_cast<unsigned>(atoi(argv[5]));
  const int n = vox + 1, nv = n * n * n;
  // stiff, soft and two actuated phases (k, dL0, omega, phi)
  MaterialSet set({{1e4f, 0.f, 12.566371f, 0.f}, {2e3f, 0.f, 12.566371f, 0.f},
                   {5e3f, 0.14f, 12.566371f, 0.f}, {5e3f, 0.14f, 12.566371f, 3.1415927f}});
  std::mt19937 rng(seed);
  std::vector<Element> robots;
  for (int r = 0; r < R; r++) {
    std::vector<uint32_t> enc(nv);
    for (int i = 0; i < nv; i++) {
      const int x = i % n, y = (i / n) % n, z = i / (n * n);
      const bool edge = x == n - 1 || y == n - 1 || z == n - 1;
      const int m = static_cast<int>(rng() % 5);  // 0 = empty
      enc[i] = edge || m == 0 ? 1u : (1u << m);
    }
    std::vector<float> pos(3 * nv), len(mh_voxel_max_springs(n, n, n));
    std::vector<uint8_t> mid(nv), sid(len.size());
    std::vector<uint16_t> pairs(2 * len.size());
    int ns = 0;
    check(mh_voxel_build(set.get(), n, n, n, 10.f, enc.data(), pos.data(), mid.data(), pairs.data(),
                         len.data(), sid.data(), nullptr, static_cast<int>(len.size()), &ns, nullptr));
    Element e;
    for (int i = 0; i < nv; i++) {
      Mass m;
      m.id = static_cast<uint32_t>(i);
      for (int c = 0; c < 3; c++) m.pos[c] = m.protoPos[c] = pos[3 * i + c];
      m.material = set.by_id(mid[i]);
      m.mass = 0.001f;
      e.masses.push_back(m);
    }
    for (int k = 0; k < ns; k++) {
      Spring s;
      s.m0 = pairs[2 * k];
      s.m1 = pairs[2 * k + 1];
      s.rest_length = s.mean_length = len[k];
      s.material = set.by_id(sid[k]);
      e.springs.push_back(s);
    }
    robots.push_back(e);
  }
  mh_config cfg;
  mh_config_default(&cfg);
  cfg.solver = MH_SOLVER_FORCE;
  Simulator sim(set);
  sim.Initialize(cfg);
  sim.Reset();
  std::vector<ElementTracker> trackers = sim.SetElements(robots);
  sim.Simulate(duration);
  std::vector<Element> res = sim.Collect(trackers);
  std::vector<float> fit = sim.Fitness(MH_FITNESS_DISPLACEMENT);
  FILE* f = fopen(out, "wb");
  if (!f) return 3;
  for (const Element& e : res)
    for (const Mass& m : e.masses) fwrite(m.pos, sizeof(float), 3, f);
  fwrite(fit.data(), sizeof(float), fit.size(), f);
  fclose(f);
  printf("shim_demo: %d robots, t=%g s, fitness[0]=%g\n", R, sim.getTotalTime(), fit.empty() ? 0.f : fit[0]);
  return 0;
}
