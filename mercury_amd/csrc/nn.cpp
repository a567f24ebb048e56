// nn.cpp -- NN-robot producer: per-mass material from a small MLP and the
// KNN spring topology.
//
// Reference: NNRobot::forward (common/evolvables/NNRobot.h:41-92) -- a
// bias-free MLP, ReLU on hidden layers, material = index of the largest
// output (softmax is monotonic; the first maximum wins, as Eigen's
// maxCoeff); NNRobot::Build (NNRobot.cpp:282-301) -- one spring per
// (mass, neighbour) edge, skipped when shorter than min_length (1e-3 in the
// reference), material decode(enc_a | enc_b), rest = mean = distance.
// Matrix-vector products here are sequential float dot products in input
// order (Eigen's blocked GEMM order is not part of the reference tree).
#include <cmath>
#include <cstring>
#include <vector>

#include "mh_internal.h"

extern "C" {

int mh_nn_materials(int n_layers, const int* sizes, const float* weights, const float* pos, int n,
                    uint8_t* out_index) {
  try {
    if (n_layers < 1 || !sizes || !weights || (n > 0 && (!pos || !out_index)) || n < 0)
      throw mh::Error(MH_ERR_ARG, "null argument");
    if (sizes[0] != 3) throw mh::Error(MH_ERR_ARG, "input size must be 3 (position)");
    int widest = 0;
    for (int l = 0; l <= n_layers; l++) {
      if (sizes[l] < 1) throw mh::Error(MH_ERR_ARG, "layer sizes must be >= 1");
      widest = sizes[l] > widest ? sizes[l] : widest;
    }
    if (sizes[n_layers] > 255) throw mh::Error(MH_ERR_RANGE, "at most 255 outputs");
    std::vector<float> a(widest), b(widest);
    for (int i = 0; i < n; i++) {
      a[0] = pos[3 * i];
      a[1] = pos[3 * i + 1];
      a[2] = pos[3 * i + 2];
      const float* w = weights;
      for (int l = 0; l < n_layers; l++) {
        const int ni = sizes[l], no = sizes[l + 1];
        for (int o = 0; o < no; o++) {
          float s = 0.f;
          for (int k = 0; k < ni; k++) s = s + w[o * ni + k] * a[k];
          b[o] = (l < n_layers - 1 && s < 0.f) ? 0.f : s;  // ReLU on hidden layers
        }
        w += static_cast<size_t>(ni) * no;
        std::swap(a, b);
      }
      int best = 0;
      for (int o = 1; o < sizes[n_layers]; o++)
        if (a[o] > a[best]) best = o;
      out_index[i] = static_cast<uint8_t>(best);
    }
    return MH_OK;
  } catch (const mh::Error& e) {
    mh::set_error(e.what());
    return e.code;
  }
}

int mh_knn_springs(const mh_material_set* set, int n, int K, const uint16_t* nbr, const float* dist,
                   const uint32_t* mass_encoding, float min_length, uint16_t* pairs, float* length,
                   uint8_t* material_id, int max_springs, int* n_springs) {
  try {
    if (!set || n < 0 || K < 1 || (n > 0 && (!nbr || !dist || !mass_encoding || !pairs || !length ||
                                              !material_id)) || !n_springs)
      throw mh::Error(MH_ERR_ARG, "null argument");
    int ns = 0;
    for (int i = 0; i < n; i++)
      for (int r = 0; r < K; r++) {
        const float d = dist[static_cast<size_t>(i) * K + r];
        const uint16_t j = nbr[static_cast<size_t>(i) * K + r];
        // the reference skips d < 1e-3 (NNRobot.cpp:290); a non-finite
        // distance means "no K-th neighbour" (coincident masses) and is
        // skipped as well
        if (!(d >= min_length) || !std::isfinite(d) || j >= n) continue;
        if (ns >= max_springs) throw mh::Error(MH_ERR_RANGE, "max_springs too small");
        pairs[2 * ns] = static_cast<uint16_t>(i);
        pairs[2 * ns + 1] = j;
        length[ns] = d;
        material_id[ns] = static_cast<uint8_t>(set->decode(mass_encoding[i] | mass_encoding[j]).id);
        ns++;
      }
    *n_springs = ns;
    return MH_OK;
  } catch (const mh::Error& e) {
    mh::set_error(e.what());
    return e.code;
  }
}

}  // extern "C"
