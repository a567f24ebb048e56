// materials.cpp -- composite material tables.
//
// Same rules as the reference's materials class (material.h:109-322):
//   * base material i has id i and encoding bit i; bit 0 is air;
//   * a composite of bases i < j gets id n_base + 1 + (rank of the pair
//     (i-1, j-1) among all pairs a < b in row-major order), a single base
//     keeps its id, anything with air is air (get_composite_id, 192-213);
//   * composite parameters are (float sum) / 2.0 evaluated in double and
//     rounded to float (avg, 215-227);
//   * decode(encoding) uses the two lowest set bits, bit 0 short-circuits to
//     air (encodedCompositeIdx, 301-322), and indexes a table laid out by
//     1 + j(j-1)/2 + i for bits i <= j (decode, 275-299).
#include <cmath>
#include <memory>

#include "mh_internal.h"

namespace mh {

namespace {

uint32_t pair_id(int n_base, uint32_t id1, uint32_t id2) {
  if (id1 == 0 || id2 == 0) return 0;
  if (id1 == id2) return id1;
  uint32_t a = std::min(id1, id2) - 1, b = std::max(id1, id2) - 1;
  // pairs (p, q), p < q < n_base, row-major: rows 0..a-1 hold
  // sum_{p<a} (n_base - 1 - p) entries, then q - a - 1 within row a
  uint32_t before = a * (2u * n_base - a - 1u) / 2u;
  return n_base + 1u + before + (b - a - 1u);
}

Material average(int n_base, const Material& m1, const Material& m2) {
  Material r;
  r.id = pair_id(n_base, m1.id, m2.id);
  const float sk = m1.k + m2.k;
  const float sd = m1.dL0 + m2.dL0;
  const float so = m1.omega + m2.omega;
  const float sp = m1.phi + m2.phi;
  r.k = static_cast<float>(static_cast<double>(sk) / 2.0);
  r.dL0 = static_cast<float>(static_cast<double>(sd) / 2.0);
  r.omega = static_cast<float>(static_cast<double>(so) / 2.0);
  r.phi = static_cast<float>(static_cast<double>(sp) / 2.0);
  r.encoding = m1.encoding | m2.encoding;
  return r;
}

}  // namespace

std::vector<Material> reference_base() {
  // material.h:9-10, 111-133; OMEGA is the double 2*pi, stored as float.
  const double w = 1.0 * (2.0 * M_PI);
  const float amp = 0.14f;
  std::vector<Material> b(23);
  auto set = [&](int i, float k, float a, double om, double ph) {
    b[i].id = static_cast<uint32_t>(i);
    b[i].k = k;
    b[i].dL0 = a;
    b[i].omega = static_cast<float>(om);
    b[i].phi = static_cast<float>(ph);
    b[i].encoding = 1u << i;
  };
  set(0, 0.f, 0.f, 0.0, 0.0);
  const double add_phase[10] = {0.0, M_PI / 8, M_PI / 4, 3 * M_PI / 8, M_PI / 2,
                                5 * M_PI / 8, 3 * M_PI / 4, 7 * M_PI / 8, 0.0, 0.0};
  const double ab_phase[10] = {M_PI, 9 * M_PI / 8, 5 * M_PI / 4, 11 * M_PI / 8, 3 * M_PI / 2,
                               13 * M_PI / 8, 7 * M_PI / 4, 15 * M_PI / 8, M_PI, M_PI};
  const double freq[10] = {w, w, w, w, w, w, w, w, w * 2, w / 2};
  for (int m = 0; m < 10; m++) {
    set(1 + m, 50.f, amp, freq[m], add_phase[m]);
    set(11 + m, 50.f, amp, freq[m], ab_phase[m]);
  }
  set(21, 40.f, 0.f, w, 0.0);   // tissue
  set(22, 100.f, 0.f, w, 0.0);  // bone
  return b;
}

std::unique_ptr<mh_material_set> make_material_set(const std::vector<Material>& base) {
  const int n_total = static_cast<int>(base.size());
  if (n_total < 1 || n_total > 32)
    throw Error(MH_ERR_RANGE, "material set needs 1..31 non-air base materials");
  const int nb = n_total - 1;
  auto s = std::make_unique<mh_material_set>();
  s->base = base;
  s->composite_count = nb * n_total / 2 + 1;
  if (s->composite_count > 256)
    throw Error(MH_ERR_RANGE, "more than 256 composite materials (ids are 8-bit)");
  s->by_id.assign(s->composite_count, Material{});
  s->by_enc.assign(s->composite_count, Material{});
  for (const Material& m : base) s->by_id[m.id] = m;
  s->by_enc[0] = base[0];
  for (int j = 1; j < n_total; j++) {
    s->by_enc[1 + j * (j - 1) / 2] = base[j];
    for (int i = 1; i < j; i++) {
      Material m = average(nb, base[i], base[j]);
      s->by_id[m.id] = m;
      s->by_enc[1 + j * (j - 1) / 2 + i] = m;
    }
  }
  return s;
}

}  // namespace mh

int mh_material_set::encoded_index(uint32_t enc) const {
  // lowest and second-lowest set bit, limited to the 32 bits of the word
  if (enc == 0 || (enc & 1u)) return 0;
  const int lo = __builtin_ctz(enc);
  const uint32_t rest = enc & (enc - 1u);
  if (rest == 0) return 1 + lo * (lo - 1) / 2;
  const int hi = __builtin_ctz(rest);
  return 1 + hi * (hi - 1) / 2 + lo;
}
