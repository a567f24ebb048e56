/*
 * materials.c -- oracle restatement of the reference material tables.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows common/simulator/material.h:
 *   base set            111-133  (floats initialised from double expressions)
 *   get_composite_id    192-213  (diagonal ids first, then pairs i<j)
 *   avg                 215-227  ((float sum) / 2.0 in double, OR of encodings)
 *   id_lookup           253-273
 *   decode              275-299
 *   encodedCompositeIdx 301-322  (two lowest set bits; bit 0 = air)
 * Pinned by: tests/golden/materials_reference.json (written by the compiled
 * reference header, oracle/ref/ref_dump.cpp), the four TestMatEncoding known
 * answers (common/tests/test_simulator.cpp:65-81), composite count 254.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

static orc_material mk(uint32_t id, float k, float dL0, double omega,
                       double phi) {
  orc_material m;
  m.id = id;
  m.k = k;
  m.dL0 = dL0;
  m.omega = (float)omega;
  m.phi = (float)phi;
  m.encoding = 1u << id;
  return m;
}

int orc_reference_base(orc_material* out) {
  /* OMEGA is (float)1.0f * (2.0f*M_PI): a double expression. */
  const double W = 1.0 * (2.0 * M_PI);
  const float A = 0.14f;
  int i = 0;
  out[i] = mk(0, 0, 0, 0, 0); i++;
  /* adductor muscles 0..9 (ids 1..10) */
  const double aphi[10] = {0, M_PI / 8, M_PI / 4, 3 * M_PI / 8, M_PI / 2,
                           5 * M_PI / 8, 3 * M_PI / 4, 7 * M_PI / 8, 0, 0};
  const double aw[10] = {W, W, W, W, W, W, W, W, W * 2, W / 2};
  for (int j = 0; j < 10; j++, i++) out[i] = mk(i, 50, A, aw[j], aphi[j]);
  /* abductor muscles 0..9 (ids 11..20) */
  const double bphi[10] = {M_PI, 9 * M_PI / 8, 5 * M_PI / 4, 11 * M_PI / 8,
                           3 * M_PI / 2, 13 * M_PI / 8, 7 * M_PI / 4,
                           15 * M_PI / 8, M_PI, M_PI};
  for (int j = 0; j < 10; j++, i++) out[i] = mk(i, 50, A, aw[j], bphi[j]);
  out[i] = mk(21, 40, 0, W, 0); i++;  /* tissue */
  out[i] = mk(22, 100, 0, W, 0); i++; /* bone */
  return i;
}

uint32_t orc_composite_id(int n_total, uint32_t id1, uint32_t id2) {
  if (id1 == 0 || id2 == 0) return 0;
  int nb = n_total - 1;
  uint32_t a = id1 - 1, b = id2 - 1;
  if (a == b) return a + 1;
  if (a > b) { uint32_t t = a; a = b; b = t; }
  /* ids nb+1.. enumerate pairs (a<b) row by row */
  uint32_t id = (uint32_t)nb + 1;
  for (uint32_t i = 0; i < (uint32_t)nb; i++)
    for (uint32_t j = i + 1; j < (uint32_t)nb; j++) {
      if (i == a && j == b) return id;
      id++;
    }
  return 0;
}

static orc_material avg2(int n_total, orc_material m1, orc_material m2) {
  orc_material r;
  r.id = orc_composite_id(n_total, m1.id, m2.id);
  float sk = m1.k + m2.k, sd = m1.dL0 + m2.dL0, so = m1.omega + m2.omega,
        sp = m1.phi + m2.phi;
  r.k = (float)((double)sk / 2.0);
  r.dL0 = (float)((double)sd / 2.0);
  r.omega = (float)((double)so / 2.0);
  r.phi = (float)((double)sp / 2.0);
  r.encoding = m1.encoding | m2.encoding;
  return r;
}

int orc_encoded_idx(uint32_t encoding, int composite_count) {
  unsigned idx[2] = {0, 0}, count = 0, bitmask = 1u;
  for (int i = 0; i < composite_count; i++) {
    if (encoding & bitmask) {
      idx[count] = (unsigned)i;
      count++;
      if (i == 0 || count == 2) break;
    }
    bitmask <<= 1;
  }
  if (idx[0] == 0) return 0;
  if (idx[1] == 0) return (int)(1 + idx[0] * (idx[0] - 1) / 2);
  return (int)(1 + idx[1] * (idx[1] - 1) / 2 + idx[0]);
}

int orc_composite_tables(const orc_material* base, int n_total,
                         orc_material* by_id, orc_material* by_enc) {
  int cc = (n_total - 1) * n_total / 2 + 1;
  memset(by_id, 0, sizeof(orc_material) * (size_t)cc);
  memset(by_enc, 0, sizeof(orc_material) * (size_t)cc);
  /* id_lookup */
  for (int i = 0; i < n_total; i++) by_id[base[i].id] = base[i];
  for (int i = 1; i < n_total; i++)
    for (int j = i + 1; j < n_total; j++) {
      orc_material m = avg2(n_total, base[i], base[j]);
      by_id[m.id] = m;
    }
  /* decode */
  by_enc[0] = base[0];
  for (int i = 1; i < n_total; i++) {
    by_enc[i * (i - 1) / 2 + 1] = base[i];
    for (int j = i + 1; j < n_total; j++)
      by_enc[j * (j - 1) / 2 + i + 1] = avg2(n_total, base[i], base[j]);
  }
  return cc;
}
