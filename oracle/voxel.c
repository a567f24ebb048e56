/*
 * voxel.c -- oracle restatement of the voxel-robot topology build.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows common/evolvables/VoxelRobot.cpp (recursive, like the reference):
 *   ShiftY 30-42, ShiftX 44-56, BuildSpringsRecurse 60-193,
 *   ConnectedVoxelRecurse 196-232, CopyFromVoxelRecurse 234-265,
 *   Strip 267-283, Build 310-331, Initialize 352-381 (voxel base positions)
 * and VoxelRobot.h:113-126 (index <-> (x,y,z)).
 * Lengths follow Eigen 3.4's Vector3f::norm(): sqrtf(x*x + (y*y + z*z))
 * (fixed-size redux splits 3 = 1 + 2).
 * Pinned by: tests/golden/voxel_*.bin written by the compiled reference
 * (oracle/ref/ref_dump.cpp builds VoxelRobot::Build from the reference
 * sources), spring-count invariants (n=5 -> 1036, n=9 -> 7448,
 * n=12 -> 18788) and per-direction counts.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct { int x, y, z; } bidx;

typedef struct {
  int nx, ny, nz;
  const orc_material* by_enc;
  int cc;
  orc_material* mat;  /* per voxel */
  float (*base)[3];   /* per voxel */
  uint8_t* visit;
  uint16_t* pairs;
  float* len;
  uint32_t* sid;
  uint32_t* senc;
  int ns, max_springs, overflow;
} vctx;

static int valid(const vctx* c, bidx i) {
  return i.x < c->nx && i.x >= 0 && i.y < c->ny && i.y >= 0 && i.z < c->nz &&
         i.z >= 0;
}
static int vid(const vctx* c, bidx i) {
  return i.x + i.y * c->nx + i.z * (c->nx * c->ny);
}
static bidx basis(const vctx* c, int idx) {
  bidx b;
  b.x = idx % c->nx;
  b.y = (idx / c->nx) % c->ny;
  b.z = (idx / c->nx / c->ny) % c->nz;
  return b;
}
static bidx add(bidx a, bidx b) { bidx r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static bidx sub(bidx a, bidx b) { bidx r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }

static int is_air(const orc_material* m) {
  /* Material::operator== compares k, dL0, omega, phi (material.h:51-53) */
  return m->k == 0.0f && m->dL0 == 0.0f && m->omega == 0.0f && m->phi == 0.0f;
}
static orc_material decode(const vctx* c, uint32_t enc) {
  return c->by_enc[orc_encoded_idx(enc, c->cc)];
}

static void spring_rec(vctx* c, bidx ind, int src) {
  if (!valid(c, ind)) return;
  int v = vid(c, ind);
  bidx sind = basis(c, src);
  if (src != v) {
    float dx = c->base[src][0] - c->base[v][0];
    float dy = c->base[src][1] - c->base[v][1];
    float dz = c->base[src][2] - c->base[v][2];
    float yz = dy * dy + dz * dz;
    float L = sqrtf(dx * dx + yz);
    bidx path = sub(ind, sind);
    int l1 = abs(path.x) + abs(path.y) + abs(path.z);
    orc_material mat;
    uint32_t enc = 0;
    const orc_material* vs = &c->mat[src];
    if (l1 == 1) {
      bidx nb[4];
      for (int i = 0; i < 4; i++) {
        bidx np;
        if (abs(path.x) == 1) { np.x = 0; np.y = (i / 2) % 2; np.z = i % 2; }
        else if (abs(path.y) == 1) { np.x = (i / 2) % 2; np.y = 0; np.z = i % 2; }
        else { np.x = (i / 2) % 2; np.y = i % 2; np.z = 0; }
        nb[i] = sub(sind, np);
      }
      for (int i = 0; i < 4; i++) {
        if (!valid(c, nb[i])) continue;
        const orc_material* n = &c->mat[vid(c, nb[i])];
        if (n->encoding != 1u) enc |= n->encoding;
      }
      mat = decode(c, enc);
    } else if (l1 == 2) {
      bidx np = {!path.x, !path.y, !path.z};
      if (path.x >= 0 && path.y >= 0 && path.z >= 0) {
        bidx ni = sub(sind, np);
        if (!valid(c, ni)) {
          mat = *vs;
        } else {
          const orc_material* n = &c->mat[vid(c, ni)];
          if (n->encoding != 1u) enc |= n->encoding;
          if (vs->encoding != 1u) enc |= vs->encoding;
          mat = decode(c, enc);
        }
      } else {
        bidx sp = {path.x == -1, path.y == -1, path.z == -1};
        bidx ms = sub(sind, sp);
        const orc_material* vm = &c->mat[vid(c, ms)];
        bidx ni = sub(ms, np);
        if (!valid(c, ni)) {
          mat = *vm;
        } else {
          const orc_material* n = &c->mat[vid(c, ni)];
          if (n->encoding != 1u) enc |= n->encoding;
          if (vm->encoding != 1u) enc |= vm->encoding;
          mat = decode(c, enc);
        }
      }
    } else {
      if (path.x >= 0 && path.y >= 0 && path.z >= 0) {
        mat = *vs;
      } else {
        bidx sp = {path.x == -1, path.y == -1, path.z == -1};
        mat = c->mat[vid(c, sub(sind, sp))];
      }
    }
    if (c->ns < c->max_springs) {
      c->pairs[2 * c->ns] = (uint16_t)src;
      c->pairs[2 * c->ns + 1] = (uint16_t)v;
      c->len[c->ns] = L;
      c->sid[c->ns] = mat.id;
      c->senc[c->ns] = mat.encoding;
    } else {
      c->overflow = 1;
    }
    c->ns++;
  }
  if (c->visit[v]) return;
  c->visit[v] = 1;
  for (int k = 1; k < 8; k++) {
    bidx o = {k % 2, (k / 2) % 2, (k / 4) % 2};
    spring_rec(c, add(ind, o), v);
  }
  int x = 1, y = -1, z = 0, tx, ty;
  for (int k = 0; k < 3; k++) {
    bidx o = {x, y, z};
    spring_rec(c, add(ind, o), v);
    tx = x; ty = y; x = z; y = tx; z = ty;
  }
  x = -1; y = 1; z = 1;
  for (int k = 0; k < 3; k++) {
    bidx o = {x, y, z};
    spring_rec(c, add(ind, o), v);
    tx = x; ty = y; x = z; y = tx; z = ty;
  }
}

static int connected_rec(vctx* c, bidx ind) {
  if (!valid(c, ind)) return 0;
  int idx = vid(c, ind);
  if (c->visit[idx]) return 0;
  c->visit[idx] = 1;
  if (is_air(&c->mat[idx])) return 0;
  int n = 1;
  int x = 1, y = 0, z = 0, tx, ty;
  for (int k = 0; k < 3; k++) {
    bidx p = {x, y, z}, q = {-x, -y, -z};
    n += connected_rec(c, add(ind, p));
    n += connected_rec(c, add(ind, q));
    tx = x; ty = y; x = z; y = tx; z = ty;
  }
  return n;
}

static void copy_rec(vctx* c, bidx ind, orc_material* dst) {
  if (!valid(c, ind)) return;
  int idx = vid(c, ind);
  if (c->visit[idx]) return;
  c->visit[idx] = 1;
  if (is_air(&c->mat[idx])) return;
  dst[idx] = c->mat[idx];
  int x = 1, y = 0, z = 0, tx, ty;
  for (int k = 0; k < 3; k++) {
    bidx p = {x, y, z}, q = {-x, -y, -z};
    copy_rec(c, add(ind, p), dst);
    copy_rec(c, add(ind, q), dst);
    tx = x; ty = y; x = z; y = tx; z = ty;
  }
}

int orc_voxel_build(int nx, int ny, int nz, float resolution,
                    const uint32_t* voxel_enc, const orc_material* by_enc,
                    int composite_count, float* mass_pos, uint32_t* mass_id,
                    uint16_t* pairs, float* spring_len, uint32_t* spring_id,
                    uint32_t* spring_enc, int max_springs,
                    uint32_t* stripped_enc) {
  vctx c;
  memset(&c, 0, sizeof c);
  c.nx = nx; c.ny = ny; c.nz = nz;
  c.by_enc = by_enc;
  c.cc = composite_count;
  int n = nx * ny * nz;
  c.mat = (orc_material*)malloc(sizeof(orc_material) * (size_t)n);
  c.base = (float(*)[3])malloc(sizeof(float) * 3 * (size_t)n);
  c.visit = (uint8_t*)calloc((size_t)n, 1);
  orc_material* stripped = (orc_material*)malloc(sizeof(orc_material) * (size_t)n);
  float inv = 1.0f / resolution;
  for (int i = 0; i < n; i++) {
    bidx b = basis(&c, i);
    c.base[i][0] = inv * (float)b.x;
    c.base[i][1] = inv * (float)b.y;
    c.base[i][2] = inv * (float)b.z;
    c.mat[i] = decode(&c, voxel_enc[i]);
  }
  /* Strip: keep the first largest face-connected non-air component. */
  int* cnt = (int*)calloc((size_t)n, sizeof(int));
  int maxi = 0;
  for (int i = 0; i < n; i++) {
    cnt[i] = connected_rec(&c, basis(&c, i));
    if (cnt[i] > cnt[maxi]) maxi = i;
  }
  for (int i = 0; i < n; i++) stripped[i] = by_enc[0]; /* air */
  memset(c.visit, 0, (size_t)n);
  copy_rec(&c, basis(&c, maxi), stripped);
  memcpy(c.mat, stripped, sizeof(orc_material) * (size_t)n);
  if (stripped_enc)
    for (int i = 0; i < n; i++) stripped_enc[i] = c.mat[i].encoding;
  /* masses */
  for (int i = 0; i < n; i++) {
    mass_pos[3 * i] = c.base[i][0];
    mass_pos[3 * i + 1] = c.base[i][1];
    mass_pos[3 * i + 2] = c.base[i][2];
    mass_id[i] = c.mat[i].id;
  }
  /* springs: DFS from (0,0,0) */
  memset(c.visit, 0, (size_t)n);
  c.pairs = pairs; c.len = spring_len; c.sid = spring_id; c.senc = spring_enc;
  c.max_springs = max_springs;
  bidx root = {0, 0, 0};
  spring_rec(&c, root, 0);
  /* ShiftX: translate the max x (over all masses) to 0 */
  int set = 0;
  float maxX = 0;
  for (int i = 0; i < n; i++)
    if (!set || mass_pos[3 * i] > maxX) { set = 1; maxX = mass_pos[3 * i]; }
  for (int i = 0; i < n; i++) mass_pos[3 * i] = mass_pos[3 * i] + -maxX;
  /* ShiftY: translate the min y to 0 */
  set = 0;
  float minY = -1;
  for (int i = 0; i < n; i++)
    if (!set || mass_pos[3 * i + 1] < minY) { set = 1; minY = mass_pos[3 * i + 1]; }
  for (int i = 0; i < n; i++) mass_pos[3 * i + 1] = mass_pos[3 * i + 1] + -minY;
  free(cnt); free(stripped); free(c.visit); free(c.base); free(c.mat);
  return c.overflow ? -1 : c.ns;
}
