/*
 * clay_oracle_common.c -- precision-independent parts of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see clay_oracle.h).  Only tests/, bench.py's
 * cpu_baseline leg and __graft_entry__.smoke() load this code.
 */
#include "clay_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* hash1's integer core (lib/src/math/gpu.ts:103-112): the "lowbias32"
 * multiply-xorshift.  Pinned by the published constants and by bijectivity
 * (every step is invertible on u32), checked in tests/test_oracle_reference.py. */
uint32_t orc_hash_u32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

/* hash1 -> [0,1) (math/gpu.ts:103-112): 24 low bits times 2^-24, exact in f32. */
float orc_hash1(uint32_t seed) {
  return (float)(orc_hash_u32(seed) & 0xffffffu) * (1.0f / 16777216.0f);
}

/* SdfVolume mip-chain derivation and resolution check (lib/src/field/volume.ts:139-179).
 * Pinned by the reference's own GPU test volume.gpu.ts:28-67 (160 -> 3 levels with
 * tiles 20/10/5, 40 -> 1, 64/128/256 -> 3/4/5, and both error messages). */
int orc_volume_derive(int resolution, int mips_req, int* mips_out, int* tile_res_out,
                      int* cell_off_out, int* total_cells_out, char* err, int errlen) {
  const int TILE = 8;
  int halvings = 0;
  double n = (double)resolution / TILE;
  while (n == floor(n) && fmod(n, 2.0) == 0.0 && n != 0.0) {
    halvings++;
    n /= 2.0;
  }
  int mips = mips_req;
  if (mips <= 0) {
    double l = log2((double)resolution / TILE);
    int li = (int)l; /* JS `| 0` truncates toward zero */
    int cap = halvings + 1;
    mips = li < cap ? li : cap;
    if (mips < 1) mips = 1;
  }
  long long need = (long long)TILE << (mips - 1);
  if (resolution % need != 0) {
    if (err && errlen > 0)
      snprintf(err, (size_t)errlen,
               "SdfVolume: resolution %d must be a multiple of %lld for %d mip levels",
               resolution, need, mips);
    return -1;
  }
  int tile_res = resolution / TILE;
  int off = 0;
  for (int m = 0; m < mips; m++) {
    int tr = tile_res >> m;
    if (tile_res_out) tile_res_out[m] = tr;
    if (cell_off_out) cell_off_out[m] = off;
    off += tr * tr * tr;
  }
  if (mips_out) *mips_out = mips;
  if (total_cells_out) *total_cells_out = off;
  return 0;
}

/* FluidSim constructor arithmetic (lib/src/sim/fluid.ts:141-203, 239-247, 444-447),
 * all in f64 as JS does it.  The rest-density lattice sum is pinned by the
 * reference's own JS mirror test (fluid.test.mjs:115-145), ported in
 * tests/test_oracle_reference.py. */
void orc_fluid_derive(const orc_fluid_opts* o, orc_fluid_k* k) {
  memset(k, 0, sizeof(*k));
  const double spacing = o->spacing;
  const double h = o->smoothing > 0 ? o->smoothing : spacing * 2.2;
  const double density = o->density > 0 ? o->density : 1.0;
  k->h = h;
  k->bucket = o->bucket > 0 ? o->bucket : 64;
  k->radius = o->radius >= 0 ? o->radius : spacing * 0.5;
  k->substeps = o->substeps > 0 ? o->substeps : 3;
  k->poly6 = 315.0 / (64.0 * M_PI * pow(h, 9));
  k->spiky_grad = -45.0 / (M_PI * pow(h, 6));
  k->mass = density * pow(spacing, 3);
  double rest = 0.0;
  const int reach = (int)ceil(h / spacing);
  for (int z = -reach; z <= reach; z++)
    for (int y = -reach; y <= reach; y++)
      for (int x = -reach; x <= reach; x++) {
        const double r2 = (double)(x * x + y * y + z * z) * spacing * spacing;
        if (r2 < h * h) rest += k->mass * k->poly6 * pow(h * h - r2, 3);
      }
  k->rest_density = rest;
  k->dt = 1.0 / 60.0 / k->substeps;
  k->contact_depth = k->radius * 1.5;
  k->grid_res = (int)ceil(o->world_size / h) + 2;
  for (int c = 0; c < 3; c++) {
    k->grid_origin[c] = o->origin[c] - h;
    k->wall_lo[c] = o->origin[c] + k->radius;
    k->wall_hi[c] = o->origin[c] + o->world_size - k->radius;
    k->gravity[c] = o->gravity[c];
  }
  k->stiffness = o->stiffness;
  k->viscosity = o->viscosity;
  k->damping = o->damping;
  k->max_speed = o->max_speed;
  k->hh = h * h;
  k->mass_poly6 = k->mass * k->poly6;
  k->h6 = pow(h, 6);
  k->inv_h = 1.0 / h;
  k->rest_floor = rest * 0.1;
}

/* packTile/unpackTile (lib/src/field/tilegrid.ts:6-14); pinned by march.test.mjs:176-182. */
uint32_t orc_pack_tile(uint32_t x, uint32_t y, uint32_t z) { return x | (y << 10) | (z << 20); }
void orc_unpack_tile(uint32_t v, uint32_t* xyz) {
  xyz[0] = v & 1023u;
  xyz[1] = (v >> 10) & 1023u;
  xyz[2] = (v >> 20) & 1023u;
}

/* IEEE binary32 -> binary16, round to nearest even (the `rgba16float` store).
 * Pinned against numpy's float16 conversion over all f16 boundaries in tests. */
uint16_t orc_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t exp = (x >> 23) & 0xffu;
  uint32_t mant = x & 0x7fffffu;
  if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
  const int e = (int)exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    const int shift = 14 - e;
    uint32_t hm = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (hm & 1u))) hm++;
    return (uint16_t)(sign | hm);
  }
  uint32_t hm = mant >> 13;
  const uint32_t rem = mant & 0x1fffu;
  uint32_t h = sign | ((uint32_t)e << 10) | hm;
  if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) h++;
  return (uint16_t)h;
}

float orc_f16_to_f32(uint16_t h) {
  const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  const uint32_t exp = ((uint32_t)h >> 10) & 0x1fu;
  uint32_t mant = (uint32_t)h & 0x3ffu;
  uint32_t x;
  if (exp == 0) {
    if (mant == 0) {
      x = sign;
    } else {
      int e = -1;
      do {
        e++;
        mant <<= 1;
      } while (!(mant & 0x400u));
      x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
    }
  } else if (exp == 31) {
    x = sign | 0x7f800000u | (mant << 13);
  } else {
    x = sign | ((exp - 15 + 127) << 23) | (mant << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

/* ENC (lib/src/sim/splat.ts:108-110, 247): u32((clamp(v,-1,1)+1)*32767.5).
 * Pinned by fluid.test.mjs:93-113 (ported). */
uint32_t orc_enc_f32(float v) {
  float c = v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v);
  return (uint32_t)((c + 1.0f) * 32767.5f);
}

/* Resolve decode (splat.ts:272): f32(e) * (1/32767.5) - 1. */
float orc_dec_f32(uint32_t e) { return (float)e * (float)(1.0 / 32767.5) - 1.0f; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Stable sort of particle ids by cell key: LSD radix sort, 8 bits per pass.
 * The result is the permutation "ascending key, ties by ascending id", which is
 * the build's definition of the bucket fill order (the reference's atomicAdd
 * order, fluid.ts:323, is unordered).  Pinned in tests against numpy's
 * argsort(kind="stable"). */
void orc_stable_sort_keys(int n, const uint32_t* keys, uint32_t* perm_out) {
  uint32_t* a = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  uint32_t* b = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) a[i] = (uint32_t)i;
  for (int pass = 0; pass < 4; pass++) {
    const int shift = pass * 8;
    size_t count[257];
    memset(count, 0, sizeof(count));
    for (int i = 0; i < n; i++) count[((keys[a[i]] >> shift) & 0xffu) + 1]++;
    for (int d = 0; d < 256; d++) count[d + 1] += count[d];
    for (int i = 0; i < n; i++) b[count[(keys[a[i]] >> shift) & 0xffu]++] = a[i];
    uint32_t* t = a;
    a = b;
    b = t;
  }
  memcpy(perm_out, a, sizeof(uint32_t) * (size_t)n);
  free(a);
  free(b);
}

/* TileGrid.dilate (lib/src/field/tilegrid.ts:207-242).  Parity: pinned against an
 * independent numpy formulation (upsampled mask, shifted-window maximum) in
 * tests/test_oracle_edit.py. */
void orc_tile_dilate(int res, int mips, int m, uint32_t* counts) {
  int tile_res[32], off[32], total = 0, nm = 0;
  char err[8];
  if (m <= 0 || orc_volume_derive(res, mips, &nm, tile_res, off, &total, err, (int)sizeof(err)) != 0 || m >= nm)
    return;
  const int tr = tile_res[m], ptr = tile_res[m - 1];
  for (int i = 0; i < tr * tr * tr; i++) {
    const int bx = (i % tr) * 2, by = ((i / tr) % tr) * 2, bz = (i / (tr * tr)) * 2;
    int live = 0;
    for (int dz = -1; dz <= 2 && !live; dz++)
      for (int dy = -1; dy <= 2 && !live; dy++)
        for (int dx = -1; dx <= 2 && !live; dx++) {
          const int qx = bx + dx, qy = by + dy, qz = bz + dz;
          if (qx < 0 || qy < 0 || qz < 0 || qx >= ptr || qy >= ptr || qz >= ptr) continue;
          if (counts[off[m - 1] + qx + ptr * (qy + ptr * qz)] != 0) live = 1;
        }
    if (live) counts[off[m] + i] = 1;
  }
}

/* Contacts -> erosion edits: the record BrushSet.make (brush.ts:539-564) builds
 * for {kind: 'sphere', size: radius, pos, op: 'cut', smooth: blend} -- scale 1,
 * identity rotation, bound = radius + blend, mask -1 -- for the first
 * min(raw_count, 32) contacts whose speed exceeds min_speed, in list order.
 * Pinned against clayhip.make_brush of the same description in
 * tests/test_oracle_edit.py. */
int orc_contacts_to_edits(uint32_t raw_count, const float* points4, double min_speed, double radius,
                          double blend, int max_edits, orc_brush* out) {
  const uint32_t n = raw_count < ORC_CONTACTS ? raw_count : ORC_CONTACTS;
  int k = 0;
  for (uint32_t i = 0; i < n && k < max_edits; i++) {
    if (!(points4[4 * i + 3] > (float)min_speed)) continue;
    orc_brush b;
    memset(&b, 0, sizeof(b));
    b.pos[0] = points4[4 * i];
    b.pos[1] = points4[4 * i + 1];
    b.pos[2] = points4[4 * i + 2];
    b.scale = 1.0f;
    b.rot[3] = 1.0f;
    b.size[0] = (float)radius;
    b.size[1] = (float)radius;
    b.size[2] = (float)radius;
    b.radius = 0.0f;
    b.material = 0.0f;
    b.blend = (float)blend;
    b.bound = (float)(radius + blend);
    b.mask = -1.0f;
    b.kind = 0;
    b.op = 1;
    out[k++] = b;
  }
  return k;
}
