/*
 * clay_oracle_impl.c -- precision-generic body of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see clay_oracle.h).  Compiled twice by
 * oracle/Makefile: -DORC_T=float -DORC_S=f32 and -DORC_T=double -DORC_S=f64,
 * always with -ffp-contract=off so every a*b+c rounds twice, as a WGSL
 * implementation without fused multiply-add does.
 *
 * Operation order follows the reference shader text left to right; vector
 * `length`/`dot` sum x, then y, then z; `pow(t, 3)` is (t*t)*t and `mix(a,b,t)`
 * is a*(1-t)+b*t (the WGSL definition).  Where the reference leaves a choice to
 * the WGSL implementation (pow lowering, length, hardware filter weights) the
 * choice made here is the one the HIP kernels make, and the functions carry
 * "parity unpinned" for that choice.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "clay_oracle.h"

#ifndef ORC_T
#define ORC_T float
#define ORC_S f32
#endif

typedef ORC_T T;
#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, ORC_S)

#define IS_F32 (sizeof(T) == 4)
static inline T t_sqrt(T x) { return IS_F32 ? (T)sqrtf((float)x) : (T)sqrt((double)x); }
static inline T t_floor(T x) { return IS_F32 ? (T)floorf((float)x) : (T)floor((double)x); }
static inline T t_fabs(T x) { return (T)fabs((double)x); }
static inline T t_min(T a, T b) { return b < a ? b : a; }
static inline T t_max(T a, T b) { return a < b ? b : a; }
static inline T t_clamp(T x, T lo, T hi) { return t_min(t_max(x, lo), hi); }
static inline T t_mix(T a, T b, T t) { return a * ((T)1 - t) + b * t; }

typedef struct { T x, y, z; } v3;
static inline v3 V(T x, T y, T z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmuls(v3 a, T s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vabs(v3 a) { return V(t_fabs(a.x), t_fabs(a.y), t_fabs(a.z)); }
static inline v3 vmax0(v3 a) { return V(t_max(a.x, 0), t_max(a.y, 0), t_max(a.z, 0)); }
static inline T vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline T vlen(v3 a) { return t_sqrt(vdot(a, a)); }
static inline v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* f32 -> i32 as the hardware converts: truncate, saturate, NaN -> 0. */
static inline int32_t sat_i32(T x) {
  if (x != x) return 0;
  if (x >= (T)2147483648.0) return INT32_MAX;
  if (x <= (T)-2147483648.0) return INT32_MIN;
  return (int32_t)x;
}
static inline T rnd(double d) { return (T)d; }
/* A shader literal: the f32 literal in f32 mode, the exact decimal in f64 mode. */
#define KF(x) (IS_F32 ? (T)(x##f) : (T)(x))

/* ---------------------------------------------------------------------------
 * Primitives and the brush fold (lib/src/field/brush.ts:238-349, 474-511;
 * lib/src/math/gpu.ts:21-25, 87-92, 132-137).  sdSphere/sdBox3d/sdCapsule are
 * Quilez's formulas (@typegpu/sdf@0.12.0, not vendored); pinned by the
 * reference's brush.gpu.ts:34-91 closed-form answers, ported in tests.
 * sdRoundedBox3d: parity unpinned (no reference test; Quilez's
 * q = |p| - b + r form is assumed).
 * ------------------------------------------------------------------------- */
static T sd_box(v3 p, v3 s) {
  v3 q = vsub(vabs(p), s);
  return vlen(vmax0(q)) + t_min(t_max(q.x, t_max(q.y, q.z)), 0);
}
static T sd_round_box(v3 p, v3 s, T r) {
  v3 q = V(t_fabs(p.x) - s.x + r, t_fabs(p.y) - s.y + r, t_fabs(p.z) - s.z + r);
  return vlen(vmax0(q)) + t_min(t_max(q.x, t_max(q.y, q.z)), 0) - r;
}
static T sd_capsule(v3 p, v3 a, v3 b, T r) {
  v3 pa = vsub(p, a), ba = vsub(b, a);
  T h = t_clamp(vdot(pa, ba) / vdot(ba, ba), 0, 1);
  return vlen(vsub(pa, vmuls(ba, h))) - r;
}

static T eval_local(uint32_t kind, v3 p, v3 size, T radius) {
  switch (kind) {
    case 0: return vlen(p) - size.x;                                  /* sphere */
    case 1: return sd_box(p, size);                                   /* box */
    case 2: return sd_round_box(p, size, radius);                     /* roundBox */
    case 3: return sd_capsule(p, V(0, -size.y, 0), V(0, size.y, 0), radius);
    case 4: {                                                         /* torus */
      T qx = t_sqrt(p.x * p.x + p.z * p.z) - size.x, qy = p.y;
      return t_sqrt(qx * qx + qy * qy) - radius;
    }
    case 5: {                                                         /* cylinder */
      T dx = t_sqrt(p.x * p.x + p.z * p.z) - size.x, dy = t_fabs(p.y) - size.y;
      T mx = t_max(dx, 0), my = t_max(dy, 0);
      return t_min(t_max(dx, dy), 0) + t_sqrt(mx * mx + my * my);
    }
    case 6: return p.y - radius;                                      /* plane */
    default: return (T)1e9;  /* kinds outside the collider slice read as empty */
  }
}

/* evalBrush (brush.ts:466-471) with quatRotate/quatConj (math/gpu.ts:15-25). */
static T eval_brush(const orc_brush* b, v3 p) {
  T inv_scale = (T)1 / t_max((T)b->scale, KF(1e-5));
  v3 q = V(-(T)b->rot[0], -(T)b->rot[1], -(T)b->rot[2]);
  T qw = (T)b->rot[3];
  v3 v = vsub(p, V(b->pos[0], b->pos[1], b->pos[2]));
  v3 t = vmuls(vcross(q, v), 2);
  v3 rot = vadd(vadd(v, vmuls(t, qw)), vcross(q, t));
  v3 local = vmuls(rot, inv_scale);
  return eval_local(b->kind, local, V(b->size[0], b->size[1], b->size[2]), (T)b->radius) *
         (T)b->scale;
}

static T smoothstep(T lo, T hi, T x) {
  T t = t_clamp((x - lo) / (hi - lo), 0, 1);
  return t * t * ((T)3 - (T)2 * t);
}

/* applyBrush (brush.ts:474-511), sminPolyWeighted/smaxPolyWeighted (math/gpu.ts:87-92,132-137). */
static void apply_brush(T acc[2], const orc_brush* b, v3 p) {
  T dist = eval_brush(b, p);
  T keep = (T)b->mask >= 0 ? (T)1 - smoothstep((T)0.25, (T)0.75, t_fabs(acc[1] - (T)b->mask)) : (T)1;
  T kk = t_max((T)b->blend, KF(1e-6));
  if (b->op == 0) {
    T h = t_clamp((T)0.5 + (T)0.5 * (acc[0] - dist) / kk, 0, 1);
    T rx = t_mix(acc[0], dist, h) - kk * h * ((T)1 - h);
    T ry = h;
    T nx = rx, ny = t_mix(acc[1], (T)b->material, ry);
    acc[0] = t_mix(acc[0], nx, keep);
    acc[1] = t_mix(acc[1], ny, keep);
  } else if (b->op == 1) {
    T a = -dist;
    T h = t_clamp((T)0.5 - (T)0.5 * (acc[0] - a) / kk, 0, 1);
    T rx = t_mix(acc[0], a, h) + kk * h * ((T)1 - h);
    T ry = h;
    T nx = rx, ny = t_mix(acc[1], (T)b->material, ry * KF(0.85));
    acc[0] = t_mix(acc[0], nx, keep);
    acc[1] = t_mix(acc[1], ny, keep);
  } else {
    T w = (T)1 - smoothstep(-t_max((T)b->blend, KF(1e-4)), 0, dist);
    acc[1] = t_mix(acc[1], (T)b->material, w * keep);
  }
}

/* ---------------------------------------------------------------------------
 * Collider (lib/src/trace/field.ts).  analyticField: sample/normal
 * (field.ts:196-236), the closure being a brush fold from (1e9, 0).
 * volumeField over SdfVolume (volume.ts:262-313, field.ts:102-156): trilinear
 * over f16 texels with clamp-to-edge.  The fold's filter weights are full f32
 * (hardware filter weights are implementation defined): parity unpinned
 * against a WebGPU sampler; HIP parity is bit-exact by construction.
 * ------------------------------------------------------------------------- */
static T analytic_dist(const orc_leaf* l, v3 p) {
  T acc[2] = {(T)1e9, 0};
  for (int i = 0; i < l->nbrush; i++) apply_brush(acc, &l->brushes[i], p);
  return acc[0];
}

typedef struct {
  T origin[3], extent, voxel, band, center[3], half, grad_h;
  int res;
} volp;

static volp vol_params(const orc_leaf* l) {
  volp v;
  double half = l->world_size / 2;
  for (int c = 0; c < 3; c++) {
    v.origin[c] = rnd(l->origin[c]);
    v.center[c] = rnd(l->origin[c] + half);
  }
  v.extent = rnd(l->world_size);
  v.voxel = rnd(l->world_size / l->res);
  v.band = rnd(l->band);
  v.half = rnd(half);
  v.grad_h = v.voxel * rnd(l->grad_voxels);
  v.res = l->res;
  return v;
}

static T texel_of(const uint16_t* plane, int n, int x, int y, int z) {
  x = x < 0 ? 0 : (x >= n ? n - 1 : x);
  y = y < 0 ? 0 : (y >= n ? n - 1 : y);
  z = z < 0 ? 0 : (z >= n ? n - 1 : z);
  return (T)orc_f16_to_f32(plane[(size_t)x + (size_t)n * ((size_t)y + (size_t)n * z)]);
}

#define texel(l, x, y, z) texel_of(plane, (l)->res, x, y, z)

/* textureSampleLevel(linear, clamp-to-edge) of one channel at mip 0. */
static T sample_plane(const orc_leaf* l, const uint16_t* plane, const volp* v, v3 p) {
  T u = (p.x - v->origin[0]) / v->extent;
  T w = (p.y - v->origin[1]) / v->extent;
  T s = (p.z - v->origin[2]) / v->extent;
  T fn = (T)v->res;
  T tx = u * fn - (T)0.5, ty = w * fn - (T)0.5, tz = s * fn - (T)0.5;
  T fx0 = t_floor(tx), fy0 = t_floor(ty), fz0 = t_floor(tz);
  int x0 = sat_i32(fx0), y0 = sat_i32(fy0), z0 = sat_i32(fz0);
  T fx = tx - fx0, fy = ty - fy0, fz = tz - fz0;
  T c00 = t_mix(texel(l, x0, y0, z0), texel(l, x0 + 1, y0, z0), fx);
  T c10 = t_mix(texel(l, x0, y0 + 1, z0), texel(l, x0 + 1, y0 + 1, z0), fx);
  T c01 = t_mix(texel(l, x0, y0, z0 + 1), texel(l, x0 + 1, y0, z0 + 1), fx);
  T c11 = t_mix(texel(l, x0, y0 + 1, z0 + 1), texel(l, x0 + 1, y0 + 1, z0 + 1), fx);
  T c0 = t_mix(c00, c10, fy);
  T c1 = t_mix(c01, c11, fy);
  return t_mix(c0, c1, fz);
}

#undef texel

static T sample_raw(const orc_leaf* l, const volp* v, v3 p) { return sample_plane(l, l->dist16, v, p); }

/* SdfVolume.sampleDist at mip 0 (volume.ts:277-286); equals volumeField.sample(p,0).x. */
static T vol_dist(const orc_leaf* l, const volp* v, v3 p) {
  v3 rel = V(p.x - v->center[0], p.y - v->center[1], p.z - v->center[2]);
  T dbox = sd_box(rel, V(v->half, v->half, v->half));
  if (dbox > 0) return dbox;
  T band_world = v->band * v->voxel * (T)1;  /* band * voxelSize * exp2(0) */
  return sample_raw(l, v, p) * band_world;
}

/* offsetField: the child reads p - offset (field.ts:405-427). */
static v3 leaf_point(const orc_leaf* l, v3 p) {
  return l->has_offset ? V(p.x - rnd(l->offset[0]), p.y - rnd(l->offset[1]), p.z - rnd(l->offset[2])) : p;
}

/* lerpField's k = clamp(t(), 0, 1) (field.ts:348). */
static T lerp_k(const orc_leaf* l) { return t_clamp(rnd(l->lerp_t), 0, 1); }

static T vol_leaf_dist(const orc_leaf* l, v3 p) {
  volp v = vol_params(l);
  return vol_dist(l, &v, p);
}

/* volumeField.normal (field.ts:178-180): central differences of sampleDist. */
static v3 vol_normal(const orc_leaf* l, v3 p) {
  volp v = vol_params(l);
  T h = v.grad_h;
  v3 g = V(vol_dist(l, &v, V(p.x + h, p.y, p.z)) - vol_dist(l, &v, V(p.x - h, p.y, p.z)),
           vol_dist(l, &v, V(p.x, p.y + h, p.z)) - vol_dist(l, &v, V(p.x, p.y - h, p.z)),
           vol_dist(l, &v, V(p.x, p.y, p.z + h)) - vol_dist(l, &v, V(p.x, p.y, p.z - h)));
  T len = vlen(g);
  return len > KF(1e-8) ? vmuls(g, (T)1 / len) : V(0, 1, 0);
}

static T leaf_dist(const orc_leaf* l, v3 p) {
  p = leaf_point(l, p);
  if (l->kind == 0) return analytic_dist(l, p);
  const T a = vol_leaf_dist(l, p);
  return l->kind == 1 ? a : t_mix(a, vol_leaf_dist(l->lerp_b, p), lerp_k(l));
}

static v3 leaf_normal(const orc_leaf* l, v3 p) {
  p = leaf_point(l, p);
  if (l->kind == 2) {
    /* normalize(mix(a.normal, b.normal, k)) (field.ts:379-383); a zero mix reads up. */
    const T k = lerp_k(l);
    const v3 na = vol_normal(l, p), nb = vol_normal(l->lerp_b, p);
    const v3 g = V(t_mix(na.x, nb.x, k), t_mix(na.y, nb.y, k), t_mix(na.z, nb.z, k));
    const T len = vlen(g);
    return len > KF(1e-9) ? vmuls(g, (T)1 / len) : V(0, 1, 0);
  }
  if (l->kind == 0) {
    T e = rnd(l->eps);
    v3 g = V(analytic_dist(l, V(p.x + e, p.y, p.z)) - analytic_dist(l, V(p.x - e, p.y, p.z)),
             analytic_dist(l, V(p.x, p.y + e, p.z)) - analytic_dist(l, V(p.x, p.y - e, p.z)),
             analytic_dist(l, V(p.x, p.y, p.z + e)) - analytic_dist(l, V(p.x, p.y, p.z - e)));
    T len = vlen(g);
    return len > KF(1e-9) ? vmuls(g, (T)1 / len) : V(0, 1, 0);
  }
  return vol_normal(l, p);
}

/* field(p, 0).y of one leaf: analyticField's constant material, or volumeField's
 * trilinear material at mip 0 and 0 outside the volume box (volume.ts:288-296). */
static T vol_material(const orc_leaf* l, v3 p);

static T leaf_material(const orc_leaf* l, v3 p) {
  p = leaf_point(l, p);
  if (l->kind == 0) return rnd(l->material);
  const T a = vol_material(l, p);
  return l->kind == 1 ? a : t_mix(a, vol_material(l->lerp_b, p), lerp_k(l));
}

static T vol_material(const orc_leaf* l, v3 p) {
  volp v = vol_params(l);
  v3 rel = V(p.x - v.center[0], p.y - v.center[1], p.z - v.center[2]);
  if (sd_box(rel, V(v.half, v.half, v.half)) > 0) return 0;
  if (!l->mat16) return 0;
  return sample_plane(l, l->mat16, &v, p);
}

/* unionField fold: keep the left child unless the right one is strictly nearer
 * (field.ts:264-313: select(sb, sa, sa.x < sb.x)). */
static int collider_pick(const orc_collider* c, v3 p, T* dist) {
  int best = 0;
  T bd = leaf_dist(&c->leaf[0], p);
  for (int i = 1; i < c->nleaf; i++) {
    T d = leaf_dist(&c->leaf[i], p);
    if (!(bd < d)) {
      best = i;
      bd = d;
    }
  }
  *dist = bd;
  return best;
}

void FN(orc_collider_eval)(const orc_collider* col, int n, const T* p3, T* dist, T* nrm3) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    v3 p = V(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]);
    T d;
    int leaf = collider_pick(col, p, &d);
    v3 nn = leaf_normal(&col->leaf[leaf], p);
    dist[i] = d;
    nrm3[3 * i] = nn.x;
    nrm3[3 * i + 1] = nn.y;
    nrm3[3 * i + 2] = nn.z;
  }
}

/* ---------------------------------------------------------------------------
 * FluidSim hash grid (fluid.ts:239-266, 308-331).  cellOf = floor((p - o) * (1/h))
 * as vec3i; inGrid = all in [0, R); key = x + R (y + R z).  The bucket keeps the
 * first `bucket` ids of each cell in ascending id order (one admissible outcome
 * of the reference's atomicAdd order).  Completeness of the 3x3x3 scan is pinned
 * by fluid.test.mjs:30-65 (ported with its LCG in tests).
 * ------------------------------------------------------------------------- */
typedef struct {
  T go[3], inv_h;
  int R, bucket;
  T hh, mass_poly6, rest, rest_floor;
  T h, h6, mass, spiky;
  T g[3], dt, vmax, stiff, visc, damp, cdepth;
  T radius, lo[3], hi[3];
} kk_t;

static kk_t kconv(const orc_fluid_k* k) {
  kk_t c;
  for (int i = 0; i < 3; i++) {
    c.go[i] = rnd(k->grid_origin[i]);
    c.g[i] = rnd(k->gravity[i]);
    c.lo[i] = rnd(k->wall_lo[i]);
    c.hi[i] = rnd(k->wall_hi[i]);
  }
  c.inv_h = rnd(k->inv_h);
  c.R = k->grid_res;
  c.bucket = k->bucket;
  c.hh = rnd(k->hh);
  c.mass_poly6 = rnd(k->mass_poly6);
  c.rest = rnd(k->rest_density);
  c.rest_floor = rnd(k->rest_floor);
  c.h = rnd(k->h);
  c.h6 = rnd(k->h6);
  c.mass = rnd(k->mass);
  c.spiky = rnd(k->spiky_grad);
  c.dt = rnd(k->dt);
  c.vmax = rnd(k->max_speed);
  c.stiff = rnd(k->stiffness);
  c.visc = rnd(k->viscosity);
  c.damp = rnd(k->damping);
  c.cdepth = rnd(k->contact_depth);
  c.radius = rnd(k->radius);
  return c;
}

static inline void cell_of(const kk_t* c, v3 p, int32_t q[3]) {
  q[0] = sat_i32(t_floor((p.x - c->go[0]) * c->inv_h));
  q[1] = sat_i32(t_floor((p.y - c->go[1]) * c->inv_h));
  q[2] = sat_i32(t_floor((p.z - c->go[2]) * c->inv_h));
}
static inline int in_grid(const kk_t* c, const int32_t q[3]) {
  return q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < c->R && q[1] < c->R && q[2] < c->R;
}
static inline uint32_t cell_index(const kk_t* c, const int32_t q[3]) {
  return (uint32_t)q[0] + (uint32_t)c->R * ((uint32_t)q[1] + (uint32_t)c->R * (uint32_t)q[2]);
}

typedef struct {
  int n;
  uint32_t* perm;   /* sorted slot -> particle id */
  uint32_t* skeys;  /* sorted keys */
} grid_t;

static void grid_build(const kk_t* c, int n, const T* px, const T* py, const T* pz,
                       uint32_t* keys, grid_t* g) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    int32_t q[3];
    cell_of(c, V(px[i], py[i], pz[i]), q);
    keys[i] = in_grid(c, q) ? cell_index(c, q) : ORC_INVALID_KEY;
  }
  g->n = n;
  g->perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  g->skeys = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  orc_stable_sort_keys(n, keys, g->perm);
  for (int i = 0; i < n; i++) g->skeys[i] = keys[g->perm[i]];
}
static void grid_free(grid_t* g) {
  free(g->perm);
  free(g->skeys);
}
/* First sorted slot with key >= k. */
static int lower(const grid_t* g, uint32_t k) {
  int lo = 0, hi = g->n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (g->skeys[mid] < k) lo = mid + 1; else hi = mid;
  }
  return lo;
}
static inline void cell_range(const kk_t* c, const grid_t* g, uint32_t key, int* s, int* e) {
  int a = lower(g, key);
  int b = a;
  while (b < g->n && g->skeys[b] == key) b++;
  if (b - a > c->bucket) b = a + c->bucket;
  *s = a;
  *e = b;
}

void FN(orc_keys)(const orc_fluid_k* k, int n, const T* px, const T* py, const T* pz,
                  uint32_t* keys) {
  kk_t c = kconv(k);
  for (int i = 0; i < n; i++) {
    int32_t q[3];
    cell_of(&c, V(px[i], py[i], pz[i]), q);
    keys[i] = in_grid(&c, q) ? cell_index(&c, q) : ORC_INVALID_KEY;
  }
}

/* Ids the 27-cell scan visits for particle i with r2 < h^2 (the set the density
 * pass sums over); returns how many, writing at most `cap`. */
int FN(orc_neighbors)(const orc_fluid_k* k, int n, const T* px, const T* py, const T* pz, int i,
                      uint32_t* out, int cap) {
  kk_t c = kconv(k);
  grid_t g;
  uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  grid_build(&c, n, px, py, pz, keys, &g);
  v3 pi = V(px[i], py[i], pz[i]);
  int32_t base[3];
  cell_of(&c, pi, base);
  int found = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int32_t q[3] = {base[0] + dx, base[1] + dy, base[2] + dz};
        if (!in_grid(&c, q)) continue;
        int s, e;
        cell_range(&c, &g, cell_index(&c, q), &s, &e);
        for (int t = s; t < e; t++) {
          uint32_t j = g.perm[t];
          v3 d = V(pi.x - px[j], pi.y - py[j], pi.z - pz[j]);
          if (vdot(d, d) < c.hh) {
            if (found < cap) out[found] = j;
            found++;
          }
        }
      }
  grid_free(&g);
  free(keys);
  return found;
}

#define ST(f) ((T*)st->f)

/* density pass (fluid.ts:333-368).  Summation order: dz, dy, dx ascending, then
 * the cell's ids ascending.  pow(t,3) as (t*t)*t: parity unpinned against WGSL
 * pow (exp2/log2 lowering); the lattice-at-rest identity rho = rho0 is checked. */
static void density_core(const kk_t* c, const grid_t* g, int n, orc_state* st) {
  const T* px = ST(px); const T* py = ST(py); const T* pz = ST(pz);
  T* rho_out = ST(rho); T* pres_out = ST(pres);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    v3 pi = V(px[i], py[i], pz[i]);
    int32_t base[3];
    cell_of(c, pi, base);
    T rho = 0;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int32_t q[3] = {(int32_t)((uint32_t)base[0] + (uint32_t)dx),
                          (int32_t)((uint32_t)base[1] + (uint32_t)dy),
                          (int32_t)((uint32_t)base[2] + (uint32_t)dz)};
          if (!in_grid(c, q)) continue;
          int s, e;
          cell_range(c, g, cell_index(c, q), &s, &e);
          for (int t = s; t < e; t++) {
            uint32_t j = g->perm[t];
            v3 d = V(pi.x - px[j], pi.y - py[j], pi.z - pz[j]);
            T r2 = vdot(d, d);
            if (r2 < c->hh) {
              T w = c->hh - r2;
              rho = rho + c->mass_poly6 * (w * w * w);
            }
          }
        }
    rho_out[i] = t_max(rho, c->rest_floor);
    pres_out[i] = t_max(c->stiff * (rho - c->rest), 0);
  }
}

/* forces pass (fluid.ts:370-439): spiky pressure gradient and XSPH blend. */
static void forces_core(const kk_t* c, const grid_t* g, int n, orc_state* st) {
  const T* px = ST(px); const T* py = ST(py); const T* pz = ST(pz);
  const T* vx = ST(vx); const T* vy = ST(vy); const T* vz = ST(vz);
  const T* rho = ST(rho); const T* pres = ST(pres);
  T* nvx = ST(nvx); T* nvy = ST(nvy); T* nvz = ST(nvz);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    v3 pi = V(px[i], py[i], pz[i]);
    v3 vi = V(vx[i], vy[i], vz[i]);
    T rhoi = rho[i], pri = pres[i];
    int32_t base[3];
    cell_of(c, pi, base);
    v3 fp = V(0, 0, 0);
    v3 vsum = vmuls(vi, c->h6);
    T wsum = c->h6;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int32_t q[3] = {(int32_t)((uint32_t)base[0] + (uint32_t)dx),
                          (int32_t)((uint32_t)base[1] + (uint32_t)dy),
                          (int32_t)((uint32_t)base[2] + (uint32_t)dz)};
          if (!in_grid(c, q)) continue;
          int s, e;
          cell_range(c, g, cell_index(c, q), &s, &e);
          for (int t = s; t < e; t++) {
            uint32_t j = g->perm[t];
            if ((int)j == i) continue;
            v3 dv = V(pi.x - px[j], pi.y - py[j], pi.z - pz[j]);
            T r2 = vdot(dv, dv);
            if (r2 >= c->hh) continue;
            T inv = (T)1 / rho[j];
            T w0 = c->hh - r2;
            T w = w0 * w0 * w0;
            vsum = V(vsum.x + vx[j] * w, vsum.y + vy[j] * w, vsum.z + vz[j] * w);
            wsum = wsum + w;
            T r = t_sqrt(r2);
            if (r > KF(1e-6)) {
              T hr = c->h - r;
              T s2 = c->mass * (pri + pres[j]) * (T)0.5 * inv * c->spiky * (hr * hr);
              T ir = (T)1 / r;
              fp = V(fp.x - dv.x * ir * s2, fp.y - dv.y * ir * s2, fp.z - dv.z * ir * s2);
            }
          }
        }
    T irho = (T)1 / rhoi;
    v3 acc = V(fp.x * irho + c->g[0], fp.y * irho + c->g[1], fp.z * irho + c->g[2]);
    T t = t_clamp(c->visc, 0, 1);
    T iw = (T)1 / wsum;
    v3 vs = vmuls(vsum, iw);
    v3 v = V(t_mix(vi.x, vs.x, t), t_mix(vi.y, vs.y, t), t_mix(vi.z, vs.z, t));
    v = vmuls(V(v.x + acc.x * c->dt, v.y + acc.y * c->dt, v.z + acc.z * c->dt), c->damp);
    T speed = vlen(v);
    if (speed > c->vmax) v = vmuls(v, c->vmax / speed);
    nvx[i] = v.x;
    nvy[i] = v.y;
    nvz[i] = v.z;
  }
}

/* integrate pass (fluid.ts:449-491): advance, collide against the world field,
 * record contacts (<= 32, counter keeps counting), clamp to the walls. */
static void integrate_core(const kk_t* c, const orc_collider* col, int n, orc_state* st,
                           orc_contacts* contacts) {
  T* px = ST(px); T* py = ST(py); T* pz = ST(pz);
  T* vx = ST(vx); T* vy = ST(vy); T* vz = ST(vz);
  const T* nvx = ST(nvx); const T* nvy = ST(nvy); const T* nvz = ST(nvz);
  unsigned char* hit = (unsigned char*)calloc((size_t)(n > 0 ? n : 1), 1);
  T* hp = (T*)malloc(sizeof(T) * 4 * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    v3 v = V(nvx[i], nvy[i], nvz[i]);
    v3 p = V(px[i] + v.x * c->dt, py[i] + v.y * c->dt, pz[i] + v.z * c->dt);
    if (col && col->nleaf > 0) {
      T dw;
      int leaf = collider_pick(col, p, &dw);
      if (dw < c->radius) {
        v3 nr = leaf_normal(&col->leaf[leaf], p);
        T push = c->radius - dw;
        p = V(p.x + nr.x * push, p.y + nr.y * push, p.z + nr.z * push);
        T vn = t_min(vdot(v, nr), 0);
        v = V(v.x - nr.x * vn, v.y - nr.y * vn, v.z - nr.z * vn);
        if (dw < c->cdepth) {
          hit[i] = 1;
          hp[4 * i] = p.x;
          hp[4 * i + 1] = p.y;
          hp[4 * i + 2] = p.z;
          hp[4 * i + 3] = vlen(v);
        }
      }
    }
    T cl[3] = {t_min(t_max(p.x, c->lo[0]), c->hi[0]), t_min(t_max(p.y, c->lo[1]), c->hi[1]),
               t_min(t_max(p.z, c->lo[2]), c->hi[2])};
    vx[i] = cl[0] == p.x ? v.x : 0;
    vy[i] = cl[1] == p.y ? v.y : 0;
    vz[i] = cl[2] == p.z ? v.z : 0;
    px[i] = cl[0];
    py[i] = cl[1];
    pz[i] = cl[2];
  }
  for (int i = 0; i < n; i++) {
    if (!hit[i]) continue;
    uint32_t slot = contacts->count++;
    if (slot < ORC_CONTACTS)
      for (int q = 0; q < 4; q++) contacts->points[slot][q] = (double)hp[4 * i + q];
  }
  free(hit);
  free(hp);
}

void FN(orc_density)(const orc_fluid_k* k, int n, orc_state* st) {
  kk_t c = kconv(k);
  grid_t g;
  uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  grid_build(&c, n, ST(px), ST(py), ST(pz), keys, &g);
  density_core(&c, &g, n, st);
  grid_free(&g);
  free(keys);
}

void FN(orc_forces)(const orc_fluid_k* k, int n, orc_state* st) {
  kk_t c = kconv(k);
  grid_t g;
  uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  grid_build(&c, n, ST(px), ST(py), ST(pz), keys, &g);
  forces_core(&c, &g, n, st);
  grid_free(&g);
  free(keys);
}

void FN(orc_integrate)(const orc_fluid_k* k, const orc_collider* col, int n, orc_state* st,
                       orc_contacts* contacts) {
  kk_t c = kconv(k);
  contacts->count = 0;
  integrate_core(&c, col, n, st, contacts);
}

/* emit (fluid.ts:268-291): `count` particles on a jittered ball around `pos` into
 * the ring at (base + i) % cap, velocity and staged velocity `vel`.  pow(x, 1/3)
 * goes through the C library: parity unpinned against WGSL pow (tolerance-graded). */
void FN(orc_emit)(int cap, int base, int count, uint32_t seed, const double* pos3, double radius,
                  const double* vel3, orc_state* st) {
  const T ep[3] = {rnd(pos3[0]), rnd(pos3[1]), rnd(pos3[2])};
  const T ev[3] = {rnd(vel3[0]), rnd(vel3[1]), rnd(vel3[2])};
  const T er = rnd(radius);
  const T third = IS_F32 ? (T)(1.0f / 3.0f) : (T)(1.0 / 3.0);
  for (int i = 0; i < count; i++) {
    const int slot = (int)(((uint32_t)base + (uint32_t)i) % (uint32_t)cap);
    const uint32_t s = seed * 9781u + (uint32_t)i;
    const T r[3] = {(T)orc_hash1(s * 3u + 0u), (T)orc_hash1(s * 3u + 1u), (T)orc_hash1(s * 3u + 2u)};
    v3 v = V(r[0] * 2 - 1 + KF(1e-4), r[1] * 2 - 1 + KF(1e-4), r[2] * 2 - 1 + KF(1e-4));
    const T len = vlen(v);
    const v3 dir = V(v.x / len, v.y / len, v.z / len);
    const T sc = er * (IS_F32 ? (T)powf((float)r[0], (float)third) : (T)pow((double)r[0], (double)third));
    ST(px)[slot] = ep[0] + dir.x * sc;
    ST(py)[slot] = ep[1] + dir.y * sc;
    ST(pz)[slot] = ep[2] + dir.z * sc;
    ST(vx)[slot] = ev[0];
    ST(vy)[slot] = ev[1];
    ST(vz)[slot] = ev[2];
    ST(rho)[slot] = 0;
    ST(pres)[slot] = 0;
    ST(nvx)[slot] = ev[0];
    ST(nvy)[slot] = ev[1];
    ST(nvz)[slot] = ev[2];
  }
}

/* One substep: clearGrid + insert + density + forces + integrate (fluid.ts:550-556). */
void FN(orc_substep)(const orc_fluid_k* k, const orc_collider* col, int n, orc_state* st,
                     orc_contacts* contacts, uint32_t* keys_out, uint32_t* perm_out) {
  kk_t c = kconv(k);
  grid_t g;
  uint32_t* keys = keys_out ? keys_out : (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  grid_build(&c, n, ST(px), ST(py), ST(pz), keys, &g);
  if (perm_out) memcpy(perm_out, g.perm, sizeof(uint32_t) * (size_t)n);
  contacts->count = 0;
  density_core(&c, &g, n, st);
  forces_core(&c, &g, n, st);
  integrate_core(&c, col, n, st, contacts);
  grid_free(&g);
  if (!keys_out) free(keys);
}

/* ---------------------------------------------------------------------------
 * SplatField bake (lib/src/sim/splat.ts:186-304): clear to 65535, per mip splat
 * with atomicMin of the monotone code, resolve to rgba16float.  The voxel box
 * [floor((p-reach)/v), floor((p+reach)/v)] is pinned by fluid.test.mjs:67-91
 * (ported); the grid against a brute-force min over particles in tests.
 * ------------------------------------------------------------------------- */
static void atomic_min_u32(uint32_t* a, uint32_t v) {
  uint32_t cur = __atomic_load_n(a, __ATOMIC_RELAXED);
  while (v < cur && !__atomic_compare_exchange_n(a, &cur, v, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}

void FN(orc_splat_bake)(const orc_splat_opts* o, int n, const T* px, const T* py, const T* pz,
                        const T* nx, const T* ny, const T* nz, int live_first, int live_count,
                        uint32_t* grid) {
  size_t total = 0;
  for (int m = 0; m < o->mips; m++) {
    size_t r = (size_t)(o->res >> m);
    total += r * r * r;
  }
  for (size_t i = 0; i < total; i++) grid[i] = 65535u;
  const double voxel0 = o->world_size / o->res;
  const T org[3] = {rnd(o->origin[0]), rnd(o->origin[1]), rnd(o->origin[2])};
  const T rad = rnd(o->radius);
  int first = live_first < 0 ? 0 : live_first;
  int last = live_first + live_count;
  if (last > n) last = n;
  size_t off = 0;
  for (int m = 0; m < o->mips; m++) {
    const int r = o->res >> m;
    const double voxel = voxel0 * (double)(1 << m);
    const double bw = o->band * voxel;
    const T tv = rnd(voxel), tbw = rnd(bw), reach = rnd(o->radius + bw), inv = rnd(1.0 / voxel);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = first; i < last; i++) {
      v3 p = V(px[i] - org[0], py[i] - org[1], pz[i] - org[2]);
      int32_t lo[3] = {sat_i32(t_floor((p.x - reach) * inv)), sat_i32(t_floor((p.y - reach) * inv)),
                       sat_i32(t_floor((p.z - reach) * inv))};
      int32_t hi[3] = {sat_i32(t_floor((p.x + reach) * inv)), sat_i32(t_floor((p.y + reach) * inv)),
                       sat_i32(t_floor((p.z + reach) * inv))};
      int32_t z0 = lo[2] > 0 ? lo[2] : 0, z1 = hi[2] < r - 1 ? hi[2] : r - 1;
      int32_t y0 = lo[1] > 0 ? lo[1] : 0, y1 = hi[1] < r - 1 ? hi[1] : r - 1;
      int32_t x0 = lo[0] > 0 ? lo[0] : 0, x1 = hi[0] < r - 1 ? hi[0] : r - 1;
      for (int32_t z = z0; z <= z1; z++)
        for (int32_t y = y0; y <= y1; y++)
          for (int32_t x = x0; x <= x1; x++) {
            v3 cc = V(((T)x + (T)0.5) * tv, ((T)y + (T)0.5) * tv, ((T)z + (T)0.5) * tv);
            v3 d = vsub(cc, p);
            T sd = vlen(d) - rad;
            if (nx) {
              T pl = d.x * nx[i] + d.y * ny[i] + d.z * nz[i];
              sd = t_max(pl, sd);
            }
            if (sd < tbw) {
              T q = (t_clamp(sd / tbw, -1, 1) + 1) * (T)32767.5;
              uint32_t e = (uint32_t)q;
              size_t idx = off + (size_t)x + (size_t)r * ((size_t)y + (size_t)r * (size_t)z);
              atomic_min_u32(&grid[idx], e);
            }
          }
    }
    off += (size_t)r * r * r;
  }
}

void FN(orc_splat_resolve)(const orc_splat_opts* o, const uint32_t* grid, uint16_t* texels) {
  size_t off = 0;
  const T scale = rnd(1.0 / 32767.5);
  const uint16_t mat = orc_f32_to_f16((float)rnd(o->material));
  const uint16_t one = orc_f32_to_f16(1.0f);
  for (int m = 0; m < o->mips; m++) {
    const size_t r = (size_t)(o->res >> m);
    const size_t cells = r * r * r;
    for (size_t i = 0; i < cells; i++) {
      T v = (T)grid[off + i] * scale - (T)1;
      texels[4 * (off + i)] = orc_f32_to_f16((float)v);
      texels[4 * (off + i) + 1] = mat;
      texels[4 * (off + i) + 2] = 0;
      texels[4 * (off + i) + 3] = one;
    }
    off += cells;
  }
}

/* SdfBuilder.gen (lib/src/field/builder.ts:87-142): per-tile brush cull (<= 64,
 * in brush order) then the fold from (4 * bandWorld, 0) at every voxel centre,
 * stored as clamp(d / bandWorld, -1, 1) and material. */
void FN(orc_bake_brushes)(int res, int mips, double world_size, const double* origin,
                          double band, const orc_brush* b, int nb, uint16_t* texels) {
  const double voxel0 = world_size / res;
  const uint16_t one = orc_f32_to_f16(1.0f);
  size_t off = 0;
  for (int m = 0; m < mips; m++) {
    const int r = res >> m;
    const int tiles = r / 8;
    const double voxel = voxel0 * (double)(1 << m);
    const double tile_size = voxel * 8;
    const double band_world = band * voxel;
    const T tv = rnd(voxel), tts = rnd(tile_size), thalf = rnd(tile_size / 2),
            tbw = rnd(band_world), empty = rnd(band_world * 4);
    const T org[3] = {rnd(origin[0]), rnd(origin[1]), rnd(origin[2])};
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < tiles * tiles * tiles; t++) {
      const int wx = t % tiles, wy = (t / tiles) % tiles, wz = t / (tiles * tiles);
      v3 tmin = V(org[0] + (T)wx * tts, org[1] + (T)wy * tts, org[2] + (T)wz * tts);
      v3 tc = V(tmin.x + thalf, tmin.y + thalf, tmin.z + thalf);
      int ids[64];
      int cnt = 0;
      for (int i = 0; i < nb; i++) {
        v3 bp = V(b[i].pos[0], b[i].pos[1], b[i].pos[2]);
        v3 nr = vmax0(V(t_fabs(bp.x - tc.x) - thalf, t_fabs(bp.y - tc.y) - thalf, t_fabs(bp.z - tc.z) - thalf));
        if (vlen(nr) < (T)b[i].bound + tbw && cnt < 64) ids[cnt++] = i;
      }
      for (int lz = 0; lz < 8; lz++)
        for (int ly = 0; ly < 8; ly++)
          for (int lx = 0; lx < 8; lx++) {
            v3 p = V(tmin.x + ((T)lx + (T)0.5) * tv, tmin.y + ((T)ly + (T)0.5) * tv,
                     tmin.z + ((T)lz + (T)0.5) * tv);
            T acc[2] = {empty, 0};
            for (int i = 0; i < cnt; i++) apply_brush(acc, &b[ids[i]], p);
            const size_t x = (size_t)(wx * 8 + lx), y = (size_t)(wy * 8 + ly), z = (size_t)(wz * 8 + lz);
            const size_t idx = off + x + (size_t)r * (y + (size_t)r * z);
            texels[4 * idx] = orc_f32_to_f16((float)t_clamp(acc[0] / tbw, -1, 1));
            texels[4 * idx + 1] = orc_f32_to_f16((float)acc[1]);
            texels[4 * idx + 2] = 0;
            texels[4 * idx + 3] = one;
          }
    }
    off += (size_t)r * r * r;
  }
}

/* ---------------------------------------------------------------------------
 * SdfEditor (lib/src/field/modify.ts:98-238) over the TileGrid
 * (lib/src/field/tilegrid.ts) and MipRefiner (lib/src/field/mips.ts).
 * Volume planes hold f16 bits, every mip concatenated, x fastest.  Parity:
 * bin, dilate and refresh are pinned against independent numpy formulations
 * and a closed form in tests/test_oracle_edit.py; apply reuses the brush fold
 * pinned by the reference's brush tests.  Choices the reference leaves to the
 * WebGPU implementation: pack2x16float rounds to nearest even, and the mip
 * downsample is the trilinear filter at the children's shared corner with
 * weights of exactly one half.
 * ------------------------------------------------------------------------- */
static size_t plane_off(int res, int m) {
  size_t off = 0;
  for (int i = 0; i < m; i++) {
    const size_t r = (size_t)(res >> i);
    off += r * r * r;
  }
  return off;
}

/* bin (modify.ts:98-128): one pass per mip-0 tile over the edits in order;
 * an edit reaches a tile when the distance from its origin to the tile box is
 * below bound + bandWorld.  At most `capacity` ids are kept (writeCell drops the
 * rest, setCellCount clamps). */
void FN(orc_edit_bin)(int res, double world_size, const double* origin, double band, int capacity,
                      const orc_brush* e, int ne, uint32_t* counts0, uint32_t* ids0) {
  const int tr = res / 8;
  const double voxel = world_size / res;
  const T tsize = rnd(voxel * 8), thalf = rnd(voxel * 8 / 2), bw = rnd(band * voxel);
  const v3 org = V(rnd(origin[0]), rnd(origin[1]), rnd(origin[2]));
  for (int i = 0; i < tr * tr * tr; i++) {
    const int tx = i % tr, ty = (i / tr) % tr, tz = i / (tr * tr);
    const v3 tc = vadd(org, V((T)tx * tsize + thalf, (T)ty * tsize + thalf, (T)tz * tsize + thalf));
    uint32_t n = 0;
    for (int b = 0; b < ne; b++) {
      const v3 bp = V(e[b].pos[0], e[b].pos[1], e[b].pos[2]);
      const v3 nr = vmax0(V(t_fabs(bp.x - tc.x) - thalf, t_fabs(bp.y - tc.y) - thalf, t_fabs(bp.z - tc.z) - thalf));
      if (vlen(nr) < (T)e[b].bound + bw) {
        if ((int)n < capacity) ids0[(size_t)i * capacity + n] = (uint32_t)b;
        n++;
      }
    }
    counts0[i] = n < (uint32_t)capacity ? n : (uint32_t)capacity;
  }
}

/* apply (modify.ts:130-168) fused with copy (:170-198): every voxel of a binned
 * mip-0 tile folds the tile's edits into (raw.x * bandWorld, raw.y) at its centre
 * origin + (coord + 0.5) * voxel and stores clamp(d / bandWorld, -1, 1), material. */
void FN(orc_edit_apply)(int res, double world_size, const double* origin, double band, int capacity,
                        const orc_brush* e, const uint32_t* counts0, const uint32_t* ids0, uint16_t* dist,
                        uint16_t* mat) {
  const int tr = res / 8;
  const double voxel = world_size / res;
  const T tv = rnd(voxel), bw = rnd(band * voxel);
  const v3 org = V(rnd(origin[0]), rnd(origin[1]), rnd(origin[2]));
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < tr * tr * tr; i++) {
    const uint32_t n = counts0[i];
    if (n == 0) continue;
    const int tx = i % tr, ty = (i / tr) % tr, tz = i / (tr * tr);
    for (int lz = 0; lz < 8; lz++)
      for (int ly = 0; ly < 8; ly++)
        for (int lx = 0; lx < 8; lx++) {
          const int x = tx * 8 + lx, y = ty * 8 + ly, z = tz * 8 + lz;
          const v3 p = vadd(org, vmuls(V((T)x + (T)0.5, (T)y + (T)0.5, (T)z + (T)0.5), tv));
          const size_t idx = (size_t)x + (size_t)res * ((size_t)y + (size_t)res * (size_t)z);
          T acc[2] = {(T)orc_f16_to_f32(dist[idx]) * bw, (T)orc_f16_to_f32(mat[idx])};
          for (uint32_t k = 0; k < n; k++) apply_brush(acc, &e[ids0[(size_t)i * capacity + k]], p);
          dist[idx] = orc_f32_to_f16((float)t_clamp(acc[0] / bw, -1, 1));
          mat[idx] = orc_f32_to_f16((float)acc[1]);
        }
  }
}

/* Trilinear sample of a mip plane at the corner shared by the 8 children of
 * coarse voxel (cx, cy, cz): every weight is one half. */
static T child_average(const uint16_t* plane, int fr, int cx, int cy, int cz) {
  const T h = (T)0.5;
  const int x0 = 2 * cx, y0 = 2 * cy, z0 = 2 * cz;
#define TX(x, y, z) ((T)orc_f16_to_f32(plane[(size_t)(x) + (size_t)fr * ((size_t)(y) + (size_t)fr * (size_t)(z))]))
  const T c00 = t_mix(TX(x0, y0, z0), TX(x0 + 1, y0, z0), h);
  const T c10 = t_mix(TX(x0, y0 + 1, z0), TX(x0 + 1, y0 + 1, z0), h);
  const T c01 = t_mix(TX(x0, y0, z0 + 1), TX(x0 + 1, y0, z0 + 1), h);
  const T c11 = t_mix(TX(x0, y0 + 1, z0 + 1), TX(x0 + 1, y0 + 1, z0 + 1), h);
#undef TX
  return t_mix(t_mix(c00, c10, h), t_mix(c01, c11, h), h);
}

/* MipRefiner.refresh (mips.ts:78-177) over the tiles of mip m with a non-zero
 * count: 12^3 gather of the 2x2x2 downsample (times one half; coordinates
 * clamped to the level), two eikonal sweeps on the 26-neighbour chamfer stencil
 * with weights |d| * (1 / band) * safety, shrinking by one shell per sweep, and
 * the tile core stored clamped to +-1 with the downsampled material. */
void FN(orc_mip_refresh)(int res, int mips, int m, double band, double safety, const uint32_t* counts,
                         uint16_t* dist, uint16_t* mat) {
  if (m <= 0 || m >= mips) return;
  int tile_res[32], coff[32], total = 0, nm = 0;
  char err[8];
  if (orc_volume_derive(res, mips, &nm, tile_res, coff, &total, err, (int)sizeof(err)) != 0) return;
  const int mr = res >> m, fr = res >> (m - 1), tr = tile_res[m];
  const size_t fo = plane_off(res, m - 1), co = plane_off(res, m);
  const T step = rnd((1.0 / band) * safety);
  T wts[27];
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++)
        wts[(dx + 1) + 3 * ((dy + 1) + 3 * (dz + 1))] = vlen(V((T)dx, (T)dy, (T)dz)) * step;
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < tr * tr * tr; i++) {
    if (counts[coff[m] + i] == 0) continue;
    const int E = 12, B = 2;
    T g[2][12 * 12 * 12];
    const int tx = (i % tr) * 8, ty = ((i / tr) % tr) * 8, tz = (i / (tr * tr)) * 8;
    for (int c = 0; c < E * E * E; c++) {
      int gx = tx + c % E - B, gy = ty + (c / E) % E - B, gz = tz + c / (E * E) - B;
      gx = gx < 0 ? 0 : (gx > mr - 1 ? mr - 1 : gx);
      gy = gy < 0 ? 0 : (gy > mr - 1 ? mr - 1 : gy);
      gz = gz < 0 ? 0 : (gz > mr - 1 ? mr - 1 : gz);
      g[0][c] = child_average(dist + fo, fr, gx, gy, gz) * (T)0.5;
    }
    for (int k = 0; k < B; k++) {
      const T* src = g[k % 2];
      T* dst = g[(k + 1) % 2];
      const int margin = k + 1;
      for (int c = 0; c < E * E * E; c++) {
        const int x = c % E, y = (c / E) % E, z = c / (E * E);
        T v = src[c];
        if (x >= margin && y >= margin && z >= margin && x < E - margin && y < E - margin && z < E - margin) {
          T lo = (T)-1e9, hi = (T)1e9;
          for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
              for (int dx = -1; dx <= 1; dx++) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                const T w = wts[(dx + 1) + 3 * ((dy + 1) + 3 * (dz + 1))];
                const T nv = src[(x + dx) + E * ((y + dy) + E * (z + dz))];
                hi = t_min(hi, nv + w);
                lo = t_max(lo, nv - w);
              }
          v = t_clamp(v, lo, hi);
        }
        dst[c] = v;
      }
    }
    const T* out = g[B % 2];
    for (int lz = 0; lz < 8; lz++)
      for (int ly = 0; ly < 8; ly++)
        for (int lx = 0; lx < 8; lx++) {
          const int x = tx + lx, y = ty + ly, z = tz + lz;
          const size_t idx = co + (size_t)x + (size_t)mr * ((size_t)y + (size_t)mr * (size_t)z);
          dist[idx] = orc_f32_to_f16((float)t_clamp(out[(lx + B) + E * ((ly + B) + E * (lz + B))], -1, 1));
          mat[idx] = orc_f32_to_f16((float)child_average(mat + fo, fr, x, y, z));
        }
  }
}

/* SdfEditor.flush (modify.ts:214-238): reset, bin, apply, then per coarser mip
 * dilate and refresh.  `counts` (the volume's total cell count) returns the final
 * tile counts of every mip. */
void FN(orc_edit_flush)(int res, int mips, double world_size, const double* origin, double band, double safety,
                        int capacity, const orc_brush* e, int ne, uint16_t* dist, uint16_t* mat,
                        uint32_t* counts) {
  int tile_res[32], coff[32], total = 0, nm = 0;
  char err[8];
  if (ne <= 0 || orc_volume_derive(res, mips, &nm, tile_res, coff, &total, err, (int)sizeof(err)) != 0) return;
  memset(counts, 0, sizeof(uint32_t) * (size_t)total);
  const int t0 = tile_res[0] * tile_res[0] * tile_res[0];
  uint32_t* ids = (uint32_t*)calloc((size_t)t0 * (size_t)capacity, sizeof(uint32_t));
  FN(orc_edit_bin)(res, world_size, origin, band, capacity, e, ne, counts, ids);
  FN(orc_edit_apply)(res, world_size, origin, band, capacity, e, counts, ids, dist, mat);
  free(ids);
  for (int m = 1; m < mips; m++) {
    orc_tile_dilate(res, mips, m, counts);
    FN(orc_mip_refresh)(res, mips, m, band, safety, counts, dist, mat);
  }
}

/* ClaySolver.match phase 1 (pbd.ts:240-266): per body, 256 lanes each sum the
 * body's particles t = lane, lane + 256, ... in order, then a pairwise tree
 * (s = 128 .. 1) adds lane + s into lane; centroid = sum * (1 / max(count, 1)).
 * In f32 this is the reference's reduction; in f64 it is the FP64 form the
 * library reports.  out6 per body: com xyz, restCom xyz (rest may be NULL). */
void FN(orc_body_centroids)(int nb, const uint32_t* first_count, const float* px, const float* py,
                            const float* pz, const float* rx, const float* ry, const float* rz, double* out6) {
  enum { R = 256 };
  for (int b = 0; b < nb; b++) {
    const uint32_t first = first_count[2 * b], count = first_count[2 * b + 1];
    T s[R][6];
    for (int li = 0; li < R; li++) {
      T a[6] = {0, 0, 0, 0, 0, 0};
      for (uint32_t t = (uint32_t)li; t < count; t += R) {
        const uint32_t j = first + t;
        a[0] = a[0] + (T)px[j];
        a[1] = a[1] + (T)py[j];
        a[2] = a[2] + (T)pz[j];
        if (rx) {
          a[3] = a[3] + (T)rx[j];
          a[4] = a[4] + (T)ry[j];
          a[5] = a[5] + (T)rz[j];
        }
      }
      for (int q = 0; q < 6; q++) s[li][q] = a[q];
    }
    for (int st = R / 2; st > 0; st >>= 1)
      for (int li = 0; li < st; li++)
        for (int q = 0; q < 6; q++) s[li][q] = s[li][q] + s[li + st][q];
    const T fc = (T)count;
    const T inv = (T)1 / t_max(fc, (T)1);
    for (int q = 0; q < 6; q++) out6[6 * b + q] = (double)(s[0][q] * inv);
  }
}

/* ---------------------------------------------------------------------------
 * ClaySolver (lib/src/sim/pbd.ts:144-373) over a ParticleSet
 * (lib/src/sim/particles.ts:18-68).  WGSL leaves sin/cos and normalize to the
 * implementation: sin/cos here are a Cody-Waite reduction to |r| <= pi/4 (the
 * first constant has 8 significant bits, so k * 1.5703125 is exact) and the
 * Taylor series of sin to r^9 and of cos to r^10 in Horner form (the choice the
 * HIP kernels make; parity unpinned against any WGSL implementation, pinned
 * against libm's sin/cos by tests/test_oracle_pbd.py); normalize(v) is v / length(v).
 * ------------------------------------------------------------------------- */
void FN(orc_sin_cos)(T x, T* s, T* c) {
  const T k = t_floor(x * KF(0.636619772) + KF(0.5));
  const T r = (x - k * KF(1.5703125)) - k * KF(4.83826794897e-4);
  const T r2 = r * r;
  const T sp = r + r * r2 * (KF(-0.166666667) + r2 * (KF(8.33333333e-3) + r2 * (KF(-1.98412698e-4) + r2 * KF(2.75573192e-6))));
  const T cp = KF(1.0) + r2 * (KF(-0.5) + r2 * (KF(4.16666667e-2) + r2 * (KF(-1.38888889e-3) +
                                                                          r2 * (KF(2.48015873e-5) - r2 * KF(2.75573192e-7)))));
  const int q = ((int)k) & 3;
  const T ss = (q & 1) ? cp : sp, cc = (q & 1) ? sp : cp;
  *s = (q == 2 || q == 3) ? -ss : ss;
  *c = (q == 1 || q == 2) ? -cc : cc;
}

typedef struct { T x, y, z, w; } q4;

static q4 quat_mul(q4 a, q4 b) { /* math/gpu.ts:5-14 */
  q4 r = {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
  return r;
}

static v3 quat_rotate(q4 q, v3 v) { /* math/gpu.ts:15-25 */
  v3 qv = V(q.x, q.y, q.z);
  v3 t = vmuls(vcross(qv, v), 2);
  return vadd(vadd(v, vmuls(t, q.w)), vcross(qv, t));
}

/* extractRotation, math/gpu.ts:55-78 (quatToMat3 :32-45, quatFromAxisAngle :27-30) */
static q4 extract_rotation(v3 a0, v3 a1, v3 a2, q4 q, int iterations) {
  for (int i = 0; i < iterations; i++) {
    T x = q.x, y = q.y, z = q.z, w = q.w;
    v3 c0 = V(1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w));
    v3 c1 = V(2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w));
    v3 c2 = V(2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y));
    T denom = t_fabs(vdot(c0, a0) + vdot(c1, a1) + vdot(c2, a2)) + KF(1e-9);
    v3 omega = vmuls(vadd(vadd(vcross(c0, a0), vcross(c1, a1)), vcross(c2, a2)), (T)1 / denom);
    T wl = vlen(omega);
    if (wl < KF(1e-9)) break;
    v3 axis = vmuls(omega, (T)1 / wl);
    T half = wl * KF(0.5), sh, ch;
    FN(orc_sin_cos)(half, &sh, &ch);
    q4 a = {axis.x * sh, axis.y * sh, axis.z * sh, ch};
    q4 m = quat_mul(a, q);
    T l = t_sqrt(m.x * m.x + m.y * m.y + m.z * m.z + m.w * m.w);
    q4 n = {m.x / l, m.y / l, m.z / l, m.w / l};
    q = n;
  }
  return q;
}

typedef struct {
  T *p[3], *q[3], *v[3], *u[3], *r[3];
} pbd_planes;

static v3 pl_ld(T* const a[3], int i) { return V(a[0][i], a[1][i], a[2][i]); }
static void pl_st(T* const a[3], int i, v3 v) {
  a[0][i] = v.x;
  a[1][i] = v.y;
  a[2][i] = v.z;
}

/* ownerOf, pbd.ts:131-140 */
static uint32_t pbd_owner(int max_bodies, int nbodies, const uint32_t* body3, const uint32_t* owner, uint32_t i) {
  const uint32_t b = owner[i];
  if (b >= (uint32_t)max_bodies) return (uint32_t)max_bodies;
  if (b >= (uint32_t)nbodies) return (uint32_t)max_bodies; /* undeclared: a zero record, count 0 */
  const uint32_t first = body3[3 * b], count = body3[3 * b + 1];
  if (i < first || i >= first + count) return (uint32_t)max_bodies;
  return b;
}

static v3 v3_mix(v3 a, v3 b, T t) { return V(t_mix(a.x, b.x, t), t_mix(a.y, b.y, t), t_mix(a.z, b.z, t)); }

void FN(orc_pbd_step)(const orc_pbd_params* k, const orc_collider* col, int capacity, int max_bodies, int nbodies,
                      const uint32_t* body3, const float* body2, const uint32_t* owner, T* planes, T* state,
                      int kicking, int substeps) {
  pbd_planes P;
  T** dst[5] = {P.p, P.q, P.v, P.u, P.r};
  for (int f = 0; f < 5; f++)
    for (int a = 0; a < 3; a++) dst[f][a] = planes + (size_t)(3 * f + a) * capacity;
  const T h = rnd(k->h), radius = rnd(k->radius), friction = rnd(k->friction), damping = rnd(k->damping);
  const T max_speed = rnd(k->max_speed);
  const v3 g = V(rnd(k->gravity[0]), rnd(k->gravity[1]), rnd(k->gravity[2]));
  const v3 push = V(rnd(k->push[0]), rnd(k->push[1]), rnd(k->push[2]));
  const v3 kick = V(rnd(k->kick[0]), rnd(k->kick[1]), rnd(k->kick[2]));
  const int use_col = col && col->nleaf > 0;
  if (kicking) { /* pbd.ts:144-170 */
    for (int i = 0; i < capacity; i++) {
      uint32_t b = pbd_owner(max_bodies, nbodies, body3, owner, (uint32_t)i);
      if (b != k->push_body || b >= (uint32_t)max_bodies) continue;
      if (body3[3 * b + 2] & 1u) continue;
      pl_st(P.v, i, vadd(pl_ld(P.v, i), kick));
      pl_st(P.u, i, vadd(pl_ld(P.u, i), kick));
      pl_st(P.q, i, vsub(pl_ld(P.q, i), vmuls(kick, h)));
    }
  }
  for (int sub = 0; sub < substeps; sub++) {
    /* predict, pbd.ts:172-223 */
    for (int i = 0; i < capacity; i++) {
      uint32_t b = pbd_owner(max_bodies, nbodies, body3, owner, (uint32_t)i);
      if (b >= (uint32_t)max_bodies) continue;
      uint32_t flags = body3[3 * b + 2];
      if (flags & 1u) continue;
      v3 acc = (flags & 2u) ? V(0, 0, 0) : g;
      if (b == k->push_body) acc = vadd(acc, push);
      v3 pos = pl_ld(P.p, i), prev = pl_ld(P.q, i), vel = pl_ld(P.v, i), velp = pl_ld(P.u, i);
      const T third = (T)1 / (T)3;
      const T h23 = (T)2 * h / (T)3;
      v3 vpred = vadd(vmuls(vsub(vmuls(vel, 4), velp), third), vmuls(acc, h23));
      v3 q = vadd(vmuls(vsub(vmuls(pos, 4), prev), third), vmuls(vpred, h23));
      v3 prev_out = pos;
      if (use_col) {
        T dw;
        int leaf = collider_pick(col, q, &dw);
        if (dw < radius) {
          v3 n = leaf_normal(&col->leaf[leaf], q);
          T ps = radius - dw;
          q = vadd(q, vmuls(n, ps));
          prev_out = vadd(pos, vmuls(n, ps));
          v3 dp = vsub(q, pos);
          q = vsub(q, vmuls(vsub(dp, vmuls(n, vdot(dp, n))), friction));
        }
      }
      pl_st(P.q, i, prev_out);
      pl_st(P.p, i, q);
    }
    /* match, pbd.ts:227-297: 256 lanes per body, lane-strided sums, pairwise tree */
    enum { R = 256 };
    for (int b = 0; b < nbodies; b++) {
      const uint32_t first = body3[3 * b], count = body3[3 * b + 1];
      static T sh[R][9];
      for (int li = 0; li < R; li++) {
        v3 sp = V(0, 0, 0), sr = V(0, 0, 0);
        for (uint32_t t = (uint32_t)li; t < count; t += R) {
          sp = vadd(sp, pl_ld(P.p, (int)(first + t)));
          sr = vadd(sr, pl_ld(P.r, (int)(first + t)));
        }
        sh[li][0] = sp.x; sh[li][1] = sp.y; sh[li][2] = sp.z;
        sh[li][3] = sr.x; sh[li][4] = sr.y; sh[li][5] = sr.z;
      }
      for (int s = R / 2; s > 0; s >>= 1)
        for (int li = 0; li < s; li++)
          for (int c = 0; c < 6; c++) sh[li][c] = sh[li][c] + sh[li + s][c];
      const T fc = (T)count;
      const T inv = (T)1 / (fc < (T)1 ? (T)1 : fc);
      const v3 com = vmuls(V(sh[0][0], sh[0][1], sh[0][2]), inv);
      const v3 rcom = vmuls(V(sh[0][3], sh[0][4], sh[0][5]), inv);
      for (int li = 0; li < R; li++) {
        v3 a0 = V(0, 0, 0), a1 = V(0, 0, 0), a2 = V(0, 0, 0);
        for (uint32_t t = (uint32_t)li; t < count; t += R) {
          v3 pr = vsub(pl_ld(P.p, (int)(first + t)), com);
          v3 rr = vsub(pl_ld(P.r, (int)(first + t)), rcom);
          a0 = vadd(a0, vmuls(pr, rr.x));
          a1 = vadd(a1, vmuls(pr, rr.y));
          a2 = vadd(a2, vmuls(pr, rr.z));
        }
        sh[li][0] = a0.x; sh[li][1] = a0.y; sh[li][2] = a0.z;
        sh[li][3] = a1.x; sh[li][4] = a1.y; sh[li][5] = a1.z;
        sh[li][6] = a2.x; sh[li][7] = a2.y; sh[li][8] = a2.z;
      }
      for (int s = R / 2; s > 0; s >>= 1)
        for (int li = 0; li < s; li++)
          for (int c = 0; c < 9; c++) sh[li][c] = sh[li][c] + sh[li + s][c];
      if (count > 0) {
        T* st = state + 10 * b;
        st[0] = com.x; st[1] = com.y; st[2] = com.z;
        st[3] = rcom.x; st[4] = rcom.y; st[5] = rcom.z;
        q4 q0 = {st[6], st[7], st[8], st[9]};
        q4 q = extract_rotation(V(sh[0][0], sh[0][1], sh[0][2]), V(sh[0][3], sh[0][4], sh[0][5]),
                                V(sh[0][6], sh[0][7], sh[0][8]), q0, k->rot_iters);
        st[6] = q.x; st[7] = q.y; st[8] = q.z; st[9] = q.w;
      }
    }
    /* apply, pbd.ts:299-340 */
    for (int i = 0; i < capacity; i++) {
      uint32_t b = pbd_owner(max_bodies, nbodies, body3, owner, (uint32_t)i);
      if (b >= (uint32_t)max_bodies) continue;
      if (body3[3 * b + 2] & 1u) continue;
      const T* st = state + 10 * b;
      q4 rot = {st[6], st[7], st[8], st[9]};
      v3 com = V(st[0], st[1], st[2]), rcom = V(st[3], st[4], st[5]);
      v3 pos = pl_ld(P.p, i), rest = pl_ld(P.r, i);
      v3 target = vadd(com, quat_rotate(rot, vsub(rest, rcom)));
      v3 q = v3_mix(pos, target, t_clamp((T)body2[2 * b], 0, 1));
      if (use_col) {
        T dw;
        int leaf = collider_pick(col, q, &dw);
        if (dw < radius) {
          v3 n = leaf_normal(&col->leaf[leaf], q);
          T ps = radius - dw;
          q = vadd(q, vmuls(n, ps));
          pl_st(P.q, i, vadd(pl_ld(P.q, i), vmuls(n, ps)));
        }
      }
      pl_st(P.p, i, q);
      const T plast = (T)body2[2 * b + 1];
      if (plast > 0) {
        q4 cj = {-rot.x, -rot.y, -rot.z, rot.w};
        v3 rnew = vadd(quat_rotate(cj, vsub(q, com)), rcom);
        pl_st(P.r, i, v3_mix(rest, rnew, t_clamp(plast, 0, 1)));
      }
    }
    /* finalize, pbd.ts:342-373 */
    for (int i = 0; i < capacity; i++) {
      uint32_t b = pbd_owner(max_bodies, nbodies, body3, owner, (uint32_t)i);
      if (b >= (uint32_t)max_bodies) continue;
      if (body3[3 * b + 2] & 1u) continue;
      v3 v = vmuls(vmuls(vsub(pl_ld(P.p, i), pl_ld(P.q, i)), (T)1 / t_max(h, KF(1e-6))), damping);
      T speed = vlen(v);
      if (speed > max_speed) v = vmuls(v, max_speed / speed);
      pl_st(P.u, i, pl_ld(P.v, i));
      pl_st(P.v, i, v);
    }
  }
}

/* ---------------------------------------------------------------------------
 * SurfaceExtractor.extract (lib/src/sim/extract.ts:193-441): reset, cells,
 * quads, finish, with particle slots taken in ascending cell index and quads in
 * ascending lattice point, then edge axis x, y, z.
 * ------------------------------------------------------------------------- */
typedef struct {
  T x, y, z, w;
} acc4;

/* crossing (extract.ts:157-164): the edge's zero crossing, weight 1, or nothing. */
static acc4 ext_crossing(v3 pa, T da, v3 pb, T db) {
  acc4 r = {0, 0, 0, 0};
  if (da * db >= 0) return r;
  const T t = t_clamp(da / (da - db), 0, 1);
  r.x = t_mix(pa.x, pb.x, t);
  r.y = t_mix(pa.y, pb.y, t);
  r.z = t_mix(pa.z, pb.z, t);
  r.w = 1;
  return r;
}

static acc4 acc_add(acc4 a, acc4 b) {
  acc4 r = {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
  return r;
}

static T ext_dist(const orc_collider* col, v3 p) {
  T d;
  (void)collider_pick(col, p, &d);
  return d;
}

void FN(orc_extract)(const orc_extract_params* k, const orc_collider* col, int capacity, uint32_t* body3,
                     T* planes, T* aux, uint32_t* owner, T* state, T* motion, uint32_t* ids, uint32_t* indices,
                     uint32_t* counts, uint32_t* draw) {
  pbd_planes P;
  T** dst[5] = {P.p, P.q, P.v, P.u, P.r};
  for (int f = 0; f < 5; f++)
    for (int a = 0; a < 3; a++) dst[f][a] = planes + (size_t)(3 * f + a) * capacity;
  const int res = k->res;
  const size_t cells = (size_t)res * res * res;
  const uint32_t body = k->body;
  /* reset (extract.ts:197-247): the body's mean position and velocity, summed by
   * 256 lanes strided then a pairwise tree, as the match pass sums. */
  if (motion && k->preserve) {
    const uint32_t first = body3[3 * body], count = body3[3 * body + 1];
    v3 ps[256], vs[256];
    for (int li = 0; li < 256; li++) {
      ps[li] = V(0, 0, 0);
      vs[li] = V(0, 0, 0);
      for (uint32_t t = (uint32_t)li; t < count && first + t < (uint32_t)capacity; t += 256) {
        ps[li] = vadd(ps[li], pl_ld(P.p, (int)(first + t)));
        vs[li] = vadd(vs[li], pl_ld(P.v, (int)(first + t)));
      }
    }
    for (int s2 = 128; s2 > 0; s2 >>= 1)
      for (int li = 0; li < s2; li++) {
        ps[li] = vadd(ps[li], ps[li + s2]);
        vs[li] = vadd(vs[li], vs[li + s2]);
      }
    if (count > 0) {
      const T inv = (T)1 / (T)count;
      motion[0] = ps[0].x * inv;
      motion[1] = ps[0].y * inv;
      motion[2] = ps[0].z * inv;
      motion[3] = 0;
      motion[4] = vs[0].x * inv;
      motion[5] = vs[0].y * inv;
      motion[6] = vs[0].z * inv;
      motion[7] = 0;
    }
  }
  v3 org = V(rnd(k->origin[0]), rnd(k->origin[1]), rnd(k->origin[2]));
  if (motion) org = vadd(org, V(motion[0], motion[1], motion[2]));
  const v3 vel = motion ? V(motion[4], motion[5], motion[6]) : V(rnd(k->vel[0]), rnd(k->vel[1]), rnd(k->vel[2]));
  const T c = rnd(k->cell);
  /* cells (extract.ts:256-329) */
  uint32_t found = 0;
  for (size_t idx = 0; idx < cells; idx++) {
    const int i = (int)(idx % res), j = (int)((idx / res) % res), l = (int)(idx / ((size_t)res * res));
    const v3 p0 = vadd(org, vmuls(V((T)i, (T)j, (T)l), c));
    const v3 px = vadd(p0, V(c, 0, 0)), py = vadd(p0, V(0, c, 0)), pz = vadd(p0, V(0, 0, c));
    const v3 pxy = vadd(p0, V(c, c, 0)), pxz = vadd(p0, V(c, 0, c)), pyz = vadd(p0, V(0, c, c));
    const v3 pxyz = vadd(p0, V(c, c, c));
    const T d0 = ext_dist(col, p0), dx = ext_dist(col, px), dy = ext_dist(col, py), dz = ext_dist(col, pz);
    const T dxy = ext_dist(col, pxy), dxz = ext_dist(col, pxz), dyz = ext_dist(col, pyz);
    const T dxyz = ext_dist(col, pxyz);
    acc4 acc = ext_crossing(p0, d0, px, dx);
    acc = acc_add(acc, ext_crossing(py, dy, pxy, dxy));
    acc = acc_add(acc, ext_crossing(pz, dz, pxz, dxz));
    acc = acc_add(acc, ext_crossing(pyz, dyz, pxyz, dxyz));
    acc = acc_add(acc, ext_crossing(p0, d0, py, dy));
    acc = acc_add(acc, ext_crossing(px, dx, pxy, dxy));
    acc = acc_add(acc, ext_crossing(pz, dz, pyz, dyz));
    acc = acc_add(acc, ext_crossing(pxz, dxz, pxyz, dxyz));
    acc = acc_add(acc, ext_crossing(p0, d0, pz, dz));
    acc = acc_add(acc, ext_crossing(px, dx, pxz, dxz));
    acc = acc_add(acc, ext_crossing(py, dy, pyz, dyz));
    acc = acc_add(acc, ext_crossing(pxy, dxy, pxyz, dxyz));
    if (acc.w < KF(0.5)) {
      ids[idx] = 0;
      continue;
    }
    const T inv = (T)1 / acc.w;
    const v3 v = V(acc.x * inv, acc.y * inv, acc.z * inv);
    T dv;
    const int lv = collider_pick(col, v, &dv);
    const v3 n = leaf_normal(&col->leaf[lv], v);
    const v3 p = vsub(v, vmuls(n, dv));
    const uint32_t slot = found++;
    if (slot >= k->capacity) {
      ids[idx] = 0;
      continue;
    }
    T dp;
    const int lp = collider_pick(col, p, &dp);
    const uint32_t id = k->base + slot;
    pl_st(P.p, (int)id, p);
    pl_st(P.q, (int)id, p);
    pl_st(P.v, (int)id, vel);
    pl_st(P.u, (int)id, vel);
    pl_st(P.r, (int)id, p);
    aux[id] = n.x;
    aux[(size_t)capacity + id] = n.y;
    aux[2 * (size_t)capacity + id] = n.z;
    aux[3 * (size_t)capacity + id] = leaf_material(&col->leaf[lp], p);
    owner[id] = body;
    ids[idx] = id + 1;
  }
  /* quads (extract.ts:331-404) */
  uint32_t quads = 0;
#define CID(a, b, cc) ids[(size_t)(a) + (size_t)res * ((size_t)(b) + (size_t)res * (size_t)(cc))]
  for (size_t idx = 0; idx < cells; idx++) {
    const int i = (int)(idx % res), j = (int)((idx / res) % res), l = (int)(idx / ((size_t)res * res));
    const v3 o = vadd(org, vmuls(V((T)i, (T)j, (T)l), c));
    const T dl = ext_dist(col, o);
    const int flip = !(dl < 0);
    for (int axis = 0; axis < 3; axis++) {
      uint32_t a, b, cc, dd;
      T de;
      if (axis == 0) {
        if (!(j >= 1 && l >= 1)) continue;
        de = ext_dist(col, vadd(o, V(c, 0, 0)));
        a = CID(i, j - 1, l - 1); b = CID(i, j, l - 1); cc = CID(i, j, l); dd = CID(i, j - 1, l);
      } else if (axis == 1) {
        if (!(i >= 1 && l >= 1)) continue;
        de = ext_dist(col, vadd(o, V(0, c, 0)));
        a = CID(i - 1, j, l - 1); b = CID(i - 1, j, l); cc = CID(i, j, l); dd = CID(i, j, l - 1);
      } else {
        if (!(i >= 1 && j >= 1)) continue;
        de = ext_dist(col, vadd(o, V(0, 0, c)));
        a = CID(i - 1, j - 1, l); b = CID(i, j - 1, l); cc = CID(i, j, l); dd = CID(i - 1, j, l);
      }
      if (!(de * dl < 0)) continue;
      if (a == 0 || b == 0 || cc == 0 || dd == 0) continue;
      const uint32_t q = quads++;
      if (q >= k->quad_capacity) continue;
      uint32_t* out = indices + (size_t)q * 6;
      out[0] = a - 1;
      out[1] = (flip ? dd : b) - 1;
      out[2] = cc - 1;
      out[3] = a - 1;
      out[4] = cc - 1;
      out[5] = (flip ? b : dd) - 1;
    }
  }
#undef CID
  /* finish (extract.ts:407-441) */
  const uint32_t n = found < k->capacity ? found : k->capacity;
  v3 sum[256];
  for (int li = 0; li < 256; li++) {
    sum[li] = V(0, 0, 0);
    for (uint32_t t = (uint32_t)li; t < n; t += 256) sum[li] = vadd(sum[li], pl_ld(P.r, (int)(k->base + t)));
  }
  for (int s2 = 128; s2 > 0; s2 >>= 1)
    for (int li = 0; li < s2; li++) sum[li] = vadd(sum[li], sum[li + s2]);
  const T inv = (T)1 / t_max((T)n, 1);
  const v3 com = vmuls(sum[0], inv);
  body3[3 * body] = k->base;
  body3[3 * body + 1] = n;
  T* st = state + 10 * (size_t)body;
  st[0] = com.x; st[1] = com.y; st[2] = com.z;
  st[3] = com.x; st[4] = com.y; st[5] = com.z;
  st[6] = 0; st[7] = 0; st[8] = 0; st[9] = 1;
  counts[0] = found;
  counts[1] = quads;
  const uint32_t kept = quads < k->quad_capacity ? quads : k->quad_capacity;
  draw[0] = kept * 6;
  draw[1] = 1;
  draw[2] = 0;
  draw[3] = 0;
}


/* ---------------------------------------------------------------------------
 * What a tracer reads from a field (lib/src/trace/field.ts:20-315) at any mip.
 * ------------------------------------------------------------------------- */
static T sample_mip(const orc_leaf* l, const uint16_t* plane, int res, const volp* v, v3 p) {
  orc_leaf lm = *l;
  volp vm = *v;
  lm.res = res;
  vm.res = res;
  return sample_plane(&lm, plane, &vm, p);
}

/* sample(p, mip) (what 0) or field(p, mip) (what 1) of one leaf: field.ts:135-149,
 * volume.ts:277-296; analytic field.ts:208-216. */
static void vol_read(const orc_leaf* l, int what, int mip, v3 p, T* x, T* y);

static void leaf_read(const orc_leaf* l, int what, int mip, v3 p, T* x, T* y) {
  p = leaf_point(l, p);
  if (l->kind == 2) {
    /* lerpField (field.ts:343-365): sample mixes x and keeps the smaller y; field
     * mixes both. */
    T ax, ay, bx, by;
    vol_read(l, what, mip, p, &ax, &ay);
    vol_read(l->lerp_b, what, mip, p, &bx, &by);
    const T k = lerp_k(l);
    *x = t_mix(ax, bx, k);
    *y = what ? t_mix(ay, by, k) : t_min(ay, by);
    return;
  }
  if (l->kind == 0) {
    const T d = analytic_dist(l, p);
    *x = d;
    *y = what ? rnd(l->material) : t_clamp(d * rnd(1.0 / l->analytic_band), -1, 1);
    return;
  }
  vol_read(l, what, mip, p, x, y);
}

static void vol_read(const orc_leaf* l, int what, int mip, v3 p, T* x, T* y) {
  volp v = vol_params(l);
  v3 rel = V(p.x - v.center[0], p.y - v.center[1], p.z - v.center[2]);
  T dbox = sd_box(rel, V(v.half, v.half, v.half));
  if (dbox > 0) {
    *x = dbox;
    *y = what ? 0 : 1;
    return;
  }
  const int mips = l->mips > 0 ? l->mips : 1;
  const int m = mip < mips - 1 ? mip : mips - 1;
  const T band_world = v.band * v.voxel * (T)ldexp(1.0, m);
  const uint16_t* dp = m == 0 && !l->dist_m[0] ? l->dist16 : l->dist_m[m];
  const T raw = sample_mip(l, dp, l->res >> m, &v, p);
  *x = raw * band_world;
  if (!what) {
    *y = raw;
  } else {
    const uint16_t* mp = m == 0 && !l->mat_m[0] ? l->mat16 : l->mat_m[m];
    *y = mp ? sample_mip(l, mp, l->res >> m, &v, p) : 0;
  }
}

void FN(orc_field_query)(const orc_collider* col, int what, int mip, int n, const T* p3, T* out2) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    v3 p = V(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]);
    T bx, by;
    leaf_read(&col->leaf[0], what, mip, p, &bx, &by);
    for (int k = 1; k < col->nleaf; k++) {
      T x, y;
      leaf_read(&col->leaf[k], what, mip, p, &x, &y);
      if (!(bx < x)) {
        bx = x;
        by = y;
      }
    }
    out2[2 * i] = bx;
    out2[2 * i + 1] = by;
  }
}

/* slabRange (field.ts:90-100) */
static void slab_range(v3 rel, v3 half, v3 inv, T* lo_out, T* hi_out) {
  v3 a = V((-half.x - rel.x) * inv.x, (-half.y - rel.y) * inv.y, (-half.z - rel.z) * inv.z);
  v3 b = V((half.x - rel.x) * inv.x, (half.y - rel.y) * inv.y, (half.z - rel.z) * inv.z);
  v3 lo = V(t_min(a.x, b.x), t_min(a.y, b.y), t_min(a.z, b.z));
  v3 hi = V(t_max(a.x, b.x), t_max(a.y, b.y), t_max(a.z, b.z));
  *lo_out = t_max(t_max(lo.x, lo.y), lo.z);
  *hi_out = t_min(t_min(hi.x, hi.y), hi.z);
}

/* volumeField.entry (field.ts:150-177) */
static T vol_entry(const orc_leaf* l, v3 ro, v3 rd, T tmin, T radius, T aperture) {
  volp v = vol_params(l);
  v3 rel = V(ro.x - v.center[0], ro.y - v.center[1], ro.z - v.center[2]);
  v3 half = V(v.half, v.half, v.half);
  const T eps = KF(1e-8);
  v3 r = V(t_fabs(rd.x) < eps ? eps : rd.x, t_fabs(rd.y) < eps ? eps : rd.y, t_fabs(rd.z) < eps ? eps : rd.z);
  v3 inv = V((T)1 / r.x, (T)1 / r.y, (T)1 / r.z);
  T blo, bhi, glo, ghi;
  slab_range(rel, half, inv, &blo, &bhi);
  const T g = radius + aperture * t_max(blo, tmin);
  slab_range(rel, V(half.x + g, half.y + g, half.z + g), inv, &glo, &ghi);
  const T t = t_max(glo, tmin);
  return ghi >= t ? t : KF(1e9);
}

/* entry of one leaf: offsetField moves the ray origin (field.ts:401-404), lerpField
 * takes the nearer of its two volumes' entries (field.ts:336-342). */
static T leaf_entry(const orc_leaf* l, v3 ro, v3 rd, T tmin, T radius, T aperture) {
  ro = leaf_point(l, ro);
  const T t = vol_entry(l, ro, rd, tmin, radius, aperture);
  return l->kind == 2 ? t_min(t, vol_entry(l->lerp_b, ro, rd, tmin, radius, aperture)) : t;
}

void FN(orc_field_entry)(const orc_collider* col, int n, const T* rays9, T* out) {
  for (int i = 0; i < n; i++) {
    const T* r = rays9 + 9 * (size_t)i;
    v3 ro = V(r[0], r[1], r[2]), rd = V(r[3], r[4], r[5]);
    T t = leaf_entry(&col->leaf[0], ro, rd, r[6], r[7], r[8]);
    for (int k = 1; k < col->nleaf; k++) t = t_min(t, leaf_entry(&col->leaf[k], ro, rd, r[6], r[7], r[8]));
    out[i] = t;
  }
}

/* ---------------------------------------------------------------------------
 * makeTracer(field, options).sweep (lib/src/trace/march.ts:71-218).
 * ------------------------------------------------------------------------- */
/* maxMip of the field: a volume's mipLevels - 1, 0 for an analytic leaf, the
 * deeper of a lerp pair and of a union's children (field.ts:116, 206, 247, 324). */
static int leaf_max_mip(const orc_leaf* l) {
  if (l->kind == 0) return 0;
  int m = (l->mips > 0 ? l->mips : 1) - 1;
  if (l->kind == 2) {
    const int b = (l->lerp_b->mips > 0 ? l->lerp_b->mips : 1) - 1;
    m = b > m ? b : m;
  }
  return m;
}

/* voxelWorld(mip): voxelSize * exp2(min(mip, maxMip)) of a volume, the epsilon of
 * an analytic field, the larger of a lerp pair's and of a union's children's
 * (field.ts:150-153, 220-223, 300-305, 366-371). */
static T vol_voxel_world(const orc_leaf* l, int mip) {
  const int mips = l->mips > 0 ? l->mips : 1;
  const int m = mip < mips - 1 ? mip : mips - 1;
  return rnd(l->world_size / l->res) * (T)ldexp(1.0, m);
}
static T field_voxel_world(const orc_collider* col, int mip) {
  T best = 0;
  for (int k = 0; k < col->nleaf; k++) {
    const orc_leaf* l = &col->leaf[k];
    T v = l->kind == 0 ? rnd(l->eps) : vol_voxel_world(l, mip);
    if (l->kind == 2) v = t_max(v, vol_voxel_world(l->lerp_b, mip));
    best = k == 0 ? v : t_max(best, v);
  }
  return best;
}

/* sample(p, mip) of the union, folded left (field.ts:270-298). */
static void field_sample(const orc_collider* col, int mip, v3 p, T* x, T* y) {
  leaf_read(&col->leaf[0], 0, mip, p, x, y);
  for (int k = 1; k < col->nleaf; k++) {
    T a, b;
    leaf_read(&col->leaf[k], 0, mip, p, &a, &b);
    if (!(*x < a)) {
      *x = a;
      *y = b;
    }
  }
}

void FN(orc_field_sweep)(const orc_collider* col, const orc_tracer_opts* o, int n, const T* rays10, T* t_dist2,
                         uint32_t* hit_steps2) {
  int max_mip = 0;
  for (int k = 0; k < col->nleaf; k++) {
    const int m = leaf_max_mip(&col->leaf[k]);
    max_mip = m > max_mip ? m : max_mip;
  }
  const int min_mip = o->min_mip < max_mip ? o->min_mip : max_mip;
  const int start_mip = min_mip + (max_mip - min_mip - ((max_mip - min_mip) % 2));
  const T hit_mip = (T)min_mip + (T)0.5;
  int bounded = 1; /* entry exists only when every leaf bounds itself (field.ts:250-264) */
  for (int k = 0; k < col->nleaf; k++)
    if (col->leaf[k].kind == 0) bounded = 0;
  const T slack = KF(0.866); /* INTERP_SLACK */
  T vw[ORC_MAX_MIPS];
  for (int m = 0; m <= max_mip && m < ORC_MAX_MIPS; m++) vw[m] = field_voxel_world(col, m);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    const T* r = rays10 + 10 * (size_t)i;
    const v3 ro = V(r[0], r[1], r[2]), rd = V(r[3], r[4], r[5]);
    const T tmin = r[6], tmax = r[7], radius = r[8], aperture = r[9];
    const T inv_growth = (T)1 / ((T)1 + aperture);
    const T min_step = vw[0] * rnd(o->min_step);
    const T hit_eps = vw[min_mip] * rnd(o->hit_eps);
    const T inv_eps_ramp = (T)1 / t_max((T)2 * hit_eps, KF(1e-6));
    T t_start = tmin;
    if (bounded) {
      t_start = leaf_entry(&col->leaf[0], ro, rd, tmin, radius, aperture);
      for (int k = 1; k < col->nleaf; k++)
        t_start = t_min(t_start, leaf_entry(&col->leaf[k], ro, rd, tmin, radius, aperture));
    }
    T t = t_start, mip = (T)start_mip, dist = 0;
    uint32_t hit = 0, steps = 0;
    for (int it = 0; it < o->max_steps; it++) {
      steps = (uint32_t)it + 1;
      T sx, sy;
      field_sample(col, (int)mip, V(ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t), &sx, &sy);
      dist = sx;
      const T safe = dist - slack * (vw[(int)mip] - vw[min_mip]);
      const T target = radius + aperture * t;
      const T eps = hit_eps * t_min((t - t_start) * inv_eps_ramp, 1);
      if (safe <= 0 && mip >= hit_mip) {
        mip = t_max(mip - 2, (T)min_mip);
        t = t_max(t - (T)0.5 * vw[(int)mip], t_start);
        continue;
      }
      if (safe < target + eps && sy < KF(0.999)) {
        if (mip >= hit_mip) {
          mip = t_max(mip - 2, (T)min_mip);
          t = t_max(t - (T)0.5 * vw[(int)mip], t_start);
          continue;
        }
        hit = 1;
        t = t + t_max(dist - target, 0);
        break;
      }
      t = t_max((t + safe - radius) * inv_growth, t + min_step);
      if (sy >= KF(0.999)) {
        const T up = t_min(mip + 2, (T)start_mip);
        if (dist > slack * (vw[(int)up] - vw[min_mip])) mip = up;
      }
      if (t > tmax) break;
    }
    t_dist2[2 * i] = t_min(t, tmax);
    t_dist2[2 * i + 1] = dist;
    hit_steps2[2 * i] = hit;
    hit_steps2[2 * i + 1] = steps;
  }
}
