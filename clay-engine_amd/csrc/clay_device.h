// clay_device.h -- device-side data layout and inline math for the gfx950 kernels.
//
// Everything here is compiled with -ffp-contract=off: each f32 operation rounds
// once, in the order the reference shader text gives, so the plain kernels agree
// bit for bit with an IEEE-f32 evaluation of the same expressions.  min/max/clamp
// are written as selects (not v_min/v_max) and float->int conversion saturates
// explicitly, so signed zeros and out-of-range values behave identically on
// every path.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "clay_hip.h"

namespace clay {

constexpr int kTile = 8;             // voxels per tile edge (volume.ts:6)
constexpr uint32_t kEncSat = 65535u; // splat.ts:110
constexpr float kEncScale = 32767.5f;

// ---------------------------------------------------------------------------
// Device-resident records
// ---------------------------------------------------------------------------

// SdfVolume's `VolumeParams` uniform (volume.ts:10-23, 220-230), f32 as the shader
// sees it.  Lives in device memory so setOrigin (host or device side) is picked up
// by kernels already enqueued or captured.
struct VolDev {
  float origin[3];
  float voxel;      // worldSize / resolution
  float center[3];
  float extent;     // worldSize
  float half;       // worldSize / 2
  float band;
  float grad_h;     // voxel * gradientVoxels
  int32_t res;
  // host-side f64 origin is kept by the owner; these are what kernels read
};

// One collider leaf: analytic brush fold or volume (mip 0).
struct LeafDev {
  int32_t kind;  // 0 analytic, 1 volume, 2 lerp of two volumes
  int32_t nbrush;
  const clay_brush* brushes;
  float eps;
  uint32_t scene_stride;  // volume leaf of a batch: one instance per scene, planes this far apart
  const VolDev* vol;      // scene_stride != 0: an array of one record per scene
  const uint16_t* dist0;  // mip-0 distance plane (f16 bits), x fastest
  const uint16_t* mat0;   // mip-0 material plane (f16 bits), same layout
  float material;         // analytic: the material `field` reports (field.ts:213-216)
  // offsetField (field.ts:388-429): a device vec3 subtracted from p before the leaf
  // is read (the caller may move it on the device between launches), or null.
  const float* offset;
  // lerpField (field.ts:316-386), kind 2: the second volume of the morph pair (same
  // resolution, mips and instances as the first) and the device blend t.
  const VolDev* vol_b;
  const uint16_t* dist0_b;
  const uint16_t* mat0_b;
  const float* lerp_t;
};

// unionField folded left over up to CLAY_MAX_LEAVES leaves (field.ts:246-315).
struct ColliderDev {
  int32_t nleaf;
  int32_t pad;
  LeafDev leaf[CLAY_MAX_LEAVES];
};

// How one substep launch finds its scenes.  A batch sorts keys (scene << shift | cell),
// so the sorted key of a particle names its scene; a single simulation is scene 0.
// Each scene owns `cells_per_scene` cell records and `contact_stride` words of
// contact list.  The substep's stamp is read from device memory so that a frame
// captured once in a graph advances it on every replay.
constexpr int kMaxSceneShift = 24;  // cell bits of a batch key at most
struct SceneK {
  const uint32_t* epoch;
  uint32_t batched;
  uint32_t shift;      // batch keys: scene << shift | cell; the cell field all ones = outside the grid
  uint32_t cell_mask;  // (1 << shift) - 1
  uint32_t cells_per_scene;
  uint32_t contact_stride;
  uint32_t xcd;     // renumber workgroups XCD-contiguously (a launch-shape choice)
  uint32_t* cmask;  // one bit per particle id: in contact this substep
  float4* cdata;    // per particle id: the contact point and speed
};

// Workgroups are dispatched round-robin over the chip's 8 XCDs (workgroup b runs
// on XCD b % 8), each with its own L2.  Renumbering them so that every XCD gets a
// contiguous run of logical workgroups keeps neighbouring particles (and a splat
// brick's chunks) in one L2.  A bijection on [0, n).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
  const uint32_t per = n / kXcds, rem = n % kXcds;
  const uint32_t x = b % kXcds, i = b / kXcds;
  return x * per + (x < rem ? x : rem) + i;
}

// FluidSim constants as f32 shader literals, derived in f64 on the host.
struct FluidK {
  float go[3];        // grid origin (volume.origin - h)
  float inv_h;
  int32_t R;          // gridRes
  int32_t bucket;
  float hh, mass_poly6, rest, rest_floor;
  float h, h6, mass, spiky;
  float g[3], dt, vmax, stiff, visc, damp, cdepth;
  float radius, lo[3], hi[3];
};

// ---------------------------------------------------------------------------
// Scalar helpers
// ---------------------------------------------------------------------------
// Words between two scenes' bound slots (six used): one 128-byte line per scene,
// so the atomics of different scenes do not queue on one cache line.
constexpr uint32_t kBoundsStride = 32;

// An order-preserving u32 encoding of a float (and back): integer min/max of the
// encodings is the float min/max, so atomics fold bounds exactly.
__device__ __forceinline__ uint32_t ordered_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
// A uniform value copied into a VGPR.  gfx950 issues a VALU instruction that reads
// an SGPR at half the rate of its all-VGPR form, so loop-invariant constants that
// inner loops combine with per-lane values are held in VGPRs.  Only for values the
// compiler holds in SGPRs (kernel arguments and what is computed from them).
__device__ __forceinline__ float vreg(float x) {
  float r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
  return r;
}
__device__ __forceinline__ uint32_t vreg_u(uint32_t x) {
  uint32_t r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
  return r;
}
// All ones when the f32 x has its sign bit set (x < 0 or x = -0), else zero: the
// test "b > a" is neg_mask(a - b) exactly, since an IEEE difference is negative
// exactly when b > a and equal operands give +0.
__device__ __forceinline__ int32_t neg_mask(float x) { return __float_as_int(x) >> 31; }

__device__ __forceinline__ float fmin_s(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmax_s(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float fclamp_s(float x, float lo, float hi) {
  return fmin_s(fmax_s(x, lo), hi);
}
__device__ __forceinline__ float fmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// Truncating f32 -> i32 with saturation; NaN -> 0.
__device__ __forceinline__ int32_t sat_i32(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}

__device__ __forceinline__ float h2f(uint16_t h) {
  return __half2float(__ushort_as_half(h));
}
__device__ __forceinline__ uint16_t f2h(float f) {
  return __half_as_ushort(__float2half_rn(f));
}

// Host-side IEEE binary32 -> binary16, round to nearest even.
inline uint16_t f2h_host(float f) {
  uint32_t x;
  __builtin_memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t exp = (x >> 23) & 0xffu;
  uint32_t mant = x & 0x7fffffu;
  if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
  const int e = (int)exp - 112;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    const int sh = 14 - e;
    uint32_t hm = mant >> sh;
    const uint32_t rem = mant & ((1u << sh) - 1u), half = 1u << (sh - 1);
    if (rem > half || (rem == half && (hm & 1u))) hm++;
    return (uint16_t)(sign | hm);
  }
  uint32_t h = sign | ((uint32_t)e << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)h;
}

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 abs3(f3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__device__ __forceinline__ f3 max30(f3 a) {
  return mk(fmax_s(a.x, 0.0f), fmax_s(a.y, 0.0f), fmax_s(a.z, 0.0f));
}
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// ---------------------------------------------------------------------------
// Primitives and the brush fold (brush.ts:238-349, 466-511; math/gpu.ts:15-25,
// 87-92, 132-137; @typegpu/sdf sdBox3d/sdRoundedBox3d/sdCapsule as Quilez's forms)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sd_box(f3 p, f3 s) {
  f3 q = abs3(p) - s;
  return len3(max30(q)) + fmin_s(fmax_s(q.x, fmax_s(q.y, q.z)), 0.0f);
}
__device__ __forceinline__ float sd_round_box(f3 p, f3 s, float r) {
  f3 q = mk(fabsf(p.x) - s.x + r, fabsf(p.y) - s.y + r, fabsf(p.z) - s.z + r);
  return len3(max30(q)) + fmin_s(fmax_s(q.x, fmax_s(q.y, q.z)), 0.0f) - r;
}
__device__ __forceinline__ float sd_capsule(f3 p, f3 a, f3 b, float r) {
  f3 pa = p - a, ba = b - a;
  float h = fclamp_s(dot3(pa, ba) / dot3(ba, ba), 0.0f, 1.0f);
  return len3(pa - ba * h) - r;
}

__device__ inline float eval_local(uint32_t kind, f3 p, f3 size, float radius) {
  switch (kind) {
    case 0: return len3(p) - size.x;
    case 1: return sd_box(p, size);
    case 2: return sd_round_box(p, size, radius);
    case 3: return sd_capsule(p, mk(0.0f, -size.y, 0.0f), mk(0.0f, size.y, 0.0f), radius);
    case 4: {
      float qx = sqrtf(p.x * p.x + p.z * p.z) - size.x, qy = p.y;
      return sqrtf(qx * qx + qy * qy) - radius;
    }
    case 5: {
      float dx = sqrtf(p.x * p.x + p.z * p.z) - size.x, dy = fabsf(p.y) - size.y;
      float mx = fmax_s(dx, 0.0f), my = fmax_s(dy, 0.0f);
      return fmin_s(fmax_s(dx, dy), 0.0f) + sqrtf(mx * mx + my * my);
    }
    case 6: return p.y - radius;
    default: return 1e9f;
  }
}

__device__ inline float eval_brush(const clay_brush& b, f3 p) {
  float inv_scale = 1.0f / fmax_s(b.scale, 1e-5f);
  f3 q = mk(-b.rot[0], -b.rot[1], -b.rot[2]);
  float qw = b.rot[3];
  f3 v = p - mk(b.pos[0], b.pos[1], b.pos[2]);
  f3 t = cross3(q, v) * 2.0f;
  f3 rot = (v + t * qw) + cross3(q, t);
  f3 local = rot * inv_scale;
  return eval_local(b.kind, local, mk(b.size[0], b.size[1], b.size[2]), b.radius) * b.scale;
}

__device__ __forceinline__ float smoothstep_s(float lo, float hi, float x) {
  float t = fclamp_s((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

__device__ inline void apply_brush(float acc[2], const clay_brush& b, f3 p) {
  float dist = eval_brush(b, p);
  float keep = b.mask >= 0.0f ? 1.0f - smoothstep_s(0.25f, 0.75f, fabsf(acc[1] - b.mask)) : 1.0f;
  float kk = fmax_s(b.blend, 1e-6f);
  if (b.op == 0) {
    float h = fclamp_s(0.5f + 0.5f * (acc[0] - dist) / kk, 0.0f, 1.0f);
    float rx = fmix(acc[0], dist, h) - kk * h * (1.0f - h);
    float ny = fmix(acc[1], b.material, h);
    acc[0] = fmix(acc[0], rx, keep);
    acc[1] = fmix(acc[1], ny, keep);
  } else if (b.op == 1) {
    float a = -dist;
    float h = fclamp_s(0.5f - 0.5f * (acc[0] - a) / kk, 0.0f, 1.0f);
    float rx = fmix(acc[0], a, h) + kk * h * (1.0f - h);
    float ny = fmix(acc[1], b.material, h * 0.85f);
    acc[0] = fmix(acc[0], rx, keep);
    acc[1] = fmix(acc[1], ny, keep);
  } else {
    float w = 1.0f - smoothstep_s(-fmax_s(b.blend, 1e-4f), 0.0f, dist);
    acc[1] = fmix(acc[1], b.material, w * keep);
  }
}

// ---------------------------------------------------------------------------
// Collider: analyticField / volumeField / unionField (field.ts:102-315)
// ---------------------------------------------------------------------------
__device__ inline float analytic_dist(const LeafDev& l, f3 p) {
  float acc[2] = {1e9f, 0.0f};
  for (int i = 0; i < l.nbrush; i++) apply_brush(acc, l.brushes[i], p);
  return acc[0];
}

__device__ __forceinline__ int clamp_texel(int x, int n) { return x < 0 ? 0 : (x >= n ? n - 1 : x); }

// The trilinear blend of the 2 x 2 x 2 texels at (x0, y0, z0) .. (x0 + 1, y0 + 1, z0 + 1),
// each coordinate clamped to the level of n^3 texels (clamp-to-edge), x then y then z.
// The two planes' offsets are 64-bit (n^3 may pass 2^32), the offsets inside a plane
// 32-bit (n^2 < 2^32 for n <= 65535).
__device__ __forceinline__ float trilinear(const uint16_t* d, int n, int x0, int y0, int z0, float fx, float fy,
                                           float fz) {
  const uint32_t xa = (uint32_t)clamp_texel(x0, n), xb = (uint32_t)clamp_texel(x0 + 1, n);
  const uint32_t ra = (uint32_t)clamp_texel(y0, n) * (uint32_t)n, rb = (uint32_t)clamp_texel(y0 + 1, n) * (uint32_t)n;
  const uint32_t nn = (uint32_t)n * (uint32_t)n;
  const uint16_t* pa = d + (size_t)clamp_texel(z0, n) * nn;
  const uint16_t* pb = d + (size_t)clamp_texel(z0 + 1, n) * nn;
  const float c00 = fmix(h2f(pa[ra + xa]), h2f(pa[ra + xb]), fx);
  const float c10 = fmix(h2f(pa[rb + xa]), h2f(pa[rb + xb]), fx);
  const float c01 = fmix(h2f(pb[ra + xa]), h2f(pb[ra + xb]), fx);
  const float c11 = fmix(h2f(pb[rb + xa]), h2f(pb[rb + xb]), fx);
  const float c0 = fmix(c00, c10, fy);
  const float c1 = fmix(c01, c11, fy);
  return fmix(c0, c1, fz);
}

// textureSampleLevel(linear, clamp-to-edge) of one channel of a level of n^3 texels
// with f32 weights (volume.ts:262-267).
__device__ __forceinline__ float sample_level_n(const VolDev& v, const uint16_t* d, int n, f3 p) {
  const float u = (p.x - v.origin[0]) / v.extent;
  const float w = (p.y - v.origin[1]) / v.extent;
  const float s = (p.z - v.origin[2]) / v.extent;
  const float fn = (float)n;
  const float tx = u * fn - 0.5f, ty = w * fn - 0.5f, tz = s * fn - 0.5f;
  const float fx0 = floorf(tx), fy0 = floorf(ty), fz0 = floorf(tz);
  return trilinear(d, n, sat_i32(fx0), sat_i32(fy0), sat_i32(fz0), tx - fx0, ty - fy0, tz - fz0);
}

// The same at mip 0.
__device__ inline float sample_raw0(const VolDev& v, const uint16_t* d, f3 p) { return sample_level_n(v, d, v.res, p); }

// SdfVolume.sampleDist at mip 0 (volume.ts:277-286) == volumeField.sample(p, 0).x.
__device__ inline float vol_dist(const VolDev& v, const uint16_t* d, f3 p) {
  f3 rel = mk(p.x - v.center[0], p.y - v.center[1], p.z - v.center[2]);
  float dbox = sd_box(rel, mk(v.half, v.half, v.half));
  if (dbox > 0.0f) return dbox;
  float band_world = v.band * v.voxel * 1.0f;
  return sample_raw0(v, d, p) * band_world;
}

__device__ __forceinline__ const VolDev& leaf_vol(const LeafDev& l, uint32_t scene) {
  return l.scene_stride ? l.vol[scene] : *l.vol;
}
__device__ __forceinline__ const uint16_t* leaf_plane(const LeafDev& l, uint32_t scene) {
  return l.dist0 + (size_t)l.scene_stride * scene;
}

__device__ __forceinline__ const VolDev& leaf_vol_b(const LeafDev& l, uint32_t scene) {
  return l.scene_stride ? l.vol_b[scene] : *l.vol_b;
}

// offsetField: the child reads p - offset (field.ts:405-427).
__device__ __forceinline__ f3 leaf_point(const LeafDev& l, f3 p) {
  return l.offset ? mk(p.x - l.offset[0], p.y - l.offset[1], p.z - l.offset[2]) : p;
}

// lerpField's k = clamp(t(), 0, 1) (field.ts:348).
__device__ __forceinline__ float lerp_k(const LeafDev& l) { return fclamp_s(*l.lerp_t, 0.0f, 1.0f); }

__device__ inline float leaf_dist(const LeafDev& l, uint32_t scene, f3 p) {
  p = leaf_point(l, p);
  if (l.kind == 0) return analytic_dist(l, p);
  const float a = vol_dist(leaf_vol(l, scene), leaf_plane(l, scene), p);
  if (l.kind == 1) return a;
  return fmix(a, vol_dist(leaf_vol_b(l, scene), l.dist0_b + (size_t)l.scene_stride * scene, p), lerp_k(l));
}

// volumeField.normal (field.ts:178-180): central differences of sampleDist.
__device__ inline f3 vol_normal(const VolDev& v, const uint16_t* d, f3 p) {
  const float h = v.grad_h;
  const f3 g = mk(vol_dist(v, d, mk(p.x + h, p.y, p.z)) - vol_dist(v, d, mk(p.x - h, p.y, p.z)),
                  vol_dist(v, d, mk(p.x, p.y + h, p.z)) - vol_dist(v, d, mk(p.x, p.y - h, p.z)),
                  vol_dist(v, d, mk(p.x, p.y, p.z + h)) - vol_dist(v, d, mk(p.x, p.y, p.z - h)));
  const float len = len3(g);
  return len > 1e-8f ? g * (1.0f / len) : mk(0.0f, 1.0f, 0.0f);
}

__device__ inline f3 leaf_normal(const LeafDev& l, uint32_t scene, f3 p) {
  p = leaf_point(l, p);
  if (l.kind == 2) {
    // normalize(mix(a.normal, b.normal, k)) (field.ts:379-383); a zero mix reads up.
    const float k = lerp_k(l);
    const f3 na = vol_normal(leaf_vol(l, scene), leaf_plane(l, scene), p);
    const f3 nb = vol_normal(leaf_vol_b(l, scene), l.dist0_b + (size_t)l.scene_stride * scene, p);
    const f3 g = mk(fmix(na.x, nb.x, k), fmix(na.y, nb.y, k), fmix(na.z, nb.z, k));
    const float len = len3(g);
    return len > 1e-9f ? g * (1.0f / len) : mk(0.0f, 1.0f, 0.0f);
  }
  if (l.kind == 0) {
    float e = l.eps;
    f3 g = mk(analytic_dist(l, mk(p.x + e, p.y, p.z)) - analytic_dist(l, mk(p.x - e, p.y, p.z)),
              analytic_dist(l, mk(p.x, p.y + e, p.z)) - analytic_dist(l, mk(p.x, p.y - e, p.z)),
              analytic_dist(l, mk(p.x, p.y, p.z + e)) - analytic_dist(l, mk(p.x, p.y, p.z - e)));
    float len = len3(g);
    return len > 1e-9f ? g * (1.0f / len) : mk(0.0f, 1.0f, 0.0f);
  }
  return vol_normal(leaf_vol(l, scene), leaf_plane(l, scene), p);
}

// volumeField.field(p, 0).y: the trilinear material at mip 0, 0 outside the box
// (volume.ts:288-296).
__device__ inline float vol_material(const VolDev& v, const uint16_t* m, f3 p) {
  f3 rel = mk(p.x - v.center[0], p.y - v.center[1], p.z - v.center[2]);
  if (sd_box(rel, mk(v.half, v.half, v.half)) > 0.0f) return 0.0f;
  return sample_raw0(v, m, p);
}

// field(p, 0).y of one leaf: analyticField's constant material, volumeField's, or
// lerpField's mix of its two volumes' (field.ts:355-359).
__device__ inline float leaf_material(const LeafDev& l, uint32_t scene, f3 p) {
  p = leaf_point(l, p);
  if (l.kind == 0) return l.material;
  const float a = vol_material(leaf_vol(l, scene), l.mat0 + (size_t)l.scene_stride * scene, p);
  if (l.kind == 1) return a;
  return fmix(a, vol_material(leaf_vol_b(l, scene), l.mat0_b + (size_t)l.scene_stride * scene, p), lerp_k(l));
}

// Returns the leaf the union selects (right child wins unless left is strictly nearer).
__device__ inline int collider_pick(const ColliderDev& c, uint32_t scene, f3 p, float* dist) {
  int best = 0;
  float bd = leaf_dist(c.leaf[0], scene, p);
  for (int i = 1; i < c.nleaf; i++) {
    float d = leaf_dist(c.leaf[i], scene, p);
    if (!(bd < d)) {
      best = i;
      bd = d;
    }
  }
  *dist = bd;
  return best;
}

__device__ inline f3 collider_normal(const ColliderDev& c, uint32_t scene, int leaf, f3 p) {
  // Unrolled so the leaf record stays in registers/SGPRs.
  for (int i = 0; i < CLAY_MAX_LEAVES; i++)
    if (i == leaf) return leaf_normal(c.leaf[i], scene, p);
  return mk(0.0f, 1.0f, 0.0f);
}

// unionField.field(p, 0).y: the material of the leaf the union selects at p.
__device__ inline float collider_material(const ColliderDev& c, uint32_t scene, int leaf, f3 p) {
  for (int i = 0; i < CLAY_MAX_LEAVES; i++)
    if (i == leaf) return leaf_material(c.leaf[i], scene, p);
  return 0.0f;
}

}  // namespace clay
