// query.hip -- what a tracer reads from a TracedField (lib/src/trace/field.ts:20-315),
// at any mip: sample (world distance, band-normalised value), field (world distance,
// material), the box entry of a swept ray, bandWorld and voxelWorld, and the tracer's
// hierarchical sweep over them (lib/src/trace/march.ts:71-218).  The fluid's
// collider reads mip 0 only (clay_device.h); these read every level of a volume
// leaf, so their per-leaf description is built on the host for each call and does
// not enlarge the collider that every fluid kernel takes.  A volume with several
// instances is read at instance 0.
#include <cmath>

#include "clay_internal.h"

namespace clay {

namespace {

constexpr float kNever = 1e9f;  // NEVER (field.ts:82)

struct QLeaf {
  int32_t kind;    // 0 analytic, 1 volume, 2 lerp of two volumes
  int32_t mips;
  float inv_band;  // analytic: 1 / band
  const VolDev* vol;
  const uint16_t* dist;
  const uint16_t* mat;
  const VolDev* vol_b;  // lerp: the second volume (same levels as the first)
  const uint16_t* dist_b;
  const uint16_t* mat_b;
  uint32_t off[CLAY_MAX_MIPS];
  int32_t res[CLAY_MAX_MIPS];
};

struct QueryK {
  int32_t nleaf;
  QLeaf leaf[CLAY_MAX_LEAVES];
};

// textureSampleLevel(linear, clamp-to-edge) of one channel at a level of res^3
// texels (volume.ts:262-267), with f32 weights.
__device__ inline float sample_level(const VolDev& v, const uint16_t* d, int res, f3 p) {
  return sample_level_n(v, d, res, p);
}

// sample(p, mip) or field(p, mip) of one leaf.  Volume: outside the box (dBox, 1) or
// (dBox, 0); inside, the level min(mip, maxMip) read as (raw * bandWorld, raw) or
// (dist * bandWorld, material) (field.ts:135-149, volume.ts:277-296).  Analytic:
// (v, clamp(v / band, -1, 1)) or (v, material) (field.ts:208-216).
__device__ inline float2 vol_read(const QLeaf& q, const VolDev& v, const uint16_t* dist, const uint16_t* mat,
                                  f3 p, int mip, bool field) {
  const f3 rel = mk(p.x - v.center[0], p.y - v.center[1], p.z - v.center[2]);
  const float dbox = sd_box(rel, mk(v.half, v.half, v.half));
  if (dbox > 0.0f) return make_float2(dbox, field ? 0.0f : 1.0f);
  const int m = mip < q.mips - 1 ? mip : q.mips - 1;
  const float band_world = v.band * v.voxel * ldexpf(1.0f, m);
  const float raw = sample_level(v, dist + q.off[m], q.res[m], p);
  if (!field) return make_float2(raw * band_world, raw);
  return make_float2(raw * band_world, sample_level(v, mat + q.off[m], q.res[m], p));
}

// offsetField reads its child at p - offset (field.ts:405-423); lerpField mixes its two
// volumes' reads, keeping the smaller band value for sample (field.ts:343-359).
__device__ inline float2 leaf_read(const QLeaf& q, const LeafDev& l, f3 p, int mip, bool field) {
  p = leaf_point(l, p);
  if (q.kind == 0) {
    const float v = analytic_dist(l, p);
    return field ? make_float2(v, l.material) : make_float2(v, fclamp_s(v * q.inv_band, -1.0f, 1.0f));
  }
  const float2 a = vol_read(q, *q.vol, q.dist, q.mat, p, mip, field);
  if (q.kind == 1) return a;
  const float2 b = vol_read(q, *q.vol_b, q.dist_b, q.mat_b, p, mip, field);
  const float k = lerp_k(l);
  return make_float2(fmix(a.x, b.x, k), field ? fmix(a.y, b.y, k) : fmin_s(a.y, b.y));
}

// slabRange (field.ts:90-100): (tNear, tFar) against the box rel +- half.
__device__ inline float2 slab_range(f3 rel, f3 half, f3 inv) {
  const f3 a = mk((-half.x - rel.x) * inv.x, (-half.y - rel.y) * inv.y, (-half.z - rel.z) * inv.z);
  const f3 b = mk((half.x - rel.x) * inv.x, (half.y - rel.y) * inv.y, (half.z - rel.z) * inv.z);
  const f3 lo = mk(fmin_s(a.x, b.x), fmin_s(a.y, b.y), fmin_s(a.z, b.z));
  const f3 hi = mk(fmax_s(a.x, b.x), fmax_s(a.y, b.y), fmax_s(a.z, b.z));
  return make_float2(fmax_s(fmax_s(lo.x, lo.y), lo.z), fmin_s(fmin_s(hi.x, hi.y), hi.z));
}

// volumeField.entry (field.ts:150-177).
__device__ inline float vol_entry(const VolDev& v, f3 ro, f3 rd, float tmin, float radius, float aperture) {
  const f3 rel = mk(ro.x - v.center[0], ro.y - v.center[1], ro.z - v.center[2]);
  const f3 half = mk(v.half, v.half, v.half);
  const f3 r = mk(fabsf(rd.x) < 1e-8f ? 1e-8f : rd.x, fabsf(rd.y) < 1e-8f ? 1e-8f : rd.y,
                  fabsf(rd.z) < 1e-8f ? 1e-8f : rd.z);
  const f3 inv = mk(1.0f / r.x, 1.0f / r.y, 1.0f / r.z);
  const float2 bare = slab_range(rel, half, inv);
  const float g = radius + aperture * fmax_s(bare.x, tmin);
  const float2 grown = slab_range(rel, mk(half.x + g, half.y + g, half.z + g), inv);
  const float t = fmax_s(grown.x, tmin);
  return grown.y >= t ? t : kNever;
}

// sample / field of the whole field.  unionField: the whole pair from the nearer
// child, left unless right is strictly nearer (field.ts:270-298).
__device__ inline float2 field_read(const QueryK& q, const ColliderDev& col, f3 p, int mip, bool field) {
  float2 best = leaf_read(q.leaf[0], col.leaf[0], p, mip, field);
  for (int l = 1; l < q.nleaf; l++) {
    const float2 r = leaf_read(q.leaf[l], col.leaf[l], p, mip, field);
    if (!(best.x < r.x)) best = r;
  }
  return best;
}

__global__ void __launch_bounds__(256)
    k_field_query(QueryK q, ColliderDev col, int n, int mip, int field, const float* __restrict__ pts,
                  float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const f3 p = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  const float2 best = field_read(q, col, p, mip, field != 0);
  out[2 * i] = best.x;
  out[2 * i + 1] = best.y;
}

// entry of one leaf: offsetField moves the ray origin (field.ts:401-404), lerpField takes
// the nearer of its two volumes' entries (field.ts:336-342).
__device__ inline float leaf_entry(const QLeaf& q, const LeafDev& l, f3 ro, f3 rd, float tmin, float radius,
                                   float aperture) {
  ro = leaf_point(l, ro);
  const float t = vol_entry(*q.vol, ro, rd, tmin, radius, aperture);
  return q.kind == 2 ? fmin_s(t, vol_entry(*q.vol_b, ro, rd, tmin, radius, aperture)) : t;
}

// unionField.entry (field.ts:255-264): the nearest of the children's entries.
__global__ void __launch_bounds__(256)
    k_field_entry(QueryK q, ColliderDev col, int n, const float* __restrict__ rays, float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* r = rays + 9 * (size_t)i;
  const f3 ro = mk(r[0], r[1], r[2]), rd = mk(r[3], r[4], r[5]);
  float t = leaf_entry(q.leaf[0], col.leaf[0], ro, rd, r[6], r[7], r[8]);
  for (int l = 1; l < q.nleaf; l++) t = fmin_s(t, leaf_entry(q.leaf[l], col.leaf[l], ro, rd, r[6], r[7], r[8]));
  out[i] = t;
}

// makeTracer's constants (march.ts:71-95), as the shader holds them: the levels as
// f32 values, the step and hit thresholds in voxels as f32 literals, and
// voxelWorld(m) of the field for m = 0..maxMip (the loop never reads a level above
// its start level, which is at most maxMip).
struct SweepK {
  int32_t max_steps;
  int32_t has_entry;  // every leaf bounds itself (field.ts:250-264): start at entry
  int32_t min_mip;
  float min_mip_f, start_mip_f, hit_mip_f;
  float min_step_vox, hit_eps_vox;
  float vw[CLAY_MAX_MIPS];
};

constexpr float kInterpSlack = 0.866f;  // INTERP_SLACK (march.ts:20)

// makeTracer(field, options).sweep (march.ts:97-218) per ray: hierarchical sphere
// tracing that starts on an even coarse level, drops two levels (and half a finer
// voxel back) wherever the level can no longer promise empty space, and climbs two
// while readings stay saturated.  The swept radius is radius + aperture * t (ray,
// sphere sweep or cone).  Every expression is the reference's, in its order and in
// f32.  A lane runs as many iterations as its own ray takes; the wave runs as long
// as its slowest ray, so callers keep neighbouring rays together.
// rays: 10 floats per ray (ro xyz, rd xyz, tMin, tMax, radius, aperture).
__global__ void __launch_bounds__(256)
    k_field_sweep(QueryK q, ColliderDev col, SweepK k, int n, const float* __restrict__ rays,
                  clay_sweep_hit* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* r = rays + 10 * (size_t)i;
  const f3 ro = mk(r[0], r[1], r[2]), rd = mk(r[3], r[4], r[5]);
  const float tmin = r[6], tmax = r[7], radius = r[8], aperture = r[9];
  const float inv_growth = 1.0f / (1.0f + aperture);
  const float min_step = k.vw[0] * k.min_step_vox;
  const float hit_eps = k.vw[k.min_mip] * k.hit_eps_vox;
  const float inv_eps_ramp = 1.0f / fmax_s(2.0f * hit_eps, 1e-6f);
  float t_start = tmin;
  if (k.has_entry) {
    t_start = leaf_entry(q.leaf[0], col.leaf[0], ro, rd, tmin, radius, aperture);
    for (int l = 1; l < q.nleaf; l++)
      t_start = fmin_s(t_start, leaf_entry(q.leaf[l], col.leaf[l], ro, rd, tmin, radius, aperture));
  }
  const float vw_min = k.vw[k.min_mip];
  float t = t_start, mip = k.start_mip_f, dist = 0.0f;
  uint32_t hit = 0, steps = 0;
  for (int it = 0; it < k.max_steps; it++) {
    steps = (uint32_t)it + 1u;
    const f3 p = mk(ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t);
    const float2 s = field_read(q, col, p, (int)mip, false);
    dist = s.x;
    const float safe = dist - kInterpSlack * (k.vw[(int)mip] - vw_min);
    const float target = radius + aperture * t;
    const float eps = hit_eps * fmin_s((t - t_start) * inv_eps_ramp, 1.0f);
    const bool coarse = mip >= k.hit_mip_f;
    const bool close = safe < target + eps && s.y < 0.999f;
    if (coarse && (safe <= 0.0f || close)) {
      // Not a promise of empty space at this level, or a maybe-hit on a coarse one:
      // two levels finer and half of that level's voxel back.
      mip = fmax_s(mip - 2.0f, k.min_mip_f);
      t = fmax_s(t - 0.5f * k.vw[(int)mip], t_start);
      continue;
    }
    if (close) {
      hit = 1u;
      t = t + fmax_s(dist - target, 0.0f);  // one Lipschitz step closes the hitEps gap
      break;
    }
    t = fmax_s((t + safe - radius) * inv_growth, t + min_step);
    if (s.y >= 0.999f) {
      const float up = fmin_s(mip + 2.0f, k.start_mip_f);
      if (dist > kInterpSlack * (k.vw[(int)up] - vw_min)) mip = up;
    }
    if (t > tmax) break;
  }
  clay_sweep_hit h;
  h.t = fmin_s(t, tmax);
  h.hit = hit;
  h.steps = steps;
  h.dist = dist;
  out[i] = h;
}

QueryK query_k(const Field* f) {
  QueryK q{};
  q.nleaf = f->col.nleaf;
  for (int l = 0; l < q.nleaf; l++) {
    QLeaf& d = q.leaf[l];
    const Volume* v = l < (int)f->vols.size() ? f->vols[l] : nullptr;
    const Volume* vb = l < (int)f->vols_b.size() ? f->vols_b[l] : nullptr;
    d.kind = v ? (vb ? 2 : 1) : 0;
    if (v) {
      d.mips = v->info.mip_levels;
      d.vol = v->dev;
      d.dist = v->dist;
      d.mat = v->mat;
      if (vb) {
        d.vol_b = vb->dev;
        d.dist_b = vb->dist;
        d.mat_b = vb->mat;
      }
      for (int m = 0; m < d.mips; m++) {
        d.off[m] = (uint32_t)v->voxel_off[m];
        d.res[m] = v->mip_res(m);
      }
    } else {
      d.mips = 1;
      d.inv_band = (float)(1.0 / (l < (int)f->bands.size() ? f->bands[l] : 64.0));
    }
  }
  return q;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() { dfree(p); }
};

}  // namespace

}  // namespace clay

using namespace clay;

extern "C" int clay_field_query(clay_field* fh, int32_t what, int32_t mip, const float* pts, int32_t n, float* out2) {
  Field* f = reinterpret_cast<Field*>(fh);
  if (!f || (n > 0 && (!pts || !out2))) return fail(CLAY_ERR_INVALID, "clay_field_query: NULL argument");
  if (what != CLAY_QUERY_SAMPLE && what != CLAY_QUERY_FIELD) return fail(CLAY_ERR_INVALID, "clay_field_query: unknown read");
  if (mip < 0) return fail(CLAY_ERR_INVALID, "clay_field_query: negative mip");
  if (n <= 0) return CLAY_OK;
  DevBuf<float> dp, dout;
  int rc;
  if ((rc = dalloc(&dp.p, (size_t)n * 3)) || (rc = dalloc(&dout.p, (size_t)n * 2))) return rc;
  hipStream_t s = f->ctx->stream;
  CLAY_HIP_CHECK(hipMemcpyAsync(dp.p, pts, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
  {
    LaunchTimer t(f->ctx, s, "field_query");
    hipLaunchKernelGGL(k_field_query, dim3((n + 255) / 256), dim3(256), 0, s, query_k(f), f->col, n, (int)mip,
                       what == CLAY_QUERY_FIELD ? 1 : 0, dp.p, dout.p);
  }
  CLAY_HIP_CHECK(hipGetLastError());
  CLAY_HIP_CHECK(hipMemcpyAsync(out2, dout.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s));
  CLAY_HIP_CHECK(hipStreamSynchronize(s));
  return CLAY_OK;
}

extern "C" int clay_field_entry(clay_field* fh, const float* rays9, int32_t n, float* t_out) {
  Field* f = reinterpret_cast<Field*>(fh);
  if (!f || (n > 0 && (!rays9 || !t_out))) return fail(CLAY_ERR_INVALID, "clay_field_entry: NULL argument");
  for (const Volume* v : f->vols)
    if (!v) return fail(CLAY_ERR_INVALID, "TracedField.entry: an analytic leaf bounds nothing, so the field has no entry");
  if (n <= 0) return CLAY_OK;
  DevBuf<float> dr, dout;
  int rc;
  if ((rc = dalloc(&dr.p, (size_t)n * 9)) || (rc = dalloc(&dout.p, (size_t)n))) return rc;
  hipStream_t s = f->ctx->stream;
  CLAY_HIP_CHECK(hipMemcpyAsync(dr.p, rays9, sizeof(float) * 9 * n, hipMemcpyHostToDevice, s));
  {
    LaunchTimer t(f->ctx, s, "field_entry");
    hipLaunchKernelGGL(k_field_entry, dim3((n + 255) / 256), dim3(256), 0, s, query_k(f), f->col, n, dr.p,
                       dout.p);
  }
  CLAY_HIP_CHECK(hipGetLastError());
  CLAY_HIP_CHECK(hipMemcpyAsync(t_out, dout.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  CLAY_HIP_CHECK(hipStreamSynchronize(s));
  return CLAY_OK;
}

// bandWorld / voxelWorld (field.ts:46-48, 150-158, 218-224, 300-317), in f32 as a
// shader evaluates them: volume band * voxel * 2^m and voxel * 2^m at m = min(mip,
// maxMip); analytic band and epsilon; a union the smaller band and the larger voxel.
static void field_scale(const Field* f, int32_t mip, float* band_world, float* voxel_world) {
  float band = 0.0f, voxel = 0.0f;
  for (size_t l = 0; l < f->vols.size(); l++) {
    float b, vx;
    if (const Volume* v = f->vols[l]) {
      const int m = std::min<int>(mip, v->info.mip_levels - 1);
      const float p2 = std::ldexp(1.0f, m);
      b = (float)v->info.band * (float)(v->info.world_size / v->info.resolution) * p2;
      vx = (float)(v->info.world_size / v->info.resolution) * p2;
      if (const Volume* w = l < f->vols_b.size() ? f->vols_b[l] : nullptr) {
        // lerpField: the smaller band and the larger voxel of the pair (field.ts:366-377).
        const int mb = std::min<int>(mip, w->info.mip_levels - 1);
        const float q2 = std::ldexp(1.0f, mb);
        b = std::min(b, (float)w->info.band * (float)(w->info.world_size / w->info.resolution) * q2);
        vx = std::max(vx, (float)(w->info.world_size / w->info.resolution) * q2);
      }
    } else {
      b = (float)f->bands[l];
      vx = (float)f->eps[l];
    }
    band = l == 0 ? b : std::min(band, b);
    voxel = l == 0 ? vx : std::max(voxel, vx);
  }
  if (band_world) *band_world = band;
  if (voxel_world) *voxel_world = voxel;
}

extern "C" int clay_field_scale(const clay_field* fh, int32_t mip, float* band_world, float* voxel_world) {
  const Field* f = reinterpret_cast<const Field*>(fh);
  if (!f || mip < 0) return fail(CLAY_ERR_INVALID, "clay_field_scale: NULL field or negative mip");
  field_scale(f, mip, band_world, voxel_world);
  return CLAY_OK;
}

extern "C" void clay_tracer_opts_default(clay_tracer_opts* o) {
  if (!o) return;
  o->max_steps = 192;
  o->min_step = 0.25f;
  o->hit_eps = 0.5f;
  o->min_mip = 0;
}

// makeTracer (march.ts:71-95): the levels from the field's maxMip and the options,
// voxelWorld at every level the loop can read, and whether the field bounds itself.
static int sweep_k(const Field* f, const clay_tracer_opts* o, const char* who, SweepK* k) {
  clay_tracer_opts d;
  clay_tracer_opts_default(&d);
  if (!o) o = &d;
  if (o->max_steps < 0 || o->min_mip < 0)
    return fail(CLAY_ERR_INVALID, std::string(who) + ": negative maxSteps or minMip");
  *k = SweepK{};
  const int max_mip = f->max_mip;
  if (max_mip + 1 > CLAY_MAX_MIPS) return fail(CLAY_ERR_INVALID, std::string(who) + ": field deeper than CLAY_MAX_MIPS");
  const int min_mip = std::min<int>(o->min_mip, max_mip);
  // Hops move by 2, so the start level lands exactly on minMip on the way down.
  const int start_mip = min_mip + (max_mip - min_mip - ((max_mip - min_mip) % 2));
  k->max_steps = o->max_steps;
  k->min_mip = min_mip;
  k->min_mip_f = (float)min_mip;
  k->start_mip_f = (float)start_mip;
  k->hit_mip_f = (float)min_mip + 0.5f;
  k->min_step_vox = o->min_step;
  k->hit_eps_vox = o->hit_eps;
  for (int m = 0; m <= max_mip; m++) field_scale(f, m, nullptr, &k->vw[m]);
  k->has_entry = 1;
  for (const Volume* v : f->vols)
    if (!v) k->has_entry = 0;  // an analytic leaf may hold a surface anywhere
  return CLAY_OK;
}

static int launch_sweep(Field* f, const SweepK& k, const float* rays, int32_t n, clay_sweep_hit* hits,
                        hipStream_t s) {
  LaunchTimer t(f->ctx, s, "field_sweep");
  hipLaunchKernelGGL(k_field_sweep, dim3((n + 255) / 256), dim3(256), 0, s, query_k(f), f->col, k, (int)n, rays,
                     hits);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

extern "C" int clay_field_sweep(clay_field* fh, const clay_tracer_opts* opts, const float* rays10, int32_t n,
                                clay_sweep_hit* hits) {
  Field* f = reinterpret_cast<Field*>(fh);
  if (!f || (n > 0 && (!rays10 || !hits))) return fail(CLAY_ERR_INVALID, "clay_field_sweep: NULL argument");
  SweepK k;
  int rc = sweep_k(f, opts, "clay_field_sweep", &k);
  if (rc || n <= 0) return rc;
  DevBuf<float> dr;
  DevBuf<clay_sweep_hit> dh;
  if ((rc = dalloc(&dr.p, (size_t)n * 10)) || (rc = dalloc(&dh.p, (size_t)n))) return rc;
  hipStream_t s = f->ctx->stream;
  CLAY_HIP_CHECK(hipMemcpyAsync(dr.p, rays10, sizeof(float) * 10 * n, hipMemcpyHostToDevice, s));
  if ((rc = launch_sweep(f, k, dr.p, n, dh.p, s))) return rc;
  CLAY_HIP_CHECK(hipMemcpyAsync(hits, dh.p, sizeof(clay_sweep_hit) * n, hipMemcpyDeviceToHost, s));
  CLAY_HIP_CHECK(hipStreamSynchronize(s));
  return CLAY_OK;
}

extern "C" int clay_field_sweep_device(clay_field* fh, const clay_tracer_opts* opts, const float* rays10, int32_t n,
                                       clay_sweep_hit* hits, clay_stream stream) {
  Field* f = reinterpret_cast<Field*>(fh);
  if (!f || (n > 0 && (!rays10 || !hits))) return fail(CLAY_ERR_INVALID, "clay_field_sweep_device: NULL argument");
  SweepK k;
  int rc = sweep_k(f, opts, "clay_field_sweep_device", &k);
  if (rc || n <= 0) return rc;
  return launch_sweep(f, k, rays10, n, hits, stream ? (hipStream_t)stream : f->ctx->stream);
}
