// api.hip -- the C ABI (include/clay_hip.h): handles, option defaults and the
// orchestration of FluidSim.step / SplatField.bake on one HIP stream.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "clay_internal.h"

namespace clay {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

// ---- timing ------------------------------------------------------------------
int Ctx::stat_index(const char* name) {
  for (size_t i = 0; i < stats.size(); i++)
    if (stats[i].name == name) return (int)i;
  KernelStat k;
  k.name = name;
  stats.push_back(k);
  return (int)stats.size() - 1;
}

hipEvent_t Ctx::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void Ctx::collect() {
  for (auto& p : pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(p.stop) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      stats[p.stat].total_ms += ms;
      stats[p.stat].launches += 1;
    }
    event_pool.push_back(p.start);
    event_pool.push_back(p.stop);
  }
  pending.clear();
}

LaunchTimer::LaunchTimer(Ctx* ctx, hipStream_t s, const char* name) : ctx_(ctx), s_(s) {
  if (!ctx_ || !ctx_->timing || ctx_->capturing) return;
  stat_ = ctx_->stat_index(name);
  a_ = ctx_->take_event();
  b_ = ctx_->take_event();
  if (a_ && b_) (void)hipEventRecord(a_, s_);
}

LaunchTimer::~LaunchTimer() {
  if (stat_ < 0 || !a_ || !b_) return;
  (void)hipEventRecord(b_, s_);
  ctx_->pending.push_back(PendingTiming{stat_, a_, b_});
  if (ctx_->pending.size() > 4096) ctx_->collect();
}

namespace {

int volume_geometry(int32_t resolution, int32_t mips_req, clay_volume_info* out) {
  // volume.ts:139-179
  memset(out, 0, sizeof(*out));
  if (resolution <= 0) return fail(CLAY_ERR_INVALID, "SdfVolume: resolution must be positive");
  const int TILE = kTile;
  int halvings = 0;
  double n = (double)resolution / TILE;
  while (n != 0.0 && n == floor(n) && fmod(n, 2.0) == 0.0) {
    halvings++;
    n /= 2.0;
  }
  int mips = mips_req;
  if (mips <= 0) {
    const int li = (int)log2((double)resolution / TILE);
    mips = std::max(1, std::min(li, halvings + 1));
  }
  if (mips > CLAY_MAX_MIPS)
    return fail(CLAY_ERR_INVALID, "SdfVolume: at most " + std::to_string(CLAY_MAX_MIPS) + " mip levels");
  const long long need = (long long)TILE << (mips - 1);
  if (resolution % need != 0) {
    char buf[160];
    snprintf(buf, sizeof(buf), "SdfVolume: resolution %d must be a multiple of %lld for %d mip levels",
             resolution, need, mips);
    return fail(CLAY_ERR_INVALID, buf);
  }
  out->resolution = resolution;
  out->mip_levels = mips;
  const int tile_res = resolution / TILE;
  int off = 0;
  for (int m = 0; m < mips; m++) {
    const int tr = tile_res >> m;
    out->tile_res[m] = tr;
    out->cell_offsets[m] = off;
    off += tr * tr * tr;
  }
  out->total_cells = off;
  return CLAY_OK;
}

void write_voldev(Volume* v) {
  const double half = v->info.world_size / 2;
  VolDev& d = v->host;
  for (int a = 0; a < 3; a++) {
    d.origin[a] = (float)v->origin[a];
    d.center[a] = (float)(v->origin[a] + half);
  }
  d.voxel = (float)v->info.voxel_size;
  d.extent = (float)v->info.world_size;
  d.half = (float)half;
  d.band = (float)v->info.band;
  d.grad_h = d.voxel * (float)v->grad_voxels;
  d.res = v->info.resolution;
}

__global__ void k_fill_u16(uint16_t* p, size_t n, uint16_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int fill_u16(Ctx* c, hipStream_t s, uint16_t* p, size_t n, uint16_t v) {
  if (!n) return CLAY_OK;
  size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  LaunchTimer t(c, s, "fill_u16");
  hipLaunchKernelGGL(k_fill_u16, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int volume_create_impl(Ctx* ctx, const clay_volume_desc* d, Volume** out, int instances = 1) {
  const int32_t res = d->resolution > 0 ? d->resolution : 128;
  clay_volume_info info;
  int rc = volume_geometry(res, d->mip_levels, &info);
  if (rc) return rc;
  Volume* v = new (std::nothrow) Volume();
  if (!v) return fail(CLAY_ERR_INVALID, "out of host memory");
  v->ctx = ctx;
  v->instances = instances;
  v->info = info;
  v->info.world_size = d->world_size > 0 ? d->world_size : 16.0;
  v->info.band = d->band > 0 ? d->band : 4.0;
  v->info.voxel_size = v->info.world_size / res;
  for (int a = 0; a < 3; a++) v->origin[a] = v->info.origin[a] = d->origin[a];
  v->grad_voxels = d->gradient_voxels > 0 ? d->gradient_voxels : 1.0;
  size_t total = 0;
  for (int m = 0; m < info.mip_levels; m++) {
    const size_t r = (size_t)v->mip_res(m);
    v->voxel_off.push_back(total);
    total += r * r * r;
  }
  v->total_voxels = total;
  write_voldev(v);
  const size_t planes = total * (size_t)instances;
  if ((rc = dalloc(&v->dev, (size_t)instances)) || (rc = dalloc(&v->dist, planes)) ||
      (rc = dalloc(&v->mat, planes))) {
    dfree(v->dev);
    dfree(v->dist);
    dfree(v->mat);
    delete v;
    return rc;
  }
  hipStream_t s = ctx->stream;
  const std::vector<VolDev> recs((size_t)instances, v->host);
  if (hipMemcpyAsync(v->dev, recs.data(), sizeof(VolDev) * instances, hipMemcpyHostToDevice, s) != hipSuccess ||
      fill_u16(ctx, s, v->dist, planes, f2h_host(1.0f)) || fill_u16(ctx, s, v->mat, planes, 0) ||
      hipStreamSynchronize(s) != hipSuccess) {
    dfree(v->dev);
    dfree(v->dist);
    dfree(v->mat);
    delete v;
    return fail(CLAY_ERR_HIP, "SdfVolume: device initialisation failed");
  }
  *out = v;
  return CLAY_OK;
}

void volume_destroy_impl(Volume* v) {
  if (!v) return;
  dfree(v->wire_scratch);
  dfree(v->wire_flag);
  dfree(v->dev);
  dfree(v->dist);
  dfree(v->mat);
  delete v;
}

int volume_set_origin_impl(Volume* v, const double o[3], hipStream_t s) {
  for (int a = 0; a < 3; a++) v->origin[a] = v->info.origin[a] = o[a];
  write_voldev(v);
  // The record is small; a host->device copy on the stream keeps it ordered with
  // the bakes that read it, as the reference's queue-ordered writePartial does.
  VolDev* staged = new VolDev(v->host);
  hipError_t e = hipMemcpyAsync(v->dev, staged, sizeof(VolDev), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  delete staged;
  if (e != hipSuccess) return fail(CLAY_ERR_HIP, std::string("setOrigin: ") + hipGetErrorString(e));
  return CLAY_OK;
}

Field* make_volume_field(Ctx* ctx, Volume* v) {
  Field* f = new (std::nothrow) Field();
  if (!f) return nullptr;
  f->ctx = ctx;
  f->max_mip = v->info.mip_levels - 1;
  f->col.nleaf = 1;
  LeafDev& l = f->col.leaf[0];
  l.kind = 1;
  l.vol = v->dev;
  l.dist0 = v->dist;
  l.mat0 = v->mat;
  f->vols = {v};
  f->vols_b = {nullptr};
  f->bands = {0.0};
  f->eps = {0.0};
  // One instance per scene: the collider reads the scene's own record and planes.
  l.scene_stride = v->instances > 1 ? (uint32_t)v->total_voxels : 0u;
  return f;
}

void fluid_build_k(Fluid* f) {
  const clay_fluid_consts& c = f->consts;
  const clay_fluid_params& p = f->params;
  FluidK& k = f->k;
  for (int a = 0; a < 3; a++) {
    k.go[a] = (float)c.grid_origin[a];
    k.g[a] = (float)p.gravity[a];
    k.lo[a] = (float)c.wall_lo[a];
    k.hi[a] = (float)c.wall_hi[a];
  }
  k.inv_h = (float)f->inv_h;
  k.R = c.grid_res;
  k.bucket = c.bucket;
  k.hh = (float)f->hh;
  k.mass_poly6 = (float)f->mass_poly6;
  k.rest = (float)c.rest_density;
  k.rest_floor = (float)f->rest_floor;
  k.h = (float)c.h;
  k.h6 = (float)f->h6;
  k.mass = (float)c.mass;
  k.spiky = (float)c.spiky_grad;
  k.dt = (float)p.dt;
  k.vmax = (float)p.max_speed;
  k.stiff = (float)p.stiffness;
  k.visc = (float)p.viscosity;
  k.damp = (float)p.damping;
  k.cdepth = (float)p.contact_depth;
  k.radius = (float)c.radius;
}

}  // namespace

int launch_emit(Ctx* c, hipStream_t s, Fluid* f, uint32_t base, uint32_t count, uint32_t seed);

}  // namespace clay

using namespace clay;

// ============================================================================
// library / context
// ============================================================================
extern "C" const char* clay_last_error(void) { return g_error.c_str(); }
extern "C" int clay_version(void) { return 1; }

extern "C" int clay_ctx_create(int device, clay_ctx** out) {
  if (!out) return fail(CLAY_ERR_INVALID, "clay_ctx_create: out is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(CLAY_ERR_NODEVICE, "clay_ctx_create: no HIP device is visible");
  if (device < 0 || device >= n) return fail(CLAY_ERR_INVALID, "clay_ctx_create: bad device index");
  CLAY_HIP_CHECK(hipSetDevice(device));
  Ctx* c = new (std::nothrow) Ctx();
  if (!c) return fail(CLAY_ERR_INVALID, "out of host memory");
  c->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    c->cus = cus;
  // Eight 256-thread workgroups per CU fill the wave slots of a gfx950 CU.
  c->persistent_blocks = c->cus * 8;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(CLAY_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  if (const int rc = splat_static_init(device)) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = reinterpret_cast<clay_ctx*>(c);
  return CLAY_OK;
}

extern "C" int clay_ctx_destroy(clay_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return CLAY_OK;
  (void)hipStreamSynchronize(c->stream);
  c->collect();
  for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return CLAY_OK;
}

extern "C" int clay_ctx_synchronize(clay_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(CLAY_ERR_INVALID, "clay_ctx_synchronize: NULL context");
  CLAY_HIP_CHECK(hipStreamSynchronize(c->stream));
  return CLAY_OK;
}

extern "C" int clay_ctx_set_timing(clay_ctx* ctx, int enabled) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(CLAY_ERR_INVALID, "clay_ctx_set_timing: NULL context");
  c->timing = enabled != 0;
  return CLAY_OK;
}

extern "C" int clay_ctx_kernel_stats(clay_ctx* ctx, int index, char* name, int name_len,
                                     double* total_ms, uint64_t* launches) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(CLAY_ERR_INVALID, "clay_ctx_kernel_stats: NULL context");
  c->collect();
  if (index < 0 || index >= (int)c->stats.size()) return fail(CLAY_ERR_INVALID, "no such kernel stat");
  const KernelStat& k = c->stats[index];
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", k.name.c_str());
  if (total_ms) *total_ms = k.total_ms;
  if (launches) *launches = k.launches;
  return CLAY_OK;
}

extern "C" int clay_ctx_reset_stats(clay_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(CLAY_ERR_INVALID, "clay_ctx_reset_stats: NULL context");
  c->collect();
  c->stats.clear();
  return CLAY_OK;
}

// ============================================================================
// SdfVolume
// ============================================================================
extern "C" int clay_volume_derive(int32_t resolution, int32_t mip_levels, clay_volume_info* out) {
  if (!out) return fail(CLAY_ERR_INVALID, "clay_volume_derive: out is NULL");
  return volume_geometry(resolution, mip_levels, out);
}

extern "C" int clay_volume_create(clay_ctx* ctx, const clay_volume_desc* desc, clay_volume** out) {
  if (!ctx || !desc || !out) return fail(CLAY_ERR_INVALID, "clay_volume_create: NULL argument");
  Volume* v = nullptr;
  int rc = volume_create_impl(reinterpret_cast<Ctx*>(ctx), desc, &v);
  if (rc) return rc;
  *out = reinterpret_cast<clay_volume*>(v);
  return CLAY_OK;
}

extern "C" int clay_volume_destroy(clay_volume* vol) {
  volume_destroy_impl(reinterpret_cast<Volume*>(vol));
  return CLAY_OK;
}

extern "C" int clay_volume_get_info(const clay_volume* vol, clay_volume_info* out) {
  if (!vol || !out) return fail(CLAY_ERR_INVALID, "clay_volume_get_info: NULL argument");
  *out = reinterpret_cast<const Volume*>(vol)->info;
  return CLAY_OK;
}

extern "C" int clay_volume_set_origin(clay_volume* vol, const double origin[3], clay_stream stream) {
  if (!vol || !origin) return fail(CLAY_ERR_INVALID, "clay_volume_set_origin: NULL argument");
  Volume* v = reinterpret_cast<Volume*>(vol);
  return volume_set_origin_impl(v, origin, pick(v->ctx, stream));
}

extern "C" int clay_volume_set_origin_device(clay_volume* vol, const float* origin3, clay_stream stream) {
  if (!vol || !origin3) return fail(CLAY_ERR_INVALID, "clay_volume_set_origin_device: NULL argument");
  Volume* v = reinterpret_cast<Volume*>(vol);
  return launch_set_origin_dev(v->ctx, pick(v->ctx, stream), v->dev, origin3, v->info.world_size / 2);
}

extern "C" int clay_volume_voxel_size_at(const clay_volume* vol, int32_t mip, double* out) {
  if (!vol || !out) return fail(CLAY_ERR_INVALID, "clay_volume_voxel_size_at: NULL argument");
  const Volume* v = reinterpret_cast<const Volume*>(vol);
  *out = v->info.voxel_size * pow(2.0, mip);  // volume.ts:336-338
  return CLAY_OK;
}

extern "C" int clay_volume_download(clay_volume* vol, int32_t mip, uint16_t* rgba) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v || !rgba) return fail(CLAY_ERR_INVALID, "clay_volume_download: NULL argument");
  if (mip < 0 || mip >= v->info.mip_levels) return fail(CLAY_ERR_INVALID, "clay_volume_download: bad mip");
  const size_t r = (size_t)v->mip_res(mip), n = r * r * r;
  std::vector<uint16_t> d(n), m(n);
  CLAY_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(d.data(), v->dist + v->voxel_off[mip], n * 2, hipMemcpyDeviceToHost));
  CLAY_HIP_CHECK(hipMemcpy(m.data(), v->mat + v->voxel_off[mip], n * 2, hipMemcpyDeviceToHost));
  const uint16_t one = f2h_host(1.0f);
  for (size_t i = 0; i < n; i++) {
    rgba[4 * i] = d[i];
    rgba[4 * i + 1] = m[i];
    rgba[4 * i + 2] = 0;
    rgba[4 * i + 3] = one;
  }
  return CLAY_OK;
}

extern "C" int clay_volume_upload(clay_volume* vol, int32_t mip, const uint16_t* rgba) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v || !rgba) return fail(CLAY_ERR_INVALID, "clay_volume_upload: NULL argument");
  if (mip < 0 || mip >= v->info.mip_levels) return fail(CLAY_ERR_INVALID, "clay_volume_upload: bad mip");
  const size_t r = (size_t)v->mip_res(mip), n = r * r * r;
  std::vector<uint16_t> d(n), m(n);
  for (size_t i = 0; i < n; i++) {
    d[i] = rgba[4 * i];
    m[i] = rgba[4 * i + 1];
  }
  CLAY_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(v->dist + v->voxel_off[mip], d.data(), n * 2, hipMemcpyHostToDevice));
  CLAY_HIP_CHECK(hipMemcpy(v->mat + v->voxel_off[mip], m.data(), n * 2, hipMemcpyHostToDevice));
  return CLAY_OK;
}

extern "C" int clay_volume_bake_brushes(clay_volume* vol, const clay_brush* brushes, int32_t n,
                                        clay_stream stream) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v || (n > 0 && !brushes)) return fail(CLAY_ERR_INVALID, "clay_volume_bake_brushes: NULL argument");
  if (n < 0) return fail(CLAY_ERR_INVALID, "clay_volume_bake_brushes: negative brush count");
  for (int i = 0; i < n; i++)
    if (brushes[i].kind > 6)
      return fail(CLAY_ERR_INVALID, "clay_volume_bake_brushes: brush kind " + std::to_string(brushes[i].kind) +
                                        " is outside the supported set (sphere, box, roundBox, capsule, torus, cylinder, plane)");
  hipStream_t s = pick(v->ctx, stream);
  clay_brush* dev = nullptr;
  int rc = dalloc(&dev, (size_t)std::max(n, 1));
  if (rc) return rc;
  if (n > 0) CLAY_HIP_CHECK(hipMemcpyAsync(dev, brushes, sizeof(clay_brush) * n, hipMemcpyHostToDevice, s));
  rc = launch_bake_brushes(v->ctx, s, v, dev, n);
  (void)hipStreamSynchronize(s);
  dfree(dev);
  return rc;
}

// ============================================================================
// TracedField
// ============================================================================
extern "C" int clay_field_analytic(clay_ctx* ctx, const clay_brush* brushes, int32_t n,
                                   const clay_analytic_opts* opts, clay_field** out) {
  if (!ctx || !out || (n > 0 && !brushes)) return fail(CLAY_ERR_INVALID, "clay_field_analytic: NULL argument");
  if (n <= 0) return fail(CLAY_ERR_INVALID, "analyticField: the brush list is empty");
  for (int i = 0; i < n; i++)
    if (brushes[i].kind > 6)
      return fail(CLAY_ERR_INVALID, "analyticField: brush kind " + std::to_string(brushes[i].kind) +
                                        " is outside the supported set (sphere, box, roundBox, capsule, torus, cylinder, plane)");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  Field* f = new (std::nothrow) Field();
  if (!f) return fail(CLAY_ERR_INVALID, "out of host memory");
  f->ctx = c;
  int rc = dalloc(&f->brushes_dev, (size_t)n);
  if (rc) {
    delete f;
    return rc;
  }
  if (hipMemcpy(f->brushes_dev, brushes, sizeof(clay_brush) * n, hipMemcpyHostToDevice) != hipSuccess) {
    dfree(f->brushes_dev);
    delete f;
    return fail(CLAY_ERR_HIP, "analyticField: brush upload failed");
  }
  f->max_mip = 0;
  f->col.nleaf = 1;
  LeafDev& l = f->col.leaf[0];
  l.kind = 0;
  l.nbrush = n;
  l.brushes = f->brushes_dev;
  l.eps = (float)(opts && opts->epsilon > 0 ? opts->epsilon : 1e-3);
  l.material = (float)(opts ? opts->material : 0.0);
  f->vols = {nullptr};
  f->vols_b = {nullptr};
  f->bands = {opts && opts->band > 0 ? opts->band : 64.0};
  f->eps = {opts && opts->epsilon > 0 ? opts->epsilon : 1e-3};
  *out = reinterpret_cast<clay_field*>(f);
  return CLAY_OK;
}

extern "C" int clay_field_volume(clay_ctx* ctx, clay_volume* vol, clay_field** out) {
  if (!ctx || !vol || !out) return fail(CLAY_ERR_INVALID, "clay_field_volume: NULL argument");
  Field* f = make_volume_field(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<Volume*>(vol));
  if (!f) return fail(CLAY_ERR_INVALID, "out of host memory");
  *out = reinterpret_cast<clay_field*>(f);
  return CLAY_OK;
}

extern "C" int clay_field_union(clay_field* a, clay_field* b, clay_field** out) {
  Field* fa = reinterpret_cast<Field*>(a);
  Field* fb = reinterpret_cast<Field*>(b);
  if (!fa || !fb || !out) return fail(CLAY_ERR_INVALID, "clay_field_union: NULL argument");
  if (fa->col.nleaf + fb->col.nleaf > CLAY_MAX_LEAVES)
    return fail(CLAY_ERR_INVALID, "unionField: more than " + std::to_string(CLAY_MAX_LEAVES) + " leaves");
  Field* f = new (std::nothrow) Field();
  if (!f) return fail(CLAY_ERR_INVALID, "out of host memory");
  f->ctx = fa->ctx;
  f->max_mip = std::max(fa->max_mip, fb->max_mip);
  f->col.nleaf = 0;
  for (int i = 0; i < fa->col.nleaf; i++) f->col.leaf[f->col.nleaf++] = fa->col.leaf[i];
  for (int i = 0; i < fb->col.nleaf; i++) f->col.leaf[f->col.nleaf++] = fb->col.leaf[i];
  for (const Field* c : {fa, fb}) {
    f->vols.insert(f->vols.end(), c->vols.begin(), c->vols.end());
    f->vols_b.insert(f->vols_b.end(), c->vols_b.begin(), c->vols_b.end());
    f->bands.insert(f->bands.end(), c->bands.begin(), c->bands.end());
    f->eps.insert(f->eps.end(), c->eps.begin(), c->eps.end());
  }
  f->children = {fa, fb};
  *out = reinterpret_cast<clay_field*>(f);
  return CLAY_OK;
}

// offsetField (field.ts:388-429): every leaf of `a` reads p - *offset_dev.  The
// offset is read on the device at each evaluation, so a caller can move the field
// without a host round trip.  One offset per leaf.
extern "C" int clay_field_offset(clay_field* a, const float* offset_dev, clay_field** out) {
  Field* fa = reinterpret_cast<Field*>(a);
  if (!fa || !offset_dev || !out) return fail(CLAY_ERR_INVALID, "clay_field_offset: NULL argument");
  for (int i = 0; i < fa->col.nleaf; i++)
    if (fa->col.leaf[i].offset)
      return fail(CLAY_ERR_INVALID, "offsetField: the field is already offset (add the offsets on the device)");
  Field* f = new (std::nothrow) Field();
  if (!f) return fail(CLAY_ERR_INVALID, "out of host memory");
  f->ctx = fa->ctx;
  f->max_mip = fa->max_mip;
  f->col = fa->col;
  for (int i = 0; i < f->col.nleaf; i++) f->col.leaf[i].offset = offset_dev;
  f->vols = fa->vols;
  f->vols_b = fa->vols_b;
  f->bands = fa->bands;
  f->eps = fa->eps;
  f->children = {fa};
  *out = reinterpret_cast<clay_field*>(f);
  return CLAY_OK;
}

// lerpField (field.ts:316-386) of a morph pair: two volume fields of the same
// resolution, mips and instances, blended by clamp(*t_dev, 0, 1) read on the device.
extern "C" int clay_field_lerp(clay_field* a, clay_field* b, const float* t_dev, clay_field** out) {
  Field* fa = reinterpret_cast<Field*>(a);
  Field* fb = reinterpret_cast<Field*>(b);
  if (!fa || !fb || !t_dev || !out) return fail(CLAY_ERR_INVALID, "clay_field_lerp: NULL argument");
  if (fa->col.nleaf != 1 || fb->col.nleaf != 1 || fa->col.leaf[0].kind != 1 || fb->col.leaf[0].kind != 1 ||
      fa->col.leaf[0].offset || fb->col.leaf[0].offset)
    return fail(CLAY_ERR_INVALID, "lerpField: both fields must be plain volume fields (a morph pair)");
  const Volume* va = fa->vols[0];
  const Volume* vb = fb->vols[0];
  if (va->info.resolution != vb->info.resolution || va->info.mip_levels != vb->info.mip_levels ||
      va->instances != vb->instances)
    return fail(CLAY_ERR_INVALID, "lerpField: the two volumes must have the same resolution, mips and instances");
  Field* f = new (std::nothrow) Field();
  if (!f) return fail(CLAY_ERR_INVALID, "out of host memory");
  f->ctx = fa->ctx;
  f->max_mip = std::max(fa->max_mip, fb->max_mip);
  f->col.nleaf = 1;
  LeafDev& l = f->col.leaf[0];
  l = fa->col.leaf[0];
  const LeafDev& lb = fb->col.leaf[0];
  l.kind = 2;
  l.vol_b = lb.vol;
  l.dist0_b = lb.dist0;
  l.mat0_b = lb.mat0;
  l.lerp_t = t_dev;
  f->vols = {fa->vols[0]};
  f->vols_b = {fb->vols[0]};
  f->bands = {0.0};
  f->eps = {0.0};
  f->children = {fa, fb};
  *out = reinterpret_cast<clay_field*>(f);
  return CLAY_OK;
}

extern "C" int clay_field_destroy(clay_field* fh) {
  Field* f = reinterpret_cast<Field*>(fh);
  if (!f) return CLAY_OK;
  dfree(f->brushes_dev);
  delete f;
  return CLAY_OK;
}

extern "C" int clay_field_max_mip(const clay_field* fh, int32_t* out) {
  if (!fh || !out) return fail(CLAY_ERR_INVALID, "clay_field_max_mip: NULL argument");
  *out = reinterpret_cast<const Field*>(fh)->max_mip;
  return CLAY_OK;
}

extern "C" int clay_field_eval(clay_field* fh, const float* pts, int32_t n, float* dist, float* nrm) {
  Field* f = reinterpret_cast<Field*>(fh);
  if (!f || (n > 0 && (!pts || !dist || !nrm))) return fail(CLAY_ERR_INVALID, "clay_field_eval: NULL argument");
  if (n <= 0) return CLAY_OK;
  float *dp = nullptr, *dd = nullptr, *dn = nullptr;
  int rc;
  if ((rc = dalloc(&dp, (size_t)n * 3)) || (rc = dalloc(&dd, (size_t)n)) || (rc = dalloc(&dn, (size_t)n * 3))) {
    dfree(dp);
    dfree(dd);
    dfree(dn);
    return rc;
  }
  hipStream_t s = f->ctx->stream;
  hipError_t e = hipMemcpyAsync(dp, pts, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) rc = launch_field_eval(f->ctx, s, f->col, n, dp, dd, dn);
  if (e == hipSuccess && !rc) e = hipMemcpyAsync(dist, dd, sizeof(float) * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && !rc) e = hipMemcpyAsync(nrm, dn, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(dp);
  dfree(dd);
  dfree(dn);
  if (e != hipSuccess) return fail(CLAY_ERR_HIP, std::string("clay_field_eval: ") + hipGetErrorString(e));
  return rc;
}

// ============================================================================
// SplatField
// ============================================================================
namespace {

void splat_free(Splat* sp) {
  if (!sp) return;
  if (sp->field) delete sp->field;
  volume_destroy_impl(sp->vol);
  dfree(sp->grid);
  dfree(sp->brick_count);
  dfree(sp->brick_start);
  dfree(sp->brick_cursor);
  dfree(sp->brick_sub);
  dfree(sp->brick_list);
  dfree(sp->brick_total_dev);
  dfree(sp->item_start);
  dfree(sp->tile_live);
  dfree(sp->own_scene_of);
  dfree(sp->own_scene_base);
  dfree(sp->follow_slots);
  delete sp;
}

// One SplatField, or (scenes > 1, or a follow splat) one volume instance per scene
// of a batch whose particle scenes are `scene_of` with prefix offsets `scene_base`.
struct SplatScenes {
  int scenes = 1;
  const uint32_t* scene_of = nullptr;
  const uint32_t* scene_base = nullptr;
  bool follow = false;
};

int splat_create_impl(Ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc, Splat** out,
                      const SplatScenes& sc = SplatScenes()) {
  if (!(desc->radius > 0)) return fail(CLAY_ERR_INVALID, "SplatField: radius must be positive");
  if (sc.scenes > 1 && desc->mode != CLAY_SPLAT_TILED)
    return fail(CLAY_ERR_INVALID, "SplatField: a batch of scenes needs the tiled schedule");
  if (desc->mode != CLAY_SPLAT_PLAIN && desc->mode != CLAY_SPLAT_TILED)
    return fail(CLAY_ERR_INVALID, "SplatField: unknown splat mode");
  if (!cloud->pos[0] || !cloud->pos[1] || !cloud->pos[2])
    return fail(CLAY_ERR_INVALID, "SplatField: the cloud has no position arrays");
  const bool n0 = cloud->nrm[0] != nullptr;
  if ((cloud->nrm[1] != nullptr) != n0 || (cloud->nrm[2] != nullptr) != n0)
    return fail(CLAY_ERR_INVALID, "SplatField: normals must be given for all three axes or none");
  Splat* sp = new (std::nothrow) Splat();
  if (!sp) return fail(CLAY_ERR_INVALID, "out of host memory");
  sp->ctx = ctx;
  sp->cloud = *cloud;
  sp->desc = *desc;
  sp->scenes = sc.scenes;
  sp->scene_of = sc.scene_of;
  sp->scene_base = sc.scene_base;
  sp->follow = sc.follow;
  sp->persistent_blocks = ctx->persistent_blocks;
  // splat.ts:148-153
  const int res = desc->resolution > 0 ? desc->resolution : 64;
  const int mips = desc->mip_levels > 0 ? desc->mip_levels : std::max(1, (int)log2((double)res / kTile));
  sp->desc.resolution = res;
  sp->desc.mip_levels = mips;
  sp->desc.band = desc->band > 0 ? desc->band : 2.0;
  sp->desc.world_size = desc->world_size > 0 ? desc->world_size : 24.0;
  clay_volume_desc vd{};
  vd.resolution = res;
  vd.mip_levels = mips;
  vd.world_size = sp->desc.world_size;
  vd.band = sp->desc.band;
  for (int a = 0; a < 3; a++) vd.origin[a] = desc->origin[a];
  int rc = volume_create_impl(ctx, &vd, &sp->vol, sc.scenes);
  if (rc) {
    delete sp;
    return rc;
  }
  sp->mips = sp->vol->info.mip_levels;
  sp->field = make_volume_field(ctx, sp->vol);
  size_t total = 0;
  for (int m = 0; m < sp->mips; m++) {
    const size_t r = (size_t)sp->vol->mip_res(m);
    sp->grid_off.push_back(total);
    total += r * r * r;
  }
  sp->grid_total = total;
  hipStream_t s = ctx->stream;
  // Both schedules keep the u32 code grid: plain accumulates codes in it, tiled
  // accumulates squared distances of the bricks it touches and resolves them to
  // codes in place, so after every bake it holds the codes of every voxel.
  const size_t S = (size_t)sc.scenes;
  if ((rc = dalloc(&sp->grid, total * S)) || (rc = launch_fill_u32(ctx, s, sp->grid, total * S, kEncSat))) {
    splat_free(sp);
    return rc;
  }
  if (desc->mode == CLAY_SPLAT_TILED) {
    splat_brick_geometry(res, sp->mips, &sp->brick_res, &sp->brick_off, &sp->brick_total, &sp->tile_off,
                         &sp->tile_total);
    const size_t per = splat_max_bricks_per_particle(sp->desc.radius, sp->desc.band,
                                                     sp->vol->info.voxel_size, sp->mips);
    sp->brick_list_cap = (size_t)std::max<uint32_t>(cloud->capacity, 1) * per;
    const size_t nb = sp->brick_total * S;
    if ((rc = dalloc(&sp->brick_count, nb)) || (rc = dalloc(&sp->brick_start, nb)) ||
        (rc = dalloc(&sp->brick_cursor, nb * splat_bin_slices())) ||
        (rc = dalloc(&sp->brick_sub, nb * splat_bin_slices())) ||
        (rc = dalloc(&sp->brick_list, sp->brick_list_cap)) ||
        (rc = dalloc(&sp->brick_total_dev, 4)) || (rc = dalloc(&sp->item_start, nb)) ||
        (rc = dalloc(&sp->tile_live, sp->tile_total * S))) {
      splat_free(sp);
      return rc;
    }
    if (hipMemsetAsync(sp->tile_live, 0, sp->tile_total * S, s) != hipSuccess) {
      splat_free(sp);
      return fail(CLAY_ERR_HIP, "SplatField: tile mask initialisation failed");
    }
  }
  if (sc.follow && ((rc = dalloc(&sp->follow_slots, (size_t)kBoundsStride * S)) || (rc = splat_bounds_reset(sp, s)))) {
    splat_free(sp);
    return rc;
  }
  // The material channel is constant for a splat (splat.ts:273): written once.
  if ((rc = fill_u16(ctx, s, sp->vol->mat, sp->vol->total_voxels * S, f2h_host((float)desc->material)))) {
    splat_free(sp);
    return rc;
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    splat_free(sp);
    return fail(CLAY_ERR_HIP, "SplatField: initialisation failed");
  }
  *out = sp;
  return CLAY_OK;
}

}  // namespace

extern "C" int clay_splat_create(clay_ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc,
                                 clay_splat** out) {
  if (!ctx || !cloud || !desc || !out) return fail(CLAY_ERR_INVALID, "clay_splat_create: NULL argument");
  Splat* sp = nullptr;
  int rc = splat_create_impl(reinterpret_cast<Ctx*>(ctx), cloud, desc, &sp);
  if (rc) return rc;
  *out = reinterpret_cast<clay_splat*>(sp);
  return CLAY_OK;
}

extern "C" int clay_splat_destroy(clay_splat* s) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (sp) (void)hipStreamSynchronize(sp->ctx->stream);
  splat_free(sp);
  return CLAY_OK;
}

extern "C" int clay_splat_set_live(clay_splat* s, uint32_t first, uint32_t count) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp) return fail(CLAY_ERR_INVALID, "clay_splat_set_live: NULL splat");
  sp->cloud.live_first = first;
  sp->cloud.live_count = count;
  return CLAY_OK;
}

extern "C" int clay_splat_bake(clay_splat* s, clay_stream stream) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp) return fail(CLAY_ERR_INVALID, "clay_splat_bake: NULL splat");
  return splat_bake(sp, pick(sp->ctx, stream));
}

extern "C" int clay_splat_bake_static(clay_ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc,
                                      uint16_t* dist_dev, clay_stream stream) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !cloud || !desc || !dist_dev) return fail(CLAY_ERR_INVALID, "clay_splat_bake_static: NULL argument");
  return splat_bake_static(c, cloud, desc, dist_dev, pick(c, stream));
}

extern "C" int clay_splat_set_center(clay_splat* s, const double c[3], clay_stream stream) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp || !c) return fail(CLAY_ERR_INVALID, "clay_splat_set_center: NULL argument");
  const double half = sp->vol->info.world_size / 2;  // splat.ts:284-287
  const double o[3] = {c[0] - half, c[1] - half, c[2] - half};
  return volume_set_origin_impl(sp->vol, o, pick(sp->ctx, stream));
}

extern "C" int clay_splat_set_tuning(clay_splat* s, const clay_tuning* t) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp || !t) return fail(CLAY_ERR_INVALID, "clay_splat_set_tuning: NULL argument");
  if (t->chunk < 256 || t->chunk > 65536) return fail(CLAY_ERR_INVALID, "clay_tuning: chunk must be in [256, 65536]");
  if (t->groups_per_cu < 1 || t->groups_per_cu > 32)
    return fail(CLAY_ERR_INVALID, "clay_tuning: groups_per_cu must be in [1, 32]");
  if ((t->splat_xcd != 0 && t->splat_xcd != 1) || (t->splat_owner != 0 && t->splat_owner != 1))
    return fail(CLAY_ERR_INVALID, "clay_tuning: splat_xcd and splat_owner must be 0 or 1");
  sp->chunk = (uint32_t)t->chunk;
  sp->persistent_blocks = sp->ctx->cus * t->groups_per_cu;
  sp->xcd = t->splat_xcd != 0;
  sp->owner = t->splat_owner != 0;
  return CLAY_OK;
}

extern "C" int clay_splat_volume(clay_splat* s, clay_volume** out) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp || !out) return fail(CLAY_ERR_INVALID, "clay_splat_volume: NULL argument");
  *out = reinterpret_cast<clay_volume*>(sp->vol);
  return CLAY_OK;
}

extern "C" int clay_splat_field(clay_splat* s, clay_field** out) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp || !out) return fail(CLAY_ERR_INVALID, "clay_splat_field: NULL argument");
  *out = reinterpret_cast<clay_field*>(sp->field);
  return CLAY_OK;
}

extern "C" int clay_splat_grid_download(clay_splat* s, int32_t mip, uint32_t* host) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp || !host) return fail(CLAY_ERR_INVALID, "clay_splat_grid_download: NULL argument");
  if (mip < 0 || mip >= sp->mips) return fail(CLAY_ERR_INVALID, "clay_splat_grid_download: bad mip");
  if (!sp->grid) return fail(CLAY_ERR_STATE, "clay_splat_grid_download: tiled splat without keep_codes has no code grid");
  const size_t r = (size_t)sp->vol->mip_res(mip), n = r * r * r;
  CLAY_HIP_CHECK(hipStreamSynchronize(sp->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(host, sp->grid + sp->grid_off[mip], n * 4, hipMemcpyDeviceToHost));
  return CLAY_OK;
}

extern "C" int clay_splat_tiles_download(clay_splat* s, int32_t mip, uint8_t* host) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp || !host) return fail(CLAY_ERR_INVALID, "clay_splat_tiles_download: NULL argument");
  if (mip < 0 || mip >= sp->mips) return fail(CLAY_ERR_INVALID, "clay_splat_tiles_download: bad mip");
  const int r = sp->vol->mip_res(mip), tr = r / kTile;
  const size_t nt = (size_t)tr * tr * tr;
  CLAY_HIP_CHECK(hipStreamSynchronize(sp->ctx->stream));
  if (sp->desc.mode == CLAY_SPLAT_TILED) {
    CLAY_HIP_CHECK(hipMemcpy(host, sp->tile_live + sp->tile_off[mip], nt, hipMemcpyDeviceToHost));
    return CLAY_OK;
  }
  // Plain mode: derive the exact mask from the dense code grid.
  std::vector<uint32_t> g((size_t)r * r * r);
  CLAY_HIP_CHECK(hipMemcpy(g.data(), sp->grid + sp->grid_off[mip], g.size() * 4, hipMemcpyDeviceToHost));
  memset(host, 0, nt);
  for (int z = 0; z < r; z++)
    for (int y = 0; y < r; y++)
      for (int x = 0; x < r; x++)
        if (g[(size_t)x + (size_t)r * ((size_t)y + (size_t)r * z)] != kEncSat)
          host[(size_t)(x / kTile) + (size_t)tr * ((size_t)(y / kTile) + (size_t)tr * (z / kTile))] = 1;
  return CLAY_OK;
}

// ============================================================================
// FluidSim
// ============================================================================
extern "C" void clay_fluid_desc_default(clay_fluid_desc* d) {
  if (!d) return;
  memset(d, 0, sizeof(*d));
  d->capacity = 8192;
  d->bucket = 64;
  d->spacing = 0.45;
  d->smoothing = -1;
  d->density = 1.0;
  d->radius = -1;
  d->surface_radius = -1;
  d->material = 0;
  d->bake_resolution = 64;
  d->bake_mip_levels = 0;
  d->bake_band = 2;
  d->origin[0] = -12;
  d->origin[1] = -4;
  d->origin[2] = -12;
  d->world_size = 24;
  d->substeps = 3;
  d->splat_mode = CLAY_SPLAT_PLAIN;
  d->gravity[0] = 0;
  d->gravity[1] = -9.81;
  d->gravity[2] = 0;
  d->stiffness = 220;
  d->viscosity = 0.4;
  d->damping = 0.995;
  d->max_speed = 15;
}

namespace {

void fluid_free(Fluid* f) {
  if (!f) return;
  splat_free(f->surface);
  dfree(f->st);
  dfree(f->ss);
  dfree(f->keys);
  dfree(f->vals);
  dfree(f->keys2);
  dfree(f->vals2);
  dfree(f->hist);
  dfree(f->cells);
  dfree(f->contacts);
  dfree(f->cmask);
  dfree(f->cdata);
  dfree(f->cover);
  dfree(f->spos);
  dfree(f->frec);
  dfree(f->reduce);
  dfree(f->epoch_dev);
  dfree(f->scene_of);
  dfree(f->scene_base_dev);
  if (f->graph) (void)hipGraphExecDestroy(f->graph);
  if (f->ev_fork) (void)hipEventDestroy(f->ev_fork);
  if (f->ev_join) (void)hipEventDestroy(f->ev_join);
  if (f->side) (void)hipStreamDestroy(f->side);
  delete f;
}

size_t contact_words(const Fluid* f) { return (size_t)f->scenes * (4 + 4 * CLAY_CONTACTS); }

// Waits for everything a fluid may still have in flight: the context stream, the
// stream of the latest step (a caller's stream) and the stream a graph replays on.
hipError_t fluid_quiesce(const Fluid* f) {
  hipError_t e = hipStreamSynchronize(f->ctx->stream);
  if (e == hipSuccess && f->last_stream && f->last_stream != f->ctx->stream) e = hipStreamSynchronize(f->last_stream);
  if (e == hipSuccess && f->graph_stream && f->graph_stream != f->ctx->stream && f->graph_stream != f->last_stream)
    e = hipStreamSynchronize(f->graph_stream);
  return e;
}

// First half of a substep (fluid.ts:293-368): cell stamp, keys, sort, cell
// table and reordered state, then densities and pressures.  With `cover`, the
// density pass also writes each particle's interior radius for the surface bake.
// The cell sort of a substep: stamp, keys, sort, reorder into the sorted planes.
int fluid_reorder(Fluid* f, hipStream_t s);

// The cell sort of a substep: stamp, keys, sort, and (unless `reorder` is false, when
// the caller runs fluid_reorder itself) the reorder into the sorted planes.
int fluid_sort(Fluid* f, hipStream_t s, uint32_t* bounds = nullptr, bool reorder = true) {
  Ctx* c = f->ctx;
  const uint32_t n = f->live;
  const uint32_t invalid = (uint32_t)f->grid_cells;
  int rc;
  if (n == 0) return CLAY_OK;
  // A new stamp per substep makes every older cell record stale at once.  After
  // 2^32 - 1 substeps the stamps restart, and the table is wiped once so that no
  // record from the previous cycle can carry a reused stamp.
  // The keys pass advances the device stamp (except on the wrap, which sets it).
  uint32_t* bump = f->epoch_dev;
  if (++f->epoch == 0) {
    if ((rc = launch_fill_u32(c, s, f->cells, 4 * f->grid_cells * f->scenes, 0u)) ||
        (rc = launch_fill_u32(c, s, f->epoch_dev, 1, 1u)))
      return rc;
    f->epoch = 1;
    bump = nullptr;
  }
  const bool in_tmp = radix_input_in_tmp(f->key_bits);
  if ((rc = launch_keys(c, s, f->k, n, f->px, f->py, f->pz, f->scene_of, in_tmp ? f->keys2 : f->keys,
                        in_tmp ? f->vals2 : f->vals, invalid, f->scene_shift, bounds, bump)))
    return rc;
  if ((rc = radix_sort_pairs(c, s, n, f->key_bits, f->keys, f->vals, f->keys2, f->vals2, f->hist))) return rc;
  return reorder ? fluid_reorder(f, s) : CLAY_OK;
}

// Cell records and the sorted snapshot of the state, from the sort's permutation.
int fluid_reorder(Fluid* f, hipStream_t s) {
  Ctx* c = f->ctx;
  const uint32_t n = f->live;
  if (n == 0) return CLAY_OK;
  const uint32_t invalid = (uint32_t)f->grid_cells;
  const SceneK sk = f->scene_k();
  const float* src[6] = {f->px, f->py, f->pz, f->vx, f->vy, f->vz};
  float* dst[6] = {f->spx, f->spy, f->spz, f->svx, f->svy, f->svz};
  const bool lds = f->desc.precision == CLAY_PRECISION_FP16LDS;
  return launch_reorder(c, s, n, f->keys, f->vals, sk, invalid, f->cells, src, dst, 6, f->k,
                        lds ? f->spos : nullptr);
}

// The density pass over the sorted planes (with `cover`, also the interior radii).
int fluid_density(Fluid* f, hipStream_t s, float* cover) {
  const uint32_t n = f->live;
  if (n == 0) return CLAY_OK;
  const SceneK sk = f->scene_k();
  const bool lds = f->desc.precision == CLAY_PRECISION_FP16LDS;
  HalfState half;
  half.spos = f->spos;
  half.svx = f->svx;
  half.svy = f->svy;
  half.svz = f->svz;
  half.frec = f->frec;
  return launch_density(f->ctx, s, f->k, sk, f->tune, lds, n, f->spx, f->spy, f->spz, f->keys, f->cells,
                        f->srho, f->spres, f->vals, cover, half);
}

int fluid_sort_density(Fluid* f, hipStream_t s, float* cover) {
  const int rc = fluid_sort(f, s);
  return rc ? rc : fluid_density(f, s, cover);
}

// Second half of a substep (fluid.ts:370-491): clearGrid's contacts reset
// (fluid.ts:302-304) for every scene and the per-id flags, then forces and
// integrate over the sorted state and cell table of fluid_sort_density.
int fluid_forces(Fluid* f, hipStream_t s) {
  Ctx* c = f->ctx;
  const uint32_t n = f->live;
  int rc;
  if ((rc = launch_fill2_u32(c, s, f->contacts, contact_words(f), f->cmask, ((size_t)f->capacity + 31) / 32, 0u)))
    return rc;
  if (n == 0) return CLAY_OK;
  const SceneK sk = f->scene_k();
  const bool lds = f->desc.precision == CLAY_PRECISION_FP16LDS;
  const float* sorted[8] = {f->spx, f->spy, f->spz, f->svx, f->svy, f->svz, f->srho, f->spres};
  float* out[11] = {f->px, f->py, f->pz, f->vx, f->vy, f->vz, f->rho, f->pres, f->nvx, f->nvy, f->nvz};
  ColliderDev none{};
  const ColliderDev& col = f->world ? f->world->col : none;
  if ((rc = launch_forces_integrate(c, s, f->k, sk, f->tune, lds, col, n, sorted, f->vals, f->keys, f->cells, out,
                                    f->contacts, f->frec)))
    return rc;
  return launch_contacts_collect(c, s, f->contacts, 4 + 4 * CLAY_CONTACTS, f->cmask, f->cdata,
                                 f->scene_of ? f->scene_base_dev : nullptr, n, f->scenes);
}

// FluidSim.step (fluid.ts:536-558): emit, `substeps` substeps, bake every mip.
int fluid_frame_impl(Fluid* f, hipStream_t s) {
  const uint32_t emit_count = f->pending_emit;
  f->pending_emit = 0;
  const uint32_t emit_base = f->cursor;
  if (emit_count > 0) {
    f->cursor = (emit_base + emit_count) % f->capacity;
    f->live = std::min<uint32_t>(f->capacity, std::max<uint32_t>(f->live, emit_base + emit_count));
    int rc = launch_emit(f->ctx, s, f, emit_base, emit_count, f->emit_seed);
    if (rc) return rc;
  }
  if (emit_count > 0) f->presorted = false;
  if (f->live > 0 && f->consts.substeps > 0) {
    // Substep i sorts and computes densities from the positions substep i - 1 left;
    // the first one's were computed at the end of the previous frame unless the
    // state has changed since.  The frame ends with the next frame's first sort
    // and densities, from the positions the bake sees (the same values the next
    // frame would compute first), so the bake can use its cell table.
    int rc;
    if (!f->presorted && (rc = fluid_sort_density(f, s, nullptr))) return rc;
    for (int i = 0; i < f->consts.substeps; i++) {
      if (i > 0 && (rc = fluid_sort_density(f, s, nullptr))) return rc;
      if ((rc = fluid_forces(f, s))) return rc;
    }
    // The frame-end sort, then its density pass beside the bake's binning, which
    // reads the id-order positions and the sort's permutation but not the sorted
    // planes or the interior radii: it forks after the reorder, or (tuning early_bin)
    // as soon as the permutation is final, beside the reorder too.  The permutation
    // is a spatially coherent visiting order for the splat; the codes do not depend
    // on it.
    // The keys pass of this sort also folds the bounds a following surface reads.
    uint32_t* bounds = splat_bounds_begin(f->surface, s);
    const bool early = f->tune.early_bin;
    if ((rc = fluid_sort(f, s, bounds, !early))) return rc;
    f->surface->cloud.live_count = f->live;
    CLAY_HIP_CHECK(hipEventRecord(f->ev_fork, s));
    CLAY_HIP_CHECK(hipStreamWaitEvent(f->side, f->ev_fork, 0));
    if ((rc = splat_bake_prepare(f->surface, f->side, f->vals, bounds != nullptr))) return rc;
    if (early && (rc = fluid_reorder(f, s))) return rc;
    if ((rc = fluid_density(f, s, f->cover))) return rc;
    CLAY_HIP_CHECK(hipEventRecord(f->ev_join, f->side));
    CLAY_HIP_CHECK(hipStreamWaitEvent(s, f->ev_join, 0));
    f->presorted = true;
    return splat_bake_finish(f->surface, s, f->vals, f->cover);
  }
  f->surface->cloud.live_count = f->live;
  return splat_bake_ordered(f->surface, s, nullptr, nullptr);
}

// Shared by FluidSim (one scene, `counts` null) and a batch of `scenes` scenes
// holding counts[i] particles each, scene-major.
int fluid_create_impl(Ctx* c, Field* world, const clay_fluid_desc* d, uint32_t scenes, const uint32_t* counts,
                      Fluid** out) {
  if (!(d->spacing > 0)) return fail(CLAY_ERR_INVALID, "FluidSim: spacing must be positive");
  if (d->precision != CLAY_PRECISION_F32 && d->precision != CLAY_PRECISION_FP16LDS)
    return fail(CLAY_ERR_INVALID, "FluidSim: unknown precision mode");
  if (d->follow_size < 0) return fail(CLAY_ERR_INVALID, "FluidSim: followSize must not be negative");
  const bool batch = counts != nullptr;
  uint64_t total = d->capacity;
  if (batch) {
    total = 0;
    for (uint32_t i = 0; i < scenes; i++) total += counts[i];
    if (total == 0) return fail(CLAY_ERR_INVALID, "FluidBatch: the scenes hold no particles");
    if (total > 0x7FFFFFFFull) return fail(CLAY_ERR_INVALID, "FluidBatch: more than 2^31 particles");
  }
  if (total == 0) return fail(CLAY_ERR_INVALID, "FluidSim: capacity must be positive");
  Fluid* f = new (std::nothrow) Fluid();
  if (!f) return fail(CLAY_ERR_INVALID, "out of host memory");
  f->ctx = c;
  f->world = world;
  f->desc = *d;
  f->capacity = (uint32_t)total;
  f->scenes = batch ? scenes : 1;
  f->use_graph = batch;
  // fluid.ts:141-203
  const double spacing = d->spacing;
  const double h = d->smoothing > 0 ? d->smoothing : spacing * 2.2;
  const int bucket = d->bucket > 0 ? d->bucket : 64;
  const double radius = d->radius >= 0 ? d->radius : spacing * 0.5;
  const double surf_r = d->surface_radius >= 0 ? d->surface_radius : spacing * 1.1;
  const int bake_res = d->bake_resolution > 0 ? d->bake_resolution : 64;
  const int bake_mips = d->bake_mip_levels > 0 ? d->bake_mip_levels : std::max(1, (int)log2((double)bake_res / kTile));
  const double band = d->bake_band > 0 ? d->bake_band : 2.0;
  const int substeps = d->substeps > 0 ? d->substeps : 3;
  const double density = d->density > 0 ? d->density : 1.0;
  clay_fluid_consts& k = f->consts;
  k.h = h;
  k.bucket = bucket;
  k.substeps = substeps;
  k.radius = radius;
  k.poly6 = 315.0 / (64.0 * M_PI * pow(h, 9));
  k.spiky_grad = -45.0 / (M_PI * pow(h, 6));
  k.mass = density * pow(spacing, 3);
  double rest = 0;
  const int reach = (int)ceil(h / spacing);
  for (int z = -reach; z <= reach; z++)
    for (int y = -reach; y <= reach; y++)
      for (int x = -reach; x <= reach; x++) {
        const double r2 = (double)(x * x + y * y + z * z) * spacing * spacing;
        if (r2 < h * h) rest += k.mass * k.poly6 * pow(h * h - r2, 3);
      }
  k.rest_density = rest;
  k.dt = 1.0 / 60.0 / substeps;
  k.contact_depth = radius * 1.5;
  f->hh = h * h;
  f->mass_poly6 = k.mass * k.poly6;
  f->h6 = pow(h, 6);
  f->inv_h = 1.0 / h;
  f->rest_floor = rest * 0.1;
  clay_fluid_params& p = f->params;
  for (int a = 0; a < 3; a++) {
    p.gravity[a] = d->gravity[a];
    p.emit_pos[a] = 0;
    p.emit_vel[a] = 0;
  }
  p.dt = k.dt;
  p.emit_radius = spacing;
  p.max_speed = d->max_speed;
  p.stiffness = d->stiffness;
  p.viscosity = d->viscosity;
  p.damping = d->damping;
  p.contact_depth = k.contact_depth;

  int rc;
  const uint32_t P = f->capacity;
  f->scene_base.assign(f->scenes + 1, 0);
  if (batch)
    for (uint32_t i = 0; i < scenes; i++) f->scene_base[i + 1] = f->scene_base[i] + counts[i];
  else
    f->scene_base[1] = P;
  if ((rc = dalloc(&f->st, (size_t)P * 11)) || (rc = dalloc(&f->ss, (size_t)P * 8)) ||
      (rc = dalloc(&f->keys, P)) || (rc = dalloc(&f->vals, P)) || (rc = dalloc(&f->keys2, P)) ||
      (rc = dalloc(&f->vals2, P)) || (rc = dalloc(&f->contacts, contact_words(f))) ||
      (rc = dalloc(&f->epoch_dev, 1)) || (rc = dalloc(&f->cmask, ((size_t)P + 31) / 32)) ||
      (rc = dalloc(&f->cdata, P)) || (rc = dalloc(&f->cover, P))) {
    fluid_free(f);
    return rc;
  }
  if (d->precision == CLAY_PRECISION_FP16LDS &&
      ((rc = dalloc(&f->spos, P)) || (rc = dalloc(&f->frec, (size_t)P * 2)))) {
    fluid_free(f);
    return rc;
  }
  if (batch) {
    std::vector<uint32_t> so(P);
    for (uint32_t i = 0; i < scenes; i++)
      for (uint32_t j = f->scene_base[i]; j < f->scene_base[i + 1]; j++) so[j] = i;
    if ((rc = dalloc(&f->scene_of, P)) || (rc = dalloc(&f->scene_base_dev, (size_t)scenes + 1))) {
      fluid_free(f);
      return rc;
    }
    if (hipMemcpy(f->scene_of, so.data(), sizeof(uint32_t) * P, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->scene_base_dev, f->scene_base.data(), sizeof(uint32_t) * (scenes + 1),
                  hipMemcpyHostToDevice) != hipSuccess) {
      fluid_free(f);
      return fail(CLAY_ERR_HIP, "FluidBatch: scene table upload failed");
    }
  }
  f->reduce_blocks = std::min<uint32_t>(1024, std::max<uint32_t>(1, (P + 255) / 256));
  if ((rc = dalloc(&f->reduce, (size_t)f->reduce_blocks * 6 + 8))) {
    fluid_free(f);
    return rc;
  }
  float** sp_[11] = {&f->px, &f->py, &f->pz, &f->vx, &f->vy, &f->vz, &f->rho, &f->pres, &f->nvx, &f->nvy, &f->nvz};
  for (int i = 0; i < 11; i++) *sp_[i] = f->st + (size_t)i * P;
  float** ss_[8] = {&f->spx, &f->spy, &f->spz, &f->svx, &f->svy, &f->svz, &f->srho, &f->spres};
  for (int i = 0; i < 8; i++) *ss_[i] = f->ss + (size_t)i * P;
  f->hist_cap = radix_hist_size(P);
  if ((rc = dalloc(&f->hist, f->hist_cap))) {
    fluid_free(f);
    return rc;
  }
  // The renderable surface (fluid.ts:215-234): a SplatField over the live prefix.
  // With followSize set it is a volume of that size centred on the particles'
  // bounds before every bake (one per scene in a batch), as SoftBody does for a body
  // (softbody.ts:424-428).
  clay_cloud_desc cloud{};
  cloud.capacity = P;
  cloud.live_rule = CLAY_LIVE_PREFIX;
  cloud.live_count = batch ? P : 0;
  cloud.pos[0] = f->px;
  cloud.pos[1] = f->py;
  cloud.pos[2] = f->pz;
  clay_splat_desc sd{};
  sd.radius = surf_r;
  sd.material = d->material;
  sd.resolution = bake_res;
  sd.mip_levels = bake_mips;
  sd.band = band;
  for (int a = 0; a < 3; a++) sd.origin[a] = d->origin[a];
  sd.world_size = d->world_size > 0 ? d->world_size : 24.0;
  if (d->follow_size > 0) sd.world_size = d->follow_size;
  sd.mode = d->splat_mode;
  sd.keep_codes = 0;
  SplatScenes sc;
  sc.scenes = (int)f->scenes;
  sc.scene_of = f->scene_of;
  sc.scene_base = f->scene_base_dev;
  sc.follow = d->follow_size > 0;
  if ((rc = splat_create_impl(c, &cloud, &sd, &f->surface, sc))) {
    fluid_free(f);
    return rc;
  }
  // Hash grid (fluid.ts:239-247) and walls (fluid.ts:444-447) over the world box.
  const double ws = d->world_size > 0 ? d->world_size : 24.0;
  k.grid_res = (int)ceil(ws / h) + 2;
  for (int a = 0; a < 3; a++) {
    k.grid_origin[a] = d->origin[a] - h;
    k.wall_lo[a] = d->origin[a] + radius;
    k.wall_hi[a] = d->origin[a] + ws - radius;
  }
  const uint64_t G = (uint64_t)k.grid_res * k.grid_res * k.grid_res;
  if (G >= 0xFFFFFFFFull) {
    fluid_free(f);
    return fail(CLAY_ERR_INVALID, "FluidSim: worldSize / smoothing gives a hash grid beyond 2^32 cells");
  }
  if (batch && G > (1ull << kMaxSceneShift) - 1) {
    fluid_free(f);
    return fail(CLAY_ERR_INVALID, "FluidBatch: worldSize / smoothing gives a hash grid beyond 2^24 - 1 cells");
  }
  f->grid_cells = (size_t)G;
  // A batch key is scene << shift | cell, with the all-ones cell field kept free
  // for particles outside the grid: the fewest bits with 2^shift - 1 >= G.
  if (batch) {
    uint32_t sh = 1;
    while ((1ull << sh) - 1 < G) sh++;
    f->scene_shift = sh;
  }
  const uint64_t max_key = batch ? ((uint64_t)(scenes - 1) << f->scene_shift) | ((1ull << f->scene_shift) - 1) : G;
  int bits = 1;
  while (bits < 32 && (1ull << bits) <= max_key) bits++;
  f->key_bits = bits;
  if ((rc = dalloc(&f->cells, 4 * G * f->scenes))) {
    fluid_free(f);
    return rc;
  }
  hipStream_t s = c->stream;
  if ((rc = launch_fill_u32(c, s, f->cells, 4 * G * f->scenes, 0u)) ||
      (rc = launch_fill_u32(c, s, reinterpret_cast<uint32_t*>(f->st), (size_t)P * 11, 0u)) ||
      (rc = launch_fill_u32(c, s, f->contacts, contact_words(f), 0u)) ||
      (rc = launch_fill_u32(c, s, f->epoch_dev, 1, 0u))) {
    fluid_free(f);
    return rc;
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    fluid_free(f);
    return fail(CLAY_ERR_HIP, "FluidSim: initialisation failed");
  }
  if (hipStreamCreateWithFlags(&f->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&f->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&f->ev_join, hipEventDisableTiming) != hipSuccess) {
    fluid_free(f);
    return fail(CLAY_ERR_HIP, "FluidSim: stream creation failed");
  }
  if (batch) f->live = P;
  fluid_build_k(f);
  *out = f;
  return CLAY_OK;
}

}  // namespace

extern "C" int clay_fluid_create(clay_ctx* ctx, clay_field* world, const clay_fluid_desc* d, clay_fluid** out) {
  if (!ctx || !d || !out) return fail(CLAY_ERR_INVALID, "clay_fluid_create: NULL argument");
  if (d->capacity == 0) return fail(CLAY_ERR_INVALID, "FluidSim: capacity must be positive");
  Fluid* f = nullptr;
  int rc = fluid_create_impl(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<Field*>(world), d, 1, nullptr, &f);
  if (rc) return rc;
  *out = reinterpret_cast<clay_fluid*>(f);
  return CLAY_OK;
}

extern "C" int clay_fluid_destroy(clay_fluid* fh) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (f) (void)fluid_quiesce(f);
  fluid_free(f);
  return CLAY_OK;
}

extern "C" int clay_fluid_constants(const clay_fluid* fh, clay_fluid_consts* out) {
  if (!fh || !out) return fail(CLAY_ERR_INVALID, "clay_fluid_constants: NULL argument");
  *out = reinterpret_cast<const Fluid*>(fh)->consts;
  return CLAY_OK;
}

namespace {
void drop_graph(Fluid* f);
int batch_frame(Fluid* f, hipStream_t s);
}  // namespace

extern "C" int clay_fluid_set_state(clay_fluid* fh, const float* pos, const float* vel, uint32_t n) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f || (n > 0 && !pos)) return fail(CLAY_ERR_INVALID, "clay_fluid_set_state: NULL argument");
  if (n > f->capacity) return fail(CLAY_ERR_INVALID, "clay_fluid_set_state: more particles than capacity");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  float* dp[3] = {f->px, f->py, f->pz};
  float* dv[3] = {f->vx, f->vy, f->vz};
  for (int a = 0; a < 3; a++) {
    if (n) CLAY_HIP_CHECK(hipMemcpy(dp[a], pos + (size_t)a * n, sizeof(float) * n, hipMemcpyHostToDevice));
    if (n && vel) CLAY_HIP_CHECK(hipMemcpy(dv[a], vel + (size_t)a * n, sizeof(float) * n, hipMemcpyHostToDevice));
    if (n && !vel) CLAY_HIP_CHECK(hipMemset(dv[a], 0, sizeof(float) * n));
  }
  f->live = n;
  f->cursor = n % f->capacity;
  f->surface->cloud.live_count = n;
  f->presorted = false;
  return CLAY_OK;
}

extern "C" int clay_fluid_get_state(clay_fluid* fh, int32_t what, float* host) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f || !host) return fail(CLAY_ERR_INVALID, "clay_fluid_get_state: NULL argument");
  const uint32_t n = f->live;
  const float* planes[5][3] = {{f->px, f->py, f->pz},
                               {f->vx, f->vy, f->vz},
                               {f->rho, nullptr, nullptr},
                               {f->pres, nullptr, nullptr},
                               {f->nvx, f->nvy, f->nvz}};
  if (what < 0 || what > 4) return fail(CLAY_ERR_INVALID, "clay_fluid_get_state: unknown field");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  for (int a = 0; a < 3; a++) {
    if (!planes[what][a] || !n) continue;
    CLAY_HIP_CHECK(hipMemcpy(host + (size_t)a * n, planes[what][a], sizeof(float) * n, hipMemcpyDeviceToHost));
  }
  return CLAY_OK;
}

extern "C" int clay_fluid_spawn(clay_fluid* fh, uint32_t n, const double pos[3], const double vel[3], double spread) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f || !pos || !vel) return fail(CLAY_ERR_INVALID, "clay_fluid_spawn: NULL argument");
  // fluid.ts:511-525
  if (spread < 0) spread = f->desc.spacing * cbrt((double)std::max<uint32_t>(n, 1)) * 0.62;
  f->pending_emit = std::min(n, f->capacity);
  for (int a = 0; a < 3; a++) {
    f->params.emit_pos[a] = pos[a];
    f->params.emit_vel[a] = vel[a];
  }
  f->params.emit_radius = spread;
  f->emit_seed = f->seed++;  // emitSeed: this.seed++ (fluid.ts:523)
  return CLAY_OK;
}

extern "C" int clay_fluid_configure(clay_fluid* fh, const clay_fluid_params* p) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f || !p) return fail(CLAY_ERR_INVALID, "clay_fluid_configure: NULL argument");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  drop_graph(f);  // the parameters are launch arguments of a captured frame
  f->params = *p;
  fluid_build_k(f);
  f->presorted = false;  // densities of the next substep depend on the parameters
  return CLAY_OK;
}

extern "C" int clay_fluid_get_params(const clay_fluid* fh, clay_fluid_params* out) {
  if (!fh || !out) return fail(CLAY_ERR_INVALID, "clay_fluid_get_params: NULL argument");
  *out = reinterpret_cast<const Fluid*>(fh)->params;
  return CLAY_OK;
}

extern "C" int clay_fluid_substep(clay_fluid* fh, clay_stream stream) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_fluid_substep: NULL fluid");
  fluid_build_k(f);
  hipStream_t s = pick(f->ctx, stream);
  f->last_stream = s;
  int rc;
  if (f->live > 0 && !f->presorted && (rc = fluid_sort_density(f, s, nullptr))) return rc;
  f->presorted = false;  // the forces pass moves the particles
  return fluid_forces(f, s);
}

// With a graph (clay_fluid_set_graph), a frame is captured once and replayed, as a
// batch's is.  Launch sizes follow the live count, so a frame that emits runs
// eagerly and a changed live count recaptures.
extern "C" int clay_fluid_step(clay_fluid* fh, clay_stream stream) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_fluid_step: NULL fluid");
  fluid_build_k(f);
  hipStream_t s = pick(f->ctx, stream);
  f->last_stream = s;
  if (f->use_graph && f->pending_emit == 0 && f->live > 0) {
    if (f->graph && f->graph_live != f->live) {
      CLAY_HIP_CHECK(hipStreamSynchronize(f->graph_stream));
      drop_graph(f);
    }
    const int rc = batch_frame(f, s);
    if (!rc) f->graph_live = f->live;
    return rc;
  }
  return fluid_frame_impl(f, s);
}

namespace {
// The surface bake of the current state again, with the owner walk tallying its
// atomics.  Needs the frame-end pass of a step (its order and interior radii).
int fluid_splat_atomics(Fluid* f, uint64_t out[2], const char* who) {
  if (!out) return fail(CLAY_ERR_INVALID, std::string(who) + ": NULL argument");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  if (f->surface->desc.mode != CLAY_SPLAT_TILED)
    return fail(CLAY_ERR_INVALID, std::string(who) + ": the counts come from the tiled schedule's owner walk");
  if (!f->presorted || f->live == 0)
    return fail(CLAY_ERR_STATE, std::string(who) + ": count after a step (the bake reads its frame-end pass)");
  unsigned long long* d = nullptr;
  int rc = dalloc(&d, 2);
  if (rc) return rc;
  hipStream_t s = f->ctx->stream;
  hipError_t e = hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), s);
  f->surface->atomic_counts = d;
  if (e == hipSuccess) {
    rc = splat_bake_prepare(f->surface, s, f->vals, false);
    if (!rc) rc = splat_bake_finish(f->surface, s, f->vals, f->cover);
  }
  f->surface->atomic_counts = nullptr;
  unsigned long long h[2] = {0, 0};
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  dfree(d);
  if (e != hipSuccess) return fail(CLAY_ERR_HIP, std::string(who) + ": " + hipGetErrorString(e));
  if (rc) return rc;
  out[0] = h[0];
  out[1] = h[1];
  return CLAY_OK;
}
}  // namespace

extern "C" int clay_fluid_splat_atomics(clay_fluid* fh, uint64_t out[2]) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_fluid_splat_atomics: NULL fluid");
  return fluid_splat_atomics(f, out, "clay_fluid_splat_atomics");
}

extern "C" int clay_batch_splat_atomics(clay_batch* b, uint64_t out[2]) {
  Fluid* f = reinterpret_cast<Fluid*>(b);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_batch_splat_atomics: NULL batch");
  return fluid_splat_atomics(f, out, "clay_batch_splat_atomics");
}

extern "C" int clay_fluid_set_graph(clay_fluid* fh, int32_t enabled) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_fluid_set_graph: NULL fluid");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  drop_graph(f);
  f->use_graph = enabled != 0;
  return CLAY_OK;
}

extern "C" int clay_fluid_read_contacts(clay_fluid* fh, float* out4, uint32_t* n, uint32_t* raw) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_fluid_read_contacts: NULL fluid");
  uint32_t buf[4 + 4 * CLAY_CONTACTS];
  CLAY_HIP_CHECK(fluid_quiesce(f));
  CLAY_HIP_CHECK(hipMemcpy(buf, f->contacts, sizeof(buf), hipMemcpyDeviceToHost));
  const uint32_t cnt = std::min<uint32_t>(buf[0], CLAY_CONTACTS);
  if (out4) memcpy(out4, buf + 4, sizeof(float) * 4 * cnt);
  if (n) *n = cnt;
  if (raw) *raw = buf[0];
  return CLAY_OK;
}

extern "C" int clay_fluid_conserved(clay_fluid* fh, double out[7]) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f || !out) return fail(CLAY_ERR_INVALID, "clay_fluid_conserved: NULL argument");
  hipStream_t s = f->ctx->stream;
  const float* pv[6] = {f->px, f->py, f->pz, f->vx, f->vy, f->vz};
  double* res = f->reduce + (size_t)f->reduce_blocks * 6;
  int rc = launch_conserved(f->ctx, s, f->live, pv, f->consts.mass, f->reduce, f->reduce_blocks, res);
  if (rc) return rc;
  CLAY_HIP_CHECK(hipMemcpyAsync(out, res, sizeof(double) * 7, hipMemcpyDeviceToHost, s));
  CLAY_HIP_CHECK(hipStreamSynchronize(s));
  return CLAY_OK;
}

extern "C" int clay_fluid_particle_count(const clay_fluid* fh, uint32_t* out) {
  if (!fh || !out) return fail(CLAY_ERR_INVALID, "clay_fluid_particle_count: NULL argument");
  *out = reinterpret_cast<const Fluid*>(fh)->live;
  return CLAY_OK;
}

extern "C" int clay_fluid_surface(clay_fluid* fh, clay_splat** out) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f || !out) return fail(CLAY_ERR_INVALID, "clay_fluid_surface: NULL argument");
  *out = reinterpret_cast<clay_splat*>(f->surface);
  return CLAY_OK;
}

extern "C" int clay_fluid_debug_download(clay_fluid* fh, int32_t what, uint32_t* host) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f || !host) return fail(CLAY_ERR_INVALID, "clay_fluid_debug_download: NULL argument");
  if (what != CLAY_DEBUG_KEYS && what != CLAY_DEBUG_PERM)
    return fail(CLAY_ERR_INVALID, "clay_fluid_debug_download: unknown buffer");
  CLAY_HIP_CHECK(hipStreamSynchronize(f->ctx->stream));
  if (f->live)
    CLAY_HIP_CHECK(hipMemcpy(host, what == CLAY_DEBUG_KEYS ? f->keys : f->vals, sizeof(uint32_t) * f->live,
                             hipMemcpyDeviceToHost));
  return CLAY_OK;
}

// ============================================================================
// Launch tuning
// ============================================================================
namespace {

int apply_tuning(Fluid* f, const clay_tuning* t) {
  FluidTune ft;
  ft.block = t->block;
  ft.stage = t->stage;
  if (!fluid_tune_valid(ft))
    return fail(CLAY_ERR_INVALID, "clay_tuning: block must be 128 or 256 and stage 64, 128 or 256");
  if (t->chunk < 256 || t->chunk > 65536) return fail(CLAY_ERR_INVALID, "clay_tuning: chunk must be in [256, 65536]");
  if (t->groups_per_cu < 1 || t->groups_per_cu > 32)
    return fail(CLAY_ERR_INVALID, "clay_tuning: groups_per_cu must be in [1, 32]");
  if ((t->fluid_xcd != 0 && t->fluid_xcd != 1) || (t->splat_xcd != 0 && t->splat_xcd != 1) ||
      (t->splat_owner != 0 && t->splat_owner != 1) || (t->exact_lds != 0 && t->exact_lds != 1) ||
      (t->early_bin != 0 && t->early_bin != 1))
    return fail(CLAY_ERR_INVALID,
                "clay_tuning: fluid_xcd, splat_xcd, splat_owner, exact_lds and early_bin must be 0 or 1");
  ft.xcd = t->fluid_xcd != 0;
  ft.exact_lds = t->exact_lds != 0;
  ft.early_bin = t->early_bin != 0;
  f->tune = ft;
  f->surface->xcd = t->splat_xcd != 0;
  f->surface->owner = t->splat_owner != 0;
  f->surface->chunk = (uint32_t)t->chunk;
  f->surface->persistent_blocks = f->ctx->cus * t->groups_per_cu;
  if (f->graph) {
    (void)hipGraphExecDestroy(f->graph);
    f->graph = nullptr;
  }
  return CLAY_OK;
}

void read_tuning(const Fluid* f, clay_tuning* out) {
  out->block = f->tune.block;
  out->stage = f->tune.stage;
  out->chunk = (int32_t)f->surface->chunk;
  out->groups_per_cu = f->surface->persistent_blocks / std::max(1, f->ctx->cus);
  out->fluid_xcd = f->tune.xcd ? 1 : 0;
  out->splat_xcd = f->surface->xcd ? 1 : 0;
  out->splat_owner = f->surface->owner ? 1 : 0;
  out->exact_lds = f->tune.exact_lds ? 1 : 0;
  out->early_bin = f->tune.early_bin ? 1 : 0;
}

}  // namespace

extern "C" int clay_fluid_set_tuning(clay_fluid* fh, const clay_tuning* t) {
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!f || !t) return fail(CLAY_ERR_INVALID, "clay_fluid_set_tuning: NULL argument");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  drop_graph(f);
  return apply_tuning(f, t);
}

extern "C" int clay_fluid_get_tuning(const clay_fluid* fh, clay_tuning* out) {
  if (!fh || !out) return fail(CLAY_ERR_INVALID, "clay_fluid_get_tuning: NULL argument");
  read_tuning(reinterpret_cast<const Fluid*>(fh), out);
  return CLAY_OK;
}

// ============================================================================
// Batched scenes
// ============================================================================
namespace {

Fluid* as_batch(clay_batch* b) { return reinterpret_cast<Fluid*>(b); }
const Fluid* as_batch(const clay_batch* b) { return reinterpret_cast<const Fluid*>(b); }

void drop_graph(Fluid* f) {
  if (f->graph) (void)hipGraphExecDestroy(f->graph);
  f->graph = nullptr;
  f->graph_stream = nullptr;
}

// One frame of every scene.  The frame is captured once per stream and replayed:
// every per-substep value it needs (the cell-record stamp, the scan totals, the
// following origins) lives in device memory, and the fills are kernels, so a
// replay repeats exactly what an eager frame enqueues.
int batch_frame(Fluid* f, hipStream_t s) {
  fluid_build_k(f);
  const uint32_t subs = (uint32_t)f->consts.substeps;
  // A stamp wrap wipes the cell table, which only an eager frame does.
  if (!f->use_graph || f->epoch > 0xFFFFFFFFu - subs - 2u) return fluid_frame_impl(f, s);
  // The graph holds a frame that starts from presorted state; after a state
  // upload the first sort and densities run eagerly here.
  if (!f->presorted && f->live > 0 && subs > 0) {
    const int rc = fluid_sort_density(f, s, nullptr);
    if (rc) return rc;
    f->presorted = true;
  }
  if (!f->graph || f->graph_stream != s) {
    drop_graph(f);
    const uint32_t epoch0 = f->epoch;
    f->ctx->capturing = true;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      f->ctx->capturing = false;
      return fail(CLAY_ERR_HIP, std::string("FluidBatch: graph capture: ") + hipGetErrorString(e));
    }
    const int rc = fluid_frame_impl(f, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    f->ctx->capturing = false;
    f->epoch = epoch0;
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : fail(CLAY_ERR_HIP, std::string("FluidBatch: graph capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&f->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      f->graph = nullptr;
      return fail(CLAY_ERR_HIP, std::string("FluidBatch: graph instantiation: ") + hipGetErrorString(e));
    }
    f->graph_stream = s;
  }
  CLAY_HIP_CHECK(hipGraphLaunch(f->graph, s));
  f->epoch += subs;
  return CLAY_OK;
}

int check_scene(const Fluid* f, uint32_t scene, const char* who) {
  if (scene >= f->scenes) return fail(CLAY_ERR_INVALID, std::string(who) + ": scene out of range");
  return CLAY_OK;
}

}  // namespace

extern "C" int clay_batch_create(clay_ctx* ctx, clay_field* world, const clay_batch_desc* d, clay_batch** out) {
  if (!ctx || !d || !out || !d->counts) return fail(CLAY_ERR_INVALID, "clay_batch_create: NULL argument");
  if (d->scenes < 1 || d->scenes > 256) return fail(CLAY_ERR_INVALID, "FluidBatch: scenes must be in [1, 256]");
  if (!(d->fluid.follow_size > 0))
    return fail(CLAY_ERR_INVALID, "FluidBatch: every scene bakes a following volume (follow_size > 0)");
  if (d->fluid.splat_mode != CLAY_SPLAT_TILED)
    return fail(CLAY_ERR_INVALID, "FluidBatch: the batched bake uses the tiled schedule");
  Fluid* f = nullptr;
  int rc = fluid_create_impl(reinterpret_cast<Ctx*>(ctx), reinterpret_cast<Field*>(world), &d->fluid, d->scenes,
                             d->counts, &f);
  if (rc) return rc;
  *out = reinterpret_cast<clay_batch*>(f);
  return CLAY_OK;
}

extern "C" int clay_batch_destroy(clay_batch* b) {
  Fluid* f = as_batch(b);
  if (f) (void)fluid_quiesce(f);
  fluid_free(f);
  return CLAY_OK;
}

extern "C" int clay_batch_set_state(clay_batch* b, const float* pos, const float* vel) {
  Fluid* f = as_batch(b);
  if (!f || !pos) return fail(CLAY_ERR_INVALID, "clay_batch_set_state: NULL argument");
  const uint32_t n = f->capacity;
  CLAY_HIP_CHECK(fluid_quiesce(f));
  float* dp[3] = {f->px, f->py, f->pz};
  float* dv[3] = {f->vx, f->vy, f->vz};
  for (int a = 0; a < 3; a++) {
    CLAY_HIP_CHECK(hipMemcpy(dp[a], pos + (size_t)a * n, sizeof(float) * n, hipMemcpyHostToDevice));
    if (vel)
      CLAY_HIP_CHECK(hipMemcpy(dv[a], vel + (size_t)a * n, sizeof(float) * n, hipMemcpyHostToDevice));
    else
      CLAY_HIP_CHECK(hipMemset(dv[a], 0, sizeof(float) * n));
  }
  f->presorted = false;
  return CLAY_OK;
}

extern "C" int clay_batch_get_state(clay_batch* b, int32_t what, float* host) {
  Fluid* f = as_batch(b);
  if (f) CLAY_HIP_CHECK(fluid_quiesce(f));
  return clay_fluid_get_state(reinterpret_cast<clay_fluid*>(b), what, host);
}

extern "C" int clay_batch_step(clay_batch* b, clay_stream stream) {
  Fluid* f = as_batch(b);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_batch_step: NULL batch");
  f->last_stream = pick(f->ctx, stream);
  return batch_frame(f, f->last_stream);
}

extern "C" int clay_batch_set_graph(clay_batch* b, int32_t enabled) {
  Fluid* f = as_batch(b);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_batch_set_graph: NULL batch");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  drop_graph(f);
  f->use_graph = enabled != 0;
  return CLAY_OK;
}

extern "C" int clay_batch_configure(clay_batch* b, const clay_fluid_params* p) {
  Fluid* f = as_batch(b);
  if (!f || !p) return fail(CLAY_ERR_INVALID, "clay_batch_configure: NULL argument");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  drop_graph(f);
  f->params = *p;
  fluid_build_k(f);
  f->presorted = false;
  return CLAY_OK;
}

extern "C" int clay_batch_set_tuning(clay_batch* b, const clay_tuning* t) {
  Fluid* f = as_batch(b);
  if (!f || !t) return fail(CLAY_ERR_INVALID, "clay_batch_set_tuning: NULL argument");
  CLAY_HIP_CHECK(fluid_quiesce(f));
  drop_graph(f);
  return apply_tuning(f, t);
}

extern "C" int clay_batch_get_tuning(const clay_batch* b, clay_tuning* out) {
  if (!b || !out) return fail(CLAY_ERR_INVALID, "clay_batch_get_tuning: NULL argument");
  read_tuning(as_batch(b), out);
  return CLAY_OK;
}

extern "C" int clay_batch_constants(const clay_batch* b, clay_fluid_consts* out) {
  if (!b || !out) return fail(CLAY_ERR_INVALID, "clay_batch_constants: NULL argument");
  *out = as_batch(b)->consts;
  return CLAY_OK;
}

extern "C" int clay_batch_scene(const clay_batch* b, uint32_t scene, uint32_t* first, uint32_t* count) {
  const Fluid* f = as_batch(b);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_batch_scene: NULL batch");
  if (int rc = check_scene(f, scene, "clay_batch_scene")) return rc;
  if (first) *first = f->scene_base[scene];
  if (count) *count = f->scene_base[scene + 1] - f->scene_base[scene];
  return CLAY_OK;
}

extern "C" int clay_batch_read_contacts(clay_batch* b, uint32_t scene, float* out4, uint32_t* n,
                                        uint32_t* raw) {
  Fluid* f = as_batch(b);
  if (!f) return fail(CLAY_ERR_INVALID, "clay_batch_read_contacts: NULL batch");
  if (int rc = check_scene(f, scene, "clay_batch_read_contacts")) return rc;
  uint32_t buf[4 + 4 * CLAY_CONTACTS];
  CLAY_HIP_CHECK(fluid_quiesce(f));
  CLAY_HIP_CHECK(fluid_quiesce(f));
  CLAY_HIP_CHECK(hipMemcpy(buf, f->contacts + (size_t)scene * (4 + 4 * CLAY_CONTACTS), sizeof(buf),
                           hipMemcpyDeviceToHost));
  const uint32_t cnt = std::min<uint32_t>(buf[0], CLAY_CONTACTS);
  if (out4) memcpy(out4, buf + 4, sizeof(float) * 4 * cnt);
  if (n) *n = cnt;
  if (raw) *raw = buf[0];
  return CLAY_OK;
}

extern "C" int clay_batch_volume_download(clay_batch* b, uint32_t scene, int32_t mip, uint16_t* rgba) {
  Fluid* f = as_batch(b);
  if (!f || !rgba) return fail(CLAY_ERR_INVALID, "clay_batch_volume_download: NULL argument");
  if (int rc = check_scene(f, scene, "clay_batch_volume_download")) return rc;
  Volume* v = f->surface->vol;
  if (mip < 0 || mip >= v->info.mip_levels) return fail(CLAY_ERR_INVALID, "clay_batch_volume_download: bad mip");
  const size_t r = (size_t)v->mip_res(mip), n = r * r * r;
  const size_t off = (size_t)scene * v->total_voxels + v->voxel_off[mip];
  std::vector<uint16_t> d(n), m(n);
  CLAY_HIP_CHECK(fluid_quiesce(f));
  CLAY_HIP_CHECK(hipMemcpy(d.data(), v->dist + off, n * 2, hipMemcpyDeviceToHost));
  CLAY_HIP_CHECK(hipMemcpy(m.data(), v->mat + off, n * 2, hipMemcpyDeviceToHost));
  const uint16_t one = f2h_host(1.0f);
  for (size_t i = 0; i < n; i++) {
    rgba[4 * i] = d[i];
    rgba[4 * i + 1] = m[i];
    rgba[4 * i + 2] = 0;
    rgba[4 * i + 3] = one;
  }
  return CLAY_OK;
}

extern "C" int clay_batch_grid_download(clay_batch* b, uint32_t scene, int32_t mip, uint32_t* host) {
  Fluid* f = as_batch(b);
  if (!f || !host) return fail(CLAY_ERR_INVALID, "clay_batch_grid_download: NULL argument");
  if (int rc = check_scene(f, scene, "clay_batch_grid_download")) return rc;
  Splat* sp = f->surface;
  if (mip < 0 || mip >= sp->mips) return fail(CLAY_ERR_INVALID, "clay_batch_grid_download: bad mip");
  const size_t r = (size_t)sp->vol->mip_res(mip), n = r * r * r;
  CLAY_HIP_CHECK(fluid_quiesce(f));
  CLAY_HIP_CHECK(hipMemcpy(host, sp->grid + (size_t)scene * sp->grid_total + sp->grid_off[mip], n * 4,
                           hipMemcpyDeviceToHost));
  return CLAY_OK;
}

extern "C" int clay_batch_tiles_download(clay_batch* b, uint32_t scene, int32_t mip, uint8_t* host) {
  Fluid* f = as_batch(b);
  if (!f || !host) return fail(CLAY_ERR_INVALID, "clay_batch_tiles_download: NULL argument");
  if (int rc = check_scene(f, scene, "clay_batch_tiles_download")) return rc;
  Splat* sp = f->surface;
  if (mip < 0 || mip >= sp->mips) return fail(CLAY_ERR_INVALID, "clay_batch_tiles_download: bad mip");
  const int tr = sp->vol->mip_res(mip) / kTile;
  const size_t nt = (size_t)tr * tr * tr;
  CLAY_HIP_CHECK(fluid_quiesce(f));
  CLAY_HIP_CHECK(hipMemcpy(host, sp->tile_live + (size_t)scene * sp->tile_total + sp->tile_off[mip], nt,
                           hipMemcpyDeviceToHost));
  return CLAY_OK;
}

extern "C" int clay_batch_origin(clay_batch* b, uint32_t scene, float origin[3]) {
  Fluid* f = as_batch(b);
  if (!f || !origin) return fail(CLAY_ERR_INVALID, "clay_batch_origin: NULL argument");
  if (int rc = check_scene(f, scene, "clay_batch_origin")) return rc;
  VolDev rec;
  CLAY_HIP_CHECK(fluid_quiesce(f));
  CLAY_HIP_CHECK(hipMemcpy(&rec, f->surface->vol->dev + scene, sizeof(VolDev), hipMemcpyDeviceToHost));
  for (int a = 0; a < 3; a++) origin[a] = rec.origin[a];
  return CLAY_OK;
}

extern "C" int clay_batch_conserved(clay_batch* b, uint32_t scene, double out[7]) {
  Fluid* f = as_batch(b);
  if (!f || !out) return fail(CLAY_ERR_INVALID, "clay_batch_conserved: NULL argument");
  if (int rc = check_scene(f, scene, "clay_batch_conserved")) return rc;
  hipStream_t s = f->ctx->stream;
  CLAY_HIP_CHECK(fluid_quiesce(f));
  const size_t first = f->scene_base[scene];
  const uint32_t n = f->scene_base[scene + 1] - f->scene_base[scene];
  const float* pv[6] = {f->px + first, f->py + first, f->pz + first, f->vx + first, f->vy + first, f->vz + first};
  double* res = f->reduce + (size_t)f->reduce_blocks * 6;
  int rc = launch_conserved(f->ctx, s, n, pv, f->consts.mass, f->reduce, f->reduce_blocks, res);
  if (rc) return rc;
  CLAY_HIP_CHECK(hipMemcpyAsync(out, res, sizeof(double) * 7, hipMemcpyDeviceToHost, s));
  CLAY_HIP_CHECK(hipStreamSynchronize(s));
  return CLAY_OK;
}

// ============================================================================
// SdfEditor
// ============================================================================
namespace {

void editor_free(Editor* e) {
  if (!e) return;
  dfree(e->edits_dev);
  dfree(e->nedit_dev);
  dfree(e->counts);
  dfree(e->ids);
  dfree(e->list);
  dfree(e->args);
  delete e;
}

uint32_t editor_scenes(const Editor* e) { return (uint32_t)e->vol->instances; }

size_t editor_pending(const Editor* e) {
  size_t n = 0;
  for (const auto& q : e->pending) n += q.size();
  return n;
}

// Uploads every scene's host queue as its device edit list (counts included) and
// clears the queues.  With nothing queued it only resets the device counts (a
// kernel), so an erosion flush can be captured in a graph.
int editor_stage(Editor* e, hipStream_t s) {
  const uint32_t S = editor_scenes(e);
  if (editor_pending(e) == 0) return launch_fill_u32(e->ctx, s, e->nedit_dev, S, 0u);
  e->staged_counts.assign(S, 0u);
  for (uint32_t sc = 0; sc < S; sc++) {
    const auto& q = e->pending[sc];
    e->staged_counts[sc] = (uint32_t)q.size();
    if (!q.empty())
      CLAY_HIP_CHECK(hipMemcpyAsync(e->edits_dev + (size_t)sc * e->max_edits, q.data(), sizeof(clay_brush) * q.size(),
                                    hipMemcpyHostToDevice, s));
  }
  CLAY_HIP_CHECK(hipMemcpyAsync(e->nedit_dev, e->staged_counts.data(), sizeof(uint32_t) * S, hipMemcpyHostToDevice, s));
  // The queues' storage must outlive the copies: wait for them before reusing it.
  CLAY_HIP_CHECK(hipStreamSynchronize(s));
  for (auto& q : e->pending) q.clear();
  return CLAY_OK;
}

constexpr uint32_t kContactWords = 4 + 4 * CLAY_CONTACTS;

// Contacts of fluid scenes [first, first + n) -> edits of editor scenes [dst, dst + n).
int editor_erode_impl(Editor* e, Fluid* f, uint32_t first, uint32_t n, uint32_t dst, double min_speed, double radius,
                      double blend, hipStream_t s) {
  if (!(radius > 0)) return fail(CLAY_ERR_INVALID, "clay_editor_erode: radius must be positive");
  if (first + n > f->scenes || dst + n > editor_scenes(e))
    return fail(CLAY_ERR_INVALID, "clay_editor_erode: scene out of range");
  if (f->graph_stream && f->graph_stream != s) CLAY_HIP_CHECK(hipStreamSynchronize(f->graph_stream));
  int rc = editor_stage(e, s);
  if (rc) return rc;
  rc = launch_contacts_to_edits(e->ctx, s, e, f->contacts + (size_t)first * kContactWords, kContactWords, n, dst,
                                min_speed, radius, blend);
  if (rc) return rc;
  return editor_flush_device(e, s, false);
}

}  // namespace

extern "C" int clay_editor_create(clay_volume* vol, const clay_editor_desc* desc, clay_editor** out) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v || !out) return fail(CLAY_ERR_INVALID, "clay_editor_create: NULL argument");
  Editor* e = new (std::nothrow) Editor();
  if (!e) return fail(CLAY_ERR_INVALID, "out of host memory");
  e->ctx = v->ctx;
  e->vol = v;
  if (desc) {
    if (desc->max_edits > 0) e->max_edits = desc->max_edits;
    if (desc->tile_capacity > 0) e->capacity = desc->tile_capacity;
    if (desc->safety > 0) e->safety = desc->safety;
  }
  const size_t S = (size_t)v->instances;
  e->pending.assign(S, {});
  const size_t total = (size_t)v->info.total_cells;
  const size_t t0 = (size_t)v->info.tile_res[0] * v->info.tile_res[0] * v->info.tile_res[0];
  int rc;
  if ((rc = dalloc(&e->edits_dev, S * (size_t)e->max_edits)) || (rc = dalloc(&e->nedit_dev, S)) ||
      (rc = dalloc(&e->counts, S * total)) || (rc = dalloc(&e->ids, S * t0 * (size_t)e->capacity)) ||
      (rc = dalloc(&e->list, S * total)) || (rc = dalloc(&e->args, S * 4 * (size_t)v->info.mip_levels))) {
    editor_free(e);
    return rc;
  }
  hipStream_t s = e->ctx->stream;
  if ((rc = launch_fill_u32(e->ctx, s, e->counts, S * total, 0u)) ||
      (rc = launch_fill_u32(e->ctx, s, e->nedit_dev, S, 0u)) ||
      (rc = launch_fill_u32(e->ctx, s, e->args, S * 4 * (size_t)v->info.mip_levels, 0u))) {
    editor_free(e);
    return rc;
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    editor_free(e);
    return fail(CLAY_ERR_HIP, "SdfEditor: initialisation failed");
  }
  *out = reinterpret_cast<clay_editor*>(e);
  return CLAY_OK;
}

extern "C" int clay_editor_destroy(clay_editor* eh) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  if (e) (void)hipStreamSynchronize(e->ctx->stream);
  editor_free(e);
  return CLAY_OK;
}

extern "C" int clay_editor_push_scene(clay_editor* eh, uint32_t scene, const clay_brush* edits, int32_t n,
                                      int32_t* accepted) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  if (!e || (n > 0 && !edits)) return fail(CLAY_ERR_INVALID, "clay_editor_push: NULL argument");
  if (scene >= editor_scenes(e)) return fail(CLAY_ERR_INVALID, "clay_editor_push: scene out of range");
  auto& q = e->pending[scene];
  int32_t k = 0;
  for (int32_t i = 0; i < n; i++) {
    if (edits[i].kind > 6)
      return fail(CLAY_ERR_INVALID, "SdfEditor: brush kind " + std::to_string(edits[i].kind) +
                                        " is outside the supported set (sphere, box, roundBox, capsule, torus, cylinder, plane)");
    if ((int)q.size() >= e->max_edits) break;  // modify.ts:203-209 drops the rest
    q.push_back(edits[i]);
    k++;
  }
  if (accepted) *accepted = k;
  if (k < n)
    fprintf(stderr, "SdfEditor: dropping %d edit(s), %d already queued this frame\n", n - k, e->max_edits);
  return CLAY_OK;
}

extern "C" int clay_editor_push(clay_editor* eh, const clay_brush* edits, int32_t n, int32_t* accepted) {
  return clay_editor_push_scene(eh, 0, edits, n, accepted);
}

extern "C" int clay_editor_pending(const clay_editor* eh, int32_t* n) {
  if (!eh || !n) return fail(CLAY_ERR_INVALID, "clay_editor_pending: NULL argument");
  *n = (int32_t)editor_pending(reinterpret_cast<const Editor*>(eh));
  return CLAY_OK;
}

extern "C" int clay_editor_flush(clay_editor* eh, clay_stream stream, int32_t* flushed) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  if (!e) return fail(CLAY_ERR_INVALID, "clay_editor_flush: NULL editor");
  if (flushed) *flushed = 0;
  if (editor_pending(e) == 0) return CLAY_OK;  // modify.ts:215-217
  hipStream_t s = pick(e->ctx, stream);
  int rc = editor_stage(e, s);
  if (rc) return rc;
  rc = editor_flush_device(e, s, false);
  if (!rc && flushed) *flushed = 1;
  return rc;
}

extern "C" int clay_editor_erode(clay_editor* eh, clay_fluid* fh, double min_speed, double radius, double blend,
                                 clay_stream stream) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  Fluid* f = reinterpret_cast<Fluid*>(fh);
  if (!e || !f) return fail(CLAY_ERR_INVALID, "clay_editor_erode: NULL argument");
  return editor_erode_impl(e, f, 0, 1, 0, min_speed, radius, blend, pick(e->ctx, stream));
}

extern "C" int clay_editor_erode_scene(clay_editor* eh, clay_batch* bh, uint32_t scene, double min_speed,
                                       double radius, double blend, clay_stream stream) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  Fluid* f = reinterpret_cast<Fluid*>(bh);
  if (!e || !f) return fail(CLAY_ERR_INVALID, "clay_editor_erode_scene: NULL argument");
  const uint32_t dst = editor_scenes(e) == 1 ? 0u : scene;
  return editor_erode_impl(e, f, scene, 1, dst, min_speed, radius, blend, pick(e->ctx, stream));
}

extern "C" int clay_editor_erode_batch(clay_editor* eh, clay_batch* bh, double min_speed, double radius,
                                       double blend, clay_stream stream) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  Fluid* f = reinterpret_cast<Fluid*>(bh);
  if (!e || !f) return fail(CLAY_ERR_INVALID, "clay_editor_erode_batch: NULL argument");
  if (editor_scenes(e) != f->scenes)
    return fail(CLAY_ERR_INVALID, "clay_editor_erode_batch: the volume needs one instance per scene of the batch");
  return editor_erode_impl(e, f, 0, f->scenes, 0, min_speed, radius, blend, pick(e->ctx, stream));
}

extern "C" int clay_editor_refresh_all(clay_editor* eh, clay_stream stream) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  if (!e) return fail(CLAY_ERR_INVALID, "clay_editor_refresh_all: NULL editor");
  return editor_flush_device(e, pick(e->ctx, stream), true);
}

extern "C" int clay_editor_tile_counts_scene(clay_editor* eh, uint32_t scene, uint32_t* per_mip) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  if (!e || !per_mip) return fail(CLAY_ERR_INVALID, "clay_editor_tile_counts: NULL argument");
  if (scene >= editor_scenes(e)) return fail(CLAY_ERR_INVALID, "clay_editor_tile_counts: scene out of range");
  const int mips = e->vol->info.mip_levels;
  std::vector<uint32_t> args(4 * (size_t)mips);
  CLAY_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(args.data(), e->args + (size_t)scene * 4 * mips, sizeof(uint32_t) * args.size(),
                           hipMemcpyDeviceToHost));
  for (int m = 0; m < mips; m++) per_mip[m] = args[4 * (size_t)m];
  return CLAY_OK;
}

extern "C" int clay_editor_tile_counts(clay_editor* eh, uint32_t* per_mip) {
  return clay_editor_tile_counts_scene(eh, 0, per_mip);
}

extern "C" int clay_editor_cells_download_scene(clay_editor* eh, uint32_t scene, uint32_t* counts) {
  Editor* e = reinterpret_cast<Editor*>(eh);
  if (!e || !counts) return fail(CLAY_ERR_INVALID, "clay_editor_cells_download: NULL argument");
  if (scene >= editor_scenes(e)) return fail(CLAY_ERR_INVALID, "clay_editor_cells_download: scene out of range");
  const size_t total = (size_t)e->vol->info.total_cells;
  CLAY_HIP_CHECK(hipStreamSynchronize(e->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(counts, e->counts + (size_t)scene * total, sizeof(uint32_t) * total, hipMemcpyDeviceToHost));
  return CLAY_OK;
}

extern "C" int clay_editor_cells_download(clay_editor* eh, uint32_t* counts) {
  return clay_editor_cells_download_scene(eh, 0, counts);
}

// ============================================================================
// Body centroids
// ============================================================================
extern "C" int clay_body_centroids(clay_ctx* ctx, const float* const pos[3], const float* const rest[3],
                                   const uint32_t* first_count, uint32_t nbodies, int32_t mode, double* out6,
                                   clay_stream stream) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !pos || !pos[0] || !pos[1] || !pos[2] || (nbodies && (!first_count || !out6)))
    return fail(CLAY_ERR_INVALID, "clay_body_centroids: NULL argument");
  if (rest && (!rest[0] || !rest[1] || !rest[2]))
    return fail(CLAY_ERR_INVALID, "clay_body_centroids: rest must give all three axes or be NULL");
  if (mode != CLAY_REDUCE_F32_TREE && mode != CLAY_REDUCE_F64)
    return fail(CLAY_ERR_INVALID, "clay_body_centroids: unknown reduction mode");
  if (nbodies == 0) return CLAY_OK;
  hipStream_t s = pick(c, stream);
  uint32_t* fc = nullptr;
  double* out = nullptr;
  int rc;
  if ((rc = dalloc(&fc, 2 * (size_t)nbodies)) || (rc = dalloc(&out, 6 * (size_t)nbodies))) {
    dfree(fc);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(fc, first_count, sizeof(uint32_t) * 2 * nbodies, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    const float* p3[3] = {pos[0], pos[1], pos[2]};
    const float* r3[3] = {rest ? rest[0] : nullptr, rest ? rest[1] : nullptr, rest ? rest[2] : nullptr};
    rc = launch_body_centroids(c, s, mode == CLAY_REDUCE_F64, nbodies, fc, p3, rest ? r3 : nullptr, out);
    if (!rc) e = hipMemcpyAsync(out6, out, sizeof(double) * 6 * nbodies, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(fc);
  dfree(out);
  if (e != hipSuccess) return fail(CLAY_ERR_HIP, std::string("clay_body_centroids: ") + hipGetErrorString(e));
  return rc;
}

// ============================================================================
// Scenes: volume instances, splats over scene-partitioned clouds, async centroids
// ============================================================================
extern "C" int clay_volume_create_instances(clay_ctx* ctx, const clay_volume_desc* desc, uint32_t instances,
                                            clay_volume** out) {
  if (!ctx || !desc || !out) return fail(CLAY_ERR_INVALID, "clay_volume_create_instances: NULL argument");
  if (instances < 1 || instances > 256) return fail(CLAY_ERR_INVALID, "SdfVolume: instances must be in [1, 256]");
  Volume* v = nullptr;
  int rc = volume_create_impl(reinterpret_cast<Ctx*>(ctx), desc, &v, (int)instances);
  if (rc) return rc;
  *out = reinterpret_cast<clay_volume*>(v);
  return CLAY_OK;
}

extern "C" int clay_volume_replicate(clay_volume* vol, clay_stream stream) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v) return fail(CLAY_ERR_INVALID, "clay_volume_replicate: NULL volume");
  hipStream_t s = pick(v->ctx, stream);
  const size_t n = v->total_voxels;
  for (int i = 1; i < v->instances; i++) {
    CLAY_HIP_CHECK(hipMemcpyAsync(v->dist + (size_t)i * n, v->dist, n * 2, hipMemcpyDeviceToDevice, s));
    CLAY_HIP_CHECK(hipMemcpyAsync(v->mat + (size_t)i * n, v->mat, n * 2, hipMemcpyDeviceToDevice, s));
    CLAY_HIP_CHECK(hipMemcpyAsync(v->dev + i, v->dev, sizeof(VolDev), hipMemcpyDeviceToDevice, s));
  }
  CLAY_HIP_CHECK(hipStreamSynchronize(s));
  return CLAY_OK;
}

extern "C" int clay_volume_instance_download(clay_volume* vol, uint32_t instance, int32_t mip, uint16_t* rgba) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v || !rgba) return fail(CLAY_ERR_INVALID, "clay_volume_instance_download: NULL argument");
  if ((int)instance >= v->instances) return fail(CLAY_ERR_INVALID, "clay_volume_instance_download: bad instance");
  if (mip < 0 || mip >= v->info.mip_levels) return fail(CLAY_ERR_INVALID, "clay_volume_instance_download: bad mip");
  const size_t r = (size_t)v->mip_res(mip), n = r * r * r;
  const size_t off = (size_t)instance * v->total_voxels + v->voxel_off[mip];
  std::vector<uint16_t> d(n), m(n);
  CLAY_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(d.data(), v->dist + off, n * 2, hipMemcpyDeviceToHost));
  CLAY_HIP_CHECK(hipMemcpy(m.data(), v->mat + off, n * 2, hipMemcpyDeviceToHost));
  const uint16_t one = f2h_host(1.0f);
  for (size_t i = 0; i < n; i++) {
    rgba[4 * i] = d[i];
    rgba[4 * i + 1] = m[i];
    rgba[4 * i + 2] = 0;
    rgba[4 * i + 3] = one;
  }
  return CLAY_OK;
}

extern "C" int clay_splat_create_scenes(clay_ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc,
                                        uint32_t scenes, const uint32_t* counts, int32_t follow,
                                        const double* centers_dev, clay_splat** out) {
  if (!ctx || !cloud || !desc || !out || !counts) return fail(CLAY_ERR_INVALID, "clay_splat_create_scenes: NULL argument");
  if (scenes < 1 || scenes > 256) return fail(CLAY_ERR_INVALID, "SplatField: scenes must be in [1, 256]");
  if (follow != CLAY_FOLLOW_NONE && follow != CLAY_FOLLOW_BOUNDS && follow != CLAY_FOLLOW_CENTERS)
    return fail(CLAY_ERR_INVALID, "SplatField: unknown follow mode");
  if (follow == CLAY_FOLLOW_CENTERS && !centers_dev)
    return fail(CLAY_ERR_INVALID, "SplatField: following centres needs the device centre array");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  std::vector<uint32_t> base(scenes + 1, 0);
  for (uint32_t i = 0; i < scenes; i++) base[i + 1] = base[i] + counts[i];
  const uint32_t total = base[scenes];
  if (total > cloud->capacity) return fail(CLAY_ERR_INVALID, "SplatField: the scenes hold more particles than the cloud");
  std::vector<uint32_t> so(std::max<uint32_t>(total, 1), 0);
  for (uint32_t i = 0; i < scenes; i++)
    for (uint32_t j = base[i]; j < base[i + 1]; j++) so[j] = i;
  uint32_t *dso = nullptr, *dbase = nullptr;
  int rc;
  if ((rc = dalloc(&dso, so.size())) || (rc = dalloc(&dbase, base.size()))) {
    dfree(dso);
    return rc;
  }
  if (hipMemcpy(dso, so.data(), sizeof(uint32_t) * so.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dbase, base.data(), sizeof(uint32_t) * base.size(), hipMemcpyHostToDevice) != hipSuccess) {
    dfree(dso);
    dfree(dbase);
    return fail(CLAY_ERR_HIP, "SplatField: scene table upload failed");
  }
  clay_cloud_desc cl = *cloud;
  cl.live_rule = CLAY_LIVE_PREFIX;
  cl.live_count = total;
  SplatScenes sc;
  sc.scenes = (int)scenes;
  sc.scene_of = scenes > 1 ? dso : nullptr;
  sc.scene_base = dbase;
  sc.follow = follow != CLAY_FOLLOW_NONE;
  Splat* sp = nullptr;
  rc = splat_create_impl(c, &cl, desc, &sp, sc);
  if (rc) {
    dfree(dso);
    dfree(dbase);
    return rc;
  }
  sp->own_scene_of = dso;
  sp->own_scene_base = dbase;
  sp->follow_centers = follow == CLAY_FOLLOW_CENTERS ? centers_dev : nullptr;
  *out = reinterpret_cast<clay_splat*>(sp);
  return CLAY_OK;
}

extern "C" int clay_splat_instance_download(clay_splat* s, uint32_t scene, int32_t mip, uint32_t* codes_host) {
  Splat* sp = reinterpret_cast<Splat*>(s);
  if (!sp || !codes_host) return fail(CLAY_ERR_INVALID, "clay_splat_instance_download: NULL argument");
  if ((int)scene >= sp->scenes) return fail(CLAY_ERR_INVALID, "clay_splat_instance_download: bad scene");
  if (mip < 0 || mip >= sp->mips) return fail(CLAY_ERR_INVALID, "clay_splat_instance_download: bad mip");
  const size_t r = (size_t)sp->vol->mip_res(mip), n = r * r * r;
  CLAY_HIP_CHECK(hipStreamSynchronize(sp->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(codes_host, sp->grid + (size_t)scene * sp->grid_total + sp->grid_off[mip], n * 4,
                           hipMemcpyDeviceToHost));
  return CLAY_OK;
}

extern "C" int clay_body_centroids_device(clay_ctx* ctx, const float* const pos[3], const float* const rest[3],
                                          const uint32_t* first_count_dev, uint32_t nbodies, int32_t mode,
                                          double* out6_dev, clay_stream stream) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !pos || !pos[0] || !pos[1] || !pos[2] || (nbodies && (!first_count_dev || !out6_dev)))
    return fail(CLAY_ERR_INVALID, "clay_body_centroids_device: NULL argument");
  if (rest && (!rest[0] || !rest[1] || !rest[2]))
    return fail(CLAY_ERR_INVALID, "clay_body_centroids_device: rest must give all three axes or be NULL");
  if (mode != CLAY_REDUCE_F32_TREE && mode != CLAY_REDUCE_F64)
    return fail(CLAY_ERR_INVALID, "clay_body_centroids_device: unknown reduction mode");
  const float* p3[3] = {pos[0], pos[1], pos[2]};
  const float* r3[3] = {rest ? rest[0] : nullptr, rest ? rest[1] : nullptr, rest ? rest[2] : nullptr};
  return launch_body_centroids(c, pick(c, stream), mode == CLAY_REDUCE_F64, nbodies, first_count_dev, p3,
                               rest ? r3 : nullptr, out6_dev);
}
