// clay_napi.cc -- N-API binding of the C ABI (include/clay_hip.h) for the
// reference's TypeScript host (lib/src/sim/fluid.ts:111-582, lib/src/field/volume.ts:91-343,
// lib/src/trace/field.ts:196-240).  napi/index.js wraps these functions in classes
// with the reference's names; every failure throws a JS Error carrying the
// library's message (the reference's own constructor messages where it has them).
//
// Handles cross into JS as externals.  Particle state crosses as Float32Arrays in
// structure-of-arrays order (x[], y[], z[]), which is how the library stores it.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "clay_hip.h"

namespace {

napi_value throw_msg(napi_env env, const std::string& msg) {
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

napi_value throw_last(napi_env env) { return throw_msg(env, clay_last_error()); }

bool ok(napi_env env, napi_status st, const char* what) {
  if (st == napi_ok) return true;
  bool pending = false;
  napi_is_exception_pending(env, &pending);
  if (!pending) napi_throw_type_error(env, nullptr, what);
  return false;
}

struct Args {
  size_t argc = 8;
  napi_value argv[8];
  bool get(napi_env env, napi_callback_info info, size_t need) {
    if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
    if (argc < need) {
      napi_throw_type_error(env, nullptr, "too few arguments");
      return false;
    }
    return true;
  }
};

napi_value num(napi_env env, double x) {
  napi_value v;
  napi_create_double(env, x, &v);
  return v;
}

bool get_num(napi_env env, napi_value v, double* out) { return ok(env, napi_get_value_double(env, v, out), "expected a number"); }

template <typename T>
bool get_external(napi_env env, napi_value v, T** out) {
  void* p = nullptr;
  if (!ok(env, napi_get_value_external(env, v, &p), "expected a native handle")) return false;
  *out = static_cast<T*>(p);
  return *out != nullptr || (napi_throw_error(env, nullptr, "handle already destroyed"), false);
}

napi_value external(napi_env env, void* p) {
  napi_value v;
  napi_create_external(env, p, nullptr, nullptr, &v);
  return v;
}

// Optional named number / vec3 properties of an options object.
bool opt_num(napi_env env, napi_value obj, const char* key, double* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return false;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return false;
  return napi_get_value_double(env, v, out) == napi_ok;
}

bool opt_vec3(napi_env env, napi_value obj, const char* key, double out[3]) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return false;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  bool arr = false;
  napi_is_array(env, v, &arr);
  if (!arr) return false;
  for (uint32_t i = 0; i < 3; i++) {
    napi_value e;
    if (napi_get_element(env, v, i, &e) != napi_ok || napi_get_value_double(env, e, &out[i]) != napi_ok) return false;
  }
  return true;
}

bool opt_str(napi_env env, napi_value obj, const char* key, std::string* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return false;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  char buf[64];
  size_t n = 0;
  if (napi_get_value_string_utf8(env, v, buf, sizeof(buf), &n) != napi_ok) return false;
  *out = std::string(buf, n);
  return true;
}

bool float_array(napi_env env, napi_value v, float** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return napi_throw_type_error(env, nullptr, "expected a Float32Array"), false;
  napi_typedarray_type t;
  void* p = nullptr;
  napi_value ab;
  size_t off = 0;
  napi_get_typedarray_info(env, v, &t, len, &p, &ab, &off);
  if (t != napi_float32_array) return napi_throw_type_error(env, nullptr, "expected a Float32Array"), false;
  *data = static_cast<float*>(p);
  return true;
}

napi_value new_float_array(napi_env env, size_t n, float** data) {
  napi_value ab, out;
  void* p = nullptr;
  napi_create_arraybuffer(env, n * sizeof(float), &p, &ab);
  napi_create_typedarray(env, napi_float32_array, n, ab, 0, &out);
  *data = static_cast<float*>(p);
  return out;
}

// ---- library --------------------------------------------------------------------
napi_value version(napi_env env, napi_callback_info) { return num(env, clay_version()); }

napi_value last_error(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_string_utf8(env, clay_last_error(), NAPI_AUTO_LENGTH, &v);
  return v;
}

// volumeDerive(resolution, mipLevels) -> {resolution, mipLevels, tileResPerMip,
// cellOffsets, totalCells}: SdfVolume's geometry (volume.ts:139-179), host only.
napi_value volume_derive(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.get(env, info, 2)) return nullptr;
  double res = 0, mips = 0;
  if (!get_num(env, a.argv[0], &res) || !get_num(env, a.argv[1], &mips)) return nullptr;
  clay_volume_info vi;
  if (clay_volume_derive((int32_t)res, (int32_t)mips, &vi) != CLAY_OK) return throw_last(env);
  napi_value out, tiles, offs;
  napi_create_object(env, &out);
  napi_create_array_with_length(env, vi.mip_levels, &tiles);
  napi_create_array_with_length(env, vi.mip_levels, &offs);
  for (int m = 0; m < vi.mip_levels; m++) {
    napi_set_element(env, tiles, m, num(env, vi.tile_res[m]));
    napi_set_element(env, offs, m, num(env, vi.cell_offsets[m]));
  }
  napi_set_named_property(env, out, "resolution", num(env, vi.resolution));
  napi_set_named_property(env, out, "mipLevels", num(env, vi.mip_levels));
  napi_set_named_property(env, out, "tileResPerMip", tiles);
  napi_set_named_property(env, out, "cellOffsets", offs);
  napi_set_named_property(env, out, "totalCells", num(env, vi.total_cells));
  return out;
}

// fluidDefaults() -> the FluidSimOptions defaults the library applies (fluid.ts:141-203).
napi_value fluid_defaults(napi_env env, napi_callback_info) {
  clay_fluid_desc d;
  clay_fluid_desc_default(&d);
  napi_value out;
  napi_create_object(env, &out);
  auto set = [&](const char* k, double x) { napi_set_named_property(env, out, k, num(env, x)); };
  set("capacity", d.capacity);
  set("bucket", d.bucket);
  set("spacing", d.spacing);
  set("smoothing", d.smoothing);
  set("radius", d.radius);
  set("surfaceRadius", d.surface_radius);
  set("material", d.material);
  set("bakeResolution", d.bake_resolution);
  set("bakeMipLevels", d.bake_mip_levels);
  set("bakeBand", d.bake_band);
  set("worldSize", d.world_size);
  set("substeps", d.substeps);
  set("stiffness", d.stiffness);
  set("viscosity", d.viscosity);
  set("damping", d.damping);
  set("maxSpeed", d.max_speed);
  napi_value o, g;
  napi_create_array_with_length(env, 3, &o);
  napi_create_array_with_length(env, 3, &g);
  for (int i = 0; i < 3; i++) {
    napi_set_element(env, o, i, num(env, d.origin[i]));
    napi_set_element(env, g, i, num(env, d.gravity[i]));
  }
  napi_set_named_property(env, out, "origin", o);
  napi_set_named_property(env, out, "gravity", g);
  return out;
}

// ---- context, fields -------------------------------------------------------------------
napi_value ctx_create(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.get(env, info, 1)) return nullptr;
  double dev = 0;
  if (!get_num(env, a.argv[0], &dev)) return nullptr;
  clay_ctx* c = nullptr;
  if (clay_ctx_create((int)dev, &c) != CLAY_OK) return throw_last(env);
  return external(env, c);
}

napi_value ctx_destroy(napi_env env, napi_callback_info info) {
  Args a;
  clay_ctx* c = nullptr;
  if (!a.get(env, info, 1) || !get_external(env, a.argv[0], &c)) return nullptr;
  clay_ctx_destroy(c);
  return nullptr;
}

// fieldAnalytic(ctx, records: ArrayBuffer of 80-byte Brush records, {band, epsilon,
// material}) -> external.  index.js builds the records (BrushSet.make, brush.ts:539-564).
napi_value field_analytic(napi_env env, napi_callback_info info) {
  Args a;
  clay_ctx* c = nullptr;
  if (!a.get(env, info, 3) || !get_external(env, a.argv[0], &c)) return nullptr;
  clay_analytic_opts o{64.0, 1e-3, 0.0};
  opt_num(env, a.argv[2], "band", &o.band);
  opt_num(env, a.argv[2], "epsilon", &o.epsilon);
  opt_num(env, a.argv[2], "material", &o.material);
  void* p = nullptr;
  size_t bytes = 0;
  if (!ok(env, napi_get_arraybuffer_info(env, a.argv[1], &p, &bytes), "expected an ArrayBuffer of brush records"))
    return nullptr;
  if (bytes % sizeof(clay_brush) != 0) return throw_msg(env, "brush records must be 80 bytes each");
  clay_field* f = nullptr;
  if (clay_field_analytic(c, static_cast<const clay_brush*>(p), (int32_t)(bytes / sizeof(clay_brush)), &o, &f) !=
      CLAY_OK)
    return throw_last(env);
  return external(env, f);
}

napi_value field_destroy(napi_env env, napi_callback_info info) {
  Args a;
  clay_field* f = nullptr;
  if (!a.get(env, info, 1) || !get_external(env, a.argv[0], &f)) return nullptr;
  clay_field_destroy(f);
  return nullptr;
}

// ---- FluidSim -------------------------------------------------------------------------------
// fluidCreate(ctx, field | null, options) -> external (new FluidSim, fluid.ts:140).
napi_value fluid_create(napi_env env, napi_callback_info info) {
  Args a;
  clay_ctx* c = nullptr;
  if (!a.get(env, info, 3) || !get_external(env, a.argv[0], &c)) return nullptr;
  clay_field* world = nullptr;
  napi_valuetype t;
  napi_typeof(env, a.argv[1], &t);
  if (t != napi_null && t != napi_undefined && !get_external(env, a.argv[1], &world)) return nullptr;
  clay_fluid_desc d;
  clay_fluid_desc_default(&d);
  const napi_value o = a.argv[2];
  double x;
  if (opt_num(env, o, "capacity", &x)) d.capacity = (uint32_t)x;
  if (opt_num(env, o, "bucket", &x)) d.bucket = (int32_t)x;
  if (opt_num(env, o, "spacing", &x)) d.spacing = x;
  if (opt_num(env, o, "smoothing", &x)) d.smoothing = x;
  if (opt_num(env, o, "density", &x)) d.density = x;
  if (opt_num(env, o, "radius", &x)) d.radius = x;
  if (opt_num(env, o, "surfaceRadius", &x)) d.surface_radius = x;
  if (opt_num(env, o, "material", &x)) d.material = x;
  if (opt_num(env, o, "bakeResolution", &x)) d.bake_resolution = (int32_t)x;
  if (opt_num(env, o, "bakeMipLevels", &x)) d.bake_mip_levels = (int32_t)x;
  if (opt_num(env, o, "bakeBand", &x)) d.bake_band = x;
  if (opt_num(env, o, "worldSize", &x)) d.world_size = x;
  if (opt_num(env, o, "substeps", &x)) d.substeps = (int32_t)x;
  if (opt_num(env, o, "stiffness", &x)) d.stiffness = x;
  if (opt_num(env, o, "viscosity", &x)) d.viscosity = x;
  if (opt_num(env, o, "damping", &x)) d.damping = x;
  if (opt_num(env, o, "maxSpeed", &x)) d.max_speed = x;
  if (opt_num(env, o, "followSize", &x)) d.follow_size = x;
  opt_vec3(env, o, "origin", d.origin);
  opt_vec3(env, o, "gravity", d.gravity);
  std::string s;
  if (opt_str(env, o, "splatMode", &s)) d.splat_mode = s == "tiled" ? CLAY_SPLAT_TILED : CLAY_SPLAT_PLAIN;
  if (opt_str(env, o, "precision", &s)) d.precision = s == "fp16lds" ? CLAY_PRECISION_FP16LDS : CLAY_PRECISION_F32;
  clay_fluid* f = nullptr;
  if (clay_fluid_create(c, world, &d, &f) != CLAY_OK) return throw_last(env);
  return external(env, f);
}

napi_value fluid_destroy(napi_env env, napi_callback_info info) {
  Args a;
  clay_fluid* f = nullptr;
  if (!a.get(env, info, 1) || !get_external(env, a.argv[0], &f)) return nullptr;
  clay_fluid_destroy(f);
  return nullptr;
}

// fluidSetState(fluid, pos SoA Float32Array(3n), vel Float32Array(3n) | null)
napi_value fluid_set_state(napi_env env, napi_callback_info info) {
  Args a;
  clay_fluid* f = nullptr;
  if (!a.get(env, info, 2) || !get_external(env, a.argv[0], &f)) return nullptr;
  float* pos = nullptr;
  size_t n3 = 0;
  if (!float_array(env, a.argv[1], &pos, &n3)) return nullptr;
  if (n3 % 3) return throw_msg(env, "fluidSetState: positions are 3 planes of n values");
  float* vel = nullptr;
  if (a.argc > 2) {
    napi_valuetype t;
    napi_typeof(env, a.argv[2], &t);
    size_t v3 = 0;
    if (t != napi_null && t != napi_undefined) {
      if (!float_array(env, a.argv[2], &vel, &v3)) return nullptr;
      if (v3 != n3) return throw_msg(env, "fluidSetState: velocities must match the positions");
    }
  }
  if (clay_fluid_set_state(f, pos, vel, (uint32_t)(n3 / 3)) != CLAY_OK) return throw_last(env);
  return nullptr;
}

// fluidGetState(fluid, what: 0 pos, 1 vel, 2 density, 3 pressure, 4 nvel) -> Float32Array SoA
napi_value fluid_get_state(napi_env env, napi_callback_info info) {
  Args a;
  clay_fluid* f = nullptr;
  if (!a.get(env, info, 2) || !get_external(env, a.argv[0], &f)) return nullptr;
  double what = 0;
  if (!get_num(env, a.argv[1], &what)) return nullptr;
  uint32_t n = 0;
  if (clay_fluid_particle_count(f, &n) != CLAY_OK) return throw_last(env);
  const int w = (int)what;
  const size_t planes = (w == 2 || w == 3) ? 1 : 3;
  float* data = nullptr;
  napi_value out = new_float_array(env, planes * n, &data);
  if (clay_fluid_get_state(f, w, data) != CLAY_OK) return throw_last(env);
  return out;
}

napi_value fluid_step(napi_env env, napi_callback_info info) {
  Args a;
  clay_fluid* f = nullptr;
  if (!a.get(env, info, 1) || !get_external(env, a.argv[0], &f)) return nullptr;
  if (clay_fluid_step(f, nullptr) != CLAY_OK) return throw_last(env);
  return nullptr;
}

napi_value fluid_spawn(napi_env env, napi_callback_info info) {
  Args a;
  clay_fluid* f = nullptr;
  if (!a.get(env, info, 5) || !get_external(env, a.argv[0], &f)) return nullptr;
  double n = 0, spread = -1;
  double pos[3], vel[3];
  napi_value holder;
  napi_create_object(env, &holder);
  napi_set_named_property(env, holder, "p", a.argv[2]);
  napi_set_named_property(env, holder, "v", a.argv[3]);
  if (!get_num(env, a.argv[1], &n) || !opt_vec3(env, holder, "p", pos) || !opt_vec3(env, holder, "v", vel))
    return throw_msg(env, "fluidSpawn(fluid, n, [x, y, z], [vx, vy, vz], spread)");
  napi_valuetype t;
  napi_typeof(env, a.argv[4], &t);
  if (t == napi_number) napi_get_value_double(env, a.argv[4], &spread);
  if (clay_fluid_spawn(f, (uint32_t)n, pos, vel, spread) != CLAY_OK) return throw_last(env);
  return nullptr;
}

napi_value fluid_particle_count(napi_env env, napi_callback_info info) {
  Args a;
  clay_fluid* f = nullptr;
  if (!a.get(env, info, 1) || !get_external(env, a.argv[0], &f)) return nullptr;
  uint32_t n = 0;
  if (clay_fluid_particle_count(f, &n) != CLAY_OK) return throw_last(env);
  return num(env, n);
}

// fluidReadContacts(fluid) -> {count, points: [[x, y, z, speed], ...]}
napi_value fluid_read_contacts(napi_env env, napi_callback_info info) {
  Args a;
  clay_fluid* f = nullptr;
  if (!a.get(env, info, 1) || !get_external(env, a.argv[0], &f)) return nullptr;
  float pts[CLAY_CONTACTS * 4];
  uint32_t n = 0, raw = 0;
  if (clay_fluid_read_contacts(f, pts, &n, &raw) != CLAY_OK) return throw_last(env);
  napi_value out, arr;
  napi_create_object(env, &out);
  napi_create_array_with_length(env, n, &arr);
  for (uint32_t i = 0; i < n; i++) {
    napi_value p;
    napi_create_array_with_length(env, 4, &p);
    for (uint32_t c = 0; c < 4; c++) napi_set_element(env, p, c, num(env, pts[4 * i + c]));
    napi_set_element(env, arr, i, p);
  }
  napi_set_named_property(env, out, "count", num(env, raw));
  napi_set_named_property(env, out, "points", arr);
  return out;
}

// fluidSurfaceCodes(fluid, mip) -> Uint32Array of the surface bake's codes.
napi_value fluid_surface_codes(napi_env env, napi_callback_info info) {
  Args a;
  clay_fluid* f = nullptr;
  if (!a.get(env, info, 2) || !get_external(env, a.argv[0], &f)) return nullptr;
  double mip = 0;
  if (!get_num(env, a.argv[1], &mip)) return nullptr;
  clay_splat* sp = nullptr;
  clay_volume* vol = nullptr;
  clay_volume_info vi;
  if (clay_fluid_surface(f, &sp) != CLAY_OK || clay_splat_volume(sp, &vol) != CLAY_OK ||
      clay_volume_get_info(vol, &vi) != CLAY_OK)
    return throw_last(env);
  if (mip < 0 || mip >= vi.mip_levels) return throw_msg(env, "fluidSurfaceCodes: bad mip");
  const size_t r = (size_t)(vi.resolution >> (int)mip), n = r * r * r;
  napi_value ab, out;
  void* p = nullptr;
  napi_create_arraybuffer(env, n * 4, &p, &ab);
  napi_create_typedarray(env, napi_uint32_array, n, ab, 0, &out);
  if (clay_splat_grid_download(sp, (int32_t)mip, static_cast<uint32_t*>(p)) != CLAY_OK) return throw_last(env);
  return out;
}

}  // namespace

NAPI_MODULE_INIT() {
  const napi_property_descriptor d[] = {
      {"version", nullptr, version, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"lastError", nullptr, last_error, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"volumeDerive", nullptr, volume_derive, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidDefaults", nullptr, fluid_defaults, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"ctxCreate", nullptr, ctx_create, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"ctxDestroy", nullptr, ctx_destroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldAnalytic", nullptr, field_analytic, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldDestroy", nullptr, field_destroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidCreate", nullptr, fluid_create, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidDestroy", nullptr, fluid_destroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidSetState", nullptr, fluid_set_state, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidGetState", nullptr, fluid_get_state, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidStep", nullptr, fluid_step, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidSpawn", nullptr, fluid_spawn, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidParticleCount", nullptr, fluid_particle_count, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidReadContacts", nullptr, fluid_read_contacts, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fluidSurfaceCodes", nullptr, fluid_surface_codes, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof(d) / sizeof(d[0]), d);
  return exports;
}
