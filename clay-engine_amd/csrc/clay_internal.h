// clay_internal.h -- host-side objects behind the C ABI handles, and the kernel
// launchers each .hip translation unit exports to api.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "clay_device.h"
#include "clay_hip.h"

namespace clay {

// ---- errors ----------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define CLAY_HIP_CHECK(expr)                                                        \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return ::clay::fail(CLAY_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---- per-kernel timing -----------------------------------------------------
struct KernelStat {
  std::string name;
  double total_ms = 0.0;
  uint64_t launches = 0;
};

struct PendingTiming {
  int stat;
  hipEvent_t start, stop;
};

struct Ctx;

// Brackets one launch with HIP events when the context's timing is on.
class LaunchTimer {
 public:
  LaunchTimer(Ctx* ctx, hipStream_t s, const char* name);
  ~LaunchTimer();

 private:
  Ctx* ctx_;
  hipStream_t s_;
  int stat_ = -1;
  hipEvent_t a_ = nullptr, b_ = nullptr;
};

struct Ctx {
  int device = 0;
  int cus = 256;                 // compute units of the device
  int persistent_blocks = 2048;  // grid of persistent work-queue kernels
  hipStream_t stream = nullptr;
  bool timing = false;
  bool capturing = false;        // a graph capture is open: launch timers stay off
  std::vector<KernelStat> stats;
  std::vector<PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
  int stat_index(const char* name);
  hipEvent_t take_event();
  void collect();  // resolves pending timings (synchronises their events)
};

inline hipStream_t pick(Ctx* c, clay_stream s) {
  return s ? reinterpret_cast<hipStream_t>(s) : c->stream;
}

// Device allocation of `count` elements (at least one), failing with the message.
template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count);
  if (e != hipSuccess)
    return fail(CLAY_ERR_HIP, std::string("hipMalloc(") + std::to_string(sizeof(T) * count) +
                                   " bytes): " + hipGetErrorString(e));
  return CLAY_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// ---- objects ---------------------------------------------------------------
struct Volume {
  Ctx* ctx = nullptr;
  int instances = 1;             // a batch keeps one instance per scene, planes back to back
  clay_volume_info info{};
  double origin[3] = {0, 0, 0};
  double grad_voxels = 1.0;
  VolDev host{};                 // mirror of the device record (of every instance at creation)
  VolDev* dev = nullptr;         // device records, one per instance (kernels read these)
  uint16_t* dist = nullptr;      // all mips of every instance, distance plane (f16 bits)
  uint16_t* mat = nullptr;       // all mips of every instance, material plane (f16 bits)
  std::vector<size_t> voxel_off; // per mip offset into the planes
  size_t total_voxels = 0;
  uint32_t* wire_scratch = nullptr;  // pack: per tile row counts and offsets (allocated on first use)
  size_t wire_scratch_n = 0;
  uint32_t* wire_flag = nullptr;     // unpack: set when a stream names a payload tile it lacks
  int mip_res(int m) const { return info.resolution >> m; }
};

struct Field {
  Ctx* ctx = nullptr;
  int max_mip = 0;
  ColliderDev col{};             // leaves point at device memory owned below / by volumes
  clay_brush* brushes_dev = nullptr;
  std::vector<Field*> children;  // union operands (not owned)
  // Per leaf, for the reads at every mip (clay_field_query / entry): the volume of a
  // volume leaf (else null), and an analytic leaf's band and epsilon options.
  std::vector<Volume*> vols;
  std::vector<Volume*> vols_b;  // a lerp leaf's second volume, else null
  std::vector<double> bands, eps;
};

struct Splat {
  Ctx* ctx = nullptr;
  clay_cloud_desc cloud{};
  clay_splat_desc desc{};
  Volume* vol = nullptr;         // owned
  Field* field = nullptr;        // owned volumeField of `vol`
  int mips = 0;
  std::vector<size_t> grid_off;  // per mip offset into `grid`
  size_t grid_total = 0;
  uint32_t* grid = nullptr;      // u32 codes (plain mode, or tiled + keep_codes)
  // tiled mode
  std::vector<int> brick_res;    // bricks per axis per mip
  std::vector<size_t> brick_off;
  size_t brick_total = 0;
  uint32_t* brick_count = nullptr;   // per brick
  uint32_t* brick_start = nullptr;   // exclusive scan of counts
  uint32_t* brick_cursor = nullptr;  // per brick slice (binning fill cursors)
  uint32_t* brick_sub = nullptr;     // per brick slice (binning counts)
  uint32_t* brick_list = nullptr;    // particle ids per brick
  size_t brick_list_cap = 0;
  uint32_t* brick_total_dev = nullptr;  // scan results: total entries, total work items
  uint32_t* item_start = nullptr;       // first work item (chunk) of each brick
  uint8_t* tile_live = nullptr;      // per 8^3 tile, all mips: live after last bake
  std::vector<size_t> tile_off;
  size_t tile_total = 0;             // per scene
  bool has_baked = false;
  // scenes (a batch): brick, grid, tile and texel blocks per scene, scene-major
  int scenes = 1;
  const uint32_t* scene_of = nullptr;    // device, per particle id (not owned)
  const uint32_t* scene_base = nullptr;  // device, scenes + 1 prefix offsets (not owned)
  uint32_t* own_scene_of = nullptr;      // the two tables above when the splat owns them
  uint32_t* own_scene_base = nullptr;
  bool follow = false;                   // centre each volume before a bake:
  const double* follow_centers = nullptr;  // on these (6 doubles per scene, xyz first), else on the AABB
  uint32_t* follow_slots = nullptr;      // per scene: the AABB's six ordered-float slots
  bool bounds_pending = false;           // slots handed to a keys pass, not yet read and reset
  unsigned long long* atomic_counts = nullptr;  // set for a bake that tallies its atomics
  // tuning
  uint32_t chunk = 1024;                 // particles per splat work item
  int persistent_blocks = 2048;          // grid of the splat work-item kernel
  bool xcd = false;                      // XCD-contiguous work-item numbering
  bool owner = true;                     // owner schedule (brick + halo staged in LDS) where it fits
};

// Tunable launch shapes of the neighbour kernels (results do not depend on them).
struct FluidTune {
  int block = 256;         // threads per workgroup: 128 or 256
  int stage = 128;         // LDS staging slice per wave: 64, 128 or 256 neighbours
  bool xcd = false;        // XCD-contiguous workgroup numbering in the neighbour kernels
  bool exact_lds = true;   // f32 precision: LDS-staged exact kernels (else global-memory walks)
  bool early_bin = false;  // the bake's binning forks before the frame-end reorder
};
bool fluid_tune_valid(const FluidTune& t);

struct Fluid {
  Ctx* ctx = nullptr;
  Field* world = nullptr;        // not owned
  clay_fluid_desc desc{};
  clay_fluid_consts consts{};
  clay_fluid_params params{};
  double rest_density = 0, mass_poly6 = 0, h6 = 0, inv_h = 0, rest_floor = 0, hh = 0;
  uint32_t capacity = 0;
  uint32_t live = 0;
  uint32_t cursor = 0;
  uint32_t pending_emit = 0;
  uint32_t seed = 1;
  uint32_t emit_seed = 1;
  // id-order state (SoA)
  float* st = nullptr;           // 11 planes of capacity floats
  float *px, *py, *pz, *vx, *vy, *vz, *rho, *pres, *nvx, *nvy, *nvz;
  // sorted snapshot
  float* ss = nullptr;           // 8 planes
  float *spx, *spy, *spz, *svx, *svy, *svz, *srho, *spres;
  // sort
  uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr;
  uint32_t* hist = nullptr;      // digit histograms
  size_t hist_cap = 0;
  uint32_t* cells = nullptr;     // dense over gridRes^3 per scene: {start, end, epoch, 0} per cell
  uint32_t epoch = 0;            // host count of the device stamp (wrap bookkeeping)
  uint32_t* epoch_dev = nullptr; // stamp of the current substep's cell records
  size_t grid_cells = 0;
  int key_bits = 0;
  uint32_t scene_shift = 0;      // batches: cell bits of a key (scene << scene_shift | cell)
  uint32_t* contacts = nullptr;  // per scene: [0] = count, then 32 x float4 (as bits)
  uint32_t* cmask = nullptr;     // one bit per particle id: in contact this substep
  float4* cdata = nullptr;       // per particle id: contact point and speed
  double* reduce = nullptr;      // FP64 reduction partials (+ 8 results at the end)
  uint32_t reduce_blocks = 0;
  Splat* surface = nullptr;      // owned
  float* cover = nullptr;        // per particle id: squared interior radius for the surface bake
  uint2* spos = nullptr;         // FP16-LDS: per sorted particle, x and the f16 offsets of y, z in its cell
  float4* frec = nullptr;        // FP16-LDS: per sorted particle, 2 float4: the forces pass's staged record
  // The sort, cell table and densities of the next substep are computed at the end
  // of a frame, from the positions the bake sees; `presorted` says they are still
  // valid (no state upload, emission or configure since).
  bool presorted = false;
  FluidK k{};
  FluidTune tune{};
  // batch of independent scenes (one scene for a FluidSim)
  uint32_t scenes = 1;
  std::vector<uint32_t> scene_base;  // host, scenes + 1 prefix offsets
  uint32_t* scene_of = nullptr;      // device, per particle (batches only)
  uint32_t* scene_base_dev = nullptr;
  // one frame captured as a graph and replayed (batches)
  bool use_graph = false;
  hipGraphExec_t graph = nullptr;
  hipStream_t graph_stream = nullptr;  // the stream the graph was captured on
  uint32_t graph_live = 0;             // a FluidSim's live count when its graph was captured
  hipStream_t last_stream = nullptr;   // the stream of the latest step or substep
  // The bake's binning runs beside the frame-end density pass: forked after the
  // sort onto `side`, joined before the walk (default priority).
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  SceneK scene_k() const {
    SceneK sk{};
    sk.xcd = tune.xcd ? 1u : 0u;
    sk.epoch = epoch_dev;
    sk.batched = scene_of != nullptr;
    sk.shift = scene_shift;
    sk.cell_mask = scene_shift ? (1u << scene_shift) - 1u : 0xFFFFFFFFu;
    sk.cells_per_scene = (uint32_t)grid_cells;
    sk.contact_stride = 4 + 4 * CLAY_CONTACTS;
    sk.cmask = cmask;
    sk.cdata = cdata;
    return sk;
  }
};

// SdfEditor over a volume: its TileGrid (counts, ids, list, per-mip counters)
// and the device edit list, whose count lives on the device so that edits made
// by kernels (erosion) and by the host share one flush.  A volume with one
// instance per scene gets one of each per scene, scene-major.
struct Editor {
  Ctx* ctx = nullptr;
  Volume* vol = nullptr;         // not owned
  int max_edits = 64;
  int capacity = 8;
  double safety = 0.98;
  std::vector<std::vector<clay_brush>> pending;  // per scene
  std::vector<uint32_t> staged_counts;           // host copy source of the per-scene counts
  clay_brush* edits_dev = nullptr;   // scenes x max_edits
  uint32_t* nedit_dev = nullptr;     // scenes
  uint32_t* counts = nullptr;    // per scene: every mip, at the volume's cell offsets
  uint32_t* ids = nullptr;       // per scene: mip-0 tiles x capacity
  uint32_t* list = nullptr;      // per scene: every mip, at the volume's cell offsets
  uint32_t* args = nullptr;      // per scene and mip: live count, 1, 1, 0
};

// ---- launchers (implemented in the .hip files) ------------------------------
// edit.hip
int editor_flush_device(Editor* e, hipStream_t s, bool all_tiles);
int launch_contacts_to_edits(Ctx* c, hipStream_t s, Editor* e, const uint32_t* contacts, uint32_t contact_stride,
                             uint32_t scenes, uint32_t first_scene, double min_speed, double radius, double blend);
int launch_body_centroids(Ctx* c, hipStream_t s, bool f64, uint32_t nbodies, const uint32_t* first_count,
                          const float* const* pos, const float* const* rest, double* out);
// fluid.hip
// With `bounds`, also folds each scene's particle bounds into its six slots.
int launch_keys(Ctx* c, hipStream_t s, const FluidK& k, uint32_t n, const float* px, const float* py,
                const float* pz, const uint32_t* scene_of, uint32_t* keys, uint32_t* vals, uint32_t invalid_key,
                uint32_t shift, uint32_t* bounds, uint32_t* bump = nullptr);
// With `spos` (FP16-LDS), also each sorted particle's staged position.
int launch_reorder(Ctx* c, hipStream_t s, uint32_t n, const uint32_t* skeys, const uint32_t* svals,
                   const SceneK& sk, uint32_t invalid_key, uint32_t* cells, const float* const* src,
                   float* const* dst, int planes, const FluidK& k, uint2* spos);
// The FP16-LDS passes' per-particle records: staged positions (from the reorder),
// the sorted velocities, and the records the density pass writes for the forces pass.
struct HalfState {
  const uint2* spos = nullptr;
  const float *svx = nullptr, *svy = nullptr, *svz = nullptr;
  float4* frec = nullptr;
};
// With `cover` (by particle id, via svals), also each particle's squared interior
// radius for the splat (+inf when its neighbours do not surround it).
int launch_density(Ctx* c, hipStream_t s, const FluidK& k, const SceneK& sk, const FluidTune& tune, bool lds,
                   uint32_t n, const float* spx, const float* spy, const float* spz, const uint32_t* skeys,
                   const uint32_t* cells, float* srho, float* spres, const uint32_t* svals, float* cover,
                   const HalfState& half);
int launch_forces_integrate(Ctx* c, hipStream_t s, const FluidK& k, const SceneK& sk, const FluidTune& tune,
                            bool lds, const ColliderDev& col, uint32_t n, const float* const* sorted,
                            const uint32_t* svals, const uint32_t* skeys, const uint32_t* cells,
                            float* const* out, uint32_t* contacts, const float4* frec);
int launch_contacts_collect(Ctx* c, hipStream_t s, uint32_t* contacts, uint32_t stride, const uint32_t* cmask,
                            const float4* cdata, const uint32_t* base, uint32_t n, uint32_t scenes);
int launch_fill2_u32(Ctx* c, hipStream_t s, uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t v);
int launch_fill_u32(Ctx* c, hipStream_t s, uint32_t* p, size_t n, uint32_t v);
int launch_conserved(Ctx* c, hipStream_t s, uint32_t n, const float* const* pv, double mass,
                     double* partial, uint32_t partial_blocks, double* out);
// sort.hip
size_t radix_hist_size(uint32_t n);
// Stable sort of n (key, value) pairs of `key_bits`-bit keys; the result always
// lands in (keys, vals).  The input must be written to (keys_tmp, vals_tmp) when
// radix_input_in_tmp(key_bits), else to (keys, vals).
bool radix_input_in_tmp(int key_bits);
int radix_sort_pairs(Ctx* c, hipStream_t s, uint32_t n, int key_bits, uint32_t* keys, uint32_t* vals,
                     uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t* hist);
// splat.hip
int splat_bake(Splat* sp, hipStream_t s);
// `order`: visiting order (codes do not depend on it).  `cover`: per particle id,
// a squared radius beyond which the particle holds no voxel's minimum (a fluid
// surface's interior radius), or null.
int splat_bake_ordered(Splat* sp, hipStream_t s, const uint32_t* order, const float* cover);
// splat_bake_ordered in two parts: the binning (positions and order only), then the
// walk and the resolve (which read `cover`).
// `bounds_given`: the slots splat_bounds_begin returned were filled with this
// bake's particle bounds (by the frame-end k_keys), so the follow reads them.
int splat_bake_prepare(Splat* sp, hipStream_t s, const uint32_t* order, bool bounds_given);
// Resets and returns the per-scene bounds slots of a splat that follows its
// particles' bounds (else null), for a producer to fill before the bake.
int splat_bounds_reset(Splat* sp, hipStream_t s);
uint32_t* splat_bounds_begin(Splat* sp, hipStream_t s);
int splat_bake_finish(Splat* sp, hipStream_t s, const uint32_t* order, const float* cover);
int splat_brick_geometry(int res, int mips, std::vector<int>* brick_res,
                         std::vector<size_t>* brick_off, size_t* brick_total,
                         std::vector<size_t>* tile_off, size_t* tile_total);
size_t splat_max_bricks_per_particle(double radius, double band, double voxel0, int mips);
size_t splat_bin_slices();
// The static scratch of clay_splat_bake_static: looks up this device's copy
// (the device must be current), once per device.
int splat_static_init(int device);
size_t splat_static_bytes();
// peaks.hip
int atomic_peak(Ctx* c, int kind, uint32_t rounds, double* ops_per_s);
int valu_peak(Ctx* c, int form, uint32_t rounds, double* wave_insts_per_s);
int splat_bake_static(Ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc, uint16_t* dist,
                      hipStream_t s);
// volume.hip
int launch_bake_brushes(Ctx* c, hipStream_t s, Volume* v, const clay_brush* brushes_dev, int n);
int launch_field_eval(Ctx* c, hipStream_t s, const ColliderDev& col, int n, const float* pts,
                      float* dist, float* nrm);
int launch_set_origin_dev(Ctx* c, hipStream_t s, VolDev* dev, const float* origin3, double half);

}  // namespace clay
