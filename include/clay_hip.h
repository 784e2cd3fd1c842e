/*
 * clay_hip.h -- C ABI of the MI355X (gfx950) fluid-to-field path.
 *
 * This is the drop-in boundary for clay-engine's FluidSim step and SplatField
 * bake.  Each entry point replaces one member of the reference's TypeScript
 * surface (paths relative to the reference root); the reference passes shader
 * closures, which a C ABI cannot carry, so colliders and particle clouds are
 * passed as descriptors instead.
 *
 * Conventions
 *  - Every function returns an int status: CLAY_OK (0) or a negative code.  On
 *    failure clay_last_error() returns a thread-local message; constructor
 *    messages are the reference's own (e.g. volume.ts:156-160).
 *  - Handles are opaque and owned by the caller; *_destroy frees them.  A
 *    handle may be used from one host thread at a time.
 *  - Host-side option values are doubles, as JS numbers are: values the
 *    reference derives in JS (f64) and hands to the shaders as f32 are derived
 *    here the same way and rounded once.
 *  - clay_stream is a hipStream_t (NULL = the context's own stream).  step,
 *    bake and set_origin_device only enqueue work; the download, upload,
 *    get_state, set_state and read_contacts calls synchronise the stream.
 *  - Device pointers in descriptors are plain `float*` into HBM, laid out as
 *    structure-of-arrays (x[], y[], z[]).
 */
#ifndef CLAY_HIP_H
#define CLAY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLAY_OK 0
#define CLAY_ERR_INVALID (-1) /* bad argument or reference-rejected option */
#define CLAY_ERR_HIP (-2)     /* HIP runtime error */
#define CLAY_ERR_NODEVICE (-3)
#define CLAY_ERR_STATE (-4)   /* call not valid in the object's current state */

#define CLAY_MAX_MIPS 16
#define CLAY_MAX_LEAVES 4
#define CLAY_CONTACTS 32      /* ContactList capacity, fluid.ts:55-58 */

typedef struct clay_ctx clay_ctx;
typedef struct clay_volume clay_volume;
typedef struct clay_field clay_field;
typedef struct clay_splat clay_splat;
typedef struct clay_fluid clay_fluid;
typedef void* clay_stream;

/* ---- library ----------------------------------------------------------- */
const char* clay_last_error(void);
int clay_version(void);

/* The TgpuRoot equivalent: one device, one default stream (tgpu.init). */
int clay_ctx_create(int device, clay_ctx** out);
int clay_ctx_destroy(clay_ctx* ctx);
int clay_ctx_synchronize(clay_ctx* ctx);
/* Optional per-kernel HIP-event timing of everything the context launches. */
int clay_ctx_set_timing(clay_ctx* ctx, int enabled);
int clay_ctx_kernel_stats(clay_ctx* ctx, int index, char* name, int name_len,
                          double* total_ms, uint64_t* launches);
int clay_ctx_reset_stats(clay_ctx* ctx);

/* ---- brushes ------------------------------------------------------------ */
/* Same field order and 80-byte layout as `Brush` (lib/src/field/brush.ts:53-80). */
typedef struct {
  float pos[3];
  float scale;
  float rot[4];   /* xyzw unit quaternion */
  float size[3];
  float radius;
  float material;
  float blend;
  float bound;
  float mask;     /* material restriction, -1 = none */
  uint32_t kind;  /* BuiltinBrushKind (brush.ts:19-35): 0..6 supported */
  uint32_t op;    /* BrushOp (brush.ts:41-48): 0 add, 1 cut, 2 paint */
  uint32_t slot;
  uint32_t pad;
} clay_brush;

/* ---- SdfVolume (lib/src/field/volume.ts:91-343) ----------------------- */
typedef struct {
  int32_t resolution;     /* 0 -> 128 (volume.ts:140) */
  int32_t mip_levels;     /* 0 -> derived (volume.ts:148-151) */
  double world_size;      /* <= 0 -> 16 */
  double band;            /* <= 0 -> 4 */
  double origin[3];
  double gradient_voxels; /* <= 0 -> 1 (volume.ts:299) */
} clay_volume_desc;

typedef struct {
  int32_t resolution;
  int32_t mip_levels;
  int32_t tile_res[CLAY_MAX_MIPS];     /* tileResPerMip */
  int32_t cell_offsets[CLAY_MAX_MIPS]; /* cellOffsets */
  int32_t total_cells;                 /* totalCells */
  int32_t pad;
  double world_size;
  double voxel_size;
  double band;
  double origin[3];
} clay_volume_info;

/* Geometry derivation and the reference's resolution check, host only (no GPU):
 * volume.ts:139-179. */
int clay_volume_derive(int32_t resolution, int32_t mip_levels, clay_volume_info* out);
/* new SdfVolume(root, options): volume.ts:139. Storage starts saturated (+1, material 0). */
int clay_volume_create(clay_ctx* ctx, const clay_volume_desc* desc, clay_volume** out);
int clay_volume_destroy(clay_volume* vol);
int clay_volume_get_info(const clay_volume* vol, clay_volume_info* out);
/* SdfVolume.setOrigin: volume.ts:327-334 (origin and centre, f64 -> f32). */
int clay_volume_set_origin(clay_volume* vol, const double origin[3], clay_stream stream);
/* setOrigin from a device-resident vec3 (no host readback; graph-capturable). */
int clay_volume_set_origin_device(clay_volume* vol, const float* origin3_device,
                                  clay_stream stream);
int clay_volume_voxel_size_at(const clay_volume* vol, int32_t mip, double* out);
/* Texels of one mip as rgba16float (r = distance / band, g = material, b = 0, a = 1),
 * x fastest, (res >> mip)^3 * 4 halves. */
int clay_volume_download(clay_volume* vol, int32_t mip, uint16_t* rgba16_host);
int clay_volume_upload(clay_volume* vol, int32_t mip, const uint16_t* rgba16_host);

/* ---- The baked volume as a stream (a wire / on-disk format) ----------------
 * One instance, every mip, as 8^3 tiles in mip-major, then tile z, y, x order.  A
 * tile whose 512 texels are one (distance, material) bit pattern is stored as that
 * texel; any other tile is stored whole as 512 rgba16float texels (distance/band,
 * material, 0, 1; lib/src/field/volume.ts:77-90), x fastest, then y, then z.
 *   [0, 96)                    clay_wire_header
 *   [table_offset, +4 tiles)   u32 per tile: (material16 << 16) | dist16 for a
 *                              constant tile, the payload index of a stored one
 *   [bitmap_offset, ...)       1 bit per tile (u32 words, tile t: word t/32, bit t%32),
 *                              set for stored tiles
 *   [payload_offset, total)    4096 bytes per stored tile; payload_offset is 256-aligned
 * All fields little-endian.  This format has no counterpart in the reference, whose
 * bake lives only in a GPU texture; it carries exactly the reference's texels. */
#define CLAY_WIRE_VERSION 1u
typedef struct {
  uint32_t magic;          /* 0x42564C43, "CLVB" */
  uint32_t version;        /* CLAY_WIRE_VERSION */
  uint32_t resolution;     /* mip-0 texels per axis */
  uint32_t mip_levels;
  float band;              /* voxels; texels hold distance / (band * voxel * 2^m) */
  float world_size;
  float gradient_voxels;
  uint32_t tiles;          /* sum over mips of (resolution / 2^m / 8)^3 */
  double origin[3];        /* the volume origin (setOrigin) */
  uint32_t stored_tiles;
  uint32_t reserved;
  uint64_t table_offset, bitmap_offset, payload_offset, total_bytes;
} clay_wire_header;

/* Largest stream a volume of this shape can produce (every tile stored). */
int clay_volume_pack_bound(const clay_volume* vol, uint64_t* bytes);
/* Writes instance `instance` as a stream into device memory `dst` of `capacity`
 * bytes (at least clay_volume_pack_bound), enqueued on the stream.  With `bytes`
 * non-NULL it synchronises and returns the stream's length (also header.total_bytes). */
int clay_volume_pack(clay_volume* vol, uint32_t instance, void* dst_device, uint64_t capacity, uint64_t* bytes,
                     clay_stream stream);
/* Checks a stream header against a volume: magic, version, resolution, mip count,
 * band and section offsets must agree.  Host memory. */
int clay_wire_check(const clay_wire_header* header, const clay_volume* vol);
/* Replaces instance `instance`'s texels, every mip, with a stream in device memory
 * (`bytes` long).  With `header` NULL it synchronises: it reads the header from the
 * stream, and fails on a mismatch (clay_wire_check), a truncated stream, or a table
 * word naming a payload tile the stream does not hold (such tiles are skipped; the
 * others are written).  With the caller's copy of the header it only enqueues the
 * kernel; a bad table word then sets a flag that clay_volume_wire_status reports.  The
 * volume's origin and geometry are left as they are. */
int clay_volume_unpack(clay_volume* vol, uint32_t instance, const void* src_device, uint64_t bytes,
                       const clay_wire_header* header, clay_stream stream);
/* Synchronises the device; *bad = 1 (and an error) when an unpack since the last call
 * met a table word outside its payload; clears the flag. */
int clay_volume_wire_status(clay_volume* vol, uint32_t* bad);
/* SdfBuilder.setBrushes + rebuild (lib/src/field/builder.ts:87-172): dense bake of
 * a brush list into every mip. */
int clay_volume_bake_brushes(clay_volume* vol, const clay_brush* brushes, int32_t n,
                             clay_stream stream);

/* ---- TracedField (lib/src/trace/field.ts:20-73) ------------------------ */
typedef struct {
  double band;     /* <= 0 -> 64 (field.ts:200) */
  double epsilon;  /* <= 0 -> 1e-3 (field.ts:201) */
  double material;
} clay_analytic_opts;

/* analyticField (field.ts:196-236) whose closure is the fold of `brushes` from
 * (1e9, 0) with applyBrush (brush.ts:474-511). */
int clay_field_analytic(clay_ctx* ctx, const clay_brush* brushes, int32_t n,
                        const clay_analytic_opts* opts, clay_field** out);
/* volumeField (field.ts:102-180). The volume must outlive the field. */
int clay_field_volume(clay_ctx* ctx, clay_volume* vol, clay_field** out);
/* unionField (field.ts:246-315); a and b must outlive the union. */
int clay_field_union(clay_field* a, clay_field* b, clay_field** out);
/* offsetField (field.ts:388-429): the field read at p - offset, with offset a device
 * vec3 (3 floats) read at every evaluation, so the caller can move the field on the
 * device.  a and offset_dev must outlive the result.  A field offset once already is
 * refused. */
int clay_field_offset(clay_field* a, const float* offset_dev, clay_field** out);
/* lerpField (field.ts:316-386) of a morph pair: two volumeFields of the same resolution,
 * mips and instances, mixed by clamp(*t_dev, 0, 1) (a device float read at every
 * evaluation).  a, b and t_dev must outlive the result. */
int clay_field_lerp(clay_field* a, clay_field* b, const float* t_dev, clay_field** out);
int clay_field_destroy(clay_field* f);
int clay_field_max_mip(const clay_field* f, int32_t* out);
/* sample(p, 0).x and normal(p) at n host points (xyz interleaved); runs on the GPU. */
int clay_field_eval(clay_field* f, const float* points_host, int32_t n, float* dist_host,
                    float* normal_host);
/* What a tracer reads (field.ts:20-73), at n host points (xyz interleaved), on the
 * GPU: CLAY_QUERY_SAMPLE -> sample(p, mip) = (world distance, band-normalised value);
 * CLAY_QUERY_FIELD -> field(p, mip) = (world distance, material).  A volume leaf reads
 * level min(mip, maxMip) (instance 0 of a multi-instance volume); a union returns
 * the nearer child's pair.  out2: 2 floats per point. */
#define CLAY_QUERY_SAMPLE 0
#define CLAY_QUERY_FIELD 1
int clay_field_query(clay_field* f, int32_t what, int32_t mip, const float* points_host, int32_t n,
                     float* out2_host);
/* entry(ro, rd, tMin, radius, aperture) per ray (9 floats: ro xyz, rd xyz, tMin,
 * radius, aperture; field.ts:150-177 and 255-264): the first t >= tMin at which a
 * shape of radius radius + aperture * t could touch the field's box, or 1e9 (NEVER).
 * Fails for a field with an analytic leaf, which has no entry. */
int clay_field_entry(clay_field* f, const float* rays9_host, int32_t n, float* t_host);
/* bandWorld(mip) and voxelWorld(mip) in f32 (either pointer may be NULL). */
int clay_field_scale(const clay_field* f, int32_t mip, float* band_world, float* voxel_world);

/* makeTracer options (march.ts:22-40, 71-75): iteration cap, minimum forward step
 * in mip-0 voxels, hit threshold in voxels of min_mip, and the finest level the
 * trace may reach (clamped to the field's maxMip). */
typedef struct clay_tracer_opts {
  int32_t max_steps;  /* maxSteps, default 192 */
  float min_step;     /* minStep, default 0.25 */
  float hit_eps;      /* hitEps, default 0.5 */
  int32_t min_mip;    /* minMip, default 0 */
} clay_tracer_opts;
void clay_tracer_opts_default(clay_tracer_opts* opts);
/* SweepHit (march.ts:4-12): t, hit (0 = left the interval untouched), iterations
 * consumed, and the field distance at the stop point. */
typedef struct clay_sweep_hit {
  float t;
  uint32_t hit;
  uint32_t steps;
  float dist;
} clay_sweep_hit;
/* makeTracer(field, opts).sweep(ro, rd, tMin, tMax, radius, aperture) per ray
 * (march.ts:97-218): radius 0 and aperture 0 is .ray, aperture alone .cone, radius
 * alone .sphere.  rays10: 10 floats per ray (ro xyz, rd xyz, tMin, tMax, radius,
 * aperture); rd is used as given.  The sweep starts at entry when every leaf
 * bounds itself (volumes, lerps of volumes and their offsets), else at tMin.
 * Host buffers; synchronous. */
int clay_field_sweep(clay_field* f, const clay_tracer_opts* opts, const float* rays10_host, int32_t n,
                     clay_sweep_hit* hits_host);
/* The same over device buffers, enqueued on `stream`. */
int clay_field_sweep_device(clay_field* f, const clay_tracer_opts* opts, const float* rays10_dev, int32_t n,
                            clay_sweep_hit* hits_dev, clay_stream stream);

/* ---- ParticleCloud + SplatField (lib/src/sim/splat.ts) ----------------- */
#define CLAY_LIVE_ALL 0     /* occluder.gpu.ts:98-101 */
#define CLAY_LIVE_PREFIX 1  /* i < count (fluid.ts:221-224) */
#define CLAY_LIVE_RANGE 2   /* first <= i < first + count (splat.ts:56-60) */

/* ParticleCloud (splat.ts:15-38) as device SoA pointers plus a live rule. */
typedef struct {
  uint32_t capacity;
  int32_t live_rule;
  uint32_t live_first;
  uint32_t live_count;
  const float* pos[3];
  const float* nrm[3];  /* normalAt (splat.ts:37); all NULL for a union of spheres */
} clay_cloud_desc;

#define CLAY_SPLAT_PLAIN 0  /* dense u32 grid, global atomicMin: the reference's schedule */
#define CLAY_SPLAT_TILED 1  /* particles binned to LDS bricks; only live 8^3 tiles written */

typedef struct {
  double radius;          /* required (splat.ts:84) */
  double material;
  int32_t resolution;     /* 0 -> 64 */
  int32_t mip_levels;     /* 0 -> log2(res / 8) (splat.ts:149) */
  double band;            /* <= 0 -> 2 */
  double origin[3];       /* default (-12, -4, -12) is applied by the host layer */
  double world_size;      /* <= 0 -> 24 */
  int32_t mode;           /* CLAY_SPLAT_PLAIN | CLAY_SPLAT_TILED */
  int32_t keep_codes;     /* reserved: both schedules keep the u32 code grid */
} clay_splat_desc;

/* new SplatField(root, cloud, options): splat.ts:147-278. */
int clay_splat_create(clay_ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc,
                      clay_splat** out);
int clay_splat_destroy(clay_splat* s);
/* Change the live count of a PREFIX/RANGE cloud (the reference's params.count). */
int clay_splat_set_live(clay_splat* s, uint32_t first, uint32_t count);
/* SplatField.bake: splat.ts:296-304. */
int clay_splat_bake(clay_splat* s, clay_stream stream);
/* SplatField.setCenter: splat.ts:284-287. */
int clay_splat_set_center(clay_splat* s, const double center[3], clay_stream stream);
int clay_splat_volume(clay_splat* s, clay_volume** out);
int clay_splat_field(clay_splat* s, clay_field** out);
/* The u32 code grid of one mip after the last bake. */
int clay_splat_grid_download(clay_splat* s, int32_t mip, uint32_t* codes_host);
/* Exact live-tile mask of one mip after the last bake: 1 byte per 8^3 tile. */
int clay_splat_tiles_download(clay_splat* s, int32_t mip, uint8_t* live_host);

/* ---- Atomic throughput of the splat ------------------------------------
 * The bake's walk is a stream of unsigned min-atomics.  clay_atomic_peak measures
 * the issue rate of each kind on this device (ops/s): LDS ds_min_u32 with every
 * lane on its own bank, or global atomicMin with a wave on one 256-byte span of a
 * 256 MiB buffer; `rounds` (1..65536) atomics per lane of 8 512-lane workgroups per
 * CU.  Synchronous.  No counterpart in the reference (splat.ts:241-253 issues the
 * atomics; nothing there measures them). */
#define CLAY_ATOMIC_LDS_MIN 0
#define CLAY_ATOMIC_GLOBAL_MIN 1
int clay_atomic_peak(clay_ctx* ctx, int32_t kind, uint32_t rounds, double* ops_per_s);
/* Measured VALU issue rate of this device for one instruction form: wave64
 * instructions retired per second, from eight independent register chains per lane
 * at full occupancy, `rounds` (1..2^20) per chain.  Synchronous.  No counterpart in
 * the reference: the two-operand all-VGPR forms (CLAY_VALU_FMAC, _ADD_U32, _MUL_F32)
 * give the full-rate ceiling the neighbour and splat kernels' instruction counts are
 * set against, and the other forms are the slower encodings those loops contain. */
#define CLAY_VALU_FMAC 0      /* v_fmac_f32 v, v, v */
#define CLAY_VALU_FMAC_SGPR 1 /* v_fmac_f32 v, s, v */
#define CLAY_VALU_FMA_VOP3 2  /* v_fma_f32 v, v, v, v */
#define CLAY_VALU_FMA_MIX 3   /* v_fma_mix_f32 (f16 operand widened) */
#define CLAY_VALU_MAX 4       /* v_max_f32 */
#define CLAY_VALU_CNDMASK 5   /* v_cndmask_b32 with an SGPR-pair mask */
#define CLAY_VALU_RSQ 6       /* v_rsq_f32 (transcendental unit) */
#define CLAY_VALU_ADD_U32 7   /* v_add_u32 */
#define CLAY_VALU_MUL_F32 8   /* v_mul_f32 */
int clay_valu_peak(clay_ctx* ctx, int32_t form, uint32_t rounds, double* wave_insts_per_s);
/* Bakes the surface of the current state again (SplatField.bake, splat.ts:296-304,
 * with the same codes) and counts the atomics of its owner walk: out[0] the LDS
 * ds_min_u32 (one per walked voxel), out[1] the global atomicMin of the merge.
 * Call after a step; the surface must use the tiled schedule (zeros when the
 * brick walk runs instead of the owner walk).  Synchronous. */
int clay_fluid_splat_atomics(clay_fluid* f, uint64_t out[2]);

/* A bake that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It is clay_splat_bake with
 * the plain schedule (desc->mode must be CLAY_SPLAT_PLAIN) of a splat created
 * from the same cloud and desc, but nothing is created: the caller passes the
 * cloud (device SoA) and the output, the f16 distance plane (.r of the
 * reference's rgba16float texels, distance / band) of every mip concatenated,
 * mip-major and x fastest: sum over m of (res >> m)^3 uint16_t in device memory.
 * The dense u32 code grid and the origin record live in a static device array
 * of 32 MiB per GPU, held until the process ends; a bake needing more (above
 * 192^3 in 3 mips, for example) is refused with CLAY_ERR_INVALID.  Enqueued only.
 * Every call on one GPU shares that array: calls that overlap on one GPU (two
 * streams, two threads) race, and must be ordered by the caller. */
int clay_splat_bake_static(clay_ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc,
                           uint16_t* dist_dev, clay_stream stream);

/* ---- FluidSim (lib/src/sim/fluid.ts:111-582) --------------------------- */
typedef struct {
  uint32_t capacity;        /* 8192 */
  int32_t bucket;           /* 64 */
  double spacing;           /* 0.45 */
  double smoothing;         /* <= 0 -> 2.2 * spacing */
  double density;           /* mass = density * spacing^3; reference = 1 */
  double radius;            /* < 0 -> 0.5 * spacing */
  double surface_radius;    /* < 0 -> 1.1 * spacing */
  double material;
  int32_t bake_resolution;  /* 64 */
  int32_t bake_mip_levels;  /* 0 -> max(1, log2(res / 8)) */
  double bake_band;         /* 2 */
  double origin[3];         /* (-12, -4, -12) */
  double world_size;        /* 24 */
  int32_t substeps;         /* 3 */
  int32_t splat_mode;       /* CLAY_SPLAT_PLAIN */
  double gravity[3];        /* (0, -9.81, 0) */
  double stiffness;         /* 220 */
  double viscosity;         /* 0.4 */
  double damping;           /* 0.995 */
  double max_speed;         /* 15 */
  int32_t precision;        /* CLAY_PRECISION_F32 (bit-exact) | CLAY_PRECISION_FP16LDS */
  int32_t pad;
  /* > 0: the surface is a followSize-wide volume re-centred on the particles'
   * bounds before every bake (SplatField.setCenter as SoftBody uses it,
   * softbody.ts:424-428), computed on the device; 0: the reference's
   * world-covering surface (fluid.ts:215-234). */
  double follow_size;
} clay_fluid_desc;

/* Arithmetic of the density and force passes.  F32: every operation rounds once
 * in the reference's order (bit-exact against an IEEE-f32 evaluation).  FP16LDS:
 * neighbour positions staged in LDS as f16 offsets from their cell's centre,
 * compensated f32 density sums, fused multiply-add allowed, pair distances from
 * the hardware reciprocal square root (1 ulp), 1/density once per neighbour. */
#define CLAY_PRECISION_F32 0
#define CLAY_PRECISION_FP16LDS 1

/* The reference's defaults (fluid.ts:141-203). */
void clay_fluid_desc_default(clay_fluid_desc* d);

/* FluidParams minus the emission cursor: what FluidSim.configure may patch (fluid.ts:528). */
typedef struct {
  double gravity[3];
  double dt;
  double emit_pos[3];
  double emit_radius;
  double emit_vel[3];
  double max_speed;
  double stiffness;
  double viscosity;
  double damping;
  double contact_depth;
} clay_fluid_params;

/* Derived constants (fluid.ts:141-203, 239-247, 444-447), f64. */
typedef struct {
  double h, mass, poly6, spiky_grad, rest_density, dt, radius, contact_depth;
  double grid_origin[3];
  int32_t grid_res;
  int32_t bucket;
  int32_t substeps;
  int32_t pad;
  double wall_lo[3];
  double wall_hi[3];
} clay_fluid_consts;

/* new FluidSim(root, world, options): fluid.ts:140-493. */
int clay_fluid_create(clay_ctx* ctx, clay_field* world, const clay_fluid_desc* desc,
                      clay_fluid** out);
int clay_fluid_destroy(clay_fluid* f);
int clay_fluid_constants(const clay_fluid* f, clay_fluid_consts* out);
/* Replaces the first n particles (SoA host arrays: x[n], y[n], z[n]); the live count
 * becomes n.  vel may be NULL (zero).  The reference fills particles only by emit;
 * this is the seeded-lattice entry the configurations use. */
int clay_fluid_set_state(clay_fluid* f, const float* pos_soa, const float* vel_soa, uint32_t n);
#define CLAY_STATE_POS 0
#define CLAY_STATE_VEL 1
#define CLAY_STATE_DENSITY 2
#define CLAY_STATE_PRESSURE 3
#define CLAY_STATE_NVEL 4
/* Reads `what` for the live particles: 3 * count floats (SoA) or count floats. */
int clay_fluid_get_state(clay_fluid* f, int32_t what, float* host);
/* FluidSim.spawn: fluid.ts:511-525 (spread < 0 -> spacing * cbrt(max(n,1)) * 0.62). */
int clay_fluid_spawn(clay_fluid* f, uint32_t n, const double pos[3], const double vel[3],
                     double spread);
/* FluidSim.configure: fluid.ts:528-530. */
int clay_fluid_configure(clay_fluid* f, const clay_fluid_params* p);
int clay_fluid_get_params(const clay_fluid* f, clay_fluid_params* out);
/* FluidSim.step: emit, substeps x (bin, density, forces, integrate), bake (fluid.ts:536-558). */
int clay_fluid_step(clay_fluid* f, clay_stream stream);
/* Capture a frame once as a HIP graph and replay it on later steps (off by default).
 * Frames that emit run eagerly; a changed live count, configure and set_tuning
 * recapture.  No counterpart in the reference (WebGPU has no graphs); for callers
 * that step one FluidSim many times with the same settings. */
int clay_fluid_set_graph(clay_fluid* f, int32_t enabled);
/* One substep without emit or bake (for per-substep parity checks). */
int clay_fluid_substep(clay_fluid* f, clay_stream stream);
/* FluidSim.readContacts: fluid.ts:565-569.  out4 holds min(raw, 32) points
 * (x, y, z, speed); raw_count is the atomic counter, which may exceed 32. */
int clay_fluid_read_contacts(clay_fluid* f, float* out4, uint32_t* n, uint32_t* raw_count);
/* Conserved quantities of the live particles in FP64, reduced in a fixed order:
 * out = {total mass, momentum x, y, z, centre of mass x, y, z}.  The reference
 * reduces its body centre of mass in f32 (pbd.ts:240-297); this is the FP64 form
 * for any particle set.  Synchronises the stream. */
int clay_fluid_conserved(clay_fluid* f, double out[7]);
/* FluidSim.particleCount: fluid.ts:496-498. */
int clay_fluid_particle_count(const clay_fluid* f, uint32_t* out);
/* FluidSim.surface (the owned SplatField) and FluidSim.field. */
int clay_fluid_surface(clay_fluid* f, clay_splat** out);
/* Cell keys and the stable sort permutation of the last substep (count entries each). */
#define CLAY_DEBUG_KEYS 0
#define CLAY_DEBUG_PERM 1
int clay_fluid_debug_download(clay_fluid* f, int32_t what, uint32_t* host);

typedef struct clay_batch clay_batch;

/* ---- SdfEditor (lib/src/field/modify.ts:41-238) ---------------------------
 * Brush edits applied in place to a volume's mip-0 tiles they reach, then the
 * dirty-tile mask walked down the mip chain (TileGrid, tilegrid.ts:43-316) and
 * each coarser level rebuilt over its live tiles (MipRefiner, mips.ts:36-195).
 * The edit list and every tile count stay on the device: a flush needs no host
 * readback.  The editor binds the volume's origin at creation, as the
 * reference's does (modify.ts:90-97). */
typedef struct clay_editor clay_editor;
typedef struct {
  int32_t max_edits;      /* 64: edits per flush (modify.ts:73) */
  int32_t tile_capacity;  /* 8: edits recorded per tile; later ones are dropped (modify.ts:78) */
  double safety;          /* 0.98: chamfer safety factor of the band re-expansion (mips.ts:46) */
} clay_editor_desc;
int clay_editor_create(clay_volume* vol, const clay_editor_desc* desc, clay_editor** out);
int clay_editor_destroy(clay_editor* e);
/* SdfEditor.push (modify.ts:200-212): queues edits; *accepted says how many fit
 * before max_edits (the rest are dropped, as the reference drops them). */
int clay_editor_push(clay_editor* e, const clay_brush* edits, int32_t n, int32_t* accepted);
int clay_editor_pending(const clay_editor* e, int32_t* n);
/* SdfEditor.flush (modify.ts:214-238); *flushed = 0 when nothing was queued. */
int clay_editor_flush(clay_editor* e, clay_stream stream, int32_t* flushed);
/* Erosion: every contact of the fluid's last substep whose speed exceeds
 * min_speed becomes a cut sphere (BrushSet.make {kind 'sphere', size radius, op
 * 'cut', smooth blend}) appended after the queued edits, then the editor flushes
 * -- all on the device (the policy game/fluid.ts:86-95 leaves to the caller). */
int clay_editor_erode(clay_editor* e, clay_fluid* fluid, double min_speed, double radius, double blend,
                      clay_stream stream);
/* The same from one scene of a batch. */
int clay_editor_erode_scene(clay_editor* e, clay_batch* batch, uint32_t scene, double min_speed, double radius,
                            double blend, clay_stream stream);
/* Several scenes: a volume with one instance per scene (clay_volume_create_instances)
 * gets one queue, tile grid and edit list per scene; every scene flushes in the same
 * launches. */
int clay_editor_push_scene(clay_editor* e, uint32_t scene, const clay_brush* edits, int32_t n, int32_t* accepted);
/* Every scene's contacts -> that scene's edits -> flush, for an editor with one
 * instance per scene of `batch`. */
int clay_editor_erode_batch(clay_editor* e, clay_batch* batch, double min_speed, double radius, double blend,
                            clay_stream stream);
int clay_editor_tile_counts_scene(clay_editor* e, uint32_t scene, uint32_t* per_mip);
int clay_editor_cells_download_scene(clay_editor* e, uint32_t scene, uint32_t* counts);
/* Rebuilds every coarser mip from mip 0 over all tiles (MipRefiner on a full grid). */
int clay_editor_refresh_all(clay_editor* e, clay_stream stream);
/* TileGrid.readTileCounts (tilegrid.ts:281-297): live tiles per mip of the last flush. */
int clay_editor_tile_counts(clay_editor* e, uint32_t* per_mip);
/* The tile counts of every mip (the volume's total_cells entries) after the last flush. */
int clay_editor_cells_download(clay_editor* e, uint32_t* counts);

/* ---- Scenes: per-scene volume instances and splats ------------------------
 * A volume can hold one instance per scene (the same geometry, its own planes and
 * origin record, scene-major).  clay_volume_bake_brushes and the download/upload
 * calls address instance 0; clay_volume_replicate copies it to every instance.  A
 * volumeField over it reads, for a particle of batch scene s, instance s.  A splat
 * over a cloud partitioned into scenes bakes scene s's particles into instance s;
 * each instance can follow its particles' bounds, or device-side centres (6 doubles
 * per scene, xyz first, as clay_body_centroids_device writes them), before a bake. */
int clay_volume_create_instances(clay_ctx* ctx, const clay_volume_desc* desc, uint32_t instances, clay_volume** out);
int clay_volume_replicate(clay_volume* vol, clay_stream stream);
int clay_volume_instance_download(clay_volume* vol, uint32_t instance, int32_t mip, uint16_t* rgba);
#define CLAY_FOLLOW_NONE 0
#define CLAY_FOLLOW_BOUNDS 1
#define CLAY_FOLLOW_CENTERS 2
int clay_splat_create_scenes(clay_ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc,
                             uint32_t scenes, const uint32_t* counts, int32_t follow, const double* centers_dev,
                             clay_splat** out);
int clay_splat_instance_download(clay_splat* s, uint32_t scene, int32_t mip, uint32_t* codes_host);

/* ---- Body centroids (ClaySolver.match, lib/src/sim/pbd.ts:240-266) -------
 * Per body [first, first + count) of a particle set: com = sum(pos) / count and
 * restCom = sum(rest) / count, reduced by one 256-lane workgroup per body
 * (lane-strided sums, then a pairwise tree).  F32_TREE is the reference's f32
 * reduction; F64 is the same order in double, the form BodyTracker.sync
 * (particles.ts:264-294) should read.  rest may be NULL (restCom = 0). */
#define CLAY_REDUCE_F32_TREE 0
#define CLAY_REDUCE_F64 1
int clay_body_centroids(clay_ctx* ctx, const float* const pos_dev[3], const float* const rest_dev[3],
                        const uint32_t* first_count, uint32_t nbodies, int32_t mode, double* out6,
                        clay_stream stream);
/* The same, all on the device (table and result in device memory), enqueued only:
 * what a captured frame calls before a centre-following bake. */
int clay_body_centroids_device(clay_ctx* ctx, const float* const pos_dev[3], const float* const rest_dev[3],
                               const uint32_t* first_count_dev, uint32_t nbodies, int32_t mode, double* out6_dev,
                               clay_stream stream);

/* Launch shapes of the neighbour and splat kernels.  Results do not depend on
 * them; clay-engine_amd/clayhip/autotune.py picks them per device and workload. */
typedef struct {
  int32_t block;          /* density / forces workgroup: 128 or 256 threads */
  int32_t stage;          /* FP16-LDS staging slice per wave: 64, 128 or 256 neighbours */
  int32_t chunk;          /* splat work item: 256..65536 particles of one brick */
  int32_t groups_per_cu;  /* persistent splat grid: 1..32 workgroups per compute unit */
  int32_t fluid_xcd;      /* 1: neighbour kernels number workgroups XCD-contiguously (one L2 per run) */
  int32_t splat_xcd;      /* 1: the same for splat work items */
  int32_t splat_owner;    /* 1: each particle binned once per mip to the brick holding it, which
                             stages itself plus a 5-voxel shell in LDS (when reach <= 5 voxels) */
  int32_t exact_lds;      /* F32 precision: 1 stages neighbour rows in LDS as f32 (results are the
                             global-memory kernels' bit for bit), 0 walks them in global memory */
  int32_t early_bin;      /* 1: the bake's binning forks onto its second stream before the frame-end
                             reorder (beside it and the density pass), 0: after it */
} clay_tuning;
int clay_fluid_set_tuning(clay_fluid* f, const clay_tuning* t);
/* The splat fields of clay_tuning (chunk, groups_per_cu, splat_xcd, splat_owner) for a
 * SplatField; the others are ignored. */
int clay_splat_set_tuning(clay_splat* s, const clay_tuning* t);
int clay_fluid_get_tuning(const clay_fluid* f, clay_tuning* out);

/* ---- Batched scenes ------------------------------------------------------
 * Independent FluidSims of one option set (the world collider and constants
 * shared) stepped by the same launches: the reference runs one FluidSim per
 * scene, each with its own dispatches.  Scene i holds counts[i] particles, ids
 * [first_i, first_i + counts[i]) scene-major; each scene keeps its own cell
 * table, contact list and surface volume (follow_size wide, following the
 * scene's bounds).  A scene's results are bit-identical to a clay_fluid of the
 * same options holding only that scene's particles. */
typedef struct {
  uint32_t scenes;          /* 1..256 */
  uint32_t pad;
  const uint32_t* counts;   /* host, scenes entries; the sum is the batch size */
  clay_fluid_desc fluid;    /* shared options; capacity is ignored; follow_size > 0;
                               splat_mode must be CLAY_SPLAT_TILED */
} clay_batch_desc;
int clay_batch_create(clay_ctx* ctx, clay_field* world, const clay_batch_desc* desc, clay_batch** out);
int clay_batch_destroy(clay_batch* b);
/* Replaces every particle (SoA host arrays of the batch size, scene-major). */
int clay_batch_set_state(clay_batch* b, const float* pos_soa, const float* vel_soa);
/* Reads `what` (CLAY_STATE_*) for every particle, as clay_fluid_get_state. */
int clay_batch_get_state(clay_batch* b, int32_t what, float* host);
/* One FluidSim.step of every scene.  With graphs on (the default) the frame is
 * captured on its first call on a stream and replayed afterwards; configure,
 * set_tuning and set_graph recapture. */
int clay_batch_step(clay_batch* b, clay_stream stream);
int clay_batch_set_graph(clay_batch* b, int32_t enabled);
int clay_batch_configure(clay_batch* b, const clay_fluid_params* p);
int clay_batch_set_tuning(clay_batch* b, const clay_tuning* t);
int clay_batch_get_tuning(const clay_batch* b, clay_tuning* out);
int clay_batch_constants(const clay_batch* b, clay_fluid_consts* out);
int clay_batch_scene(const clay_batch* b, uint32_t scene, uint32_t* first, uint32_t* count);
int clay_batch_read_contacts(clay_batch* b, uint32_t scene, float* out4, uint32_t* n, uint32_t* raw_count);
/* A scene's surface after the last bake: texels (as clay_volume_download), u32
 * codes (as clay_splat_grid_download), live tiles, and the device-side origin
 * its following volume was baked at. */
int clay_batch_volume_download(clay_batch* b, uint32_t scene, int32_t mip, uint16_t* rgba);
int clay_batch_grid_download(clay_batch* b, uint32_t scene, int32_t mip, uint32_t* codes_host);
int clay_batch_tiles_download(clay_batch* b, uint32_t scene, int32_t mip, uint8_t* live_host);
int clay_batch_origin(clay_batch* b, uint32_t scene, float origin[3]);
/* clay_fluid_conserved over one scene's particles. */
int clay_batch_conserved(clay_batch* b, uint32_t scene, double out[7]);
/* clay_fluid_splat_atomics over every scene's following bake. */
int clay_batch_splat_atomics(clay_batch* b, uint64_t out[2]);

/* ---- ClaySolver: shape-matched particle bodies (lib/src/sim/pbd.ts:78-424) ----
 * A ParticleSet (lib/src/sim/particles.ts:18-206) of `capacity` particles and up
 * to `max_bodies` bodies, stepped against a TracedField.  Per substep: predict
 * (BDF2 + field collision with positional friction), match (centroids and the
 * covariance, reduced per body in the reference's 256-lane order, then the
 * warm-started polar rotation of Mueller et al. 2016), apply (shape matching,
 * collision projection, plasticity), finalize (velocity from positions).  Body b
 * collides with instance b of a per-scene world field (or the one world). */
typedef struct clay_solver clay_solver;

typedef struct {          /* ClaySolverOptions, pbd.ts:44-55 (defaults pbd.ts:102-115) */
  double gravity[3];      /* (0, -9.81, 0) */
  double dt;              /* 1/60 */
  int32_t substeps;       /* 2 */
  int32_t rotation_iterations; /* 2 */
  double radius;          /* 0.1 */
  double friction;        /* 0.35 */
  double damping;         /* 0.995 */
  double max_speed;       /* 40 */
} clay_solver_desc;
void clay_solver_desc_default(clay_solver_desc* d);

#define CLAY_BODY_KINEMATIC 1u  /* BodyFlags.kinematic (particles.ts:38-43) */
#define CLAY_BODY_WEIGHTLESS 2u /* BodyFlags.weightless */

typedef struct {          /* Body (particles.ts:52-68), the host-owned fields */
  uint32_t first, count;  /* the body's slice of particles */
  uint32_t flags;
  float stiffness;        /* ParticleSet.addBody default 0.5 */
  float plasticity;       /* per substep; default 0 */
  float pad;
} clay_body_desc;

/* new ParticleSet({capacity, maxBodies}) + new ClaySolver(set, world, options). */
int clay_solver_create(clay_ctx* ctx, clay_field* world, const clay_solver_desc* desc, uint32_t capacity,
                       uint32_t max_bodies, clay_solver** out);
int clay_solver_destroy(clay_solver* s);
/* ParticleSet.addBody / setBody for bodies [0, n): com = restCom = 0 and rot =
 * identity, as addBody writes them; count of declared bodies = n. */
int clay_solver_set_bodies(clay_solver* s, const clay_body_desc* bodies, uint32_t n);
/* ParticleSet.addBody (particles.ts:146-168): declares one more body and writes
 * only its record; *index = its body index. Fails past max_bodies with the
 * reference's message. */
int clay_solver_add_body(clay_solver* s, const clay_body_desc* body, uint32_t* index);
/* Particle slots [0, n): pos and rest (SoA host, 3*n floats each), owning body
 * (0xFFFFFFFF = free); prev = pos, vel = velPrev = 0, restNormal = 0, material 0. */
int clay_solver_set_particles(clay_solver* s, const float* pos, const float* rest, const uint32_t* body,
                              uint32_t n);
/* ClaySolver.setForce (pbd.ts:383-394): acceleration for this frame, impulse once. */
int clay_solver_set_force(clay_solver* s, uint32_t body, const double accel[3], const double impulse[3]);
/* ClaySolver.step (pbd.ts:400-415), enqueued on the stream. */
int clay_solver_step(clay_solver* s, clay_stream stream);
/* Particle planes (SoA host, 3 * capacity floats): 0 pos, 1 prev, 2 vel, 3 velPrev,
 * 4 rest, 5 restNormal; 6 material (capacity floats). */
int clay_solver_get_particles(clay_solver* s, int32_t what, float* host);
/* Owning body per particle (capacity u32). */
int clay_solver_get_owners(clay_solver* s, uint32_t* host);
/* Per body: com xyz, restCom xyz, rot xyzw (10 floats). */
int clay_solver_get_bodies(clay_solver* s, float* host);
/* Per body: first, count, flags (3 u32), as the GPU holds them. */
int clay_solver_get_slices(clay_solver* s, uint32_t* host);
/* Device pointers of the position planes (a ParticleCloud over the set). */
int clay_solver_positions(clay_solver* s, const float* out[3]);

/* ---- SurfaceExtractor (lib/src/sim/extract.ts:85-484) ----------------------
 * Naive surface nets of a field's zero level set on a res^3 lattice: per cell the
 * average of the 12 edge crossings, snapped once along the normal; the particles go
 * to slots [base, base + capacity) of the set's particle planes (pos = prev = rest,
 * vel = velPrev = the region's or the motion record's velocity, restNormal,
 * material, owner = body) and one quad per lattice edge that straddles the surface
 * to an index list of ceil(1.5 capacity) quads.  The reference takes slots with
 * atomic counters; here particles are numbered in ascending cell index and quads in
 * ascending lattice point then edge axis x, y, z, one of the orders it allows.
 * `finish` sets the body's first/count, com = restCom = the centroid of the rest
 * positions, rot = identity.  The set and the field must outlive the extractor. */
typedef struct clay_extractor clay_extractor;

typedef struct {
  uint32_t resolution; /* lattice cells per axis; 0 -> 32 (extract.ts:105), at most 256 */
  uint32_t body;       /* body index of the emitted particles */
  uint32_t base;       /* first particle slot of the body's slice */
  uint32_t capacity;   /* particles the extractor may emit */
  int32_t motion;      /* nonzero: keep an ExtractMotion record (extract.ts:37-42) on the device */
  uint32_t instance;   /* the instance an instanced field is read at (0 otherwise) */
} clay_extractor_desc;

/* new SurfaceExtractor(set, field, options) (extract.ts:103-444); fails with the
 * reference's message when the slice exceeds the set capacity. */
int clay_extractor_create(clay_solver* set, clay_field* field, const clay_extractor_desc* desc,
                          clay_extractor** out);
int clay_extractor_destroy(clay_extractor* e);
/* SurfaceExtractor.setRegion (extract.ts:451-463): lattice origin (relative to the
 * motion centre when there is a motion record), size = res * cell, velocity for
 * emitted particles, and whether the next extract snapshots the body's motion. */
int clay_extractor_set_region(clay_extractor* e, const double origin[3], double size, const double vel[3],
                              int32_t preserve_motion);
/* SurfaceExtractor.extract (extract.ts:469-475), enqueued on the stream. */
int clay_extractor_extract(clay_extractor* e, clay_stream stream);
/* Synchronises; counts = particles and quads found (before the capacities cut them),
 * draw_args = (6 * quads kept, 1, 0, 0), indices = 6 * quad_capacity u32. Each may be NULL. */
int clay_extractor_read(clay_extractor* e, uint32_t counts[2], uint32_t draw_args[4], uint32_t* indices);
int clay_extractor_quad_capacity(const clay_extractor* e, uint32_t* out);
/* The ExtractMotion record {centre xyz, 0, velocity xyz, 0}: written from `in`, then
 * read into `out` (either may be NULL). Synchronises. */
int clay_extractor_motion(clay_extractor* e, const float* in, float* out);

#ifdef __cplusplus
}
#endif
#endif /* CLAY_HIP_H */
