/*
 * clay_oracle.h -- CPU oracle for clay-engine's fluid-to-field path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * code: it exists so that tests/, bench.py's cpu_baseline leg and
 * __graft_entry__.smoke() can check the HIP library against an independent,
 * plain-C restatement of the reference algorithm.
 *
 * Every entry point exists twice: `_f32` ("WGSL-f32" mode: IEEE binary32, no
 * FMA contraction, operations in the reference shader's order) and `_f64`
 * (the same algorithm in binary64, the accuracy yardstick).  Constants that
 * the reference computes in JS (f64) and bakes into WGSL as f32 literals are
 * derived here in double and rounded once to the working type.
 *
 * Reference files (paths relative to the reference root):
 *   lib/src/sim/fluid.ts      FluidSim: constants, grid, density, forces, integrate
 *   lib/src/sim/splat.ts      SplatField: clear, splat[m], resolve[m]
 *   lib/src/field/volume.ts   SdfVolume: mip derivation, sampling, gradient
 *   lib/src/trace/field.ts    volumeField / analyticField / unionField
 *   lib/src/field/brush.ts    primitives, evalBrush, applyBrush
 *   lib/src/field/builder.ts  SdfBuilder.gen (dense brush bake)
 *   lib/src/field/tilegrid.ts packTile, compact, dilate
 *   lib/src/field/mips.ts     MipRefiner (eikonal band re-expansion)
 *   lib/src/field/modify.ts   SdfEditor bin/apply
 *   lib/src/sim/pbd.ts        ClaySolver.match centroids (body centre of mass)
 *   lib/src/math/gpu.ts       hash1/hash3, smin/smax
 *
 * Parity status of each function is recorded in the comment above it in
 * clay_oracle.c; a function with no independent check says "parity unpinned".
 */
#ifndef CLAY_ORACLE_H
#define CLAY_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_INVALID_KEY 0xFFFFFFFFu
#define ORC_MAX_LEAVES 4
#define ORC_MAX_MIPS 8
#define ORC_CONTACTS 32

/* Brush record: same field order and 80-byte size as the reference's `Brush`
 * struct (lib/src/field/brush.ts:53-80). */
typedef struct {
  float pos[3];
  float scale;
  float rot[4];
  float size[3];
  float radius;
  float material;
  float blend;
  float bound;
  float mask;
  uint32_t kind;
  uint32_t op;
  uint32_t slot;
  uint32_t pad;
} orc_brush;

/* FluidSim options (lib/src/sim/fluid.ts:60-91, defaults :141-203).
 * `density` scales the particle mass (mass = density * spacing^3); the
 * reference has density 1. Fields set to a negative value take the default. */
typedef struct {
  double spacing;
  double smoothing;
  double density;
  double radius;
  int32_t bucket;
  int32_t substeps;
  double origin[3];
  double world_size;
  double gravity[3];
  double stiffness;
  double viscosity;
  double damping;
  double max_speed;
} orc_fluid_opts;

/* Derived FluidSim constants, in f64 as the reference's JS computes them
 * (fluid.ts:141-203, 239-266, 444-447). Products that the shader sees as a
 * single literal are stored pre-multiplied. */
typedef struct {
  double h, mass, poly6, spiky_grad, rest_density, dt, radius, contact_depth;
  double grid_origin[3];
  int32_t grid_res;
  int32_t bucket;
  int32_t substeps;
  int32_t pad0;
  double wall_lo[3], wall_hi[3];
  double gravity[3];
  double stiffness, viscosity, damping, max_speed;
  double hh, mass_poly6, h6, inv_h, rest_floor;
} orc_fluid_k;

/* One collider leaf: an analytic brush fold or a baked volume (mip 0). */
typedef struct orc_leaf_s {
  int32_t kind;            /* 0 = analytic brush list, 1 = volume, 2 = lerp of two volumes */
  int32_t nbrush;
  const orc_brush* brushes;
  double eps;              /* analytic: central-difference epsilon (field.ts:201) */
  /* volume (lib/src/field/volume.ts uniform VolumeParams) */
  const uint16_t* dist16;  /* mip-0 `.r` channel as f16 bits, x-fastest, res^3 */
  int32_t res;
  int32_t pad1;
  double origin[3];
  double world_size;
  double band;
  double grad_voxels;
  const uint16_t* mat16;   /* volume: mip-0 `.g` channel (f16 bits), NULL reads 0 */
  double material;         /* analytic: the material `field` reports (field.ts:213-216) */
  /* Every mip, for the reads at any level (orc_field_query): level m has res >> m
   * texels per axis; NULL planes beyond `mips`.  Analytic: the band option. */
  int32_t mips;
  int32_t pad2;
  const uint16_t* dist_m[ORC_MAX_MIPS];
  const uint16_t* mat_m[ORC_MAX_MIPS];
  double analytic_band;
  /* offsetField (field.ts:388-429): when has_offset, p - offset is read. */
  int32_t has_offset;
  int32_t pad3;
  double offset[3];
  /* lerpField (field.ts:316-386), kind 2: this leaf's volume is `a`, lerp_b the
   * second volume (same resolution, mips and box), mixed by clamp(lerp_t, 0, 1). */
  const struct orc_leaf_s* lerp_b;
  double lerp_t;
} orc_leaf;

/* unionField over leaves, folded left: union(union(a, b), c) (field.ts:246-315). */
typedef struct {
  int32_t nleaf;
  int32_t pad;
  orc_leaf leaf[ORC_MAX_LEAVES];
} orc_collider;

/* makeTracer options (lib/src/trace/march.ts:22-40, 71-75). */
typedef struct {
  int32_t max_steps;
  int32_t min_mip;
  double min_step;   /* mip-0 voxels */
  double hit_eps;    /* voxels of min_mip */
} orc_tracer_opts;

/* SplatField options (lib/src/sim/splat.ts:74-106, 147-162). `origin` is the
 * volume origin at bake time (SdfVolume.setOrigin). */
typedef struct {
  int32_t res;
  int32_t mips;
  double world_size;
  double origin[3];
  double band;
  double radius;
  double material;
} orc_splat_opts;

/* Fluid state, structure-of-arrays, in the working precision. */
typedef struct {
  void* px; void* py; void* pz;
  void* vx; void* vy; void* vz;
  void* rho; void* pres;
  void* nvx; void* nvy; void* nvz;
} orc_state;

typedef struct {
  uint32_t count;        /* atomic counter value (may exceed 32) */
  uint32_t pad;
  double points[ORC_CONTACTS][4];
} orc_contacts;

/* ---- shared, precision independent ------------------------------------ */
uint32_t orc_hash_u32(uint32_t x);
float orc_hash1(uint32_t seed);
int orc_volume_derive(int resolution, int mips_req, int* mips_out, int* tile_res_out,
                      int* cell_off_out, int* total_cells_out, char* err, int errlen);
void orc_fluid_derive(const orc_fluid_opts* o, orc_fluid_k* k);
uint32_t orc_pack_tile(uint32_t x, uint32_t y, uint32_t z);
void orc_unpack_tile(uint32_t v, uint32_t* xyz);
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
uint32_t orc_enc_f32(float v);
float orc_dec_f32(uint32_t e);
void orc_set_threads(int n);
int orc_get_threads(void);
void orc_stable_sort_keys(int n, const uint32_t* keys, uint32_t* perm_out);
/* TileGrid.dilate (tilegrid.ts:207-242): cells of mip m whose children in
 * [2c-1, 2c+2]^3 at mip m-1 have a non-zero count get count 1.  `counts` holds
 * every mip at the volume's cell offsets. */
void orc_tile_dilate(int res, int mips, int m, uint32_t* counts);
/* Drops contacts into erosion edits: every contact whose speed exceeds
 * min_speed becomes a cut sphere of `radius` at the contact (the policy the
 * reference leaves to the caller, game/fluid.ts:86-95).  Returns the count. */
int orc_contacts_to_edits(uint32_t raw_count, const float* points4, double min_speed, double radius,
                          double blend, int max_edits, orc_brush* out);

/* ClaySolver (lib/src/sim/pbd.ts:12-40, 78-424): SolveParams in f64, rounded to the
 * working precision once; `h` is dt / substeps.  Bodies: first, count, flags per
 * body (3 u32) and stiffness, plasticity (2 f32, as the Body record holds them). */
typedef struct {
  double gravity[3];
  double h;
  double radius, friction, damping, max_speed;
  double push[3];
  double kick[3];
  uint32_t push_body;
  int32_t rot_iters;
} orc_pbd_params;

/* SurfaceExtractor (lib/src/sim/extract.ts:85-484): the region (f64, rounded
 * once), the body and slice, and the quad capacity ceil(1.5 capacity). */
typedef struct {
  double origin[3];
  double cell;             /* size / resolution */
  double vel[3];
  int32_t preserve;
  int32_t res;
  uint32_t body, base, capacity, quad_capacity;
} orc_extract_params;

#define ORC_DECLARE(T, S)                                                              \
  void orc_keys_##S(const orc_fluid_k* k, int n, const T* px, const T* py, const T* pz, \
                    uint32_t* keys);                                                   \
  void orc_substep_##S(const orc_fluid_k* k, const orc_collider* col, int n,           \
                       orc_state* st, orc_contacts* contacts, uint32_t* keys_out,      \
                       uint32_t* perm_out);                                            \
  int orc_neighbors_##S(const orc_fluid_k* k, int n, const T* px, const T* py,         \
                        const T* pz, int i, uint32_t* out, int cap);                   \
  void orc_emit_##S(int cap, int base, int count, uint32_t seed, const double* pos3,    \
                    double radius, const double* vel3, orc_state* st);                 \
  void orc_density_##S(const orc_fluid_k* k, int n, orc_state* st);                    \
  void orc_forces_##S(const orc_fluid_k* k, int n, orc_state* st);                     \
  void orc_integrate_##S(const orc_fluid_k* k, const orc_collider* col, int n,         \
                         orc_state* st, orc_contacts* contacts);                       \
  void orc_collider_eval_##S(const orc_collider* col, int n, const T* p3, T* dist,     \
                             T* nrm3);                                                 \
  void orc_splat_bake_##S(const orc_splat_opts* o, int n, const T* px, const T* py,    \
                          const T* pz, const T* nx, const T* ny, const T* nz,          \
                          int live_first, int live_count, uint32_t* grid);             \
  void orc_splat_resolve_##S(const orc_splat_opts* o, const uint32_t* grid,            \
                             uint16_t* texels);                                        \
  void orc_bake_brushes_##S(int res, int mips, double world_size, const double* origin,\
                            double band, const orc_brush* b, int nb, uint16_t* texels);   \
  void orc_edit_bin_##S(int res, double world_size, const double* origin, double band, \
                        int capacity, const orc_brush* e, int ne, uint32_t* counts0,   \
                        uint32_t* ids0);                                               \
  void orc_edit_apply_##S(int res, double world_size, const double* origin,            \
                          double band, int capacity, const orc_brush* e,               \
                          const uint32_t* counts0, const uint32_t* ids0,               \
                          uint16_t* dist, uint16_t* mat);                              \
  void orc_mip_refresh_##S(int res, int mips, int m, double band, double safety,       \
                           const uint32_t* counts, uint16_t* dist, uint16_t* mat);     \
  void orc_edit_flush_##S(int res, int mips, double world_size, const double* origin,  \
                          double band, double safety, int capacity, const orc_brush* e,\
                          int ne, uint16_t* dist, uint16_t* mat, uint32_t* counts);    \
  void orc_body_centroids_##S(int nb, const uint32_t* first_count, const float* px,    \
                              const float* py, const float* pz, const float* rx,       \
                              const float* ry, const float* rz, double* out6);         \
  /* One ClaySolver.step: the kick when `kicking`, then `substeps` x (predict, match, \
   * apply, finalize).  planes: 15 * capacity values, SoA pos, prev, vel, velPrev,    \
   * rest (x, y, z each); state: 10 per body (com, restCom, rot xyzw). */             \
  void orc_pbd_step_##S(const orc_pbd_params* k, const orc_collider* col, int capacity,\
                        int max_bodies, int nbodies, const uint32_t* body3,             \
                        const float* body2, const uint32_t* owner, T* planes, T* state, \
                        int kicking, int substeps);                                     \
  void orc_sin_cos_##S(T x, T* s, T* c);                                               \
  /* One SurfaceExtractor.extract: reset (motion snapshot when `motion`), cells,      \
   * quads, finish.  aux: 4 * capacity values (restNormal xyz, material); ids: res^3; \
   * indices: 6 * quad_capacity; counts: particles and quads found; draw: 4 u32.      \
   * Slots are taken in ascending cell index, quads in ascending lattice point then   \
   * edge axis. */                                                                    \
  /* TracedField.sample (what 0) or field (what 1) at `mip`, and entry per ray      \
   * (ro, rd, tMin, radius, aperture): field.ts:135-177, 208-216, 255-296. */          \
  void orc_field_query_##S(const orc_collider* col, int what, int mip, int n,           \
                           const T* p3, T* out2);                                      \
  void orc_field_entry_##S(const orc_collider* col, int n, const T* rays9, T* out);     \
  /* makeTracer(field, o).sweep per ray (march.ts:97-218); rays10: ro xyz, rd xyz,    \
   * tMin, tMax, radius, aperture.  Out: (t, dist) and (hit, steps) per ray. */        \
  void orc_field_sweep_##S(const orc_collider* col, const orc_tracer_opts* o, int n,    \
                           const T* rays10, T* t_dist2, uint32_t* hit_steps2);         \
  void orc_extract_##S(const orc_extract_params* k, const orc_collider* col,          \
                       int capacity, uint32_t* body3, T* planes, T* aux,               \
                       uint32_t* owner, T* state, T* motion, uint32_t* ids,            \
                       uint32_t* indices, uint32_t* counts, uint32_t* draw);

ORC_DECLARE(float, f32)
ORC_DECLARE(double, f64)

#ifdef __cplusplus
}
#endif
#endif
