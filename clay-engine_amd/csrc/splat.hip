// splat.hip -- SplatField bake for gfx950 (lib/src/sim/splat.ts:186-304).
//
// Two schedules produce the same u32 codes, voxel for voxel:
//
// PLAIN (the reference's schedule): clear a dense per-mip u32 grid to 65535, one
//   thread per particle walks its voxel box and does a global atomicMin of
//   e = u32((clamp(sd / bandWorld, -1, 1) + 1) * 32767.5), then resolve every
//   voxel to f16.
//
// TILED:
//   1. k_bin<count>, k_brick_sum/scan/cursor, k_bin<fill>: particles are binned to 16^3-voxel
//      bricks of every mip (a particle lands in each brick its voxel box overlaps;
//      the counters are hit once per wave and distinct brick).  The scan also
//      cuts every brick's list into chunks (2048 particles unless tuned): a chunk
//      is one work item, so a coarse-mip brick holding most of the cloud is spread
//      over many workgroups instead of serialising on one.
//   2. k_brick_clear: the grid of every brick that received particles is reset.
//   3. k_brick_splat: a persistent grid walks the work items.  Each item keeps a
//      16 KB LDS copy of its brick and, for every (particle, voxel) pair inside the
//      reference's voxel box, takes ds_min_u32 of the f32 bits of the squared
//      distance s = (dx*dx + dy*dy) + dz*dz.  Then it merges the brick into the
//      grid (plain stores when the brick is one chunk, atomicMin otherwise).
//   4. k_brick_resolve: per brick, code = enc(sqrt(s_min)) once per voxel, the exact
//      live-tile mask, and f16 texels for tiles that are live now or were live
//      after the previous bake.  Every other tile already reads the saturated +1.
//
// Why step 3 may keep only min s: the code is a non-decreasing function of s in
// IEEE arithmetic (sqrt, subtracting the radius, dividing by the band, clamping,
// adding, scaling and truncating are each monotone), so min over particles of
// code(s_i) == code(min over particles of s_i), bit for bit.  Non-negative f32
// values order like their bit patterns, so the min is an integer atomicMin.
// Clouds with normals (sd = max(dot(d, n), |d| - R)) are not a function of s
// alone; they take the min of the code itself, as the reference does.
//
// Voxels whose centre is farther than `reach` from the particle are skipped row
// by row; they fail sd < bandWorld, so skipping them cannot change a code.
//
// A fluid's surface bake also gets each particle's interior radius (from the
// density pass over the same positions): when a particle's neighbours surround it
// in all 26 lattice directions, every voxel centre beyond that radius has another
// particle strictly nearer, so the particle cannot hold that voxel's minimum and
// the walk stops at the radius.  Only particles at the fluid's surface walk their
// whole reach.  The codes are those of the full walk.
//
// A batch bakes one volume instance per scene in the same launches: brick ids run
// scene-major (scene * bricks-per-scene + brick), a particle's scene comes from a
// per-particle table, and each scene has its own origin record, code grid block,
// texel planes and tile mask.  A one-volume splat is the batch of one scene.
#include <mutex>

#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kBrick = 16;                  // voxels per brick edge
constexpr int kBrickVox = kBrick * kBrick * kBrick;
constexpr int kThreads = 256;
constexpr int kSlices = 32;                 // binning counters per brick
constexpr uint32_t kSInit = 0xFFFFFFFFu;    // "no particle" for the squared distance
constexpr int kHalo = 5;                    // owner schedule: voxels staged around a brick
constexpr int kExt = kBrick + 2 * kHalo;    // staged edge (26)
constexpr int kExtVox = kExt * kExt * kExt;
static_assert(kExtVox < (1 << 15) && kExt <= 26, "the owner merge's multiply-high division needs these bounds");
constexpr int kPart = 2048;                 // owner schedule: particles partitioned per LDS pass
// Owner schedule workgroup: the staged region takes ~78 KB of LDS, so two workgroups
// fit a CU; 512 threads each keep 4 waves per SIMD in flight (1024-thread
// workgroups leave most lanes idle on the small bricks of following volumes).
constexpr int kOwnerBits = 9;
constexpr int kOwnerThreads = 1 << kOwnerBits;

struct SplatMipK {
  int32_t r;          // res >> m
  int32_t nb;         // bricks per axis (tiled)
  uint32_t off;       // offset of this mip in the code grid
  uint32_t voff;      // offset of this mip in the volume planes
  uint32_t boff;      // first brick of this mip (tiled)
  uint32_t toff;      // first 8^3 tile of this mip (tiled)
  float voxel, bw, reach, inv, radius;
  float cull2;        // (1.001 reach)^2: rows beyond it hold no voxel with sd < bw
};

struct CloudK {
  const float* pos[3];
  const float* nrm[3];
  const uint32_t* order;     // optional visiting order (any permutation gives the same codes)
  const uint32_t* scene_of;  // batch: scene of each particle id (else every particle is scene 0)
  const float* cover;        // optional, per id: squared radius beyond which the particle holds
                             // no voxel's minimum (no normals); the walk stops there
  uint32_t first, end;       // live particles: ids in [first, end)
  uint32_t count;            // end - first
};

struct MipsK {
  int32_t mips;
  float inv_enc;          // f32(1 / 32767.5)
  uint16_t sat_half;      // f16(+1.0)
  uint32_t chunk;         // particles per work item
  uint32_t bps;           // bricks per scene (all mips)
  uint32_t gps;           // code-grid cells per scene
  uint32_t vps;           // texels per scene (all mips)
  uint32_t tps;           // 8^3 tiles per scene (all mips)
  uint32_t xcd;           // XCD-contiguous work-item numbering
  uint32_t owner;         // owner schedule: one list entry per particle and mip
  SplatMipK m[CLAY_MAX_MIPS];
};

// Device-side results of the brick scan.
struct ScanOut {
  uint32_t entries;  // total brick-list entries
  uint32_t items;    // total work items (chunks)
};

__device__ __forceinline__ uint32_t particle_at(const CloudK& c, uint32_t k) {
  return c.order ? c.order[k] : c.first + k;
}

__device__ __forceinline__ bool live_id(const CloudK& c, uint32_t i) {
  return i >= c.first && i < c.end;
}

// The reference's voxel box (splat.ts:234-243), clipped to the grid.  Returns false
// when it is empty.
__device__ __forceinline__ bool voxel_box(const SplatMipK& m, f3 p, int32_t lo[3], int32_t hi[3]) {
  lo[0] = sat_i32(floorf((p.x - m.reach) * m.inv));
  lo[1] = sat_i32(floorf((p.y - m.reach) * m.inv));
  lo[2] = sat_i32(floorf((p.z - m.reach) * m.inv));
  hi[0] = sat_i32(floorf((p.x + m.reach) * m.inv));
  hi[1] = sat_i32(floorf((p.y + m.reach) * m.inv));
  hi[2] = sat_i32(floorf((p.z + m.reach) * m.inv));
  const int32_t last = m.r - 1;
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    lo[a] = lo[a] > 0 ? lo[a] : 0;
    hi[a] = hi[a] < last ? hi[a] : last;
    ok = ok && lo[a] <= hi[a];
  }
  return ok;
}

// Voxels [a, b] of one axis whose centres (i + 0.5) * voxel lie within s of c,
// padded by 1/64 voxel each side and clipped to [lo, hi].  s is the chord of the
// cull ball, whose radius is 1.001 reach: a voxel left out lies beyond the reach by
// more than the pad, far more than the f32 error of this test (~1e-4 voxel at
// 2048 voxels), so its code is saturated and leaving it out changes nothing.
// The clamps run in f32 (lo and hi are exact there) and keep a in [lo, hi + 1] and
// b in [lo - 1, hi], so the conversions are exact and branch-free; a NaN bound
// gives the whole clipped span (maxnum/minnum return the number).
__device__ __forceinline__ void axis_span(float c, float s, float inv, int32_t lo, int32_t hi, int32_t& a,
                                          int32_t& b) {
  const float flo = (float)lo, fhi = (float)hi;
  a = (int32_t)fminf(fmaxf(ceilf((c - s) * inv - 0.515625f), flo), fhi + 1.0f);
  b = (int32_t)fmaxf(fminf(floorf((c + s) * inv - 0.484375f), fhi), flo - 1.0f);
}

// enc of a signed distance (splat.ts:246-247), or 65535 outside the band.
__device__ __forceinline__ uint32_t encode_sd(const SplatMipK& m, float sd) {
  if (sd < m.bw) {
    const float q = (fclamp_s(sd / m.bw, -1.0f, 1.0f) + 1.0f) * kEncScale;
    return (uint32_t)q;
  }
  return kEncSat;
}

// Code of voxel (x,y,z) for a particle at p (relative to the volume origin).
// splat.ts:244-250 (and :170-184 with normals).
__device__ __forceinline__ uint32_t voxel_code(const SplatMipK& m, f3 p, bool has_n, f3 n,
                                               int32_t x, int32_t y, int32_t z) {
  const f3 cc = mk(((float)x + 0.5f) * m.voxel, ((float)y + 0.5f) * m.voxel,
                   ((float)z + 0.5f) * m.voxel);
  const f3 d = cc - p;
  float sd = len3(d) - m.radius;
  if (has_n) sd = fmax_s(dot3(d, n), sd);
  return encode_sd(m, sd);
}

// Code of the nearest particle given the min squared distance's bits.
__device__ __forceinline__ uint32_t code_from_s(const SplatMipK& m, uint32_t sbits) {
  if (sbits == kSInit) return kEncSat;
  const float sd = sqrtf(__uint_as_float(sbits)) - m.radius;
  return encode_sd(m, sd);
}

__device__ __forceinline__ f3 load_rel(const CloudK& c, const VolDev* vol, uint32_t i) {
  return mk(c.pos[0][i] - vol->origin[0], c.pos[1][i] - vol->origin[1], c.pos[2][i] - vol->origin[2]);
}
__device__ __forceinline__ f3 load_nrm(const CloudK& c, uint32_t i) {
  return c.nrm[0] ? mk(c.nrm[0][i], c.nrm[1][i], c.nrm[2][i]) : mk(0.0f, 0.0f, 0.0f);
}

// One row of a particle's ball in LDS (the voxel loop of splat.ts:243-250): voxels
// rp..re (inclusive; none when re < rp) whose centres have x-coordinate fx, fx + 1,
// ... in voxels.  fx = x + 0.5 and its steps are exact (integers and halves below
// 2^23), so each dx is the value ((float)x + 0.5f) * voxel - p.x gives.  Voxels go
// in pairs, with one more when the count is odd.  Without normals the squared
// distance is min-merged (its f32 bits order as the value); with normals the code.
template <bool kNormals>
__device__ __forceinline__ void splat_row(const SplatMipK& m, f3 p, f3 nr, float dy, float dz, float dy2,
                                          float dz2, uint32_t* rp, uint32_t* re, float fx) {
  auto voxel = [&](uint32_t* q, float f) {
    const float dx = f * m.voxel - p.x;
    if (kNormals) {
      const float len = sqrtf(dx * dx + dy2 + dz2);
      float sd = len - m.radius;
      sd = fmax_s(dx * nr.x + dy * nr.y + dz * nr.z, sd);
      const uint32_t code = encode_sd(m, sd);
      if (code != kEncSat) atomicMin(q, code);
    } else {
      atomicMin(q, __float_as_uint(dx * dx + dy2 + dz2));
    }
  };
  for (; rp < re; rp += 2, fx += 2.0f) {
    voxel(rp, fx);
    voxel(rp + 1, fx + 1.0f);
  }
  if (rp == re) voxel(rp, fx);
}

// ---------------------------------------------------------------------------
// PLAIN
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
    k_splat_plain(SplatMipK m, CloudK c, const VolDev* __restrict__ vol, uint32_t* __restrict__ grid) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= c.count) return;
  const uint32_t i = particle_at(c, k);
  if (!live_id(c, i)) return;
  const f3 p = load_rel(c, vol, i);
  const bool has_n = c.nrm[0] != nullptr;
  const f3 n = load_nrm(c, i);
  int32_t lo[3], hi[3];
  if (!voxel_box(m, p, lo, hi)) return;
  uint32_t* g = grid + m.off;
  for (int32_t z = lo[2]; z <= hi[2]; z++)
    for (int32_t y = lo[1]; y <= hi[1]; y++)
      for (int32_t x = lo[0]; x <= hi[0]; x++) {
        const uint32_t e = voxel_code(m, p, has_n, n, x, y, z);
        if (e != kEncSat) atomicMin(&g[(uint32_t)x + (uint32_t)m.r * ((uint32_t)y + (uint32_t)m.r * (uint32_t)z)], e);
      }
}

__global__ void __launch_bounds__(kThreads)
    k_resolve_plain(SplatMipK m, float inv_enc, const uint32_t* __restrict__ grid,
                    uint16_t* __restrict__ dist) {
  const uint32_t cells = (uint32_t)m.r * (uint32_t)m.r * (uint32_t)m.r;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < cells; i += gridDim.x * kThreads)
    dist[m.voff + i] = f2h((float)grid[m.off + i] * inv_enc - 1.0f);
}

// ---------------------------------------------------------------------------
// TILED
// ---------------------------------------------------------------------------
// Per-wave aggregated atomicAdd: lanes with the same counter index share one
// atomic; each active lane gets old value + its rank among them.
__device__ __forceinline__ uint32_t wave_add(uint32_t* counters, uint32_t idx, bool active) {
  uint64_t pending = __ballot(active);
  const int lane = __lane_id();
  uint32_t result = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lead_idx = __shfl(idx, leader);
    const uint64_t match = __ballot(active && idx == lead_idx) & pending;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&counters[lead_idx], (uint32_t)__popcll(match));
    base = __shfl(base, leader);
    if ((match >> lane) & 1ull) result = base + (uint32_t)__popcll(match & __lanemask_lt());
    pending &= ~match;
  }
  return result;
}

// Brick range of a particle at mip `m`; returns the number of bricks it overlaps.
__device__ __forceinline__ int brick_range(const SplatMipK& m, f3 p, int32_t blo[3], int32_t bext[3]) {
  int32_t lo[3], hi[3];
  if (!voxel_box(m, p, lo, hi)) return 0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    blo[a] = lo[a] / kBrick;
    bext[a] = hi[a] / kBrick - blo[a] + 1;
  }
  return bext[0] * bext[1] * bext[2];
}

// Binning pass 1 (count, kFill = false) and pass 3 (fill, kFill = true).
template <bool kFill>
__global__ void __launch_bounds__(kThreads)
    k_bin(MipsK mk_, CloudK c, const VolDev* __restrict__ vol, uint32_t* __restrict__ counters,
          uint32_t* __restrict__ list) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  bool valid = k < c.count;
  uint32_t i = 0;
  if (valid) {
    i = particle_at(c, k);
    valid = live_id(c, i);
  }
  uint32_t scene = 0;
  if (valid && c.scene_of) scene = c.scene_of[i];
  f3 p = mk(0.0f, 0.0f, 0.0f);
  if (valid) p = load_rel(c, vol + scene, i);
  const uint32_t sb = scene * mk_.bps;
  for (int mi = 0; mi < mk_.mips; mi++) {
    const SplatMipK& m = mk_.m[mi];
    int32_t blo[3] = {0, 0, 0}, bext[3] = {1, 1, 1};
    int nb3 = valid ? brick_range(m, p, blo, bext) : 0;
    if (mk_.owner && nb3 > 0) {
      // The brick holding the particle's own voxel (clamped to the level) owns it.
      const float pc[3] = {p.x, p.y, p.z};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        int32_t v = sat_i32(floorf(pc[a] * m.inv));
        v = v < 0 ? 0 : (v > m.r - 1 ? m.r - 1 : v);
        blo[a] = v / kBrick;
        bext[a] = 1;
      }
      nb3 = 1;
    }
    for (int q = 0;; q++) {
      const bool act = q < nb3;
      if (!__ballot(act)) break;
      uint32_t b = 0;
      if (act) {
        const int32_t bx = blo[0] + q % bext[0];
        const int32_t by = blo[1] + (q / bext[0]) % bext[1];
        const int32_t bz = blo[2] + q / (bext[0] * bext[1]);
        b = sb + m.boff + (uint32_t)bx + (uint32_t)m.nb * ((uint32_t)by + (uint32_t)m.nb * (uint32_t)bz);
      }
      // Each brick has kSlices counters; a workgroup uses slice blockIdx % kSlices,
      // which spreads a coarse-mip brick's tens of thousands of wave atomics over
      // kSlices addresses instead of serialising them on one.
      const uint32_t slot = wave_add(counters, b * kSlices + blockIdx.x % kSlices, act);
      if (kFill && act) list[slot] = i;
    }
  }
}

// Binning pass 2, in three launches: per brick, the total of its slice counters
// (-> count); over all bricks, exclusive scans of the totals (-> list starts) and of
// the per-brick chunk counts (-> first work item of each brick); per slice, its fill
// cursor.  A brick's kSlices counters are read and written by kSlices lanes at once.
static_assert(kSlices == 32, "a brick's slices are one half-wave");

__device__ __forceinline__ uint32_t half_wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Totals: two bricks per wave.
__global__ void __launch_bounds__(kThreads)
    k_brick_sum(const uint32_t* __restrict__ sub, uint32_t nbricks, uint32_t* __restrict__ count) {
  const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t b = g / kSlices;
  const uint32_t v = b < nbricks ? sub[g] : 0u;
  const uint32_t tot = half_wave_sum(v);
  if (b < nbricks && (g & (kSlices - 1)) == 0) count[b] = tot;
}

// Inclusive scan of v over the wave's 64 lanes.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = __shfl_up(v, d);
    if (lane >= d) v += u;
  }
  return v;
}

// The scans over bricks: one workgroup of four waves, each thread a run of
// kScanRun consecutive bricks per round, run totals joined by wave scans and a scan
// of the four wave totals.  The binning chain that ends in this launch runs on a
// second stream beside the frame-end density pass, which holds every CU: a
// four-wave workgroup is placed as soon as one of that pass's workgroups retires,
// where a 1024-thread one waited for a CU with 16 free wave slots (5 us alone, about
// 0.2 ms beside the density pass).
constexpr int kScanThreads = 256, kScanRun = 8;
__global__ void __launch_bounds__(kScanThreads)
    k_brick_scan(uint32_t nbricks, uint32_t chunk, const uint32_t* __restrict__ count,
                 uint32_t* __restrict__ start, uint32_t* __restrict__ item_start, ScanOut* __restrict__ out) {
  __shared__ uint32_t wa[kScanThreads / 64], wb[kScanThreads / 64];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  uint32_t carry_a = 0, carry_b = 0;  // the same in every thread
  for (uint32_t c0 = 0; c0 < nbricks; c0 += kScanThreads * kScanRun) {
    const uint32_t i0 = c0 + (uint32_t)t * kScanRun;
    uint32_t va[kScanRun], vb[kScanRun], sa = 0, sb = 0;
#pragma unroll
    for (int q = 0; q < kScanRun; q++) {
      va[q] = i0 + q < nbricks ? count[i0 + q] : 0u;
      vb[q] = (va[q] + chunk - 1) / chunk;
      sa += va[q];
      sb += vb[q];
    }
    const uint32_t ia = wave_incl_scan(sa, lane), ib = wave_incl_scan(sb, lane);
    if (lane == 63) {
      wa[wave] = ia;
      wb[wave] = ib;
    }
    __syncthreads();
    uint32_t pa = carry_a + ia - sa, pb = carry_b + ib - sb;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; w++) {
      if (w < wave) {
        pa += wa[w];
        pb += wb[w];
      }
      carry_a += wa[w];
      carry_b += wb[w];
    }
#pragma unroll
    for (int q = 0; q < kScanRun; q++) {
      if (i0 + q < nbricks) {
        start[i0 + q] = pa;
        item_start[i0 + q] = pb;
      }
      pa += va[q];
      pb += vb[q];
    }
    __syncthreads();  // wa and wb are written again by the next round
  }
  if (t == 0) {
    out->entries = carry_a;
    out->items = carry_b;
  }
}

// Fill cursors: each slice starts after its brick's earlier slices.
__global__ void __launch_bounds__(kThreads)
    k_brick_cursor(const uint32_t* __restrict__ sub, uint32_t nbricks, const uint32_t* __restrict__ start,
                   uint32_t* __restrict__ cursor) {
  const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t b = g / kSlices;
  const int l = (int)(g & (kSlices - 1));
  const uint32_t v = b < nbricks ? sub[g] : 0u;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < kSlices; d <<= 1) {
    const uint32_t u = __shfl_up(x, d, kSlices);
    if (l >= d) x += u;
  }
  if (b < nbricks) cursor[g] = start[b] + x - v;
}

struct BrickPos {
  int32_t mi, ox, oy, oz, bx, by, bz;
};

__device__ __forceinline__ BrickPos brick_pos(const MipsK& mk_, uint32_t b) {
  BrickPos q;
  int mi = 0;
  while (mi + 1 < mk_.mips && b >= mk_.m[mi + 1].boff) mi++;
  const SplatMipK& m = mk_.m[mi];
  const uint32_t lb = b - m.boff;
  q.mi = mi;
  q.bx = (int32_t)(lb % (uint32_t)m.nb);
  q.by = (int32_t)((lb / (uint32_t)m.nb) % (uint32_t)m.nb);
  q.bz = (int32_t)(lb / ((uint32_t)m.nb * (uint32_t)m.nb));
  q.ox = q.bx * kBrick;
  q.oy = q.by * kBrick;
  q.oz = q.bz * kBrick;
  return q;
}

// Whether any particle can have written brick b: under the owner schedule a
// brick also receives the halos of its 26 neighbours' particles.
__device__ __forceinline__ bool brick_touched(const MipsK& mk_, const uint32_t* __restrict__ count, uint32_t b) {
  if (count[b] != 0) return true;
  if (!mk_.owner) return false;
  const uint32_t scene = b / mk_.bps;
  const BrickPos q = brick_pos(mk_, b - scene * mk_.bps);
  const SplatMipK& m = mk_.m[q.mi];
  const uint32_t base = scene * mk_.bps + m.boff;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int32_t x = q.bx + dx, y = q.by + dy, z = q.bz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= m.nb || y >= m.nb || z >= m.nb) continue;
        if (count[base + (uint32_t)x + (uint32_t)m.nb * ((uint32_t)y + (uint32_t)m.nb * (uint32_t)z)] != 0)
          return true;
      }
  return false;
}

// Step 2: reset the grid of bricks that received particles.
template <bool kNormals>
__global__ void __launch_bounds__(kThreads)
    k_brick_clear(MipsK mk_, const uint32_t* __restrict__ count, uint32_t* __restrict__ grid) {
  const uint32_t b = blockIdx.x;
  if (!brick_touched(mk_, count, b)) return;
  const uint32_t scene = b / mk_.bps;
  const BrickPos q = brick_pos(mk_, b - scene * mk_.bps);
  const SplatMipK& m = mk_.m[q.mi];
  const uint32_t init = kNormals ? kEncSat : kSInit;
  uint32_t* g = grid + (size_t)scene * mk_.gps + m.off;
  for (int v = threadIdx.x; v < kBrickVox; v += kThreads) {
    const int32_t x = q.ox + (v & 15), y = q.oy + ((v >> 4) & 15), z = q.oz + (v >> 8);
    if (x >= m.r || y >= m.r || z >= m.r) continue;
    g[(uint32_t)x + (uint32_t)m.r * ((uint32_t)y + (uint32_t)m.r * (uint32_t)z)] = init;
  }
}

// Step 3: persistent walk over (brick, chunk) work items.
template <bool kNormals>
__global__ void __launch_bounds__(kThreads)
    k_brick_splat(MipsK mk_, CloudK c, const VolDev* __restrict__ vol,
                  const uint32_t* __restrict__ start, const uint32_t* __restrict__ count,
                  const uint32_t* __restrict__ item_start, const ScanOut* __restrict__ scan,
                  uint32_t nbricks, const uint32_t* __restrict__ list, uint32_t* __restrict__ grid) {
  __shared__ uint32_t lds[kBrickVox];
  __shared__ uint32_t sh_brick;
  const int t = threadIdx.x;
  const uint32_t init = kNormals ? kEncSat : kSInit;
  // Lane order inside a pass: the 8-bit reversal of the thread index, so that
  // neighbouring lanes take list entries 128 apart (spatially apart) and their
  // ds_min_u32 rarely meet on one address.
  const uint32_t rev = __brev((uint32_t)t) >> 24;
  const uint32_t items = scan->items;
  for (uint32_t item = mk_.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x; item < items; item += gridDim.x) {
    if (t == 0) {
      // Last brick whose first item is <= item (a brick with no items shares its
      // first item with the next brick, so the last one always has items).
      uint32_t lo = 0, hi = nbricks;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (item_start[mid] <= item) lo = mid; else hi = mid;
      }
      sh_brick = lo;
    }
    for (int v = t; v < kBrickVox; v += kThreads) lds[v] = init;
    __syncthreads();
    const uint32_t b = sh_brick;
    const uint32_t scene = b / mk_.bps;
    const BrickPos q = brick_pos(mk_, b - scene * mk_.bps);
    const SplatMipK& m = mk_.m[q.mi];
    const VolDev* sv = vol + scene;
    uint32_t* sgrid = grid + (size_t)scene * mk_.gps;
    const uint32_t n_b = count[b];
    const uint32_t chunk = item - item_start[b];
    const uint32_t e_lo = chunk * mk_.chunk;
    const uint32_t n_e = (n_b - e_lo) < mk_.chunk ? (n_b - e_lo) : mk_.chunk;
    const uint32_t base = start[b] + e_lo;
    const int32_t ox = q.ox, oy = q.oy, oz = q.oz;
    for (uint32_t pass = 0; pass * kThreads < n_e; pass++) {
      const uint32_t e = pass * kThreads + rev;
      if (e >= n_e) continue;
      const uint32_t i = list[base + e];
      const f3 p = load_rel(c, sv, i);
      const f3 nr = kNormals ? load_nrm(c, i) : mk(0.0f, 0.0f, 0.0f);
      const float cull2 = (!kNormals && c.cover) ? fminf(m.cull2, c.cover[i]) : m.cull2;
      int32_t lo[3], hi[3];
      voxel_box(m, p, lo, hi);  // non-empty: the particle was binned here
      const int32_t x0 = lo[0] > ox ? lo[0] : ox, x1 = hi[0] < ox + kBrick - 1 ? hi[0] : ox + kBrick - 1;
      const int32_t y0 = lo[1] > oy ? lo[1] : oy, y1 = hi[1] < oy + kBrick - 1 ? hi[1] : oy + kBrick - 1;
      const int32_t z0 = lo[2] > oz ? lo[2] : oz, z1 = hi[2] < oz + kBrick - 1 ? hi[2] : oz + kBrick - 1;
      // Planes, rows and row spans of the cull ball (axis_span).  The hardware
      // square root (within 1 ulp) is enough: the spans carry a margin.
      int32_t za, zb;
      axis_span(p.z, __builtin_amdgcn_sqrtf(cull2), m.inv, z0, z1, za, zb);
      for (int32_t z = za; z <= zb; z++) {
        const float dz = ((float)z + 0.5f) * m.voxel - p.z;
        const float dz2 = dz * dz;
        const float rz = cull2 - dz2;
        if (rz < 0.0f) continue;
        int32_t ya, yb;
        axis_span(p.y, __builtin_amdgcn_sqrtf(rz), m.inv, y0, y1, ya, yb);
        for (int32_t y = ya; y <= yb; y++) {
          const float dy = ((float)y + 0.5f) * m.voxel - p.y;
          const float dy2 = dy * dy;
          const float rem = rz - dy2;
          if (rem < 0.0f) continue;
          int32_t xa, xb;
          axis_span(p.x, __builtin_amdgcn_sqrtf(rem), m.inv, x0, x1, xa, xb);
          uint32_t* row = lds + (y - oy) * kBrick + (z - oz) * kBrick * kBrick - ox;
          splat_row<kNormals>(m, p, nr, dy, dz, dy2, dz2, row + xa, row + xb, (float)xa + 0.5f);
        }
      }
    }
    __syncthreads();
    // Merge the brick into the grid.
    const bool single = n_b <= mk_.chunk;
    for (int v = t; v < kBrickVox; v += kThreads) {
      const int32_t x = ox + (v & 15), y = oy + ((v >> 4) & 15), z = oz + (v >> 8);
      if (x >= m.r || y >= m.r || z >= m.r) continue;
      const uint32_t val = lds[v];
      uint32_t* g = sgrid + m.off + (uint32_t)x + (uint32_t)m.r * ((uint32_t)y + (uint32_t)m.r * (uint32_t)z);
      if (single) {
        *g = val;
      } else if (val != init) {
        atomicMin(g, val);
      }
    }
    __syncthreads();
  }
}

// Step 3, owner schedule: a work item is (brick, chunk of the particles it owns).
// The workgroup stages the brick plus a kHalo-voxel shell (clipped to the level)
// in LDS, so every particle's whole voxel box lies inside it and no lane's box is
// clipped by the brick: all lanes walk boxes of the same size, which keeps the
// wave's lanes busy.  The staged region is merged into the code grid with
// atomicMin, because neighbouring bricks' shells overlap it.
// kCount: also tallies the atomics the walk issues (ds_min_u32 into LDS, one per
// walked voxel) and the merge issues (global atomicMin, one per voxel below the
// initial value) into counts[0] and counts[1]; the codes are the same.
template <bool kNormals, bool kCount = false>
__global__ void __launch_bounds__(kOwnerThreads)
    k_owner_splat(MipsK mk_, CloudK c, const VolDev* __restrict__ vol, const uint32_t* __restrict__ start,
                  const uint32_t* __restrict__ count, const uint32_t* __restrict__ item_start,
                  const ScanOut* __restrict__ scan, uint32_t nbricks, const uint32_t* __restrict__ list,
                  uint32_t* __restrict__ grid, unsigned long long* __restrict__ counts = nullptr) {
  unsigned long long n_lds = 0, n_glob = 0;
  // The staged region, sized at launch for the largest brick-plus-shell of these
  // levels (owner_stage_voxels): at most kExtVox.
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t part[kPart];    // interior-radius schedule: the sub-chunk, surface particles first
  __shared__ uint32_t part_n[2];      // surface and interior counts
  __shared__ uint32_t sh_brick;
  const int t = threadIdx.x;
  const uint32_t init = kNormals ? kEncSat : kSInit;
  const uint32_t items = scan->items;
  for (uint32_t item = mk_.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x; item < items; item += gridDim.x) {
    if (t == 0) {
      uint32_t lo = 0, hi = nbricks;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (item_start[mid] <= item) lo = mid; else hi = mid;
      }
      sh_brick = lo;
    }
    __syncthreads();
    const uint32_t b = sh_brick;
    const uint32_t scene = b / mk_.bps;
    const BrickPos q = brick_pos(mk_, b - scene * mk_.bps);
    const SplatMipK& m = mk_.m[q.mi];
    const VolDev* sv = vol + scene;
    uint32_t* sgrid = grid + (size_t)scene * mk_.gps + m.off;
    const int32_t ex0 = q.ox - kHalo > 0 ? q.ox - kHalo : 0, ey0 = q.oy - kHalo > 0 ? q.oy - kHalo : 0,
                  ez0 = q.oz - kHalo > 0 ? q.oz - kHalo : 0;
    const int32_t ex1 = q.ox + kBrick + kHalo < m.r ? q.ox + kBrick + kHalo : m.r;
    const int32_t ey1 = q.oy + kBrick + kHalo < m.r ? q.oy + kBrick + kHalo : m.r;
    const int32_t ez1 = q.oz + kBrick + kHalo < m.r ? q.oz + kBrick + kHalo : m.r;
    const int32_t nx = ex1 - ex0, ny = ey1 - ey0, nz = ez1 - ez0;
    const int32_t nvox = nx * ny * nz;
    for (int v = t; v < nvox; v += kOwnerThreads) lds[v] = init;
    __syncthreads();
    const uint32_t n_b = count[b];
    const uint32_t chunk = item - item_start[b];
    const uint32_t e_lo = chunk * mk_.chunk;
    const uint32_t n_e = (n_b - e_lo) < mk_.chunk ? (n_b - e_lo) : mk_.chunk;
    const uint32_t base = start[b] + e_lo;
    auto walk = [&](uint32_t i) {
      const f3 p = load_rel(c, sv, i);
      const f3 nr = kNormals ? load_nrm(c, i) : mk(0.0f, 0.0f, 0.0f);
      const float cull2 = (!kNormals && c.cover) ? fminf(m.cull2, c.cover[i]) : m.cull2;
      int32_t lo[3], hi[3];
      voxel_box(m, p, lo, hi);  // non-empty: the particle was binned
      // Inside the staged region by construction (reach <= kHalo voxels at every
      // level); clipping to it keeps every LDS access in bounds regardless.
      const int32_t x0 = lo[0] > ex0 ? lo[0] : ex0, x1 = hi[0] < ex1 - 1 ? hi[0] : ex1 - 1;
      const int32_t y0 = lo[1] > ey0 ? lo[1] : ey0, y1 = hi[1] < ey1 - 1 ? hi[1] : ey1 - 1;
      const int32_t z0 = lo[2] > ez0 ? lo[2] : ez0, z1 = hi[2] < ez1 - 1 ? hi[2] : ez1 - 1;
      int32_t za, zb;
      axis_span(p.z, __builtin_amdgcn_sqrtf(cull2), m.inv, z0, z1, za, zb);
      for (int32_t z = za; z <= zb; z++) {
        const float dz = ((float)z + 0.5f) * m.voxel - p.z;
        const float dz2 = dz * dz;
        const float rz = cull2 - dz2;
        if (rz < 0.0f) continue;
        int32_t ya, yb;
        axis_span(p.y, __builtin_amdgcn_sqrtf(rz), m.inv, y0, y1, ya, yb);
        for (int32_t y = ya; y <= yb; y++) {
          const float dy = ((float)y + 0.5f) * m.voxel - p.y;
          const float dy2 = dy * dy;
          const float rem = rz - dy2;
          if (rem < 0.0f) continue;
          int32_t xa, xb;
          axis_span(p.x, __builtin_amdgcn_sqrtf(rem), m.inv, x0, x1, xa, xb);
          uint32_t* row = lds + (y - ey0) * nx + (z - ez0) * nx * ny - ex0;
          if (kCount && xb >= xa) n_lds += (unsigned long long)(xb - xa + 1);
          splat_row<kNormals>(m, p, nr, dy, dz, dy2, dz2, row + xa, row + xb, (float)xa + 0.5f);
        }
      }
    };
    // Entries [0, count) of a list, a pass of up to kOwnerThreads at a time: a pass
    // of m entries goes to the fewest waves that hold it (W = ceil(m / 64)), lane l
    // of wave w taking entry w + W rev6(l) (rev6: the 6-bit reversal).  The waves'
    // lanes are full however short the list, and neighbouring lanes hold entries
    // far apart (spatially apart in the binned order, so their ds_min_u32 rarely
    // meet on one address); waves without work skip the pass.
    auto packed = [&](uint32_t count, auto&& visit) {
      const uint32_t w = (uint32_t)t >> 6, l = __brev((uint32_t)t & 63u) >> 26;
      for (uint32_t e0 = 0; e0 < count; e0 += kOwnerThreads) {
        const uint32_t m = count - e0 < (uint32_t)kOwnerThreads ? count - e0 : (uint32_t)kOwnerThreads;
        const uint32_t W = (m + 63u) >> 6;
        const uint32_t e = w + l * W;
        if (w < W && e < m) visit(e0 + e);
      }
    };
    if (!kNormals && c.cover) {
      // A wave takes as long as its longest walk, so particles that walk their whole
      // reach (the fluid's surface) and interior ones (a voxel or so) go to different
      // waves: each sub-chunk is partitioned in LDS, surface particles first, and
      // each part is walked packed into the fewest waves, so a wave's lanes hold
      // particles of one part.
      for (uint32_t sub = 0; sub < n_e; sub += kPart) {
        const uint32_t mm = n_e - sub < (uint32_t)kPart ? n_e - sub : (uint32_t)kPart;
        if (t < 2) part_n[t] = 0;
        __syncthreads();
        for (uint32_t e = t; e < mm; e += kOwnerThreads) {
          const uint32_t i = list[base + sub + e];
          if (c.cover[i] < m.cull2)
            part[mm - 1u - atomicAdd(&part_n[1], 1u)] = i;
          else
            part[atomicAdd(&part_n[0], 1u)] = i;
        }
        __syncthreads();
        const uint32_t ns = part_n[0], ni = mm - ns;
        packed(ns, [&](uint32_t e) { walk(part[e]); });
        packed(ni, [&](uint32_t e) { walk(part[ns + e]); });
        __syncthreads();
      }
    } else {
      packed(n_e, [&](uint32_t e) { walk(list[base + e]); });
    }
    __syncthreads();
    // Voxel coordinates from the flat index by multiply-high with ceil(2^32 / d):
    // exact for indices below 2^15 and divisors up to 26 (checked exhaustively).
    const uint32_t mx = 0xFFFFFFFFu / (uint32_t)nx + 1u, my = 0xFFFFFFFFu / (uint32_t)ny + 1u;
    for (int v = t; v < nvox; v += kOwnerThreads) {
      const uint32_t val = lds[v];
      if (val == init) continue;
      const uint32_t qx = __umulhi((uint32_t)v, mx), qy = __umulhi(qx, my);
      const int32_t x = ex0 + (v - (int32_t)qx * nx), y = ey0 + ((int32_t)qx - (int32_t)qy * ny), z = ez0 + (int32_t)qy;
      atomicMin(sgrid + (uint32_t)x + (uint32_t)m.r * ((uint32_t)y + (uint32_t)m.r * (uint32_t)z), val);
      if (kCount) n_glob++;
    }
    __syncthreads();
  }
  if (kCount) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      n_lds += __shfl_xor(n_lds, off);
      n_glob += __shfl_xor(n_glob, off);
    }
    if ((t & 63) == 0) {
      atomicAdd(&counts[0], n_lds);
      atomicAdd(&counts[1], n_glob);
    }
  }
}

// Step 4: codes, exact live-tile mask and texels, one workgroup per brick.
template <bool kNormals>
__global__ void __launch_bounds__(kThreads)
    k_brick_resolve(MipsK mk_, const uint32_t* __restrict__ count, uint8_t* __restrict__ tile_live,
                    uint16_t* __restrict__ dist, uint32_t* __restrict__ grid) {
  __shared__ uint32_t tflag;
  const uint32_t b = blockIdx.x;
  const uint32_t scene = b / mk_.bps;
  const BrickPos q = brick_pos(mk_, b - scene * mk_.bps);
  const SplatMipK& m = mk_.m[q.mi];
  tile_live += (size_t)scene * mk_.tps;
  dist += (size_t)scene * mk_.vps;
  grid += (size_t)scene * mk_.gps;
  const int32_t tres = m.r / kTile;  // tiles per axis at this mip
  const int t = threadIdx.x;
  auto tile_index = [&](int tx, int ty, int tz) -> int32_t {
    const int32_t gx = q.bx * 2 + tx, gy = q.by * 2 + ty, gz = q.bz * 2 + tz;
    if (gx >= tres || gy >= tres || gz >= tres) return -1;
    return (int32_t)(m.toff + (uint32_t)gx + (uint32_t)tres * ((uint32_t)gy + (uint32_t)tres * (uint32_t)gz));
  };
  // Tiles of this brick that were live after the previous bake.
  uint32_t stale = 0;
  for (int k = 0; k < 8; k++) {
    const int32_t ti = tile_index(k & 1, (k >> 1) & 1, k >> 2);
    if (ti >= 0 && tile_live[ti]) stale |= 1u << k;
  }
  const bool has = brick_touched(mk_, count, b);
  if (!has && !stale) return;
  if (t == 0) tflag = 0;
  __syncthreads();
  uint32_t codes[kBrickVox / kThreads];
  uint32_t mask = 0;
#pragma unroll
  for (int j = 0; j < kBrickVox / kThreads; j++) {
    const int v = t + j * kThreads;
    const int32_t x = q.ox + (v & 15), y = q.oy + ((v >> 4) & 15), z = q.oz + (v >> 8);
    uint32_t code = kEncSat;
    if (has && x < m.r && y < m.r && z < m.r) {
      uint32_t* g = grid + m.off + (uint32_t)x + (uint32_t)m.r * ((uint32_t)y + (uint32_t)m.r * (uint32_t)z);
      const uint32_t acc = *g;
      code = kNormals ? acc : code_from_s(m, acc);
      *g = code;
      if (code != kEncSat) mask |= 1u << (((v & 15) >> 3) | ((((v >> 4) & 15) >> 3) << 1) | (((v >> 8) >> 3) << 2));
    }
    codes[j] = code;
  }
  if (mask) atomicOr(&tflag, mask);
  __syncthreads();
  const uint32_t live = tflag;
  const uint32_t write = live | stale;
#pragma unroll
  for (int j = 0; j < kBrickVox / kThreads; j++) {
    const int v = t + j * kThreads;
    const int lx = v & 15, ly = (v >> 4) & 15, lz = v >> 8;
    const int k = (lx >> 3) | ((ly >> 3) << 1) | ((lz >> 3) << 2);
    if (!((write >> k) & 1u)) continue;
    const int32_t x = q.ox + lx, y = q.oy + ly, z = q.oz + lz;
    if (x >= m.r || y >= m.r || z >= m.r) continue;
    const uint32_t gi = (uint32_t)x + (uint32_t)m.r * ((uint32_t)y + (uint32_t)m.r * (uint32_t)z);
    dist[m.voff + gi] = codes[j] == kEncSat ? mk_.sat_half : f2h((float)codes[j] * mk_.inv_enc - 1.0f);
    if (!has) grid[m.off + gi] = kEncSat;
  }
  if (t < 8) {
    const int32_t ti = tile_index(t & 1, (t >> 1) & 1, t >> 2);
    if (ti >= 0) tile_live[ti] = (uint8_t)((live >> t) & 1u);
  }
}

// Follows each scene: origin = centre - worldSize / 2, as SplatField.setCenter
// followed by SdfVolume.setOrigin computes it (splat.ts:284-287, volume.ts:327-334:
// f64 arithmetic, stored as f32).  The centre is `centers` (6 doubles per scene,
// xyz first: a body centroid) when given, else the f32 midpoint of the f32 bounds
// of the scene's particles.  A scene with no particles keeps its origin.  One
// workgroup per scene; `base` holds scenes + 1 prefix offsets, or is null for one
// scene of `n` particles.
__global__ void __launch_bounds__(kThreads)
    k_follow(const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
             const uint32_t* __restrict__ base, uint32_t n, const double* __restrict__ centers,
             VolDev* __restrict__ vols, double half) {
  __shared__ float sh[6][kThreads];
  const uint32_t sc = blockIdx.x;
  if (centers) {
    if (threadIdx.x == 0) {
      for (int q = 0; q < 3; q++) {
        const double o = centers[6 * sc + q] - half;
        vols[sc].origin[q] = (float)o;
        vols[sc].center[q] = (float)(o + half);
      }
    }
    return;
  }
  const uint32_t lo = base ? base[sc] : 0u, hi = base ? base[sc + 1] : n;
  if (lo >= hi) return;
  const int t = threadIdx.x;
  const float inf = __uint_as_float(0x7f800000u);
  float a[6] = {inf, inf, inf, -inf, -inf, -inf};
  for (uint32_t i = lo + t; i < hi; i += kThreads) {
    const float x = px[i], y = py[i], z = pz[i];
    a[0] = fmin_s(a[0], x);
    a[1] = fmin_s(a[1], y);
    a[2] = fmin_s(a[2], z);
    a[3] = fmax_s(a[3], x);
    a[4] = fmax_s(a[4], y);
    a[5] = fmax_s(a[5], z);
  }
  for (int q = 0; q < 6; q++) sh[q][t] = a[q];
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (t < off) {
      for (int q = 0; q < 3; q++) sh[q][t] = fmin_s(sh[q][t], sh[q][t + off]);
      for (int q = 3; q < 6; q++) sh[q][t] = fmax_s(sh[q][t], sh[q][t + off]);
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int q = 0; q < 3; q++) {
      const float c = (sh[q][0] + sh[q + 3][0]) * 0.5f;
      const double o = (double)c - half;
      vols[sc].origin[q] = (float)o;
      vols[sc].center[q] = (float)(o + half);
    }
  }
}

// Bounds of each scene's particles for the bounds-following bake, over the whole
// grid: every workgroup reduces a slice of one scene (blockIdx.y) in LDS, then
// folds it into the scene's six slots with integer atomics on order-preserving
// encodings of the floats (exact, so the result does not depend on timing).

__global__ void __launch_bounds__(kThreads)
    k_bounds_partial(const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
                     const uint32_t* __restrict__ base, uint32_t n, uint32_t* __restrict__ slots) {
  __shared__ float sh[6][kThreads];
  const uint32_t sc = blockIdx.y;
  const uint32_t lo = base ? base[sc] : 0u, hi = base ? base[sc + 1] : n;
  const int t = threadIdx.x;
  const float inf = __uint_as_float(0x7f800000u);
  float a[6] = {inf, inf, inf, -inf, -inf, -inf};
  for (uint32_t i = lo + blockIdx.x * kThreads + t; i < hi; i += gridDim.x * kThreads) {
    const float x = px[i], y = py[i], z = pz[i];
    a[0] = fmin_s(a[0], x);
    a[1] = fmin_s(a[1], y);
    a[2] = fmin_s(a[2], z);
    a[3] = fmax_s(a[3], x);
    a[4] = fmax_s(a[4], y);
    a[5] = fmax_s(a[5], z);
  }
  for (int q = 0; q < 6; q++) sh[q][t] = a[q];
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (t < off) {
      for (int q = 0; q < 3; q++) sh[q][t] = fmin_s(sh[q][t], sh[q][t + off]);
      for (int q = 3; q < 6; q++) sh[q][t] = fmax_s(sh[q][t], sh[q][t + off]);
    }
    __syncthreads();
  }
  if (t < 3) atomicMin(&slots[kBoundsStride * sc + t], ordered_bits(sh[t][0]));
  else if (t < 6) atomicMax(&slots[kBoundsStride * sc + t], ordered_bits(sh[t][0]));
}

// Slots start at (max, max, max, 0, 0, 0) encoded: an empty scene keeps its origin.
// k_bounds_follow puts them back there once it has read them, so the keys pass of
// the next frame-end sort folds into fresh slots with no reset launch before it.
__global__ void k_bounds_init(uint32_t* __restrict__ slots, uint32_t scenes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 6 * scenes) slots[kBoundsStride * (i / 6) + i % 6] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
}

__global__ void k_bounds_follow(uint32_t* __restrict__ slots, const uint32_t* __restrict__ base, uint32_t n,
                                VolDev* __restrict__ vols, double half, uint32_t scenes) {
  const uint32_t sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= scenes) return;
  const uint32_t lo = base ? base[sc] : 0u, hi = base ? base[sc + 1] : n;
  if (lo >= hi) return;  // no particle folded into this scene's slots: still fresh
  for (int q = 0; q < 3; q++) {
    const float c =
        (from_ordered(slots[kBoundsStride * sc + q]) + from_ordered(slots[kBoundsStride * sc + 3 + q])) * 0.5f;
    const double o = (double)c - half;
    vols[sc].origin[q] = (float)o;
    vols[sc].center[q] = (float)(o + half);
  }
  for (int q = 0; q < 6; q++) slots[kBoundsStride * sc + q] = q < 3 ? 0xFFFFFFFFu : 0u;
}

}  // namespace

// The owner schedule needs every particle's voxel box within kHalo voxels of its
// own voxel at every level: reach / voxel = radius / voxel + band < kHalo.
static bool splat_owner_fits(const Splat* sp) {
  if (!sp->owner || sp->brick_res.empty()) return false;
  const double voxel0 = sp->vol->info.world_size / sp->vol->info.resolution;
  for (int m = 0; m < sp->mips; m++) {
    const double voxel = voxel0 * (double)(1 << m);
    if (sp->desc.radius / voxel + sp->desc.band > (double)kHalo - 0.01) return false;
  }
  return true;
}

static void fill_mips(Splat* sp, MipsK* mk_) {
  const Volume* v = sp->vol;
  const double voxel0 = v->info.world_size / v->info.resolution;
  mk_->mips = sp->mips;
  mk_->inv_enc = (float)(1.0 / 32767.5);
  mk_->sat_half = f2h_host(1.0f);
  mk_->chunk = sp->chunk;
  mk_->bps = (uint32_t)sp->brick_total;
  mk_->gps = (uint32_t)sp->grid_total;
  mk_->vps = (uint32_t)v->total_voxels;
  mk_->tps = (uint32_t)sp->tile_total;
  mk_->xcd = sp->xcd ? 1u : 0u;
  mk_->owner = splat_owner_fits(sp) ? 1u : 0u;
  for (int m = 0; m < sp->mips; m++) {
    SplatMipK& k = mk_->m[m];
    k.r = v->mip_res(m);
    k.nb = sp->brick_res.empty() ? 0 : sp->brick_res[m];
    k.off = (uint32_t)sp->grid_off[m];
    k.voff = (uint32_t)v->voxel_off[m];
    k.boff = sp->brick_off.empty() ? 0 : (uint32_t)sp->brick_off[m];
    k.toff = sp->tile_off.empty() ? 0 : (uint32_t)sp->tile_off[m];
    const double voxel = voxel0 * (double)(1 << m);
    const double bw = sp->desc.band * voxel;
    k.voxel = (float)voxel;
    k.bw = (float)bw;
    k.reach = (float)(sp->desc.radius + bw);
    k.inv = (float)(1.0 / voxel);
    k.radius = (float)sp->desc.radius;
    const double cr = (double)k.reach * 1.001 + 1e-6;
    k.cull2 = (float)(cr * cr);
  }
}

static CloudK cloud_k(const Splat* sp, const uint32_t* order, const float* cover) {
  CloudK c{};
  for (int a = 0; a < 3; a++) {
    c.pos[a] = sp->cloud.pos[a];
    c.nrm[a] = sp->cloud.nrm[a];
  }
  uint32_t first = 0, count = sp->cloud.capacity;
  if (sp->cloud.live_rule == CLAY_LIVE_PREFIX) {
    count = sp->cloud.live_count;
  } else if (sp->cloud.live_rule == CLAY_LIVE_RANGE) {
    first = sp->cloud.live_first;
    count = sp->cloud.live_count;
  }
  if (first > sp->cloud.capacity) first = sp->cloud.capacity;
  if (count > sp->cloud.capacity - first) count = sp->cloud.capacity - first;
  c.first = first;
  c.end = first + count;
  c.count = count;
  c.order = order;
  c.scene_of = sp->scene_of;
  c.cover = c.nrm[0] ? nullptr : cover;
  return c;
}

// Voxels of the largest region the owner schedule stages for these levels: per
// axis a brick and its kHalo shell, clipped to the level.  A 32^3 level has two
// bricks per axis, each clipped on one side (21 voxels), so its workgroups need
// half the LDS of a brick inside a large level (26).
static uint32_t owner_stage_voxels(const MipsK& mk_) {
  int32_t ext = 1;
  for (int mi = 0; mi < mk_.mips; mi++) {
    const int32_t r = mk_.m[mi].r;
    for (int32_t o = 0; o < r; o += kBrick) {
      const int32_t lo = o - kHalo > 0 ? o - kHalo : 0;
      const int32_t hi = o + kBrick + kHalo < r ? o + kBrick + kHalo : r;
      ext = std::max(ext, hi - lo);
    }
  }
  return (uint32_t)(ext * ext * ext);
}

// Binning: everything before the walk, which reads positions and the visiting
// order only (not the interior radii).
template <bool kNormals>
static int tiled_prepare(Splat* sp, hipStream_t s, const MipsK& mk_, const CloudK& c) {
  Ctx* ctx = sp->ctx;
  const unsigned pgrid = (c.count + kThreads - 1) / kThreads;
  const unsigned nbricks = (unsigned)(sp->brick_total * sp->scenes);
  ScanOut* scan = reinterpret_cast<ScanOut*>(sp->brick_total_dev);
  int rc = launch_fill_u32(ctx, s, sp->brick_sub, (size_t)nbricks * kSlices, 0u);
  if (rc) return rc;
  if (c.count) {
    LaunchTimer t(ctx, s, "splat_bin_count");
    hipLaunchKernelGGL(k_bin<false>, dim3(pgrid), dim3(kThreads), 0, s, mk_, c, sp->vol->dev,
                       sp->brick_sub, nullptr);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  {
    LaunchTimer t(ctx, s, "splat_brick_scan");
    const unsigned sgrid = (unsigned)(((size_t)nbricks * kSlices + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_brick_sum, dim3(sgrid), dim3(kThreads), 0, s, sp->brick_sub, (uint32_t)nbricks,
                       sp->brick_count);
    hipLaunchKernelGGL(k_brick_scan, dim3(1), dim3(kScanThreads), 0, s, (uint32_t)nbricks, sp->chunk, sp->brick_count,
                       sp->brick_start, sp->item_start, scan);
    hipLaunchKernelGGL(k_brick_cursor, dim3(sgrid), dim3(kThreads), 0, s, sp->brick_sub, (uint32_t)nbricks,
                       sp->brick_start, sp->brick_cursor);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  if (c.count) {
    LaunchTimer t(ctx, s, "splat_bin_fill");
    hipLaunchKernelGGL(k_bin<true>, dim3(pgrid), dim3(kThreads), 0, s, mk_, c, sp->vol->dev,
                       sp->brick_cursor, sp->brick_list);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  {
    LaunchTimer t(ctx, s, "splat_brick_clear");
    hipLaunchKernelGGL(k_brick_clear<kNormals>, dim3(nbricks), dim3(kThreads), 0, s, mk_, sp->brick_count,
                       sp->grid);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  return CLAY_OK;
}

// The walk and the resolve, after tiled_prepare.
template <bool kNormals>
static int tiled_finish(Splat* sp, hipStream_t s, const MipsK& mk_, const CloudK& c) {
  Ctx* ctx = sp->ctx;
  const unsigned nbricks = (unsigned)(sp->brick_total * sp->scenes);
  ScanOut* scan = reinterpret_cast<ScanOut*>(sp->brick_total_dev);
  if (c.count) {
    LaunchTimer t(ctx, s, "splat_brick");
    if (mk_.owner && sp->atomic_counts)
      hipLaunchKernelGGL((k_owner_splat<kNormals, true>),
                         dim3(std::max(1, sp->persistent_blocks * kThreads / kOwnerThreads)), dim3(kOwnerThreads),
                         sizeof(uint32_t) * owner_stage_voxels(mk_), s, mk_, c,
                         sp->vol->dev, sp->brick_start, sp->brick_count, sp->item_start, scan,
                         (uint32_t)nbricks, sp->brick_list, sp->grid, sp->atomic_counts);
    else if (mk_.owner)
      hipLaunchKernelGGL((k_owner_splat<kNormals, false>),
                         dim3(std::max(1, sp->persistent_blocks * kThreads / kOwnerThreads)), dim3(kOwnerThreads),
                         sizeof(uint32_t) * owner_stage_voxels(mk_), s, mk_, c,
                         sp->vol->dev, sp->brick_start, sp->brick_count, sp->item_start, scan,
                         (uint32_t)nbricks, sp->brick_list, sp->grid, nullptr);
    else
      hipLaunchKernelGGL(k_brick_splat<kNormals>, dim3(sp->persistent_blocks), dim3(kThreads), 0, s, mk_, c,
                         sp->vol->dev, sp->brick_start, sp->brick_count, sp->item_start, scan,
                         (uint32_t)nbricks, sp->brick_list, sp->grid);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  {
    LaunchTimer t(ctx, s, "splat_brick_resolve");
    hipLaunchKernelGGL(k_brick_resolve<kNormals>, dim3(nbricks), dim3(kThreads), 0, s, mk_, sp->brick_count,
                       sp->tile_live, sp->vol->dist, sp->grid);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  return CLAY_OK;
}

int splat_bounds_reset(Splat* sp, hipStream_t s) {
  if (!sp->follow_slots) return CLAY_OK;
  const unsigned S = (unsigned)sp->scenes;
  hipLaunchKernelGGL(k_bounds_init, dim3((6 * S + 255) / 256), dim3(256), 0, s, sp->follow_slots, S);
  CLAY_HIP_CHECK(hipGetLastError());
  sp->bounds_pending = false;
  return CLAY_OK;
}

// The slots a keys pass may fold into.  They are fresh (k_bounds_follow of the
// previous bake reset them) unless a bake that was handed them never reached its
// k_bounds_follow (an error between the two), which a reset here repairs.
uint32_t* splat_bounds_begin(Splat* sp, hipStream_t s) {
  if (!sp->follow || sp->follow_centers || !sp->follow_slots) return nullptr;
  if (sp->bounds_pending && splat_bounds_reset(sp, s) != CLAY_OK) return nullptr;
  sp->bounds_pending = true;
  return sp->follow_slots;
}

int splat_bake_prepare(Splat* sp, hipStream_t s, const uint32_t* order, bool bounds_given) {
  Ctx* ctx = sp->ctx;
  MipsK mk_{};
  fill_mips(sp, &mk_);
  const CloudK c = cloud_k(sp, order, nullptr);
  if (sp->follow) {
    LaunchTimer t(ctx, s, "splat_follow");
    const double half = sp->vol->info.world_size / 2;
    if (!sp->follow_centers && bounds_given) {
      const unsigned S = (unsigned)sp->scenes;
      hipLaunchKernelGGL(k_bounds_follow, dim3((S + 63) / 64), dim3(64), 0, s, sp->follow_slots, sp->scene_base,
                         c.end, sp->vol->dev, half, S);
      sp->bounds_pending = false;
    } else if (sp->follow_centers) {
      hipLaunchKernelGGL(k_follow, dim3(sp->scenes), dim3(kThreads), 0, s, c.pos[0], c.pos[1], c.pos[2],
                         sp->scene_base, c.end, sp->follow_centers, sp->vol->dev, half);
    } else {
      const unsigned S = (unsigned)sp->scenes;
      // Enough workgroups to fill the chip once over all scenes.
      const unsigned per = (unsigned)std::max<int>(1, std::min<int>(64, (sp->persistent_blocks + S - 1) / S));
      hipLaunchKernelGGL(k_bounds_init, dim3((6 * S + 255) / 256), dim3(256), 0, s, sp->follow_slots, S);
      hipLaunchKernelGGL(k_bounds_partial, dim3(per, S), dim3(kThreads), 0, s, c.pos[0], c.pos[1], c.pos[2],
                         sp->scene_base, c.end, sp->follow_slots);
      hipLaunchKernelGGL(k_bounds_follow, dim3((S + 63) / 64), dim3(64), 0, s, sp->follow_slots, sp->scene_base,
                         c.end, sp->vol->dev, half, S);
    }
    CLAY_HIP_CHECK(hipGetLastError());
  }
  if (sp->desc.mode == CLAY_SPLAT_PLAIN) return CLAY_OK;
  return c.nrm[0] ? tiled_prepare<true>(sp, s, mk_, c) : tiled_prepare<false>(sp, s, mk_, c);
}

int splat_bake_finish(Splat* sp, hipStream_t s, const uint32_t* order, const float* cover) {
  Ctx* ctx = sp->ctx;
  MipsK mk_{};
  fill_mips(sp, &mk_);
  const CloudK c = cloud_k(sp, order, cover);
  const unsigned pgrid = (c.count + kThreads - 1) / kThreads;
  if (sp->desc.mode == CLAY_SPLAT_PLAIN) {
    int rc = launch_fill_u32(ctx, s, sp->grid, sp->grid_total, kEncSat);
    if (rc) return rc;
    for (int m = 0; m < sp->mips; m++) {
      if (c.count) {
        LaunchTimer t(ctx, s, "splat_plain");
        hipLaunchKernelGGL(k_splat_plain, dim3(pgrid), dim3(kThreads), 0, s, mk_.m[m], c,
                           sp->vol->dev, sp->grid);
        CLAY_HIP_CHECK(hipGetLastError());
      }
      const uint32_t cells = (uint32_t)mk_.m[m].r * mk_.m[m].r * mk_.m[m].r;
      unsigned rg = (cells + kThreads - 1) / kThreads;
      if (rg > 16384) rg = 16384;
      LaunchTimer t(ctx, s, "splat_resolve");
      hipLaunchKernelGGL(k_resolve_plain, dim3(rg), dim3(kThreads), 0, s, mk_.m[m], mk_.inv_enc,
                         sp->grid, sp->vol->dist);
      CLAY_HIP_CHECK(hipGetLastError());
    }
    sp->has_baked = true;
    return CLAY_OK;
  }
  const int rc = c.nrm[0] ? tiled_finish<true>(sp, s, mk_, c) : tiled_finish<false>(sp, s, mk_, c);
  if (rc) return rc;
  sp->has_baked = true;
  return CLAY_OK;
}

int splat_bake_ordered(Splat* sp, hipStream_t s, const uint32_t* order, const float* cover) {
  const int rc = splat_bake_prepare(sp, s, order, false);
  return rc ? rc : splat_bake_finish(sp, s, order, cover);
}

int splat_bake(Splat* sp, hipStream_t s) { return splat_bake_ordered(sp, s, nullptr, nullptr); }

int splat_brick_geometry(int res, int mips, std::vector<int>* brick_res, std::vector<size_t>* brick_off,
                         size_t* brick_total, std::vector<size_t>* tile_off, size_t* tile_total) {
  size_t bo = 0, to = 0;
  for (int m = 0; m < mips; m++) {
    const int r = res >> m;
    const int nb = (r + kBrick - 1) / kBrick;
    brick_res->push_back(nb);
    brick_off->push_back(bo);
    bo += (size_t)nb * nb * nb;
    const int tr = r / kTile;
    tile_off->push_back(to);
    to += (size_t)tr * tr * tr;
  }
  *brick_total = bo;
  *tile_total = to;
  return CLAY_OK;
}

size_t splat_bin_slices() { return kSlices; }

// ---------------------------------------------------------------------------
// A bake that allocates nothing (clay_splat_bake_static)
// ---------------------------------------------------------------------------
// The plain schedule's working memory (the origin record and the dense u32 code
// grid of every mip) lives in one file-scope device array.  Each device holds its
// own copy for the life of the process.  Host code never passes `scratch` itself
// to a kernel: it passes the address hipGetSymbolAddress returned for the current
// device, looked up once per device when a context is created.
constexpr size_t kStaticScratchBytes = (size_t)32 << 20;
constexpr uint32_t kStaticHeadWords = 64;  // the origin record, padded to 256 bytes
static_assert(sizeof(VolDev) <= 4 * kStaticHeadWords, "the origin record must fit the scratch head");
static __device__ __align__(256) uint32_t scratch[kStaticScratchBytes / 4];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
uint32_t* g_static_ptr[kMaxDevices] = {};

__global__ void k_put_voldev(VolDev* __restrict__ dst, VolDev v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

__global__ void __launch_bounds__(kThreads) k_fill_codes(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) p[i] = v;
}
}  // namespace

int splat_static_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(CLAY_ERR_INVALID, "clay_splat_bake_static: device index beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (g_static_ptr[device]) return CLAY_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(CLAY_ERR_HIP, std::string("clay_splat_bake_static: hipGetSymbolAddress: ") + hipGetErrorString(e));
  g_static_ptr[device] = static_cast<uint32_t*>(p);
  return CLAY_OK;
}

size_t splat_static_bytes() { return kStaticScratchBytes; }

int splat_bake_static(Ctx* ctx, const clay_cloud_desc* cloud, const clay_splat_desc* desc, uint16_t* dist,
                      hipStream_t s) {
  if (!(desc->radius > 0)) return fail(CLAY_ERR_INVALID, "SplatField: radius must be positive");
  if (desc->mode != CLAY_SPLAT_PLAIN)
    return fail(CLAY_ERR_INVALID, "clay_splat_bake_static: only the plain schedule runs without allocation");
  if (cloud->live_rule != CLAY_LIVE_ALL && cloud->live_rule != CLAY_LIVE_PREFIX && cloud->live_rule != CLAY_LIVE_RANGE)
    return fail(CLAY_ERR_INVALID, "ParticleCloud: unknown live rule");
  if (cloud->capacity > 0 && (!cloud->pos[0] || !cloud->pos[1] || !cloud->pos[2]))
    return fail(CLAY_ERR_INVALID, "ParticleCloud: positions are required");
  const bool nrm = cloud->nrm[0] != nullptr;
  if (nrm && (!cloud->nrm[1] || !cloud->nrm[2]))
    return fail(CLAY_ERR_INVALID, "ParticleCloud: normals must give all three axes or none");
  // Defaults and geometry as clay_splat_create derives them (splat.ts:147-160).
  const int res = desc->resolution > 0 ? desc->resolution : 64;
  const int mips_req = desc->mip_levels > 0 ? desc->mip_levels : std::max(1, (int)log2((double)res / kTile));
  clay_volume_info info;
  int rc = clay_volume_derive(res, mips_req, &info);
  if (rc) return rc;
  const double band = desc->band > 0 ? desc->band : 2.0;
  const double world = desc->world_size > 0 ? desc->world_size : 24.0;
  size_t cells = 0;
  for (int m = 0; m < info.mip_levels; m++) {
    const size_t r = (size_t)(res >> m);
    cells += r * r * r;
  }
  const size_t need = 4 * (kStaticHeadWords + cells);
  if (need > kStaticScratchBytes) {
    char buf[200];
    snprintf(buf, sizeof(buf),
             "clay_splat_bake_static: a %d^3 bake in %d mip levels needs %zu bytes of scratch; the static "
             "scratch holds %zu",
             res, info.mip_levels, need, kStaticScratchBytes);
    return fail(CLAY_ERR_INVALID, buf);
  }
  uint32_t* base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_static_ptr[ctx->device] : nullptr;
  if (!base) return fail(CLAY_ERR_STATE, "clay_splat_bake_static: the static scratch of this device is not set up");
  VolDev* vdev = reinterpret_cast<VolDev*>(base);
  uint32_t* grid = base + kStaticHeadWords;
  // The origin record the kernels read (volume.ts:221-230), as SdfVolume writes it.
  VolDev vh{};
  const double half = world / 2;
  for (int a = 0; a < 3; a++) {
    vh.origin[a] = (float)desc->origin[a];
    vh.center[a] = (float)(desc->origin[a] + half);
  }
  vh.voxel = (float)(world / res);
  vh.extent = (float)world;
  vh.half = (float)half;
  vh.band = (float)band;
  vh.grad_h = vh.voxel;
  vh.res = res;
  MipsK mk_{};
  mk_.mips = info.mip_levels;
  mk_.inv_enc = (float)(1.0 / 32767.5);
  const double voxel0 = world / res;
  size_t off = 0;
  for (int m = 0; m < info.mip_levels; m++) {
    SplatMipK& k = mk_.m[m];
    k.r = res >> m;
    k.off = (uint32_t)off;
    k.voff = (uint32_t)off;
    const double voxel = voxel0 * (double)(1 << m);
    const double bw = band * voxel;
    k.voxel = (float)voxel;
    k.bw = (float)bw;
    k.reach = (float)(desc->radius + bw);
    k.inv = (float)(1.0 / voxel);
    k.radius = (float)desc->radius;
    const double cr = (double)k.reach * 1.001 + 1e-6;
    k.cull2 = (float)(cr * cr);
    off += (size_t)k.r * k.r * k.r;
  }
  CloudK c{};
  for (int a = 0; a < 3; a++) {
    c.pos[a] = cloud->pos[a];
    c.nrm[a] = cloud->nrm[a];
  }
  uint32_t first = 0, count = cloud->capacity;
  if (cloud->live_rule == CLAY_LIVE_PREFIX) {
    count = cloud->live_count;
  } else if (cloud->live_rule == CLAY_LIVE_RANGE) {
    first = cloud->live_first;
    count = cloud->live_count;
  }
  if (first > cloud->capacity) first = cloud->capacity;
  if (count > cloud->capacity - first) count = cloud->capacity - first;
  c.first = first;
  c.end = first + count;
  c.count = count;
  // Every call rewrites the record and refills the whole grid: the array keeps the
  // previous call's contents.
  hipLaunchKernelGGL(k_put_voldev, dim3(1), dim3(64), 0, s, vdev, vh);
  const unsigned fill_blocks = (unsigned)std::min<size_t>((cells + kThreads - 1) / kThreads, 16384);
  hipLaunchKernelGGL(k_fill_codes, dim3(fill_blocks), dim3(kThreads), 0, s, grid, (uint32_t)cells, kEncSat);
  const unsigned pgrid = (c.count + kThreads - 1) / kThreads;
  for (int m = 0; m < info.mip_levels; m++) {
    if (c.count)
      hipLaunchKernelGGL(k_splat_plain, dim3(pgrid), dim3(kThreads), 0, s, mk_.m[m], c, vdev, grid);
    const uint32_t mc = (uint32_t)mk_.m[m].r * mk_.m[m].r * mk_.m[m].r;
    const unsigned rg = std::min<unsigned>((mc + kThreads - 1) / kThreads, 16384u);
    hipLaunchKernelGGL(k_resolve_plain, dim3(rg), dim3(kThreads), 0, s, mk_.m[m], mk_.inv_enc, grid, dist);
  }
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

size_t splat_max_bricks_per_particle(double radius, double band, double voxel0, int mips) {
  size_t worst = 1;
  for (int m = 0; m < mips; m++) {
    const double voxel = voxel0 * (double)(1 << m);
    const double span = 2.0 * (radius + band * voxel) / voxel + 2.0;  // voxels covered per axis
    const size_t per_axis = (size_t)(span / kBrick) + 2;
    const size_t b3 = per_axis * per_axis * per_axis;
    worst += b3;
  }
  return worst;
}

}  // namespace clay
