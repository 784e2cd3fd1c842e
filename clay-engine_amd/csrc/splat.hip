// splat.hip -- SplatField bake for gfx950 (lib/src/sim/splat.ts:186-304).
//
// Two schedules produce the same u32 codes, voxel for voxel:
//
// PLAIN (the reference's schedule): clear a dense per-mip u32 grid to 65535, one
//   thread per particle walks its voxel box and does a global atomicMin of
//   e = u32((clamp(sd / bandWorld, -1, 1) + 1) * 32767.5), then resolve every
//   voxel to f16.
//
// TILED: particles are binned to 16^3-voxel bricks of every mip (a particle
//   lands in each brick its voxel box overlaps; the per-brick counters are hit
//   once per wave and distinct brick, not once per lane).  One workgroup per
//   brick owns a 16 KB u32 LDS copy of the brick, does the atomicMin there
//   (ds_min_u32), finds which 8^3 tiles came out live (some code < 65535) and
//   writes only tiles that are live now or were live after the previous bake;
//   every other tile already reads the saturated +1 the reference's clear would
//   give it.  Voxels whose centre is farther than `reach` from the particle are
//   skipped row by row, which cannot change a code: they fail sd < bandWorld.
//
// min is order independent, so both schedules give bit-identical codes whatever
// the atomic order; the f32 arithmetic per voxel is the reference expression
// evaluated in IEEE order (contraction off).
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kBrick = 16;                  // voxels per brick edge
constexpr int kBrickVox = kBrick * kBrick * kBrick;
constexpr int kThreads = 256;

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
  const uint32_t* order;  // optional visiting order (any permutation gives the same codes)
  uint32_t first, end;    // live particles: ids in [first, end)
  uint32_t count;         // end - first
};

struct MipsK {
  int32_t mips;
  float inv_enc;          // f32(1 / 32767.5)
  uint16_t sat_half;      // f16(+1.0)
  SplatMipK m[CLAY_MAX_MIPS];
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

// Code of voxel (x,y,z) for a particle at p (relative to the volume origin), or
// 65535 when the voxel is outside the band.  splat.ts:244-250 (and :170-184 with normals).
__device__ __forceinline__ uint32_t voxel_code(const SplatMipK& m, f3 p, bool has_n, f3 n,
                                               int32_t x, int32_t y, int32_t z) {
  const f3 cc = mk(((float)x + 0.5f) * m.voxel, ((float)y + 0.5f) * m.voxel,
                   ((float)z + 0.5f) * m.voxel);
  const f3 d = cc - p;
  float sd = len3(d) - m.radius;
  if (has_n) sd = fmax_s(dot3(d, n), sd);
  if (sd < m.bw) {
    const float q = (fclamp_s(sd / m.bw, -1.0f, 1.0f) + 1.0f) * kEncScale;
    return (uint32_t)q;
  }
  return kEncSat;
}

__device__ __forceinline__ f3 load_rel(const CloudK& c, const VolDev* vol, uint32_t i) {
  return mk(c.pos[0][i] - vol->origin[0], c.pos[1][i] - vol->origin[1], c.pos[2][i] - vol->origin[2]);
}
__device__ __forceinline__ f3 load_nrm(const CloudK& c, uint32_t i) {
  return c.nrm[0] ? mk(c.nrm[0][i], c.nrm[1][i], c.nrm[2][i]) : mk(0.0f, 0.0f, 0.0f);
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

// Brick range of particle `i` at mip `m`; nb3 = number of bricks it overlaps (0 if none).
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

// Pass 1 (count, fill=false) and pass 3 (fill=true) of the binning.
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
  f3 p = mk(0.0f, 0.0f, 0.0f);
  if (valid) p = load_rel(c, vol, i);
  for (int mi = 0; mi < mk_.mips; mi++) {
    const SplatMipK& m = mk_.m[mi];
    int32_t blo[3] = {0, 0, 0}, bext[3] = {1, 1, 1};
    const int nb3 = valid ? brick_range(m, p, blo, bext) : 0;
    for (int q = 0;; q++) {
      const bool act = q < nb3;
      if (!__ballot(act)) break;
      uint32_t b = 0;
      if (act) {
        const int32_t bx = blo[0] + q % bext[0];
        const int32_t by = blo[1] + (q / bext[0]) % bext[1];
        const int32_t bz = blo[2] + q / (bext[0] * bext[1]);
        b = m.boff + (uint32_t)bx + (uint32_t)m.nb * ((uint32_t)by + (uint32_t)m.nb * (uint32_t)bz);
      }
      const uint32_t slot = wave_add(counters, b, act);
      if (kFill && act) list[slot] = i;
    }
  }
}

// Pass 2: exclusive scan of brick counts -> start; cursor = start; total.
__global__ void __launch_bounds__(1024)
    k_brick_scan(const uint32_t* __restrict__ count, uint32_t nbricks, uint32_t* __restrict__ start,
                 uint32_t* __restrict__ cursor, uint32_t* __restrict__ total) {
  __shared__ uint32_t sh[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < nbricks; c0 += 1024) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < nbricks ? count[i] : 0u;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const uint32_t add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0u;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    const uint32_t incl = sh[threadIdx.x];
    if (i < nbricks) {
      start[i] = carry + incl - v;
      cursor[i] = carry + incl - v;
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Pass 4: one workgroup per brick of every mip.
__global__ void __launch_bounds__(kThreads)
    k_brick_splat(MipsK mk_, CloudK c, const VolDev* __restrict__ vol,
                  const uint32_t* __restrict__ start, const uint32_t* __restrict__ count,
                  const uint32_t* __restrict__ list, uint8_t* __restrict__ tile_live,
                  uint16_t* __restrict__ dist, uint32_t* __restrict__ codes) {
  __shared__ uint32_t lds[kBrickVox];
  __shared__ uint32_t tflag;
  const uint32_t b = blockIdx.x;
  int mi = 0;
  while (mi + 1 < mk_.mips && b >= mk_.m[mi + 1].boff) mi++;
  const SplatMipK& m = mk_.m[mi];
  const uint32_t lb = b - m.boff;
  const int32_t bx = (int32_t)(lb % (uint32_t)m.nb);
  const int32_t by = (int32_t)((lb / (uint32_t)m.nb) % (uint32_t)m.nb);
  const int32_t bz = (int32_t)(lb / ((uint32_t)m.nb * (uint32_t)m.nb));
  const int32_t ox = bx * kBrick, oy = by * kBrick, oz = bz * kBrick;
  const int32_t tres = m.r / kTile;  // tiles per axis at this mip
  const uint32_t n = count[b];
  const int t = threadIdx.x;

  // Which of the brick's 2x2x2 tiles exist (a brick may hang over the grid edge).
  auto tile_index = [&](int tx, int ty, int tz) -> int32_t {
    const int32_t gx = bx * 2 + tx, gy = by * 2 + ty, gz = bz * 2 + tz;
    if (gx >= tres || gy >= tres || gz >= tres) return -1;
    return (int32_t)(m.toff + (uint32_t)gx + (uint32_t)tres * ((uint32_t)gy + (uint32_t)tres * (uint32_t)gz));
  };

  if (n == 0) {
    // Nothing lands here: only tiles still holding last bake's surface need the +1.
    uint32_t stale = 0;
    for (int q = 0; q < 8; q++) {
      const int32_t ti = tile_index(q & 1, (q >> 1) & 1, q >> 2);
      if (ti >= 0 && tile_live[ti]) stale |= 1u << q;
    }
    if (!stale) return;
    __syncthreads();
    for (int v = t; v < kBrickVox; v += kThreads) {
      const int lx = v & 15, ly = (v >> 4) & 15, lz = v >> 8;
      const int q = (lx >> 3) | ((ly >> 3) << 1) | ((lz >> 3) << 2);
      if (!((stale >> q) & 1u)) continue;
      const uint32_t x = (uint32_t)(ox + lx), y = (uint32_t)(oy + ly), z = (uint32_t)(oz + lz);
      const uint32_t gi = x + (uint32_t)m.r * (y + (uint32_t)m.r * z);
      dist[m.voff + gi] = mk_.sat_half;
      if (codes) codes[m.off + gi] = kEncSat;
    }
    __syncthreads();
    if (t < 8 && ((stale >> t) & 1u)) tile_live[tile_index(t & 1, (t >> 1) & 1, t >> 2)] = 0;
    return;
  }

  for (int v = t; v < kBrickVox; v += kThreads) lds[v] = kEncSat;
  if (t == 0) tflag = 0;
  __syncthreads();

  const bool has_n = c.nrm[0] != nullptr;
  const uint32_t s0 = start[b];
  // Blocked assignment: the list is in spatial order, so lane t takes the t-th
  // contiguous chunk.  Neighbouring lanes then splat particles that lie far apart
  // and their ds_min_u32 hit different addresses, instead of 64 lanes walking
  // nearly the same voxels in lockstep.
  const uint32_t chunk = (n + kThreads - 1) / kThreads;
  const uint32_t e0 = (uint32_t)t * chunk;
  const uint32_t e1 = e0 + chunk < n ? e0 + chunk : n;
  for (uint32_t e = e0; e < e1; e++) {
    const uint32_t i = list[s0 + e];
    const f3 p = load_rel(c, vol, i);
    const f3 nr = load_nrm(c, i);
    int32_t lo[3], hi[3];
    voxel_box(m, p, lo, hi);  // non-empty: the particle was binned here
    const int32_t x0 = lo[0] > ox ? lo[0] : ox, x1 = hi[0] < ox + kBrick - 1 ? hi[0] : ox + kBrick - 1;
    const int32_t y0 = lo[1] > oy ? lo[1] : oy, y1 = hi[1] < oy + kBrick - 1 ? hi[1] : oy + kBrick - 1;
    const int32_t z0 = lo[2] > oz ? lo[2] : oz, z1 = hi[2] < oz + kBrick - 1 ? hi[2] : oz + kBrick - 1;
    for (int32_t z = z0; z <= z1; z++) {
      const float dz = ((float)z + 0.5f) * m.voxel - p.z;
      const float dz2 = dz * dz;
      if (dz2 > m.cull2) continue;
      for (int32_t y = y0; y <= y1; y++) {
        const float dy = ((float)y + 0.5f) * m.voxel - p.y;
        const float rem = m.cull2 - dz2 - dy * dy;
        if (rem < 0.0f) continue;
        // x-span of the reach ball on this row, widened by one voxel each side.
        const float s = sqrtf(rem);
        int32_t xa = sat_i32(floorf((p.x - s) * m.inv - 0.5f)) - 1;
        int32_t xb = sat_i32(floorf((p.x + s) * m.inv - 0.5f)) + 1;
        xa = xa > x0 ? xa : x0;
        xb = xb < x1 ? xb : x1;
        for (int32_t x = xa; x <= xb; x++) {
          const uint32_t code = voxel_code(m, p, has_n, nr, x, y, z);
          if (code != kEncSat)
            atomicMin(&lds[(x - ox) + kBrick * ((y - oy) + kBrick * (z - oz))], code);
        }
      }
    }
  }
  __syncthreads();

  // Exact live-tile mask of this brick.
  uint32_t mask = 0;
  for (int v = t; v < kBrickVox; v += kThreads) {
    const int lx = v & 15, ly = (v >> 4) & 15, lz = v >> 8;
    if (lds[v] != kEncSat) mask |= 1u << ((lx >> 3) | ((ly >> 3) << 1) | ((lz >> 3) << 2));
  }
  if (mask) atomicOr(&tflag, mask);
  __syncthreads();
  const uint32_t live = tflag;
  uint32_t write = live;
  for (int q = 0; q < 8; q++) {
    const int32_t ti = tile_index(q & 1, (q >> 1) & 1, q >> 2);
    if (ti < 0) {
      write &= ~(1u << q);
    } else if (tile_live[ti]) {
      write |= 1u << q;
    }
  }
  __syncthreads();
  if (write) {
    for (int v = t; v < kBrickVox; v += kThreads) {
      const int lx = v & 15, ly = (v >> 4) & 15, lz = v >> 8;
      const int q = (lx >> 3) | ((ly >> 3) << 1) | ((lz >> 3) << 2);
      if (!((write >> q) & 1u)) continue;
      const uint32_t x = (uint32_t)(ox + lx), y = (uint32_t)(oy + ly), z = (uint32_t)(oz + lz);
      const uint32_t gi = x + (uint32_t)m.r * (y + (uint32_t)m.r * z);
      const uint32_t code = lds[v];
      dist[m.voff + gi] = f2h((float)code * mk_.inv_enc - 1.0f);
      if (codes) codes[m.off + gi] = code;
    }
  }
  if (t < 8) {
    const int32_t ti = tile_index(t & 1, (t >> 1) & 1, t >> 2);
    if (ti >= 0) tile_live[ti] = (uint8_t)((live >> t) & 1u);
  }
}

}  // namespace

static void fill_mips(Splat* sp, MipsK* mk_) {
  const Volume* v = sp->vol;
  const double voxel0 = v->info.world_size / v->info.resolution;
  mk_->mips = sp->mips;
  mk_->inv_enc = (float)(1.0 / 32767.5);
  mk_->sat_half = f2h_host(1.0f);
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

static CloudK cloud_k(const Splat* sp, const uint32_t* order) {
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
  return c;
}

int splat_bake_ordered(Splat* sp, hipStream_t s, const uint32_t* order) {
  Ctx* ctx = sp->ctx;
  MipsK mk_{};
  fill_mips(sp, &mk_);
  const CloudK c = cloud_k(sp, order);
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
  // TILED
  int rc = launch_fill_u32(ctx, s, sp->brick_count, sp->brick_total, 0u);
  if (rc) return rc;
  if (c.count) {
    LaunchTimer t(ctx, s, "splat_bin_count");
    hipLaunchKernelGGL(k_bin<false>, dim3(pgrid), dim3(kThreads), 0, s, mk_, c, sp->vol->dev,
                       sp->brick_count, nullptr);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  {
    LaunchTimer t(ctx, s, "splat_brick_scan");
    hipLaunchKernelGGL(k_brick_scan, dim3(1), dim3(1024), 0, s, sp->brick_count,
                       (uint32_t)sp->brick_total, sp->brick_start, sp->brick_cursor,
                       sp->brick_total_dev);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  if (c.count) {
    LaunchTimer t(ctx, s, "splat_bin_fill");
    hipLaunchKernelGGL(k_bin<true>, dim3(pgrid), dim3(kThreads), 0, s, mk_, c, sp->vol->dev,
                       sp->brick_cursor, sp->brick_list);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  {
    LaunchTimer t(ctx, s, "splat_brick");
    hipLaunchKernelGGL(k_brick_splat, dim3((unsigned)sp->brick_total), dim3(kThreads), 0, s, mk_, c,
                       sp->vol->dev, sp->brick_start, sp->brick_count, sp->brick_list,
                       sp->tile_live, sp->vol->dist, sp->desc.keep_codes ? sp->grid : nullptr);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  sp->has_baked = true;
  return CLAY_OK;
}

int splat_bake(Splat* sp, hipStream_t s) { return splat_bake_ordered(sp, s, nullptr); }

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
