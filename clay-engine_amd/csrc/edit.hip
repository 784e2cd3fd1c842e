// edit.hip -- SdfEditor, TileGrid and MipRefiner for gfx950, and body centroids.
//
//   SdfEditor.flush (modify.ts:214-238):
//     k_grid_reset     zero every mip's tile counts, re-arm the per-mip list counters
//     k_edit_bin       per mip-0 tile: the edits whose bound reaches it, in edit order
//                      (at most `capacity`; tilegrid.ts:100-121 writeCell/setCellCount)
//     k_compact        live tiles of one mip -> dense list (tilegrid.ts:179-198)
//     k_edit_apply     per live mip-0 tile: the brush fold over the tile's edits, in place
//     then for every coarser mip m:
//     k_dilate         mip m live where a child in [2c-1, 2c+2]^3 is (tilegrid.ts:207-242)
//     k_compact        (mip m)
//     k_mip_refresh    MipRefiner (mips.ts:78-177): 12^3 downsample gather in LDS, two
//                      eikonal sweeps on the 26-neighbour chamfer stencil, tile core store
//
// The reference stages apply through a buffer and copies it back because WebGPU cannot
// read and write an rgba16float texture in one dispatch.  The volume here is a pair of
// f16 planes in HBM, and each voxel is read and written by the one thread that owns it,
// so the fold writes in place.  The live-tile lists are built with device counters and
// consumed by grid-stride workgroups reading those counters, so a flush needs no host
// readback and can be captured in a graph (nothing here is a memset).
//
// Erosion: k_contacts_to_edits turns a fluid's contact list into cut spheres on the
// device (the policy the reference leaves to its caller, game/fluid.ts:86-95).
//
// Body centroids (pbd.ts:240-266): one 256-lane workgroup per body sums its particles
// lane-strided, then a pairwise tree; in f32 this is the reference's reduction, in f64
// the same order at double precision.
//
// A volume with one instance per scene (a batch's worlds) gets one editor state per
// scene: every kernel takes its scene from blockIdx.y and offsets the edit list, tile
// counts, list, counters and planes by it, so all scenes flush in the same launches.
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kThreads = 256;
constexpr int kE = 12;               // gathered edge: tile + 2-voxel border each side
constexpr int kB = 2;                // border = number of eikonal sweeps
constexpr int kCells = kE * kE * kE;

struct LevelK {
  int32_t r;        // voxels per axis
  int32_t tr;       // tiles per axis
  uint32_t voff;    // offset of the level in the planes
  uint32_t coff;    // offset of the level in the tile counts / list
};

struct EditK {
  float org[3];
  float voxel, tsize, thalf, bw;
  int32_t capacity;
  int32_t max_edits;
  uint32_t total_cells;   // tile counts / list entries per scene
  uint32_t tiles0;        // mip-0 tiles per scene
  uint32_t vps;           // plane entries per scene
  int32_t mips;
  LevelK lv[CLAY_MAX_MIPS];
};


__device__ __forceinline__ uint32_t pack_tile(uint32_t x, uint32_t y, uint32_t z) {
  return x | (y << 10) | (z << 20);
}
__device__ __forceinline__ void unpack_tile(uint32_t v, int32_t t[3]) {
  t[0] = (int32_t)(v & 1023u);
  t[1] = (int32_t)((v >> 10) & 1023u);
  t[2] = (int32_t)((v >> 20) & 1023u);
}

__global__ void k_grid_reset(uint32_t* __restrict__ counts, uint32_t total, uint32_t* __restrict__ args,
                             int mips) {
  const uint32_t sc = blockIdx.y;
  counts += (size_t)sc * total;
  args += (size_t)sc * mips * 4u;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) counts[i] = 0;
  if (i < (uint32_t)mips * 4u) {
    const uint32_t lane = i & 3u;
    args[i] = (lane == 1u || lane == 2u) ? 1u : 0u;
  }
}

__global__ void k_edit_bin(EditK k, const clay_brush* __restrict__ edits, const uint32_t* __restrict__ nedit,
                           uint32_t* __restrict__ counts, uint32_t* __restrict__ ids) {
  const uint32_t sc = blockIdx.y;
  edits += (size_t)sc * k.max_edits;
  nedit += sc;
  counts += (size_t)sc * k.total_cells;
  ids += (size_t)sc * k.tiles0 * k.capacity;
  const LevelK& l = k.lv[0];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint32_t)(l.tr * l.tr * l.tr)) return;
  const int32_t tx = (int32_t)(i % (uint32_t)l.tr), ty = (int32_t)((i / (uint32_t)l.tr) % (uint32_t)l.tr),
                tz = (int32_t)(i / ((uint32_t)l.tr * (uint32_t)l.tr));
  const f3 tc = mk(k.org[0] + ((float)tx * k.tsize + k.thalf), k.org[1] + ((float)ty * k.tsize + k.thalf),
                   k.org[2] + ((float)tz * k.tsize + k.thalf));
  const uint32_t ne = *nedit < (uint32_t)k.max_edits ? *nedit : (uint32_t)k.max_edits;
  uint32_t n = 0;
  for (uint32_t b = 0; b < ne; b++) {
    const clay_brush& e = edits[b];
    const f3 nr = max30(mk(fabsf(e.pos[0] - tc.x) - k.thalf, fabsf(e.pos[1] - tc.y) - k.thalf,
                           fabsf(e.pos[2] - tc.z) - k.thalf));
    if (len3(nr) < e.bound + k.bw) {
      if (n < (uint32_t)k.capacity) ids[(size_t)i * k.capacity + n] = b;
      n++;
    }
  }
  if (n > 0) counts[l.coff + i] = n < (uint32_t)k.capacity ? n : (uint32_t)k.capacity;
}

__global__ void k_compact(EditK k, int m, const uint32_t* __restrict__ counts, uint32_t* __restrict__ list,
                          uint32_t* __restrict__ args) {
  const uint32_t sc = blockIdx.y;
  counts += (size_t)sc * k.total_cells;
  list += (size_t)sc * k.total_cells;
  args += (size_t)sc * k.mips * 4u;
  const LevelK& l = k.lv[m];
  const uint32_t cells = (uint32_t)(l.tr * l.tr * l.tr);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cells || counts[l.coff + i] == 0) return;
  const uint32_t slot = atomicAdd(&args[m * 4], 1u);
  list[l.coff + slot] = pack_tile(i % (uint32_t)l.tr, (i / (uint32_t)l.tr) % (uint32_t)l.tr,
                                  i / ((uint32_t)l.tr * (uint32_t)l.tr));
}

__global__ void k_dilate(EditK k, int m, uint32_t* __restrict__ counts) {
  counts += (size_t)blockIdx.y * k.total_cells;
  const LevelK& l = k.lv[m];
  const LevelK& p = k.lv[m - 1];
  const uint32_t cells = (uint32_t)(l.tr * l.tr * l.tr);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  const int32_t bx = (int32_t)(i % (uint32_t)l.tr) * 2, by = (int32_t)((i / (uint32_t)l.tr) % (uint32_t)l.tr) * 2,
                bz = (int32_t)(i / ((uint32_t)l.tr * (uint32_t)l.tr)) * 2;
  bool live = false;
  for (int dz = -1; dz <= 2; dz++)
    for (int dy = -1; dy <= 2; dy++)
      for (int dx = -1; dx <= 2; dx++) {
        const int32_t qx = bx + dx, qy = by + dy, qz = bz + dz;
        if (qx < 0 || qy < 0 || qz < 0 || qx >= p.tr || qy >= p.tr || qz >= p.tr) continue;
        if (counts[p.coff + (uint32_t)qx + (uint32_t)p.tr * ((uint32_t)qy + (uint32_t)p.tr * (uint32_t)qz)] != 0)
          live = true;
      }
  if (live) counts[l.coff + i] = 1;
}

// One workgroup per live mip-0 tile (grid-stride over the device count); each
// thread owns two voxels (x, y, 2z) and (x, y, 2z + 1) as the reference's 8x8x4
// workgroup does.
__global__ void __launch_bounds__(kThreads)
    k_edit_apply(EditK k, const clay_brush* __restrict__ edits, const uint32_t* __restrict__ counts,
                 const uint32_t* __restrict__ ids, const uint32_t* __restrict__ list,
                 const uint32_t* __restrict__ args, uint16_t* __restrict__ dist, uint16_t* __restrict__ mat) {
  const uint32_t sc = blockIdx.y;
  edits += (size_t)sc * k.max_edits;
  counts += (size_t)sc * k.total_cells;
  ids += (size_t)sc * k.tiles0 * k.capacity;
  list += (size_t)sc * k.total_cells;
  args += (size_t)sc * k.mips * 4u;
  dist += (size_t)sc * k.vps;
  mat += (size_t)sc * k.vps;
  const LevelK& l = k.lv[0];
  const uint32_t live = args[0];
  const int t = threadIdx.x;
  const int lx = t & 7, ly = (t >> 3) & 7, lz2 = t >> 6;
  for (uint32_t slot = blockIdx.x; slot < live; slot += gridDim.x) {
    int32_t tile[3];
    unpack_tile(list[l.coff + slot], tile);
    const uint32_t cell = (uint32_t)tile[0] + (uint32_t)l.tr * ((uint32_t)tile[1] + (uint32_t)l.tr * (uint32_t)tile[2]);
    const uint32_t n = counts[l.coff + cell];
    for (int half = 0; half < 2; half++) {
      const int32_t x = tile[0] * kTile + lx, y = tile[1] * kTile + ly, z = tile[2] * kTile + lz2 * 2 + half;
      const f3 p = mk(k.org[0] + ((float)x + 0.5f) * k.voxel, k.org[1] + ((float)y + 0.5f) * k.voxel,
                      k.org[2] + ((float)z + 0.5f) * k.voxel);
      const size_t idx = l.voff + (size_t)x + (size_t)l.r * ((size_t)y + (size_t)l.r * (size_t)z);
      float acc[2] = {h2f(dist[idx]) * k.bw, h2f(mat[idx])};
      for (uint32_t j = 0; j < n; j++) apply_brush(acc, edits[ids[(size_t)cell * k.capacity + j]], p);
      dist[idx] = f2h(fclamp_s(acc[0] / k.bw, -1.0f, 1.0f));
      mat[idx] = f2h(acc[1]);
    }
  }
}

// The trilinear filter at the corner shared by the 8 children of coarse voxel c:
// every weight is one half.
__device__ __forceinline__ float child_average(const uint16_t* __restrict__ plane, int32_t fr, int32_t cx,
                                               int32_t cy, int32_t cz) {
  const int32_t x0 = 2 * cx, y0 = 2 * cy, z0 = 2 * cz;
  auto tx = [&](int32_t x, int32_t y, int32_t z) {
    return h2f(plane[(size_t)x + (size_t)fr * ((size_t)y + (size_t)fr * (size_t)z)]);
  };
  const float h = 0.5f;
  const float c00 = fmix(tx(x0, y0, z0), tx(x0 + 1, y0, z0), h);
  const float c10 = fmix(tx(x0, y0 + 1, z0), tx(x0 + 1, y0 + 1, z0), h);
  const float c01 = fmix(tx(x0, y0, z0 + 1), tx(x0 + 1, y0, z0 + 1), h);
  const float c11 = fmix(tx(x0, y0 + 1, z0 + 1), tx(x0 + 1, y0 + 1, z0 + 1), h);
  return fmix(fmix(c00, c10, h), fmix(c01, c11, h), h);
}

__global__ void __launch_bounds__(kThreads)
    k_mip_refresh(EditK k, int m, float step, const uint32_t* __restrict__ list, const uint32_t* __restrict__ args,
                  uint16_t* __restrict__ dist, uint16_t* __restrict__ mat) {
  __shared__ float g[2][kCells];
  const uint32_t sc = blockIdx.y;
  list += (size_t)sc * k.total_cells;
  args += (size_t)sc * k.mips * 4u;
  dist += (size_t)sc * k.vps;
  mat += (size_t)sc * k.vps;
  const LevelK& l = k.lv[m];
  const LevelK& f = k.lv[m - 1];
  const uint32_t live = args[m * 4];
  const int t = threadIdx.x;
  float w[27];
#pragma unroll
  for (int q = 0; q < 27; q++) {
    const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
    w[q] = len3(mk((float)dx, (float)dy, (float)dz)) * step;
  }
  for (uint32_t slot = blockIdx.x; slot < live; slot += gridDim.x) {
    int32_t tile[3];
    unpack_tile(list[l.coff + slot], tile);
    const int32_t ox = tile[0] * kTile, oy = tile[1] * kTile, oz = tile[2] * kTile;
    for (int c = t; c < kCells; c += kThreads) {
      int32_t gx = ox + c % kE - kB, gy = oy + (c / kE) % kE - kB, gz = oz + c / (kE * kE) - kB;
      gx = gx < 0 ? 0 : (gx > l.r - 1 ? l.r - 1 : gx);
      gy = gy < 0 ? 0 : (gy > l.r - 1 ? l.r - 1 : gy);
      gz = gz < 0 ? 0 : (gz > l.r - 1 ? l.r - 1 : gz);
      g[0][c] = child_average(dist + f.voff, f.r, gx, gy, gz) * 0.5f;
    }
    for (int sweep = 0; sweep < kB; sweep++) {
      const float* src = g[sweep & 1];
      float* dst = g[(sweep + 1) & 1];
      const int margin = sweep + 1;
      __syncthreads();
      for (int c = t; c < kCells; c += kThreads) {
        const int x = c % kE, y = (c / kE) % kE, z = c / (kE * kE);
        float v = src[c];
        if (x >= margin && y >= margin && z >= margin && x < kE - margin && y < kE - margin && z < kE - margin) {
          float lo = -1e9f, hi = 1e9f;
#pragma unroll
          for (int q = 0; q < 27; q++) {
            if (q == 13) continue;
            const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
            const float nv = src[(x + dx) + kE * ((y + dy) + kE * (z + dz))];
            hi = fmin_s(hi, nv + w[q]);
            lo = fmax_s(lo, nv - w[q]);
          }
          v = fclamp_s(v, lo, hi);
        }
        dst[c] = v;
      }
    }
    __syncthreads();
    const float* out = g[kB & 1];
    const int lx = t & 7, ly = (t >> 3) & 7, lz2 = t >> 6;
    for (int half = 0; half < 2; half++) {
      const int lz = lz2 * 2 + half;
      const int32_t x = ox + lx, y = oy + ly, z = oz + lz;
      const size_t idx = l.voff + (size_t)x + (size_t)l.r * ((size_t)y + (size_t)l.r * (size_t)z);
      dist[idx] = f2h(fclamp_s(out[(lx + kB) + kE * ((ly + kB) + kE * (lz + kB))], -1.0f, 1.0f));
      mat[idx] = f2h(child_average(mat + f.voff, f.r, x, y, z));
    }
    __syncthreads();
  }
}

// Contacts -> cut spheres appended to the device edit list (at most max_edits).
__global__ void k_contacts_to_edits(const uint32_t* __restrict__ contacts, uint32_t contact_stride,
                                    float min_speed, float radius, float blend, float bound, uint32_t max_edits,
                                    clay_brush* __restrict__ edits, uint32_t* __restrict__ nedit) {
  if (threadIdx.x != 0) return;
  const uint32_t sc = blockIdx.x;
  contacts += (size_t)sc * contact_stride;
  edits += (size_t)sc * max_edits;
  nedit += sc;
  const uint32_t raw = contacts[0];
  const uint32_t n = raw < (uint32_t)CLAY_CONTACTS ? raw : (uint32_t)CLAY_CONTACTS;
  const float4* pts = reinterpret_cast<const float4*>(contacts + 4);
  uint32_t k = *nedit;
  for (uint32_t i = 0; i < n && k < max_edits; i++) {
    const float4 c = pts[i];
    if (!(c.w > min_speed)) continue;
    clay_brush b{};
    b.pos[0] = c.x;
    b.pos[1] = c.y;
    b.pos[2] = c.z;
    b.scale = 1.0f;
    b.rot[3] = 1.0f;
    b.size[0] = radius;
    b.size[1] = radius;
    b.size[2] = radius;
    b.blend = blend;
    b.bound = bound;
    b.mask = -1.0f;
    b.kind = 0;
    b.op = 1;
    edits[k++] = b;
  }
  *nedit = k;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_body_centroids(const uint32_t* __restrict__ first_count, const float* __restrict__ px,
                     const float* __restrict__ py, const float* __restrict__ pz, const float* __restrict__ rx,
                     const float* __restrict__ ry, const float* __restrict__ rz, double* __restrict__ out) {
  __shared__ T sh[6][kThreads];
  const uint32_t b = blockIdx.x;
  const uint32_t first = first_count[2 * b], count = first_count[2 * b + 1];
  const int li = threadIdx.x;
  T a[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t t = (uint32_t)li; t < count; t += kThreads) {
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
  for (int q = 0; q < 6; q++) sh[q][li] = a[q];
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (li < s)
      for (int q = 0; q < 6; q++) sh[q][li] = sh[q][li] + sh[q][li + s];
    __syncthreads();
  }
  if (li < 6) {
    const T fc = (T)count;
    const T inv = (T)1 / (fc < (T)1 ? (T)1 : fc);
    out[6 * b + li] = (double)(sh[li][0] * inv);
  }
}

inline unsigned groups(uint32_t n, int block) { return (n + block - 1) / block; }

EditK edit_k(const Editor* e) {
  const Volume* v = e->vol;
  EditK k{};
  const double voxel = v->info.voxel_size;
  for (int a = 0; a < 3; a++) k.org[a] = (float)v->origin[a];
  k.voxel = (float)voxel;
  k.tsize = (float)(voxel * kTile);
  k.thalf = (float)(voxel * kTile / 2);
  k.bw = (float)(v->info.band * voxel);
  k.capacity = e->capacity;
  k.max_edits = e->max_edits;
  k.total_cells = (uint32_t)v->info.total_cells;
  k.tiles0 = (uint32_t)v->info.tile_res[0] * v->info.tile_res[0] * v->info.tile_res[0];
  k.vps = (uint32_t)v->total_voxels;
  k.mips = v->info.mip_levels;
  for (int m = 0; m < v->info.mip_levels; m++) {
    k.lv[m].r = v->mip_res(m);
    k.lv[m].tr = v->info.tile_res[m];
    k.lv[m].voff = (uint32_t)v->voxel_off[m];
    k.lv[m].coff = (uint32_t)v->info.cell_offsets[m];
  }
  return k;
}

}  // namespace

int editor_flush_device(Editor* e, hipStream_t s, bool all_tiles) {
  Ctx* c = e->ctx;
  const Volume* v = e->vol;
  const EditK k = edit_k(e);
  const int mips = v->info.mip_levels;
  const uint32_t total = (uint32_t)v->info.total_cells;
  const unsigned S = (unsigned)v->instances;
  {
    LaunchTimer t(c, s, "edit_grid_reset");
    hipLaunchKernelGGL(k_grid_reset, dim3(groups(std::max<uint32_t>(total, (uint32_t)mips * 4), 64), S), dim3(64), 0,
                       s, e->counts, total, e->args, mips);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  const uint32_t tiles0 = (uint32_t)(k.lv[0].tr * k.lv[0].tr * k.lv[0].tr);
  if (all_tiles) {
    int rc = launch_fill_u32(c, s, e->counts, (size_t)total * S, 1u);
    if (rc) return rc;
  } else {
    LaunchTimer t(c, s, "edit_bin");
    hipLaunchKernelGGL(k_edit_bin, dim3(groups(tiles0, 64), S), dim3(64), 0, s, k, e->edits_dev, e->nedit_dev,
                       e->counts, e->ids);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  const unsigned tile_groups =
      (unsigned)std::max<uint32_t>(1, std::min<uint32_t>(tiles0, (uint32_t)c->persistent_blocks / S));
  {
    LaunchTimer t(c, s, "edit_compact");
    hipLaunchKernelGGL(k_compact, dim3(groups(tiles0, 64), S), dim3(64), 0, s, k, 0, e->counts, e->list, e->args);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  if (!all_tiles) {
    LaunchTimer t(c, s, "edit_apply");
    hipLaunchKernelGGL(k_edit_apply, dim3(tile_groups, S), dim3(kThreads), 0, s, k, e->edits_dev, e->counts, e->ids,
                       e->list, e->args, v->dist, v->mat);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  const float step = (float)((1.0 / v->info.band) * e->safety);
  for (int m = 1; m < mips; m++) {
    const uint32_t cells = (uint32_t)(k.lv[m].tr * k.lv[m].tr * k.lv[m].tr);
    if (!all_tiles) {
      LaunchTimer t(c, s, "edit_dilate");
      hipLaunchKernelGGL(k_dilate, dim3(groups(cells, 64), S), dim3(64), 0, s, k, m, e->counts);
      CLAY_HIP_CHECK(hipGetLastError());
    }
    {
      LaunchTimer t(c, s, "edit_compact");
      hipLaunchKernelGGL(k_compact, dim3(groups(cells, 64), S), dim3(64), 0, s, k, m, e->counts, e->list, e->args);
      CLAY_HIP_CHECK(hipGetLastError());
    }
    {
      LaunchTimer t(c, s, "mip_refresh");
      const unsigned g = (unsigned)std::max<uint32_t>(1, std::min<uint32_t>(cells, (uint32_t)c->persistent_blocks / S));
      hipLaunchKernelGGL(k_mip_refresh, dim3(g, S), dim3(kThreads), 0, s, k, m, step, e->list, e->args, v->dist,
                         v->mat);
      CLAY_HIP_CHECK(hipGetLastError());
    }
  }
  return CLAY_OK;
}

int launch_contacts_to_edits(Ctx* c, hipStream_t s, Editor* e, const uint32_t* contacts, uint32_t contact_stride,
                             uint32_t scenes, uint32_t first_scene, double min_speed, double radius, double blend) {
  LaunchTimer t(c, s, "edit_from_contacts");
  hipLaunchKernelGGL(k_contacts_to_edits, dim3(scenes), dim3(64), 0, s, contacts, contact_stride, (float)min_speed,
                     (float)radius, (float)blend, (float)(radius + blend), (uint32_t)e->max_edits,
                     e->edits_dev + (size_t)first_scene * e->max_edits, e->nedit_dev + first_scene);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_body_centroids(Ctx* c, hipStream_t s, bool f64, uint32_t nbodies, const uint32_t* first_count,
                          const float* const* pos, const float* const* rest, double* out) {
  if (nbodies == 0) return CLAY_OK;
  LaunchTimer t(c, s, f64 ? "body_centroids_fp64" : "body_centroids_f32");
  const float* r0 = rest ? rest[0] : nullptr;
  const float* r1 = rest ? rest[1] : nullptr;
  const float* r2 = rest ? rest[2] : nullptr;
  if (f64)
    hipLaunchKernelGGL(k_body_centroids<double>, dim3(nbodies), dim3(kThreads), 0, s, first_count, pos[0], pos[1],
                       pos[2], r0, r1, r2, out);
  else
    hipLaunchKernelGGL(k_body_centroids<float>, dim3(nbodies), dim3(kThreads), 0, s, first_count, pos[0], pos[1],
                       pos[2], r0, r1, r2, out);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

}  // namespace clay
