// wire.hip -- the baked volume as a tile-sparse stream of rgba16float texels.
//
// A baked SdfVolume is, per mip, an rgba16float 3D texture whose texel is
// (distance / band, material, 0, 1) (lib/src/field/volume.ts:77-90).  A stream holds
// one instance, every mip, as 8^3 tiles (mip-major, then tile z, y, x):
//
//   header   96 bytes (clay_wire_header)
//   table    one u32 per tile: a constant tile's texel as (material16 << 16) | dist16,
//            or a stored tile's index into the payload
//   bitmap   one bit per tile, set for stored tiles (u32 words, tile t at word t/32,
//            bit t%32)
//   payload  from a 256-byte aligned offset: 4096 bytes per stored tile, its 512
//            rgba16float texels with x fastest, then y, then z
//
// A tile is constant when its 512 (distance, material) pairs are the same bits; far
// from the surface a band-limited field is saturated, so most tiles are.  Stored tiles
// keep the reference's texel as it is, so a consumer uploads each one with a single
// texture write.
//
// Each pass takes one workgroup per row of tiles (fixed mip, tile y and z): its lanes
// walk the row's 64 texel lines, 16 bytes of each plane (one tile line) per lane and
// line, so a wave reads whole lines.  pack = classify (constant or not, per tile,
// which sets the bitmap), scan of the stored-tile counts per row, write (table,
// payload through LDS, header); unpack = one pass back into the planes.
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kWireGroup = 256;
constexpr int kMaxRowTiles = 128;  // resolution <= 1024
constexpr uint32_t kWireMagic = 0x42564C43u;  // "CLVB"
static_assert(sizeof(clay_wire_header) == 96, "the stream header is 96 bytes");

// Where one row of tiles lives.
struct RowMap {
  uint32_t mips;
  uint32_t row_start[CLAY_MAX_MIPS + 1];   // first row of each mip (rows = tile_res^2)
  uint32_t tile_start[CLAY_MAX_MIPS + 1];  // first tile of each mip
  uint32_t tile_res[CLAY_MAX_MIPS];
  uint64_t voxel_off[CLAY_MAX_MIPS];       // plane offset of each mip (elements)
};

struct RowPos {
  uint32_t m, ty, tz, tr, res, first_tile;
  uint64_t plane;  // element offset of the mip in the planes
};

__device__ __forceinline__ RowPos row_pos(const RowMap& rm, uint32_t row) {
  RowPos r{};
  uint32_t m = 0;
  while (m + 1 < rm.mips && row >= rm.row_start[m + 1]) m++;
  const uint32_t local = row - rm.row_start[m];
  r.m = m;
  r.tr = rm.tile_res[m];
  r.ty = local % r.tr;
  r.tz = local / r.tr;
  r.res = r.tr * 8u;
  r.first_tile = rm.tile_start[m] + local * r.tr;
  r.plane = rm.voxel_off[m];
  return r;
}

// Element offset of texel (8 tile, y, z) of a row line: line = dy + 8 dz.
__device__ __forceinline__ uint64_t line_off(const RowPos& r, uint32_t tile, uint32_t line) {
  const uint64_t y = r.ty * 8u + (line & 7u), z = r.tz * 8u + (line >> 3);
  return r.plane + (uint64_t)tile * 8u + (uint64_t)r.res * (y + (uint64_t)r.res * z);
}

__device__ __forceinline__ uint32_t texel_of(uint32_t d2, uint32_t m2, int hi) {
  return hi ? ((d2 >> 16) | (m2 & 0xFFFF0000u)) : ((d2 & 0xFFFFu) | (m2 << 16));
}

__global__ void __launch_bounds__(kWireGroup)
    k_wire_classify(RowMap rm, const uint16_t* __restrict__ dist, const uint16_t* __restrict__ mat,
                    uint32_t* __restrict__ table, uint32_t* __restrict__ bitmap, uint32_t* __restrict__ row_count) {
  __shared__ uint32_t ref[kMaxRowTiles];
  __shared__ uint32_t same[kMaxRowTiles];
  const RowPos r = row_pos(rm, blockIdx.x);
  const uint32_t t = threadIdx.x;
  if (t < r.tr) {
    const uint64_t o = line_off(r, t, 0);
    ref[t] = (uint32_t)dist[o] | ((uint32_t)mat[o] << 16);
    same[t] = 1u;
  }
  __syncthreads();
  const uint32_t units = 64u * r.tr;
  for (uint32_t u = t; u < units; u += kWireGroup) {
    const uint32_t line = u / r.tr, tile = u % r.tr;
    const uint64_t o = line_off(r, tile, line);
    const uint4 d = *reinterpret_cast<const uint4*>(dist + o);
    const uint4 m = *reinterpret_cast<const uint4*>(mat + o);
    const uint32_t want = ref[tile];
    const uint32_t w2d = (want & 0xFFFFu) * 0x10001u, w2m = (want >> 16) * 0x10001u;
    const bool eq = d.x == w2d && d.y == w2d && d.z == w2d && d.w == w2d && m.x == w2m && m.y == w2m &&
                    m.z == w2m && m.w == w2m;
    if (!eq) same[tile] = 0u;  // every writer stores the same value
  }
  __syncthreads();
  if (t < r.tr) {
    const uint32_t g = r.first_tile + t;
    if (same[t])
      table[g] = ref[t];
    else
      atomicOr(bitmap + g / 32u, 1u << (g % 32u));  // stored; its index is written with the payload
  }
  __syncthreads();
  if (t == 0) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < r.tr; i++) n += same[i] ? 0u : 1u;
    row_count[blockIdx.x] = n;
  }
}

__global__ void __launch_bounds__(1024) k_wire_scan(const uint32_t* __restrict__ cnt, uint32_t* __restrict__ off,
                                                     uint32_t n, uint32_t* __restrict__ total) {
  __shared__ uint32_t buf[1024];
  __shared__ uint32_t carry;
  const uint32_t t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += 1024) {
    const uint32_t v = base + t < n ? cnt[base + t] : 0u;
    buf[t] = v;
    __syncthreads();
    for (uint32_t s = 1; s < 1024; s <<= 1) {
      const uint32_t add = t >= s ? buf[t - s] : 0u;
      __syncthreads();
      buf[t] += add;
      __syncthreads();
    }
    if (base + t < n) off[base + t] = carry + buf[t] - v;
    __syncthreads();
    if (t == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

// Stream sections (bytes), from the tile count.
struct Layout {
  uint64_t table, bitmap, payload;
};
__host__ __device__ inline Layout wire_layout(uint32_t tiles) {
  Layout l;
  l.table = sizeof(clay_wire_header);
  l.bitmap = l.table + 4ull * tiles;
  l.payload = (l.bitmap + 4ull * ((tiles + 31u) / 32u) + 255ull) & ~255ull;
  return l;
}

// The stored tiles of a row have consecutive payload indices, so their payload is
// one contiguous run.  Batches of up to kBatch tiles go through LDS: lanes read
// 16-byte tile lines of both planes (tile fastest, so a wave reads along a texel
// line), expand them to rgba16float in LDS, then the workgroup writes the batch out
// as contiguous 16-byte chunks.  Each staged tile is padded by 16 bytes so the lanes
// of one line land in different LDS banks.
constexpr int kBatch = 8;
constexpr int kTileChunks = 256;  // 4096 bytes of payload per tile / 16

__global__ void __launch_bounds__(kWireGroup)
    k_wire_write(RowMap rm, clay_wire_header hdr, const uint16_t* __restrict__ dist, const uint16_t* __restrict__ mat,
                 const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ stored_total,
                 uint8_t* __restrict__ out) {
  __shared__ uint32_t list[kMaxRowTiles];
  __shared__ uint32_t nstored;
  __shared__ uint4 stage[kBatch * (kTileChunks + 1)];
  const RowPos r = row_pos(rm, blockIdx.x);
  const Layout L = wire_layout(hdr.tiles);
  uint32_t* table = reinterpret_cast<uint32_t*>(out + L.table);
  const uint32_t* bitmap = reinterpret_cast<const uint32_t*>(out + L.bitmap);
  const uint32_t t = threadIdx.x;
  const uint32_t first = row_off[blockIdx.x];
  // Rank the row's stored tiles in order: one lane per tile, wave ballots, per-wave counts.
  __shared__ uint32_t wave_count[kWireGroup / 64];
  uint32_t bit = 0;
  if (t < r.tr) {
    const uint32_t g = r.first_tile + t;
    bit = (bitmap[g / 32u] >> (g % 32u)) & 1u;
  }
  const uint32_t lane = t & 63u, wave = t >> 6;
  const uint64_t ballot = __ballot(bit);
  if (lane == 0) wave_count[wave] = (uint32_t)__popcll(ballot);
  __syncthreads();
  uint32_t rank = (uint32_t)__popcll(ballot & (lane ? (~0ull >> (64u - lane)) : 0ull));
  for (uint32_t w = 0; w < wave; w++) rank += wave_count[w];
  if (bit) {
    list[rank] = t;
    table[r.first_tile + t] = first + rank;
  }
  if (t == 0) {
    uint32_t n = 0;
    for (uint32_t w = 0; w < kWireGroup / 64; w++) n += wave_count[w];
    nstored = n;
  }
  __syncthreads();
  const uint32_t ns = nstored;
  uint4* payload = reinterpret_cast<uint4*>(out + L.payload);
  for (uint32_t b0 = 0; b0 < ns; b0 += kBatch) {
    const uint32_t nb = ns - b0 < (uint32_t)kBatch ? ns - b0 : (uint32_t)kBatch;
    for (uint32_t u = t; u < 64u * nb; u += kWireGroup) {
      const uint32_t j = u % nb, line = u / nb;
      const uint64_t o = line_off(r, list[b0 + j], line);
      const uint4 d = *reinterpret_cast<const uint4*>(dist + o);
      const uint4 m = *reinterpret_cast<const uint4*>(mat + o);
      const uint32_t dd[4] = {d.x, d.y, d.z, d.w}, mm[4] = {m.x, m.y, m.z, m.w};
      uint4* st = stage + j * (kTileChunks + 1) + line * 4u;
      // Two texels per 16 bytes: (d, m) then (0, 1.0) as f16 bits.
#pragma unroll
      for (int q = 0; q < 4; q++)
        st[q] = make_uint4(texel_of(dd[q], mm[q], 0), 0x3C000000u, texel_of(dd[q], mm[q], 1), 0x3C000000u);
    }
    __syncthreads();
    uint4* dst = payload + (uint64_t)(first + b0) * kTileChunks;
    for (uint32_t c = t; c < nb * (uint32_t)kTileChunks; c += kWireGroup)
      dst[c] = stage[(c / kTileChunks) * (kTileChunks + 1) + c % kTileChunks];
    __syncthreads();
  }
  if (blockIdx.x == 0 && t == 0) {
    clay_wire_header h = hdr;
    h.stored_tiles = *stored_total;
    h.table_offset = L.table;
    h.bitmap_offset = L.bitmap;
    h.payload_offset = L.payload;
    h.total_bytes = L.payload + 4096ull * h.stored_tiles;
    *reinterpret_cast<clay_wire_header*>(out) = h;
  }
}

// A stored tile whose table word is not below `stored` sets *bad and is skipped.
// Constant tiles are written line by line across the row (tile fastest); stored
// tiles go in batches of kBatch: their 4096-byte payloads are read whole into LDS
// (each tile contiguous, wherever its index points), then written out as 16-byte
// tile lines of both planes.
__global__ void __launch_bounds__(kWireGroup)
    k_wire_unpack(RowMap rm, uint32_t tiles, uint32_t stored_tiles, const uint8_t* __restrict__ in,
                  uint16_t* __restrict__ dist, uint16_t* __restrict__ mat, uint32_t* __restrict__ bad) {
  __shared__ uint32_t word[kMaxRowTiles];
  __shared__ uint32_t kind[kMaxRowTiles];  // 0 constant, 1 stored, 2 stored with a bad index
  __shared__ uint32_t list[kMaxRowTiles];
  __shared__ uint32_t wave_count[kWireGroup / 64];
  __shared__ uint4 stage[kBatch * (kTileChunks + 1)];
  const RowPos r = row_pos(rm, blockIdx.x);
  const Layout L = wire_layout(tiles);
  const uint32_t* table = reinterpret_cast<const uint32_t*>(in + L.table);
  const uint32_t* bitmap = reinterpret_cast<const uint32_t*>(in + L.bitmap);
  const uint4* payload = reinterpret_cast<const uint4*>(in + L.payload);
  const uint32_t t = threadIdx.x;
  uint32_t stored = 0;
  if (t < r.tr) {
    const uint32_t g = r.first_tile + t;
    word[t] = table[g];
    uint32_t k = (bitmap[g / 32u] >> (g % 32u)) & 1u;
    if (k && word[t] >= stored_tiles) {
      k = 2u;
      atomicOr(bad, 1u);
    }
    kind[t] = k;
    stored = k == 1u;
  }
  const uint32_t lane = t & 63u, wave = t >> 6;
  const uint64_t ballot = __ballot(stored);
  if (lane == 0) wave_count[wave] = (uint32_t)__popcll(ballot);
  __syncthreads();
  uint32_t rank = (uint32_t)__popcll(ballot & (lane ? (~0ull >> (64u - lane)) : 0ull));
  uint32_t ns = 0;
  for (uint32_t w = 0; w < kWireGroup / 64; w++) {
    if (w < wave) rank += wave_count[w];
    ns += wave_count[w];
  }
  if (stored) list[rank] = t;
  __syncthreads();
  const uint32_t units = 64u * r.tr;
  for (uint32_t u = t; u < units; u += kWireGroup) {
    const uint32_t line = u / r.tr, tile = u % r.tr;
    if (kind[tile] != 0u) continue;
    const uint32_t w = word[tile];
    const uint32_t d2 = (w & 0xFFFFu) * 0x10001u, m2 = (w >> 16) * 0x10001u;
    const uint64_t o = line_off(r, tile, line);
    *reinterpret_cast<uint4*>(dist + o) = make_uint4(d2, d2, d2, d2);
    *reinterpret_cast<uint4*>(mat + o) = make_uint4(m2, m2, m2, m2);
  }
  for (uint32_t b0 = 0; b0 < ns; b0 += kBatch) {
    const uint32_t nb = ns - b0 < (uint32_t)kBatch ? ns - b0 : (uint32_t)kBatch;
    for (uint32_t c = t; c < nb * (uint32_t)kTileChunks; c += kWireGroup) {
      const uint32_t j = c / kTileChunks;
      stage[j * (kTileChunks + 1) + c % kTileChunks] =
          payload[(uint64_t)word[list[b0 + j]] * kTileChunks + c % kTileChunks];
    }
    __syncthreads();
    for (uint32_t u = t; u < 64u * nb; u += kWireGroup) {
      const uint32_t j = u % nb, line = u / nb;
      const uint4* src = stage + j * (kTileChunks + 1) + line * 4u;
      uint32_t dd[4], mm[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint4 sv = src[q];  // (d0 m0) (b0 a0) (d1 m1) (b1 a1)
        dd[q] = (sv.x & 0xFFFFu) | (sv.z << 16);
        mm[q] = (sv.x >> 16) | (sv.z & 0xFFFF0000u);
      }
      const uint64_t o = line_off(r, list[b0 + j], line);
      *reinterpret_cast<uint4*>(dist + o) = make_uint4(dd[0], dd[1], dd[2], dd[3]);
      *reinterpret_cast<uint4*>(mat + o) = make_uint4(mm[0], mm[1], mm[2], mm[3]);
    }
    __syncthreads();
  }
}

RowMap row_map(const Volume* v, uint32_t* rows, uint32_t* tiles) {
  RowMap rm{};
  rm.mips = (uint32_t)v->info.mip_levels;
  uint32_t row = 0, tile = 0;
  for (uint32_t m = 0; m < rm.mips; m++) {
    const uint32_t tr = (uint32_t)v->mip_res((int)m) / 8u;
    rm.row_start[m] = row;
    rm.tile_start[m] = tile;
    rm.tile_res[m] = tr;
    rm.voxel_off[m] = v->voxel_off[m];
    row += tr * tr;
    tile += tr * tr * tr;
  }
  rm.row_start[rm.mips] = row;
  rm.tile_start[rm.mips] = tile;
  *rows = row;
  *tiles = tile;
  return rm;
}

int check_volume(const Volume* v, uint32_t instance, const char* who) {
  if ((int)instance >= v->instances) return fail(CLAY_ERR_INVALID, std::string(who) + ": bad instance");
  if (v->info.resolution > 8 * kMaxRowTiles)
    return fail(CLAY_ERR_INVALID, std::string(who) + ": resolution above " + std::to_string(8 * kMaxRowTiles));
  if (v->info.mip_levels > CLAY_MAX_MIPS) return fail(CLAY_ERR_INVALID, std::string(who) + ": too many mips");
  return CLAY_OK;
}

}  // namespace

}  // namespace clay

using namespace clay;

extern "C" int clay_volume_pack_bound(const clay_volume* vol, uint64_t* bytes) {
  const Volume* v = reinterpret_cast<const Volume*>(vol);
  if (!v || !bytes) return fail(CLAY_ERR_INVALID, "clay_volume_pack_bound: NULL argument");
  uint32_t rows, tiles;
  (void)row_map(v, &rows, &tiles);
  *bytes = wire_layout(tiles).payload + 4096ull * tiles;
  return CLAY_OK;
}

extern "C" int clay_volume_pack(clay_volume* vol, uint32_t instance, void* dst, uint64_t capacity, uint64_t* bytes,
                                clay_stream stream) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v || !dst) return fail(CLAY_ERR_INVALID, "clay_volume_pack: NULL argument");
  int rc = check_volume(v, instance, "clay_volume_pack");
  if (rc) return rc;
  uint32_t rows, tiles;
  const RowMap rm = row_map(v, &rows, &tiles);
  const Layout L = wire_layout(tiles);
  if (capacity < L.payload + 4096ull * tiles)
    return fail(CLAY_ERR_INVALID, "clay_volume_pack: the buffer is smaller than clay_volume_pack_bound");
  const size_t need = 2 * (size_t)rows + 2;
  if (v->wire_scratch_n < need) {
    dfree(v->wire_scratch);
    v->wire_scratch = nullptr;
    v->wire_scratch_n = 0;
    if ((rc = dalloc(&v->wire_scratch, need))) return rc;
    v->wire_scratch_n = need;
  }
  uint32_t* row_count = v->wire_scratch;
  uint32_t* row_off = v->wire_scratch + rows;
  uint32_t* total = v->wire_scratch + 2 * (size_t)rows;
  hipStream_t s = pick(v->ctx, stream);
  const size_t base = (size_t)instance * v->total_voxels;
  const uint16_t* dist = v->dist + base;
  const uint16_t* mat = v->mat + base;
  uint8_t* out = static_cast<uint8_t*>(dst);
  clay_wire_header h{};
  h.magic = kWireMagic;
  h.version = CLAY_WIRE_VERSION;
  h.resolution = (uint32_t)v->info.resolution;
  h.mip_levels = (uint32_t)v->info.mip_levels;
  h.band = (float)v->info.band;
  h.world_size = (float)v->info.world_size;
  h.gradient_voxels = (float)v->grad_voxels;
  h.tiles = tiles;
  for (int a = 0; a < 3; a++) h.origin[a] = v->origin[a];
  // The bitmap is built with atomic ORs: clear it first (a fill kernel, not a memset).
  uint32_t* bitmap = reinterpret_cast<uint32_t*>(out + L.bitmap);
  if ((rc = launch_fill_u32(v->ctx, s, bitmap, (tiles + 31u) / 32u, 0u))) return rc;
  {
    LaunchTimer lt(v->ctx, s, "wire_classify");
    hipLaunchKernelGGL(k_wire_classify, dim3(rows), dim3(kWireGroup), 0, s, rm, dist, mat,
                       reinterpret_cast<uint32_t*>(out + L.table), bitmap, row_count);
  }
  hipLaunchKernelGGL(k_wire_scan, dim3(1), dim3(1024), 0, s, row_count, row_off, rows, total);
  {
    LaunchTimer lt(v->ctx, s, "wire_write");
    hipLaunchKernelGGL(k_wire_write, dim3(rows), dim3(kWireGroup), 0, s, rm, h, dist, mat, row_off, total, out);
  }
  CLAY_HIP_CHECK(hipGetLastError());
  if (bytes) {
    uint32_t stored = 0;
    CLAY_HIP_CHECK(hipMemcpyAsync(&stored, total, sizeof(stored), hipMemcpyDeviceToHost, s));
    CLAY_HIP_CHECK(hipStreamSynchronize(s));
    *bytes = L.payload + 4096ull * stored;
  }
  return CLAY_OK;
}

extern "C" int clay_wire_check(const clay_wire_header* h, const clay_volume* vol) {
  const Volume* v = reinterpret_cast<const Volume*>(vol);
  if (!h || !v) return fail(CLAY_ERR_INVALID, "clay_wire_check: NULL argument");
  if (h->magic != kWireMagic) return fail(CLAY_ERR_INVALID, "volume stream: bad magic");
  if (h->version != CLAY_WIRE_VERSION)
    return fail(CLAY_ERR_INVALID, "volume stream: version " + std::to_string(h->version) + " is not supported");
  if (h->resolution != (uint32_t)v->info.resolution || h->mip_levels != (uint32_t)v->info.mip_levels)
    return fail(CLAY_ERR_INVALID, "volume stream: " + std::to_string(h->resolution) + "^3 x " +
                                      std::to_string(h->mip_levels) + " mips does not fit a " +
                                      std::to_string(v->info.resolution) + "^3 x " +
                                      std::to_string(v->info.mip_levels) + " volume");
  if (h->band != (float)v->info.band)
    return fail(CLAY_ERR_INVALID, "volume stream: band " + std::to_string(h->band) + " differs from the volume's " +
                                      std::to_string(v->info.band));
  uint32_t rows, tiles;
  (void)row_map(v, &rows, &tiles);
  const Layout L = wire_layout(tiles);
  if (h->tiles != tiles || h->table_offset != L.table || h->bitmap_offset != L.bitmap ||
      h->payload_offset != L.payload || h->stored_tiles > tiles || h->total_bytes != L.payload + 4096ull * h->stored_tiles)
    return fail(CLAY_ERR_INVALID, "volume stream: inconsistent header");
  return CLAY_OK;
}

extern "C" int clay_volume_unpack(clay_volume* vol, uint32_t instance, const void* src, uint64_t bytes,
                                  const clay_wire_header* header, clay_stream stream) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v || !src) return fail(CLAY_ERR_INVALID, "clay_volume_unpack: NULL argument");
  int rc = check_volume(v, instance, "clay_volume_unpack");
  if (rc) return rc;
  hipStream_t s = pick(v->ctx, stream);
  clay_wire_header h{};
  if (header) {
    h = *header;
  } else {
    if (bytes < sizeof(h)) return fail(CLAY_ERR_INVALID, "volume stream: shorter than its header");
    CLAY_HIP_CHECK(hipMemcpyAsync(&h, src, sizeof(h), hipMemcpyDeviceToHost, s));
    CLAY_HIP_CHECK(hipStreamSynchronize(s));
  }
  if ((rc = clay_wire_check(&h, vol))) return rc;
  if (bytes < h.total_bytes) return fail(CLAY_ERR_INVALID, "volume stream: truncated");
  uint32_t rows, tiles;
  const RowMap rm = row_map(v, &rows, &tiles);
  if (!v->wire_flag) {
    if ((rc = dalloc(&v->wire_flag, 1))) return rc;
    if ((rc = launch_fill_u32(v->ctx, s, v->wire_flag, 1, 0u))) return rc;
  }
  // The flag is sticky: set by a bad stream, cleared when clay_volume_wire_status reads it.
  uint32_t* bad = v->wire_flag;
  const size_t base = (size_t)instance * v->total_voxels;
  {
    LaunchTimer lt(v->ctx, s, "wire_unpack");
    hipLaunchKernelGGL(k_wire_unpack, dim3(rows), dim3(kWireGroup), 0, s, rm, tiles, h.stored_tiles,
                       static_cast<const uint8_t*>(src), v->dist + base, v->mat + base, bad);
  }
  CLAY_HIP_CHECK(hipGetLastError());
  if (header) return CLAY_OK;
  uint32_t flag = 0;
  return clay_volume_wire_status(vol, &flag) ? CLAY_ERR_INVALID : CLAY_OK;
}

extern "C" int clay_volume_wire_status(clay_volume* vol, uint32_t* bad) {
  Volume* v = reinterpret_cast<Volume*>(vol);
  if (!v || !bad) return fail(CLAY_ERR_INVALID, "clay_volume_wire_status: NULL argument");
  *bad = 0;
  if (!v->wire_flag) return CLAY_OK;
  CLAY_HIP_CHECK(hipDeviceSynchronize());
  CLAY_HIP_CHECK(hipMemcpy(bad, v->wire_flag, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (*bad) {
    int rc = launch_fill_u32(v->ctx, v->ctx->stream, v->wire_flag, 1, 0u);
    if (rc) return rc;
    CLAY_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
    return fail(CLAY_ERR_INVALID, "volume stream: a stored tile's index lies outside the payload");
  }
  return CLAY_OK;
}
