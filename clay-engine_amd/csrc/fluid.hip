// fluid.hip -- FluidSim substep kernels for gfx950.
//
// One substep of the reference (fluid.ts:550-556: clearGrid, insert, density,
// forces, integrate) becomes:
//   k_keys            cell key per particle (cellOf / inGrid / cellIndex, fluid.ts:253-266)
//   radix sort        stable by (key, id)  -> the bucket fill order        (sort.hip)
//   k_reorder         cell [start, end) table + sorted snapshot of pos/vel
//   k_density         rho, p over the 27-cell scan              (fluid.ts:333-368)
//   k_forces_integrate  pressure + XSPH -> nvel, then advance, collide, contain
//                     (fluid.ts:370-439 fused with :449-491)
// The dense per-cell count grid of the reference (cleared every substep) is gone:
// a cell record {start, end, stamp} is written only for occupied cells, stamped
// with the substep's epoch, and trusted only when its stamp is the current epoch,
// so stale records are never read and nothing grid-sized is cleared per substep.
//
// Forces and integrate can share one launch because the neighbour reads all hit
// the sorted snapshot while the writes go to the id-ordered state: the snapshot is
// the reference's `nvel` staging double buffer (fluid.ts:14-19) turned around.
//
// The same launches serve a batch of independent scenes: keys carry the scene in
// their top bits, so one sort orders every scene's particles by (scene, cell, id),
// each scene has its own block of cell records and contact list, and a scene's
// result is the one its own single simulation would compute.
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kBlock = 256;  // kernels that are not tuned

__device__ __forceinline__ void cell_of(const FluidK& k, f3 p, int32_t q[3]) {
  q[0] = sat_i32(floorf((p.x - k.go[0]) * k.inv_h));
  q[1] = sat_i32(floorf((p.y - k.go[1]) * k.inv_h));
  q[2] = sat_i32(floorf((p.z - k.go[2]) * k.inv_h));
}
__device__ __forceinline__ bool in_grid(const FluidK& k, int32_t x, int32_t y, int32_t z) {
  return x >= 0 && y >= 0 && z >= 0 && x < k.R && y < k.R && z < k.R;
}
__device__ __forceinline__ uint32_t cell_index(const FluidK& k, int32_t x, int32_t y, int32_t z) {
  return (uint32_t)x + (uint32_t)k.R * ((uint32_t)y + (uint32_t)k.R * (uint32_t)z);
}
__device__ __forceinline__ int32_t wrap_add(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}

// Scene of sorted slot s and the first of its cell records.
__device__ __forceinline__ uint32_t scene_at(const SceneK& sk, const uint32_t* __restrict__ skeys, uint32_t s) {
  return sk.batched ? skeys[s] >> sk.shift : 0u;
}
__device__ __forceinline__ const uint4* scene_cells(const SceneK& sk, const uint4* __restrict__ cells,
                                                    uint32_t scene) {
  return cells + (size_t)scene * sk.cells_per_scene;
}

// Walks the 27-cell neighbourhood of cell b in the reference's order (dz, dy, dx
// ascending, ids ascending inside a cell), calling visit(t) for the sorted slot t
// of each of the first `bucket` particles of every cell.
//
// A cell record is {start, end, stamp}: it describes the current substep only if
// its stamp equals this substep's epoch, so the table is never cleared.  The three
// cells of one grid row have consecutive keys, so when none of them holds more
// than `bucket` particles their particles form one contiguous run of the sorted
// order and are walked as one loop (same order, same result).
template <typename F>
__device__ __forceinline__ void for_neighbours(const FluidK& k, const uint4* __restrict__ cells,
                                               uint32_t epoch, const int32_t b[3], F&& visit) {
  for (int dz = -1; dz <= 1; dz++) {
    const int32_t z = wrap_add(b[2], dz);
    for (int dy = -1; dy <= 1; dy++) {
      const int32_t y = wrap_add(b[1], dy);
      uint32_t st[3], en[3];
      bool ok[3];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int32_t x = wrap_add(b[0], j - 1);
        const bool in = in_grid(k, x, y, z);
        uint4 c = make_uint4(0u, 0u, 0u, 0u);
        if (in) c = cells[cell_index(k, x, y, z)];
        ok[j] = in && c.z == epoch;
        st[j] = c.x;
        en[j] = c.y;
      }
      uint32_t lo = 0xFFFFFFFFu, hi = 0u;
      bool capped = false;
#pragma unroll
      for (int j = 0; j < 3; j++) {
        if (!ok[j]) continue;
        lo = st[j] < lo ? st[j] : lo;
        hi = en[j] > hi ? en[j] : hi;
        capped = capped || (en[j] - st[j] > (uint32_t)k.bucket);
      }
      if (!capped) {
        for (uint32_t t = lo; t < hi; t++) visit(t);
      } else {
#pragma unroll
        for (int j = 0; j < 3; j++) {
          if (!ok[j]) continue;
          const uint32_t cap = st[j] + (uint32_t)k.bucket;
          const uint32_t e = en[j] < cap ? en[j] : cap;
          for (uint32_t t = st[j]; t < e; t++) visit(t);
        }
      }
    }
  }
}

// Wave-wide min / max of a u32, returned to every lane: DPP row shifts build each
// 16-lane row's result in its last lane, two row broadcasts carry it across the
// rows, and lane 63 holds the wave's result (no LDS round trips).  Lanes a shift
// or broadcast does not reach keep their own value.  All 64 lanes must be active.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t keep, uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)keep, (int)x, kCtrl, kRowMask, 0xF, false);
}
template <bool kMax>
__device__ __forceinline__ uint32_t wave_reduce_u32(uint32_t v) {
  auto pick = [](uint32_t a, uint32_t b) { return kMax ? (a > b ? a : b) : (a < b ? a : b); };
  v = pick(v, dpp_u32<0x111, 0xF>(v, v));  // row_shr:1
  v = pick(v, dpp_u32<0x112, 0xF>(v, v));  // row_shr:2
  v = pick(v, dpp_u32<0x114, 0xF>(v, v));  // row_shr:4
  v = pick(v, dpp_u32<0x118, 0xF>(v, v));  // row_shr:8
  v = pick(v, dpp_u32<0x142, 0xA>(v, v));  // row_bcast:15 into rows 1 and 3
  v = pick(v, dpp_u32<0x143, 0xC>(v, v));  // row_bcast:31 into rows 2 and 3
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) { return wave_reduce_u32<false>(v); }
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) { return wave_reduce_u32<true>(v); }

// Cell key per particle.  One scene: the cell index, or `invalid` outside the
// grid.  A batch: (scene << shift) | cell, with all `shift` cell bits set outside
// the grid, so a scene's outside particles sort after its cells; `shift` is the
// fewest bits that leave that value free, so the keys stay short for the sort.
//
// With `bounds` (the frame-end sort of a fluid whose surface follows its
// particles), the same position reads also fold each scene's bounds into its six
// slots (min x, y, z, max x, y, z as ordered bits), so the bake needs no second
// pass over the positions.  A workgroup inside one scene (the usual case: ids run
// scene-major and scenes hold >= 1024 particles) reduces to one set of six
// atomics.  Otherwise the wave's lanes are grouped by scene and each group takes a
// DPP min/max, its first lane doing the atomics.
__global__ void __launch_bounds__(kBlock) k_keys(FluidK k, uint32_t n, const float* __restrict__ px,
                                                 const float* __restrict__ py,
                                                 const float* __restrict__ pz,
                                                 const uint32_t* __restrict__ scene_of,
                                                 uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals, uint32_t invalid, uint32_t shift,
                                                 uint32_t* __restrict__ bounds, uint32_t* __restrict__ bump) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool active = i < n;
  // The substep stamp advances here, not in a launch of its own: nothing in this
  // pass or the sort reads it, and the reorder that stamps the cells runs after.
  if (bump && i == 0) bump[0] = bump[0] + 1u;
  f3 p = mk(0.0f, 0.0f, 0.0f);
  uint32_t sc = 0;
  if (active) {
    p = mk(px[i], py[i], pz[i]);
    if (scene_of) sc = scene_of[i];
    int32_t q[3];
    cell_of(k, p, q);
    const bool in = in_grid(k, q[0], q[1], q[2]);
    if (scene_of)
      keys[i] = (sc << shift) | (in ? cell_index(k, q[0], q[1], q[2]) : (1u << shift) - 1u);
    else
      keys[i] = in ? cell_index(k, q[0], q[1], q[2]) : invalid;
    vals[i] = i;
  }
  if (!bounds) return;  // uniform over the launch
  // Every lane of every wave reaches here (inactive lanes carry the identities), as
  // the DPP reductions need.  The bounds are folded as ordered bits (ordered_bits),
  // whose integer min / max is the float min / max.
  const int lane = (int)__lane_id();
  const uint32_t first = blockIdx.x * kBlock;  // < n
  const uint32_t last = first + kBlock - 1 < n ? first + kBlock - 1 : n - 1;
  const uint32_t pb[3] = {ordered_bits(p.x), ordered_bits(p.y), ordered_bits(p.z)};
  if (!scene_of || scene_of[first] == scene_of[last]) {  // uniform over the workgroup
    __shared__ uint32_t part[kBlock / 64][6];
    uint32_t a[6];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      a[q] = wave_min_u32(active ? pb[q] : 0xFFFFFFFFu);
      a[q + 3] = wave_max_u32(active ? pb[q] : 0u);
    }
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 6; q++) part[threadIdx.x >> 6][q] = a[q];
    __syncthreads();
    if (threadIdx.x < 6) {
      const int q = threadIdx.x;
      uint32_t v = part[0][q];
      for (int w = 1; w < kBlock / 64; w++) v = q < 3 ? (part[w][q] < v ? part[w][q] : v) : (part[w][q] > v ? part[w][q] : v);
      uint32_t* slot = bounds + (size_t)kBoundsStride * (scene_of ? scene_of[first] : 0u) + q;
      if (q < 3)
        atomicMin(slot, v);
      else
        atomicMax(slot, v);
    }
    return;
  }
  uint64_t pending = __ballot(active);
  while (pending) {
    const int leader = __builtin_ctzll(pending);
    const uint32_t lead = (uint32_t)__builtin_amdgcn_readlane((int)sc, leader);
    const uint64_t same = __ballot(active && sc == lead);
    const bool mine = (same >> lane) & 1ull;
    uint32_t a[6];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      a[q] = wave_min_u32(mine ? pb[q] : 0xFFFFFFFFu);
      a[q + 3] = wave_max_u32(mine ? pb[q] : 0u);
    }
    if (lane == leader) {
      uint32_t* slot = bounds + (size_t)kBoundsStride * lead;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        atomicMin(slot + q, a[q]);
        atomicMax(slot + 3 + q, a[q + 3]);
      }
    }
    pending &= ~same;
  }
}

struct Planes8 {
  const float* src[8];
  float* dst[8];
};

// Cell records of this substep (run boundaries of the sorted keys) and the sorted
// snapshot of the state planes.
// A staged neighbour position: x in f32, y and z as f16 offsets from the centre of
// the neighbour's own cell (|offset| <= h/2).  All candidates of one lane's row lie
// in cells of the same y and z, so for that lane y_i - y_j = (o_iy - dy h) - o_jy, and
// x_i - x_j is a plain f32 difference.
__device__ __forceinline__ uint2 pack_stage(const FluidK& k, f3 p) {
  int32_t c[3];
  cell_of(k, p, c);
  const float oy = p.y - (k.go[1] + ((float)c[1] + 0.5f) * k.h);
  const float oz = p.z - (k.go[2] + ((float)c[2] + 0.5f) * k.h);
  return make_uint2(__float_as_uint(p.x), (uint32_t)f2h(oy) | ((uint32_t)f2h(oz) << 16));
}

// The FP16-LDS forces pass stages each neighbour as this record, which the density
// pass writes for its own particle once its density is known.
struct StagedNbr {
  float4 vr;   // velocity and 1 / density
  uint2 pos;   // x (f32) and the y, z offsets from its cell centre (f16), as pack_stage
  float pinv;  // pressure / density
  float pad;
};

// With `spos` (FP16-LDS), also each sorted particle's staged position (pack_stage).
__global__ void __launch_bounds__(kBlock) k_reorder(uint32_t n, const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ svals, SceneK sk,
                                                    uint32_t invalid, uint32_t* __restrict__ cells,
                                                    Planes8 pl, int planes, FluidK k, uint2* __restrict__ spos) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t key = skeys[s];
  const uint32_t cell = sk.batched ? (key & sk.cell_mask) : key;
  const bool inside = sk.batched ? cell != sk.cell_mask : key != invalid;
  if (inside) {
    const size_t rec = 4 * ((size_t)(sk.batched ? key >> sk.shift : 0u) * sk.cells_per_scene + cell);
    if (s == 0 || skeys[s - 1] != key) {
      cells[rec + 0] = s;
      cells[rec + 2] = *sk.epoch;
    }
    if (s == n - 1 || skeys[s + 1] != key) cells[rec + 1] = s + 1;
  }
  const uint32_t j = svals[s];
#pragma unroll
  for (int p = 0; p < 8; p++)
    if (p < planes) pl.dst[p][s] = pl.src[p][j];
  if (spos) spos[s] = pack_stage(k, mk(pl.src[0][j], pl.src[1][j], pl.src[2][j]));
}

// ---------------------------------------------------------------------------
// Interior radius of a particle, for the splat (a voxel's code only needs the
// particle nearest its centre).  The offset d between the particle and each
// neighbour found by the density pass is classed by the signs of its components
// that exceed 0.35 |d| into one of the 26 lattice directions: every direction
// falls in exactly one class, within 30.5 degrees of its axis (measured over 4M
// random directions).  When every class holds a neighbour q at distance in
// (0.1875 h, dmax), a voxel centre c at distance r > 1.118 dmax from the
// particle p has, in the class of c - p, a neighbour within 61 degrees of that
// direction, so |c - q|^2 <= r^2 - |q - p| (2 r cos 62deg - |q - p|) <
// r^2 - 0.049 |q - p| dmax: q is strictly nearer c, by far more than the rounding
// of the f32 squared distances and of the f16 offsets (<= 5e-5 in d, which also
// bends the class angle by < 0.2 degree).  Such a pair cannot hold c's minimum,
// so the splat skips voxels beyond that radius.  Levels dmax = h, 1.25h, 1.5h
// over every candidate of the 27-cell walk (a fluid at rest spreads past h on
// the diagonals within a few frames); the smallest level whose 26 classes are
// all filled gives the radius.
// ---------------------------------------------------------------------------
// Direction classes of a particle's neighbours for the interior radius: a
// neighbour with 0.1875 h < |d| < 1.5 h sets the bit of its lattice direction
// (per axis, the sign of d_a when |d_a| >= 0.35 |d|, else 0) in the masks of
// the distance levels it is within (h, 1.25 h, 1.5 h).  Candidates outside the
// shell are branched over; inside it, the level tests are sign bits of IEEE
// differences (neg_mask), exact and on full-rate integer operations.
struct Cover {
  uint32_t m0 = 0, m1 = 0, m2 = 0;
  float lo, hi, l0, l1, c07, c15;
  uint32_t one;
  __device__ __forceinline__ explicit Cover(float hh)
      : lo(vreg(0.03515625f * hh)), hi(vreg(2.25f * hh)), l0(vreg(hh)), l1(vreg(1.5625f * hh)),
        c07(vreg(1.0f / 0.7f)), c15(vreg(1.5f)), one(vreg_u(1u)) {}
  // Per axis, floor(d_a / (0.7 |d|) + 1.5) is 0, 1 or 2 as the unit component
  // d_a / |d| is below -0.35, within [-0.35, 0.35) or at least 0.35.  Since
  // |d_a / |d|| <= 1 the argument lies in [0.07, 2.93], where the truncating
  // conversion is the floor.  1 / |d| comes from the hardware reciprocal square
  // root, whose 1-ulp error only moves directions that lie on a class border.
  __device__ __forceinline__ void add(f3 d, float r2) {
    if (!(r2 > lo) || !(r2 < hi)) return;
    const float k = __builtin_amdgcn_rsqf(r2) * c07;
    const int32_t sx = (int32_t)__builtin_fmaf(d.x, k, c15);
    const int32_t sy = (int32_t)__builtin_fmaf(d.y, k, c15);
    const int32_t sz = (int32_t)__builtin_fmaf(d.z, k, c15);
    const uint32_t b = one << (uint32_t)(sx + 3 * sy + 9 * sz);  // bit 13 + (sx - 1) + 3 (sy - 1) + 9 (sz - 1)
    m0 |= b & (uint32_t)neg_mask(r2 - l0);
    m1 |= b & (uint32_t)neg_mask(r2 - l1);
    m2 |= b;
  }
  // Squared radius (1.118 dmax)^2 = 1.25 dmax^2, or +inf when no level is full.
  __device__ __forceinline__ float radius2(float hh) const {
    constexpr uint32_t kAll = 0x7FFFFFFu & ~(1u << 13);
    if (m0 == kAll) return 1.25f * hh;
    if (m1 == kAll) return 1.953125f * hh;
    if (m2 == kAll) return 2.8125f * hh;
    return __uint_as_float(0x7f800000u);
  }
};

// Where the density pass writes the interior radius: by particle id.
struct CoverOut {
  const uint32_t* svals;
  float* r2;
};

template <int kB, bool kCover>
__global__ void __launch_bounds__(kB)
    k_density(FluidK k, SceneK sk, uint32_t n, const float* __restrict__ spx, const float* __restrict__ spy,
              const float* __restrict__ spz, const uint32_t* __restrict__ skeys,
              const uint4* __restrict__ cells, float* __restrict__ srho, float* __restrict__ spres, CoverOut co) {
  const uint32_t s = (sk.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x) * kB + threadIdx.x;
  if (s >= n) return;
  const uint32_t epoch = *sk.epoch;
  const uint4* cs = scene_cells(sk, cells, scene_at(sk, skeys, s));
  const f3 pi = mk(spx[s], spy[s], spz[s]);
  int32_t b[3];
  cell_of(k, pi, b);
  float rho = 0.0f;
  Cover cov(k.hh);
  for_neighbours(k, cs, epoch, b, [&](uint32_t t) {
    const f3 d = mk(pi.x - spx[t], pi.y - spy[t], pi.z - spz[t]);
    const float r2 = dot3(d, d);
    if (r2 < k.hh) {
      const float w = k.hh - r2;
      rho = rho + k.mass_poly6 * (w * w * w);
    }
    if (kCover) cov.add(d, r2);
  });
  srho[s] = fmax_s(rho, k.rest_floor);
  spres[s] = fmax_s(k.stiff * (rho - k.rest), 0.0f);
  if (kCover) co.r2[co.svals[s]] = cov.radius2(k.hh);
}

struct SortedIn {
  const float* p[8];  // spx spy spz svx svy svz srho spres
};
struct StateOut {
  float* p[11];  // px py pz vx vy vz rho pres nvx nvy nvz
};

// The reference's contact counter (fluid.ts:469-477), advanced once per wave and
// counter: the wave's contacting lanes are grouped by the counter they add to (a
// wave can straddle scenes), and one lane of each group adds the group's size.
// One atomic per particle on a scene's single counter serialises in L2 once a
// scene holds 1e5 contacts.  Called by every active lane of the wave.
__device__ __forceinline__ void add_contacts(uint32_t* counter, bool hit) {
  // readlane returns a signed int: both halves are zero-extended before they are
  // joined, and the leader's bit is cleared whatever the compare gives, so every
  // iteration retires at least one lane.
  const uint64_t addr = (uint64_t)(uintptr_t)counter;
  uint64_t pending = __ballot(hit);
  while (pending) {
    const int leader = __builtin_ctzll(pending);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)addr, leader);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(addr >> 32), leader);
    const uint64_t want = (uint64_t)lo | ((uint64_t)hi << 32);
    const uint64_t same = __ballot(hit && addr == want) | (1ull << leader);
    if ((int)__lane_id() == leader) atomicAdd(counter, (uint32_t)__popcll(same));
    pending &= ~same;
  }
}

// The end of the force pass (fluid.ts:430-437) and the integrate pass
// (fluid.ts:449-491) for one particle, writing its id-ordered state.
__device__ __forceinline__ void finish_particle(const FluidK& k, const SceneK& sk, const ColliderDev& col,
                                                uint32_t scene, uint32_t s, f3 pi, f3 vi, float rhoi, float pri,
                                                f3 fp, f3 vsum, float wsum, const uint32_t* __restrict__ svals,
                                                const StateOut& out, uint32_t* __restrict__ contacts) {
  const uint32_t j = svals[s];
  const float irho = 1.0f / rhoi;
  const f3 acc = mk(fp.x * irho + k.g[0], fp.y * irho + k.g[1], fp.z * irho + k.g[2]);
  const float tv = fclamp_s(k.visc, 0.0f, 1.0f);
  const float iw = 1.0f / wsum;
  const f3 vs = vsum * iw;
  f3 v = mk(fmix(vi.x, vs.x, tv), fmix(vi.y, vs.y, tv), fmix(vi.z, vs.z, tv));
  v = mk(v.x + acc.x * k.dt, v.y + acc.y * k.dt, v.z + acc.z * k.dt) * k.damp;
  const float speed = len3(v);
  if (speed > k.vmax) v = v * (k.vmax / speed);
  const f3 nv = v;

  // ---- integrate (fluid.ts:449-491) ----
  f3 p = mk(pi.x + v.x * k.dt, pi.y + v.y * k.dt, pi.z + v.z * k.dt);
  if (col.nleaf > 0) {
    float dw;
    bool hit = false;
    const int leaf = collider_pick(col, scene, p, &dw);
    if (dw < k.radius) {
      const f3 nr = collider_normal(col, scene, leaf, p);
      const float push = k.radius - dw;
      p = mk(p.x + nr.x * push, p.y + nr.y * push, p.z + nr.z * push);
      const float vn = fmin_s(dot3(v, nr), 0.0f);
      v = mk(v.x - nr.x * vn, v.y - nr.y * vn, v.z - nr.z * vn);
      if (dw < k.cdepth) {
        // Which 32 contacts the list keeps is decided afterwards, by particle id
        // (k_contacts_collect).
        hit = true;
        sk.cdata[j] = make_float4(p.x, p.y, p.z, len3(v));
        atomicOr(&sk.cmask[j >> 5], 1u << (j & 31u));
      }
    }
    add_contacts(contacts, hit);
  }
  const float clx = fmin_s(fmax_s(p.x, k.lo[0]), k.hi[0]);
  const float cly = fmin_s(fmax_s(p.y, k.lo[1]), k.hi[1]);
  const float clz = fmin_s(fmax_s(p.z, k.lo[2]), k.hi[2]);
  out.p[0][j] = clx;
  out.p[1][j] = cly;
  out.p[2][j] = clz;
  out.p[3][j] = clx == p.x ? v.x : 0.0f;
  out.p[4][j] = cly == p.y ? v.y : 0.0f;
  out.p[5][j] = clz == p.z ? v.z : 0.0f;
  out.p[6][j] = rhoi;
  out.p[7][j] = pri;
  out.p[8][j] = nv.x;
  out.p[9][j] = nv.y;
  out.p[10][j] = nv.z;
}

template <int kB>
__global__ void __launch_bounds__(kB)
    k_forces_integrate(FluidK k, SceneK sk, ColliderDev col, uint32_t n, SortedIn in,
                       const uint32_t* __restrict__ svals, const uint32_t* __restrict__ skeys,
                       const uint4* __restrict__ cells, StateOut out, uint32_t* __restrict__ contacts) {
  const uint32_t s = (sk.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x) * kB + threadIdx.x;
  if (s >= n) return;
  const float* __restrict__ spx = in.p[0];
  const float* __restrict__ spy = in.p[1];
  const float* __restrict__ spz = in.p[2];
  const float* __restrict__ svx = in.p[3];
  const float* __restrict__ svy = in.p[4];
  const float* __restrict__ svz = in.p[5];
  const float* __restrict__ srho = in.p[6];
  const float* __restrict__ spres = in.p[7];
  const uint32_t epoch = *sk.epoch;
  const uint32_t scene = scene_at(sk, skeys, s);
  const uint4* cs = scene_cells(sk, cells, scene);

  const f3 pi = mk(spx[s], spy[s], spz[s]);
  const f3 vi = mk(svx[s], svy[s], svz[s]);
  const float rhoi = srho[s], pri = spres[s];
  int32_t b[3];
  cell_of(k, pi, b);
  f3 fp = mk(0.0f, 0.0f, 0.0f);
  f3 vsum = vi * k.h6;
  float wsum = k.h6;
  for_neighbours(k, cs, epoch, b, [&](uint32_t t) {
    if (t == s) return;
    const f3 dv = mk(pi.x - spx[t], pi.y - spy[t], pi.z - spz[t]);
    const float r2 = dot3(dv, dv);
    if (r2 >= k.hh) return;
    const float inv = 1.0f / srho[t];
    const float w0 = k.hh - r2;
    const float w = w0 * w0 * w0;
    vsum = mk(vsum.x + svx[t] * w, vsum.y + svy[t] * w, vsum.z + svz[t] * w);
    wsum = wsum + w;
    const float r = sqrtf(r2);
    if (r > 1e-6f) {
      const float hr = k.h - r;
      const float s2 = k.mass * (pri + spres[t]) * 0.5f * inv * k.spiky * (hr * hr);
      const float ir = 1.0f / r;
      fp = mk(fp.x - dv.x * ir * s2, fp.y - dv.y * ir * s2, fp.z - dv.z * ir * s2);
    }
  });
  finish_particle(k, sk, col, scene, s, pi, vi, rhoi, pri, fp, vsum, wsum, svals, out,
                  contacts + (size_t)scene * sk.contact_stride);
}

// ---------------------------------------------------------------------------
// Precision variant: FP16 cell-relative offsets staged in LDS.  Tolerance-graded
// against the f64 oracle; FMA contraction is allowed in the neighbour loops.
//
// For each of the 9 rows of the 27-cell neighbourhood, every lane first drops the
// neighbour cells of the row whose box lies beyond the kernel's reach from its
// particle (row_span), leaving one sorted range.  The wave takes the union of its
// lanes' ranges (consecutive sorted particles share most of them) and stages it in
// its LDS slice -- x as f32, y and z as f16 offsets from the centre of each
// particle's own cell (|offset| <= h/2), the other fields as f32 -- and every lane
// walks its range out of LDS.  The row fixes the cells' y and z, so a lane rebuilds
// p_i - p_j = (x_i - x_j, (o_iy - dy h) - o_jy, (o_iz - dz h) - o_jz).  A row
// whose union is wider than the slice, or whose cells overflow the bucket, is
// walked from global memory instead.  In a batch a wave may straddle two scenes:
// each lane reads its own scene's cell records, and the union of ranges is still
// one span of the sorted order.
// ---------------------------------------------------------------------------

// Orders a wave's LDS accesses around this point (writes before it are seen by
// every lane after it), without waiting for the workgroup's other waves: each
// wave stages into a slice of its own.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One row (dz, dy) of a particle's neighbourhood: its three cells' run [lo, hi)
// when no cell overflows the bucket (capped = false), else the per-cell spans.
struct RowRun {
  uint32_t st[3], en[3];
  bool ok[3];
  uint32_t lo, hi;
  bool capped;
};

__device__ __forceinline__ RowRun row_run(const FluidK& k, const uint4* __restrict__ cells, uint32_t epoch,
                                          const int32_t b[3], int32_t y, int32_t z, bool active) {
  RowRun r;
  r.lo = 0xFFFFFFFFu;
  r.hi = 0u;
  r.capped = false;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int32_t x = wrap_add(b[0], j - 1);
    const bool in = active && in_grid(k, x, y, z);
    uint4 c = make_uint4(0u, 0u, 0u, 0u);
    if (in) c = cells[cell_index(k, x, y, z)];
    r.ok[j] = in && c.z == epoch;
    r.st[j] = c.x;
    r.en[j] = c.y;
    if (r.ok[j]) {
      r.lo = c.x < r.lo ? c.x : r.lo;
      r.hi = c.y > r.hi ? c.y : r.hi;
      r.capped = r.capped || (c.y - c.x > (uint32_t)k.bucket);
    }
  }
  return r;
}

template <typename F>
__device__ __forceinline__ void walk_row(const FluidK& k, const RowRun& r, F&& visit) {
  if (!r.capped) {
    for (uint32_t t = r.lo; t < r.hi; t++) visit(t);
  } else {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (!r.ok[j]) continue;
      const uint32_t cap = r.st[j] + (uint32_t)k.bucket;
      const uint32_t e = r.en[j] < cap ? r.en[j] : cap;
      for (uint32_t t = r.st[j]; t < e; t++) visit(t);
    }
  }
}

// Offset of p from the centre of its own cell b.
__device__ __forceinline__ f3 cell_rel(const FluidK& k, f3 p, const int32_t b[3]) {
  return mk(p.x - (k.go[0] + ((float)b[0] + 0.5f) * k.h), p.y - (k.go[1] + ((float)b[1] + 0.5f) * k.h),
            p.z - (k.go[2] + ((float)b[2] + 0.5f) * k.h));
}

// Which of a row's three cells (dy, dz fixed, dx = -1, 0, 1) can hold a particle
// within `reach` of p_i (o_i: its offset from its cell centre): a cell is dropped
// when the gap between p_i and the cell's box is at least reach, by a relative
// margin on the squared reach that covers the rounding of o_i and of the cell
// assignment.  Every candidate in a dropped cell lies beyond reach, so dropping it
// changes no sum.
__device__ __forceinline__ void row_keep(f3 oi, int dy, int dz, float h, float reach2, float margin, bool keep[3]) {
  const float hh = 0.5f * h;
  const float gy = dy == 0 ? 0.0f : fmax_s(dy < 0 ? oi.y + hh : hh - oi.y, 0.0f);
  const float gz = dz == 0 ? 0.0f : fmax_s(dz < 0 ? oi.z + hh : hh - oi.z, 0.0f);
  const float gyz = gy * gy + gz * gz;
  const float gl = fmax_s(oi.x + hh, 0.0f), gr = fmax_s(hh - oi.x, 0.0f);
  const float lim = reach2 * margin;
  keep[0] = gl * gl + gyz < lim;
  keep[1] = gyz < lim;
  keep[2] = gr * gr + gyz < lim;
}

// One lane's share of a row: the sorted range of the kept cells.  The row's three
// cells are consecutive keys, so what remains is one range.
struct RowSpan {
  uint32_t lo, hi;
};
__device__ __forceinline__ RowSpan kept_span(const RowRun& r, const bool keep[3]) {
  RowSpan sp{0xFFFFFFFFu, 0u};
#pragma unroll
  for (int j = 0; j < 3; j++)
    if (r.ok[j] && keep[j]) {
      sp.lo = r.st[j] < sp.lo ? r.st[j] : sp.lo;
      sp.hi = r.en[j] > sp.hi ? r.en[j] : sp.hi;
    }
  return sp;
}
__device__ __forceinline__ RowSpan row_span(const RowRun& r, f3 oi, int dy, int dz, float h, float reach2) {
  bool keep[3];
  row_keep(oi, dy, dz, h, reach2, 1.0001f, keep);
  return kept_span(r, keep);
}

// The range of the sorted order a wave stages for one row: the union of its
// lanes' kept spans.  For particles inside the grid a row's run [lo, hi) (the
// first occupied cell's start, the last occupied cell's end) never decreases
// along the sorted order: a later particle's cell has a key at least as large,
// and so do the first and the last cell of its row.  Every kept span lies in
// its lane's run, so the union lies in [lo of the first lane whose row is
// occupied, hi of the last one): two readlanes instead of two wave reductions.
// A wave holding a particle outside the grid (`mono` false: its rows can lie
// anywhere) reduces the kept spans themselves.
__device__ __forceinline__ void wave_union(const RowRun& r, const RowSpan& sp, bool any, bool mono, uint32_t& ulo,
                                           uint32_t& uhi) {
  if (mono) {
    const uint64_t m = __ballot(r.hi > r.lo);
    ulo = 0xFFFFFFFFu;
    uhi = 0u;
    if (m) {
      ulo = (uint32_t)__builtin_amdgcn_readlane((int)r.lo, __builtin_ctzll(m));
      uhi = (uint32_t)__builtin_amdgcn_readlane((int)r.hi, 63 - __builtin_clzll(m));
    }
  } else {
    ulo = wave_min_u32(any ? sp.lo : 0xFFFFFFFFu);
    uhi = wave_max_u32(any ? sp.hi : 0u);
  }
}

// base - (the staged f16 offset): each component is one v_fma_mix_f32, which widens
// the f16 half (exact) and computes base - x with one rounding -- the result of a
// convert followed by a subtract, in one instruction instead of two.
__device__ __forceinline__ float sub_f16_lo(float base, uint32_t h2) {
  float r;
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h2), "v"(base));
  return r;
}
__device__ __forceinline__ float sub_f16_hi(float base, uint32_t h2) {
  float r;
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h2), "v"(base));
  return r;
}
__device__ __forceinline__ float dot3_fma(f3 d) { return __builtin_fmaf(d.x, d.x, __builtin_fmaf(d.y, d.y, d.z * d.z)); }

// Waves per SIMD the forces kernels are compiled for.  Their staging slice takes
// 32 bytes per neighbour per wave: at 256 neighbours two 64-lane waves hold 16 KB,
// so LDS (160 KB per CU) admits 5 waves per SIMD whatever the registers allow, and
// asking for 7 would only make the register allocator spill.
template <int kS>
constexpr int kForcesWaves = kS >= 256 ? 5 : 7;

// Occupancy hints: the FP16 density kernel fits 64 VGPRs (8 waves per SIMD) with no
// spill; its frame-end form (interior-radius classes) needs a few more and runs at
// 7; the forces kernels with 128-entry staging are held to 7 waves, at the cost of
// 48 bytes of scratch per lane, which the extra wave more than pays for (FP16
// forces: 1.11 -> 1.02 ms per batched frame).
template <int kB, int kS, bool kCover>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(kCover ? 7 : 8)))
    k_density_lds(FluidK k, SceneK sk, uint32_t n, const float* __restrict__ spx, const float* __restrict__ spy,
                  const float* __restrict__ spz, const uint32_t* __restrict__ skeys,
                  const uint4* __restrict__ cells, float* __restrict__ srho, float* __restrict__ spres, CoverOut co,
                  const uint2* __restrict__ spos, const float* __restrict__ svx, const float* __restrict__ svy,
                  const float* __restrict__ svz, StagedNbr* __restrict__ frec) {
#pragma clang fp contract(fast)
  __shared__ uint2 stage[kB / 64][kS];
  const uint32_t s = (sk.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x) * kB + threadIdx.x;
  const bool active = s < n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t epoch = *sk.epoch;
  const uint4* cs = scene_cells(sk, cells, active ? scene_at(sk, skeys, s) : 0u);
  f3 pi = mk(0.0f, 0.0f, 0.0f);
  if (active) pi = mk(spx[s], spy[s], spz[s]);
  int32_t b[3] = {0, 0, 0};
  if (active) cell_of(k, pi, b);
  const f3 oi = cell_rel(k, pi, b);
  const bool mono = !__ballot(active && !in_grid(k, b[0], b[1], b[2]));  // wave_union
  // Compensated density: each row's terms (at most a few dozen) are summed plainly,
  // and the nine row sums are added with Neumaier's compensation, so the rounding
  // of the total is that of one row rather than of every term.  The sums are of
  // (hh - r2)^3, clamped at zero beyond h instead of branched on; the mass and
  // kernel factor multiply the total once.
  float rho_hi = 0.0f, rho_lo = 0.0f, rho = 0.0f;
  const float hh_v = vreg(k.hh);
  Cover cov(k.hh);
  auto term = [&](f3 d) {
    const float r2 = dot3_fma(d);
    const float w = __builtin_fmaxf(hh_v - r2, 0.0f);
    rho = __builtin_fmaf(w * w, w, rho);
    if (kCover) cov.add(d, r2);
  };
  auto fold_row = [&]() {
    const float t = rho_hi + rho;
    rho_lo += fabsf(rho_hi) >= fabsf(rho) ? (rho_hi - t) + rho : (rho - t) + rho_hi;
    rho_hi = t;
    rho = 0.0f;
  };
  // The frame-end pass classes neighbours out to 1.5 h for the interior radius.
  const float reach2 = kCover ? 2.25f * k.hh : k.hh;
  for (int dz = -1; dz <= 1; dz++) {
    for (int dy = -1; dy <= 1; dy++) {
      const RowRun r = row_run(k, cs, epoch, b, wrap_add(b[1], dy), wrap_add(b[2], dz), active);
      const RowSpan sp = row_span(r, oi, dy, dz, k.h, reach2);
      const bool any = sp.hi > sp.lo;
      uint32_t ulo, uhi;
      wave_union(r, sp, any, mono, ulo, uhi);
      const bool staged = !__ballot(r.capped) && uhi > ulo && uhi - ulo <= (uint32_t)kS;
      if (staged)
        for (uint32_t t = ulo + lane; t < uhi; t += 64) stage[wave][t - ulo] = spos[t];
      wave_sync();
      if (staged) {
        const float by = oi.y - (float)dy * k.h, bz = oi.z - (float)dz * k.h;
        auto staged_term = [&](uint2 u) {
          term(mk(pi.x - __uint_as_float(u.x), sub_f16_lo(by, u.y), sub_f16_hi(bz, u.y)));
        };
        // Four candidates per trip (then two, then one): a quarter of the loop
        // counting and exit tests.
        const uint2* q = &stage[wave][0] + (any ? sp.lo - ulo : 0u);
        const uint32_t cnt = any ? sp.hi - sp.lo : 0u;
        uint32_t j = 0;
        for (; j + 4 <= cnt; j += 4) {
          const uint2 u0 = q[j], u1 = q[j + 1], u2 = q[j + 2], u3 = q[j + 3];
          staged_term(u0);
          staged_term(u1);
          staged_term(u2);
          staged_term(u3);
        }
        if (j + 2 <= cnt) {
          const uint2 u0 = q[j], u1 = q[j + 1];
          staged_term(u0);
          staged_term(u1);
          j += 2;
        }
        if (j < cnt) staged_term(q[j]);
      } else {
        walk_row(k, r, [&](uint32_t t) { term(pi - mk(spx[t], spy[t], spz[t])); });
      }
      fold_row();
      wave_sync();
    }
  }
  if (!active) return;
  rho = (rho_hi + rho_lo) * k.mass_poly6;
  const float rr = fmax_s(rho, k.rest_floor), pr = fmax_s(k.stiff * (rho - k.rest), 0.0f);
  srho[s] = rr;
  spres[s] = pr;
  if (kCover) co.r2[co.svals[s]] = cov.radius2(k.hh);
  // The record the forces pass stages for this particle: 1 / density taken once here.
  const float inv = 1.0f / rr;
  StagedNbr e;
  e.vr = make_float4(svx[s], svy[s], svz[s], inv);
  e.pos = spos[s];
  e.pinv = pr * inv;
  e.pad = 0.0f;
  frec[s] = e;
}

// ---------------------------------------------------------------------------
// Exact LDS-staged variants of k_density / k_forces_integrate: the same wave row
// staging as the FP16 kernels, but the staged copy is each neighbour's f32 data
// and every term is the plain kernel's, evaluated in the same order with no
// contraction, so the results are the plain path's bit for bit.  A row's three
// cells are contiguous in sorted order, so walking them cell by cell from LDS
// visits the candidates in the order of the plain walk.  1/rho of a neighbour is
// taken once at staging: the same IEEE quotient the plain kernel takes per pair.
// ---------------------------------------------------------------------------
template <int kB, int kS, bool kCover>
__global__ void __launch_bounds__(kB)
    k_density_x(FluidK k, SceneK sk, uint32_t n, const float* __restrict__ spx, const float* __restrict__ spy,
                const float* __restrict__ spz, const uint32_t* __restrict__ skeys,
                const uint4* __restrict__ cells, float* __restrict__ srho, float* __restrict__ spres, CoverOut co) {
  __shared__ float4 stage[kB / 64][kS];
  const uint32_t s = (sk.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x) * kB + threadIdx.x;
  const bool active = s < n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t epoch = *sk.epoch;
  const uint4* cs = scene_cells(sk, cells, active ? scene_at(sk, skeys, s) : 0u);
  f3 pi = mk(0.0f, 0.0f, 0.0f);
  if (active) pi = mk(spx[s], spy[s], spz[s]);
  int32_t b[3] = {0, 0, 0};
  if (active) cell_of(k, pi, b);
  float rho = 0.0f;
  Cover cov(k.hh);
  auto term = [&](f3 d) {
    const float r2 = dot3(d, d);
    if (r2 < k.hh) {
      const float w = k.hh - r2;
      rho = rho + k.mass_poly6 * (w * w * w);
    }
    if (kCover) cov.add(d, r2);
  };
  // Cells beyond reach hold only terms that add nothing (the interior radius
  // classes neighbours out to 1.5 h): skipping them keeps every sum bit for bit.
  const f3 oi = cell_rel(k, pi, b);
  const bool mono = !__ballot(active && !in_grid(k, b[0], b[1], b[2]));  // wave_union
  const float reach2 = kCover ? 2.25f * k.hh : k.hh;
  for (int dz = -1; dz <= 1; dz++) {
    for (int dy = -1; dy <= 1; dy++) {
      const RowRun r = row_run(k, cs, epoch, b, wrap_add(b[1], dy), wrap_add(b[2], dz), active);
      bool keep[3];
      row_keep(oi, dy, dz, k.h, reach2, 1.002f, keep);
      const RowSpan sp = kept_span(r, keep);
      const bool any = sp.hi > sp.lo;
      uint32_t ulo, uhi;
      wave_union(r, sp, any, mono, ulo, uhi);
      const bool staged = !__ballot(r.capped) && uhi > ulo && uhi - ulo <= (uint32_t)kS;
      if (staged)
        for (uint32_t t = ulo + lane; t < uhi; t += 64) stage[wave][t - ulo] = make_float4(spx[t], spy[t], spz[t], 0.0f);
      wave_sync();
      if (staged) {
        // The kept cells' candidates in increasing sorted index: the plain walk's
        // order, two per trip with both loaded first.
        uint32_t t = sp.lo;
        for (; t + 2 <= sp.hi; t += 2) {
          const float4 q0 = stage[wave][t - ulo], q1 = stage[wave][t + 1 - ulo];
          term(mk(pi.x - q0.x, pi.y - q0.y, pi.z - q0.z));
          term(mk(pi.x - q1.x, pi.y - q1.y, pi.z - q1.z));
        }
        if (t < sp.hi) {
          const float4 q = stage[wave][t - ulo];
          term(mk(pi.x - q.x, pi.y - q.y, pi.z - q.z));
        }
      } else {
        walk_row(k, r, [&](uint32_t t) { term(mk(pi.x - spx[t], pi.y - spy[t], pi.z - spz[t])); });
      }
      wave_sync();
    }
  }
  if (!active) return;
  srho[s] = fmax_s(rho, k.rest_floor);
  spres[s] = fmax_s(k.stiff * (rho - k.rest), 0.0f);
  if (kCover) co.r2[co.svals[s]] = cov.radius2(k.hh);
}

struct StagedX {
  float4 pr;  // position and 1 / density
  float4 vp;  // velocity and pressure
};

template <int kB, int kS>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(kForcesWaves<kS>)))
    k_forces_integrate_x(FluidK k, SceneK sk, ColliderDev col, uint32_t n, SortedIn in,
                         const uint32_t* __restrict__ svals, const uint32_t* __restrict__ skeys,
                         const uint4* __restrict__ cells, StateOut out, uint32_t* __restrict__ contacts) {
  __shared__ StagedX stage[kB / 64][kS];
  const uint32_t s = (sk.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x) * kB + threadIdx.x;
  const bool active = s < n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float* __restrict__ spx = in.p[0];
  const float* __restrict__ spy = in.p[1];
  const float* __restrict__ spz = in.p[2];
  const float* __restrict__ svx = in.p[3];
  const float* __restrict__ svy = in.p[4];
  const float* __restrict__ svz = in.p[5];
  const float* __restrict__ srho = in.p[6];
  const float* __restrict__ spres = in.p[7];
  const uint32_t epoch = *sk.epoch;
  const uint32_t scene = active ? scene_at(sk, skeys, s) : 0u;
  const uint4* cs = scene_cells(sk, cells, scene);
  f3 pi = mk(0.0f, 0.0f, 0.0f), vi = mk(0.0f, 0.0f, 0.0f);
  float rhoi = 1.0f, pri = 0.0f;
  int32_t b[3] = {0, 0, 0};
  if (active) {
    pi = mk(spx[s], spy[s], spz[s]);
    vi = mk(svx[s], svy[s], svz[s]);
    rhoi = srho[s];
    pri = spres[s];
    cell_of(k, pi, b);
  }
  f3 fp = mk(0.0f, 0.0f, 0.0f);
  f3 vsum = vi * k.h6;
  float wsum = k.h6;
  // k_forces_integrate's term, operation for operation.
  auto term = [&](uint32_t t, f3 dv, f3 vj, float inv, float prj) {
    if (t == s) return;
    const float r2 = dot3(dv, dv);
    if (r2 >= k.hh) return;
    const float w0 = k.hh - r2;
    const float w = w0 * w0 * w0;
    vsum = mk(vsum.x + vj.x * w, vsum.y + vj.y * w, vsum.z + vj.z * w);
    wsum = wsum + w;
    const float r = sqrtf(r2);
    if (r > 1e-6f) {
      const float hr = k.h - r;
      const float s2 = k.mass * (pri + prj) * 0.5f * inv * k.spiky * (hr * hr);
      const float ir = 1.0f / r;
      fp = mk(fp.x - dv.x * ir * s2, fp.y - dv.y * ir * s2, fp.z - dv.z * ir * s2);
    }
  };
  // Cells beyond h hold only terms that add nothing: skipping them keeps every sum.
  const f3 oi = cell_rel(k, pi, b);
  const bool mono = !__ballot(active && !in_grid(k, b[0], b[1], b[2]));  // wave_union
  for (int dz = -1; dz <= 1; dz++) {
    for (int dy = -1; dy <= 1; dy++) {
      const RowRun r = row_run(k, cs, epoch, b, wrap_add(b[1], dy), wrap_add(b[2], dz), active);
      bool keep[3];
      row_keep(oi, dy, dz, k.h, k.hh, 1.002f, keep);
      const RowSpan sp = kept_span(r, keep);
      const bool any = sp.hi > sp.lo;
      uint32_t ulo, uhi;
      wave_union(r, sp, any, mono, ulo, uhi);
      const bool staged = !__ballot(r.capped) && uhi > ulo && uhi - ulo <= (uint32_t)kS;
      if (staged)
        for (uint32_t t = ulo + lane; t < uhi; t += 64) {
          StagedX e;
          e.pr = make_float4(spx[t], spy[t], spz[t], 1.0f / srho[t]);
          e.vp = make_float4(svx[t], svy[t], svz[t], spres[t]);
          stage[wave][t - ulo] = e;
        }
      wave_sync();
      if (staged) {
        // In sorted order, two per trip with both positions loaded first.
        auto cand = [&](uint32_t t, float4 pr) {
          const StagedX& e = stage[wave][t - ulo];
          term(t, mk(pi.x - pr.x, pi.y - pr.y, pi.z - pr.z), mk(e.vp.x, e.vp.y, e.vp.z), pr.w, e.vp.w);
        };
        uint32_t t = sp.lo;
        for (; t + 2 <= sp.hi; t += 2) {
          const float4 p0 = stage[wave][t - ulo].pr, p1 = stage[wave][t + 1 - ulo].pr;
          cand(t, p0);
          cand(t + 1, p1);
        }
        if (t < sp.hi) cand(t, stage[wave][t - ulo].pr);
      } else {
        walk_row(k, r, [&](uint32_t t) {
          term(t, mk(pi.x - spx[t], pi.y - spy[t], pi.z - spz[t]), mk(svx[t], svy[t], svz[t]), 1.0f / srho[t],
               spres[t]);
        });
      }
      wave_sync();
    }
  }
  if (!active) return;
  finish_particle(k, sk, col, scene, s, pi, vi, rhoi, pri, fp, vsum, wsum, svals, out,
                  contacts + (size_t)scene * sk.contact_stride);
}

template <int kB, int kS>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(kForcesWaves<kS>)))
    k_forces_integrate_lds(FluidK k, SceneK sk, ColliderDev col, uint32_t n, SortedIn in,
                           const uint32_t* __restrict__ svals, const uint32_t* __restrict__ skeys,
                           const uint4* __restrict__ cells, StateOut out, uint32_t* __restrict__ contacts,
                           const StagedNbr* __restrict__ frec) {
#pragma clang fp contract(fast)
  // The staged records as three arrays (position, velocity and 1 / density,
  // pressure / density): lanes reading different neighbours then hit different
  // banks, where 32-byte records put every fourth record on the same banks.
  __shared__ uint2 st_pos[kB / 64][kS];
  __shared__ float4 st_vr[kB / 64][kS];
  __shared__ float st_pinv[kB / 64][kS];
  const uint32_t s = (sk.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x) * kB + threadIdx.x;
  const bool active = s < n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float* __restrict__ spx = in.p[0];
  const float* __restrict__ spy = in.p[1];
  const float* __restrict__ spz = in.p[2];
  const float* __restrict__ svx = in.p[3];
  const float* __restrict__ svy = in.p[4];
  const float* __restrict__ svz = in.p[5];
  const float* __restrict__ srho = in.p[6];
  const float* __restrict__ spres = in.p[7];
  const uint32_t epoch = *sk.epoch;
  const uint32_t scene = active ? scene_at(sk, skeys, s) : 0u;
  const uint4* cs = scene_cells(sk, cells, scene);
  f3 pi = mk(0.0f, 0.0f, 0.0f), vi = mk(0.0f, 0.0f, 0.0f);
  float rhoi = 1.0f, pri = 0.0f;
  int32_t b[3] = {0, 0, 0};
  if (active) {
    pi = mk(spx[s], spy[s], spz[s]);
    vi = mk(svx[s], svy[s], svz[s]);
    rhoi = srho[s];
    pri = spres[s];
    cell_of(k, pi, b);
  }
  const f3 oi = cell_rel(k, pi, b);
  const bool mono = !__ballot(active && !in_grid(k, b[0], b[1], b[2]));  // wave_union
  f3 fp = mk(0.0f, 0.0f, 0.0f);
  f3 vsum = vi * k.h6;
  float wsum = k.h6;
  // The neighbour's 1 / density and pressure / density come in precomputed (once
  // per staged particle, not once per pair), r and 1 / r come from the hardware
  // reciprocal square root (1 ulp), r > 1e-6 is tested as r^2 > 1e-12, and the
  // constant factor -m spiky / 2 multiplies the summed pressure force once.  The
  // constants sit in VGPRs (full-rate operands).  The branch on r2 < hh is kept:
  // consecutive sorted lanes walk the same candidates in step, so a candidate
  // beyond h is usually beyond h for the whole wave and its work is skipped.
  const float hh_v = vreg(k.hh), h_v = vreg(k.h), r2_min = vreg(1e-12f), close2 = vreg(0.01f * k.hh);
  auto term = [&](f3 dv, f3 vj, float inv, float pinv, uint32_t t) {
    float r2 = dot3_fma(dv);
    if (!(r2 < hh_v)) return;
    // A near-coincident pair (closer than h / 10, where the staged offsets' f16
    // rounding, up to ~1.5e-4 h, would tilt the pair's direction by more than a
    // thousandth) takes its offset from the f32 snapshot instead.
    if (r2 < close2) {
      dv = mk(pi.x - spx[t], pi.y - spy[t], pi.z - spz[t]);
      r2 = dot3_fma(dv);
    }
    const float w0 = hh_v - r2;
    const float w = w0 * w0 * w0;
    vsum = mk(vsum.x + vj.x * w, vsum.y + vj.y * w, vsum.z + vj.z * w);
    wsum = wsum + w;
    if (r2 > r2_min) {
      const float ir = __builtin_amdgcn_rsqf(r2);
      const float hr = h_v - r2 * ir;
      const float f = (pri * inv + pinv) * (hr * hr) * ir;
      fp = mk(fp.x + dv.x * f, fp.y + dv.y * f, fp.z + dv.z * f);
    }
  };
  for (int dz = -1; dz <= 1; dz++) {
    for (int dy = -1; dy <= 1; dy++) {
      const RowRun r = row_run(k, cs, epoch, b, wrap_add(b[1], dy), wrap_add(b[2], dz), active);
      const RowSpan sp = row_span(r, oi, dy, dz, k.h, k.hh);
      const bool any = sp.hi > sp.lo;
      uint32_t ulo, uhi;
      wave_union(r, sp, any, mono, ulo, uhi);
      const bool staged = !__ballot(r.capped) && uhi > ulo && uhi - ulo <= (uint32_t)kS;
      if (staged)
        for (uint32_t t = ulo + lane; t < uhi; t += 64) {
          const StagedNbr e = frec[t];
          st_pos[wave][t - ulo] = e.pos;
          st_vr[wave][t - ulo] = e.vr;
          st_pinv[wave][t - ulo] = e.pinv;
        }
      wave_sync();
      if (staged) {
        const float by = oi.y - (float)dy * k.h, bz = oi.z - (float)dz * k.h;
        auto cand = [&](uint32_t j, uint2 pos) {
          const float4 vr = st_vr[wave][j];
          term(mk(pi.x - __uint_as_float(pos.x), sub_f16_lo(by, pos.y), sub_f16_hi(bz, pos.y)),
               mk(vr.x, vr.y, vr.z), vr.w, st_pinv[wave][j], ulo + j);
        };
        // Four candidates per trip (then two, then one), all positions loaded before
        // any is used: one LDS round trip and one loop test for four candidates.
        // [a, e_) lies in the staged [ulo, uhi) when it is not empty; an empty range
        // (a lane whose row holds no candidate has a = 0xFFFFFFFF) walks nothing.
        auto walk = [&](uint32_t a, uint32_t e_) {
          const uint32_t cnt = e_ > a ? e_ - a : 0u;
          const uint32_t base = cnt ? a - ulo : 0u;
          uint32_t j = 0;
          for (; j + 4 <= cnt; j += 4) {
            const uint2 p0 = st_pos[wave][base + j], p1 = st_pos[wave][base + j + 1];
            const uint2 p2 = st_pos[wave][base + j + 2], p3 = st_pos[wave][base + j + 3];
            cand(base + j, p0);
            cand(base + j + 1, p1);
            cand(base + j + 2, p2);
            cand(base + j + 3, p3);
          }
          if (j + 2 <= cnt) {
            const uint2 p0 = st_pos[wave][base + j], p1 = st_pos[wave][base + j + 1];
            cand(base + j, p0);
            cand(base + j + 1, p1);
            j += 2;
          }
          if (j < cnt) cand(base + j, st_pos[wave][base + j]);
        };
        // The self pair is left out by the bounds: [lo, s) and (s, hi).  A row that
        // does not hold s has s outside [lo, hi), and one of the two ranges is empty.
        walk(sp.lo, s < sp.hi ? s : sp.hi);
        walk(s + 1 > sp.lo ? s + 1 : sp.lo, sp.hi);
      } else {
        walk_row(k, r, [&](uint32_t t) {
          if (t == s) return;
          const float inv = 1.0f / srho[t];
          term(pi - mk(spx[t], spy[t], spz[t]), mk(svx[t], svy[t], svz[t]), inv, spres[t] * inv, t);
        });
      }
      wave_sync();
    }
  }
  if (!active) return;
  fp = fp * -(k.mass * 0.5f * k.spiky);
  finish_particle(k, sk, col, scene, s, pi, vi, rhoi, pri, fp, vsum, wsum, svals, out,
                  contacts + (size_t)scene * sk.contact_stride);
}

// ---------------------------------------------------------------------------
// FP64 conserved quantities: total mass, momentum and centre of mass, reduced in
// a fixed order (per-block partials, then one block) so the result is the same
// on every run.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
    k_conserved_partial(uint32_t n, const float* __restrict__ px, const float* __restrict__ py,
                        const float* __restrict__ pz, const float* __restrict__ vx,
                        const float* __restrict__ vy, const float* __restrict__ vz,
                        double* __restrict__ partial) {
  __shared__ double sh[6][kBlock];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    acc[0] += (double)vx[i];
    acc[1] += (double)vy[i];
    acc[2] += (double)vz[i];
    acc[3] += (double)px[i];
    acc[4] += (double)py[i];
    acc[5] += (double)pz[i];
  }
  for (int q = 0; q < 6; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < 6; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void __launch_bounds__(kBlock)
    k_conserved_final(uint32_t blocks, uint32_t n, double mass, const double* __restrict__ partial,
                      double* __restrict__ out) {
  __shared__ double sh[6][kBlock];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t i = threadIdx.x; i < blocks; i += kBlock)
    for (int q = 0; q < 6; q++) acc[q] += partial[i * 6 + q];
  for (int q = 0; q < 6; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < 6; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = mass * (double)n;
    out[0] = m;                                  // total mass
    for (int q = 0; q < 3; q++) out[1 + q] = mass * sh[q][0];           // momentum
    for (int q = 0; q < 3; q++) out[4 + q] = n ? sh[3 + q][0] / n : 0;  // centre of mass
  }
}

// hash1 / hash3 (math/gpu.ts:103-117): lowbias32, low 24 bits scaled to [0,1).
__device__ __forceinline__ float hash1(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return (float)(x & 0xffffffu) * (1.0f / 16777216.0f);
}

struct EmitK {
  uint32_t base, count, cap, seed;
  float pos[3], radius, vel[3];
};

// emit (fluid.ts:268-291): jittered ball around emitPos, written into the ring.
__global__ void __launch_bounds__(kBlock) k_emit(EmitK e, StateOut out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= e.count) return;
  const uint32_t slot = (e.base + i) % e.cap;
  const uint32_t sd = e.seed * 9781u + i;
  const f3 r = mk(hash1(sd * 3u + 0u), hash1(sd * 3u + 1u), hash1(sd * 3u + 2u));
  const f3 v = mk(r.x * 2.0f - 1.0f + 1e-4f, r.y * 2.0f - 1.0f + 1e-4f, r.z * 2.0f - 1.0f + 1e-4f);
  const float len = len3(v);
  const f3 dir = mk(v.x / len, v.y / len, v.z / len);
  const float s = e.radius * powf(r.x, 1.0f / 3.0f);
  out.p[0][slot] = e.pos[0] + dir.x * s;
  out.p[1][slot] = e.pos[1] + dir.y * s;
  out.p[2][slot] = e.pos[2] + dir.z * s;
  out.p[3][slot] = e.vel[0];
  out.p[4][slot] = e.vel[1];
  out.p[5][slot] = e.vel[2];
  out.p[6][slot] = 0.0f;
  out.p[7][slot] = 0.0f;
  out.p[8][slot] = e.vel[0];
  out.p[9][slot] = e.vel[1];
  out.p[10][slot] = e.vel[2];
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Two fills in one launch: words [0, na) of a, then [0, nb) of b.
__global__ void k_fill2_u32(uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < na + nb;
       i += (size_t)gridDim.x * blockDim.x) {
    if (i < na)
      a[i] = v;
    else
      b[i - na] = v;
  }
}

// The contact list (fluid.ts:54-58, 470-477) of each scene: its first 32 contacts
// by particle id, gathered from the per-id flags the integrate pass set.  The
// reference keeps whichever 32 its atomic counter admitted first; ascending id is
// one admissible order, it does not depend on scheduling, and it is the order the
// oracle appends in.  One workgroup per scene scans the scene's flag words.
__global__ void __launch_bounds__(kBlock)
    k_contacts_collect(uint32_t* __restrict__ contacts, uint32_t stride, const uint32_t* __restrict__ cmask,
                       const float4* __restrict__ cdata, const uint32_t* __restrict__ base, uint32_t n) {
  constexpr int kPer = 4;  // flag words per thread and round
  __shared__ uint32_t wsum[kBlock / 64];
  __shared__ uint32_t found;
  const uint32_t sc = blockIdx.x;
  const uint32_t lo = base ? base[sc] : 0u, hi = base ? base[sc + 1] : n;
  float4* pts = reinterpret_cast<float4*>(contacts + (size_t)sc * stride + 4);
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  if (t == 0) found = 0;
  __syncthreads();
  if (lo >= hi) return;
  const uint32_t w0 = lo >> 5, w1 = (hi + 31u) >> 5;
  for (uint32_t wb = w0; wb < w1; wb += kPer * kBlock) {
    uint32_t bits[kPer];
    uint32_t cnt = 0;
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const uint32_t w = wb + (uint32_t)(kPer * t + q);
      uint32_t m = 0;
      if (w < w1) {
        m = cmask[w];
        if (w == w0) m &= ~0u << (lo & 31u);
        const uint32_t last = hi - (w << 5);  // ids of this word below hi
        if (last < 32u) m &= (1u << last) - 1u;
      }
      bits[q] = m;
      cnt += (uint32_t)__popc(m);
    }
    // Exclusive rank of this thread's first flag: a wave scan, then the earlier waves.
    uint32_t incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t u = __shfl_up(incl, d);
      if (lane >= d) incl += u;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t r = found + incl - cnt, total = found;
    for (int w = 0; w < kBlock / 64; w++) {
      if (w < wave) r += wsum[w];
      total += wsum[w];
    }
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      uint32_t m = bits[q];
      const uint32_t w = wb + (uint32_t)(kPer * t + q);
      while (m && r < (uint32_t)CLAY_CONTACTS) {
        const uint32_t bit = (uint32_t)__ffs(m) - 1u;
        pts[r++] = cdata[(w << 5) + bit];
        m &= m - 1u;
      }
    }
    __syncthreads();
    if (t == 0) found = total;
    __syncthreads();
    if (found >= (uint32_t)CLAY_CONTACTS) break;
  }
}

inline unsigned grid_for(uint32_t n, int block = kBlock) { return (n + block - 1) / block; }

// Kernel family of a neighbour pass.
enum NbrMode { kGlobal = 0, kHalfLds = 1, kExactLds = 2 };

// What the FP16-LDS density pass reads and writes beside the plain kernel's arguments.
struct HalfRecs {
  const uint2* spos;
  const float *svx, *svy, *svz;
  StagedNbr* frec;
};

template <int kB, bool kCover>
void launch_density_t(hipStream_t s, const FluidK& k, const SceneK& sk, uint32_t n, const float* spx,
                      const float* spy, const float* spz, const uint32_t* skeys, const uint4* cells,
                      float* srho, float* spres, int mode, int stage, CoverOut co, const HalfRecs& hr) {
  if (mode == kGlobal) {
    hipLaunchKernelGGL((k_density<kB, kCover>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, n, spx, spy, spz,
                       skeys, cells, srho, spres, co);
  } else if (mode == kExactLds) {
    if (stage == 64)
      hipLaunchKernelGGL((k_density_x<kB, 64, kCover>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, n, spx, spy,
                         spz, skeys, cells, srho, spres, co);
    else if (stage == 256)
      hipLaunchKernelGGL((k_density_x<kB, 256, kCover>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, n, spx,
                         spy, spz, skeys, cells, srho, spres, co);
    else
      hipLaunchKernelGGL((k_density_x<kB, 128, kCover>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, n, spx,
                         spy, spz, skeys, cells, srho, spres, co);
  } else if (stage == 64) {
    hipLaunchKernelGGL((k_density_lds<kB, 64, kCover>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, n, spx,
                       spy, spz, skeys, cells, srho, spres, co, hr.spos, hr.svx, hr.svy, hr.svz, hr.frec);
  } else if (stage == 256) {
    hipLaunchKernelGGL((k_density_lds<kB, 256, kCover>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, n, spx,
                       spy, spz, skeys, cells, srho, spres, co, hr.spos, hr.svx, hr.svy, hr.svz, hr.frec);
  } else {
    hipLaunchKernelGGL((k_density_lds<kB, 128, kCover>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, n, spx,
                       spy, spz, skeys, cells, srho, spres, co, hr.spos, hr.svx, hr.svy, hr.svz, hr.frec);
  }
}

template <int kB>
void launch_forces_t(hipStream_t s, const FluidK& k, const SceneK& sk, const ColliderDev& col, uint32_t n,
                     const SortedIn& in, const uint32_t* svals, const uint32_t* skeys, const uint4* cells,
                     const StateOut& o, uint32_t* contacts, int mode, int stage, const StagedNbr* frec) {
  if (mode == kGlobal) {
    hipLaunchKernelGGL(k_forces_integrate<kB>, dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, col, n, in, svals,
                       skeys, cells, o, contacts);
  } else if (mode == kExactLds) {
    if (stage == 64)
      hipLaunchKernelGGL((k_forces_integrate_x<kB, 64>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, col, n, in,
                         svals, skeys, cells, o, contacts);
    else if (stage == 256)
      hipLaunchKernelGGL((k_forces_integrate_x<kB, 256>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, col, n,
                         in, svals, skeys, cells, o, contacts);
    else
      hipLaunchKernelGGL((k_forces_integrate_x<kB, 128>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, col, n,
                         in, svals, skeys, cells, o, contacts);
  } else if (stage == 64) {
    hipLaunchKernelGGL((k_forces_integrate_lds<kB, 64>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, col, n,
                       in, svals, skeys, cells, o, contacts, frec);
  } else if (stage == 256) {
    hipLaunchKernelGGL((k_forces_integrate_lds<kB, 256>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, col,
                       n, in, svals, skeys, cells, o, contacts, frec);
  } else {
    hipLaunchKernelGGL((k_forces_integrate_lds<kB, 128>), dim3(grid_for(n, kB)), dim3(kB), 0, s, k, sk, col,
                       n, in, svals, skeys, cells, o, contacts, frec);
  }
}

}  // namespace

bool fluid_tune_valid(const FluidTune& t) {
  return (t.block == 128 || t.block == 256) && (t.stage == 64 || t.stage == 128 || t.stage == 256);
}

int launch_keys(Ctx* c, hipStream_t s, const FluidK& k, uint32_t n, const float* px, const float* py,
                const float* pz, const uint32_t* scene_of, uint32_t* keys, uint32_t* vals, uint32_t invalid_key,
                uint32_t shift, uint32_t* bounds, uint32_t* bump) {
  if (n == 0) return CLAY_OK;
  LaunchTimer t(c, s, "fluid_keys");
  hipLaunchKernelGGL(k_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, k, n, px, py, pz, scene_of, keys, vals,
                     invalid_key, shift, bounds, bump);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_reorder(Ctx* c, hipStream_t s, uint32_t n, const uint32_t* skeys, const uint32_t* svals,
                   const SceneK& sk, uint32_t invalid_key, uint32_t* cells, const float* const* src,
                   float* const* dst, int planes, const FluidK& k, uint2* spos) {
  if (n == 0) return CLAY_OK;
  Planes8 pl{};
  for (int p = 0; p < planes && p < 8; p++) {
    pl.src[p] = src[p];
    pl.dst[p] = dst[p];
  }
  LaunchTimer t(c, s, "fluid_reorder");
  hipLaunchKernelGGL(k_reorder, dim3(grid_for(n)), dim3(kBlock), 0, s, n, skeys, svals, sk, invalid_key, cells,
                     pl, planes, k, spos);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_density(Ctx* c, hipStream_t s, const FluidK& k, const SceneK& sk, const FluidTune& tune, bool lds,
                   uint32_t n, const float* spx, const float* spy, const float* spz, const uint32_t* skeys,
                   const uint32_t* cells, float* srho, float* spres, const uint32_t* svals, float* cover,
                   const HalfState& half) {
  if (n == 0) return CLAY_OK;
  const uint4* cl = reinterpret_cast<const uint4*>(cells);
  const int mode = lds ? kHalfLds : (tune.exact_lds ? kExactLds : kGlobal);
  if (lds && (!half.spos || !half.frec)) return fail(CLAY_ERR_INVALID, "FP16-LDS density: staging records missing");
  const CoverOut co{svals, cover};
  const HalfRecs hr{reinterpret_cast<const uint2*>(half.spos), half.svx, half.svy, half.svz,
                    reinterpret_cast<StagedNbr*>(half.frec)};
  LaunchTimer t(c, s, lds ? "fluid_density_lds" : "fluid_density");
  if (cover) {
    if (tune.block == 128)
      launch_density_t<128, true>(s, k, sk, n, spx, spy, spz, skeys, cl, srho, spres, mode, tune.stage, co, hr);
    else
      launch_density_t<256, true>(s, k, sk, n, spx, spy, spz, skeys, cl, srho, spres, mode, tune.stage, co, hr);
  } else if (tune.block == 128) {
    launch_density_t<128, false>(s, k, sk, n, spx, spy, spz, skeys, cl, srho, spres, mode, tune.stage, co, hr);
  } else {
    launch_density_t<256, false>(s, k, sk, n, spx, spy, spz, skeys, cl, srho, spres, mode, tune.stage, co, hr);
  }
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_forces_integrate(Ctx* c, hipStream_t s, const FluidK& k, const SceneK& sk, const FluidTune& tune,
                            bool lds, const ColliderDev& col, uint32_t n, const float* const* sorted,
                            const uint32_t* svals, const uint32_t* skeys, const uint32_t* cells,
                            float* const* out, uint32_t* contacts, const float4* frec) {
  if (n == 0) return CLAY_OK;
  if (lds && !frec) return fail(CLAY_ERR_INVALID, "FP16-LDS forces: staging records missing");
  SortedIn in{};
  for (int p = 0; p < 8; p++) in.p[p] = sorted[p];
  StateOut o{};
  for (int p = 0; p < 11; p++) o.p[p] = out[p];
  const uint4* cl = reinterpret_cast<const uint4*>(cells);
  const int mode = lds ? kHalfLds : (tune.exact_lds ? kExactLds : kGlobal);
  LaunchTimer t(c, s, lds ? "fluid_forces_integrate_lds" : "fluid_forces_integrate");
  if (tune.block == 128)
    launch_forces_t<128>(s, k, sk, col, n, in, svals, skeys, cl, o, contacts, mode, tune.stage, reinterpret_cast<const StagedNbr*>(frec));
  else
    launch_forces_t<256>(s, k, sk, col, n, in, svals, skeys, cl, o, contacts, mode, tune.stage, reinterpret_cast<const StagedNbr*>(frec));
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_contacts_collect(Ctx* c, hipStream_t s, uint32_t* contacts, uint32_t stride, const uint32_t* cmask,
                            const float4* cdata, const uint32_t* base, uint32_t n, uint32_t scenes) {
  LaunchTimer t(c, s, "fluid_contacts");
  hipLaunchKernelGGL(k_contacts_collect, dim3(scenes), dim3(kBlock), 0, s, contacts, stride, cmask, cdata, base, n);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_emit(Ctx* c, hipStream_t s, Fluid* f, uint32_t base, uint32_t count, uint32_t seed) {
  if (count == 0) return CLAY_OK;
  EmitK e{};
  e.base = base;
  e.count = count;
  e.cap = f->capacity;
  e.seed = seed;
  for (int a = 0; a < 3; a++) {
    e.pos[a] = (float)f->params.emit_pos[a];
    e.vel[a] = (float)f->params.emit_vel[a];
  }
  e.radius = (float)f->params.emit_radius;
  StateOut o{};
  float* planes[11] = {f->px, f->py, f->pz, f->vx, f->vy, f->vz, f->rho, f->pres, f->nvx, f->nvy, f->nvz};
  for (int p = 0; p < 11; p++) o.p[p] = planes[p];
  LaunchTimer t(c, s, "fluid_emit");
  hipLaunchKernelGGL(k_emit, dim3(grid_for(count)), dim3(kBlock), 0, s, e, o);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_fill2_u32(Ctx* c, hipStream_t s, uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t v) {
  if (na + nb == 0) return CLAY_OK;
  size_t blocks = (na + nb + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  LaunchTimer t(c, s, "fill_u32");
  hipLaunchKernelGGL(k_fill2_u32, dim3((unsigned)blocks), dim3(kBlock), 0, s, a, na, b, nb, v);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_fill_u32(Ctx* c, hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
  if (n == 0) return CLAY_OK;
  size_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  LaunchTimer t(c, s, "fill_u32");
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, n, v);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_conserved(Ctx* c, hipStream_t s, uint32_t n, const float* const* pv, double mass,
                     double* partial, uint32_t partial_blocks, double* out) {
  LaunchTimer t(c, s, "fluid_conserved_fp64");
  hipLaunchKernelGGL(k_conserved_partial, dim3(partial_blocks), dim3(kBlock), 0, s, n, pv[0], pv[1], pv[2],
                     pv[3], pv[4], pv[5], partial);
  hipLaunchKernelGGL(k_conserved_final, dim3(1), dim3(kBlock), 0, s, partial_blocks, n, mass, partial, out);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

}  // namespace clay
