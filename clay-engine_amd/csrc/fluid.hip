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
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kBlock = 256;

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

__global__ void __launch_bounds__(kBlock) k_keys(FluidK k, uint32_t n, const float* __restrict__ px,
                                                 const float* __restrict__ py,
                                                 const float* __restrict__ pz,
                                                 uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals, uint32_t invalid) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  int32_t q[3];
  cell_of(k, mk(px[i], py[i], pz[i]), q);
  keys[i] = in_grid(k, q[0], q[1], q[2]) ? cell_index(k, q[0], q[1], q[2]) : invalid;
  vals[i] = i;
}

struct Planes8 {
  const float* src[8];
  float* dst[8];
};

// Cell records of this substep (run boundaries of the sorted keys) and the sorted
// snapshot of the state planes.
__global__ void __launch_bounds__(kBlock) k_reorder(uint32_t n, const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ svals,
                                                    uint32_t invalid, uint32_t* __restrict__ cells,
                                                    uint32_t epoch, Planes8 pl, int planes) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t key = skeys[s];
  if (key != invalid) {
    if (s == 0 || skeys[s - 1] != key) {
      cells[4 * (size_t)key + 0] = s;
      cells[4 * (size_t)key + 2] = epoch;
    }
    if (s == n - 1 || skeys[s + 1] != key) cells[4 * (size_t)key + 1] = s + 1;
  }
  const uint32_t j = svals[s];
#pragma unroll
  for (int p = 0; p < 8; p++)
    if (p < planes) pl.dst[p][s] = pl.src[p][j];
}

__global__ void __launch_bounds__(kBlock)
    k_density(FluidK k, uint32_t n, const float* __restrict__ spx, const float* __restrict__ spy,
              const float* __restrict__ spz, const uint4* __restrict__ cells, uint32_t epoch,
              float* __restrict__ srho, float* __restrict__ spres) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const f3 pi = mk(spx[s], spy[s], spz[s]);
  int32_t b[3];
  cell_of(k, pi, b);
  float rho = 0.0f;
  for_neighbours(k, cells, epoch, b, [&](uint32_t t) {
    const f3 d = mk(pi.x - spx[t], pi.y - spy[t], pi.z - spz[t]);
    const float r2 = dot3(d, d);
    if (r2 < k.hh) {
      const float w = k.hh - r2;
      rho = rho + k.mass_poly6 * (w * w * w);
    }
  });
  srho[s] = fmax_s(rho, k.rest_floor);
  spres[s] = fmax_s(k.stiff * (rho - k.rest), 0.0f);
}

struct SortedIn {
  const float* p[8];  // spx spy spz svx svy svz srho spres
};
struct StateOut {
  float* p[11];  // px py pz vx vy vz rho pres nvx nvy nvz
};

__global__ void __launch_bounds__(kBlock)
    k_forces_integrate(FluidK k, ColliderDev col, uint32_t n, SortedIn in,
                       const uint32_t* __restrict__ svals, const uint4* __restrict__ cells,
                       uint32_t epoch, StateOut out, uint32_t* __restrict__ contacts) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const float* __restrict__ spx = in.p[0];
  const float* __restrict__ spy = in.p[1];
  const float* __restrict__ spz = in.p[2];
  const float* __restrict__ svx = in.p[3];
  const float* __restrict__ svy = in.p[4];
  const float* __restrict__ svz = in.p[5];
  const float* __restrict__ srho = in.p[6];
  const float* __restrict__ spres = in.p[7];

  const f3 pi = mk(spx[s], spy[s], spz[s]);
  const f3 vi = mk(svx[s], svy[s], svz[s]);
  const float rhoi = srho[s], pri = spres[s];
  int32_t b[3];
  cell_of(k, pi, b);
  f3 fp = mk(0.0f, 0.0f, 0.0f);
  f3 vsum = vi * k.h6;
  float wsum = k.h6;
  for_neighbours(k, cells, epoch, b, [&](uint32_t t) {
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
    const int leaf = collider_pick(col, p, &dw);
    if (dw < k.radius) {
      const f3 nr = collider_normal(col, leaf, p);
      const float push = k.radius - dw;
      p = mk(p.x + nr.x * push, p.y + nr.y * push, p.z + nr.z * push);
      const float vn = fmin_s(dot3(v, nr), 0.0f);
      v = mk(v.x - nr.x * vn, v.y - nr.y * vn, v.z - nr.z * vn);
      if (dw < k.cdepth) {
        const uint32_t slot = atomicAdd(contacts, 1u);
        if (slot < CLAY_CONTACTS) {
          float4* pts = reinterpret_cast<float4*>(contacts + 4);
          pts[slot] = make_float4(p.x, p.y, p.z, len3(v));
        }
      }
    }
  }
  const float clx = fmin_s(fmax_s(p.x, k.lo[0]), k.hi[0]);
  const float cly = fmin_s(fmax_s(p.y, k.lo[1]), k.hi[1]);
  const float clz = fmin_s(fmax_s(p.z, k.lo[2]), k.hi[2]);
  const uint32_t j = svals[s];
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

inline unsigned grid_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

int launch_keys(Ctx* c, hipStream_t s, const FluidK& k, uint32_t n, const float* px,
                const float* py, const float* pz, uint32_t* keys, uint32_t* vals,
                uint32_t invalid_key) {
  if (n == 0) return CLAY_OK;
  LaunchTimer t(c, s, "fluid_keys");
  hipLaunchKernelGGL(k_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, k, n, px, py, pz, keys, vals,
                     invalid_key);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_reorder(Ctx* c, hipStream_t s, uint32_t n, const uint32_t* skeys,
                   const uint32_t* svals, uint32_t invalid_key, uint32_t* cells, uint32_t epoch,
                   const float* const* src, float* const* dst, int planes) {
  if (n == 0) return CLAY_OK;
  Planes8 pl{};
  for (int p = 0; p < planes && p < 8; p++) {
    pl.src[p] = src[p];
    pl.dst[p] = dst[p];
  }
  LaunchTimer t(c, s, "fluid_reorder");
  hipLaunchKernelGGL(k_reorder, dim3(grid_for(n)), dim3(kBlock), 0, s, n, skeys, svals,
                     invalid_key, cells, epoch, pl, planes);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_density(Ctx* c, hipStream_t s, const FluidK& k, uint32_t n, const float* spx,
                   const float* spy, const float* spz, const uint32_t* cells, uint32_t epoch,
                   float* srho, float* spres) {
  if (n == 0) return CLAY_OK;
  LaunchTimer t(c, s, "fluid_density");
  hipLaunchKernelGGL(k_density, dim3(grid_for(n)), dim3(kBlock), 0, s, k, n, spx, spy, spz,
                     reinterpret_cast<const uint4*>(cells), epoch, srho, spres);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_forces_integrate(Ctx* c, hipStream_t s, const FluidK& k, const ColliderDev& col,
                            uint32_t n, const float* const* sorted, const uint32_t* svals,
                            const uint32_t* cells, uint32_t epoch, float* const* out,
                            uint32_t* contacts) {
  if (n == 0) return CLAY_OK;
  SortedIn in{};
  for (int p = 0; p < 8; p++) in.p[p] = sorted[p];
  StateOut o{};
  for (int p = 0; p < 11; p++) o.p[p] = out[p];
  LaunchTimer t(c, s, "fluid_forces_integrate");
  hipLaunchKernelGGL(k_forces_integrate, dim3(grid_for(n)), dim3(kBlock), 0, s, k, col, n, in,
                     svals, reinterpret_cast<const uint4*>(cells), epoch, o, contacts);
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

int launch_fill_u32(Ctx* c, hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
  if (n == 0) return CLAY_OK;
  size_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  LaunchTimer t(c, s, "fill_u32");
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, n, v);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

}  // namespace clay
