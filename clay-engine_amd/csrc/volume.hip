// volume.hip -- SdfVolume-side kernels: dense brush bake (SdfBuilder.gen), collider
// evaluation for tests, and a device-side setOrigin for following volumes.
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kMaxTileBrushes = 64;  // builder.ts:7

struct BakeMipK {
  int32_t r;
  float org[3];
  float tile_size, tile_half, voxel, band_world, empty;
};

// SdfBuilder.gen[m] (builder.ts:87-142): one workgroup per 8^3 tile, 8x8x4 threads,
// lane 0 culls the brush list into LDS in brush order, every voxel folds it.
__global__ void __launch_bounds__(256)
    k_bake_brushes(BakeMipK k, const clay_brush* __restrict__ brushes, int nb,
                   uint16_t* __restrict__ dist, uint16_t* __restrict__ mat) {
  __shared__ int ids[kMaxTileBrushes];
  __shared__ int count;
  const int wx = blockIdx.x, wy = blockIdx.y, wz = blockIdx.z;
  const f3 tmin = mk(k.org[0] + (float)wx * k.tile_size, k.org[1] + (float)wy * k.tile_size,
                     k.org[2] + (float)wz * k.tile_size);
  const f3 tc = mk(tmin.x + k.tile_half, tmin.y + k.tile_half, tmin.z + k.tile_half);
  if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) {
    int n = 0;
    for (int b = 0; b < nb; b++) {
      const clay_brush& br = brushes[b];
      const f3 nr = max30(mk(fabsf(br.pos[0] - tc.x) - k.tile_half, fabsf(br.pos[1] - tc.y) - k.tile_half,
                             fabsf(br.pos[2] - tc.z) - k.tile_half));
      if (len3(nr) < br.bound + k.band_world && n < kMaxTileBrushes) ids[n++] = b;
    }
    count = n;
  }
  __syncthreads();
  const int n = count;
  for (int half = 0; half < 2; half++) {
    const int lx = threadIdx.x, ly = threadIdx.y, lz = threadIdx.z * 2 + half;
    const f3 p = mk(tmin.x + ((float)lx + 0.5f) * k.voxel, tmin.y + ((float)ly + 0.5f) * k.voxel,
                    tmin.z + ((float)lz + 0.5f) * k.voxel);
    float acc[2] = {k.empty, 0.0f};
    for (int i = 0; i < n; i++) apply_brush(acc, brushes[ids[i]], p);
    const size_t x = (size_t)wx * 8 + lx, y = (size_t)wy * 8 + ly, z = (size_t)wz * 8 + lz;
    const size_t idx = x + (size_t)k.r * (y + (size_t)k.r * z);
    dist[idx] = f2h(fclamp_s(acc[0] / k.band_world, -1.0f, 1.0f));
    mat[idx] = f2h(acc[1]);
  }
}

__global__ void k_field_eval(ColliderDev col, int n, const float* __restrict__ pts,
                             float* __restrict__ dist, float* __restrict__ nrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 p = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  float d;
  const int leaf = collider_pick(col, 0u, p, &d);
  const f3 g = collider_normal(col, 0u, leaf, p);
  dist[i] = d;
  nrm[3 * i] = g.x;
  nrm[3 * i + 1] = g.y;
  nrm[3 * i + 2] = g.z;
}

// volume.ts:327-334: centre = origin + worldSize / 2 in f64, stored as f32.
__global__ void k_set_origin(VolDev* v, const float* __restrict__ o, double half) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int c = 0; c < 3; c++) {
      const float oc = o[c];
      v->origin[c] = oc;
      v->center[c] = (float)((double)oc + half);
    }
  }
}

}  // namespace

int launch_bake_brushes(Ctx* c, hipStream_t s, Volume* v, const clay_brush* brushes_dev, int n) {
  const double voxel0 = v->info.world_size / v->info.resolution;
  for (int m = 0; m < v->info.mip_levels; m++) {
    BakeMipK k{};
    k.r = v->mip_res(m);
    const double voxel = voxel0 * (double)(1 << m);
    const double tile = voxel * kTile;
    const double bw = v->info.band * voxel;
    for (int a = 0; a < 3; a++) k.org[a] = (float)v->origin[a];
    k.tile_size = (float)tile;
    k.tile_half = (float)(tile / 2);
    k.voxel = (float)voxel;
    k.band_world = (float)bw;
    k.empty = (float)(bw * 4);
    const int tiles = k.r / kTile;
    LaunchTimer t(c, s, "volume_bake_brushes");
    hipLaunchKernelGGL(k_bake_brushes, dim3(tiles, tiles, tiles), dim3(8, 8, 4), 0, s, k,
                       brushes_dev, n, v->dist + v->voxel_off[m], v->mat + v->voxel_off[m]);
    CLAY_HIP_CHECK(hipGetLastError());
  }
  return CLAY_OK;
}

int launch_field_eval(Ctx* c, hipStream_t s, const ColliderDev& col, int n, const float* pts,
                      float* dist, float* nrm) {
  if (n <= 0) return CLAY_OK;
  LaunchTimer t(c, s, "field_eval");
  hipLaunchKernelGGL(k_field_eval, dim3((n + 255) / 256), dim3(256), 0, s, col, n, pts, dist, nrm);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

int launch_set_origin_dev(Ctx* c, hipStream_t s, VolDev* dev, const float* origin3, double half) {
  LaunchTimer t(c, s, "volume_set_origin");
  hipLaunchKernelGGL(k_set_origin, dim3(1), dim3(64), 0, s, dev, origin3, half);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

}  // namespace clay
