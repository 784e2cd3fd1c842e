// atomic_peak.hip -- issue rate of the two atomics the splat runs on, measured.
//
// The bake's walk is a stream of unsigned min-atomics: ds_min_u32 into an LDS copy
// of a brick (the tiled schedules) and global atomicMin into the code grid (the
// plain schedule, and the tiled schedules' merge).  No datasheet gives their rates
// for gfx950, so these two kernels measure them, and the splat's atomics per
// second are reported as a fraction of the measured figure.
//
//   k_lds_min_peak     every lane min-merges into its own LDS word, 64 consecutive
//                      words per wave: no bank conflicts, the best case of the
//                      walk's LDS traffic.
//   k_global_min_peak  every lane min-merges into its own word of a buffer, a
//                      wave's 64 lanes on one 256-byte span, successive rounds
//                      striding over the whole buffer (far larger than L2).
//
// Both run a fixed grid for a bounded number of rounds; every address is inside
// the allocation by construction (lane index modulo the array size).
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kPeakThreads = 512;
constexpr int kLdsWords = 8192;  // 32 KB per workgroup

__global__ void __launch_bounds__(kPeakThreads) k_lds_min_peak(uint32_t rounds, uint32_t* __restrict__ sink) {
  __shared__ uint32_t lds[kLdsWords];
  const uint32_t t = threadIdx.x;
  for (uint32_t v = t; v < (uint32_t)kLdsWords; v += kPeakThreads) lds[v] = 0xFFFFFFFFu;
  __syncthreads();
  // Each round moves the whole workgroup to the next 512-word window: consecutive
  // lanes, consecutive words, every bank once per 32 lanes.
  uint32_t x = t * 2654435761u + blockIdx.x;
  for (uint32_t r = 0; r < rounds; r++) {
    const uint32_t w = (t + r * kPeakThreads) & (kLdsWords - 1u);
    x = x * 1664525u + 1013904223u;
    atomicMin(&lds[w], x);
  }
  __syncthreads();
  if (t == 0) sink[blockIdx.x] = lds[blockIdx.x & (kLdsWords - 1u)];
}

__global__ void __launch_bounds__(kPeakThreads)
    k_global_min_peak(uint32_t rounds, uint32_t* __restrict__ buf, uint32_t words_mask) {
  const uint32_t g = blockIdx.x * kPeakThreads + threadIdx.x;
  const uint32_t stride = gridDim.x * kPeakThreads;
  uint32_t x = g * 2654435761u;
  for (uint32_t r = 0; r < rounds; r++) {
    x = x * 1664525u + 1013904223u;
    atomicMin(&buf[(g + r * stride) & words_mask], x);
  }
}

}  // namespace

int atomic_peak(Ctx* c, int kind, uint32_t rounds, double* ops_per_s) {
  const unsigned blocks = (unsigned)c->cus * 8u;
  const size_t words = (size_t)1 << 26;  // 256 MiB
  uint32_t* buf = nullptr;
  int rc = dalloc(&buf, kind == CLAY_ATOMIC_GLOBAL_MIN ? words : (size_t)blocks);
  if (rc) return rc;
  hipStream_t s = c->stream;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  float ms = 0.0f;
  for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {  // the first run warms up
    if (kind == CLAY_ATOMIC_GLOBAL_MIN)
      e = hipMemsetAsync(buf, 0xFF, words * sizeof(uint32_t), s);
    if (e == hipSuccess) e = hipEventRecord(a, s);
    if (kind == CLAY_ATOMIC_GLOBAL_MIN)
      hipLaunchKernelGGL(k_global_min_peak, dim3(blocks), dim3(kPeakThreads), 0, s, rounds, buf,
                         (uint32_t)(words - 1));
    else
      hipLaunchKernelGGL(k_lds_min_peak, dim3(blocks), dim3(kPeakThreads), 0, s, rounds, buf);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  dfree(buf);
  if (e != hipSuccess) return fail(CLAY_ERR_HIP, std::string("clay_atomic_peak: ") + hipGetErrorString(e));
  *ops_per_s = (double)blocks * kPeakThreads * rounds / (ms * 1e-3);
  return CLAY_OK;
}

}  // namespace clay

extern "C" int clay_atomic_peak(clay_ctx* ctx, int32_t kind, uint32_t rounds, double* ops_per_s) {
  clay::Ctx* c = reinterpret_cast<clay::Ctx*>(ctx);
  if (!c || !ops_per_s) return clay::fail(CLAY_ERR_INVALID, "clay_atomic_peak: NULL argument");
  if (kind != CLAY_ATOMIC_LDS_MIN && kind != CLAY_ATOMIC_GLOBAL_MIN)
    return clay::fail(CLAY_ERR_INVALID, "clay_atomic_peak: unknown atomic kind");
  if (rounds < 1 || rounds > 65536) return clay::fail(CLAY_ERR_INVALID, "clay_atomic_peak: rounds must be in [1, 65536]");
  return clay::atomic_peak(c, kind, rounds, ops_per_s);
}
