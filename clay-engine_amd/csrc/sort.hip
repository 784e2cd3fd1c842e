// sort.hip -- stable LSD radix sort of (cell key, particle id) pairs for gfx950.
//
// Replaces the reference's insert pass (fluid.ts:308-331), whose bucket order is
// the atomicAdd order.  Here the order is defined: ascending key, ties by
// ascending id.  8-bit digits, three launches per digit:
//   k_hist      per-tile digit histogram in LDS             -> hist[digit][tile]
//   k_scan_rows exclusive scan of each digit's row of tiles  -> row totals
//   k_scatter   digit base (scan of 256 row totals, in LDS) + tile offset +
//               stable rank inside the tile from wave64 ballots
// A tile is 1024 keys: 4 rounds of 256 threads, so rank order inside a tile is
// (round, lane) = input order and the sort is stable.
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kThreads = 256;
constexpr int kRounds = 4;
constexpr int kTileKeys = kThreads * kRounds;
constexpr int kWaves = kThreads / 64;

__global__ void __launch_bounds__(kThreads)
    k_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift, uint32_t* __restrict__ hist,
           uint32_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kTileKeys;
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const uint32_t i = base + r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// One block per digit: exclusive scan of hist[d][0..ntiles) in place; total -> totals[d].
__global__ void __launch_bounds__(kThreads)
    k_scan_rows(uint32_t* __restrict__ hist, uint32_t ntiles, uint32_t* __restrict__ totals) {
  __shared__ uint32_t sh[kThreads];
  __shared__ uint32_t carry;
  uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < ntiles; c0 += kThreads) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < ntiles ? row[i] : 0u;
    sh[threadIdx.x] = v;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 entries.
    for (int off = 1; off < kThreads; off <<= 1) {
      const uint32_t add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0u;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    const uint32_t incl = sh[threadIdx.x];
    if (i < ntiles) row[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == kThreads - 1) carry += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(kThreads)
    k_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n,
              int shift, const uint32_t* __restrict__ hist, uint32_t ntiles,
              const uint32_t* __restrict__ totals, uint32_t* __restrict__ keys_out,
              uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t digit_base[256];
  __shared__ uint32_t running[256];
  __shared__ uint32_t wave_cnt[kWaves][256];
  const int t = threadIdx.x;
  const int wave = t >> 6;
  // Exclusive scan of the 256 digit totals (Hillis-Steele in LDS).
  {
    const uint32_t v = totals[t];
    digit_base[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const uint32_t add = t >= off ? digit_base[t - off] : 0u;
      __syncthreads();
      digit_base[t] += add;
      __syncthreads();
    }
    const uint32_t incl = digit_base[t];
    __syncthreads();
    digit_base[t] = incl - v + hist[(size_t)t * ntiles + blockIdx.x];
    running[t] = 0;
  }
  const uint64_t lt = __lanemask_lt();
  const uint32_t base = blockIdx.x * kTileKeys;
  for (int r = 0; r < kRounds; r++) {
    const uint32_t i = base + r * kThreads + t;
    const bool active = i < n;
    uint32_t key = 0, val = 0, d = 0;
    if (active) {
      key = keys[i];
      val = vals[i];
      d = (key >> shift) & 255u;
    }
#pragma unroll
    for (int w = 0; w < kWaves; w++) wave_cnt[w][t] = 0;
    __syncthreads();
    // Lanes of this wave holding the same digit.
    uint64_t same = __ballot(active);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t ones = __ballot(bit);
      same &= bit ? ones : ~ones;
    }
    const uint32_t rank = (uint32_t)__popcll(same & lt);
    if (active && rank == 0) wave_cnt[wave][d] = (uint32_t)__popcll(same);
    __syncthreads();
    // Per digit: offsets of each wave's group after the earlier waves and rounds.
    {
      uint32_t acc = running[t];
#pragma unroll
      for (int w = 0; w < kWaves; w++) {
        const uint32_t c = wave_cnt[w][t];
        wave_cnt[w][t] = acc;
        acc += c;
      }
      running[t] = acc;
    }
    __syncthreads();
    if (active) {
      const uint32_t dst = digit_base[d] + wave_cnt[wave][d] + rank;
      keys_out[dst] = key;
      vals_out[dst] = val;
    }
    __syncthreads();
  }
}

}  // namespace

size_t radix_hist_size(uint32_t n) {
  const size_t ntiles = (n + kTileKeys - 1) / kTileKeys;
  return 256 * (ntiles > 0 ? ntiles : 1) + 256;
}

int radix_sort_pairs(Ctx* c, hipStream_t s, uint32_t n, int key_bits, uint32_t*& keys,
                     uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt, uint32_t* hist) {
  if (n == 0) return CLAY_OK;
  const uint32_t ntiles = (n + kTileKeys - 1) / kTileKeys;
  uint32_t* totals = hist + 256 * (size_t)ntiles;
  const int passes = (key_bits + 7) / 8;
  for (int p = 0; p < passes; p++) {
    const int shift = 8 * p;
    {
      LaunchTimer t(c, s, "sort_hist");
      hipLaunchKernelGGL(k_hist, dim3(ntiles), dim3(kThreads), 0, s, keys, n, shift, hist, ntiles);
    }
    {
      LaunchTimer t(c, s, "sort_scan");
      hipLaunchKernelGGL(k_scan_rows, dim3(256), dim3(kThreads), 0, s, hist, ntiles, totals);
    }
    {
      LaunchTimer t(c, s, "sort_scatter");
      hipLaunchKernelGGL(k_scatter, dim3(ntiles), dim3(kThreads), 0, s, keys, vals, n, shift, hist,
                         ntiles, totals, keys_alt, vals_alt);
    }
    CLAY_HIP_CHECK(hipGetLastError());
    uint32_t* tk = keys;
    keys = keys_alt;
    keys_alt = tk;
    uint32_t* tv = vals;
    vals = vals_alt;
    vals_alt = tv;
  }
  return CLAY_OK;
}

}  // namespace clay
