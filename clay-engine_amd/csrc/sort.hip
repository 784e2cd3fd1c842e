// sort.hip -- stable LSD radix sort of (cell key, particle id) pairs for gfx950.
//
// Replaces the reference's insert pass (fluid.ts:308-331), whose bucket order is
// the atomicAdd order.  Here the order is defined: ascending key, ties by
// ascending id.  Digits of 8 to 11 bits (the fewest passes), three launches per digit:
//   k_hist      per-tile digit histogram in LDS             -> hist[digit][tile]
//   k_scan_rows exclusive scan of each digit's row of tiles  -> row totals
//   k_scatter   digit base (scan of the row totals, in LDS) + tile offset +
//               stable rank inside the tile from wave64 ballots
// A tile is R rounds of 256 threads, so rank order inside a tile is (round, lane) =
// input order and the sort is stable.  R = 8 (2048 keys) up to 8M keys and 32 above:
// a pass reads and writes its per-tile histogram column at scattered addresses, and
// 33.5M keys in 2048-key tiles make that a 64 MB matrix at 10-bit digits.
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kThreads = 256;
constexpr int kMinRounds = 8;
constexpr int kMaxRounds = 32;
constexpr uint32_t kBigSort = 1u << 23;  // keys from which tiles take kMaxRounds rounds
constexpr int kWaves = kThreads / 64;
constexpr int kMaxDigitBits = 11;

template <int W, int kRounds>
__global__ void __launch_bounds__(kThreads)
    k_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift, uint32_t* __restrict__ hist,
           uint32_t ntiles) {
  constexpr uint32_t kBins = 1u << W;
  constexpr int kTileKeys = kThreads * kRounds;
  __shared__ uint32_t h[kBins];
  for (uint32_t b = threadIdx.x; b < kBins; b += kThreads) h[b] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kTileKeys;
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const uint32_t i = base + r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kBins - 1u)], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kBins; b += kThreads) hist[b * ntiles + blockIdx.x] = h[b];
}

// One block per digit: exclusive scan of hist[d][0..ntiles) in place; total -> totals[d].
// Each thread sums a contiguous segment of the row, one block-wide scan of the 256
// segment sums (wave shuffles, then the four wave totals through LDS) gives each
// segment its offset, and the thread rewrites its segment as running sums.
__global__ void __launch_bounds__(kThreads)
    k_scan_rows(uint32_t* __restrict__ hist, uint32_t ntiles, uint32_t* __restrict__ totals) {
  __shared__ uint32_t wsum[kWaves];
  uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t per = (ntiles + kThreads - 1) / kThreads;
  const uint32_t lo = t * per < ntiles ? t * per : ntiles;
  const uint32_t hi = lo + per < ntiles ? lo + per : ntiles;
  uint32_t sum = 0;
  for (uint32_t i = lo; i < hi; i++) sum += row[i];
  uint32_t incl = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off);
    if (lane >= (uint32_t)off) incl += v;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t acc = incl - sum;
  for (uint32_t w = 0; w < wave; w++) acc += wsum[w];
  for (uint32_t i = lo; i < hi; i++) {
    const uint32_t v = row[i];
    row[i] = acc;
    acc += v;
  }
  if (t == kThreads - 1) totals[blockIdx.x] = acc;
}

template <int W, int kRounds>
__global__ void __launch_bounds__(kThreads)
    k_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n,
              int shift, const uint32_t* __restrict__ hist, uint32_t ntiles,
              const uint32_t* __restrict__ totals, uint32_t* __restrict__ keys_out,
              uint32_t* __restrict__ vals_out) {
  constexpr uint32_t kBins = 1u << W;
  constexpr uint32_t kPer = kBins / kThreads;  // digits per thread
  constexpr int kTileKeys = kThreads * kRounds;
  __shared__ uint32_t digit_base[kBins];
  __shared__ uint32_t running[kBins];
  __shared__ uint32_t wave_cnt[kWaves][kBins];
  __shared__ uint32_t part[kThreads];
  const int t = threadIdx.x;
  const int wave = t >> 6;
  // Exclusive scan of the digit totals: each thread sums its kPer digits, a
  // Hillis-Steele scan over the 256 partial sums, then the digits in order.
  {
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < kPer; q++) sum += totals[t * kPer + q];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
      const uint32_t add = t >= off ? part[t - off] : 0u;
      __syncthreads();
      part[t] += add;
      __syncthreads();
    }
    uint32_t acc = part[t] - sum;
#pragma unroll
    for (uint32_t q = 0; q < kPer; q++) {
      const uint32_t d = t * kPer + q;
      digit_base[d] = acc + hist[(size_t)d * ntiles + blockIdx.x];
      acc += totals[d];
      running[d] = 0;
    }
  }
  const uint64_t lt = __lanemask_lt();
  const uint32_t base = blockIdx.x * kTileKeys;
  // The keys and values of kAhead rounds are loaded together, before their rounds'
  // barriers, so a group of rounds waits for memory once instead of once per round.
  constexpr int kAhead = 8;
  static_assert(kRounds % kAhead == 0, "rounds come in groups of kAhead");
  uint32_t kk[kAhead], vv[kAhead];
  for (int r = 0; r < kRounds; r++) {
    if (r % kAhead == 0) {
#pragma unroll
      for (int a = 0; a < kAhead; a++) {
        const uint32_t ia = base + (r + a) * kThreads + t;
        kk[a] = ia < n ? keys[ia] : 0u;
        vv[a] = ia < n ? vals[ia] : 0u;
      }
    }
    const uint32_t i = base + r * kThreads + t;
    const bool active = i < n;
    uint32_t key = 0, val = 0, d = 0;
    if (active) {
      key = kk[r % kAhead];
      val = vv[r % kAhead];
      d = (key >> shift) & (kBins - 1u);
    }
#pragma unroll
    for (int w = 0; w < kWaves; w++)
#pragma unroll
      for (uint32_t q = 0; q < kPer; q++) wave_cnt[w][t * kPer + q] = 0;
    __syncthreads();
    // Lanes of this wave holding the same digit.
    uint64_t same = __ballot(active);
#pragma unroll
    for (int b = 0; b < W; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t ones = __ballot(bit);
      same &= bit ? ones : ~ones;
    }
    const uint32_t rank = (uint32_t)__popcll(same & lt);
    if (active && rank == 0) wave_cnt[wave][d] = (uint32_t)__popcll(same);
    __syncthreads();
    // Per digit: offsets of each wave's group after the earlier waves and rounds.
#pragma unroll
    for (uint32_t q = 0; q < kPer; q++) {
      const uint32_t dd = t * kPer + q;
      uint32_t acc = running[dd];
#pragma unroll
      for (int w = 0; w < kWaves; w++) {
        const uint32_t c = wave_cnt[w][dd];
        wave_cnt[w][dd] = acc;
        acc += c;
      }
      running[dd] = acc;
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

template <int W, int R>
void sort_pass(Ctx* c, hipStream_t s, uint32_t n, uint32_t ntiles, int shift, const uint32_t* keys,
               const uint32_t* vals, uint32_t* hist, uint32_t* totals, uint32_t* keys_out, uint32_t* vals_out) {
  {
    LaunchTimer t(c, s, "sort_hist");
    hipLaunchKernelGGL((k_hist<W, R>), dim3(ntiles), dim3(kThreads), 0, s, keys, n, shift, hist, ntiles);
  }
  {
    LaunchTimer t(c, s, "sort_scan");
    hipLaunchKernelGGL(k_scan_rows, dim3(1u << W), dim3(kThreads), 0, s, hist, ntiles, totals);
  }
  {
    LaunchTimer t(c, s, "sort_scatter");
    hipLaunchKernelGGL((k_scatter<W, R>), dim3(ntiles), dim3(kThreads), 0, s, keys, vals, n, shift, hist, ntiles,
                       totals, keys_out, vals_out);
  }
}

}  // namespace

size_t radix_hist_size(uint32_t n) {
  // Sized for the smallest tiles, the most of them.
  const size_t tile = (size_t)kThreads * kMinRounds;
  const size_t ntiles = (n + tile - 1) / tile;
  return ((size_t)1 << kMaxDigitBits) * ((ntiles > 0 ? ntiles : 1) + 1);
}

namespace {
int digit_passes(int key_bits) { return (key_bits + kMaxDigitBits - 1) / kMaxDigitBits; }

// Ping-pongs between the two buffer pairs, starting from the one holding the input
// (the tmp pair for an odd number of passes), so that the last pass writes `keys`/`vals`.
template <int R>
void sort_passes(Ctx* c, hipStream_t s, uint32_t n, int passes, int w, uint32_t* keys, uint32_t* vals,
                 uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t* hist) {
  const uint32_t tile = (uint32_t)kThreads * R;
  const uint32_t ntiles = (n + tile - 1) / tile;
  uint32_t* totals = hist + ((size_t)1 << w) * ntiles;
  const bool odd = passes & 1;
  uint32_t* ik = odd ? keys_tmp : keys;
  uint32_t* iv = odd ? vals_tmp : vals;
  uint32_t* ok = odd ? keys : keys_tmp;
  uint32_t* ov = odd ? vals : vals_tmp;
  for (int p = 0; p < passes; p++) {
    const int shift = w * p;
    switch (w) {
      case 8:
        sort_pass<8, R>(c, s, n, ntiles, shift, ik, iv, hist, totals, ok, ov);
        break;
      case 9:
        sort_pass<9, R>(c, s, n, ntiles, shift, ik, iv, hist, totals, ok, ov);
        break;
      case 10:
        sort_pass<10, R>(c, s, n, ntiles, shift, ik, iv, hist, totals, ok, ov);
        break;
      default:
        sort_pass<11, R>(c, s, n, ntiles, shift, ik, iv, hist, totals, ok, ov);
        break;
    }
    std::swap(ik, ok);
    std::swap(iv, ov);
  }
}
}  // namespace

bool radix_input_in_tmp(int key_bits) { return digit_passes(key_bits) & 1; }

// Fewest passes of at most 11-bit digits, the digit width spread evenly: 24-bit keys
// take three 8-bit passes, 27-bit keys three 9-bit passes, 32-bit keys three 11-bit.
// The buffers never change roles, so a captured graph of a sort stays valid however
// many sorts run between its replays.
int radix_sort_pairs(Ctx* c, hipStream_t s, uint32_t n, int key_bits, uint32_t* keys, uint32_t* vals,
                     uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t* hist) {
  if (n == 0) return CLAY_OK;
  const int passes = digit_passes(key_bits);
  const int w = std::max(8, (key_bits + passes - 1) / passes);
  if (n >= kBigSort)
    sort_passes<kMaxRounds>(c, s, n, passes, w, keys, vals, keys_tmp, vals_tmp, hist);
  else
    sort_passes<kMinRounds>(c, s, n, passes, w, keys, vals, keys_tmp, vals_tmp, hist);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

}  // namespace clay
