// sort.hip -- stable LSD radix sort of (cell key, particle id) pairs for gfx950.
//
// Replaces the reference's insert pass (fluid.ts:308-331), whose bucket order is
// the atomicAdd order.  Here the order is defined: ascending key, ties by
// ascending id.  Digits of 8 to 11 bits (the fewest passes).  Per digit, a per-tile
// histogram (k_hist, in LDS), the offsets of each (tile, digit) from scans over the
// tiles, and k_scatter: digit base (scan of the digit totals, in LDS) + the tile's
// offset + stable rank inside the tile from wave64 ballots.  A tile is R rounds of
// 256 threads, so rank order inside a tile is (round, lane) = input order and the sort
// is stable.  R = 8 (2048 keys) up to 8M keys and 32 above.
// The histogram matrix has two layouts:
//  - digit-major, hist[digit][tile], for small matrices: k_hist writes and k_scatter
//    reads one word per row at scattered addresses, and k_scan_rows scans each row;
//  - tile-major, hist[tile][digit], from 2^19 entries, where those scattered words
//    cost more than a launch: k_hist writes and k_scatter reads contiguous rows, and
//    the scan is two-level, k_chunk_prefix (offsets inside chunks of kChunk tiles, in
//    place, and the chunk totals, whole rows) then k_scan_chunks (each digit's column
//    of chunk totals).
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kThreads = 256;
constexpr int kMinRounds = 8;
constexpr int kMaxRounds = 32;
constexpr uint32_t kBigSort = 1u << 23;  // keys from which tiles take kMaxRounds rounds
constexpr int kWaves = kThreads / 64;
constexpr int kMaxDigitBits = 11;

constexpr int kChunk = 8;  // tiles per chunk of the two-level scan

// The histogram of one tile: a contiguous row hist[tile][.] (tile-major) or one word
// per row hist[.][tile] (digit-major).
template <int W, int kRounds, bool kTileMajor>
__global__ void __launch_bounds__(kThreads)
    k_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift, uint32_t* __restrict__ hist, uint32_t ntiles) {
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
  for (uint32_t b = threadIdx.x; b < kBins; b += kThreads)
    hist[kTileMajor ? (size_t)blockIdx.x * kBins + b : (size_t)b * ntiles + blockIdx.x] = h[b];
}

// Digit-major: one block per digit, exclusive scan of hist[d][0..ntiles) in place;
// total -> totals[d].  Each thread sums a contiguous segment of the row, one
// block-wide scan of the 256 segment sums (wave shuffles, then the four wave totals
// through LDS) gives each segment its offset, and the thread rewrites its segment as
// running sums.
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

// Tile-major, per chunk of kChunk tiles and digit: the exclusive prefix over the chunk's tiles, in
// place, and the chunk's total -> chunk[c][d].  Threads take consecutive digits, so
// every access is a contiguous row.
template <int W>
__global__ void __launch_bounds__(kThreads)
    k_chunk_prefix(uint32_t* __restrict__ hist, uint32_t ntiles, uint32_t* __restrict__ chunk) {
  constexpr uint32_t kBins = 1u << W;
  const uint32_t t0 = blockIdx.x * kChunk;
  for (uint32_t b = threadIdx.x; b < kBins; b += kThreads) {
    uint32_t v[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; c++) v[c] = t0 + c < ntiles ? hist[(size_t)(t0 + c) * kBins + b] : 0u;
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < kChunk; c++) {
      if (t0 + c < ntiles) hist[(size_t)(t0 + c) * kBins + b] = acc;
      acc += v[c];
    }
    chunk[(size_t)blockIdx.x * kBins + b] = acc;
  }
}

// Tile-major, one block per digit: exclusive scan of its column chunk[0..nchunks)[d] in place;
// total -> totals[d].  Each thread sums a contiguous segment of the chunks, one
// block-wide scan of the 256 segment sums (wave shuffles, then the four wave totals
// through LDS) gives each segment its offset, and the thread rewrites its segment as
// running sums.
template <int W>
__global__ void __launch_bounds__(kThreads)
    k_scan_chunks(uint32_t* __restrict__ chunk, uint32_t nchunks, uint32_t* __restrict__ totals) {
  constexpr uint32_t kBins = 1u << W;
  __shared__ uint32_t wsum[kWaves];
  uint32_t* col = chunk + blockIdx.x;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t per = (nchunks + kThreads - 1) / kThreads;
  const uint32_t lo = t * per < nchunks ? t * per : nchunks;
  const uint32_t hi = lo + per < nchunks ? lo + per : nchunks;
  uint32_t sum = 0;
  for (uint32_t i = lo; i < hi; i++) sum += col[(size_t)i * kBins];
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
    const uint32_t v = col[(size_t)i * kBins];
    col[(size_t)i * kBins] = acc;
    acc += v;
  }
  if (t == kThreads - 1) totals[blockIdx.x] = acc;
}

template <int W, int kRounds, bool kTileMajor>
__global__ void __launch_bounds__(kThreads)
    k_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n,
              int shift, const uint32_t* __restrict__ hist, const uint32_t* __restrict__ chunk, uint32_t ntiles,
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
      digit_base[d] = acc + (kTileMajor ? chunk[(size_t)(blockIdx.x / kChunk) * kBins + d] +
                                             hist[(size_t)blockIdx.x * kBins + d]
                                       : hist[(size_t)d * ntiles + blockIdx.x]);
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

// Entries of the histogram matrix from which it is laid out tile-major.
constexpr size_t kTileMajorFrom = (size_t)1 << 19;

template <int W, int R>
void sort_pass(Ctx* c, hipStream_t s, uint32_t n, uint32_t ntiles, int shift, const uint32_t* keys,
               const uint32_t* vals, uint32_t* hist, uint32_t* keys_out, uint32_t* vals_out) {
  constexpr uint32_t kBins = 1u << W;
  const uint32_t nchunks = (ntiles + kChunk - 1) / kChunk;
  if ((size_t)ntiles * kBins >= kTileMajorFrom) {
    uint32_t* chunk = hist + (size_t)kBins * ntiles;
    uint32_t* totals = chunk + (size_t)kBins * nchunks;
    {
      LaunchTimer t(c, s, "sort_hist");
      hipLaunchKernelGGL((k_hist<W, R, true>), dim3(ntiles), dim3(kThreads), 0, s, keys, n, shift, hist, ntiles);
    }
    {
      LaunchTimer t(c, s, "sort_scan");
      hipLaunchKernelGGL(k_chunk_prefix<W>, dim3(nchunks), dim3(kThreads), 0, s, hist, ntiles, chunk);
      hipLaunchKernelGGL(k_scan_chunks<W>, dim3(kBins), dim3(kThreads), 0, s, chunk, nchunks, totals);
    }
    {
      LaunchTimer t(c, s, "sort_scatter");
      hipLaunchKernelGGL((k_scatter<W, R, true>), dim3(ntiles), dim3(kThreads), 0, s, keys, vals, n, shift, hist,
                         chunk, ntiles, totals, keys_out, vals_out);
    }
    return;
  }
  uint32_t* totals = hist + (size_t)kBins * ntiles;
  {
    LaunchTimer t(c, s, "sort_hist");
    hipLaunchKernelGGL((k_hist<W, R, false>), dim3(ntiles), dim3(kThreads), 0, s, keys, n, shift, hist, ntiles);
  }
  {
    LaunchTimer t(c, s, "sort_scan");
    hipLaunchKernelGGL(k_scan_rows, dim3(kBins), dim3(kThreads), 0, s, hist, ntiles, totals);
  }
  {
    LaunchTimer t(c, s, "sort_scatter");
    hipLaunchKernelGGL((k_scatter<W, R, false>), dim3(ntiles), dim3(kThreads), 0, s, keys, vals, n, shift, hist,
                       nullptr, ntiles, totals, keys_out, vals_out);
  }
}

}  // namespace

size_t radix_hist_size(uint32_t n) {
  // Sized for the smallest tiles, the most of them.
  const size_t tile = (size_t)kThreads * kMinRounds;
  const size_t ntiles = std::max<size_t>(1, (n + tile - 1) / tile);
  return ((size_t)1 << kMaxDigitBits) * (ntiles + (ntiles + kChunk - 1) / kChunk + 1);
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
  const bool odd = passes & 1;
  uint32_t* ik = odd ? keys_tmp : keys;
  uint32_t* iv = odd ? vals_tmp : vals;
  uint32_t* ok = odd ? keys : keys_tmp;
  uint32_t* ov = odd ? vals : vals_tmp;
  for (int p = 0; p < passes; p++) {
    const int shift = w * p;
    switch (w) {
      case 8:
        sort_pass<8, R>(c, s, n, ntiles, shift, ik, iv, hist, ok, ov);
        break;
      case 9:
        sort_pass<9, R>(c, s, n, ntiles, shift, ik, iv, hist, ok, ov);
        break;
      case 10:
        sort_pass<10, R>(c, s, n, ntiles, shift, ik, iv, hist, ok, ov);
        break;
      default:
        sort_pass<11, R>(c, s, n, ntiles, shift, ik, iv, hist, ok, ov);
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
