// Scan, wavefront-level LSD radix sort, key duplication and tile ranges.
//
// The reference sorts K (tile << 32 | depth) 64-bit keys with one LSD radix
// sort over bits [0, 32 + msb(#tiles)) [U].  An LSD sort processes the low
// 32 bits (depth) first; because every duplicate of a Gaussian carries the
// same depth, those passes can run over the P Gaussians *before*
// duplication.  This library therefore
//   1. sorts the P Gaussians by depth bits (4 passes over P, stable),
//   2. duplicates them in depth order into (tile, id) pairs,
//   3. stably sorts the K pairs by tile id (2 passes of <= 8 bits),
// which yields exactly the reference's order — (tile, depth bits, Gaussian
// index) lexicographic — while moving the 32 depth bits over P instead of K.
//
// Ranking inside a block is wavefront-level: each wave64 owns a contiguous
// 1024-key slice, finds lanes with equal digits with per-bit ballots and
// keeps running per-digit counters in LDS, so ties keep input order.
#include <utility>

#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;

// ---------------------------------------------------------------------------
// block-wide helpers
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// scan
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in, size_t n,
                                                                  uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t s_wave[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  // every load is issued before the first is used: offsets past n read the
  // last value and count nothing (a guarded load would wait for each one)
  const uint32_t* __restrict__ inb = in + base;
  const uint32_t last = (uint32_t)min(n - base, (size_t)kScanTile) - 1u;
  uint32_t v[kScanItems];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) v[i] = inb[min((uint32_t)(i * kScanThreads) + threadIdx.x, last)];
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) sum += (uint32_t)(i * kScanThreads) + threadIdx.x <= last ? v[i] : 0u;
  uint32_t total;
  block_exclusive_scan<kScanThreads / 64>(sum, s_wave, total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Exclusive scan of the block sums in place, one block, any length.
__global__ __launch_bounds__(kScanThreads) void scan_spine_kernel(uint32_t* __restrict__ sums, size_t nb) {
  __shared__ uint32_t s_wave[kScanThreads / 64];
  uint32_t carry = 0;
  for (size_t base = 0; base < nb; base += kScanTile) {
    uint32_t v[kScanItems];
    uint32_t tsum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const size_t idx = base + (size_t)threadIdx.x * kScanItems + i;
      v[i] = idx < nb ? sums[idx] : 0u;
      tsum += v[i];
    }
    uint32_t total;
    uint32_t run = carry + block_exclusive_scan<kScanThreads / 64>(tsum, s_wave, total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const size_t idx = base + (size_t)threadIdx.x * kScanItems + i;
      if (idx < nb) sums[idx] = run;
      run += v[i];
    }
    carry += total;
    __syncthreads();
  }
}

// LDS tile index with one padding word per 32 (thread-contiguous reads of 16
// consecutive words then hit different banks).
__device__ __forceinline__ int scan_pad(int p) { return p + (p >> 5); }

// The tile moves between HBM and registers through LDS: global loads and
// stores stay fully coalesced (consecutive lanes, consecutive words) while
// each thread scans 16 consecutive elements.
// chunk_first (inclusive scans of instance counts): also record, for every
// 4096-slot output chunk below K, the element whose [start, end) holds the
// chunk's first slot (what chunk_first_kernel computes from the offsets).
constexpr uint32_t kChunkSlots = 4096;
// sums_raw: `spine` holds the raw block sums (at most kPrefixBlocks of them)
// and each work-group adds up those before it, so no spine launch is needed;
// otherwise `spine` holds their exclusive scan.
constexpr int kPrefixItems = 8;
constexpr size_t kPrefixBlocks = (size_t)kPrefixItems * kScanThreads;  // 2048 blocks: 8M values
template <bool kExclusive>
__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t* __restrict__ in,
                                                                 uint32_t* __restrict__ out, size_t n,
                                                                 const uint32_t* __restrict__ spine, int sums_raw,
                                                                 uint32_t* __restrict__ chunk_first, uint32_t K) {
  __shared__ uint32_t s_wave[kScanThreads / 64];
  __shared__ uint32_t tile[kScanTile + kScanTile / 32];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  // all loads in flight at once (offsets past n read the last value, stored as
  // zero; block sums past this block's predecessors count nothing)
  const uint32_t* __restrict__ inb = in + base;
  const uint32_t last = (uint32_t)min(n - base, (size_t)kScanTile) - 1u;
  uint32_t pre = 0;
  {
    uint32_t ld[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) ld[i] = inb[min((uint32_t)(i * kScanThreads) + threadIdx.x, last)];
    if (sums_raw) {
      uint32_t ps[kPrefixItems];
#pragma unroll
      for (int i = 0; i < kPrefixItems; ++i) ps[i] = spine[min((uint32_t)(i * kScanThreads) + threadIdx.x, gridDim.x - 1)];
#pragma unroll
      for (int i = 0; i < kPrefixItems; ++i) pre += (uint32_t)(i * kScanThreads) + threadIdx.x < blockIdx.x ? ps[i] : 0u;
    }
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const uint32_t o = (uint32_t)(i * kScanThreads) + threadIdx.x;
      tile[scan_pad((int)o)] = o <= last ? ld[i] : 0u;
    }
  }
  uint32_t prefix;
  if (sums_raw)
    block_exclusive_scan<kScanThreads / 64>(pre, s_wave, prefix);  // its barriers also order the tile stores
  else
    prefix = spine[blockIdx.x];
  __syncthreads();
  uint32_t v[kScanItems];
  uint32_t tsum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = tile[scan_pad((int)threadIdx.x * kScanItems + i)];
    tsum += v[i];
  }
  uint32_t total;
  uint32_t run = prefix + block_exclusive_scan<kScanThreads / 64>(tsum, s_wave, total);  // ends with a barrier
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const uint32_t x = v[i];
    if (!kExclusive && chunk_first && x > 0u) {
      const uint32_t beg = min(run, K), end = min(run + x, K);
      const size_t idx = base + (size_t)threadIdx.x * kScanItems + i;
      for (uint32_t c = (beg + kChunkSlots - 1) / kChunkSlots; c * kChunkSlots < end; ++c)
        chunk_first[c] = (uint32_t)idx;
    }
    if (!kExclusive) run += x;
    tile[scan_pad((int)threadIdx.x * kScanItems + i)] = run;
    if (kExclusive) run += x;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int p = i * kScanThreads + (int)threadIdx.x;
    const size_t idx = base + p;
    if (idx < n) out[idx] = tile[scan_pad(p)];
  }
}

size_t scan_scratch_words(size_t n) { return (n + kScanTile - 1) / kScanTile + 64; }

static void scan_impl(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, bool exclusive,
                      hipStream_t s, uint32_t* chunk_first = nullptr, uint32_t K = 0) {
  if (n == 0) return;
  const size_t nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, scratch);
  const int sums_raw = nb <= kPrefixBlocks ? 1 : 0;  // few blocks: each adds up its own prefix
  if (!sums_raw) hipLaunchKernelGGL(scan_spine_kernel, dim3(1), dim3(kScanThreads), 0, s, scratch, nb);
  if (exclusive)
    hipLaunchKernelGGL(scan_apply_kernel<true>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n,
                       (const uint32_t*)scratch, sums_raw, (uint32_t*)nullptr, 0u);
  else
    hipLaunchKernelGGL(scan_apply_kernel<false>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n,
                       (const uint32_t*)scratch, sums_raw, chunk_first, K);
}

void launch_inclusive_scan_chunks(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch,
                                  uint32_t* chunk_first, uint32_t K, hipStream_t s) {
  scan_impl(in, out, n, scratch, false, s, chunk_first, K);
}

void launch_inclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, hipStream_t s) {
  scan_impl(in, out, n, scratch, false, s);
}

// ---------------------------------------------------------------------------
// radix sort
//
// One pass = histogram (per-block digit counts, digit-major) -> per-digit
// exclusive scan over blocks (one workgroup per digit) -> rank + scatter.
// The scatter derives each digit's global base from the digit totals itself,
// so a pass is three launches.  Blocks hold 256 x ITEMS keys; small inputs use
// fewer items per thread so that every CU gets work.
// ---------------------------------------------------------------------------
// Lanes of the wave whose digit equals this lane's: one ballot per digit bit
// (at most 8; the loop is unrolled with a uniform test of the width).
// Per bit: s = the lane's bit sign-extended (all ones or zero); a lane j is
// still a peer when its bit equals ours, i.e. where ballot ^ s is clear.
__device__ __forceinline__ uint64_t digit_peers_bit(uint64_t peers, uint32_t d, int b) {
  const uint32_t s = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
  const uint64_t bal = __ballot(s != 0u);
  const uint32_t lo = (uint32_t)peers & ~((uint32_t)bal ^ s);
  const uint32_t hi = (uint32_t)(peers >> 32) & ~((uint32_t)(bal >> 32) ^ s);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, int nbits, bool valid) {
  uint64_t peers = __ballot(valid);
  if (nbits == 8) {  // every depth pass and the one-pass cell sort: no per-bit branch
#pragma unroll
    for (int b = 0; b < 8; ++b) peers = digit_peers_bit(peers, d, b);
    return peers;
  }
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    if (b >= nbits) break;
    peers = digit_peers_bit(peers, d, b);
  }
  return peers;
}

// SPREAD: the digits are close to uniform (the low bytes of depth keys), so
// per-wave counters with one plain LDS atomic per key rarely collide;
// otherwise lanes with equal digits are grouped by ballots first.
// FIRST (the depth sort's first pass over the preprocess's per-entry keys):
// keys equal to kSortSkip (culled entries) are not counted, and a key in the
// view-major far-depth bucket (bits 24-27 all set) sets bit 0 of *flags.
template <int ITEMS, bool SPREAD, bool FIRST = false>
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t* __restrict__ keys, size_t n,
                                                                 int shift, int nbits,
                                                                 uint32_t* __restrict__ hist, uint32_t nb,
                                                                 const uint32_t* __restrict__ n_dev,
                                                                 uint32_t* __restrict__ flags = nullptr) {
  __shared__ uint32_t cnt[256];
  __shared__ uint32_t wcnt[SPREAD ? kSortWaves : 1][256];
  if (n_dev) n = min(n, (size_t)*n_dev);  // device-side length (blocks past it count nothing)
  const uint32_t mask = (1u << nbits) - 1u;
  cnt[threadIdx.x] = 0;
  if constexpr (SPREAD) {
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) wcnt[w][threadIdx.x] = 0;
  }
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * (kSortThreads * ITEMS);
  const int lane = threadIdx.x & 63;
  if (base < n) {  // uniform per block; then n >= 1
    // every key of the thread is requested before the first is used
    // (indices past n read the last key and count nothing)
    uint32_t k[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const size_t idx = base + (size_t)i * kSortThreads + threadIdx.x;
      k[i] = keys[idx < n ? idx : n - 1];
    }
    // Lanes holding the same digit are found with one ballot per digit bit
    // and counted by one LDS atomic from the group's highest lane: skewed
    // digits (e.g. the exponent byte of depth keys) do not serialise 64
    // atomics.
    bool far = false;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const size_t idx = base + (size_t)i * kSortThreads + threadIdx.x;
      bool valid = idx < n;
      if constexpr (FIRST) {
        valid = valid && k[i] != kSortSkip;
        far = far || (valid && (k[i] & kFarBucketBits) == kFarBucketBits);
      }
      const uint32_t d = (k[i] >> shift) & mask;
      if constexpr (SPREAD) {
        if (valid) atomicAdd(&wcnt[threadIdx.x >> 6][d], 1u);
      } else {
        const uint64_t peers = digit_peers(d, nbits, valid);
        if (valid && lane == 63 - __clzll(peers)) atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
      }
    }
    if constexpr (FIRST) {
      if (flags && __ballot(far) != 0 && lane == 0) atomicOr(flags, 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x <= mask) {
    uint32_t c = cnt[threadIdx.x];
    if constexpr (SPREAD) {
#pragma unroll
      for (int w = 0; w < kSortWaves; ++w) c += wcnt[w][threadIdx.x];
    }
    hist[(size_t)threadIdx.x * nb + blockIdx.x] = c;
  }
}

// Exclusive scan of row `blockIdx.x` (one digit, nb block counts) in place;
// the row total goes to totals[digit].
// NT threads hold NT x 16 block counts per round (512 when a row has more
// than 4096 blocks, so the cell sort's rows are scanned in one round).  The
// row moves through LDS like scan_apply_kernel's tile: consecutive lanes read
// and write consecutive counts, and each thread scans 16 consecutive ones.
template <int NT>
__global__ __launch_bounds__(NT) void radix_rowscan_kernel(uint32_t* __restrict__ hist, uint32_t nb,
                                                          uint32_t* __restrict__ totals) {
  constexpr int kTile = NT * kScanItems;
  __shared__ uint32_t s_wave[NT / 64];
  __shared__ uint32_t tile[kTile + kTile / 32];
  uint32_t* row = hist + (size_t)blockIdx.x * nb;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += (uint32_t)kTile) {
    const uint32_t last = min(nb - base, (uint32_t)kTile) - 1u;
    {
      uint32_t ld[kScanItems];
#pragma unroll
      for (int i = 0; i < kScanItems; ++i) ld[i] = row[base + min((uint32_t)(i * NT) + threadIdx.x, last)];
#pragma unroll
      for (int i = 0; i < kScanItems; ++i) {
        const uint32_t o = (uint32_t)(i * NT) + threadIdx.x;
        tile[scan_pad((int)o)] = o <= last ? ld[i] : 0u;
      }
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t tsum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      v[i] = tile[scan_pad((int)threadIdx.x * kScanItems + i)];
      tsum += v[i];
    }
    uint32_t total;
    uint32_t run = carry + block_exclusive_scan<NT / 64>(tsum, s_wave, total);  // ends with a barrier
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      tile[scan_pad((int)threadIdx.x * kScanItems + i)] = run;
      run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const uint32_t o = (uint32_t)(i * NT) + threadIdx.x;
      if (o <= last) row[base + o] = tile[scan_pad((int)o)];
    }
    carry += total;
    __syncthreads();  // the tile is read out before the next round fills it
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// FIRST: keys equal to kSortSkip are dropped, each value is its key's input
// position (vin unused), and work-group 0 stores the number of keys kept in
// *n_out.
template <int ITEMS, bool FIRST = false>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, size_t n, int shift, int nbits, const uint32_t* __restrict__ hist_excl,
    const uint32_t* __restrict__ totals, uint32_t nb, const uint32_t* __restrict__ n_dev,
    uint32_t* __restrict__ n_out = nullptr) {
  constexpr int kTile = kSortThreads * ITEMS;
  constexpr int kSlice = 64 * ITEMS;
  if (n_dev) n = min(n, (size_t)*n_dev);
  if ((size_t)blockIdx.x * kTile >= n) return;  // uniform per block
  __shared__ uint32_t wcnt[kSortWaves][256];
  __shared__ uint32_t dstart[256];  // block-local start of each digit
  __shared__ uint32_t gbase[256];   // global position of the block's first element of each digit
  // the block's keys, then its values, in block-sorted order: one buffer for
  // both keeps the LDS per work-group small enough for seven per CU
  __shared__ uint32_t sbuf[kTile];
  __shared__ uint8_t sdig[kTile];  // digit of each block-sorted slot
  __shared__ uint32_t s_wave[kSortWaves];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar addressing
  const uint32_t mask = (1u << nbits) - 1u;
  for (int w = 0; w < kSortWaves; ++w) wcnt[w][threadIdx.x] = 0;
  // global position of the block's first element of each digit (one thread
  // per digit), before any key is held in registers
  const uint32_t dg = threadIdx.x;
  {
    uint32_t all_total;
    const uint32_t digit_base = block_exclusive_scan<kSortWaves>(dg <= mask ? totals[dg] : 0u, s_wave, all_total);
    if (dg <= mask) gbase[dg] = digit_base + hist_excl[(size_t)dg * nb + blockIdx.x];
    if constexpr (FIRST) {
      if (n_out && blockIdx.x == 0 && threadIdx.x == 0) *n_out = all_total;
    }
  }

  const size_t wbase = (size_t)blockIdx.x * kTile + (size_t)wave * kSlice;
  const uint64_t lt = lanemask_lt();
  // the wave's slice through wave-uniform base pointers and 32-bit lane
  // offsets: one address register for all loads instead of a 64-bit address each
  const uint32_t* __restrict__ kw = kin + wbase;
  const uint32_t* __restrict__ vw = vin + wbase;
  const uint32_t rem = n > wbase ? (uint32_t)min(n - wbase, (size_t)kSlice) : 0u;
  // keys now; the values only once the keys are written (fewer live registers)
  uint32_t key[ITEMS], rank[ITEMS];
  if (rem == (uint32_t)kSlice) {  // a full slice (all but the last): no per-item guards
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) key[i] = kw[i * 64 + lane];
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t o = (uint32_t)(i * 64) + (uint32_t)lane;
      key[i] = o < rem ? kw[o] : 0u;
    }
  }
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    bool valid = (uint32_t)(i * 64) + (uint32_t)lane < rem;
    if constexpr (FIRST) valid = valid && key[i] != kSortSkip;
    const uint32_t d = (key[i] >> shift) & mask;
    const uint64_t peers = digit_peers(d, nbits, valid);
    const uint32_t before = wcnt[wave][d];
    rank[i] = before + (uint32_t)__popcll(peers & lt);
    const int leader = 63 - __clzll(peers);
    if (valid && lane == leader) wcnt[wave][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // per digit (one thread each): exclusive prefix over waves and block-local
  // digit start
  uint32_t tot = 0;
  if (dg <= mask) {
    for (int w = 0; w < kSortWaves; ++w) {
      const uint32_t c = wcnt[w][dg];
      wcnt[w][dg] = tot;
      tot += c;
    }
  }
  uint32_t block_total;
  dstart[dg] = block_exclusive_scan<kSortWaves>(tot, s_wave, block_total);
  __syncthreads();
  // block-local sorted position of every item (kTile: none)
  uint32_t lpos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    lpos[i] = (uint32_t)kTile;
    bool valid = (uint32_t)(i * 64) + (uint32_t)lane < rem;
    if constexpr (FIRST) valid = valid && key[i] != kSortSkip;
    if (valid) {
      const uint32_t d = (key[i] >> shift) & mask;
      lpos[i] = dstart[d] + wcnt[wave][d] + rank[i];
    }
  }
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (lpos[i] < (uint32_t)kTile) {
      sbuf[lpos[i]] = key[i];
      sdig[lpos[i]] = (uint8_t)((key[i] >> shift) & mask);
    }
  }
  __syncthreads();
  // coalesced write-out, keys then values: each digit's run is contiguous in
  // LDS and in memory
  const uint32_t count = block_total;  // the keys this block sorts (all of its slice unless skipped ones)
  for (uint32_t i = threadIdx.x; i < count; i += kSortThreads) {
    const uint32_t d = sdig[i];
    const uint32_t pos = gbase[d] + (i - dstart[d]);
    if (pos < n) kout[pos] = sbuf[i];  // always true for a consistent histogram; guards the write
  }
  uint32_t val[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if constexpr (FIRST) {
      val[i] = (uint32_t)wbase + (uint32_t)(i * 64 + lane);
    } else {
      val[i] = lpos[i] < (uint32_t)kTile ? vw[i * 64 + lane] : 0u;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if (lpos[i] < (uint32_t)kTile) sbuf[lpos[i]] = val[i];
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < count; i += kSortThreads) {
    const uint32_t d = sdig[i];
    const uint32_t pos = gbase[d] + (i - dstart[d]);
    if (pos < n) vout[pos] = sbuf[i];
  }
}

static void launch_rowscan(uint32_t* hist, uint32_t nb, uint32_t* totals, uint32_t digits, hipStream_t s) {
  if (nb > (uint32_t)kScanTile)
    hipLaunchKernelGGL(radix_rowscan_kernel<512>, dim3(digits), dim3(512), 0, s, hist, nb, totals);
  else
    hipLaunchKernelGGL(radix_rowscan_kernel<kScanThreads>, dim3(digits), dim3(kScanThreads), 0, s, hist, nb, totals);
}

static int sort_items(size_t n) { return n >= ((size_t)4 << 20) ? 16 : 4; }

size_t radix_scratch_words(size_t n) {
  const size_t tile = (size_t)kSortThreads * sort_items(n);
  const size_t nb = (n + tile - 1) / tile;
  return 256 * nb + 256 + 64;
}

// Entry-key first pass (see radix_sort_entry_keys): when `entry` is set, pass 0
// reads k0 as per-entry keys with kSortSkip holes and implicit values and
// counts the kept keys into entry->n_out, which bounds the later passes.
struct EntryPass {
  uint32_t* n_out;
  uint32_t* flags;
};

template <int ITEMS>
static void radix_passes(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                         uint32_t* scratch, hipStream_t s, const uint32_t* n_dev, int passes,
                         bool first_hist_ready = false, int spread_below = 0, const EntryPass* entry = nullptr) {
  constexpr size_t kTile = (size_t)kSortThreads * ITEMS;
  const uint32_t nb = (uint32_t)((n + kTile - 1) / kTile);
  int widths[4];
  int left = bits;
  for (int p = 0; p < passes; ++p) {
    const int w = (left + (passes - p) - 1) / (passes - p);
    widths[p] = w;
    left -= w;
  }
  uint32_t* hist = scratch;
  uint32_t* totals = scratch + 256 * (size_t)nb;
  uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
  int shift = 0;
  for (int p = 0; p < passes; ++p) {
    const int w = widths[p];
    if (p == 0 && entry) {
      hipLaunchKernelGGL((radix_hist_kernel<ITEMS, true, true>), dim3(nb), dim3(kSortThreads), 0, s,
                         (const uint32_t*)ka, n, shift, w, hist, nb, (const uint32_t*)nullptr, entry->flags);
      launch_rowscan(hist, nb, totals, 1u << w, s);
      hipLaunchKernelGGL((radix_scatter_kernel<ITEMS, true>), dim3(nb), dim3(kSortThreads), 0, s,
                         (const uint32_t*)ka, (const uint32_t*)nullptr, kb, vb, n, shift, w, (const uint32_t*)hist,
                         (const uint32_t*)totals, nb, (const uint32_t*)nullptr, entry->n_out);
      n_dev = entry->n_out;  // the later passes sort the kept keys only
      shift += w;
      std::swap(ka, kb);
      std::swap(va, vb);
      continue;
    }
    if (p == 0 && first_hist_ready) {
    } else if (shift + w <= spread_below) {
      hipLaunchKernelGGL((radix_hist_kernel<ITEMS, true>), dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ka, n,
                         shift, w, hist, nb, n_dev);
    } else {
      hipLaunchKernelGGL((radix_hist_kernel<ITEMS, false>), dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ka,
                         n, shift, w, hist, nb, n_dev);
    }
    launch_rowscan(hist, nb, totals, 1u << w, s);
    hipLaunchKernelGGL(radix_scatter_kernel<ITEMS>, dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ka,
                       (const uint32_t*)va, kb, vb, n, shift, w, (const uint32_t*)hist, (const uint32_t*)totals, nb,
                       n_dev);
    shift += w;
    uint32_t* t = ka;
    ka = kb;
    kb = t;
    t = va;
    va = vb;
    vb = t;
  }
}

static void radix_dispatch(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                           uint32_t* scratch, hipStream_t s, const uint32_t* n_dev, int passes,
                           bool first_hist_ready = false, int spread_below = 0, const EntryPass* entry = nullptr) {
  if (sort_items(n) == 16)
    radix_passes<16>(k0, v0, k1, v1, n, bits, scratch, s, n_dev, passes, first_hist_ready, spread_below, entry);
  else
    radix_passes<4>(k0, v0, k1, v1, n, bits, scratch, s, n_dev, passes, first_hist_ready, spread_below, entry);
}

void radix_sort_entry_keys(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, uint32_t* scratch,
                           hipStream_t s, uint32_t* n_out, uint32_t* flags) {
  if (n == 0) return;
  // every pass counts with per-wave LDS atomics: even the top byte (view and
  // exponent bits) spreads over enough digits that they rarely collide
  // (measured 5 us per sort faster than ballot grouping on 5.9M keys)
  const EntryPass entry{n_out, flags};
  radix_dispatch(k0, v0, k1, v1, n, 32, scratch, s, nullptr, 4, false, 32, &entry);
}

static int min_passes(int bits) { return bits <= 8 ? 1 : (bits <= 16 ? 2 : 4); }

int radix_min_passes_first_width(size_t n, int bits) {
  if (sort_items(n) != 16) return -1;  // blocks are not the duplicate's 4096-slot chunks
  bits = max(0, min(32, bits));
  const int passes = min_passes(bits);
  return (bits + passes - 1) / passes;
}

void radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                      uint32_t* scratch, hipStream_t s, const uint32_t* n_dev) {
  if (n == 0) return;
  if (bits < 0) bits = 0;
  if (bits > 32) bits = 32;
  radix_dispatch(k0, v0, k1, v1, n, bits, scratch, s, n_dev, bits <= 16 ? 2 : 4);
}

bool radix_sort_pairs_min_passes(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                                 uint32_t* scratch, hipStream_t s, const uint32_t* n_dev, bool first_hist_ready) {
  if (n == 0) return false;
  if (bits < 0) bits = 0;
  if (bits > 32) bits = 32;
  const int passes = min_passes(bits);
  radix_dispatch(k0, v0, k1, v1, n, bits, scratch, s, n_dev, passes, first_hist_ready);
  return (passes & 1) != 0;
}

// ---------------------------------------------------------------------------
// gather / duplicate / ranges
// ---------------------------------------------------------------------------
__global__ void gather_u32_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx,
                                  uint32_t* __restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

void launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t* dst, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gather_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, src, idx, dst, n);
}

// Load-balanced emission.  The K output slots are cut into chunks of
// kDupChunk; a first pass records, for every chunk, the depth-ordered
// Gaussian whose output range contains the chunk's first slot.  Each
// workgroup then fills exactly one chunk: it stages windows of up to 256
// consecutive Gaussians (end offset, rectangle) in LDS and maps each of its
// slots back to its Gaussian by binary search, so stores are coalesced and a
// few huge splats cannot serialise a workgroup.
constexpr uint32_t kDupChunk = 4096;
static_assert(kDupChunk == kChunkSlots, "the scan fills the duplicate's chunk table");

__device__ __forceinline__ uint32_t device_count(uint32_t cap, const uint32_t* dev) {
  return dev ? min(cap, *dev) : cap;
}

__global__ __launch_bounds__(256) void chunk_first_kernel(DupArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.P) return;
  const uint32_t K = device_count(a.K, a.K_dev);
  const uint32_t beg = (s == 0) ? 0u : min(a.offsets[s - 1], K);
  const uint32_t end = min(a.offsets[s], K);
  for (uint32_t c = (beg + kDupChunk - 1) / kDupChunk; c * kDupChunk < end; ++c) a.chunk_first[c] = (uint32_t)s;
}

constexpr int kDupViews = 256;  // views whose cell-grid width and base are cached in LDS

__global__ __launch_bounds__(256) void duplicate_kernel(DupArgs a) {
  __shared__ uint32_t s_end[256], s_g[256], s_tb[256];
  __shared__ int s_x0[256], s_y0[256], s_w[256], s_gx[256];
  __shared__ uint32_t s_rx[256], s_ry[256];  // multi-view: tile rectangle, [lo, hi) packed 16:16
  __shared__ int s_vcx[kDupViews];           // multi-view: cells per row of view v
  __shared__ uint32_t s_vtb[kDupViews];      // multi-view: key base of view v
  __shared__ uint32_t s_hist[256];           // digit counts of this chunk (first sort pass)
  const uint32_t K = device_count(a.K, a.K_dev);
  const uint32_t lo_slot = blockIdx.x * kDupChunk;
  const uint32_t hmask = (1u << a.hist_bits) - 1u;
  if (a.zero) {  // this work-group's share of the words a later step needs zeroed
    const size_t per = (a.zero_n + gridDim.x - 1) / gridDim.x;
    const size_t zlo = (size_t)blockIdx.x * per, zhi = min(a.zero_n, zlo + per);
    for (size_t i = zlo + threadIdx.x; i < zhi; i += 256) a.zero[i] = 0u;
  }
  if (lo_slot >= K) {  // uniform per block: an empty chunk counts nothing
    if (a.hist && threadIdx.x <= hmask) a.hist[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = 0u;
    return;
  }
  if (a.hist) s_hist[threadIdx.x] = 0u;
  const uint32_t hi_slot = min(K, lo_slot + kDupChunk);
  const uint32_t P = (uint32_t)a.P;
  const uint32_t nviews = a.views ? (P + a.per_view - 1) / a.per_view : 0u;
  auto view_tb = [&](uint32_t v) -> uint32_t {
    return a.view_major ? v << 12 : (uint32_t)a.views[v].cell_base;
  };
  for (uint32_t v = threadIdx.x; v < min(nviews, (uint32_t)kDupViews); v += 256) {
    s_vcx[v] = cells_x(a.views[v].grid_x);
    s_vtb[v] = view_tb(v);
  }
  __syncthreads();
  uint32_t first = a.chunk_first[blockIdx.x];
  uint32_t win_begin = (first == 0) ? 0u : min(a.offsets[first - 1], K);
  // A window's per-entry loads (entry, its slot range, its rectangle), issued
  // together; the next window's are in flight while this one is emitted.
  struct Fetched {
    uint32_t g, prev_end, end;
    uint2 pr;
  };
  auto fetch = [&](uint32_t s) {
    Fetched f{0u, 0u, 0u, make_uint2(0u, 0u)};
    if (s < P) {
      f.g = a.order[s];
      f.prev_end = a.offsets[s > 0 ? s - 1 : 0];  // unconditional: no branch between the loads
      f.end = a.offsets[s];
      if (a.views) f.pr = a.rects[s];
    }
    return f;
  };
  Fetched cur = fetch(first + threadIdx.x);
  while (true) {
    // stage Gaussians first .. first+255 (those past P get an empty range)
    const uint32_t s = first + threadIdx.x;
    uint32_t end = 0xFFFFFFFFu, g = 0, tb = 0, rx = 0, ry = 0;
    int x0 = 0, y0 = 0, w = 1, gx = a.grid_x;
    if (s < P) {
      g = cur.g;
      end = min(cur.end, K);
      const uint2 pr = cur.pr;
      const uint32_t beg = (s == 0) ? 0u : min(cur.prev_end, K);
      if (end > beg && g < P) {
        int2 rmin, rmax;
        if (a.views) {  // depth-ordered rectangle from the gather step; per binning cell of 4x4 tiles
          const uint32_t v = g / a.per_view;
          unpack_rect(pr, rmin, rmax);
          rx = (uint32_t)rmin.x | ((uint32_t)rmax.x << 16);
          ry = (uint32_t)rmin.y | ((uint32_t)rmax.y << 16);
          if (v < (uint32_t)kDupViews) {
            gx = s_vcx[v];
            tb = s_vtb[v];
          } else {
            gx = cells_x(a.views[v].grid_x);
            tb = view_tb(v);
          }
          x0 = rmin.x >> DMGS_CELL_SHIFT;
          y0 = rmin.y >> DMGS_CELL_SHIFT;
          w = max(1, ((rmax.x - 1) >> DMGS_CELL_SHIFT) + 1 - x0);
        } else {
          const float4 s0 = a.splats[(size_t)DMGS_SPLAT_F4 * g];
          const int r = __float_as_int(a.splats[(size_t)DMGS_SPLAT_F4 * g + 2].z);
          tile_rect(make_float2(s0.x, s0.y), r, gx, a.grid_y, rmin, rmax);
          x0 = rmin.x;
          y0 = rmin.y;
          w = max(1, rmax.x - rmin.x);
        }
      }
    }
    s_end[threadIdx.x] = end;
    s_g[threadIdx.x] = g;
    s_tb[threadIdx.x] = tb;
    s_x0[threadIdx.x] = x0;
    s_y0[threadIdx.x] = y0;
    s_w[threadIdx.x] = w;
    s_gx[threadIdx.x] = gx;
    s_rx[threadIdx.x] = rx;
    s_ry[threadIdx.x] = ry;
    __syncthreads();
    const uint32_t win_end = min(s_end[255], hi_slot);
    const uint32_t o_begin = max(lo_slot, win_begin);
    const bool more = win_end < hi_slot && first + 256 < P;  // uniform: another window follows
    if (more) cur = fetch(first + 256 + threadIdx.x);
    // Four consecutive slots per thread: one binary search per group, then
    // the slots walk the entry's cells in order (a new entry starts at its
    // first cell), and a full group leaves as one 16-byte store per array.
    for (uint32_t o4 = (o_begin & ~3u) + 4 * threadIdx.x; o4 < win_end; o4 += 1024) {
      const uint32_t o0 = max(o4, o_begin);
      int lo = 0, hi = 255;  // smallest j with s_end[j] > o0
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_end[mid] > o0) hi = mid;
        else lo = mid + 1;
      }
      const uint32_t beg = (lo == 0) ? win_begin : s_end[lo - 1];
      uint32_t end = s_end[lo], wj = (uint32_t)s_w[lo];
      uint32_t row = (o0 - beg) / wj;
      uint32_t col = (o0 - beg) - row * wj;
      uint32_t kk[4], vv[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t o = o4 + e;
        kk[e] = vv[e] = 0u;
        if (o < o0 || o >= win_end) continue;
        if (o > o0) {  // next slot: next cell of this entry, or the first cell of the next non-empty one
          if (o >= end) {
            do {
              ++lo;
            } while (s_end[lo] <= o);
            end = s_end[lo];
            wj = (uint32_t)s_w[lo];
            row = col = 0;
          } else if (++col == wj) {
            col = 0;
            ++row;
          }
        }
        const uint32_t cx = (uint32_t)s_x0[lo] + col, cy = (uint32_t)s_y0[lo] + row;
        uint32_t key = s_tb[lo] + cy * (uint32_t)s_gx[lo] + cx;
        if (a.views) {  // cell id | tile mask << 16
          const uint32_t rx = s_rx[lo], ry = s_ry[lo];
          const int2 rmin = make_int2((int)(rx & 0xFFFFu), (int)(ry & 0xFFFFu));
          const int2 rmax = make_int2((int)(rx >> 16), (int)(ry >> 16));
          key |= cell_mask(rmin, rmax, (int)cx, (int)cy) << DMGS_CELL_KEY_BITS;
        }
        kk[e] = key;
        vv[e] = s_g[lo];
        if (a.hist) atomicAdd(&s_hist[key & hmask], 1u);
      }
      if (o4 >= o_begin && o4 + 4 <= win_end) {
        *reinterpret_cast<uint4*>(a.tile_keys + o4) = make_uint4(kk[0], kk[1], kk[2], kk[3]);
        *reinterpret_cast<uint4*>(a.values + o4) = make_uint4(vv[0], vv[1], vv[2], vv[3]);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t o = o4 + e;
          if (o >= o_begin && o < win_end) {
            a.tile_keys[o] = kk[e];
            a.values[o] = vv[e];
          }
        }
      }
    }
    win_begin = s_end[255];  // the next window starts where this one ends
    __syncthreads();
    if (!more) break;
    first += 256;
  }
  // the chunk is the first sort pass's block: its digit histogram, in the
  // radix histogram's layout (digit-major over blocks)
  if (a.hist && threadIdx.x <= hmask) a.hist[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s_hist[threadIdx.x];
}

size_t duplicate_scratch_words(size_t K) { return (K + kDupChunk - 1) / kDupChunk + 64; }

void launch_duplicate(const DupArgs& a, hipStream_t s) {
  if (a.P <= 0 || a.K == 0) return;
  if (!a.chunk_ready) hipLaunchKernelGGL(chunk_first_kernel, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
  const uint32_t chunks = (a.K + kDupChunk - 1) / kDupChunk;
  hipLaunchKernelGGL(duplicate_kernel, dim3(chunks), dim3(256), 0, s, a);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (n) hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

// Sixteen sorted keys per thread (four 16-byte loads in flight), each compared with its
// predecessor: a key that differs from the previous one closes the previous
// tile's range and opens its own.  With views, a masked key is
// (cell within the view | view << 12) and indexes ranges through the view's
// cell_base.
__device__ __forceinline__ uint32_t range_slot(uint32_t t, const ViewDesc* views) {
  return views ? (uint32_t)views[t >> 12].cell_base + (t & 0xFFFu) : t;
}

constexpr int kRangeKeys = 16;  // sorted keys per thread of tile_ranges_kernel

__global__ void tile_ranges_kernel(const uint32_t* __restrict__ keys, size_t K, uint2* __restrict__ ranges,
                                   uint32_t num_tiles, const uint32_t* __restrict__ K_dev, uint32_t key_mask,
                                   const ViewDesc* __restrict__ views) {
  if (K_dev) K = min(K, (size_t)*K_dev);
  const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * kRangeKeys;
  if (i0 >= K) return;
  uint32_t k[kRangeKeys];
  if (i0 + kRangeKeys <= K) {  // every thread but the last: four 16-byte loads in flight
#pragma unroll
    for (int q = 0; q < kRangeKeys / 4; ++q) {
      const uint4 v = reinterpret_cast<const uint4*>(keys + i0)[q];
      k[4 * q] = v.x & key_mask;
      k[4 * q + 1] = v.y & key_mask;
      k[4 * q + 2] = v.z & key_mask;
      k[4 * q + 3] = v.w & key_mask;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kRangeKeys; ++j) k[j] = (i0 + j < K) ? keys[i0 + j] & key_mask : 0u;
  }
  uint32_t prev = (i0 == 0) ? 0u : keys[i0 - 1] & key_mask;
#pragma unroll
  for (int j = 0; j < kRangeKeys; ++j) {
    const size_t i = i0 + j;
    if (i >= K) break;
    // boundaries are rare: only they look up range slots (a table read with views)
    if (i == 0 || i == K - 1 || k[j] != prev) {
      const uint32_t t = range_slot(k[j], views);
      if (t < num_tiles) {  // always true for keys from the duplicate step
        if (i == 0) {
          ranges[t].x = 0;
        } else if (k[j] != prev) {
          const uint32_t pt = range_slot(prev, views);
          if (pt < num_tiles) {
            ranges[pt].y = (uint32_t)i;
            ranges[t].x = (uint32_t)i;
          }
        }
        if (i == K - 1) ranges[t].y = (uint32_t)K;
      }
    }
    prev = k[j];
  }
}

void launch_tile_ranges(const uint32_t* tile_keys, size_t K, uint2* ranges, int num_tiles, hipStream_t s,
                        const uint32_t* K_dev, uint32_t key_mask) {
  launch_fill_u32((uint32_t*)ranges, (size_t)num_tiles * 2, 0u, s);
  const size_t threads = (K + kRangeKeys - 1) / kRangeKeys;
  if (K) hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, tile_keys,
                            K, ranges, (uint32_t)num_tiles, K_dev, key_mask, (const ViewDesc*)nullptr);
}

void launch_view_cell_ranges(const uint32_t* keys, size_t K, uint2* ranges, int num_cells, const ViewDesc* views,
                             hipStream_t s, const uint32_t* K_dev, bool ranges_zeroed) {
  if (!ranges_zeroed) launch_fill_u32((uint32_t*)ranges, (size_t)num_cells * 2, 0u, s);
  const size_t threads = (K + kRangeKeys - 1) / kRangeKeys;
  if (K) hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, keys, K,
                            ranges, (uint32_t)num_cells, K_dev, (uint32_t)DMGS_CELL_KEY_MASK, views);
}

}  // namespace dmgs
