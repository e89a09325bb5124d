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
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096
constexpr int kWaveSlice = 64 * kSortItems;           // 1024

// ---------------------------------------------------------------------------
// block-wide helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    const uint32_t x = s_wave[w];
    if (w < wave) wave_off += x;
    tot += x;
  }
  total = tot;
  __syncthreads();
  return wave_off + inc - v;
}

// ---------------------------------------------------------------------------
// scan
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in, size_t n,
                                                                  uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t s_wave[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const size_t idx = base + (size_t)i * kScanThreads + threadIdx.x;
    if (idx < n) sum += in[idx];
  }
  uint32_t total;
  block_exclusive_scan(sum, s_wave, total);
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
    uint32_t run = carry + block_exclusive_scan(tsum, s_wave, total);
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

template <bool kExclusive>
__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t* __restrict__ in,
                                                                 uint32_t* __restrict__ out, size_t n,
                                                                 const uint32_t* __restrict__ spine) {
  __shared__ uint32_t s_wave[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t tsum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const size_t idx = base + i;
    v[i] = idx < n ? in[idx] : 0u;
    tsum += v[i];
  }
  uint32_t total;
  uint32_t run = spine[blockIdx.x] + block_exclusive_scan(tsum, s_wave, total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const size_t idx = base + i;
    const uint32_t x = v[i];
    if (!kExclusive) run += x;
    if (idx < n) out[idx] = run;
    if (kExclusive) run += x;
  }
}

size_t scan_scratch_words(size_t n) { return (n + kScanTile - 1) / kScanTile + 64; }

static void scan_impl(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, bool exclusive,
                      hipStream_t s) {
  if (n == 0) return;
  const size_t nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, scratch);
  hipLaunchKernelGGL(scan_spine_kernel, dim3(1), dim3(kScanThreads), 0, s, scratch, nb);
  if (exclusive)
    hipLaunchKernelGGL(scan_apply_kernel<true>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n,
                       (const uint32_t*)scratch);
  else
    hipLaunchKernelGGL(scan_apply_kernel<false>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n,
                       (const uint32_t*)scratch);
}

void launch_inclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, hipStream_t s) {
  scan_impl(in, out, n, scratch, false, s);
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t* __restrict__ keys, size_t n,
                                                                 int shift, int nbits,
                                                                 uint32_t* __restrict__ hist, uint32_t nb) {
  __shared__ uint32_t cnt[256];
  const uint32_t mask = (1u << nbits) - 1u;
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int i = 0; i < kSortItems; ++i) {
    const size_t idx = base + (size_t)i * kSortThreads + threadIdx.x;
    if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & mask], 1u);
  }
  __syncthreads();
  if (threadIdx.x <= mask) hist[(size_t)threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, size_t n, int shift, int nbits, const uint32_t* __restrict__ hist_excl,
    uint32_t nb) {
  __shared__ uint32_t wcnt[kSortWaves][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t mask = (1u << nbits) - 1u;
  for (int w = 0; w < kSortWaves; ++w) wcnt[w][threadIdx.x] = 0;
  __syncthreads();

  const size_t wbase = (size_t)blockIdx.x * kSortTile + (size_t)wave * kWaveSlice;
  const uint64_t lt = lanemask_lt();
  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const size_t idx = wbase + (size_t)i * 64 + lane;
    const bool valid = idx < n;
    key[i] = valid ? kin[idx] : 0u;
    val[i] = valid ? vin[idx] : 0u;
    const uint32_t d = (key[i] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t before = wcnt[wave][d];
    rank[i] = before + (uint32_t)__popcll(peers & lt);
    const int leader = 63 - __clzll(peers);
    if (valid && lane == leader) wcnt[wave][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // exclusive prefix over waves per digit, plus this block's global offset
  if (threadIdx.x <= mask) {
    uint32_t run = hist_excl[(size_t)threadIdx.x * nb + blockIdx.x];
    for (int w = 0; w < kSortWaves; ++w) {
      const uint32_t c = wcnt[w][threadIdx.x];
      wcnt[w][threadIdx.x] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const size_t idx = wbase + (size_t)i * 64 + lane;
    if (idx < n) {
      const uint32_t d = (key[i] >> shift) & mask;
      const uint32_t pos = wcnt[wave][d] + rank[i];
      if (pos < n) {  // always true for a consistent histogram; guards the write
        kout[pos] = key[i];
        vout[pos] = val[i];
      }
    }
  }
}

size_t radix_scratch_words(size_t n) {
  const size_t nb = (n + kSortTile - 1) / kSortTile;
  const size_t hist = 256 * nb;
  return hist + scan_scratch_words(hist) + 64;
}

void radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                      uint32_t* scratch, hipStream_t s) {
  if (n == 0) return;
  if (bits < 0) bits = 0;
  if (bits > 32) bits = 32;
  const uint32_t nb = (uint32_t)((n + kSortTile - 1) / kSortTile);
  const int passes = bits <= 16 ? 2 : 4;
  int widths[4];
  int left = bits;
  for (int p = 0; p < passes; ++p) {
    const int w = (left + (passes - p) - 1) / (passes - p);
    widths[p] = w;
    left -= w;
  }
  uint32_t* hist = scratch;
  uint32_t* scan_scratch = scratch + 256 * (size_t)nb;
  uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
  int shift = 0;
  for (int p = 0; p < passes; ++p) {
    const int w = widths[p];
    const size_t hn = (size_t)(1u << w) * nb;
    hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ka, n, shift, w,
                       hist, nb);
    scan_impl(hist, hist, hn, scan_scratch, true, s);
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ka,
                       (const uint32_t*)va, kb, vb, n, shift, w, (const uint32_t*)hist, nb);
    shift += w;
    uint32_t* t = ka;
    ka = kb;
    kb = t;
    t = va;
    va = vb;
    vb = t;
  }
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

__global__ __launch_bounds__(256) void duplicate_kernel(int P, const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ offsets,
                                                        const float4* __restrict__ splats, int gx, int gy,
                                                        uint32_t* __restrict__ tile_keys,
                                                        uint32_t* __restrict__ values, uint32_t K) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P) return;
  const uint32_t g = order[s];
  uint32_t off = (s == 0) ? 0u : offsets[s - 1];
  const uint32_t end = min(offsets[s], K);
  if (end <= off || g >= (uint32_t)P) return;  // culled: no tiles
  const float4 s0 = splats[(size_t)DMGS_SPLAT_F4 * g];
  const int r = __float_as_int(splats[(size_t)DMGS_SPLAT_F4 * g + 2].z);
  int2 rmin, rmax;
  tile_rect(make_float2(s0.x, s0.y), r, gx, gy, rmin, rmax);
  for (int y = rmin.y; y < rmax.y; ++y)
    for (int x = rmin.x; x < rmax.x && off < end; ++x) {
      tile_keys[off] = (uint32_t)(y * gx + x);
      values[off] = g;
      ++off;
    }
}

void launch_duplicate(int P, const uint32_t* order, const uint32_t* offsets, const float4* splats, int grid_x,
                      int grid_y, uint32_t* tile_keys, uint32_t* values, uint32_t K, hipStream_t s) {
  if (P <= 0) return;
  hipLaunchKernelGGL(duplicate_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, order, offsets, splats, grid_x,
                     grid_y, tile_keys, values, K);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void tile_ranges_kernel(const uint32_t* __restrict__ keys, size_t K, uint2* __restrict__ ranges,
                                   uint32_t num_tiles) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  const uint32_t t = keys[i];
  if (t >= num_tiles) return;  // cannot happen for keys from the duplicate step
  if (i == 0) {
    ranges[t].x = 0;
  } else {
    const uint32_t prev = keys[i - 1];
    if (t != prev && prev < num_tiles) {
      ranges[prev].y = (uint32_t)i;
      ranges[t].x = (uint32_t)i;
    }
  }
  if (i == K - 1) ranges[t].y = (uint32_t)K;
}

void launch_tile_ranges(const uint32_t* tile_keys, size_t K, uint2* ranges, int num_tiles, hipStream_t s) {
  const size_t words = (size_t)num_tiles * 2;
  if (words) hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s,
                                (uint32_t*)ranges, words, 0u);
  if (K) hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s, tile_keys, K,
                            ranges, (uint32_t)num_tiles);
}

}  // namespace dmgs
