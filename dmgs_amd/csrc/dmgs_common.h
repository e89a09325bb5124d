// Shared device helpers for the gfx950 rasterizer kernels.
//
// Conventions (the reference rasterizer's, diff-gaussian-rasterization [U],
// called from gaussian_renderer/__init__.py:86-94):
//   * view / projection matrices are 16 floats in the order the reference
//     stores world_view_transform / full_proj_transform (row-vector
//     convention), i.e. p_view.x = m[0] x + m[4] y + m[8] z + m[12];
//   * tiles are 16 x 16 pixels; tile id = ty * grid_x + tx;
//   * covariances are 6-vectors (xx, xy, xz, yy, yz, zz).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DMGS_BLOCK_X 16
#define DMGS_BLOCK_Y 16
#define DMGS_BLOCK_SIZE (DMGS_BLOCK_X * DMGS_BLOCK_Y)
#define DMGS_WAVE 64

// Per-Gaussian render record ("splat"): three float4 per Gaussian.
//   s0 = (pixel x, pixel y, conic a, conic b)
//   s1 = (conic c, opacity, red, green)
//   s2 = (blue, view depth, radius as int bits, 0)
// One 48-byte gather feeds both blend kernels.
#define DMGS_SPLAT_F4 3

// Per-Gaussian accumulator written by the blend backward with atomics: 9
// values per Gaussian in two planes, so each consumer reads only its part
// (the geometry backward the first, the SH backward the second):
//   plane 1, 6 floats (24 B) per Gaussian: [0] dL/dmean2D.x (NDC)
//     [1] dL/dmean2D.y  [2] dL/dconic.a  [3] dL/dconic.b (half)
//     [4] dL/dconic.c  [5] dL/dopacity
//   plane 2 (after n x 6 floats), 3 floats (12 B) per Gaussian: [6..8] dL/dcolour rgb
// Value q of Gaussian id in an accumulator of n Gaussians is at gacc_index(n, id, q).
#define DMGS_GACC_STRIDE 9  // floats per Gaussian, both planes
#define DMGS_GACC_USED 9
#define DMGS_GACC_GEO 6     // values in plane 1

namespace dmgs {

__host__ __device__ __forceinline__ size_t gacc_index(size_t n, size_t id, int q) {
  return q < DMGS_GACC_GEO ? id * DMGS_GACC_GEO + (size_t)q : n * DMGS_GACC_GEO + id * 3 + (size_t)(q - DMGS_GACC_GEO);
}

// One camera of a multi-view launch (host-filled, copied to the device).
struct ViewDesc {
  float view[16];       // world_view_transform, reference storage order
  float proj[16];       // full_proj_transform
  float campos[4];      // xyz + pad
  float bg[4];          // rgb + pad
  float tanfovx, tanfovy, focal_x, focal_y;
  int32_t W, H, grid_x, grid_y;
  int32_t tile_base;    // first global tile id of the view
  int32_t cell_base;    // first global binning-cell id of the view (multi-view lists)
  int64_t pixel_base;   // first pixel of the view in the packed per-pixel buffers
};

__device__ __forceinline__ float3 xform_point_4x3(const float3 p, const float* m) {
  return make_float3(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12],
                     m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
                     m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]);
}

__device__ __forceinline__ float4 xform_point_4x4(const float3 p, const float* m) {
  return make_float4(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12],
                     m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
                     m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14],
                     m[3] * p.x + m[7] * p.y + m[11] * p.z + m[15]);
}

__device__ __forceinline__ float ndc_to_pix(float v, int S) { return ((v + 1.0f) * S - 1.0f) * 0.5f; }

// Screen rectangle of tiles touched by a splat of integer radius r centred
// at p: [min, max) per axis, clamped to the grid.  The float operation order
// is fixed so that the host oracle can reproduce it exactly.
__device__ __forceinline__ void tile_rect(float2 p, int r, int gx, int gy, int2& rmin, int2& rmax) {
  const float fr = (float)r;
  const float b = (float)DMGS_BLOCK_X;
  rmin.x = min(gx, max(0, (int)((p.x - fr) / b)));
  rmin.y = min(gy, max(0, (int)((p.y - fr) / b)));
  rmax.x = min(gx, max(0, (int)((((p.x + fr) + b) - 1.0f) / b)));
  rmax.y = min(gy, max(0, (int)((((p.y + fr) + b) - 1.0f) / b)));
}

// --- binning cells of the multi-view lists --------------------------------
// The multi-view path bins each splat into cells of 4 x 4 blend tiles (64 x 64
// pixels).  An instance's key is  cell id | tile mask << 16,  the mask
// marking which of the cell's 16 tiles (bit (ty & 3) * 4 + (tx & 3)) the
// splat's tile rectangle covers; sorting on the low 16 bits only carries the
// mask along.  A tile's list is its cell's list filtered by its bit, in the
// same order, so it equals the per-tile list of the reference binning.
#define DMGS_CELL_SHIFT 2          // log2 of the cell edge in tiles
#define DMGS_CELL_KEY_BITS 16      // cell id bits of a key (at most 65536 cells)
#define DMGS_CELL_KEY_MASK 0xFFFFu
__device__ __forceinline__ int cells_x(int gx) { return (gx + (1 << DMGS_CELL_SHIFT) - 1) >> DMGS_CELL_SHIFT; }
// Number of cells a non-empty tile rectangle [rmin, rmax) touches.
__device__ __forceinline__ int cell_count(int2 rmin, int2 rmax) {
  return (((rmax.x - 1) >> DMGS_CELL_SHIFT) - (rmin.x >> DMGS_CELL_SHIFT) + 1) *
         (((rmax.y - 1) >> DMGS_CELL_SHIFT) - (rmin.y >> DMGS_CELL_SHIFT) + 1);
}
// Mask of the tiles of cell (cx, cy) inside the tile rectangle [rmin, rmax):
// the covered columns and rows of the cell, as an outer product of bit runs.
__device__ __forceinline__ uint32_t cell_mask(int2 rmin, int2 rmax, int cx, int cy) {
  constexpr int E = 1 << DMGS_CELL_SHIFT;
  static_assert(E == 4, "the row replication below is for cells of 4 x 4 tiles");
  const int x0 = max(rmin.x - cx * E, 0), x1 = min(rmax.x - cx * E, E);
  const int y0 = max(rmin.y - cy * E, 0), y1 = min(rmax.y - cy * E, E);
  if (x1 <= x0 || y1 <= y0) return 0u;
  const uint32_t row = ((1u << (x1 - x0)) - 1u) << x0;              // columns of one cell row
  const uint32_t rows = ((1u << (E * (y1 - y0))) - 1u) << (E * y0);  // the covered rows' bits
  return (row * 0x1111u) & rows;                                    // the row in all four rows, masked
}

// Tile rectangle [rmin, rmax) packed in 8 bytes: (x0 | x1 << 16, y0 | y1 << 16).
__device__ __forceinline__ uint2 pack_rect(int2 rmin, int2 rmax) {
  return make_uint2((uint32_t)rmin.x | ((uint32_t)rmax.x << 16), (uint32_t)rmin.y | ((uint32_t)rmax.y << 16));
}
__device__ __forceinline__ void unpack_rect(uint2 r, int2& rmin, int2& rmax) {
  rmin = make_int2((int)(r.x & 0xFFFFu), (int)(r.y & 0xFFFFu));
  rmax = make_int2((int)(r.x >> 16), (int)(r.y >> 16));
}

// --- wave64 reductions --------------------------------------------------
// Sum over the 16 lanes of each DPP row; every lane of the row gets the sum.
__device__ __forceinline__ float row16_sum(float v) {
  int x = __float_as_int(v);
  v += __int_as_float(__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true));   // quad_perm 1,0,3,2
  x = __float_as_int(v);
  v += __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true));   // quad_perm 2,3,0,1
  x = __float_as_int(v);
  v += __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x124, 0xF, 0xF, true));  // row_ror 4
  x = __float_as_int(v);
  v += __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, true));  // row_ror 8
  return v;
}

// Wave-uniform sum of v over all 64 lanes (all lanes must be active).
__device__ __forceinline__ float wave_sum(float v) {
  v = row16_sum(v);
  const int x = __float_as_int(v);
  return (__int_as_float(__builtin_amdgcn_readlane(x, 0)) + __int_as_float(__builtin_amdgcn_readlane(x, 16))) +
         (__int_as_float(__builtin_amdgcn_readlane(x, 32)) + __int_as_float(__builtin_amdgcn_readlane(x, 48)));
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  const unsigned l = __lane_id();
  return (l == 0) ? 0ull : (~0ull >> (64 - l));
}

// Inclusive prefix sum across the wave (uint32).
// Inclusive scan over the wave with DPP: prefix sums within each 16-lane row
// (row_shr 1, 2, 4, 8; lanes whose source falls outside the row add 0), then
// row_bcast:15 adds row r's last lane into row r + 1 (rows 1 and 3) and
// row_bcast:31 adds lane 31 into rows 2 and 3.  Ten VALU, no LDS.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}

// Exclusive scan of one value per thread over the work-group (s_wave: one
// word per wave); total receives the sum.  Every thread must call it.
// NW > 0: the work-group has exactly NW waves (a fixed-length wave loop).
template <int NW = 0>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
  const int nw = NW > 0 ? NW : (int)(blockDim.x >> 6);
#pragma unroll
  for (int w = 0; w < nw; ++w) {
    const uint32_t x = s_wave[w];
    if (w < wave) wave_off += x;
    tot += x;
  }
  total = tot;
  __syncthreads();
  return wave_off + inc - v;
}

// Load the first 16 SH coefficients (RGB each) of one Gaussian into registers;
// 16-byte vector loads when the per-Gaussian stride allows it.
__device__ __forceinline__ void load_sh16(const float* __restrict__ sh, int M, float (&r)[48]) {
  if (M == 16 && ((reinterpret_cast<uintptr_t>(sh) & 15) == 0)) {
    const float4* s4 = reinterpret_cast<const float4*>(sh);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      const float4 v = s4[q];
      r[4 * q] = v.x;
      r[4 * q + 1] = v.y;
      r[4 * q + 2] = v.z;
      r[4 * q + 3] = v.w;
    }
  } else {  // fewer coefficients: every load issued at once (indices past 3M read the last one, then zero)
    const int last = 3 * min(M, 16) - 1;
#pragma unroll
    for (int i = 0; i < 48; ++i) r[i] = last >= 0 ? sh[min(i, last)] : 0.f;
#pragma unroll
    for (int i = 0; i < 48; ++i) r[i] = (i / 3 < M) ? r[i] : 0.f;
  }
}

}  // namespace dmgs
