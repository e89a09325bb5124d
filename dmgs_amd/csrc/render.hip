// Front-to-back alpha blending (forward) and its reverse-traversal backward
// for gfx950.
//
// Semantics follow the reference renderer [U] (diff-gaussian-rasterization,
// called from gaussian_renderer/__init__.py:86-94): per pixel, in sorted
// order, power = -1/2 (a dx^2 + c dy^2) - b dx dy, skip if power > 0,
// alpha = min(0.99, o exp(power)), skip if alpha < 1/255, stop (without
// adding) once T (1 - alpha) < 1e-4, C += c alpha T; output C + T bg.
//
// Layout: one workgroup of 256 lanes (4 wave64) per 16x16 tile; Gaussian
// batches of 256 splat records are staged in LDS with one coalesced 48-byte
// gather per lane.
//
// Backward: the list is walked from the back, starting at the largest
// n_contrib of the tile (entries behind every pixel's last contributor
// cannot contribute), transmittance is recomputed by division instead of
// being stored, and each wave reduces its 9 per-Gaussian partial gradients
// with DPP row sums into a per-batch LDS accumulator; the workgroup flushes
// that accumulator with one global atomic per (Gaussian, value) per batch.
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

struct SplatLDS {
  float2 xy[DMGS_BLOCK_SIZE];
  float4 con_o[DMGS_BLOCK_SIZE];  // conic a, b, c, opacity
  float4 rgb_id[DMGS_BLOCK_SIZE]; // r, g, b, id (bits)
};

__device__ __forceinline__ void stage_splat(SplatLDS& L, int slot, const float4* __restrict__ splats, uint32_t id,
                                            uint32_t P) {
  if (id >= P) {  // cannot happen for lists built by the binning step; a zero-opacity entry never contributes
    L.xy[slot] = make_float2(0.f, 0.f);
    L.con_o[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    L.rgb_id[slot] = make_float4(0.f, 0.f, 0.f, __uint_as_float(0u));
    return;
  }
  const float4* sp = splats + (size_t)DMGS_SPLAT_F4 * id;
  const float4 s0 = sp[0];
  const float4 s1 = sp[1];
  const float4 s2 = sp[2];
  L.xy[slot] = make_float2(s0.x, s0.y);
  L.con_o[slot] = make_float4(s0.z, s0.w, s1.x, s1.y);
  L.rgb_id[slot] = make_float4(s1.z, s1.w, s2.x, __uint_as_float(id));
}

__global__ __launch_bounds__(DMGS_BLOCK_SIZE) void render_fwd_kernel(RenderArgs a) {
  __shared__ SplatLDS L;
  const int tile = blockIdx.x;
  const int tx = tile % a.grid_x, ty = tile / a.grid_x;
  const int lx = threadIdx.x & (DMGS_BLOCK_X - 1), ly = threadIdx.x / DMGS_BLOCK_X;
  const int px = tx * DMGS_BLOCK_X + lx, py = ty * DMGS_BLOCK_Y + ly;
  const bool inside = px < a.W && py < a.H;
  const float pxf = (float)px, pyf = (float)py;
  uint2 range = a.ranges[tile];
  range.y = min(range.y, a.K);
  range.x = min(range.x, range.y);
  const int todo_total = (int)(range.y - range.x);
  const int rounds = (todo_total + DMGS_BLOCK_SIZE - 1) / DMGS_BLOCK_SIZE;

  bool done = !inside;
  float T = 1.0f;
  float C0 = 0.f, C1 = 0.f, C2 = 0.f;
  uint32_t contributor = 0, last_contributor = 0;

  for (int r = 0; r < rounds; ++r) {
    if (__syncthreads_count(done) == DMGS_BLOCK_SIZE) break;
    const int pos = r * DMGS_BLOCK_SIZE + threadIdx.x;
    if (pos < todo_total) stage_splat(L, threadIdx.x, a.splats, a.point_list[range.x + pos], a.P);
    __syncthreads();
    const int n = min(DMGS_BLOCK_SIZE, todo_total - r * DMGS_BLOCK_SIZE);
    for (int j = 0; !done && j < n; ++j) {
      ++contributor;
      const float2 xy = L.xy[j];
      const float4 co = L.con_o[j];
      const float dx = xy.x - pxf, dy = xy.y - pyf;
      const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
      if (power > 0.0f) continue;
      const float alpha = fminf(0.99f, co.w * __expf(power));
      if (alpha < 1.0f / 255.0f) continue;
      const float test_T = T * (1.0f - alpha);
      if (test_T < 0.0001f) {
        done = true;
        continue;
      }
      const float4 c = L.rgb_id[j];
      const float w = alpha * T;
      C0 += c.x * w;
      C1 += c.y * w;
      C2 += c.z * w;
      T = test_T;
      last_contributor = contributor;
    }
  }
  if (inside) {
    const int pid = py * a.W + px;
    const int HW = a.W * a.H;
    a.final_T[pid] = T;
    a.n_contrib[pid] = last_contributor;
    a.out_color[pid] = C0 + T * a.bg[0];
    a.out_color[HW + pid] = C1 + T * a.bg[1];
    a.out_color[2 * HW + pid] = C2 + T * a.bg[2];
  }
}

__global__ __launch_bounds__(DMGS_BLOCK_SIZE) void render_bwd_kernel(RenderArgs a) {
  __shared__ SplatLDS L;
  __shared__ uint32_t s_max[DMGS_BLOCK_SIZE / DMGS_WAVE];
  __shared__ float s_acc[DMGS_BLOCK_SIZE][DMGS_GACC_USED];
  const int tile = blockIdx.x;
  const int tx = tile % a.grid_x, ty = tile / a.grid_x;
  const int lx = threadIdx.x & (DMGS_BLOCK_X - 1), ly = threadIdx.x / DMGS_BLOCK_X;
  const int px = tx * DMGS_BLOCK_X + lx, py = ty * DMGS_BLOCK_Y + ly;
  const bool inside = px < a.W && py < a.H;
  const float pxf = (float)px, pyf = (float)py;
  uint2 range = a.ranges[tile];
  range.y = min(range.y, a.K);
  range.x = min(range.x, range.y);
  const int HW = a.W * a.H;
  const int pid = py * a.W + px;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

  const float T_final = inside ? a.final_T[pid] : 0.0f;
  const uint32_t last_contributor = inside ? a.n_contrib[pid] : 0u;
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  if (inside) {
    g0 = a.dL_dpix[pid];
    g1 = a.dL_dpix[HW + pid];
    g2 = a.dL_dpix[2 * HW + pid];
  }
  const float bg_dot = a.bg[0] * g0 + a.bg[1] * g1 + a.bg[2] * g2;
  const float ddelx_dx = 0.5f * (float)a.W, ddely_dy = 0.5f * (float)a.H;

  // Entries at list positions >= max(n_contrib) of the tile contribute nothing.
  uint32_t m = last_contributor;
  for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
  if (lane == 0) s_max[wave] = m;
  __syncthreads();
  uint32_t tile_max = 0;
  for (int w = 0; w < DMGS_BLOCK_SIZE / DMGS_WAVE; ++w) tile_max = max(tile_max, s_max[w]);
  const int todo_total = min((int)(range.y - range.x), (int)tile_max);
  const int rounds = (todo_total + DMGS_BLOCK_SIZE - 1) / DMGS_BLOCK_SIZE;

  float T = T_final;
  uint32_t contributor = (uint32_t)todo_total;
  float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
  float last_alpha = 0.f, lc0 = 0.f, lc1 = 0.f, lc2 = 0.f;

  const unsigned row_lane = lane & 15;  // lane within its DPP row of 16
  for (int r = 0; r < rounds; ++r) {
    __syncthreads();
    const int pos = r * DMGS_BLOCK_SIZE + threadIdx.x;
    if (pos < todo_total) stage_splat(L, threadIdx.x, a.splats, a.point_list[range.x + todo_total - pos - 1], a.P);
#pragma unroll
    for (int k = 0; k < DMGS_GACC_USED; ++k) s_acc[threadIdx.x][k] = 0.f;
    __syncthreads();
    const int n = min(DMGS_BLOCK_SIZE, todo_total - r * DMGS_BLOCK_SIZE);
    for (int j = 0; j < n; ++j) {
      --contributor;
      float v[DMGS_GACC_USED];
#pragma unroll
      for (int k = 0; k < DMGS_GACC_USED; ++k) v[k] = 0.f;
      bool contrib = false;
      const float2 xy = L.xy[j];
      const float4 co = L.con_o[j];
      if (contributor < last_contributor) {
        const float dx = xy.x - pxf, dy = xy.y - pyf;
        const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
        const float G = __expf(power);
        const float alpha = fminf(0.99f, co.w * G);
        if (power <= 0.0f && alpha >= 1.0f / 255.0f) {
          contrib = true;
          const float4 c = L.rgb_id[j];
          T = T / (1.f - alpha);
          const float dchannel_dcolor = alpha * T;
          acc0 = last_alpha * lc0 + (1.f - last_alpha) * acc0;
          acc1 = last_alpha * lc1 + (1.f - last_alpha) * acc1;
          acc2 = last_alpha * lc2 + (1.f - last_alpha) * acc2;
          lc0 = c.x;
          lc1 = c.y;
          lc2 = c.z;
          float dL_dalpha = (c.x - acc0) * g0 + (c.y - acc1) * g1 + (c.z - acc2) * g2;
          v[6] = dchannel_dcolor * g0;
          v[7] = dchannel_dcolor * g1;
          v[8] = dchannel_dcolor * g2;
          dL_dalpha *= T;
          last_alpha = alpha;
          dL_dalpha += (-T_final / (1.f - alpha)) * bg_dot;
          const float dL_dG = co.w * dL_dalpha;
          const float gdx = G * dx, gdy = G * dy;
          const float dG_ddelx = -gdx * co.x - gdy * co.y;
          const float dG_ddely = -gdy * co.z - gdx * co.y;
          v[0] = dL_dG * dG_ddelx * ddelx_dx;
          v[1] = dL_dG * dG_ddely * ddely_dy;
          v[2] = -0.5f * gdx * dx * dL_dG;
          v[3] = -0.5f * gdx * dy * dL_dG;
          v[4] = -0.5f * gdy * dy * dL_dG;
          v[5] = G * dL_dalpha;
        }
      }
      if (__ballot(contrib) == 0) continue;  // wave-uniform
      // sum over each 16-lane row, then lane k of every row adds value k into
      // the batch accumulator in LDS (one ds_add with 4 x 9 active lanes)
      float mine = 0.f;
#pragma unroll
      for (int k = 0; k < DMGS_GACC_USED; ++k) {
        const float s = row16_sum(v[k]);
        mine = (row_lane == (unsigned)k) ? s : mine;
      }
      if (row_lane < DMGS_GACC_USED) atomicAdd(&s_acc[j][row_lane], mine);
    }
    __syncthreads();
    // flush the batch: consecutive lanes cover the 9 values of one Gaussian
    for (int i = threadIdx.x; i < n * DMGS_GACC_USED; i += DMGS_BLOCK_SIZE) {
      const int slot = i / DMGS_GACC_USED, k = i - slot * DMGS_GACC_USED;
      const float val = s_acc[slot][k];
      if (val != 0.f) {
        const uint32_t gid = __float_as_uint(L.rgb_id[slot].w);
        unsafeAtomicAdd(a.gacc + (size_t)gid * DMGS_GACC_STRIDE + k, val);
      }
    }
  }
}

void launch_render_fwd(const RenderArgs& a, hipStream_t s) {
  const int tiles = a.grid_x * a.grid_y;
  if (tiles <= 0) return;
  hipLaunchKernelGGL(render_fwd_kernel, dim3(tiles), dim3(DMGS_BLOCK_SIZE), 0, s, a);
}

void launch_render_bwd(const RenderArgs& a, hipStream_t s) {
  const int tiles = a.grid_x * a.grid_y;
  if (tiles <= 0) return;
  hipLaunchKernelGGL(render_bwd_kernel, dim3(tiles), dim3(DMGS_BLOCK_SIZE), 0, s, a);
}

}  // namespace dmgs
