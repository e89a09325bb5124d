// Front-to-back alpha blending (forward) and its reverse-traversal backward
// for gfx950.
//
// Semantics follow the reference renderer [U] (diff-gaussian-rasterization,
// called from gaussian_renderer/__init__.py:86-94): per pixel, in sorted
// order, power = -1/2 (a dx^2 + c dy^2) - b dx dy, skip if power > 0,
// alpha = min(0.99, o exp(power)), skip if alpha < 1/255, stop (without
// adding) once T (1 - alpha) < 1e-4, C += c alpha T; output C + T bg.
//
// Layout: a workgroup covers SUB x SUB pixels of a 16x16 binning tile (SUB =
// 16: the whole tile with 4 wave64; SUB = 8: one wave64 per quarter tile, four
// workgroups reading the same list).  Splat records are staged in LDS in
// batches of BATCH with one 48-byte gather per record.  Every variant
// evaluates each pixel in the same order, so they give identical images.
//
// Backward: the list is walked from the back, starting at the largest
// n_contrib of the workgroup (entries behind every pixel's last contributor
// cannot contribute), transmittance is recomputed by division instead of
// being stored, and each wave reduces its 9 per-splat partial gradients with
// DPP row sums into a per-batch LDS accumulator; the workgroup flushes that
// accumulator with one global atomic per (splat, value) per batch.
//
// Several views in one launch: a global tile id indexes the ranges; the view
// table gives each tile's image size, background and packed output offset.
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

template <int BATCH>
struct SplatLDS {
  float2 xy[BATCH];
  float4 con_o[BATCH];   // conic a, b, c, opacity
  float4 rgb_id[BATCH];  // r, g, b, id (bits)
};

template <int BATCH>
__device__ __forceinline__ void stage_splat(SplatLDS<BATCH>& L, int slot, const float4* __restrict__ splats,
                                            uint32_t id, uint32_t P) {
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

// Everything a workgroup needs to know about its pixels.
struct TileCtx {
  int tile;            // global tile id (indexes ranges)
  int px, py;          // this lane's pixel
  int W, H;
  bool inside;
  size_t pix;          // index into the per-pixel buffers
  size_t plane;        // W * H of the view
  size_t img;          // offset of the view's colour planes
  float bg0, bg1, bg2;
  uint32_t bit;        // multi-view: this tile's bit in its cell's instance tags
  uint2 range;         // single view: the tile's list; multi-view: its cell's list
};

template <int SUB>
__device__ __forceinline__ TileCtx tile_ctx(const RenderArgs& a) {
  constexpr int Q = DMGS_BLOCK_X / SUB;  // workgroups per tile edge
  TileCtx c;
  c.tile = blockIdx.x / (Q * Q);
  const int q = blockIdx.x - c.tile * Q * Q;
  int lt = c.tile, gx = a.grid_x;
  int64_t pbase = 0;
  int list = c.tile;
  c.bit = 0;
  if (a.views) {
    const ViewDesc& v = a.views[a.tile_view[c.tile]];
    lt = c.tile - v.tile_base;
    gx = v.grid_x;
    const int ltx = lt % gx, lty = lt / gx;
    list = v.cell_base + (lty >> 1) * cells_x(gx) + (ltx >> 1);
    c.bit = 1u << (((lty & 1) << 1) | (ltx & 1));
    c.W = v.W;
    c.H = v.H;
    c.bg0 = v.bg[0];
    c.bg1 = v.bg[1];
    c.bg2 = v.bg[2];
    pbase = v.pixel_base;
  } else {
    c.W = a.W;
    c.H = a.H;
    c.bg0 = a.bg[0];
    c.bg1 = a.bg[1];
    c.bg2 = a.bg[2];
  }
  const int tx = lt % gx, ty = lt / gx;
  const int lx = threadIdx.x % SUB, ly = threadIdx.x / SUB;
  c.px = tx * DMGS_BLOCK_X + (q % Q) * SUB + lx;
  c.py = ty * DMGS_BLOCK_Y + (q / Q) * SUB + ly;
  c.inside = c.px < c.W && c.py < c.H;
  c.plane = (size_t)c.W * c.H;
  c.pix = (size_t)pbase + (size_t)c.py * c.W + c.px;
  c.img = 3 * (size_t)pbase;
  uint32_t K = a.K;
  if (a.K_dev) K = min(K, *a.K_dev);
  uint2 r = a.ranges[list];
  r.y = min(r.y, K);
  r.x = min(r.x, r.y);
  c.range = r;
  return c;
}

// Stage one batch of a cell list (multi-view): visit BATCH list positions —
// forward: base, base+1, ... below len; backward: top-1, top-2, ... down to 0
// — keep the instances tagged with this tile's bit and stage them in visit
// order.  s_pos receives each staged instance's list position.  Returns the
// number staged (workgroup-uniform); ends with a barrier.
template <int NT, int BATCH, bool BACKWARD>
__device__ __forceinline__ int stage_cell_batch(SplatLDS<BATCH>& L, uint32_t* s_pos, uint32_t* s_wc,
                                                const RenderArgs& a, const TileCtx& c, int start, int limit) {
  constexpr int NW = NT / DMGS_WAVE;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  int run = 0;
#pragma unroll
  for (int k = 0; k < (BATCH + NT - 1) / NT; ++k) {
    const int s = k * NT + (int)threadIdx.x;
    const int q = BACKWARD ? start - 1 - s : start + s;
    const bool valid = s < BATCH && (BACKWARD ? q >= 0 : q < limit);
    const uint32_t val = valid ? a.point_list[c.range.x + q] : 0u;
    const bool member = valid && ((val >> DMGS_CELL_MASK_SHIFT) & c.bit);
    const uint64_t bal = __ballot(member);
    if (lane == 0) s_wc[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    int before = run, total = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const int cnt = (int)s_wc[w];
      before += (w < (int)wave) ? cnt : 0;
      total += cnt;
    }
    if (member) {
      const int slot = before + (int)__popcll(bal & lt);
      stage_splat(L, slot, a.splats, val & DMGS_CELL_ID_MASK, a.P);
      s_pos[slot] = (uint32_t)q;
    }
    run += total;
    __syncthreads();
  }
  return run;
}

template <int SUB, int BATCH>
__global__ __launch_bounds__(SUB* SUB) void render_fwd_kernel(RenderArgs a) {
  constexpr int NT = SUB * SUB;
  __shared__ SplatLDS<BATCH> L;
  const TileCtx c = tile_ctx<SUB>(a);
  const float pxf = (float)c.px, pyf = (float)c.py;
  const int todo_total = (int)(c.range.y - c.range.x);
  const int rounds = (todo_total + BATCH - 1) / BATCH;

  bool done = !c.inside;
  float T = 1.0f;
  float C0 = 0.f, C1 = 0.f, C2 = 0.f;
  uint32_t contributor = 0, last_contributor = 0;

  for (int r = 0; r < rounds; ++r) {
    if (__syncthreads_count(done) == NT) break;
    for (int s = threadIdx.x; s < BATCH; s += NT) {
      const int pos = r * BATCH + s;
      if (pos < todo_total) stage_splat(L, s, a.splats, a.point_list[c.range.x + pos], a.P);
    }
    __syncthreads();
    const int n = min(BATCH, todo_total - r * BATCH);
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
      const float4 col = L.rgb_id[j];
      const float w = alpha * T;
      C0 += col.x * w;
      C1 += col.y * w;
      C2 += col.z * w;
      T = test_T;
      last_contributor = contributor;
    }
  }
  if (c.inside) {
    const size_t pid = c.pix - c.img / 3;  // pixel index within the view
    a.final_T[c.pix] = T;
    a.n_contrib[c.pix] = last_contributor;
    float* out = a.out_color + c.img;
    out[pid] = C0 + T * c.bg0;
    out[c.plane + pid] = C1 + T * c.bg1;
    out[2 * c.plane + pid] = C2 + T * c.bg2;
  }
}

// Forward over cell lists (multi-view).  Same per-pixel arithmetic and order
// as render_fwd_kernel; the per-pixel "last contributor" is one past the
// list position of the last contributing instance.
template <int SUB, int BATCH>
__global__ __launch_bounds__(SUB* SUB) void render_fwd_cells_kernel(RenderArgs a) {
  constexpr int NT = SUB * SUB;
  __shared__ SplatLDS<BATCH> L;
  __shared__ uint32_t s_pos[BATCH];
  __shared__ uint32_t s_wc[NT / DMGS_WAVE];
  const TileCtx c = tile_ctx<SUB>(a);
  const float pxf = (float)c.px, pyf = (float)c.py;
  const int len = (int)(c.range.y - c.range.x);

  bool done = !c.inside;
  float T = 1.0f;
  float C0 = 0.f, C1 = 0.f, C2 = 0.f;
  uint32_t last_contributor = 0;

  for (int base = 0; base < len; base += BATCH) {
    if (__syncthreads_count(done) == NT) break;
    const int n = stage_cell_batch<NT, BATCH, false>(L, s_pos, s_wc, a, c, base, len);
    for (int j = 0; !done && j < n; ++j) {
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
      const float4 col = L.rgb_id[j];
      const float w = alpha * T;
      C0 += col.x * w;
      C1 += col.y * w;
      C2 += col.z * w;
      T = test_T;
      last_contributor = s_pos[j] + 1;
    }
  }
  if (c.inside) {
    const size_t pid = c.pix - c.img / 3;
    a.final_T[c.pix] = T;
    a.n_contrib[c.pix] = last_contributor;
    float* out = a.out_color + c.img;
    out[pid] = C0 + T * c.bg0;
    out[c.plane + pid] = C1 + T * c.bg1;
    out[2 * c.plane + pid] = C2 + T * c.bg2;
  }
}

// Backward.  CELLS: the lists are cell lists (multi-view); positions replace
// the per-tile counter and the staging keeps only this tile's instances.
template <int SUB, int BATCH, bool CELLS>
__global__ __launch_bounds__(SUB* SUB) void render_bwd_kernel(RenderArgs a) {
  constexpr int NT = SUB * SUB;
  constexpr int NW = NT / DMGS_WAVE;
  __shared__ SplatLDS<BATCH> L;
  __shared__ uint32_t s_max[NW];
  __shared__ float s_acc[BATCH][DMGS_GACC_USED];
  __shared__ uint32_t s_pos[CELLS ? BATCH : 1];
  __shared__ uint32_t s_wc[NW];
  const TileCtx c = tile_ctx<SUB>(a);
  const float pxf = (float)c.px, pyf = (float)c.py;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

  float T_final = 0.f, g0 = 0.f, g1 = 0.f, g2 = 0.f;
  uint32_t last_contributor = 0;
  if (c.inside) {
    const size_t pid = c.pix - c.img / 3;
    const float* dp = a.dL_dpix + c.img;
    T_final = a.final_T[c.pix];
    last_contributor = a.n_contrib[c.pix];
    g0 = dp[pid];
    g1 = dp[c.plane + pid];
    g2 = dp[2 * c.plane + pid];
  }
  const float bg_dot = c.bg0 * g0 + c.bg1 * g1 + c.bg2 * g2;
  const float ddelx_dx = 0.5f * (float)c.W, ddely_dy = 0.5f * (float)c.H;

  // Entries at list positions >= max(last contributor) of the workgroup contribute nothing.
  uint32_t m = last_contributor;
  for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
  if (lane == 0) s_max[wave] = m;
  __syncthreads();
  uint32_t tile_max = 0;
  for (int w = 0; w < NW; ++w) tile_max = max(tile_max, s_max[w]);
  const int todo_total = min((int)(c.range.y - c.range.x), (int)tile_max);
  const int rounds = (todo_total + BATCH - 1) / BATCH;

  float T = T_final;
  uint32_t contributor = (uint32_t)todo_total;
  float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
  float last_alpha = 0.f, lc0 = 0.f, lc1 = 0.f, lc2 = 0.f;

  const unsigned row_lane = lane & 15;  // lane within its DPP row of 16
  for (int r = 0; r < rounds; ++r) {
    __syncthreads();
    int n;
    if constexpr (CELLS) {
      for (int s = threadIdx.x; s < BATCH; s += NT) {
#pragma unroll
        for (int k = 0; k < DMGS_GACC_USED; ++k) s_acc[s][k] = 0.f;
      }
      n = stage_cell_batch<NT, BATCH, true>(L, s_pos, s_wc, a, c, todo_total - r * BATCH, 0);
    } else {
      for (int s = threadIdx.x; s < BATCH; s += NT) {
        const int pos = r * BATCH + s;
        if (pos < todo_total) stage_splat(L, s, a.splats, a.point_list[c.range.x + todo_total - pos - 1], a.P);
#pragma unroll
        for (int k = 0; k < DMGS_GACC_USED; ++k) s_acc[s][k] = 0.f;
      }
      __syncthreads();
      n = min(BATCH, todo_total - r * BATCH);
    }
    for (int j = 0; j < n; ++j) {
      if constexpr (CELLS) {
        contributor = s_pos[j];
      } else {
        --contributor;
      }
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
          const float4 col = L.rgb_id[j];
          T = T / (1.f - alpha);
          const float dchannel_dcolor = alpha * T;
          acc0 = last_alpha * lc0 + (1.f - last_alpha) * acc0;
          acc1 = last_alpha * lc1 + (1.f - last_alpha) * acc1;
          acc2 = last_alpha * lc2 + (1.f - last_alpha) * acc2;
          lc0 = col.x;
          lc1 = col.y;
          lc2 = col.z;
          float dL_dalpha = (col.x - acc0) * g0 + (col.y - acc1) * g1 + (col.z - acc2) * g2;
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
    // flush the batch: consecutive lanes cover the 9 values of one splat
    for (int i = threadIdx.x; i < n * DMGS_GACC_USED; i += NT) {
      const int slot = i / DMGS_GACC_USED, k = i - slot * DMGS_GACC_USED;
      const float val = s_acc[slot][k];
      if (val != 0.f) {
        const uint32_t gid = __float_as_uint(L.rgb_id[slot].w);
        unsafeAtomicAdd(a.gacc + (size_t)gid * DMGS_GACC_STRIDE + k, val);
      }
    }
  }
}

template <int SUB, int BATCH>
static void launch_fwd_t(const RenderArgs& a, hipStream_t s) {
  constexpr int Q = DMGS_BLOCK_X / SUB;
  if (a.views)
    hipLaunchKernelGGL((render_fwd_cells_kernel<SUB, BATCH>), dim3(a.num_tiles * Q * Q), dim3(SUB * SUB), 0, s, a);
  else
    hipLaunchKernelGGL((render_fwd_kernel<SUB, BATCH>), dim3(a.num_tiles * Q * Q), dim3(SUB * SUB), 0, s, a);
}

template <int SUB, int BATCH>
static void launch_bwd_t(const RenderArgs& a, hipStream_t s) {
  constexpr int Q = DMGS_BLOCK_X / SUB;
  if (a.views)
    hipLaunchKernelGGL((render_bwd_kernel<SUB, BATCH, true>), dim3(a.num_tiles * Q * Q), dim3(SUB * SUB), 0, s, a);
  else
    hipLaunchKernelGGL((render_bwd_kernel<SUB, BATCH, false>), dim3(a.num_tiles * Q * Q), dim3(SUB * SUB), 0, s,
                       a);
}

void launch_render_fwd(const RenderArgs& a, hipStream_t s, int variant) {
  if (a.num_tiles <= 0) return;
  switch (variant) {
    case kRender16x128: launch_fwd_t<16, 128>(a, s); break;
    case kRender8x64: launch_fwd_t<8, 64>(a, s); break;
    case kRender8x128: launch_fwd_t<8, 128>(a, s); break;
    default: launch_fwd_t<16, 256>(a, s); break;
  }
}

void launch_render_bwd(const RenderArgs& a, hipStream_t s, int variant) {
  if (a.num_tiles <= 0) return;
  switch (variant) {
    case kRender16x128: launch_bwd_t<16, 128>(a, s); break;
    case kRender8x64: launch_bwd_t<8, 64>(a, s); break;
    case kRender8x128: launch_bwd_t<8, 128>(a, s); break;
    default: launch_bwd_t<16, 256>(a, s); break;
  }
}

}  // namespace dmgs
