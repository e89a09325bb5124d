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
// cannot contribute), transmittance is recomputed (times 1 / (1 - alpha), one
// reciprocal) instead of being stored, and each wave reduces its 9 per-splat
// partial gradients with DPP row sums into a per-batch LDS accumulator; the
// workgroup flushes that accumulator with one global atomic per (splat,
// value) per batch.
//
// Several views in one launch: a global tile id indexes the ranges; the view
// table gives each tile's image size, background and packed output offset.
//
// Footprint skip: staging also stores the bounding box of the splat's
// alpha >= 1/255 region, the ellipse  d^T Q d <= 2 ln(255 o)  (Q the conic),
// widened by one pixel.  A wave whose 64 pixels all lie outside the box
// skips the splat: no pixel of it could pass the alpha test, so skipping
// changes nothing (the splat is still counted in the list positions).
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

// The blend evaluates alpha = o exp(power) as o exp2(p) with p = log2(e) power,
// from the conic scaled once per staged splat: q = log2(e) (-a/2, -b, -c/2).
constexpr float kLog2e = 1.44269504088896341f;

// Three 16-byte fields per staged splat, all indexed alike: one address
// register serves every field of a splat (fixed offsets between them).
template <int BATCH>
struct SplatLDS {
  float4 xym[BATCH];     // x, y, reach mask (bits, reach_mask), list position (bits)
  float4 qo[BATCH];      // scaled conic q, opacity
  float4 rgb_id[BATCH];  // r, g, b, id (bits)
};
template <int BATCH>
__device__ __forceinline__ uint32_t splat_mask(const SplatLDS<BATCH>& L, int j) {
  return __float_as_uint(L.xym[j].z);
}
template <int BATCH>
__device__ __forceinline__ uint32_t splat_pos(const SplatLDS<BATCH>& L, int j) {
  return __float_as_uint(L.xym[j].w);
}

// Which pixel rectangles of the work-group a splat can reach.  A splat passes
// the alpha test only inside the ellipse  d^T Q d <= tau,  tau = 2 ln(255 o)
// (o exp(-d^T Q d / 2) >= 1/255, Q the conic).  Staging tests that ellipse
// against each rectangle of pixel centres a wave (or a strip of the tile)
// covers and keeps one bit per rectangle; a wave skips the splat when its
// bit is clear: no pixel there could pass the alpha test, so results are
// unchanged.  Margins (rectangle +0.01 px, tau x 1.001 + 0.001) keep the test
// conservative against rounding.
struct RowRects {  // n rectangles [x0, x1] x [y0 + k h, y0 + k h + h - 1]
  float x0, x1, y0, h;
  int n;
};

// min over the rectangle of q(p - centre) <= tau, q(d) = a dx^2 + 2 b dx dy + c dy^2
// (convex: the minimum is the centre if it is inside, else on an edge, where
// the one-dimensional minimiser is clamped to the edge).
__device__ __forceinline__ bool ellipse_meets_rect(float cx, float cy, float a, float b, float c, float tau,
                                                   float x0, float x1, float y0, float y1) {
  if (cx >= x0 && cx <= x1 && cy >= y0 && cy <= y1) return true;
  // (an approximate reciprocal only moves the edge point off the exact
  // minimiser, which raises the value found by far less than the margins)
  const float ia = __builtin_amdgcn_rcpf(a), ic = __builtin_amdgcn_rcpf(c);
  float best = INFINITY;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const float dx = (e ? x1 : x0) - cx;
    const float dy = fminf(fmaxf(-b * dx * ic, y0 - cy), y1 - cy);
    best = fminf(best, a * dx * dx + 2.0f * b * dx * dy + c * dy * dy);
  }
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const float dy = (e ? y1 : y0) - cy;
    const float dx = fminf(fmaxf(-b * dy * ia, x0 - cx), x1 - cx);
    best = fminf(best, a * dx * dx + 2.0f * b * dx * dy + c * dy * dy);
  }
  return best <= tau;
}

__device__ __forceinline__ uint32_t reach_mask(float cx, float cy, float a, float b, float c, float o,
                                               const RowRects& rr) {
  const float tau = 2.0f * logf(255.0f * o);
  if (!(tau > 0.0f)) return 0u;  // o < 1/255: alpha never reaches the threshold
  if (!(a > 0.0f && c > 0.0f)) return 0xFu;
  const float t = tau * 1.001f + 0.001f, m = 0.01f;
  uint32_t mask = 0;
  for (int k = 0; k < rr.n; ++k) {
    const float y0 = rr.y0 + (float)k * rr.h;
    if (ellipse_meets_rect(cx, cy, a, b, c, t, rr.x0 - m, rr.x1 + m, y0 - m, y0 + rr.h - 1.0f + m)) mask |= 1u << k;
  }
  return mask;
}

// p = log2(e) power = (q.x dx) dx + dy (q.y dx + q.z dy) for power = -1/2 (a dx^2
// + c dy^2) - b dx dy.  The dx terms depend only on the splat and the pixel
// column, so a lane owning several pixels of one column forms them once
// (power_dx).  The rounding sequence is fixed (explicit products, no compiler
// contraction), so every kernel variant, and the backward's recomputation,
// gets bit-identical values.
struct PowerDx {
  float xx, xy;  // q.x dx dx, q.y dx
};
__device__ __forceinline__ PowerDx power_dx(const float4& q, float dx) {
  return {__fmul_rn(__fmul_rn(q.x, dx), dx), __fmul_rn(q.y, dx)};
}
__device__ __forceinline__ float power_at(const PowerDx& d, float qz, float dy) {
  return fmaf(fmaf(qz, dy, d.xy), dy, d.xx);
}
// exp2 on the hardware exponential (one instruction)
__device__ __forceinline__ float splat_exp2(float p) { return __builtin_amdgcn_exp2f(p); }

#define DMGS_DPP_ADD(v, ctrl) ((v) + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), (ctrl), 0xF, 0xF, true)))

// Wave-uniform test of bit k of a staged mask (read once, scalar branch).
__device__ __forceinline__ bool reaches(uint32_t mask_lds, int k) {
  return (__builtin_amdgcn_readfirstlane(mask_lds) >> k) & 1u;
}

// Row sums of nine per-lane values over each 16-lane DPP row; row lane k < 9
// returns the sum of value k.  Two reduce-scatter steps over lane bits 0 and
// 1 (each lane keeps the half of the values its partner gives up: 3 VALU per
// pair) leave lane m (m = lane & 3) the quad sums of values m and 4 + m;
// value 8 takes plain butterfly steps; rotating the row by 4 and by 8 then
// adds the four quads.  28 VALU instead of 54 for nine full row sums.
__device__ __forceinline__ float row_sums9(const float (&v)[DMGS_GACC_USED], unsigned row_lane) {
  const bool b0 = (row_lane & 1u) != 0u, b1 = (row_lane & 2u) != 0u;
  float p[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {  // lanes l, l ^ 1: keep value 2i + b0
    const float keep = b0 ? v[2 * i + 1] : v[2 * i];
    const float give = b0 ? v[2 * i] : v[2 * i + 1];
    p[i] = keep + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(give), 0xB1, 0xF, 0xF, true));
  }
  float e = DMGS_DPP_ADD(v[8], 0xB1);
  float r[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {  // lanes l, l ^ 2: keep value 4i + 2 b1 + b0
    const float keep = b1 ? p[2 * i + 1] : p[2 * i];
    const float give = b1 ? p[2 * i] : p[2 * i + 1];
    r[i] = keep + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(give), 0x4E, 0xF, 0xF, true));
  }
  e = DMGS_DPP_ADD(e, 0x4E);
#pragma unroll
  for (int i = 0; i < 2; ++i) r[i] = DMGS_DPP_ADD(DMGS_DPP_ADD(r[i], 0x124), 0x128);  // row_ror 4, row_ror 8
  e = DMGS_DPP_ADD(DMGS_DPP_ADD(e, 0x124), 0x128);
  return row_lane < 4u ? r[0] : (row_lane < 8u ? r[1] : e);
}

// Sums over a one-wave work-group of nine per-lane values, through LDS: each
// lane stores its values 0-7 (red[q][lane]); lane (row r, row position p)
// then adds value 2r + (p >> 3) of the eight lanes 8 (p & 7) .. 8 (p & 7) + 7
// (two 16-byte reads, seven adds), and two quad butterflies plus a shift by
// four lanes leave the sum over all 64 lanes in row positions 7 and 15.
// Value 8 takes DPP row sums and two row broadcasts (lane 63 ends with its
// total).  About 17 VALU instead of row_sums9's 28 plus its selects.
struct WaveSum9 {
  int q;       // the value this lane sums in the LDS step (0..7)
  int src;     // the first of its eight source lanes
  bool owner;  // row position 7 or 15: ends with the total of value q
  bool last;   // lane 63: ends with the total of value 8
};
__device__ __forceinline__ WaveSum9 wave_sum9_lane(unsigned lane) {
  const unsigned p = lane & 15u, r = lane >> 4;
  return {(int)(2u * r + (p >> 3)), (int)(8u * (p & 7u)), (p & 7u) == 7u, lane == 63u};
}
// Rows named in row_mask add lane 15 (ctrl 0x142) or lane 31 (0x143) of the
// wave; the other rows add zero.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float add_row_bcast(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
// acc[0..8] receive the totals by plain stores.  Called by the whole wave.
__device__ __forceinline__ void wave_sums9(const float (&v)[DMGS_GACC_USED], float (*red)[DMGS_WAVE],
                                           const WaveSum9& w, unsigned lane, float* acc) {
  __syncthreads();  // the previous call's reads of red are done
#pragma unroll
  for (int q = 0; q < 8; ++q) red[q][lane] = v[q];
  __syncthreads();
  const float4 a = *reinterpret_cast<const float4*>(&red[w.q][w.src]);
  const float4 b = *reinterpret_cast<const float4*>(&red[w.q][w.src + 4]);
  float s = ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w));
  s = DMGS_DPP_ADD(s, 0xB1);   // quad_perm [1,0,3,2]
  s = DMGS_DPP_ADD(s, 0x4E);   // quad_perm [2,3,0,1]: each quad holds its four sources' sum
  s = DMGS_DPP_ADD(s, 0x114);  // row_shr 4: position 7 (15) adds position 3 (11)
  float e = DMGS_DPP_ADD(v[8], 0xB1);
  e = DMGS_DPP_ADD(e, 0x4E);
  e = DMGS_DPP_ADD(DMGS_DPP_ADD(e, 0x124), 0x128);  // row_ror 4, row_ror 8: every lane holds its row's sum
  e = add_row_bcast<0x142, 0xA>(e);                  // rows 1, 3 add rows 0, 2
  e = add_row_bcast<0x143, 0xC>(e);                  // rows 2, 3 add rows 0 + 1
  if (w.owner) acc[w.q] = s;
  if (w.last) acc[8] = e;
}

// Backward state of one pixel, walking its list back to front.  The reference
// keeps per channel ch an accumulator acc_ch (colour behind the current splat)
// and needs only  sum_ch (col_ch - acc_ch) g_ch  (g = dL/dpixel), so one
// recursion on S = sum_ch acc_ch g_ch replaces three:
// S <- alpha' <col', g> + (1 - alpha') S, with alpha', col' the splat behind.
struct BwdPixel {
  float T;    // transmittance in front of the splat last processed
  float S;    // sum_ch acc_ch g_ch
  float pcg;  // alpha' <col', g>: the recursion's next term
  float om;   // 1 - alpha': its factor
};
__device__ __forceinline__ BwdPixel bwd_pixel(float T_final) { return {T_final, 0.f, 0.f, 1.f}; }

// One contributing splat (alpha, colour col) at a pixel with dL/dpixel =
// (g0, g1, g2) and bg_term = -T_final <bg, g>: updates the state and returns
// dL/dalpha; w = alpha T is dL/dcolour_ch / g_ch.
__device__ __forceinline__ float bwd_step(BwdPixel& st, float alpha, const float4& col, float g0, float g1,
                                          float g2, float bg_term, float& w) {
  // each state value is read before it is replaced, so it can be updated in
  // place; the FMA names its destination (S's own register) because the
  // compiler's two-address form would land it in pcg's register and cost two
  // moves per pixel to restore the loop's registers
  asm("v_fma_f32 %0, %1, %0, %2" : "+v"(st.S) : "v"(st.om), "v"(st.pcg));
  st.om = 1.f - alpha;
  const float inv = __builtin_amdgcn_rcpf(st.om);  // T before this splat = T / (1 - alpha)
  st.T *= inv;
  w = alpha * st.T;
  const float cg = fmaf(col.z, g2, fmaf(col.y, g1, col.x * g0));
  st.pcg = alpha * cg;
  return fmaf(bg_term, inv, (cg - st.S) * st.T);
}

// Gradients through G = exp(power), with t = dL/dG G = o G dL/dalpha (the
// reference ignores the 0.99 clamp): dL/d(dx) = -t (a dx + b dy), dL/d(dy) =
// -t (b dx + c dy), dL/d(a, b, c) = -t (dx^2, dx dy, dy^2) / 2, dL/do =
// t / o.  From the scaled conic q = log2(e) (-a/2, -b, -c/2):
// -(a X + b Y) = ln 2 (2 q.x X + q.y Y) and -(b X + c Y) = ln 2 (q.y X + 2 q.z Y).
// The kernels accumulate, per splat, (2 q.x X + q.y Y, q.y X + 2 q.z Y,
// sum t dx^2, sum t dx dy, sum t dy^2, sum G dL/dalpha, sum w g_ch) with
// X = sum t dx, Y = sum t dy; the flush applies the constant factors.
__device__ __forceinline__ float grad_scale(int k, float ddelx_dx, float ddely_dy) {
  constexpr float kLn2 = 0.693147180559945309f;
  return k == 0 ? kLn2 * ddelx_dx : (k == 1 ? kLn2 * ddely_dy : (k < 5 ? -0.5f : 1.0f));
}
__device__ __forceinline__ float dmean_x(const float4& q, float X, float Y) { return fmaf(q.x, X + X, q.y * Y); }
__device__ __forceinline__ float dmean_y(const float4& q, float X, float Y) { return fmaf(q.z, Y + Y, q.y * X); }

template <int BATCH>
__device__ __forceinline__ void stage_splat(SplatLDS<BATCH>& L, int slot, const float4* __restrict__ splats,
                                            uint32_t id, uint32_t P, const RowRects& rr, uint32_t pos) {
  if (id >= P) {  // cannot happen for lists built by the binning step; a zero-opacity entry never contributes
    L.xym[slot] = make_float4(0.f, 0.f, __uint_as_float(0u), __uint_as_float(pos));
    L.qo[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    L.rgb_id[slot] = make_float4(0.f, 0.f, 0.f, __uint_as_float(0u));
    return;
  }
  const float4* sp = splats + (size_t)DMGS_SPLAT_F4 * id;
  const float4 s0 = sp[0];
  const float4 s1 = sp[1];
  const float4 s2 = sp[2];
  constexpr float kHalf = -0.5f * kLog2e;
  const uint32_t mask = reach_mask(s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, rr);
  L.xym[slot] = make_float4(s0.x, s0.y, __uint_as_float(mask), __uint_as_float(pos));
  L.qo[slot] = make_float4(kHalf * s0.z, -kLog2e * s0.w, kHalf * s1.x, s1.y);
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
  uint32_t bit;        // multi-view: this tile's bit in its cell's key masks
  uint2 range;         // single view: the tile's list; multi-view: its cell's list
  RowRects rows;       // the work-group's pixel rectangles: 16x4 strips (SUB 16) or the 8x8 block
};

template <int SUB>
__device__ __forceinline__ TileCtx tile_ctx(const RenderArgs& a) {
  constexpr int Q = DMGS_BLOCK_X / SUB;  // workgroups per tile edge
  TileCtx c;
  const int slot = blockIdx.x / (Q * Q);
  const int q = blockIdx.x - slot * Q * Q;
  c.tile = slot;
  if (a.tile_order) {  // a stale or unset order entry falls back to the identity
    const uint32_t t = a.tile_order[slot];
    if (t < (uint32_t)a.num_tiles) c.tile = (int)t;
  }
  int lt = c.tile, gx = a.grid_x;
  int64_t pbase = 0;
  int list = c.tile;
  c.bit = 0;
  if (a.views) {
    const ViewDesc& v = a.views[a.tile_view[c.tile]];
    lt = c.tile - v.tile_base;
    gx = v.grid_x;
    const int ltx = lt % gx, lty = lt / gx;
    constexpr int E = 1 << DMGS_CELL_SHIFT;
    list = v.cell_base + (lty >> DMGS_CELL_SHIFT) * cells_x(gx) + (ltx >> DMGS_CELL_SHIFT);
    c.bit = 1u << ((lty & (E - 1)) * E + (ltx & (E - 1)));
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
  const int ox = tx * DMGS_BLOCK_X + (q % Q) * SUB, oy = ty * DMGS_BLOCK_Y + (q / Q) * SUB;
  c.px = ox + lx;
  c.py = oy + ly;
  // rectangle k: the rows one wave covers in a 256-lane work-group (= strip k
  // of a one-wave-per-tile work-group), or the whole 8x8 block
  c.rows.x0 = (float)ox;
  c.rows.x1 = (float)(ox + SUB - 1);
  c.rows.y0 = (float)oy;
  c.rows.h = SUB == 16 ? 4.0f : 8.0f;
  c.rows.n = SUB == 16 ? 4 : 1;
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

// A forward work-group's share of clearing the backward accumulator: run
// after its blend work, so the stores delay no load of the blend.
__device__ __forceinline__ void clear_share(const RenderArgs& a) {
  if (a.clear_words && blockIdx.x == 0)
    for (int i = threadIdx.x; i < a.clear_words_n; i += blockDim.x) a.clear_words[i] = 0u;
  if (!a.clear) return;
  const size_t per = (a.clear_n + gridDim.x - 1) / gridDim.x;
  const size_t lo = (size_t)blockIdx.x * per, hi = min(a.clear_n, lo + per);
  for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) a.clear[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Stage the next instances of this tile from its cell list (multi-view):
// visit list positions from `cursor` — forward: cursor, cursor+1, ... below
// limit; backward: cursor-1, cursor-2, ... down to 0 — min(NT, BATCH) at a
// time and keep those whose key holds this tile's bit, in visit order, while
// another chunk surely fits in the BATCH slots; then stage the kept records
// in one pass, one lane per record (the footprint test then runs once per
// NT records instead of once per chunk visited).  Advances cursor; returns
// the number staged (workgroup-uniform); ends with a barrier.  Each staged
// record carries its list position.
template <int NT, bool BACKWARD, int BATCH>
__device__ __forceinline__ int stage_cell_batch(SplatLDS<BATCH>& L, uint32_t* s_wc,
                                                const RenderArgs& a, const TileCtx& c, int& cursor, int limit) {
  constexpr int NW = NT / DMGS_WAVE;
  constexpr int CH = NT < BATCH ? NT : BATCH;  // positions per chunk
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  int run = 0;
  do {
    const int s = (int)threadIdx.x;
    const int q = BACKWARD ? cursor - 1 - s : cursor + s;
    const bool valid = s < CH && (BACKWARD ? q >= 0 : q < limit);
    // The entry id is requested with the key, both without a branch (lanes
    // past the list read its first entry and ignore it): one memory round
    // trip, not two.
    const uint32_t at = c.range.x + (valid ? (uint32_t)q : 0u);
    const uint32_t key = a.list_keys[at];
    const uint32_t id = a.point_list[at];
    const bool member = valid && ((key >> DMGS_CELL_KEY_BITS) & c.bit);
    const uint64_t bal = __ballot(member);
    int before = run, total;
    if constexpr (NW == 1) {
      total = (int)__popcll(bal);
    } else {
      if (lane == 0) s_wc[wave] = (uint32_t)__popcll(bal);
      __syncthreads();
      total = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const int cnt = (int)s_wc[w];
        before += (w < (int)wave) ? cnt : 0;
        total += cnt;
      }
    }
    if (member) {
      const int slot = before + (int)__popcll(bal & lt);
      // id and position wait in the record for the staging pass
      L.xym[slot] = make_float4(0.f, 0.f, __uint_as_float(id), __uint_as_float((uint32_t)q));
    }
    run += total;
    cursor = BACKWARD ? cursor - CH : cursor + CH;
    if constexpr (NW > 1) __syncthreads();  // s_wc is rewritten by the next chunk
  } while (run + CH <= BATCH && (BACKWARD ? cursor > 0 : cursor < limit));
  __syncthreads();
  for (int s = (int)threadIdx.x; s < run; s += NT) {
    const float4 t = L.xym[s];
    stage_splat(L, s, a.splats, __float_as_uint(t.z), a.P, c.rows, __float_as_uint(t.w));
  }
  __syncthreads();
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
  uint32_t last_contributor = 0;
  const int my_rect = SUB == 16 ? (int)(threadIdx.x >> 6) : 0;  // this wave's rectangle in the reach masks

  for (int r = 0; r < rounds; ++r) {
    if (__syncthreads_count(done) == NT) break;
    for (int s = threadIdx.x; s < BATCH; s += NT) {
      const int pos = r * BATCH + s;
      if (pos < todo_total) stage_splat(L, s, a.splats, a.point_list[c.range.x + pos], a.P, c.rows, (uint32_t)pos);
    }
    __syncthreads();
    const int n = min(BATCH, todo_total - r * BATCH);
    for (int j = 0; !done && j < n; ++j) {
      if (!reaches(splat_mask(L, j), my_rect)) continue;
      const uint32_t contributor = (uint32_t)(r * BATCH + j + 1);
      const float2 xy = make_float2(L.xym[j].x, L.xym[j].y);
      const float4 q = L.qo[j];
      const float dx = xy.x - pxf, dy = xy.y - pyf;
      const float p = power_at(power_dx(q, dx), q.z, dy);
      if (p > 0.0f) continue;
      const float alpha = fminf(0.99f, q.w * splat_exp2(p));
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
  clear_share(a);
}

// Forward over cell lists (multi-view).  Same per-pixel arithmetic and order
// as render_fwd_kernel; the per-pixel "last contributor" is one past the
// list position of the last contributing instance.
template <int SUB, int BATCH>
__global__ __launch_bounds__(SUB* SUB) void render_fwd_cells_kernel(RenderArgs a) {
  constexpr int NT = SUB * SUB;
  __shared__ SplatLDS<BATCH> L;
  __shared__ uint32_t s_wc[NT / DMGS_WAVE];
  const TileCtx c = tile_ctx<SUB>(a);
  const float pxf = (float)c.px, pyf = (float)c.py;
  const int len = (int)(c.range.y - c.range.x);

  bool done = !c.inside;
  float T = 1.0f;
  float C0 = 0.f, C1 = 0.f, C2 = 0.f;
  uint32_t last_contributor = 0;
  const int my_rect = SUB == 16 ? (int)(threadIdx.x >> 6) : 0;  // this wave's rectangle in the reach masks

  int cursor = 0;
  while (cursor < len) {
    if (__syncthreads_count(done) == NT) break;
    const int n = stage_cell_batch<NT, false>(L, s_wc, a, c, cursor, len);
    for (int j = 0; !done && j < n; ++j) {
      if (!reaches(splat_mask(L, j), my_rect)) continue;
      const float2 xy = make_float2(L.xym[j].x, L.xym[j].y);
      const float4 q = L.qo[j];
      const float dx = xy.x - pxf, dy = xy.y - pyf;
      const float p = power_at(power_dx(q, dx), q.z, dy);
      if (p > 0.0f) continue;
      const float alpha = fminf(0.99f, q.w * splat_exp2(p));
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
      last_contributor = splat_pos(L, j) + 1;
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
  clear_share(a);
}

// Forward with one wave per 16x16 tile (the layout of the strip backward):
// lane (lx, ly) owns the pixels (lx, ly + 4k), one per 16x4 strip.  Each
// staged splat is read from LDS once per wave for four pixels; a strip is
// skipped when the splat's footprint misses it or all its pixels are done.
// Per pixel the arithmetic and order are render_fwd_kernel's.
template <int BATCH, bool CELLS>
__global__ __launch_bounds__(64) void render_fwd_strips_kernel(RenderArgs a) {
  constexpr int NT = 64;
  __shared__ SplatLDS<BATCH> L;
  __shared__ uint32_t s_wc[1];
  const TileCtx c = tile_ctx<16>(a);
  const float pxf = (float)c.px;
  const int len = (int)(c.range.y - c.range.x);
  float pyf[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) pyf[k] = (float)(c.py + 4 * k);
  // T > 0 while a pixel blends; a finished pixel (saturated, or outside the
  // image) keeps -T, so one compare tells both
  float T[4], C0[4], C1[4], C2[4];
  uint32_t last_contributor[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    T[k] = (c.px < c.W && c.py + 4 * k < c.H) ? 1.0f : -1.0f;
    C0[k] = C1[k] = C2[k] = 0.f;
    last_contributor[k] = 0;
  }
  int cursor = 0;
  while (cursor < len) {
    if (__syncthreads_count(T[0] < 0.f && T[1] < 0.f && T[2] < 0.f && T[3] < 0.f) == NT) break;
    const int base = cursor;
    int n;
    if constexpr (CELLS) {
      n = stage_cell_batch<NT, false>(L, s_wc, a, c, cursor, len);
    } else {
      for (int s = threadIdx.x; s < BATCH; s += NT) {
        const int pos = base + s;
        if (pos < len) stage_splat(L, s, a.splats, a.point_list[c.range.x + pos], a.P, c.rows, (uint32_t)pos);
      }
      __syncthreads();
      n = min(BATCH, len - base);
      cursor += n;
    }
    uint32_t mask_next = n > 0 ? splat_mask(L, 0) : 0u;  // read one splat ahead
    for (int j = 0; j < n; ++j) {
      const uint32_t reach = __builtin_amdgcn_readfirstlane(mask_next);
      if (j + 1 < n) mask_next = splat_mask(L, j + 1);
      if (reach == 0u) continue;
      const float2 xy = make_float2(L.xym[j].x, L.xym[j].y);
      const float4 q = L.qo[j];
      uint32_t position;
      if constexpr (CELLS) {
        position = splat_pos(L, j);
      } else {
        position = (uint32_t)(base + j);
      }
      const PowerDx pd = power_dx(q, xy.x - pxf);  // shared by the lane's four pixels (one column)
      const float4 col = L.rgb_id[j];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!((reach >> k) & 1u) || __ballot(T[k] > 0.f) == 0) continue;  // wave-uniform
        // the per-pixel tests become selects: no branch per pixel and splat; a
        // skipped splat adds colour x 0 (exact) and keeps T
        const float p = power_at(pd, q.z, xy.y - pyf[k]);
        const float alpha = fminf(0.99f, q.w * splat_exp2(p));
        const float test_T = T[k] * (1.0f - alpha);
        const bool ok = T[k] > 0.f && p <= 0.0f && alpha >= 1.0f / 255.0f;
        const bool stop = ok && test_T < 0.0001f;
        const bool take = ok && !stop;
        const float w = take ? alpha * T[k] : 0.0f;
        C0[k] = fmaf(col.x, w, C0[k]);
        C1[k] = fmaf(col.y, w, C1[k]);
        C2[k] = fmaf(col.z, w, C2[k]);
        T[k] = take ? test_T : (stop ? -T[k] : T[k]);
        last_contributor[k] = take ? position + 1 : last_contributor[k];
      }
    }
  }
  if (a.tile_work) {  // the backward walks list positions below the tile's largest last contributor
    uint32_t m = max(max(last_contributor[0], last_contributor[1]), max(last_contributor[2], last_contributor[3]));
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
    if (threadIdx.x == 0) a.tile_work[c.tile] = m;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!(c.px < c.W && c.py + 4 * k < c.H)) continue;
    const size_t pix = c.pix + (size_t)(4 * k) * c.W;
    const size_t pid = pix - c.img / 3;
    const float Tk = fabsf(T[k]);
    a.final_T[pix] = Tk;
    a.n_contrib[pix] = last_contributor[k];
    float* out = a.out_color + c.img;
    out[pid] = C0[k] + Tk * c.bg0;
    out[c.plane + pid] = C1[k] + Tk * c.bg1;
    out[2 * c.plane + pid] = C2[k] + Tk * c.bg2;
  }
  clear_share(a);
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
  const float bg_term = -T_final * (c.bg0 * g0 + c.bg1 * g1 + c.bg2 * g2);  // -T_final <bg, dL/dpixel>
  const float ddelx_dx = 0.5f * (float)c.W, ddely_dy = 0.5f * (float)c.H;  // d pixel / d NDC

  // Entries at list positions >= max(last contributor) of the workgroup contribute nothing.
  uint32_t m = last_contributor;
  for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
  if (lane == 0) s_max[wave] = m;
  __syncthreads();
  uint32_t tile_max = 0;
  for (int w = 0; w < NW; ++w) tile_max = max(tile_max, s_max[w]);
  const int todo_total = min((int)(c.range.y - c.range.x), (int)tile_max);

  BwdPixel st = bwd_pixel(T_final);
  const unsigned row_lane = lane & 15;  // lane within its DPP row of 16
  const int my_rect = SUB == 16 ? (int)(threadIdx.x >> 6) : 0;  // this wave's rectangle in the reach masks
  int cursor = todo_total;  // list positions below cursor are still to do
  while (cursor > 0) {
    __syncthreads();
    const int top = cursor;
    int n;
    if constexpr (CELLS) {
      for (int s = threadIdx.x; s < BATCH; s += NT) {
#pragma unroll
        for (int k = 0; k < DMGS_GACC_USED; ++k) s_acc[s][k] = 0.f;
      }
      n = stage_cell_batch<NT, true>(L, s_wc, a, c, cursor, 0);
    } else {
      for (int s = threadIdx.x; s < BATCH; s += NT) {
        const int pos = top - 1 - s;
        if (pos >= 0) stage_splat(L, s, a.splats, a.point_list[c.range.x + pos], a.P, c.rows, (uint32_t)pos);
#pragma unroll
        for (int k = 0; k < DMGS_GACC_USED; ++k) s_acc[s][k] = 0.f;
      }
      __syncthreads();
      n = min(BATCH, top);
      cursor -= n;
    }
    for (int j = 0; j < n; ++j) {
      if (!reaches(splat_mask(L, j), my_rect)) continue;  // no pixel of this wave can pass the alpha test
      uint32_t contributor;
      if constexpr (CELLS) {
        contributor = splat_pos(L, j);
      } else {
        contributor = (uint32_t)(top - 1 - j);
      }
      const float2 xy = make_float2(L.xym[j].x, L.xym[j].y);
      const float4 q = L.qo[j];
      const float dx = xy.x - pxf, dy = xy.y - pyf;
      const float p = power_at(power_dx(q, dx), q.z, dy);
      const float G = splat_exp2(p);
      const float ga = q.w * G;  // o G: alpha before the 0.99 clamp
      const float alpha = fminf(0.99f, ga);
      const bool contrib = contributor < last_contributor && p <= 0.0f && alpha >= 1.0f / 255.0f;
      if (__ballot(contrib) == 0) continue;  // wave-uniform
      float v[DMGS_GACC_USED];
      if (contrib) {
        float w;
        const float dL_dalpha = bwd_step(st, alpha, L.rgb_id[j], g0, g1, g2, bg_term, w);
        const float t = ga * dL_dalpha;
        const float tdx = t * dx, tdy = t * dy;
        v[0] = dmean_x(q, tdx, tdy);
        v[1] = dmean_y(q, tdx, tdy);
        v[2] = tdx * dx;
        v[3] = tdx * dy;
        v[4] = tdy * dy;
        v[5] = G * dL_dalpha;
        v[6] = w * g0;
        v[7] = w * g1;
        v[8] = w * g2;
      } else {
#pragma unroll
        for (int k = 0; k < DMGS_GACC_USED; ++k) v[k] = 0.f;
      }
      // sum over each 16-lane row, then lane k of every row adds value k into
      // the batch accumulator in LDS (one ds_add with 4 x 9 active lanes)
      const float mine = row_sums9(v, row_lane);
      if (row_lane < DMGS_GACC_USED) atomicAdd(&s_acc[j][row_lane], mine);
    }
    __syncthreads();
    // flush the batch: consecutive lanes cover the 9 values of one splat
    for (int i = threadIdx.x; i < n * DMGS_GACC_USED; i += NT) {
      const int slot = i / DMGS_GACC_USED, k = i - slot * DMGS_GACC_USED;
      const float val = s_acc[slot][k] * grad_scale(k, ddelx_dx, ddely_dy);
      if (val != 0.f) {
        const uint32_t gid = __float_as_uint(L.rgb_id[slot].w);
        unsafeAtomicAdd(a.gacc + gacc_index(a.P, gid, k), val);
      }
    }
  }
}

// Backward with one wave per 16x16 tile: lane (lx, ly) owns the pixels
// (lx, ly + 4k), k = 0..3, i.e. one pixel in each of four 16x4 strips.  The
// footprint test runs per strip, each pixel is evaluated exactly as in
// render_bwd_kernel, and a lane first sums its four pixels' partial
// gradients, so each (tile, splat) pays for one wave reduction instead of
// four.
template <int BATCH, bool CELLS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void render_bwd_strips_kernel(RenderArgs a) {
  constexpr int NT = 64;
  __shared__ SplatLDS<BATCH> L;
  // Two accumulator sets: a batch's global atomics are issued after the next
  // batch has been staged, so no staging load waits behind them (the memory
  // counter is in order) and they complete while that batch is computed.
  __shared__ __align__(16) float s_acc[2][BATCH][DMGS_GACC_USED];
  __shared__ uint32_t s_gid[2][BATCH];
  __shared__ uint32_t s_wc[1];
  static_assert(BATCH * DMGS_GACC_USED >= 8 * DMGS_WAVE && (BATCH * DMGS_GACC_USED) % 4 == 0,
                "one accumulator set holds a splat's 8 x 64 per-lane values");
  const TileCtx c = tile_ctx<16>(a);  // pixel k = 0 of this lane: (lx, ly), ly in 0..3
  const unsigned lane = threadIdx.x & 63;
  const WaveSum9 wsum = wave_sum9_lane(lane);
  const float pxf = (float)c.px;
  const float ddelx_dx = 0.5f * (float)c.W, ddely_dy = 0.5f * (float)c.H;  // d pixel / d NDC
  float pyf[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) pyf[k] = (float)(c.py + 4 * k);

  float g0[4], g1[4], g2[4], bg_term[4];
  BwdPixel st[4];
  uint32_t last_contributor[4];
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int py = c.py + 4 * k;
    const bool inside = c.px < c.W && py < c.H;
    float T_final = 0.f;
    g0[k] = g1[k] = g2[k] = 0.f;
    last_contributor[k] = 0;
    if (inside) {
      const size_t pix = c.pix + (size_t)(4 * k) * c.W;
      const size_t pid = pix - c.img / 3;
      const float* dp = a.dL_dpix + c.img;
      T_final = a.final_T[pix];
      last_contributor[k] = a.n_contrib[pix];
      g0[k] = dp[pid];
      g1[k] = dp[c.plane + pid];
      g2[k] = dp[2 * c.plane + pid];
    }
    bg_term[k] = -T_final * (c.bg0 * g0[k] + c.bg1 * g1[k] + c.bg2 * g2[k]);
    st[k] = bwd_pixel(T_final);
  }
  // per strip, the largest last contributor of its 64 pixels (wave-uniform):
  // list positions at or past it cannot contribute to the strip
  uint32_t strip_max[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint32_t x = last_contributor[k];
    for (int off = 32; off > 0; off >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, off, 64));
    strip_max[k] = __builtin_amdgcn_readfirstlane(x);
    m = max(m, strip_max[k]);
  }
  const int todo_total = min((int)(c.range.y - c.range.x), (int)m);

  // one batch's accumulated sums -> one global atomic per (splat, value)
  auto flush = [&](int buf, int cnt) {
    for (int i = threadIdx.x; i < cnt * DMGS_GACC_USED; i += NT) {
      const int slot = i / DMGS_GACC_USED, q = i - slot * DMGS_GACC_USED;
      const float val = s_acc[buf][slot][q] * grad_scale(q, ddelx_dx, ddely_dy);
      if (val != 0.f) unsafeAtomicAdd(a.gacc + gacc_index(a.P, s_gid[buf][slot], q), val);
    }
  };
  int buf = 0, prev_n = 0;
  int cursor = todo_total;  // list positions below cursor are still to do
  while (cursor > 0) {
    __syncthreads();
    const int top = cursor;
    int n;
    for (int s = threadIdx.x; s < BATCH; s += NT) {
#pragma unroll
      for (int k = 0; k < DMGS_GACC_USED; ++k) s_acc[buf][s][k] = 0.f;
    }
    if constexpr (CELLS) {
      n = stage_cell_batch<NT, true>(L, s_wc, a, c, cursor, 0);
    } else {
      for (int s = threadIdx.x; s < BATCH; s += NT) {
        const int pos = top - 1 - s;
        if (pos >= 0) stage_splat(L, s, a.splats, a.point_list[c.range.x + pos], a.P, c.rows, (uint32_t)pos);
      }
      __syncthreads();
      n = min(BATCH, top);
      cursor -= n;
    }
    for (int s = threadIdx.x; s < n; s += NT) s_gid[buf][s] = __float_as_uint(L.rgb_id[s].w);
    // Every staging load has landed here; saying so keeps the compiler from
    // waiting later (behind the atomics below) for a load it cannot rule out.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    flush(buf ^ 1, prev_n);  // the previous batch's atomics overlap this batch
    __syncthreads();
    // the flushed set is free until the next batch zeroes it: it holds the
    // per-lane values of the wave sums (wave_sums9) meanwhile
    float(*const s_red)[DMGS_WAVE] = reinterpret_cast<float(*)[DMGS_WAVE]>(&s_acc[buf ^ 1][0][0]);
    uint32_t mask_next = n > 0 ? splat_mask(L, 0) : 0u;  // read one splat ahead
    for (int j = 0; j < n; ++j) {
      const uint32_t reach = __builtin_amdgcn_readfirstlane(mask_next);
      if (j + 1 < n) mask_next = splat_mask(L, j + 1);
      if (reach == 0u) continue;
      uint32_t contributor;
      if constexpr (CELLS) {
        contributor = __builtin_amdgcn_readfirstlane(splat_pos(L, j));
      } else {
        contributor = (uint32_t)(top - 1 - j);
      }
      uint32_t live = reach;  // strips the splat reaches and that still have pixels behind it
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (contributor >= strip_max[k]) live &= ~(1u << k);
      if (live == 0u) continue;
      const float2 xy = make_float2(L.xym[j].x, L.xym[j].y);
      const float4 q = L.qo[j];
      const float4 col = L.rgb_id[j];
      const float dx = xy.x - pxf;
      const PowerDx pd = power_dx(q, dx);  // shared by the lane's four pixels (one column)
      // per-lane sums over its pixels: sum t, sum t dy, sum t dy^2 (dx is the
      // same for the four pixels, so its powers multiply in once per splat)
      float s_t = 0.f, s_tdy = 0.f, s_tdy2 = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
      bool any = false;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!((live >> k) & 1u)) continue;  // wave-uniform
        const float dy = xy.y - pyf[k];
        const float p = power_at(pd, q.z, dy);
        const float G = splat_exp2(p);
        const float ga = q.w * G;  // o G: alpha before the 0.99 clamp
        const float alpha = fminf(0.99f, ga);
        if (contributor < last_contributor[k] && p <= 0.0f && alpha >= 1.0f / 255.0f) {
          any = true;
          float w;
          const float t = ga * bwd_step(st[k], alpha, col, g0[k], g1[k], g2[k], bg_term[k], w);
          c0 = fmaf(w, g0[k], c0);
          c1 = fmaf(w, g1[k], c1);
          c2 = fmaf(w, g2[k], c2);
          s_t += t;
          const float tdy = t * dy;
          s_tdy += tdy;
          s_tdy2 = fmaf(tdy, dy, s_tdy2);
        }
      }
      if (__ballot(any) == 0) continue;
      const float s_tdx = s_t * dx;
      const float v[DMGS_GACC_USED] = {dmean_x(q, s_tdx, s_tdy), dmean_y(q, s_tdx, s_tdy), s_tdx * dx, s_tdy * dx,
                                       s_tdy2, s_t * __builtin_amdgcn_rcpf(q.w), c0, c1, c2};
      wave_sums9(v, s_red, wsum, lane, s_acc[buf][j]);
    }
    prev_n = n;
    buf ^= 1;
  }
  __syncthreads();
  flush(buf ^ 1, prev_n);
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

template <int BATCH>
static void launch_fwd_strips_t(const RenderArgs& a, hipStream_t s) {
  if (a.views)
    hipLaunchKernelGGL((render_fwd_strips_kernel<BATCH, true>), dim3(a.num_tiles), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((render_fwd_strips_kernel<BATCH, false>), dim3(a.num_tiles), dim3(64), 0, s, a);
}

template <int BATCH>
static void launch_bwd_strips_t(const RenderArgs& a, hipStream_t s) {
  if (a.views)
    hipLaunchKernelGGL((render_bwd_strips_kernel<BATCH, true>), dim3(a.num_tiles), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((render_bwd_strips_kernel<BATCH, false>), dim3(a.num_tiles), dim3(64), 0, s, a);
}

// Longest-first scheduling of the blend: a tile's cost is about the length
// of its binning cell's list, so tiles are bucketed by log2 of that length
// (8 buckets per octave, 256 in all) and emitted heaviest bucket first.  Only
// the dispatch order changes; every tile's result is the same in any order.
__device__ __forceinline__ uint32_t tile_cost_bucket(const RenderArgs& a, int t, uint32_t K, bool use_work) {
  const ViewDesc& v = a.views[a.tile_view[t]];
  const int lt = t - v.tile_base, gx = v.grid_x;
  const int ltx = lt % gx, lty = lt / gx;
  const int cell = v.cell_base + (lty >> DMGS_CELL_SHIFT) * cells_x(gx) + (ltx >> DMGS_CELL_SHIFT);
  const uint2 r = a.ranges[cell];
  const uint32_t hi = min(r.y, K), lo = min(r.x, hi);
  uint32_t len = hi - lo;
  if (use_work) len = min(len, a.tile_work[t]);
  if (len == 0) return 255u;
  return 254u - min(254u, (uint32_t)(8.0f * __log2f((float)len)));
}

// Pass 1 (all tiles in parallel): bucket per tile, global bucket histogram.
__global__ __launch_bounds__(256) void tile_bucket_kernel(RenderArgs a, uint8_t* __restrict__ bucket,
                                                          uint32_t* __restrict__ hist, int use_work) {
  __shared__ uint32_t cnt[256];
  uint32_t K = a.K;
  if (a.K_dev) K = min(K, *a.K_dev);
  cnt[threadIdx.x] = 0u;
  __syncthreads();
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < a.num_tiles) {
    const uint32_t b = tile_cost_bucket(a, t, K, use_work != 0);
    bucket[t] = (uint8_t)b;
    atomicAdd(&cnt[b], 1u);
  }
  __syncthreads();
  if (cnt[threadIdx.x]) atomicAdd(&hist[threadIdx.x], cnt[threadIdx.x]);
}

// Pass 2 (all tiles in parallel): every work-group scans the 256-bucket
// histogram (hist[0..255]), reserves a run per bucket for its tiles from the
// cursors (hist[256..511], zero on entry) and writes its tiles there.
__global__ __launch_bounds__(256) void tile_order_kernel(const uint8_t* __restrict__ bucket, uint32_t* __restrict__ hist,
                                                         int num_tiles, uint32_t* __restrict__ order) {
  __shared__ uint32_t start[256], cnt[256], base[256];
  __shared__ uint32_t s_wave[4];
  const int t = blockIdx.x * 256 + threadIdx.x;
  const uint32_t h = hist[threadIdx.x];
  uint32_t total;
  start[threadIdx.x] = block_exclusive_scan<4>(h, s_wave, total);  // 256-lane work-group
  cnt[threadIdx.x] = 0u;
  __syncthreads();
  uint32_t b = 0, rank = 0;
  if (t < num_tiles) {
    b = bucket[t];
    rank = atomicAdd(&cnt[b], 1u);
  }
  __syncthreads();
  if (cnt[threadIdx.x]) base[threadIdx.x] = start[threadIdx.x] + atomicAdd(&hist[256 + threadIdx.x], cnt[threadIdx.x]);
  __syncthreads();
  if (t < num_tiles) order[base[b] + rank] = (uint32_t)t;
}

static void tile_order_impl(const RenderArgs& a, uint32_t* order, uint8_t* bucket, uint32_t* hist, bool use_work,
                            hipStream_t s, bool hist_zeroed) {
  if (a.num_tiles <= 0 || !a.views) return;
  const unsigned blocks = (unsigned)((a.num_tiles + 255) / 256);
  if (!hist_zeroed) launch_fill_u32(hist, 512, 0u, s);
  hipLaunchKernelGGL(tile_bucket_kernel, dim3(blocks), dim3(256), 0, s, a, bucket, hist, use_work ? 1 : 0);
  hipLaunchKernelGGL(tile_order_kernel, dim3(blocks), dim3(256), 0, s, (const uint8_t*)bucket, hist, a.num_tiles,
                     order);
}

void launch_tile_order(const RenderArgs& a, uint32_t* order, uint8_t* bucket, uint32_t* hist, hipStream_t s,
                       bool hist_zeroed) {
  tile_order_impl(a, order, bucket, hist, false, s, hist_zeroed);
}

void launch_tile_order_bwd(const RenderArgs& a, uint32_t* order, uint8_t* bucket, uint32_t* hist, hipStream_t s,
                           bool hist_zeroed) {
  if (a.tile_work) tile_order_impl(a, order, bucket, hist, true, s, hist_zeroed);
}

void launch_render_fwd(const RenderArgs& a, hipStream_t s, int variant) {
  if (a.num_tiles <= 0) return;
  switch (variant) {
    case kRender16x128: launch_fwd_t<16, 128>(a, s); break;
    case kRenderStrips64: launch_fwd_strips_t<64>(a, s); break;
    case kRenderStrips128: launch_fwd_strips_t<128>(a, s); break;
    case kRender8x64: launch_fwd_t<8, 64>(a, s); break;
    case kRender8x128: launch_fwd_t<8, 128>(a, s); break;
    default: launch_fwd_t<16, 256>(a, s); break;
  }
}

void launch_render_bwd(const RenderArgs& a, hipStream_t s, int variant) {
  if (a.num_tiles <= 0) return;
  switch (variant) {
    case kRenderStrips64:
    case kRenderMixed: launch_bwd_strips_t<64>(a, s); break;
    case kRenderStrips128: launch_bwd_strips_t<128>(a, s); break;
    case kRender16x128: launch_bwd_t<16, 128>(a, s); break;
    case kRender8x64: launch_bwd_t<8, 64>(a, s); break;
    case kRender8x128: launch_bwd_t<8, 128>(a, s); break;
    default: launch_bwd_t<16, 256>(a, s); break;
  }
}

}  // namespace dmgs
