// C-ABI of the batched multi-view pipeline (include/dmgs.h, "batched views").
//
// B views of one scene go through one launch sequence with no host
// synchronisation: fused-activation preprocess of all views, one depth sort
// over the B*P (view, Gaussian) entries, duplication into global tile ids
// (each view owns a contiguous tile range), one stable tile sort, ranges, and
// one blend launch over every tile of every view.  The instance count stays
// on the device; buffers are sized by a capacity fixed at plan time, and the
// count is exposed so the host can detect an overflow after the fact.  The
// sequence contains only kernels (no memset / memcpy nodes), so it can be
// captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>

#include "../../include/dmgs.h"
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {
void set_last_error(const char* msg);
}

static_assert(sizeof(dmgs_view) == sizeof(dmgs::ViewDesc), "dmgs_view / ViewDesc layout");
static_assert(offsetof(dmgs_view, tile_base) == offsetof(dmgs::ViewDesc, tile_base), "dmgs_view layout");
static_assert(offsetof(dmgs_view, pixel_base) == offsetof(dmgs::ViewDesc, pixel_base), "dmgs_view layout");
static_assert(offsetof(dmgs_view, tanfovx) == offsetof(dmgs::ViewDesc, tanfovx), "dmgs_view layout");
static_assert(offsetof(dmgs_view, width) == offsetof(dmgs::ViewDesc, W), "dmgs_view layout");
static_assert(offsetof(dmgs_view, bg) == offsetof(dmgs::ViewDesc, bg), "dmgs_view layout");
static_assert(sizeof(dmgs_batch) == 104 && offsetof(dmgs_batch, capacity) == 56 && offsetof(dmgs_batch, dens) == 96,
              "dmgs_batch layout");
static_assert(offsetof(dmgs_view, cell_base) == offsetof(dmgs::ViewDesc, cell_base), "dmgs_view layout");

namespace {

int bfail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int bfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  dmgs::set_last_error(buf);
  return code;
}

int bcheck(const char* what, hipStream_t s) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(DMGS_EHIP, "%s: %s", what, hipGetErrorString(e));
  (void)s;
  return DMGS_OK;
}

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
template <typename T>
T* at(void* base, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(base) + off); }

int bits_for(uint32_t n) {
  int b = 0;
  while (b < 32 && (n >> b) != 0) ++b;
  return b;
}

// K-independent state of a batch: per-entry splats / keys / order / radii /
// scan, the depth-sort scratch, per-tile ranges and per-pixel blend state.
struct StateLayout {
  size_t splats, keys_a, keys_b, order_a, order_b, rects, rects_sorted, radii, offsets, gscratch, visible,
      status, tile_order, tile_bucket, bucket_hist, tile_work, ranges, final_T, n_contrib, total;
  StateLayout(size_t N, size_t T, size_t X) {
    size_t o = 0;
    splats = o; o += align_up(N * DMGS_SPLAT_F4 * 16);
    keys_a = o; o += align_up(N * 4);
    keys_b = o; o += align_up(N * 4);
    order_a = o; o += align_up(N * 4);
    order_b = o; o += align_up(N * 4);
    rects = o; o += align_up(N * 8);
    rects_sorted = o; o += align_up(N * 8);
    radii = o; o += align_up(N * 4);
    offsets = o; o += align_up(N * 4);
    gscratch = o; o += align_up(std::max(dmgs::radix_scratch_words(N), dmgs::scan_scratch_words(N)) * 4);
    visible = o; o += align_up(4);  // visible entries of the last forward (the depth sort's length)
    status = o; o += align_up(4);
    tile_order = o; o += align_up(T * 4);
    tile_bucket = o; o += align_up(T);
    tile_work = o; o += align_up(T * 4);
    bucket_hist = o; o += 512 * 4;  // directly before the ranges: the duplicate zeroes both in one run
    ranges = o; o += align_up(T * 8);
    final_T = o; o += align_up(X * 4);
    n_contrib = o; o += align_up(X * 4);
    total = o;
  }
};

// Per-instance lists (tile keys and entry ids, ping-pong) and their scratch.
struct ListLayout {
  size_t tkeys_a, tkeys_b, vals_a, vals_b, scratch, chunk, total;
  explicit ListLayout(size_t K) {
    size_t o = 0;
    tkeys_a = o; o += align_up(K * 4);
    tkeys_b = o; o += align_up(K * 4);
    vals_a = o; o += align_up(K * 4);
    vals_b = o; o += align_up(K * 4);
    scratch = o; o += align_up(dmgs::radix_scratch_words(K) * 4);   // sort histograms
    chunk = o; o += align_up(dmgs::duplicate_scratch_words(K) * 4);  // duplicate chunk starts
    total = o;
  }
};

int check_batch(const dmgs_batch* b) {
  if (!b) return bfail(DMGS_EINVAL, "batch is null");
  if (b->num_views <= 0 || b->num_gaussians <= 0 || (int64_t)b->num_views * b->num_gaussians > (int64_t)0x7fffffff)
    return bfail(DMGS_EINVAL, "invalid batch: %d views x %lld Gaussians", b->num_views,
                 (long long)b->num_gaussians);
  // cell ids share a key with the 16-bit tile mask of the cell lists
  if (b->total_cells <= 0 || b->total_cells > b->total_tiles || b->total_cells > (1 << DMGS_CELL_KEY_BITS))
    return bfail(DMGS_EINVAL, "invalid cell total (at most 65536 cells)");
  if (b->capacity < 0 || b->capacity > (int64_t)0xffffffff) return bfail(DMGS_EINVAL, "invalid instance capacity");
  if (b->total_tiles <= 0 || b->total_pixels <= 0) return bfail(DMGS_EINVAL, "invalid view totals");
  if (b->cell_mode != 0 && b->cell_mode != 1) return bfail(DMGS_EINVAL, "cell_mode must be 0 or 1");
  if (b->cell_mode == 1 && (b->num_views > 16 || b->max_view_cells <= 0 || b->max_view_cells > 4096 ||
                            b->max_view_cells > b->total_cells))
    return bfail(DMGS_EINVAL, "cell_mode 1 needs at most 16 views and 1..4096 cells per view");
  if (b->sh_degree < 0 || b->sh_degree > 3) return bfail(DMGS_EINVAL, "sh_degree must be in [0, 3]");
  if (!b->views || !b->state) return bfail(DMGS_EINVAL, "views and state are required");
  if (b->dens && (reinterpret_cast<uintptr_t>(b->dens) & 3u)) return bfail(DMGS_EINVAL, "dens must be 4-byte aligned");
  const auto misaligned = [](const void* p) { return p && (reinterpret_cast<uintptr_t>(p) & 255u) != 0; };
  if (misaligned(b->state) || misaligned(b->lists) || misaligned(b->bwd_ws))
    return bfail(DMGS_EINVAL, "workspaces must be 256-byte aligned");
  return DMGS_OK;
}

struct Ctx {
  size_t N;
  StateLayout S;
  ListLayout L;
  explicit Ctx(const dmgs_batch* b)
      : N((size_t)b->num_views * b->num_gaussians),
        S(N, (size_t)b->total_tiles, (size_t)b->total_pixels),
        L((size_t)b->capacity), mode(b->cell_mode),
        local_bits(b->cell_mode == 1 ? std::max(1, bits_for((uint32_t)(b->max_view_cells - 1))) : 0),
        global_bits(std::max(1, bits_for((uint32_t)(b->total_cells - 1)))) {}
  uint32_t* count(const dmgs_batch* b) const { return at<uint32_t>(b->state, S.offsets) + (N - 1); }
  int mode;
  // key bits of the cell sort: the cell within the view (mode 1) or the
  // global cell id (mode 0; cell ids 0 .. total_cells - 1, so 65536 cells
  // need 16 bits and bit 16, the first tile-mask bit, stays out of the sort)
  int local_bits;
  int global_bits;
  int cell_sort_bits() const { return mode == 1 ? local_bits : global_bits; }
  // a one-pass sort (view-major lists, at most 256 cells per view) leaves the lists in the b buffers
  bool lists_in_b() const { return mode == 1 && local_bits <= 8; }
};

dmgs::MVArgs mv_args(const dmgs_batch* b, const Ctx& c, const float* params) {
  dmgs::MVArgs a{};
  a.P = (int)b->num_gaussians;
  a.B = b->num_views;
  a.sh_degree = b->sh_degree;
  a.scale_modifier = b->scale_modifier;
  a.views = reinterpret_cast<const dmgs::ViewDesc*>(b->views);
  a.params = params;
  a.splats = at<float4>(b->state, c.S.splats);
  a.radii = at<int32_t>(b->state, c.S.radii);
  a.rects = at<uint2>(b->state, c.S.rects);
  a.depth_keys = at<uint32_t>(b->state, c.S.keys_a);  // sorted in place into keys_a / order_a
  a.order = at<uint32_t>(b->state, c.S.order_a);
  a.view_major = b->cell_mode == 1 ? 1 : 0;
  a.gacc = static_cast<float*>(b->bwd_ws);
  a.dens = b->dens;
  return a;
}

dmgs::RenderArgs render_args(const dmgs_batch* b, const Ctx& c) {
  dmgs::RenderArgs r{};
  r.num_tiles = b->total_tiles;
  r.P = (uint32_t)c.N;
  r.K = (uint32_t)b->capacity;
  r.K_dev = c.count(b);
  r.views = reinterpret_cast<const dmgs::ViewDesc*>(b->views);
  r.tile_view = b->tile_view;
  r.ranges = at<uint2>(b->state, c.S.ranges);
  r.point_list = at<uint32_t>(b->lists, c.lists_in_b() ? c.L.vals_b : c.L.vals_a);
  r.list_keys = at<uint32_t>(b->lists, c.lists_in_b() ? c.L.tkeys_b : c.L.tkeys_a);
  r.splats = at<float4>(b->state, c.S.splats);
  r.final_T = at<float>(b->state, c.S.final_T);
  r.n_contrib = at<uint32_t>(b->state, c.S.n_contrib);
  r.gacc = static_cast<float*>(b->bwd_ws);
  r.tile_order = at<uint32_t>(b->state, c.S.tile_order);
  // the strip forward measures each tile's backward list length
  const bool strips_fwd = b->variant == dmgs::kRenderStrips64 || b->variant == dmgs::kRenderStrips128;
  r.tile_work = strips_fwd ? at<uint32_t>(b->state, c.S.tile_work) : nullptr;
  return r;
}

}  // namespace

extern "C" {

int dmgs_batched_prepare_views(dmgs_view* views, int32_t B, int32_t* total_tiles, int32_t* total_cells,
                               int32_t* max_view_cells, int64_t* total_pixels) {
  if (!views || B <= 0 || !total_tiles || !total_cells || !max_view_cells || !total_pixels)
    return bfail(DMGS_EINVAL, "bad arguments");
  int64_t tiles = 0, cells = 0, pixels = 0, max_cells = 0;
  for (int v = 0; v < B; ++v) {
    dmgs_view& d = views[v];
    if (d.width <= 0 || d.height <= 0 || !(d.tanfovx > 0.f) || !(d.tanfovy > 0.f))
      return bfail(DMGS_EINVAL, "view %d: invalid size or field of view", v);
    d.focal_x = (float)d.width / (2.0f * d.tanfovx);
    d.focal_y = (float)d.height / (2.0f * d.tanfovy);
    d.grid_x = (d.width + DMGS_BLOCK_X - 1) / DMGS_BLOCK_X;
    d.grid_y = (d.height + DMGS_BLOCK_Y - 1) / DMGS_BLOCK_Y;
    d.tile_base = (int32_t)tiles;
    d.cell_base = (int32_t)cells;
    d.pixel_base = pixels;
    tiles += (int64_t)d.grid_x * d.grid_y;
    const int E = 1 << DMGS_CELL_SHIFT;
    const int64_t vc = (int64_t)((d.grid_x + E - 1) / E) * ((d.grid_y + E - 1) / E);
    cells += vc;
    max_cells = std::max(max_cells, vc);
    pixels += (int64_t)d.width * d.height;
    if (tiles > (int64_t)0x7fffffff) return bfail(DMGS_EINVAL, "too many tiles");
  }
  *total_tiles = (int32_t)tiles;
  *total_cells = (int32_t)std::min<int64_t>(cells, 0x7fffffff);
  *max_view_cells = (int32_t)std::min<int64_t>(max_cells, 0x7fffffff);
  *total_pixels = pixels;
  return DMGS_OK;
}

int dmgs_batched_tile_view(const dmgs_view* views, int32_t B, int32_t* tile_view, int32_t total_tiles) {
  if (!views || B <= 0 || !tile_view || total_tiles <= 0) return bfail(DMGS_EINVAL, "bad arguments");
  for (int v = 0; v < B; ++v) {
    const int32_t n = views[v].grid_x * views[v].grid_y;
    if (views[v].tile_base < 0 || views[v].tile_base + n > total_tiles)
      return bfail(DMGS_EINVAL, "view %d: tile range outside the total (prepare the views first)", v);
    for (int32_t t = 0; t < n; ++t) tile_view[views[v].tile_base + t] = v;
  }
  return DMGS_OK;
}

size_t dmgs_batched_state_size(int32_t B, int64_t P, int32_t total_tiles, int64_t total_pixels) {
  if (B <= 0 || P <= 0 || total_tiles <= 0 || total_pixels <= 0) return 0;
  return StateLayout((size_t)B * P, (size_t)total_tiles, (size_t)total_pixels).total;
}

size_t dmgs_batched_lists_size(int64_t capacity) {
  if (capacity <= 0) return 0;
  return ListLayout((size_t)capacity).total;
}

size_t dmgs_batched_backward_size(int32_t B, int64_t P) {
  if (B <= 0 || P <= 0) return 0;
  return align_up((size_t)B * P * DMGS_GACC_STRIDE * 4);  // whole 16-byte vectors for the forward's clear
}

int dmgs_batched_offsets(int32_t B, int64_t P, int32_t total_tiles, int64_t total_pixels, int64_t capacity,
                         size_t offsets[9]) {
  if (B <= 0 || P <= 0 || total_tiles <= 0 || total_pixels <= 0 || capacity < 0 || (int64_t)B * P > (int64_t)0x7fffffff)
    return bfail(DMGS_EINVAL, "invalid batch sizes");
  if (!offsets) return bfail(DMGS_EINVAL, "offsets is null");
  const size_t N = (size_t)B * P;
  StateLayout S(N, (size_t)total_tiles, (size_t)total_pixels);
  ListLayout Ls((size_t)capacity);
  offsets[6] = Ls.tkeys_a;  // one-pass view-major sorts leave the lists at tkeys_b / vals_b (see cell_lists)
  offsets[7] = Ls.vals_a;
  offsets[0] = S.radii;                  // (B*P) int32 radii, view-major
  offsets[1] = S.offsets + 4 * (N - 1);  // uint32 instance count of the batch
  offsets[2] = S.ranges;
  offsets[3] = S.final_T;
  offsets[4] = S.n_contrib;
  offsets[5] = S.splats;
  offsets[8] = S.status;
  return DMGS_OK;
}

int dmgs_batched_list_offsets(const dmgs_batch* b, size_t* keys, size_t* ids) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!keys || !ids) return bfail(DMGS_EINVAL, "bad arguments");
  const Ctx c(b);
  *keys = c.lists_in_b() ? c.L.tkeys_b : c.L.tkeys_a;
  *ids = c.lists_in_b() ? c.L.vals_b : c.L.vals_a;
  return DMGS_OK;
}

int dmgs_batched_cell_sort_bits(const dmgs_batch* b) {
  int rc = check_batch(b);
  if (rc) return rc;
  return Ctx(b).cell_sort_bits();
}

int dmgs_batched_forward(const dmgs_batch* b, const float* params, float* images, int32_t stages,
                         dmgs_stream_t stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!params) return bfail(DMGS_EINVAL, "params is required");
  if (stages <= 0 || (stages & ~DMGS_STAGES_ALL)) return bfail(DMGS_EINVAL, "invalid stage mask %d", stages);
  if ((stages & (DMGS_STAGE_BINNING | DMGS_STAGE_BLEND)) && (b->capacity == 0 || !b->lists || !b->tile_view))
    return bfail(DMGS_EINVAL, "binning and blend need lists, a capacity and the tile map");
  if ((stages & DMGS_STAGE_BLEND) && !images) return bfail(DMGS_EINVAL, "images is required for the blend");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const Ctx c(b);
  const size_t N = c.N;
  const dmgs::MVArgs a = mv_args(b, c, params);
  uint32_t* offs = at<uint32_t>(b->state, c.S.offsets);
  uint32_t* status = at<uint32_t>(b->state, c.S.status);
  bool chunks_ready = false;  // the duplicate's chunk table was filled by this call's scan
  if (stages & DMGS_STAGE_GEOMETRY) {
    dmgs::MVArgs ap = a;
    ap.status = status;  // zeroed by the preprocess, set by the depth sort
    dmgs::launch_mv_preprocess(ap, st);
    uint32_t* gs = at<uint32_t>(b->state, c.S.gscratch);
    // the visible entries, in entry order, sorted by their keys; culled entries drop out in the first pass
    uint32_t* V_dev = at<uint32_t>(b->state, c.S.visible);
    dmgs::radix_sort_entry_keys(a.depth_keys, a.order, at<uint32_t>(b->state, c.S.keys_b),
                                at<uint32_t>(b->state, c.S.order_b), N, gs, st, V_dev,
                                b->cell_mode == 1 ? status : nullptr);
    dmgs::launch_gather_rects(a.rects, a.order, at<uint2>(b->state, c.S.rects_sorted), offs, N, st, V_dev);
    if ((stages & DMGS_STAGE_BINNING) && b->capacity > 0 && b->lists) {
      // the binning follows in this call: the scan also fills the duplicate's chunk table
      dmgs::launch_inclusive_scan_chunks(offs, offs, N, gs, at<uint32_t>(b->lists, c.L.chunk),
                                         (uint32_t)b->capacity, st);
      chunks_ready = true;
    } else {
      dmgs::launch_inclusive_scan(offs, offs, N, gs, st);
    }
    if ((rc = bcheck("batch preprocess / depth sort / scan", st))) return rc;
  }
  if (stages & DMGS_STAGE_BINNING) {
    const uint32_t K = (uint32_t)b->capacity;
    dmgs::DupArgs d{};
    d.P = (int)N;
    d.order = a.order;
    d.offsets = offs;
    d.splats = a.splats;
    d.rects = at<uint2>(b->state, c.S.rects_sorted);
    d.views = a.views;
    d.per_view = (uint32_t)b->num_gaussians;
    d.tile_keys = at<uint32_t>(b->lists, c.L.tkeys_a);
    d.values = at<uint32_t>(b->lists, c.L.vals_a);
    d.K = K;
    d.K_dev = c.count(b);
    d.chunk_first = at<uint32_t>(b->lists, c.L.chunk);
    d.chunk_ready = chunks_ready;
    d.view_major = b->cell_mode == 1;
    // the duplicate also zeroes the tile-order histogram and the cell ranges (adjacent)
    d.zero = at<uint32_t>(b->state, c.S.bucket_hist);
    d.zero_n = 512 + 2 * (size_t)b->total_cells;
    // view-major lists: the duplicate also counts the first sort pass's digits
    const int w0 = d.view_major ? dmgs::radix_min_passes_first_width(K, c.local_bits) : -1;
    if (w0 > 0) {
      d.hist = at<uint32_t>(b->lists, c.L.scratch);
      d.hist_bits = w0;
    }
    dmgs::launch_duplicate(d, st);
    uint32_t* kb = at<uint32_t>(b->lists, c.L.tkeys_b);
    uint32_t* vb = at<uint32_t>(b->lists, c.L.vals_b);
    if (d.view_major) {  // sort the cell-within-view bits only; the view order is already there
      const bool in_b = dmgs::radix_sort_pairs_min_passes(d.tile_keys, d.values, kb, vb, K,
                                                          c.local_bits,
                                                          at<uint32_t>(b->lists, c.L.scratch), st, d.K_dev,
                                                          w0 > 0);
      if (in_b != c.lists_in_b()) return bfail(DMGS_EHIP, "internal: sorted list buffer mismatch");
      dmgs::launch_view_cell_ranges(in_b ? kb : d.tile_keys, K, at<uint2>(b->state, c.S.ranges), b->total_cells,
                                    a.views, st, d.K_dev, true);
    } else {
      dmgs::radix_sort_pairs(d.tile_keys, d.values, kb, vb, K, c.global_bits,
                             at<uint32_t>(b->lists, c.L.scratch), st, d.K_dev);
      dmgs::launch_tile_ranges(d.tile_keys, K, at<uint2>(b->state, c.S.ranges), b->total_cells, st, d.K_dev,
                               DMGS_CELL_KEY_MASK);
    }
    // blend dispatch order: tiles with the longest lists first
    dmgs::launch_tile_order(render_args(b, c), at<uint32_t>(b->state, c.S.tile_order),
                            at<uint8_t>(b->state, c.S.tile_bucket), at<uint32_t>(b->state, c.S.bucket_hist), st,
                            true);
    if ((rc = bcheck("batch binning", st))) return rc;
  }
  if (stages & DMGS_STAGE_BLEND) {
    dmgs::RenderArgs r = render_args(b, c);
    r.out_color = images;
    if (b->bwd_ws) {  // the blend work-groups also clear the backward accumulator
      r.clear = static_cast<float4*>(b->bwd_ws);
      r.clear_n = (c.N * DMGS_GACC_STRIDE + 3) / 4;  // inside the 256-byte-aligned workspace
      r.clear_words = at<uint32_t>(b->state, c.S.bucket_hist);  // for the backward's tile order
      r.clear_words_n = 512;
    }
    dmgs::launch_render_fwd(r, st, b->variant);
    if ((rc = bcheck("batch render forward", st))) return rc;
  }
  return DMGS_OK;
}

int dmgs_batched_backward(const dmgs_batch* b, const float* params, const float* dL_dimages, float* grads,
                          int32_t stages, dmgs_stream_t stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!params || !b->bwd_ws || !b->lists || !b->tile_view || b->capacity == 0)
    return bfail(DMGS_EINVAL, "params, lists, a capacity, the tile map and the backward workspace are required");
  if (stages <= 0 || (stages & ~(DMGS_STAGE_GEOMETRY | DMGS_STAGE_BLEND)))
    return bfail(DMGS_EINVAL, "invalid stage mask %d", stages);
  if ((stages & DMGS_STAGE_BLEND) && !dL_dimages) return bfail(DMGS_EINVAL, "dL_dimages is required");
  if ((stages & DMGS_STAGE_GEOMETRY) && !grads) return bfail(DMGS_EINVAL, "grads is required");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const Ctx c(b);
  dmgs::MVArgs a = mv_args(b, c, params);
  if (stages & DMGS_STAGE_BLEND) {  // accumulates into bwd_ws, which holds zeros on entry
    dmgs::RenderArgs r = render_args(b, c);
    r.dL_dpix = dL_dimages;
    // re-order the tiles by the list lengths the strip forward measured
    // (the forward blend, which every backward follows, zeroed the histogram)
    dmgs::launch_tile_order_bwd(r, at<uint32_t>(b->state, c.S.tile_order), at<uint8_t>(b->state, c.S.tile_bucket),
                                at<uint32_t>(b->state, c.S.bucket_hist), st, true);
    dmgs::launch_render_bwd(r, st, b->variant);
    if ((rc = bcheck("batch render backward", st))) return rc;
  }
  if (stages & DMGS_STAGE_GEOMETRY) {
    a.grads = grads;
    dmgs::launch_mv_geometry_bwd(a, st);
    if ((rc = bcheck("batch geometry backward", st))) return rc;
  }
  return DMGS_OK;
}

}  // extern "C"
