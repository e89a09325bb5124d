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

struct BatchLayout {
  size_t splats, keys_a, keys_b, order_a, order_b, tiles, radii, offsets, gscratch;
  size_t tkeys_a, tkeys_b, vals_a, vals_b, bscratch;
  size_t ranges, final_T, n_contrib, gacc, total;
  BatchLayout(size_t N, size_t K, size_t T, size_t X) {
    size_t o = 0;
    splats = o; o += align_up(N * DMGS_SPLAT_F4 * 16);
    keys_a = o; o += align_up(N * 4);
    keys_b = o; o += align_up(N * 4);
    order_a = o; o += align_up(N * 4);
    order_b = o; o += align_up(N * 4);
    tiles = o; o += align_up(N * 4);
    radii = o; o += align_up(N * 4);
    offsets = o; o += align_up(N * 4);
    gscratch = o; o += align_up(std::max(dmgs::radix_scratch_words(N), dmgs::scan_scratch_words(N)) * 4);
    tkeys_a = o; o += align_up(K * 4);
    tkeys_b = o; o += align_up(K * 4);
    vals_a = o; o += align_up(K * 4);
    vals_b = o; o += align_up(K * 4);
    bscratch = o; o += align_up(std::max(dmgs::radix_scratch_words(K), dmgs::duplicate_scratch_words(K)) * 4);
    ranges = o; o += align_up(T * 8);
    final_T = o; o += align_up(X * 4);
    n_contrib = o; o += align_up(X * 4);
    gacc = o; o += align_up(N * DMGS_GACC_STRIDE * 4);
    total = o;
  }
};

int check_batch(int32_t B, int64_t P, int64_t K_cap, int32_t total_tiles, int64_t total_pixels) {
  if (B <= 0 || P <= 0 || (int64_t)B * P > (int64_t)0x7fffffff)
    return bfail(DMGS_EINVAL, "invalid batch: %d views x %lld Gaussians", B, (long long)P);
  if (K_cap < 0 || K_cap > (int64_t)0xffffffff) return bfail(DMGS_EINVAL, "invalid instance capacity");
  if (total_tiles <= 0 || total_pixels <= 0) return bfail(DMGS_EINVAL, "invalid view totals");
  return DMGS_OK;
}

dmgs::MVArgs mv_args(const dmgs_view* views, int32_t B, int64_t P, int32_t sh_degree, float mod,
                     const float* params, void* ws, const BatchLayout& L) {
  dmgs::MVArgs a{};
  a.P = (int)P;
  a.B = B;
  a.sh_degree = sh_degree;
  a.scale_modifier = mod;
  a.views = reinterpret_cast<const dmgs::ViewDesc*>(views);
  a.params = params;
  a.splats = at<float4>(ws, L.splats);
  a.radii = at<int32_t>(ws, L.radii);
  a.tiles = at<uint32_t>(ws, L.tiles);
  a.depth_keys = at<uint32_t>(ws, L.keys_a);
  a.order = at<uint32_t>(ws, L.order_a);
  a.gacc = at<float>(ws, L.gacc);
  return a;
}

dmgs::RenderArgs render_args(const dmgs_view* views, size_t N, int64_t K_cap, int32_t total_tiles,
                             const int32_t* tile_view, void* ws, const BatchLayout& L) {
  dmgs::RenderArgs r{};
  r.num_tiles = total_tiles;
  r.P = (uint32_t)N;
  r.K = (uint32_t)K_cap;
  r.K_dev = at<uint32_t>(ws, L.offsets) + (N - 1);
  r.views = reinterpret_cast<const dmgs::ViewDesc*>(views);
  r.tile_view = tile_view;
  r.ranges = at<uint2>(ws, L.ranges);
  r.point_list = at<uint32_t>(ws, L.vals_a);
  r.splats = at<float4>(ws, L.splats);
  r.final_T = at<float>(ws, L.final_T);
  r.n_contrib = at<uint32_t>(ws, L.n_contrib);
  r.gacc = at<float>(ws, L.gacc);
  return r;
}

}  // namespace

extern "C" {

int dmgs_batched_prepare_views(dmgs_view* views, int32_t B, int32_t* total_tiles, int64_t* total_pixels) {
  if (!views || B <= 0 || !total_tiles || !total_pixels) return bfail(DMGS_EINVAL, "bad arguments");
  int64_t tiles = 0, pixels = 0;
  for (int v = 0; v < B; ++v) {
    dmgs_view& d = views[v];
    if (d.width <= 0 || d.height <= 0 || !(d.tanfovx > 0.f) || !(d.tanfovy > 0.f))
      return bfail(DMGS_EINVAL, "view %d: invalid size or field of view", v);
    d.focal_x = (float)d.width / (2.0f * d.tanfovx);
    d.focal_y = (float)d.height / (2.0f * d.tanfovy);
    d.grid_x = (d.width + DMGS_BLOCK_X - 1) / DMGS_BLOCK_X;
    d.grid_y = (d.height + DMGS_BLOCK_Y - 1) / DMGS_BLOCK_Y;
    d.tile_base = (int32_t)tiles;
    d.pixel_base = pixels;
    tiles += (int64_t)d.grid_x * d.grid_y;
    pixels += (int64_t)d.width * d.height;
    if (tiles > (int64_t)0x7fffffff) return bfail(DMGS_EINVAL, "too many tiles");
  }
  *total_tiles = (int32_t)tiles;
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

size_t dmgs_batched_workspace_size(int32_t B, int64_t P, int64_t K_cap, int32_t total_tiles, int64_t total_pixels) {
  if (B <= 0 || P <= 0 || K_cap < 0 || total_tiles <= 0 || total_pixels <= 0) return 0;
  return BatchLayout((size_t)B * P, (size_t)K_cap, (size_t)total_tiles, (size_t)total_pixels).total;
}

int dmgs_batched_offsets(int32_t B, int64_t P, int64_t K_cap, int32_t total_tiles, int64_t total_pixels,
                       size_t offsets[6]) {
  int rc = check_batch(B, P, K_cap, total_tiles, total_pixels);
  if (rc) return rc;
  if (!offsets) return bfail(DMGS_EINVAL, "offsets is null");
  const size_t N = (size_t)B * P;
  BatchLayout L(N, (size_t)K_cap, (size_t)total_tiles, (size_t)total_pixels);
  offsets[0] = L.radii;                      // (B*P) int32 radii, view-major
  offsets[1] = L.offsets + 4 * (N - 1);      // uint32 instance count of the batch
  offsets[2] = L.ranges;
  offsets[3] = L.final_T;
  offsets[4] = L.n_contrib;
  offsets[5] = L.splats;
  return DMGS_OK;
}

int dmgs_batched_forward(const dmgs_view* views, int32_t B, int64_t P, int32_t sh_degree, float scale_modifier,
                       const float* params, int64_t K_cap, int32_t total_tiles, int64_t total_pixels,
                       const int32_t* tile_view, void* ws, float* images, int32_t variant, int32_t stages,
                       dmgs_stream_t stream) {
  int rc = check_batch(B, P, K_cap, total_tiles, total_pixels);
  if (rc) return rc;
  if (!views || !params || !ws || (K_cap > 0 && (!tile_view || !images)))
    return bfail(DMGS_EINVAL, "views, params, ws, tile_view and images are required");
  if (sh_degree < 0 || sh_degree > 3) return bfail(DMGS_EINVAL, "sh_degree must be in [0, 3]");
  if (stages <= 0 || (stages & ~DMGS_STAGES_ALL)) return bfail(DMGS_EINVAL, "invalid stage mask %d", stages);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const size_t N = (size_t)B * P;
  BatchLayout L(N, (size_t)K_cap, (size_t)total_tiles, (size_t)total_pixels);
  const dmgs::MVArgs a = mv_args(views, B, P, sh_degree, scale_modifier, params, ws, L);
  if (!(stages & DMGS_STAGE_GEOMETRY)) {
    if (K_cap == 0) return bfail(DMGS_EINVAL, "the blend stage needs a capacity");
    dmgs::RenderArgs r = render_args(views, N, K_cap, total_tiles, tile_view, ws, L);
    r.out_color = images;
    dmgs::launch_render_fwd(r, st, variant);
    return bcheck("batch render forward", st);
  }
  dmgs::launch_mv_preprocess(a, st);
  if ((rc = bcheck("batch preprocess", st))) return rc;
  uint32_t* gs = at<uint32_t>(ws, L.gscratch);
  dmgs::radix_sort_pairs(a.depth_keys, a.order, at<uint32_t>(ws, L.keys_b), at<uint32_t>(ws, L.order_b), N, 32, gs,
                         st);
  uint32_t* offs = at<uint32_t>(ws, L.offsets);
  dmgs::launch_gather_u32(a.tiles, a.order, offs, (int)N, st);
  dmgs::launch_inclusive_scan(offs, offs, N, gs, st);
  if ((rc = bcheck("batch depth sort / scan", st))) return rc;
  if (K_cap == 0) return DMGS_OK;  // count only: read it at dmgs_batched_offsets()[1]
  const uint32_t* K_dev = offs + (N - 1);
  dmgs::DupArgs d{};
  d.P = (int)N;
  d.order = a.order;
  d.offsets = offs;
  d.splats = a.splats;
  d.views = a.views;
  d.per_view = (uint32_t)P;
  d.tile_keys = at<uint32_t>(ws, L.tkeys_a);
  d.values = at<uint32_t>(ws, L.vals_a);
  d.K = (uint32_t)K_cap;
  d.K_dev = K_dev;
  d.chunk_first = at<uint32_t>(ws, L.bscratch);
  dmgs::launch_duplicate(d, st);
  dmgs::radix_sort_pairs(d.tile_keys, d.values, at<uint32_t>(ws, L.tkeys_b), at<uint32_t>(ws, L.vals_b),
                         (size_t)K_cap, bits_for((uint32_t)total_tiles), at<uint32_t>(ws, L.bscratch), st, K_dev);
  dmgs::launch_tile_ranges(d.tile_keys, (size_t)K_cap, at<uint2>(ws, L.ranges), total_tiles, st, K_dev);
  if ((rc = bcheck("batch binning", st))) return rc;
  if (!(stages & DMGS_STAGE_BLEND)) return DMGS_OK;
  dmgs::RenderArgs r = render_args(views, N, K_cap, total_tiles, tile_view, ws, L);
  r.out_color = images;
  dmgs::launch_render_fwd(r, st, variant);
  return bcheck("batch render forward", st);
}

int dmgs_batched_backward(const dmgs_view* views, int32_t B, int64_t P, int32_t sh_degree, float scale_modifier,
                        const float* params, int64_t K_cap, int32_t total_tiles, int64_t total_pixels,
                        const int32_t* tile_view, void* ws, const float* dL_dimages, float* grads, int32_t variant,
                        int32_t stages, dmgs_stream_t stream) {
  int rc = check_batch(B, P, K_cap, total_tiles, total_pixels);
  if (rc) return rc;
  if (!views || !params || !ws || !tile_view || !dL_dimages || !grads || K_cap == 0)
    return bfail(DMGS_EINVAL, "views, params, ws, tile_view, dL_dimages, grads and a capacity are required");
  if (sh_degree < 0 || sh_degree > 3) return bfail(DMGS_EINVAL, "sh_degree must be in [0, 3]");
  if (stages <= 0 || (stages & ~DMGS_STAGES_ALL)) return bfail(DMGS_EINVAL, "invalid stage mask %d", stages);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const size_t N = (size_t)B * P;
  BatchLayout L(N, (size_t)K_cap, (size_t)total_tiles, (size_t)total_pixels);
  if (stages & DMGS_STAGE_BLEND) {
    dmgs::RenderArgs r = render_args(views, N, K_cap, total_tiles, tile_view, ws, L);
    r.dL_dpix = dL_dimages;
    dmgs::launch_render_bwd(r, st, variant);
    if ((rc = bcheck("batch render backward", st))) return rc;
  }
  if (!(stages & DMGS_STAGE_GEOMETRY)) return DMGS_OK;
  dmgs::MVArgs a = mv_args(views, B, P, sh_degree, scale_modifier, params, ws, L);
  a.grads = grads;
  dmgs::launch_mv_geometry_bwd(a, st);
  return bcheck("batch geometry backward", st);
}

}  // extern "C"
