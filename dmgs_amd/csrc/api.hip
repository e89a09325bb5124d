// C-ABI entry points (include/dmgs.h): argument checks, workspace layout and
// kernel sequencing.  No allocation happens here; every buffer comes from the
// caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/dmgs.h"
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int check_launch(const char* what, hipStream_t s, bool debug) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && debug) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DMGS_EHIP, "%s: %s", what, hipGetErrorString(e));
  return DMGS_OK;
}

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct GeomLayout {
  size_t splats, keys_a, keys_b, order_a, order_b, tiles, offsets, scratch, total;
  explicit GeomLayout(size_t P) {
    size_t o = 0;
    splats = o; o += align_up(P * DMGS_SPLAT_F4 * 16);
    keys_a = o; o += align_up(P * 4);
    keys_b = o; o += align_up(P * 4);
    order_a = o; o += align_up(P * 4);
    order_b = o; o += align_up(P * 4);
    tiles = o; o += align_up(P * 4);
    offsets = o; o += align_up(P * 4);
    const size_t sw = std::max(dmgs::radix_scratch_words(P), dmgs::scan_scratch_words(P));
    scratch = o; o += align_up(sw * 4);
    total = o;
  }
};

struct BinLayout {
  size_t keys_a, keys_b, vals_a, vals_b, scratch, chunk, total;
  explicit BinLayout(size_t K) {
    size_t o = 0;
    keys_a = o; o += align_up(K * 4);
    keys_b = o; o += align_up(K * 4);
    vals_a = o; o += align_up(K * 4);
    vals_b = o; o += align_up(K * 4);
    scratch = o; o += align_up(dmgs::radix_scratch_words(K) * 4);  // the duplicate may fill the first histogram
    chunk = o; o += align_up(dmgs::duplicate_scratch_words(K) * 4);
    total = o;
  }
};

struct ImgLayout {
  size_t ranges, final_T, n_contrib, total;
  ImgLayout(int W, int H) {
    const size_t T = (size_t)((W + DMGS_BLOCK_X - 1) / DMGS_BLOCK_X) * ((H + DMGS_BLOCK_Y - 1) / DMGS_BLOCK_Y);
    const size_t HW = (size_t)W * H;
    size_t o = 0;
    ranges = o; o += align_up(T * 8);
    final_T = o; o += align_up(HW * 4);
    n_contrib = o; o += align_up(HW * 4);
    total = o;
  }
};

template <typename T>
T* at(void* base, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(base) + off); }
template <typename T>
const T* at(const void* base, size_t off) { return reinterpret_cast<const T*>(static_cast<const char*>(base) + off); }
// Sub-buffers sit at 256-byte offsets and kernels use 16-byte vector accesses.
bool misaligned(const void* p) { return p && (reinterpret_cast<uintptr_t>(p) & 255u) != 0; }

int tile_bits(int num_tiles) {
  int b = 0;
  while (b < 32 && ((uint32_t)num_tiles >> b) != 0) ++b;
  return b;
}

int check_settings(const dmgs_settings* s) {
  if (!s) return fail(DMGS_EINVAL, "settings is null");
  if (s->image_width <= 0 || s->image_height <= 0)
    return fail(DMGS_EINVAL, "image size must be positive (got %d x %d)", s->image_width, s->image_height);
  if ((int64_t)s->image_width * s->image_height > (int64_t)1 << 30)
    return fail(DMGS_EINVAL, "image too large (%d x %d)", s->image_width, s->image_height);
  if (!(s->tanfovx > 0.f) || !(s->tanfovy > 0.f))
    return fail(DMGS_EINVAL, "tanfovx / tanfovy must be positive");
  if (!s->viewmatrix || !s->projmatrix || !s->campos || !s->bg)
    return fail(DMGS_EINVAL, "viewmatrix, projmatrix, campos and bg must be device pointers");
  if (s->sh_degree < 0 || s->sh_degree > 3) return fail(DMGS_EINVAL, "sh_degree must be in [0, 3]");
  return DMGS_OK;
}

int check_inputs(const dmgs_settings* s, int64_t P, const float* means3D, const float* opacities, const float* shs,
                 const float* colors_precomp, const float* scales, const float* rotations,
                 const float* cov3D_precomp) {
  if (P < 0 || P > (int64_t)0x7fffffff) return fail(DMGS_EINVAL, "invalid Gaussian count %lld", (long long)P);
  if (P == 0) return DMGS_OK;
  if (!means3D || !opacities) return fail(DMGS_EINVAL, "means3D and opacities are required");
  if ((shs == nullptr) == (colors_precomp == nullptr))
    return fail(DMGS_EINVAL, "Please provide exactly one of either SHs or precomputed colors!");
  if (((scales == nullptr) || (rotations == nullptr)) == (cov3D_precomp == nullptr) ||
      ((scales == nullptr) != (rotations == nullptr)))
    return fail(DMGS_EINVAL, "Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!");
  if (shs) {
    const int need = (s->sh_degree + 1) * (s->sh_degree + 1);
    if (s->sh_coeffs < need)
      return fail(DMGS_EINVAL, "sh_coeffs (%d) smaller than (sh_degree+1)^2 (%d)", s->sh_coeffs, need);
  }
  return DMGS_OK;
}

void fill_geom_common(const dmgs_settings* s, float& fx, float& fy, int& gx, int& gy) {
  fy = (float)s->image_height / (2.0f * s->tanfovy);
  fx = (float)s->image_width / (2.0f * s->tanfovx);
  gx = (s->image_width + DMGS_BLOCK_X - 1) / DMGS_BLOCK_X;
  gy = (s->image_height + DMGS_BLOCK_Y - 1) / DMGS_BLOCK_Y;
}

}  // namespace

namespace dmgs {
// Error reporting for the entry points defined in other translation units.
void set_last_error(const char* msg) { g_last_error = msg; }
}  // namespace dmgs

extern "C" {

const char* dmgs_last_error(void) { return g_last_error.c_str(); }
const char* dmgs_version(void) { return "dmgs-gfx950 0.1"; }

size_t dmgs_geometry_workspace_size(int64_t P) { return P < 0 ? 0 : GeomLayout((size_t)P).total; }
size_t dmgs_binning_workspace_size(int64_t K) { return K < 0 ? 0 : BinLayout((size_t)K).total; }
size_t dmgs_image_workspace_size(int32_t width, int32_t height) {
  return (width <= 0 || height <= 0) ? 0 : ImgLayout(width, height).total;
}
size_t dmgs_backward_workspace_size(int64_t P) {
  return P < 0 ? 0 : align_up((size_t)P * DMGS_GACC_STRIDE * 4);
}

int dmgs_workspace_offsets(int64_t P, int64_t K, int32_t width, int32_t height, size_t offsets[10]) {
  if (!offsets || P < 0 || K < 0 || width <= 0 || height <= 0) return fail(DMGS_EINVAL, "bad arguments");
  GeomLayout g((size_t)P);
  BinLayout b((size_t)K);
  ImgLayout im(width, height);
  offsets[0] = g.splats;
  offsets[1] = g.keys_a;
  offsets[2] = g.order_a;
  offsets[3] = g.tiles;
  offsets[4] = g.offsets;
  offsets[5] = b.keys_a;
  offsets[6] = b.vals_a;
  offsets[7] = im.ranges;
  offsets[8] = im.final_T;
  offsets[9] = im.n_contrib;
  return DMGS_OK;
}

int dmgs_forward_geometry(const dmgs_settings* s, int64_t P, const float* means3D, const float* opacities,
                          const float* shs, const float* colors_precomp, const float* scales,
                          const float* rotations, const float* cov3D_precomp, int32_t* radii, void* geom_ws,
                          uint32_t* num_rendered_dev, dmgs_stream_t stream) {
  int rc = check_settings(s);
  if (rc) return rc;
  rc = check_inputs(s, P, means3D, opacities, shs, colors_precomp, scales, rotations, cov3D_precomp);
  if (rc) return rc;
  if (!num_rendered_dev) return fail(DMGS_EINVAL, "num_rendered_dev is null");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (P == 0) {
    if (hipMemsetAsync(num_rendered_dev, 0, 4, st) != hipSuccess) return fail(DMGS_EHIP, "memset failed");
    return DMGS_OK;
  }
  if (!radii || !geom_ws) return fail(DMGS_EINVAL, "radii and geom_ws are required");
  if (misaligned(geom_ws)) return fail(DMGS_EINVAL, "geom_ws must be 256-byte aligned");
  GeomLayout L((size_t)P);
  dmgs::PreprocessArgs a{};
  a.P = (int)P;
  a.W = s->image_width;
  a.H = s->image_height;
  fill_geom_common(s, a.focal_x, a.focal_y, a.grid_x, a.grid_y);
  a.tanfovx = s->tanfovx;
  a.tanfovy = s->tanfovy;
  a.scale_modifier = s->scale_modifier;
  a.sh_degree = s->sh_degree;
  a.sh_coeffs = shs ? s->sh_coeffs : 0;
  a.means3D = means3D;
  a.scales = scales;
  a.rotations = rotations;
  a.opacities = opacities;
  a.shs = shs;
  a.colors_precomp = colors_precomp;
  a.cov3D_precomp = cov3D_precomp;
  a.viewmatrix = s->viewmatrix;
  a.projmatrix = s->projmatrix;
  a.campos = s->campos;
  a.radii = radii;
  a.splats = at<float4>(geom_ws, L.splats);
  a.tiles_touched = at<uint32_t>(geom_ws, L.tiles);
  a.depth_keys = at<uint32_t>(geom_ws, L.keys_a);
  a.order = at<uint32_t>(geom_ws, L.order_a);
  dmgs::launch_preprocess_fwd(a, st);
  if ((rc = check_launch("preprocess", st, s->debug))) return rc;
  uint32_t* scratch = at<uint32_t>(geom_ws, L.scratch);
  dmgs::radix_sort_pairs(a.depth_keys, a.order, at<uint32_t>(geom_ws, L.keys_b), at<uint32_t>(geom_ws, L.order_b),
                         (size_t)P, 32, scratch, st);
  if ((rc = check_launch("depth sort", st, s->debug))) return rc;
  uint32_t* offs = at<uint32_t>(geom_ws, L.offsets);
  dmgs::launch_gather_u32(a.tiles_touched, a.order, offs, (int)P, st);
  dmgs::launch_inclusive_scan(offs, offs, (size_t)P, scratch, st);
  if ((rc = check_launch("scan", st, s->debug))) return rc;
  if (hipMemcpyAsync(num_rendered_dev, offs + (P - 1), 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return fail(DMGS_EHIP, "copy of the instance count failed");
  return check_launch("count copy", st, s->debug);
}

int dmgs_forward_binning(const dmgs_settings* s, int64_t P, int64_t K, const void* geom_ws, void* binning_ws,
                         void* image_ws, dmgs_stream_t stream) {
  int rc = check_settings(s);
  if (rc) return rc;
  if (P < 0 || P > (int64_t)0x7fffffff || K < 0 || K > (int64_t)0xffffffff)
    return fail(DMGS_EINVAL, "invalid P / K");
  if (!image_ws) return fail(DMGS_EINVAL, "image_ws is required");
  if (K > 0 && (!geom_ws || !binning_ws)) return fail(DMGS_EINVAL, "geom_ws and binning_ws are required");
  if (misaligned(geom_ws) || misaligned(binning_ws) || misaligned(image_ws))
    return fail(DMGS_EINVAL, "workspaces must be 256-byte aligned");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float fx, fy;
  int gx, gy;
  fill_geom_common(s, fx, fy, gx, gy);
  const int T = gx * gy;
  ImgLayout IL(s->image_width, s->image_height);
  uint2* ranges = at<uint2>(image_ws, IL.ranges);
  if (K > 0) {
    GeomLayout GL((size_t)P);
    BinLayout BL((size_t)K);
    uint32_t* ka = at<uint32_t>(binning_ws, BL.keys_a);
    uint32_t* va = at<uint32_t>(binning_ws, BL.vals_a);
    dmgs::DupArgs d{};
    d.P = (int)P;
    d.order = at<uint32_t>(geom_ws, GL.order_a);
    d.offsets = at<uint32_t>(geom_ws, GL.offsets);
    d.splats = at<float4>(geom_ws, GL.splats);
    d.grid_x = gx;
    d.grid_y = gy;
    d.tile_keys = ka;
    d.values = va;
    d.K = (uint32_t)K;
    d.chunk_first = at<uint32_t>(binning_ws, BL.chunk);
    // two tile-sort passes (9..16 tile bits) on 4096-key blocks: the duplicate
    // counts the first pass's digits per output chunk, so that pass needs no
    // histogram launch
    const int bits = tile_bits(T);
    const int w0 = (bits > 8 && bits <= 16) ? dmgs::radix_min_passes_first_width((size_t)K, bits) : -1;
    uint32_t* scratch = at<uint32_t>(binning_ws, BL.scratch);
    if (w0 > 0) {
      d.hist = scratch;
      d.hist_bits = w0;
    }
    dmgs::launch_duplicate(d, st);
    if ((rc = check_launch("duplicate", st, s->debug))) return rc;
    uint32_t* kb = at<uint32_t>(binning_ws, BL.keys_b);
    uint32_t* vb = at<uint32_t>(binning_ws, BL.vals_b);
    if (w0 > 0) {
      if (dmgs::radix_sort_pairs_min_passes(ka, va, kb, vb, (size_t)K, bits, scratch, st, nullptr, true))
        return fail(DMGS_EHIP, "internal: tile sort ended in the second buffer");
    } else {
      dmgs::radix_sort_pairs(ka, va, kb, vb, (size_t)K, bits, scratch, st);
    }
    if ((rc = check_launch("tile sort", st, s->debug))) return rc;
    dmgs::launch_tile_ranges(ka, (size_t)K, ranges, T, st);
  } else {
    dmgs::launch_tile_ranges(nullptr, 0, ranges, T, st);
  }
  return check_launch("tile ranges", st, s->debug);
}

int dmgs_render_forward(const dmgs_settings* s, int64_t P, int64_t K, const void* geom_ws, const void* binning_ws,
                        void* image_ws, float* out_color, dmgs_stream_t stream) {
  int rc = check_settings(s);
  if (rc) return rc;
  if (P < 0 || P > (int64_t)0x7fffffff || K < 0 || K > (int64_t)0xffffffff)
    return fail(DMGS_EINVAL, "invalid P / K");
  if (!image_ws || !out_color) return fail(DMGS_EINVAL, "image_ws and out_color are required");
  if (K > 0 && (!geom_ws || !binning_ws)) return fail(DMGS_EINVAL, "geom_ws and binning_ws are required");
  if (misaligned(geom_ws) || misaligned(binning_ws) || misaligned(image_ws))
    return fail(DMGS_EINVAL, "workspaces must be 256-byte aligned");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float fx, fy;
  int gx, gy;
  fill_geom_common(s, fx, fy, gx, gy);
  ImgLayout IL(s->image_width, s->image_height);
  dmgs::RenderArgs r{};
  r.W = s->image_width;
  r.H = s->image_height;
  r.grid_x = gx;
  r.grid_y = gy;
  r.num_tiles = gx * gy;
  r.P = (uint32_t)P;
  r.K = (uint32_t)K;
  r.ranges = at<uint2>(image_ws, IL.ranges);
  r.point_list = K > 0 ? at<uint32_t>(binning_ws, BinLayout((size_t)K).vals_a) : nullptr;
  r.splats = K > 0 ? at<float4>(geom_ws, GeomLayout((size_t)P).splats) : nullptr;
  r.bg = s->bg;
  r.out_color = out_color;
  r.final_T = at<float>(image_ws, IL.final_T);
  r.n_contrib = at<uint32_t>(image_ws, IL.n_contrib);
  dmgs::launch_render_fwd(r, st);
  return check_launch("render forward", st, s->debug);
}

int dmgs_forward_render(const dmgs_settings* s, int64_t P, int64_t K, void* geom_ws, void* binning_ws,
                        void* image_ws, float* out_color, dmgs_stream_t stream) {
  int rc = dmgs_forward_binning(s, P, K, geom_ws, binning_ws, image_ws, stream);
  if (rc) return rc;
  return dmgs_render_forward(s, P, K, geom_ws, binning_ws, image_ws, out_color, stream);
}

namespace {
int check_bwd_common(const dmgs_settings* s, int64_t P, int64_t K, const float* shs, const float* colors_precomp,
                     const float* scales, const float* rotations, const float* cov3D_precomp) {
  int rc = check_settings(s);
  if (rc) return rc;
  if (P < 0 || P > (int64_t)0x7fffffff) return fail(DMGS_EINVAL, "invalid Gaussian count %lld", (long long)P);
  if (K < 0 || K > (int64_t)0xffffffff) return fail(DMGS_EINVAL, "invalid K");
  if (P == 0) return DMGS_OK;
  if ((shs == nullptr) == (colors_precomp == nullptr))
    return fail(DMGS_EINVAL, "Please provide exactly one of either SHs or precomputed colors!");
  if ((scales == nullptr) != (rotations == nullptr) || (scales == nullptr) == (cov3D_precomp == nullptr))
    return fail(DMGS_EINVAL, "Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!");
  return DMGS_OK;
}
}  // namespace

int dmgs_render_backward(const dmgs_settings* s, int64_t P, int64_t K, const void* geom_ws, const void* binning_ws,
                         const void* image_ws, void* bwd_ws, const float* dL_dout, dmgs_stream_t stream) {
  int rc = check_settings(s);
  if (rc) return rc;
  if (P < 0 || P > (int64_t)0x7fffffff || K < 0 || K > (int64_t)0xffffffff)
    return fail(DMGS_EINVAL, "invalid P / K");
  if (P == 0) return DMGS_OK;
  if (!geom_ws || !image_ws || !bwd_ws || !dL_dout) return fail(DMGS_EINVAL, "missing buffers");
  if (misaligned(geom_ws) || misaligned(binning_ws) || misaligned(image_ws) || misaligned(bwd_ws))
    return fail(DMGS_EINVAL, "workspaces must be 256-byte aligned");
  if (K > 0 && !binning_ws) return fail(DMGS_EINVAL, "binning_ws is required");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float fx, fy;
  int gx, gy;
  fill_geom_common(s, fx, fy, gx, gy);
  ImgLayout IL(s->image_width, s->image_height);
  float* gacc = static_cast<float*>(bwd_ws);
  if (hipMemsetAsync(gacc, 0, (size_t)P * DMGS_GACC_STRIDE * 4, st) != hipSuccess)
    return fail(DMGS_EHIP, "memset of the gradient accumulator failed");
  if (K == 0) return check_launch("gradient accumulator clear", st, s->debug);
  dmgs::RenderArgs r{};
  r.W = s->image_width;
  r.H = s->image_height;
  r.grid_x = gx;
  r.grid_y = gy;
  r.num_tiles = gx * gy;
  r.P = (uint32_t)P;
  r.K = (uint32_t)K;
  r.ranges = at<uint2>(image_ws, IL.ranges);
  r.point_list = at<uint32_t>(binning_ws, BinLayout((size_t)K).vals_a);
  r.splats = at<float4>(geom_ws, GeomLayout((size_t)P).splats);
  r.bg = s->bg;
  r.final_T = const_cast<float*>(at<float>(image_ws, IL.final_T));
  r.n_contrib = const_cast<uint32_t*>(at<uint32_t>(image_ws, IL.n_contrib));
  r.dL_dpix = dL_dout;
  r.gacc = gacc;
  dmgs::launch_render_bwd(r, st, dmgs::kRenderMixed);
  return check_launch("render backward", st, s->debug);
}

int dmgs_geometry_backward(const dmgs_settings* s, int64_t P, const float* means3D, const float* shs,
                           const float* colors_precomp, const float* scales, const float* rotations,
                           const float* cov3D_precomp, const int32_t* radii, const void* bwd_ws, float* dL_dmeans2D,
                           float* dL_dcolors, float* dL_dopacity, float* dL_dconic, float* dL_dmeans3D,
                           float* dL_dcov3D, float* dL_dsh, float* dL_dscales, float* dL_drotations,
                           dmgs_stream_t stream) {
  int rc = check_bwd_common(s, P, 0, shs, colors_precomp, scales, rotations, cov3D_precomp);
  if (rc) return rc;
  if (P == 0) return DMGS_OK;
  if (!means3D || !radii || !bwd_ws) return fail(DMGS_EINVAL, "missing buffers");
  if (shs && !dL_dsh) return fail(DMGS_EINVAL, "dL_dsh is required with shs");
  if (scales && (!dL_dscales || !dL_drotations)) return fail(DMGS_EINVAL, "dL_dscales/dL_drotations required");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float fx, fy;
  int gx, gy;
  fill_geom_common(s, fx, fy, gx, gy);
  dmgs::GeometryBwdArgs g{};
  g.P = (int)P;
  g.W = s->image_width;
  g.H = s->image_height;
  g.tanfovx = s->tanfovx;
  g.tanfovy = s->tanfovy;
  g.focal_x = fx;
  g.focal_y = fy;
  g.scale_modifier = s->scale_modifier;
  g.sh_degree = s->sh_degree;
  g.sh_coeffs = shs ? s->sh_coeffs : 0;
  g.means3D = means3D;
  g.scales = scales;
  g.rotations = rotations;
  g.shs = shs;
  g.cov3D_precomp = cov3D_precomp;
  g.viewmatrix = s->viewmatrix;
  g.projmatrix = s->projmatrix;
  g.campos = s->campos;
  g.radii = radii;
  g.gacc = static_cast<const float*>(bwd_ws);
  g.dL_dmeans2D = dL_dmeans2D;
  g.dL_dcolors = dL_dcolors;
  g.dL_dopacity = dL_dopacity;
  g.dL_dconic = dL_dconic;
  g.dL_dmeans3D = dL_dmeans3D;
  g.dL_dcov3D = dL_dcov3D;
  g.dL_dsh = dL_dsh;
  g.dL_dscales = dL_dscales;
  g.dL_drotations = dL_drotations;
  dmgs::launch_geometry_bwd(g, st);
  return check_launch("geometry backward", st, s->debug);
}

int dmgs_backward(const dmgs_settings* s, int64_t P, int64_t K, const float* means3D, const float* shs,
                  const float* colors_precomp, const float* scales, const float* rotations,
                  const float* cov3D_precomp, const int32_t* radii, const void* geom_ws, const void* binning_ws,
                  const void* image_ws, void* bwd_ws, const float* dL_dout, float* dL_dmeans2D,
                  float* dL_dcolors, float* dL_dopacity, float* dL_dconic, float* dL_dmeans3D, float* dL_dcov3D,
                  float* dL_dsh, float* dL_dscales, float* dL_drotations, dmgs_stream_t stream) {
  int rc = check_bwd_common(s, P, K, shs, colors_precomp, scales, rotations, cov3D_precomp);
  if (rc) return rc;
  if (P == 0) return DMGS_OK;
  if (!means3D || !radii) return fail(DMGS_EINVAL, "missing buffers");
  if (shs && !dL_dsh) return fail(DMGS_EINVAL, "dL_dsh is required with shs");
  if (scales && (!dL_dscales || !dL_drotations)) return fail(DMGS_EINVAL, "dL_dscales/dL_drotations required");
  rc = dmgs_render_backward(s, P, K, geom_ws, binning_ws, image_ws, bwd_ws, dL_dout, stream);
  if (rc) return rc;
  return dmgs_geometry_backward(s, P, means3D, shs, colors_precomp, scales, rotations, cov3D_precomp, radii, bwd_ws,
                                dL_dmeans2D, dL_dcolors, dL_dopacity, dL_dconic, dL_dmeans3D, dL_dcov3D, dL_dsh,
                                dL_dscales, dL_drotations, stream);
}

int dmgs_mark_visible(int64_t P, const float* means3D, const float* viewmatrix, uint8_t* present,
                      dmgs_stream_t stream) {
  if (P < 0 || P > (int64_t)0x7fffffff) return fail(DMGS_EINVAL, "invalid P");
  if (P == 0) return DMGS_OK;
  if (!means3D || !viewmatrix || !present) return fail(DMGS_EINVAL, "null pointer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::launch_mark_visible((int)P, means3D, viewmatrix, present, st);
  return check_launch("mark visible", st, false);
}

size_t dmgs_sort_scratch_size(int64_t n) { return n < 0 ? 0 : align_up(dmgs::radix_scratch_words((size_t)n) * 4); }

int dmgs_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int32_t bits, void* scratch,
                    dmgs_stream_t stream) {
  if (n < 0 || n > (int64_t)0xffffffff) return fail(DMGS_EINVAL, "invalid n");
  if (bits < 0 || bits > 32) return fail(DMGS_EINVAL, "bits must be in [0, 32]");
  if (n == 0) return DMGS_OK;
  if (!k0 || !v0 || !k1 || !v1 || !scratch) return fail(DMGS_EINVAL, "null pointer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::radix_sort_pairs(k0, v0, k1, v1, (size_t)n, bits, static_cast<uint32_t*>(scratch), st);
  return check_launch("sort", st, false);
}

size_t dmgs_scan_scratch_size(int64_t n) { return n < 0 ? 0 : align_up(dmgs::scan_scratch_words((size_t)n) * 4); }

int dmgs_inclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* scratch, dmgs_stream_t stream) {
  if (n < 0) return fail(DMGS_EINVAL, "invalid n");
  if (n == 0) return DMGS_OK;
  if (!in || !out || !scratch) return fail(DMGS_EINVAL, "null pointer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::launch_inclusive_scan(in, out, (size_t)n, static_cast<uint32_t*>(scratch), st);
  return check_launch("scan", st, false);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// fused L1 + SSIM loss
// ---------------------------------------------------------------------------
namespace {
void ssim_window(float w[11]) {
  // the reference builds the window in float32: exp values, then divide by their sum
  float sum = 0.f;
  for (int x = 0; x < 11; ++x) {
    w[x] = (float)std::exp(-(double)((x - 5) * (x - 5)) / (2.0 * 1.5 * 1.5));
    sum += w[x];
  }
  for (int x = 0; x < 11; ++x) w[x] = w[x] / sum;
}

struct SsimLayout {
  size_t coef, partials, out, total;
  SsimLayout(int C, int H, int W) {
    size_t o = 0;
    coef = o; o += align_up((size_t)3 * C * H * W * 4);
    partials = o; o += align_up(dmgs::dssim_partials(C, H, W) * 4);
    out = o; o += align_up(4 * 4);
    total = o;
  }
};
}  // namespace

extern "C" {

size_t dmgs_dssim_l1_workspace_size(int32_t C, int32_t H, int32_t W) {
  return (C <= 0 || H <= 0 || W <= 0) ? 0 : SsimLayout(C, H, W).total;
}

int dmgs_dssim_target_stats(const float* gt, int32_t C, int32_t H, int32_t W, float* stats, dmgs_stream_t stream) {
  if (!gt || !stats) return fail(DMGS_EINVAL, "null pointer");
  if (C <= 0 || H <= 0 || W <= 0 || H >= (1 << 24) || W >= (1 << 24) || (int64_t)C * H * W > ((int64_t)1 << 31))
    return fail(DMGS_EINVAL, "invalid image shape %d x %d x %d", C, H, W);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float w[11];
  ssim_window(w);
  dmgs::launch_dssim_target(gt, C, H, W, w, stats, st);
  return check_launch("dssim target statistics", st, false);
}

int dmgs_dssim_l1_forward_groups_stats(const float* img, const float* gt, const float* gt_stats, int32_t G,
                                       int32_t Cg, int32_t H, int32_t W, float lambda_dssim, void* ws,
                                       float* loss_out, dmgs_stream_t stream) {
  if (!img || !gt || !ws || !loss_out) return fail(DMGS_EINVAL, "null pointer");
  if (G <= 0 || Cg <= 0 || H <= 0 || W <= 0 || H >= (1 << 24) || W >= (1 << 24) ||
      (int64_t)G * Cg * H * W > ((int64_t)1 << 31))
    return fail(DMGS_EINVAL, "invalid image shape %d x %d x %d x %d", G, Cg, H, W);
  if (misaligned(ws)) return fail(DMGS_EINVAL, "ws must be 256-byte aligned");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  SsimLayout L(G * Cg, H, W);
  float w[11];
  ssim_window(w);
  // the final reduction writes (loss, l1, ssim) per group straight to loss_out (no copy node when captured)
  dmgs::launch_dssim_forward(img, gt, gt_stats, G * Cg, Cg, H, W, w, lambda_dssim, at<float>(ws, L.coef),
                             at<float>(ws, L.partials), loss_out, st);
  return check_launch("dssim forward", st, false);
}

int dmgs_dssim_l1_forward_groups(const float* img, const float* gt, int32_t G, int32_t Cg, int32_t H, int32_t W,
                                 float lambda_dssim, void* ws, float* loss_out, dmgs_stream_t stream) {
  return dmgs_dssim_l1_forward_groups_stats(img, gt, nullptr, G, Cg, H, W, lambda_dssim, ws, loss_out, stream);
}

int dmgs_dssim_l1_backward_groups(const float* img, const float* gt, int32_t G, int32_t Cg, int32_t H, int32_t W,
                                  float lambda_dssim, const void* ws, const float* dloss, float* dimg,
                                  dmgs_stream_t stream) {
  if (!img || !gt || !ws || !dloss || !dimg) return fail(DMGS_EINVAL, "null pointer");
  if (G <= 0 || Cg <= 0 || H <= 0 || W <= 0 || H >= (1 << 24) || W >= (1 << 24) ||
      (int64_t)G * Cg * H * W > ((int64_t)1 << 31))
    return fail(DMGS_EINVAL, "invalid image shape %d x %d x %d x %d", G, Cg, H, W);
  if (misaligned(ws)) return fail(DMGS_EINVAL, "ws must be 256-byte aligned");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  SsimLayout L(G * Cg, H, W);
  float w[11];
  ssim_window(w);
  dmgs::launch_dssim_backward(img, gt, at<float>(ws, L.coef), G * Cg, Cg, H, W, w, lambda_dssim, dloss, dimg, st);
  return check_launch("dssim backward", st, false);
}

int dmgs_dssim_l1_forward(const float* img, const float* gt, int32_t C, int32_t H, int32_t W, float lambda_dssim,
                          void* ws, float* loss_out, dmgs_stream_t stream) {
  return dmgs_dssim_l1_forward_groups(img, gt, 1, C, H, W, lambda_dssim, ws, loss_out, stream);
}

int dmgs_dssim_l1_backward(const float* img, const float* gt, int32_t C, int32_t H, int32_t W, float lambda_dssim,
                           const void* ws, const float* dloss, float* dimg, dmgs_stream_t stream) {
  return dmgs_dssim_l1_backward_groups(img, gt, 1, C, H, W, lambda_dssim, ws, dloss, dimg, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// mesh binding
// ---------------------------------------------------------------------------
namespace {
int check_mesh(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G, const float* bc) {
  if (V <= 0 || V > (int64_t)0x7fffffff / 3) return fail(DMGS_EINVAL, "invalid vertex count %lld", (long long)V);
  if (F < 0 || G <= 0 || F * (int64_t)G > (int64_t)0x7fffffff)
    return fail(DMGS_EINVAL, "invalid face count %lld or Gaussians per face %d", (long long)F, G);
  if (!verts || !faces || !bc) return fail(DMGS_EINVAL, "verts, faces and barycentrics are required");
  return DMGS_OK;
}
}  // namespace

extern "C" {

int dmgs_bind_forward(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G, const float* bc,
                      int32_t bc_per_gaussian, const float* rot2d, const float* scale2d, float thin_z, float* xyz,
                      float* rot, float* scale, dmgs_stream_t stream) {
  int rc = check_mesh(V, verts, F, faces, G, bc);
  if (rc) return rc;
  if (!xyz) return fail(DMGS_EINVAL, "xyz output is required");
  if ((rot2d == nullptr) != (rot == nullptr) || (scale2d == nullptr) != (scale == nullptr))
    return fail(DMGS_EINVAL, "rotation / scale inputs and outputs must be given together");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::BindArgs a{};
  a.F = (int)F;
  a.G = G;
  a.verts = verts;
  a.faces = faces;
  a.bc = bc;
  a.bc_per_gaussian = bc_per_gaussian;
  a.rot2d = rot2d;
  a.scale2d = scale2d;
  a.thin_z = thin_z;
  a.xyz = xyz;
  a.rot = rot;
  a.scale = scale;
  dmgs::launch_bind_fwd(a, st);
  return check_launch("bind forward", st, false);
}

int dmgs_bind_backward(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G, const float* bc,
                       int32_t bc_per_gaussian, const float* rot2d, const float* scale2d, const float* dxyz,
                       const float* drot, const float* dscale, float* dverts, float* drot2d, float* dscale2d,
                       dmgs_stream_t stream) {
  int rc = check_mesh(V, verts, F, faces, G, bc);
  if (rc) return rc;
  if (!dxyz || !dverts) return fail(DMGS_EINVAL, "dxyz and dverts are required");
  if (drot && (!rot2d || !drot2d)) return fail(DMGS_EINVAL, "drot needs rot2d and drot2d");
  if (dscale && (!scale2d || !dscale2d)) return fail(DMGS_EINVAL, "dscale needs scale2d and dscale2d");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::BindArgs a{};
  a.F = (int)F;
  a.G = G;
  a.verts = verts;
  a.faces = faces;
  a.bc = bc;
  a.bc_per_gaussian = bc_per_gaussian;
  a.rot2d = rot2d;
  a.scale2d = scale2d;
  a.dxyz = dxyz;
  a.drot = drot;
  a.dscale = dscale;
  a.dverts = dverts;
  a.drot2d = drot2d;
  a.dscale2d = dscale2d;
  dmgs::launch_bind_bwd(a, st);
  return check_launch("bind backward", st, false);
}

int dmgs_bind_cov_forward(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G, const float* bc,
                          int32_t bc_per_gaussian, const float* k, float rad_base, float flat_scale, int32_t adaptive,
                          float* xyz, float* cov, float* Lmat, dmgs_stream_t stream) {
  int rc = check_mesh(V, verts, F, faces, G, bc);
  if (rc) return rc;
  if (!k || !xyz || !cov || !Lmat) return fail(DMGS_EINVAL, "k, xyz, cov and Lmat are required");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::BindArgs a{};
  a.F = (int)F;
  a.G = G;
  a.verts = verts;
  a.faces = faces;
  a.bc = bc;
  a.bc_per_gaussian = bc_per_gaussian;
  a.k = k;
  a.rad_base = rad_base;
  a.flat_scale = flat_scale;
  a.adaptive = adaptive;
  a.xyz = xyz;
  a.cov = cov;
  a.Lmat = Lmat;
  dmgs::launch_bind_cov_fwd(a, st);
  return check_launch("bind cov forward", st, false);
}

int dmgs_bind_cov_backward(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G,
                           const float* bc, int32_t bc_per_gaussian, const float* k, const float* Lmat,
                           const float* dxyz, const float* dcov, float* dverts, float* dk, dmgs_stream_t stream) {
  int rc = check_mesh(V, verts, F, faces, G, bc);
  if (rc) return rc;
  if (!k || !Lmat || !dxyz || !dcov || !dverts) return fail(DMGS_EINVAL, "missing buffers");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::BindArgs a{};
  a.F = (int)F;
  a.G = G;
  a.verts = verts;
  a.faces = faces;
  a.bc = bc;
  a.bc_per_gaussian = bc_per_gaussian;
  a.k = k;
  a.Lmat = const_cast<float*>(Lmat);
  a.dxyz = dxyz;
  a.dcov = dcov;
  a.dverts = dverts;
  a.dk = dk;
  dmgs::launch_bind_bwd(a, st);
  return check_launch("bind cov backward", st, false);
}

int dmgs_faces_in_frustum(int64_t V, const float* verts, int64_t F, const int32_t* faces, const float* projmatrix,
                          uint8_t* mask, dmgs_stream_t stream) {
  if (V <= 0 || F < 0 || F > (int64_t)0x7fffffff) return fail(DMGS_EINVAL, "invalid sizes");
  if (F == 0) return DMGS_OK;
  if (!verts || !faces || !projmatrix || !mask) return fail(DMGS_EINVAL, "null pointer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::launch_faces_in_frustum((int)F, verts, faces, projmatrix, mask, st);
  return check_launch("faces in frustum", st, false);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scene and frame conversions
// ---------------------------------------------------------------------------
int dmgs_ply_rows_to_params(const float* rows, int64_t P, int32_t row_floats, const int32_t channel_cols[59],
                            float* params, dmgs_stream_t stream) {
  if (P < 0 || P > (int64_t)0x7fffffff) return fail(DMGS_EINVAL, "invalid P");
  if (row_floats <= 0 || row_floats > dmgs::kMaxRowFloats)
    return fail(DMGS_EINVAL, "row_floats must be in [1, %d]", dmgs::kMaxRowFloats);
  if (P == 0) return DMGS_OK;
  if (!rows || !channel_cols || !params) return fail(DMGS_EINVAL, "null pointer");
  dmgs::ChannelMap map{};
  map.channels = DMGS_MV_CHANNELS;
  for (int c = 0; c < DMGS_MV_CHANNELS; ++c) {
    if (channel_cols[c] < -1 || channel_cols[c] >= row_floats)
      return fail(DMGS_EINVAL, "channel %d: column %d outside the row", c, channel_cols[c]);
    map.col[c] = channel_cols[c];
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::launch_rows_to_channels(rows, P, row_floats, map, params, st);
  return check_launch("ply rows to params", st, false);
}

int dmgs_params_to_ply_rows(const float* params, int64_t P, int32_t row_floats, const int32_t* column_channels,
                            float* rows, dmgs_stream_t stream) {
  if (P < 0 || P > (int64_t)0x7fffffff) return fail(DMGS_EINVAL, "invalid P");
  if (row_floats <= 0 || row_floats > dmgs::kMaxRowFloats)
    return fail(DMGS_EINVAL, "row_floats must be in [1, %d]", dmgs::kMaxRowFloats);
  if (P == 0) return DMGS_OK;
  if (!rows || !column_channels || !params) return fail(DMGS_EINVAL, "null pointer");
  dmgs::ChannelMap map{};
  map.channels = row_floats;
  for (int j = 0; j < row_floats; ++j) {
    if (column_channels[j] < -1 || column_channels[j] >= DMGS_MV_CHANNELS)
      return fail(DMGS_EINVAL, "column %d: channel %d outside the scene layout", j, column_channels[j]);
    map.col[j] = column_channels[j];
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::launch_channels_to_rows(params, P, row_floats, map, rows, st);
  return check_launch("params to ply rows", st, false);
}

int dmgs_image_rgb8(const float* chw, int32_t H, int32_t W, uint8_t* hwc, dmgs_stream_t stream) {
  if (H < 0 || W < 0) return fail(DMGS_EINVAL, "invalid image size");
  if ((int64_t)H * W == 0) return DMGS_OK;
  if (!chw || !hwc) return fail(DMGS_EINVAL, "null pointer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  dmgs::launch_rgb8(chw, (int64_t)H * W, hwc, st);
  return check_launch("image rgb8", st, false);
}
