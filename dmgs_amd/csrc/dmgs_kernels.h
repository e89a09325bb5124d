// Internal kernel argument structs and launchers (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dmgs_common.h"

namespace dmgs {

struct PreprocessArgs {
  int P;
  int W, H, grid_x, grid_y;
  float tanfovx, tanfovy, focal_x, focal_y, scale_modifier;
  int sh_degree, sh_coeffs;
  const float* means3D;
  const float* scales;
  const float* rotations;
  const float* opacities;
  const float* shs;
  const float* colors_precomp;
  const float* cov3D_precomp;
  const float* viewmatrix;
  const float* projmatrix;
  const float* campos;
  // outputs
  int32_t* radii;
  float4* splats;
  uint32_t* tiles_touched;
  uint32_t* depth_keys;
  uint32_t* order;
};

void launch_preprocess_fwd(const PreprocessArgs& a, hipStream_t s);
void launch_mark_visible(int P, const float* means3D, const float* viewmatrix, uint8_t* present, hipStream_t s);

// ---- scan / sort / binning ----
// Inclusive scan of n uint32 values; scratch must hold scan_scratch_words(n) words.
size_t scan_scratch_words(size_t n);
void launch_inclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch, hipStream_t s);
// The same over instance counts, also filling the duplicate's chunk table
// (DupArgs::chunk_first for K output slots; then set DupArgs::chunk_ready).
void launch_inclusive_scan_chunks(const uint32_t* in, uint32_t* out, size_t n, uint32_t* scratch,
                                  uint32_t* chunk_first, uint32_t K, hipStream_t s);

// Depth sort key of a culled entry (never a sort key of a visible one), and
// the far-depth bucket bits of view-major keys.
constexpr uint32_t kSortSkip = 0xFFFFFFFFu;
constexpr uint32_t kFarBucketBits = 0x0F000000u;

// Stable LSD radix sort of (key, value) pairs over bits [0, bits), in passes of
// at most 8 bits.  Data starts in (k0, v0); (k1, v1) is the ping-pong buffer.
// The number of passes is always even, so the result ends in (k0, v0).
size_t radix_scratch_words(size_t n);
// n_dev (optional) holds the actual length on the device; n is then the capacity.
void radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                      uint32_t* scratch, hipStream_t s, const uint32_t* n_dev = nullptr);
// The same with as few passes as the bits need (one pass up to 8 bits);
// returns true when the result is in (k1, v1).  first_hist_ready: the first
// pass's block histogram is already in the scratch (see DupArgs::hist).
bool radix_sort_pairs_min_passes(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, int bits,
                                 uint32_t* scratch, hipStream_t s, const uint32_t* n_dev = nullptr,
                                 bool first_hist_ready = false);
// Depth sort of a batch's entries straight from the preprocess's keys: k0
// holds n per-entry 32-bit keys, kSortSkip for culled entries.  The first of
// the four passes drops the skipped keys and takes each entry's index as its
// value, and stores the number of kept keys in *n_out, which bounds the
// other passes; the sorted (key, entry) pairs end in (k0, v0).  flags
// (optional, view-major keys only): bit 0 is set when a key lies in the
// far-depth bucket.  The three low bytes are close to uniform and are
// counted with plain LDS atomics.
void radix_sort_entry_keys(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, size_t n, uint32_t* scratch,
                           hipStream_t s, uint32_t* n_out, uint32_t* flags);
// Digit width of the first pass of radix_sort_pairs_min_passes, or -1 when
// its blocks are not 4096 keys (the duplicate's chunks) for this n.
int radix_min_passes_first_width(size_t n, int bits);

// tiles_touched gathered into depth order, ready for the scan.
void launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t* dst, int n, hipStream_t s);

// Emit (tile, splat) pairs in depth order.  Single view: grid_x/grid_y give
// the tile grid.  Several views: entry g belongs to view g / per_view, the
// pairs are (binning cell | tile mask << 16, splat) with cells numbered
// globally from that view's cell_base, or with view_major as
// (cell within the view | view << 12 | tile mask << 16).
struct DupArgs {
  int P;                        // depth-sorted entries
  const uint32_t* order;
  const uint32_t* offsets;      // inclusive prefix of tiles touched, in depth order
  const float4* splats;         // single view: radius in s2.z
  const uint2* rects;           // several views: tile rectangles in depth order (pack_rect)
  int grid_x, grid_y;
  const ViewDesc* views;
  uint32_t per_view;
  uint32_t* tile_keys;
  uint32_t* values;
  uint32_t K;                   // capacity of tile_keys / values
  const uint32_t* K_dev;        // optional device count (clamped to K)
  uint32_t* chunk_first;        // duplicate_scratch_words(K) words
  bool chunk_ready;             // chunk_first already filled (launch_inclusive_scan_chunks)
  bool view_major;
  // optional: the first radix pass's histogram (digits = the low hist_bits of
  // a key), written per 4096-slot chunk into the sort scratch
  uint32_t* hist;
  int hist_bits;
  // optional: words a later step needs zeroed (cleared in shares by the work-groups)
  uint32_t* zero;
  size_t zero_n;
};
size_t duplicate_scratch_words(size_t K);
void launch_duplicate(const DupArgs& a, hipStream_t s);
// key_mask selects the tile / cell id bits of a key (the rest is carried data).
void launch_tile_ranges(const uint32_t* tile_keys, size_t K, uint2* ranges, int num_tiles, hipStream_t s,
                        const uint32_t* K_dev = nullptr, uint32_t key_mask = 0xFFFFFFFFu);
void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
// Ranges of view-major lists sorted on the cell-within-view bits: each
// (view, cell) run is contiguous; ranges are indexed by global cell id
// (views[view].cell_base + cell).
void launch_view_cell_ranges(const uint32_t* keys, size_t K, uint2* ranges, int num_cells, const ViewDesc* views,
                             hipStream_t s, const uint32_t* K_dev, bool ranges_zeroed = false);

// ---- multi-view per-Gaussian stages ----
// Component-major scene parameters (channels x P).
#define DMGS_MV_XYZ 0
#define DMGS_MV_SCALE 3
#define DMGS_MV_ROT 6
#define DMGS_MV_OPACITY 10
#define DMGS_MV_SH 11
#define DMGS_MV_CHANNELS 59
struct MVArgs {
  int P, B, sh_degree;
  float scale_modifier;
  const ViewDesc* views;
  const float* params;       // (59, P)
  float4* splats;            // (B*P) records
  int32_t* radii;            // (B*P)
  uint2* rects;              // (B*P) tile rectangle, pack_rect; 0 when culled
  uint32_t* depth_keys;      // (B*P) depth sort keys (entry_sort_key), kSortSkip for culled entries
  uint32_t* order;           // (B*P)
  float* gacc;               // B*P x DMGS_GACC_STRIDE blend-backward accumulator (gacc_index); zero between backward passes
  float* grads;              // (59, P)
  float* dens;               // optional (3, P): densification statistics (see dmgs_batch)
  uint32_t* status;          // optional: zeroed by the preprocess (flags set by the depth sort)
  int view_major;            // depth keys also order by view (at most 16 views)
};
void launch_mv_preprocess(const MVArgs& a, hipStream_t s);
// rects_sorted[i] = rects[order[i]], counts[i] = its number of binning cells
// (0 for i >= *n_dev when given).
void launch_gather_rects(const uint2* rects, const uint32_t* order, uint2* rects_sorted, uint32_t* counts, size_t n,
                         hipStream_t s, const uint32_t* n_dev = nullptr);
// Reads the accumulator rows of visible entries (the forward blend cleared
// the accumulator; the blend backward filled it).
void launch_mv_geometry_bwd(const MVArgs& a, hipStream_t s);

// ---- blend ----
// Single view: W, H, grid_x, grid_y, bg and the image pointers describe the
// view.  Several views: `views` / `tile_view` map each global tile to its
// view; images, final_T and n_contrib are packed at each view's pixel_base.
struct RenderArgs {
  int W, H, grid_x, grid_y;
  int num_tiles;          // tiles in the launch (all views)
  uint32_t P, K;          // splat records and list capacity (guards for the gathers)
  const uint32_t* K_dev;  // optional device list length (clamped to K)
  const ViewDesc* views;  // optional view table (multi-view)
  const int32_t* tile_view;
  const uint2* ranges;
  const uint32_t* point_list;
  const uint32_t* list_keys;  // multi-view: sorted keys (cell | tile mask << 16) beside point_list
  const float4* splats;
  const float* bg;
  float* out_color;      // (3, H, W)
  float* final_T;        // (H, W)
  uint32_t* n_contrib;   // (H, W)
  // backward only
  const float* dL_dpix;  // (3, H, W)
  float* gacc;           // P x DMGS_GACC_STRIDE floats, two planes (gacc_index)
  // optional (multi-view): work-group slot -> tile, heaviest tiles first
  const uint32_t* tile_order;
  // optional (multi-view, strip forward): per tile, the largest last
  // contributor of its pixels (the backward's list length)
  uint32_t* tile_work;
  // optional (forward): the backward accumulator, zeroed by the forward's
  // work-groups in equal shares after their blend work
  float4* clear;
  size_t clear_n;  // float4 count
  // and these words (work-group 0): the tile-order histogram the backward's
  // dispatch order starts from
  uint32_t* clear_words;
  int clear_words_n;
};
// Work-group shape / LDS batch depth of the blend kernels (results are the
// same for every variant; the autotuner picks the fastest).  The strip
// variants run one wave per tile (four pixels per lane) forward and
// backward; kRenderMixed pairs the 16x256 forward with the strip backward.
enum RenderVariant {
  kRender16x256 = 0,
  kRender16x128 = 1,
  kRender8x64 = 2,
  kRender8x128 = 3,
  kRenderStrips64 = 4,
  kRenderStrips128 = 5,
  kRenderMixed = 6,
  kRenderVariants = 7
};
void launch_render_fwd(const RenderArgs& a, hipStream_t s, int variant = kRender16x256);
// Multi-view: order the a.num_tiles tiles by decreasing list length (a
// binned estimate of their cost) so the longest start first.  bucket: one
// byte per tile, hist: 512 words of scratch.
void launch_tile_order(const RenderArgs& a, uint32_t* order, uint8_t* bucket, uint32_t* hist, hipStream_t s,
                       bool hist_zeroed = false);
// The same with a.tile_work (written by the strip forward) as each tile's
// cost: the backward's exact list length.
void launch_tile_order_bwd(const RenderArgs& a, uint32_t* order, uint8_t* bucket, uint32_t* hist, hipStream_t s,
                           bool hist_zeroed = false);
void launch_render_bwd(const RenderArgs& a, hipStream_t s, int variant = kRender16x256);

// ---- geometry backward ----
struct GeometryBwdArgs {
  int P;
  int W, H;
  float tanfovx, tanfovy, focal_x, focal_y, scale_modifier;
  int sh_degree, sh_coeffs;
  const float* means3D;
  const float* scales;
  const float* rotations;
  const float* shs;
  const float* cov3D_precomp;
  const float* viewmatrix;
  const float* projmatrix;
  const float* campos;
  const int32_t* radii;
  const float* gacc;
  // outputs (each may be null when the caller does not want it)
  float* dL_dmeans2D;    // (P, 3)
  float* dL_dcolors;     // (P, 3)
  float* dL_dopacity;    // (P, 1)
  float* dL_dconic;      // (P, 4)
  float* dL_dmeans3D;    // (P, 3)
  float* dL_dcov3D;      // (P, 6)
  float* dL_dsh;         // (P, M, 3)
  float* dL_dscales;     // (P, 3)
  float* dL_drotations;  // (P, 4)
};
void launch_geometry_bwd(const GeometryBwdArgs& a, hipStream_t s);

// ---- mesh binding ----
struct BindArgs {
  int F, G;                 // faces, Gaussians per face
  const float* verts;       // (V, 3)
  const int32_t* faces;     // (F, 3)
  const float* bc;          // (G, 3) shared, or (F*G, 3) per Gaussian
  int bc_per_gaussian;
  // stage 3 parameters (may be null)
  const float* rot2d;       // (F*G, 2)
  const float* scale2d;     // (F*G, 2) log-scales
  float thin_z;
  // stage 2 parameters
  const float* k;           // device scalar tanh(scale_factor) * max_scale
  float rad_base, flat_scale;
  int adaptive;
  // forward outputs
  float* xyz;               // (F*G, 3)
  float* rot;               // (F*G, 4)
  float* scale;             // (F*G, 3)
  float* cov;               // (F*G, 6)
  float* Lmat;              // (F, 9) stage-2 triangle matrices (written by the forward)
  // backward inputs / outputs
  const float* dxyz;
  const float* drot;
  const float* dscale;
  const float* dcov;
  float* dverts;            // (V, 3), accumulated
  float* drot2d;
  float* dscale2d;
  float* dk;                // device scalar, accumulated
};
void launch_bind_fwd(const BindArgs& a, hipStream_t s);
void launch_bind_cov_fwd(const BindArgs& a, hipStream_t s);
void launch_bind_bwd(const BindArgs& a, hipStream_t s);
void launch_faces_in_frustum(int F, const float* verts, const int32_t* faces, const float* proj, uint8_t* mask,
                             hipStream_t s);

// ---- fused L1 + SSIM loss ----
size_t dssim_partials(int C, int H, int W);
// C planes in C / group loss groups of `group` consecutive planes each.
// ystats: the target's blurred maps from launch_dssim_target, or null.
void launch_dssim_forward(const float* x, const float* y, const float* ystats, int C, int group, int H, int W,
                          const float* w11, float lam, float* coef, float* partials, float* out, hipStream_t s);
// stats (2 C planes): G*y, then G*y^2.
void launch_dssim_target(const float* y, int C, int H, int W, const float* w11, float* stats, hipStream_t s);
void launch_dssim_backward(const float* x, const float* y, const float* coef, int C, int group, int H, int W,
                           const float* w11, float lam, const float* dloss, float* dx, hipStream_t s);

// ---- scene / frame conversions ----
constexpr int kMaxRowFloats = 128;   // properties per PLY vertex row
constexpr int kMaxChannels = 128;
struct ChannelMap {
  int channels;                      // entries used in col
  int col[kMaxChannels];             // source index per destination (-1: zero)
};
// params[c][g] = rows[g][map.col[c]] for c < map.channels (59 scene channels)
void launch_rows_to_channels(const float* rows, int64_t P, int row_len, const ChannelMap& map, float* params,
                             hipStream_t s);
// rows[g][j] = params[map.col[j]][g] for j < row_len
void launch_channels_to_rows(const float* params, int64_t P, int row_len, const ChannelMap& map, float* rows,
                             hipStream_t s);
// (3, H*W) float -> H*W*3 bytes of uint8(clamp(x, 0, 1) * 255)
void launch_rgb8(const float* chw, int64_t plane, uint8_t* hwc, hipStream_t s);

}  // namespace dmgs
