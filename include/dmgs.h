/*
 * dmgs.h — C-ABI of the MI355X (gfx950) Gaussian rasterizer and mesh binding.
 *
 * Plain pointers and sizes only: every tensor argument is a device pointer to
 * contiguous float32 / int32 / uint32 memory that the caller owns, and every
 * call is enqueued on the caller's hipStream_t.  The library allocates
 * nothing: each workspace is sized by a *_workspace_size() query and passed
 * in as raw device memory (two-phase, like the reference's resize callbacks),
 * 256-byte aligned (hipMalloc and torch allocations are).
 *
 * Every entry point returns 0 on success, DMGS_EINVAL (-1) for an invalid
 * argument (nothing is launched) or DMGS_EHIP (-2) when a HIP call or launch
 * failed; dmgs_last_error() then returns a message (per calling thread).
 *
 * Reference interface being replaced: the pybind module `_C` of the
 * un-vendored submodule diff-gaussian-rasterization (reference .gitmodules:4-6),
 * reached from gaussian_renderer/__init__.py:14,86-94 through
 * GaussianRasterizer / _RasterizeGaussians.  Its Python entry points map as:
 *   _C.rasterize_gaussians          -> dmgs_forward_geometry + host read of the
 *                                      count + dmgs_forward_render
 *   _C.rasterize_gaussians_backward -> dmgs_backward
 *   _C.mark_visible                 -> dmgs_mark_visible
 * The mesh binding replaces the PyTorch code of
 * scene/gaussian_geo_model_finetune.py:396-482 and
 * scene/gaussian_geo_model_mlp_flex.py:258-385.
 */
#ifndef DMGS_H
#define DMGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* dmgs_stream_t; /* == hipStream_t */

#define DMGS_OK 0
#define DMGS_EINVAL (-1)
#define DMGS_EHIP (-2)

/* Mirrors GaussianRasterizationSettings (gaussian_renderer/__init__.py:36-49). */
typedef struct dmgs_settings {
  int32_t image_width;
  int32_t image_height;
  float tanfovx;
  float tanfovy;
  float scale_modifier;
  int32_t sh_degree;   /* active SH degree D */
  int32_t sh_coeffs;   /* M: SH coefficients per channel in `shs` (0 with precomputed colours) */
  int32_t prefiltered; /* accepted for API parity; culling is identical either way */
  int32_t debug;       /* synchronise and check for errors after every kernel */
  const float* viewmatrix; /* device, 16 floats, reference storage order (world_view_transform) */
  const float* projmatrix; /* device, 16 floats (full_proj_transform) */
  const float* campos;     /* device, 3 floats */
  const float* bg;         /* device, 3 floats */
} dmgs_settings;

/* Message for the last failing call on this thread ("" if none). */
const char* dmgs_last_error(void);
/* Library version string. */
const char* dmgs_version(void);

/* ---- workspace sizing (bytes) ------------------------------------------- */
/* Per-Gaussian forward state kept until backward (the reference geomBuffer). */
size_t dmgs_geometry_workspace_size(int64_t P);
/* Per-(Gaussian, tile) sorted lists kept until backward (binningBuffer). */
size_t dmgs_binning_workspace_size(int64_t K);
/* Per-pixel and per-tile state kept until backward (imgBuffer). */
size_t dmgs_image_workspace_size(int32_t width, int32_t height);
/* Per-Gaussian gradient accumulator used only inside dmgs_backward. */
size_t dmgs_backward_workspace_size(int64_t P);
/* Byte offsets of the sub-buffers inside the workspaces, for inspection:
 * geometry: [0] splats (P x 12 f32), [1] depth keys (P u32), [2] depth order (P u32),
 *           [3] tiles touched (P u32), [4] inclusive offsets in depth order (P u32)
 * binning:  [5] tile ids, sorted (K u32), [6] Gaussian ids, sorted (K u32)
 * image:    [7] ranges (tiles x 2 u32), [8] final T (H*W f32), [9] n_contrib (H*W u32) */
int dmgs_workspace_offsets(int64_t P, int64_t K, int32_t width, int32_t height, size_t offsets[10]);

/* ---- forward -------------------------------------------------------------- */
/* Preprocess every Gaussian (cull, project, EWA conic, radius, tile count,
 * SH -> RGB), sort the Gaussians by depth and prefix-sum their tile counts.
 * Exactly one of {shs, colors_precomp} and one of {(scales, rotations),
 * cov3D_precomp} must be non-null.  radii: (P) int32 output.  The number of
 * (Gaussian, tile) instances K is written to *num_rendered_dev (device u32).
 * Replaces the first half of _C.rasterize_gaussians [U]. */
int dmgs_forward_geometry(const dmgs_settings* s, int64_t P, const float* means3D, const float* opacities,
                          const float* shs, const float* colors_precomp, const float* scales,
                          const float* rotations, const float* cov3D_precomp, int32_t* radii, void* geom_ws,
                          uint32_t* num_rendered_dev, dmgs_stream_t stream);

/* Duplicate into (tile, Gaussian) instances in depth order, stably sort by
 * tile and find the per-tile ranges.  K must be the count written by
 * dmgs_forward_geometry.  (duplicateWithKeys + SortPairs + identifyTileRanges [U]) */
int dmgs_forward_binning(const dmgs_settings* s, int64_t P, int64_t K, const void* geom_ws, void* binning_ws,
                         void* image_ws, dmgs_stream_t stream);
/* Front-to-back blend of every tile.  out_color: (3, H, W) float32.
 * (renderCUDA forward [U]) */
int dmgs_render_forward(const dmgs_settings* s, int64_t P, int64_t K, const void* geom_ws, const void* binning_ws,
                        void* image_ws, float* out_color, dmgs_stream_t stream);
/* dmgs_forward_binning followed by dmgs_render_forward.
 * Replaces the second half of _C.rasterize_gaussians [U]. */
int dmgs_forward_render(const dmgs_settings* s, int64_t P, int64_t K, void* geom_ws, void* binning_ws,
                        void* image_ws, float* out_color, dmgs_stream_t stream);

/* ---- backward ------------------------------------------------------------- */
/* Blend backward (renderCUDA backward [U]): clears the accumulator in bwd_ws
 * and adds per-Gaussian dL/d{mean2D, conic, opacity, colour} into it. */
int dmgs_render_backward(const dmgs_settings* s, int64_t P, int64_t K, const void* geom_ws, const void* binning_ws,
                         const void* image_ws, void* bwd_ws, const float* dL_dout, dmgs_stream_t stream);
/* Per-Gaussian backward from the accumulator (computeCov2DCUDA + preprocessCUDA
 * backward [U], fused); outputs as for dmgs_backward. */
int dmgs_geometry_backward(const dmgs_settings* s, int64_t P, const float* means3D, const float* shs,
                           const float* colors_precomp, const float* scales, const float* rotations,
                           const float* cov3D_precomp, const int32_t* radii, const void* bwd_ws, float* dL_dmeans2D,
                           float* dL_dcolors, float* dL_dopacity, float* dL_dconic, float* dL_dmeans3D,
                           float* dL_dcov3D, float* dL_dsh, float* dL_dscales, float* dL_drotations,
                           dmgs_stream_t stream);
/* dmgs_render_backward followed by dmgs_geometry_backward: reverse-traversal
 * blend backward + fused cov2D / projection / SH / cov3D backward.
 * Inputs are the forward's inputs and workspaces; dL_dout is
 * (3, H, W).  Outputs (caller-allocated, fully overwritten):
 *   dL_dmeans2D (P,3) NDC units, z = 0; dL_dcolors (P,3); dL_dopacity (P,1);
 *   dL_dconic (P,4) (.z = 0); dL_dmeans3D (P,3); dL_dcov3D (P,6);
 *   dL_dsh (P,M,3) (null iff shs is null); dL_dscales (P,3), dL_drotations (P,4)
 *   (null iff cov3D_precomp is given).
 * Replaces _C.rasterize_gaussians_backward [U]. */
int dmgs_backward(const dmgs_settings* s, int64_t P, int64_t K, const float* means3D, const float* shs,
                  const float* colors_precomp, const float* scales, const float* rotations,
                  const float* cov3D_precomp, const int32_t* radii, const void* geom_ws, const void* binning_ws,
                  const void* image_ws, void* bwd_ws, const float* dL_dout, float* dL_dmeans2D,
                  float* dL_dcolors, float* dL_dopacity, float* dL_dconic, float* dL_dmeans3D, float* dL_dcov3D,
                  float* dL_dsh, float* dL_dscales, float* dL_drotations, dmgs_stream_t stream);

/* ---- utilities -------------------------------------------------------------- */
/* present[i] = 1 if Gaussian i lies in front of the near plane (z_view > 0.2).
 * Replaces _C.mark_visible [U] (GaussianRasterizer.markVisible). */
int dmgs_mark_visible(int64_t P, const float* means3D, const float* viewmatrix, uint8_t* present,
                      dmgs_stream_t stream);

/* Stable LSD radix sort of (key, value) uint32 pairs on bits [0, bits).
 * (k0, v0) hold the input and receive the output; (k1, v1) are ping-pong
 * buffers of n elements; scratch holds dmgs_sort_scratch_size(n) bytes. */
size_t dmgs_sort_scratch_size(int64_t n);
int dmgs_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int32_t bits,
                    void* scratch, dmgs_stream_t stream);

/* Inclusive prefix sum of n uint32 (in == out allowed). */
size_t dmgs_scan_scratch_size(int64_t n);
int dmgs_inclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* scratch, dmgs_stream_t stream);

/* dmgs_inclusive_scan_u32 without a scratch argument: the block sums live in
 * a static device array of the library (1 MiB per GPU, held until the
 * process ends), so the call allocates nothing and takes no workspace.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  n <= 2^30 (larger n returns DMGS_EINVAL); the stream
 * must belong to the current device.  Calls that overlap on one GPU share the
 * array and race: serialise them. */
int dmgs_inclusive_scan_u32_noalloc(const uint32_t* in, uint32_t* out, int64_t n, dmgs_stream_t stream);

/* ---- mesh -> Gaussian binding ------------------------------------------------ */
/* Shared arguments: verts (V,3) float32, faces (F,3) int32, G Gaussians per
 * face in face-major order (Gaussian f*G + j), barycentrics bc either (G,3)
 * shared by every face (bc_per_gaussian = 0, as generate_barycentric_v2 gives,
 * geo/mesh_utils.py:15-40) or (F*G,3) (bc_per_gaussian = 1).
 *
 * Stage 3, replaces GaussianGeoModel.renew_gaussian + get_rotation /
 * get_scaling (scene/gaussian_geo_model_finetune.py:396-421, 445-482):
 * xyz (F*G,3); rot (F*G,4) = normalize(matrix_to_quaternion(R_t2w R_t(rot2d)));
 * scale (F*G,3) = (exp(scale2d), thin_z).  rot2d/rot and scale2d/scale are
 * optional in pairs. */
int dmgs_bind_forward(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G, const float* bc,
                      int32_t bc_per_gaussian, const float* rot2d, const float* scale2d, float thin_z, float* xyz,
                      float* rot, float* scale, dmgs_stream_t stream);
/* Gradients of the stage-3 binding: dverts (V,3) is ADDED to (clear it first);
 * drot2d, dscale2d (F*G,2) are written. */
int dmgs_bind_backward(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G, const float* bc,
                       int32_t bc_per_gaussian, const float* rot2d, const float* scale2d, const float* dxyz,
                       const float* drot, const float* dscale, float* dverts, float* drot2d, float* dscale2d,
                       dmgs_stream_t stream);
/* Stage 2, replaces GaussianGeoModel.renew_gaussian's means and
 * get_covariance_dyn (scene/gaussian_geo_model_mlp_flex.py:258-310, 370-385):
 * xyz (F*G,3); cov (F*G,6) = (R_t2w L k)(R_t2w L k)^T per face with
 * k = *k (device scalar, tanh(scale_factor) * max_scale) and L the affine
 * triangle matrix (no gradient), written to Lmat (F,9) for the backward;
 * rad_base scales |v1 - v0|, flat_scale is L[2][2] (spatial_lr_scale * 1e-6),
 * adaptive selects the affine M (else identity). */
int dmgs_bind_cov_forward(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G, const float* bc,
                          int32_t bc_per_gaussian, const float* k, float rad_base, float flat_scale, int32_t adaptive,
                          float* xyz, float* cov, float* Lmat, dmgs_stream_t stream);
/* Gradients of the stage-2 binding: dverts (V,3) and dk (device scalar,
 * dL/dk) are ADDED to. */
int dmgs_bind_cov_backward(int64_t V, const float* verts, int64_t F, const int32_t* faces, int32_t G,
                           const float* bc, int32_t bc_per_gaussian, const float* k, const float* Lmat,
                           const float* dxyz, const float* dcov, float* dverts, float* dk, dmgs_stream_t stream);
/* mask[f] = 1 if face f's centroid projects inside |ndc| < 1.05 with w > 0
 * (in_frustum, scene/gaussian_geo_model_finetune.py:33-48). */
int dmgs_faces_in_frustum(int64_t V, const float* verts, int64_t F, const int32_t* faces, const float* projmatrix,
                          uint8_t* mask, dmgs_stream_t stream);

/* ---- training objective ------------------------------------------------------ */
/* Fused (1 - lambda) * L1 + lambda * (1 - SSIM) between img and gt, both
 * (C, H, W) float32; SSIM as utils/loss_utils.py:17-62 (11x11 Gaussian window,
 * sigma 1.5, zero padding, C1 = 1e-4, C2 = 9e-4, mean over C*H*W), objective as
 * train_geo_stage3.py:122-124.  Forward writes (loss, l1, ssim) to loss_out
 * (3 device floats) and keeps per-pixel coefficients in ws for the backward;
 * backward writes dimg = dloss[0] * d(loss)/d(img).  H and W below 2^24, at
 * most 2^31 values per call (DMGS_EINVAL otherwise). */
size_t dmgs_dssim_l1_workspace_size(int32_t C, int32_t H, int32_t W);
int dmgs_dssim_l1_forward(const float* img, const float* gt, int32_t C, int32_t H, int32_t W, float lambda_dssim,
                          void* ws, float* loss_out, dmgs_stream_t stream);
int dmgs_dssim_l1_backward(const float* img, const float* gt, int32_t C, int32_t H, int32_t W, float lambda_dssim,
                           const void* ws, const float* dloss, float* dimg, dmgs_stream_t stream);
/* The same for G images of Cg channels each, stored back to back (G*Cg planes
 * of H x W; workspace dmgs_dssim_l1_workspace_size(G*Cg, H, W)): one loss per
 * image, (loss, l1, ssim) at loss_out[3g..3g+2]; the backward scales image
 * g's gradient by dloss[g].  Two launches for all G images (the per-view loss
 * of a batch of equally sized views). */
int dmgs_dssim_l1_forward_groups(const float* img, const float* gt, int32_t G, int32_t Cg, int32_t H, int32_t W,
                                 float lambda_dssim, void* ws, float* loss_out, dmgs_stream_t stream);
int dmgs_dssim_l1_backward_groups(const float* img, const float* gt, int32_t G, int32_t Cg, int32_t H, int32_t W,
                                  float lambda_dssim, const void* ws, const float* dloss, float* dimg,
                                  dmgs_stream_t stream);
/* Fixed targets: stats (2*C planes of H x W, float32) receives the target's
 * Gaussian-blurred maps G*gt (planes 0..C-1) and G*gt^2 (planes C..2C-1), the
 * two SSIM inputs that do not depend on the rendered image.  The forward
 * below then blurs three maps per plane instead of five; its results are
 * bit-identical to dmgs_dssim_l1_forward_groups' (the same loops compute the
 * maps).  No reference counterpart: utils/loss_utils.py:37-62 recomputes them
 * on every call. */
int dmgs_dssim_target_stats(const float* gt, int32_t C, int32_t H, int32_t W, float* stats, dmgs_stream_t stream);
int dmgs_dssim_l1_forward_groups_stats(const float* img, const float* gt, const float* gt_stats, int32_t G,
                                       int32_t Cg, int32_t H, int32_t W, float lambda_dssim, void* ws,
                                       float* loss_out, dmgs_stream_t stream);

/* ---- batched views --------------------------------------------------------
 * Replaces the per-view loop of the training / evaluation scripts
 * (train_geo_stage3.py:104-124: one render() + loss per camera, one
 * num_rendered host read-back per view) with one launch sequence over B views
 * of one scene, with no host synchronisation, that a caller may capture in a
 * HIP graph (kernels only: no memset or memcpy nodes).
 *
 * Scene parameters are component-major, params[c * P + g] for 59 channels:
 *   0-2 xyz, 3-5 log-scale (get_scaling = exp), 6-9 raw quaternion (w,x,y,z;
 *   get_rotation = normalise), 10 opacity logit (get_opacity = sigmoid),
 *   11 + 3k + c SH coefficient k of colour c (k = 0 is features_dc, 1..15
 *   features_rest), as scene/gaussian_model.py:26-41,95-118 store and
 *   activate them.  The activations run inside the kernels and the backward
 *   returns gradients of the raw channels (same layout), summed over views.
 * Every per-view quantity follows the single-view path exactly: images are
 * those of dmgs_forward for the same camera, and the gradients are the sum
 * of the per-view gradients chained through the activations.
 *
 * Views may have different resolutions; each owns the tile range
 * [tile_base, tile_base + grid_x*grid_y) and the pixel range
 * [pixel_base, pixel_base + width*height) of the packed per-pixel buffers.
 * Images are packed per view as (3, height, width) at offset 3*pixel_base. */
typedef struct dmgs_view {
  float viewmatrix[16]; /* world_view_transform, reference storage order */
  float projmatrix[16]; /* full_proj_transform */
  float campos[4];      /* xyz, then padding */
  float bg[4];          /* rgb, then padding */
  float tanfovx, tanfovy;
  float focal_x, focal_y;   /* filled by dmgs_batched_prepare_views */
  int32_t width, height;
  int32_t grid_x, grid_y;   /* filled */
  int32_t tile_base;        /* filled */
  int32_t cell_base;        /* filled: first binning cell (4x4 tiles) of the view */
  int64_t pixel_base;       /* filled */
} dmgs_view;

/* Host-side: fill the derived fields of B host records and return the tile,
 * binning-cell and pixel totals of the batch.  The batched lists bin splats
 * into cells of 4x4 tiles (64x64 pixels), each instance's key carrying a
 * mask of the cell's tiles it covers; a tile's list is its cell's list
 * filtered by its bit, i.e. exactly the reference per-tile list.  At most
 * 65536 cells per batch. */
int dmgs_batched_prepare_views(dmgs_view* views, int32_t B, int32_t* total_tiles, int32_t* total_cells,
                               int32_t* max_view_cells, int64_t* total_pixels);
/* Host-side: tile_view[t] = view owning global tile t (host array of total_tiles). */
int dmgs_batched_tile_view(const dmgs_view* views, int32_t B, int32_t* tile_view, int32_t total_tiles);

/* Workspaces of a batch (device memory, caller-owned):
 *   state   K-independent: per (view, Gaussian) splats, keys, radii, scan;
 *           per-tile ranges; per-pixel blend state.  Kept until backward.
 *   lists   the (Gaussian, tile) instance lists for up to `capacity`
 *           instances and their sort scratch.  Kept until backward.
 *   bwd_ws  per (view, Gaussian) gradient accumulator (backward only).  A
 *           forward whose BLEND stage runs with bwd_ws set clears it (the
 *           blend work-groups zero it after their blend work, so no clearing
 *           pass runs per step); a backward accumulates into it, so every
 *           backward needs a forward BLEND since the previous backward. */
size_t dmgs_batched_state_size(int32_t B, int64_t P, int32_t total_tiles, int64_t total_pixels);
size_t dmgs_batched_lists_size(int64_t capacity);
size_t dmgs_batched_backward_size(int32_t B, int64_t P);
/* Byte offsets inside `state`: [0] radii (B*P int32, view-major), [1] instance
 * count of the last forward (uint32), [2] cell ranges (uint2 per global
 * cell), [3] final_T, [4] last contributor (uint32 per packed pixel: one past
 * the position of its last contributing instance in its cell's list),
 * [5] splat records, [8] status word (uint32; bit 0: a visible depth was too
 * large for cell_mode 1, see dmgs_batch); inside `lists`: [6] sorted keys
 * (cell | tile mask << 16), [7] the entry ids beside them. */
int dmgs_batched_offsets(int32_t B, int64_t P, int32_t total_tiles, int64_t total_pixels, int64_t capacity,
                         size_t offsets[9]);

typedef struct dmgs_batch {
  const dmgs_view* views;   /* device copy of the B prepared records */
  const int32_t* tile_view; /* device copy of the tile -> view map */
  int32_t num_views;        /* B */
  int32_t sh_degree;        /* active SH degree (coefficients are always 16 per colour) */
  int64_t num_gaussians;    /* P */
  float scale_modifier;
  int32_t variant;          /* blend kernel shape 0..6 (same results; picked by an autotuner) */
  int32_t total_tiles;
  int32_t total_cells;
  int64_t total_pixels;
  int64_t capacity;         /* (splat, cell) instances the lists hold */
  void* state;
  void* lists;
  void* bwd_ws;
  /* List order.  0: instances keyed by global cell id (two sort passes up to
   * 65536 cells).  1: view-major — the depth sort also orders by view and
   * instances are keyed by (view, cell within the view), one sort pass when
   * every view has at most 256 cells; needs num_views <= 16, max_view_cells
   * <= 4096 and visible depths below 2^27 (a larger one sets bit 0 of the
   * status word; its order among such far splats is then not guaranteed:
   * re-run with cell_mode 0). */
  int32_t cell_mode;
  int32_t max_view_cells; /* largest cell count of one view (dmgs_batched_prepare_views) */
  /* Optional (3 x P floats, or NULL): densification statistics that a
   * backward with the GEOMETRY stage accumulates, per Gaussian and per view
   * where it is visible, in view order — the consumer of
   * gaussian_model.py:405-407 and train.py:116-118:
   *   dens[g]       += |dL/dmean2D (NDC) of that view|_2   (xyz_gradient_accum)
   *   dens[P + g]   += 1                                   (denom)
   *   dens[2P + g]   = max(dens[2P + g], radius in that view)  (max_radii2D)
   * The caller zeroes it when it resets the statistics. */
  float* dens;
} dmgs_batch;
/* Byte offsets inside `lists` of the sorted keys and entry ids that a forward
 * of this batch leaves (they depend on cell_mode and max_view_cells). */
int dmgs_batched_list_offsets(const dmgs_batch* batch, size_t* keys, size_t* ids);

/* Number of low key bits the cell-list sort of this batch covers (the cell
 * id in cell_mode 0, the cell within the view in mode 1; never the tile-mask
 * bits above bit 15), or a negative error code.  Host-side query. */
int dmgs_batched_cell_sort_bits(const dmgs_batch* batch);

/* Stage mask of the batched calls.  Forward: GEOMETRY = preprocess, depth
 * sort and the scan of tile counts (afterwards the instance count is at
 * offsets[1]); BINNING = duplicate, tile sort, ranges; BLEND = the blend.
 * Backward: BLEND = blend backward into bwd_ws, GEOMETRY = the per-Gaussian
 * backward into grads.  Stages run in that order within a call; a caller
 * may split a call (e.g. read the count after GEOMETRY and size the lists,
 * or bracket the blend for timing). */
#define DMGS_STAGE_GEOMETRY 1
#define DMGS_STAGE_BINNING 2
#define DMGS_STAGE_BLEND 4
#define DMGS_STAGES_ALL 7
/* When the count exceeds `capacity` the lists are truncated (the images are
 * then wrong, nothing is accessed out of bounds): callers read the count
 * back after the fact and re-plan. */
int dmgs_batched_forward(const dmgs_batch* batch, const float* params, float* images, int32_t stages,
                         dmgs_stream_t stream);
/* Gradients of the 59 raw channels (59 x P, overwritten) from dL/dimages
 * (packed like images), using the state and lists the forward left. */
int dmgs_batched_backward(const dmgs_batch* batch, const float* params, const float* dL_dimages, float* grads,
                          int32_t stages, dmgs_stream_t stream);

/* ---- Scene and frame conversions ---- */
/* PLY vertex rows (P x row_floats float32, row-major, as the 3DGS writer of
 * scene/gaussian_model.py:191-210 lays them out) -> the 59-channel
 * component-major scene parameters of the batched path.  channel_cols[c] is
 * the row column holding channel c (xyz 0-2, log-scale 3-5, raw quaternion
 * 6-9, opacity logit 10, SH coefficient k of colour c at 11 + 3k + c), or -1
 * for a channel the file lacks (written as 0; gaussian_model.py:214-250 reads
 * the same properties by name).  row_floats <= 128. */
int dmgs_ply_rows_to_params(const float* rows, int64_t P, int32_t row_floats, const int32_t channel_cols[59],
                            float* params, dmgs_stream_t stream);
/* The reverse: rows[g][j] = params[column_channels[j]][g], 0 where it is -1
 * (the writer's normals). */
int dmgs_params_to_ply_rows(const float* params, int64_t P, int32_t row_floats, const int32_t* column_channels,
                            float* rows, dmgs_stream_t stream);
/* A (3, H, W) float image -> H x W x 3 bytes of uint8(clamp(x, 0, 1) * 255),
 * truncating: the frame the viewer protocol sends (train.py:60,
 * gaussian_renderer/network_gui.py:51-55). */
int dmgs_image_rgb8(const float* chw, int32_t H, int32_t W, uint8_t* hwc, dmgs_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DMGS_H */
