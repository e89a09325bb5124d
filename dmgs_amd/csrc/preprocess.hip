// Per-Gaussian forward preprocessing for gfx950: near-plane cull, projection,
// 3D covariance, EWA 2D covariance and conic, screen radius, tile rectangle,
// SH -> RGB, and the 32-bit depth sort key.
//
// Semantics follow the reference rasterizer's preprocess step [U]
// (diff-gaussian-rasterization, reached from gaussian_renderer/__init__.py:86-94):
// near cull at z_view <= 0.2, p_w = 1 / (w + 1e-7), EWA with the 1.3 tan(fov)
// clamp and +0.3 dilation, radius = ceil(3 sqrt(lambda_max)), SH colour + 0.5
// clamped at 0 (sh_utils.py:26-110 holds the same polynomials).  The
// quaternion is used as given (the caller normalises).
//
// Output layout is this library's own: one 48-byte "splat" record per
// Gaussian (see dmgs_common.h) instead of separate xy / conic / rgb arrays.
#include "dmgs_common.h"
#include "dmgs_kernels.h"
#include "gs_math.h"

namespace dmgs {

__global__ __launch_bounds__(256) void preprocess_fwd_kernel(PreprocessArgs a) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.P) return;
  a.radii[idx] = 0;
  a.tiles_touched[idx] = 0;
  a.depth_keys[idx] = 0xFFFFFFFFu;
  a.order[idx] = (uint32_t)idx;

  const float3 p = make_float3(a.means3D[3 * idx], a.means3D[3 * idx + 1], a.means3D[3 * idx + 2]);
  float c3[6];
  if (a.cov3D_precomp) {
#pragma unroll
    for (int k = 0; k < 6; ++k) c3[k] = a.cov3D_precomp[6 * idx + k];
  } else {
    const float m = a.scale_modifier;
    const float3 s = make_float3(m * a.scales[3 * idx], m * a.scales[3 * idx + 1], m * a.scales[3 * idx + 2]);
    float R[9];
    quat_to_R(reinterpret_cast<const float4*>(a.rotations)[idx], R);
    cov3d_from_sR(s, R, c3);
  }
  const ViewParams v = view_params(a.viewmatrix, a.projmatrix, a.campos, a.W, a.H, a.grid_x, a.grid_y, a.focal_x,
                                   a.focal_y, a.tanfovx, a.tanfovy);
  SplatGeom o;
  if (!project_splat(p, c3, v, o)) return;

  float3 rgb;
  if (a.colors_precomp) {
    rgb = make_float3(a.colors_precomp[3 * idx], a.colors_precomp[3 * idx + 1], a.colors_precomp[3 * idx + 2]);
  } else {
    float3 dir = make_float3(p.x - v.campos[0], p.y - v.campos[1], p.z - v.campos[2]);
    const float inv_len = 1.0f / sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
    dir = make_float3(dir.x * inv_len, dir.y * inv_len, dir.z * inv_len);
    float sh[48];
    load_sh16(a.shs + (size_t)idx * a.sh_coeffs * 3, a.sh_coeffs, sh);
    rgb = sh_to_rgb(a.sh_degree, sh, dir);
    rgb = make_float3(fmaxf(rgb.x, 0.0f), fmaxf(rgb.y, 0.0f), fmaxf(rgb.z, 0.0f));
  }
  float4* sp = a.splats + (size_t)DMGS_SPLAT_F4 * idx;
  sp[0] = make_float4(o.pix.x, o.pix.y, o.conic.x, o.conic.y);
  sp[1] = make_float4(o.conic.z, a.opacities[idx], rgb.x, rgb.y);
  sp[2] = make_float4(rgb.z, o.depth, __int_as_float(o.radius), 0.0f);
  a.radii[idx] = o.radius;
  a.tiles_touched[idx] = (uint32_t)o.tiles;
  a.depth_keys[idx] = __float_as_uint(o.depth);
}

__global__ __launch_bounds__(256) void mark_visible_kernel(int P, const float* __restrict__ means3D,
                                                         const float* __restrict__ viewmatrix,
                                                         uint8_t* __restrict__ present) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P) return;
  const float3 p = make_float3(means3D[3 * idx], means3D[3 * idx + 1], means3D[3 * idx + 2]);
  const float3 t = xform_point_4x3(p, viewmatrix);
  present[idx] = t.z > 0.2f ? 1 : 0;
}

void launch_preprocess_fwd(const PreprocessArgs& a, hipStream_t s) {
  if (a.P <= 0) return;
  hipLaunchKernelGGL(preprocess_fwd_kernel, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
}

void launch_mark_visible(int P, const float* means3D, const float* viewmatrix, uint8_t* present,
                         hipStream_t s) {
  if (P <= 0) return;
  hipLaunchKernelGGL(mark_visible_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, means3D, viewmatrix,
                     present);
}

}  // namespace dmgs
