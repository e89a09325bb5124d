// Per-Gaussian forward preprocessing for gfx950: near-plane cull, projection,
// 3D covariance, EWA 2D covariance and conic, screen radius, tile rectangle,
// SH -> RGB, and the 32-bit depth sort key.
//
// Semantics follow the reference rasterizer's preprocess step [U]
// (diff-gaussian-rasterization, reached from gaussian_renderer/__init__.py:86-94):
// near cull at z_view <= 0.2, p_w = 1 / (w + 1e-7), EWA with the 1.3 tan(fov)
// clamp and +0.3 dilation, radius = ceil(3 sqrt(lambda_max)), SH colour + 0.5
// clamped at 0 (sh_utils.py:26-110 holds the same polynomials).
//
// Output layout is this library's own: one 48-byte "splat" record per
// Gaussian (see dmgs_common.h) instead of separate xy / conic / rgb arrays.
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

__constant__ float kSH_C0 = 0.28209479177387814f;
__constant__ float kSH_C1 = 0.4886025119029199f;
__constant__ float kSH_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                                -1.0925484305920792f, 0.5462742152960396f};
__constant__ float kSH_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                                0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                                -0.5900435899266435f};

// RGB from SH coefficients laid out (M, 3) per Gaussian; returns the raw
// (unclamped) value + 0.5.
__device__ __forceinline__ float3 sh_to_rgb(int deg, const float (&sh)[48], float3 dir) {
  const float x = dir.x, y = dir.y, z = dir.z;
  float3 r;
  r.x = kSH_C0 * sh[0];
  r.y = kSH_C0 * sh[1];
  r.z = kSH_C0 * sh[2];
  if (deg > 0) {
    const float b1 = -kSH_C1 * y, b2 = kSH_C1 * z, b3 = -kSH_C1 * x;
    r.x += b1 * sh[3] + b2 * sh[6] + b3 * sh[9];
    r.y += b1 * sh[4] + b2 * sh[7] + b3 * sh[10];
    r.z += b1 * sh[5] + b2 * sh[8] + b3 * sh[11];
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      const float b[5] = {kSH_C2[0] * xy, kSH_C2[1] * yz, kSH_C2[2] * (2.0f * zz - xx - yy),
                          kSH_C2[3] * xz, kSH_C2[4] * (xx - yy)};
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        r.x += b[k] * sh[3 * (4 + k) + 0];
        r.y += b[k] * sh[3 * (4 + k) + 1];
        r.z += b[k] * sh[3 * (4 + k) + 2];
      }
      if (deg > 2) {
        const float c[7] = {kSH_C3[0] * y * (3.0f * xx - yy), kSH_C3[1] * xy * z,
                            kSH_C3[2] * y * (4.0f * zz - xx - yy),
                            kSH_C3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy),
                            kSH_C3[4] * x * (4.0f * zz - xx - yy), kSH_C3[5] * z * (xx - yy),
                            kSH_C3[6] * x * (xx - 3.0f * yy)};
#pragma unroll
        for (int k = 0; k < 7; ++k) {
          r.x += c[k] * sh[3 * (9 + k) + 0];
          r.y += c[k] * sh[3 * (9 + k) + 1];
          r.z += c[k] * sh[3 * (9 + k) + 2];
        }
      }
    }
  }
  r.x += 0.5f;
  r.y += 0.5f;
  r.z += 0.5f;
  return r;
}

// Sigma = R(q) S S^T R(q)^T with q = (r, x, y, z) used as given (not normalised).
__device__ __forceinline__ void cov3d_from(float3 s, float mod, float4 q, float* c) {
  const float sx = mod * s.x, sy = mod * s.y, sz = mod * s.z;
  const float r = q.x, x = q.y, y = q.z, z = q.w;
  const float R00 = 1.f - 2.f * (y * y + z * z), R01 = 2.f * (x * y - r * z), R02 = 2.f * (x * z + r * y);
  const float R10 = 2.f * (x * y + r * z), R11 = 1.f - 2.f * (x * x + z * z), R12 = 2.f * (y * z - r * x);
  const float R20 = 2.f * (x * z - r * y), R21 = 2.f * (y * z + r * x), R22 = 1.f - 2.f * (x * x + y * y);
  // M = R diag(s)
  const float M00 = R00 * sx, M01 = R01 * sy, M02 = R02 * sz;
  const float M10 = R10 * sx, M11 = R11 * sy, M12 = R12 * sz;
  const float M20 = R20 * sx, M21 = R21 * sy, M22 = R22 * sz;
  c[0] = M00 * M00 + M01 * M01 + M02 * M02;
  c[1] = M00 * M10 + M01 * M11 + M02 * M12;
  c[2] = M00 * M20 + M01 * M21 + M02 * M22;
  c[3] = M10 * M10 + M11 * M11 + M12 * M12;
  c[4] = M10 * M20 + M11 * M21 + M12 * M22;
  c[5] = M20 * M20 + M21 * M21 + M22 * M22;
}

// EWA splatting: cov2D = (J W) Sigma (J W)^T + 0.3 I, returns (a, b, c).
__device__ __forceinline__ float3 ewa_cov2d(float3 t, float fx, float fy, float tanx, float tany,
                                            const float* c, const float* vm) {
  const float limx = 1.3f * tanx, limy = 1.3f * tany;
  const float txtz = t.x / t.z, tytz = t.y / t.z;
  const float tx = fminf(limx, fmaxf(-limx, txtz)) * t.z;
  const float ty = fminf(limy, fmaxf(-limy, tytz)) * t.z;
  const float iz = 1.0f / t.z;
  const float j00 = fx * iz, j02 = -(fx * tx) * (iz * iz);
  const float j11 = fy * iz, j12 = -(fy * ty) * (iz * iz);
  // rows of M = J W, with W[r][k] = vm[4 k + r]
  const float m00 = j00 * vm[0] + j02 * vm[2], m01 = j00 * vm[4] + j02 * vm[6], m02 = j00 * vm[8] + j02 * vm[10];
  const float m10 = j11 * vm[1] + j12 * vm[2], m11 = j11 * vm[5] + j12 * vm[6], m12 = j11 * vm[9] + j12 * vm[10];
  // Sigma m_row
  const float s0x = c[0] * m00 + c[1] * m01 + c[2] * m02;
  const float s0y = c[1] * m00 + c[3] * m01 + c[4] * m02;
  const float s0z = c[2] * m00 + c[4] * m01 + c[5] * m02;
  const float s1x = c[0] * m10 + c[1] * m11 + c[2] * m12;
  const float s1y = c[1] * m10 + c[3] * m11 + c[4] * m12;
  const float s1z = c[2] * m10 + c[4] * m11 + c[5] * m12;
  float3 r;
  r.x = m00 * s0x + m01 * s0y + m02 * s0z + 0.3f;
  r.y = m10 * s0x + m11 * s0y + m12 * s0z;
  r.z = m10 * s1x + m11 * s1y + m12 * s1z + 0.3f;
  return r;
}

__global__ __launch_bounds__(256) void preprocess_fwd_kernel(PreprocessArgs a) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.P) return;
  a.radii[idx] = 0;
  a.tiles_touched[idx] = 0;
  a.depth_keys[idx] = 0xFFFFFFFFu;
  a.order[idx] = (uint32_t)idx;

  const float* vm = a.viewmatrix;
  const float* pm = a.projmatrix;
  const float3 p = make_float3(a.means3D[3 * idx], a.means3D[3 * idx + 1], a.means3D[3 * idx + 2]);
  const float3 t = xform_point_4x3(p, vm);
  if (t.z <= 0.2f) return;

  const float4 ph = xform_point_4x4(p, pm);
  const float pw = 1.0f / (ph.w + 0.0000001f);
  const float3 pp = make_float3(ph.x * pw, ph.y * pw, ph.z * pw);

  float c3[6];
  if (a.cov3D_precomp) {
#pragma unroll
    for (int k = 0; k < 6; ++k) c3[k] = a.cov3D_precomp[6 * idx + k];
  } else {
    const float3 s = make_float3(a.scales[3 * idx], a.scales[3 * idx + 1], a.scales[3 * idx + 2]);
    const float4 q = reinterpret_cast<const float4*>(a.rotations)[idx];
    cov3d_from(s, a.scale_modifier, q, c3);
  }
  const float3 cov = ewa_cov2d(t, a.focal_x, a.focal_y, a.tanfovx, a.tanfovy, c3, vm);
  const float det = cov.x * cov.z - cov.y * cov.y;
  if (det == 0.0f) return;
  const float det_inv = 1.f / det;
  const float3 conic = make_float3(cov.z * det_inv, -cov.y * det_inv, cov.x * det_inv);
  const float mid = 0.5f * (cov.x + cov.z);
  const float disc = sqrtf(fmaxf(0.1f, mid * mid - det));
  const float lambda1 = mid + disc, lambda2 = mid - disc;
  const float radius = ceilf(3.f * sqrtf(fmaxf(lambda1, lambda2)));
  const float2 pix = make_float2(ndc_to_pix(pp.x, a.W), ndc_to_pix(pp.y, a.H));
  int2 rmin, rmax;
  tile_rect(pix, (int)radius, a.grid_x, a.grid_y, rmin, rmax);
  const int area = (rmax.x - rmin.x) * (rmax.y - rmin.y);
  if (area == 0) return;

  float3 rgb;
  if (a.colors_precomp) {
    rgb = make_float3(a.colors_precomp[3 * idx], a.colors_precomp[3 * idx + 1], a.colors_precomp[3 * idx + 2]);
  } else {
    float3 dir = make_float3(p.x - a.campos[0], p.y - a.campos[1], p.z - a.campos[2]);
    const float inv_len = 1.0f / sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
    dir.x *= inv_len;
    dir.y *= inv_len;
    dir.z *= inv_len;
    float sh[48];
    load_sh16(a.shs + (size_t)idx * a.sh_coeffs * 3, a.sh_coeffs, sh);
    rgb = sh_to_rgb(a.sh_degree, sh, dir);
    rgb.x = fmaxf(rgb.x, 0.0f);
    rgb.y = fmaxf(rgb.y, 0.0f);
    rgb.z = fmaxf(rgb.z, 0.0f);
  }
  const float opacity = a.opacities[idx];
  float4* sp = a.splats + (size_t)DMGS_SPLAT_F4 * idx;
  sp[0] = make_float4(pix.x, pix.y, conic.x, conic.y);
  sp[1] = make_float4(conic.z, opacity, rgb.x, rgb.y);
  sp[2] = make_float4(rgb.z, t.z, __int_as_float((int)radius), 0.0f);
  a.radii[idx] = (int)radius;
  a.tiles_touched[idx] = (uint32_t)area;
  a.depth_keys[idx] = __float_as_uint(t.z);
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
