// Per-Gaussian forward and backward math shared by the single-view and the
// multi-view kernels.  Semantics are the reference rasterizer's [U]
// (diff-gaussian-rasterization, reached from gaussian_renderer/__init__.py:86-94);
// the conventions are listed in preprocess.hip and geometry_bwd.hip.
#pragma once

#include "dmgs_common.h"

namespace dmgs {

static __constant__ float kSH_C0 = 0.28209479177387814f;
static __constant__ float kSH_C1 = 0.4886025119029199f;
static __constant__ float kSH_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                                -1.0925484305920792f, 0.5462742152960396f};
static __constant__ float kSH_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                                0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                                -0.5900435899266435f};

// Rotation matrix (row-major) of the quaternion (r, x, y, z) used as given.
__device__ __forceinline__ void quat_to_R(float4 q, float R[9]) {
  const float r = q.x, x = q.y, y = q.z, z = q.w;
  R[0] = 1.f - 2.f * (y * y + z * z); R[1] = 2.f * (x * y - r * z); R[2] = 2.f * (x * z + r * y);
  R[3] = 2.f * (x * y + r * z); R[4] = 1.f - 2.f * (x * x + z * z); R[5] = 2.f * (y * z - r * x);
  R[6] = 2.f * (x * z - r * y); R[7] = 2.f * (y * z + r * x); R[8] = 1.f - 2.f * (x * x + y * y);
}

// Sigma = R diag(s)^2 R^T as the 6-vector (xx, xy, xz, yy, yz, zz); s already includes the modifier.
__device__ __forceinline__ void cov3d_from_sR(float3 s, const float R[9], float c[6]) {
  const float M00 = R[0] * s.x, M01 = R[1] * s.y, M02 = R[2] * s.z;
  const float M10 = R[3] * s.x, M11 = R[4] * s.y, M12 = R[5] * s.z;
  const float M20 = R[6] * s.x, M21 = R[7] * s.y, M22 = R[8] * s.z;
  c[0] = M00 * M00 + M01 * M01 + M02 * M02;
  c[1] = M00 * M10 + M01 * M11 + M02 * M12;
  c[2] = M00 * M20 + M01 * M21 + M02 * M22;
  c[3] = M10 * M10 + M11 * M11 + M12 * M12;
  c[4] = M10 * M20 + M11 * M21 + M12 * M22;
  c[5] = M20 * M20 + M21 * M21 + M22 * M22;
}

// RGB from SH coefficients in registers ((M, 3) layout, first 16 coefficients);
// returns the raw value + 0.5 (the caller clamps at 0).
__device__ __forceinline__ float3 sh_to_rgb(int deg, const float (&sh)[48], float3 dir) {
  const float x = dir.x, y = dir.y, z = dir.z;
  float3 r = make_float3(kSH_C0 * sh[0], kSH_C0 * sh[1], kSH_C0 * sh[2]);
  if (deg > 0) {
    const float b1 = -kSH_C1 * y, b2 = kSH_C1 * z, b3 = -kSH_C1 * x;
    r.x += b1 * sh[3] + b2 * sh[6] + b3 * sh[9];
    r.y += b1 * sh[4] + b2 * sh[7] + b3 * sh[10];
    r.z += b1 * sh[5] + b2 * sh[8] + b3 * sh[11];
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      const float b[5] = {kSH_C2[0] * xy, kSH_C2[1] * yz, kSH_C2[2] * (2.0f * zz - xx - yy), kSH_C2[3] * xz,
                          kSH_C2[4] * (xx - yy)};
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

// Read-only kernel inputs seen through the constant address space: loads
// from a wave-uniform address become scalar loads (one per wave, through the
// scalar cache) instead of per-lane vector loads.
#define DMGS_CONST_AS __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const DMGS_CONST_AS T* const_ptr(const T* p) {
  return (const DMGS_CONST_AS T*)(p);
}

// One view of a camera, as the per-Gaussian kernels need it (values: with a
// uniform view they live in scalar registers).
struct ViewParams {
  float vm[16];
  float pm[16];
  float campos[3];
  int W, H, gx, gy;
  float fx, fy, tanx, tany;
};

__device__ __forceinline__ ViewParams view_params(const DMGS_CONST_AS ViewDesc& v) {
  ViewParams p;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    p.vm[k] = v.view[k];
    p.pm[k] = v.proj[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) p.campos[k] = v.campos[k];
  p.W = v.W;
  p.H = v.H;
  p.gx = v.grid_x;
  p.gy = v.grid_y;
  p.fx = v.focal_x;
  p.fy = v.focal_y;
  p.tanx = v.tanfovx;
  p.tany = v.tanfovy;
  return p;
}

// The same from the single-view kernels' separate camera arrays.
__device__ __forceinline__ ViewParams view_params(const float* viewmatrix, const float* projmatrix,
                                                  const float* campos, int W, int H, int gx, int gy, float fx,
                                                  float fy, float tanx, float tany) {
  ViewParams p;
  const DMGS_CONST_AS float* vm = const_ptr(viewmatrix);
  const DMGS_CONST_AS float* pm = const_ptr(projmatrix);
  const DMGS_CONST_AS float* cp = const_ptr(campos);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    p.vm[k] = vm[k];
    p.pm[k] = pm[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) p.campos[k] = cp[k];
  p.W = W;
  p.H = H;
  p.gx = gx;
  p.gy = gy;
  p.fx = fx;
  p.fy = fy;
  p.tanx = tanx;
  p.tany = tany;
  return p;
}

struct SplatGeom {
  float2 pix;
  float3 conic;
  float depth;
  int radius;
  int tiles;
  int2 rmin, rmax;  // tile rectangle [rmin, rmax)
};

// Cull, project, EWA cov2D (+0.3), conic, radius and tile rectangle for one
// view; false when the Gaussian is culled.
__device__ __forceinline__ bool project_splat(float3 p, const float c[6], const ViewParams& v, SplatGeom& o) {
  const float* vm = v.vm;
  const float3 t = xform_point_4x3(p, vm);
  if (t.z <= 0.2f) return false;
  const float4 ph = xform_point_4x4(p, v.pm);
  const float pw = 1.0f / (ph.w + 0.0000001f);
  const float limx = 1.3f * v.tanx, limy = 1.3f * v.tany;
  const float txtz = t.x / t.z, tytz = t.y / t.z;
  const float tx = fminf(limx, fmaxf(-limx, txtz)) * t.z;
  const float ty = fminf(limy, fmaxf(-limy, tytz)) * t.z;
  const float iz = 1.0f / t.z;
  const float j00 = v.fx * iz, j02 = -(v.fx * tx) * (iz * iz);
  const float j11 = v.fy * iz, j12 = -(v.fy * ty) * (iz * iz);
  // rows of M = J W, with W[r][k] = vm[4 k + r]
  const float m00 = j00 * vm[0] + j02 * vm[2], m01 = j00 * vm[4] + j02 * vm[6], m02 = j00 * vm[8] + j02 * vm[10];
  const float m10 = j11 * vm[1] + j12 * vm[2], m11 = j11 * vm[5] + j12 * vm[6], m12 = j11 * vm[9] + j12 * vm[10];
  const float s0x = c[0] * m00 + c[1] * m01 + c[2] * m02;
  const float s0y = c[1] * m00 + c[3] * m01 + c[4] * m02;
  const float s0z = c[2] * m00 + c[4] * m01 + c[5] * m02;
  const float s1x = c[0] * m10 + c[1] * m11 + c[2] * m12;
  const float s1y = c[1] * m10 + c[3] * m11 + c[4] * m12;
  const float s1z = c[2] * m10 + c[4] * m11 + c[5] * m12;
  const float ca = m00 * s0x + m01 * s0y + m02 * s0z + 0.3f;
  const float cb = m10 * s0x + m11 * s0y + m12 * s0z;
  const float cc = m10 * s1x + m11 * s1y + m12 * s1z + 0.3f;
  const float det = ca * cc - cb * cb;
  if (det == 0.0f) return false;
  const float det_inv = 1.f / det;
  o.conic = make_float3(cc * det_inv, -cb * det_inv, ca * det_inv);
  const float mid = 0.5f * (ca + cc);
  const float disc = sqrtf(fmaxf(0.1f, mid * mid - det));
  const float lambda1 = mid + disc, lambda2 = mid - disc;
  const float radius = ceilf(3.f * sqrtf(fmaxf(lambda1, lambda2)));
  o.pix = make_float2(ndc_to_pix(ph.x * pw, v.W), ndc_to_pix(ph.y * pw, v.H));
  int2& rmin = o.rmin;
  int2& rmax = o.rmax;
  tile_rect(o.pix, (int)radius, v.gx, v.gy, rmin, rmax);
  o.tiles = (rmax.x - rmin.x) * (rmax.y - rmin.y);
  if (o.tiles == 0) return false;
  o.radius = (int)radius;
  o.depth = t.z;
  return true;
}

// Backward of projection + EWA for one view.  Inputs: dL/dmean2D (NDC) and
// dL/dconic (a, b half-valued, c).  Adds dL/dSigma (6-vector) to dc and
// returns dL/dmean.
__device__ __forceinline__ float3 ewa_backward(float3 p, const float c[6], const ViewParams& v, float dmx, float dmy,
                                               float dqa, float dqb, float dqc, float dc[6]) {
  const float* vm = v.vm;
  const float* pm = v.pm;
  const float3 t = xform_point_4x3(p, vm);
  const float limx = 1.3f * v.tanx, limy = 1.3f * v.tany;
  const float txtz = t.x / t.z, tytz = t.y / t.z;
  const float tx = fminf(limx, fmaxf(-limx, txtz)) * t.z;
  const float ty = fminf(limy, fmaxf(-limy, tytz)) * t.z;
  const float xmul = (txtz < -limx || txtz > limx) ? 0.f : 1.f;
  const float ymul = (tytz < -limy || tytz > limy) ? 0.f : 1.f;
  // Reciprocals that only scale gradients use the hardware rcp (1 ulp): t.z
  // >= 0.2, denom^2 + 1e-7 and w > 0 are normal numbers.  The clamp tests
  // above keep the forward's IEEE quotients, so both passes decide alike.
  const float iz = __builtin_amdgcn_rcpf(t.z), iz2 = iz * iz, iz3 = iz2 * iz;
  const float fx = v.fx, fy = v.fy;
  const float j00 = fx * iz, j02 = -(fx * tx) * iz2, j11 = fy * iz, j12 = -(fy * ty) * iz2;
  const float V0[3] = {vm[0], vm[4], vm[8]}, V1[3] = {vm[1], vm[5], vm[9]}, V2[3] = {vm[2], vm[6], vm[10]};
  float m0[3], m1[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    m0[k] = j00 * V0[k] + j02 * V2[k];
    m1[k] = j11 * V1[k] + j12 * V2[k];
  }
  float S0[3], S1[3];  // Sigma m0, Sigma m1
  S0[0] = c[0] * m0[0] + c[1] * m0[1] + c[2] * m0[2];
  S0[1] = c[1] * m0[0] + c[3] * m0[1] + c[4] * m0[2];
  S0[2] = c[2] * m0[0] + c[4] * m0[1] + c[5] * m0[2];
  S1[0] = c[0] * m1[0] + c[1] * m1[1] + c[2] * m1[2];
  S1[1] = c[1] * m1[0] + c[3] * m1[1] + c[4] * m1[2];
  S1[2] = c[2] * m1[0] + c[4] * m1[1] + c[5] * m1[2];
  const float ca = m0[0] * S0[0] + m0[1] * S0[1] + m0[2] * S0[2] + 0.3f;
  const float cb = m1[0] * S0[0] + m1[1] * S0[1] + m1[2] * S0[2];
  const float cc = m1[0] * S1[0] + m1[1] * S1[1] + m1[2] * S1[2] + 0.3f;
  // conic -> cov2D
  const float denom = ca * cc - cb * cb;
  const float d2inv = __builtin_amdgcn_rcpf((denom * denom) + 0.0000001f);
  float dLa = 0.f, dLb = 0.f, dLc = 0.f;
  if (d2inv != 0.f) {
    dLa = d2inv * (-cc * cc * dqa + 2.f * cb * cc * dqb + (denom - ca * cc) * dqc);
    dLc = d2inv * (-ca * ca * dqc + 2.f * ca * cb * dqb + (denom - ca * cc) * dqa);
    dLb = d2inv * 2.f * (cb * cc * dqa - (denom + 2.f * cb * cb) * dqb + ca * cb * dqc);
  }
  // cov2D -> cov3D (6-vector)
  dc[0] += dLa * m0[0] * m0[0] + dLb * m0[0] * m1[0] + dLc * m1[0] * m1[0];
  dc[3] += dLa * m0[1] * m0[1] + dLb * m0[1] * m1[1] + dLc * m1[1] * m1[1];
  dc[5] += dLa * m0[2] * m0[2] + dLb * m0[2] * m1[2] + dLc * m1[2] * m1[2];
  dc[1] += 2.f * dLa * m0[0] * m0[1] + dLb * (m0[0] * m1[1] + m0[1] * m1[0]) + 2.f * dLc * m1[0] * m1[1];
  dc[2] += 2.f * dLa * m0[0] * m0[2] + dLb * (m0[0] * m1[2] + m0[2] * m1[0]) + 2.f * dLc * m1[0] * m1[2];
  dc[4] += 2.f * dLa * m0[1] * m0[2] + dLb * (m0[1] * m1[2] + m0[2] * m1[1]) + 2.f * dLc * m1[1] * m1[2];
  // cov2D -> J -> t -> mean
  float dm0[3], dm1[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dm0[k] = 2.f * dLa * S0[k] + dLb * S1[k];
    dm1[k] = 2.f * dLc * S1[k] + dLb * S0[k];
  }
  const float dj00 = dm0[0] * V0[0] + dm0[1] * V0[1] + dm0[2] * V0[2];
  const float dj02 = dm0[0] * V2[0] + dm0[1] * V2[1] + dm0[2] * V2[2];
  const float dj11 = dm1[0] * V1[0] + dm1[1] * V1[1] + dm1[2] * V1[2];
  const float dj12 = dm1[0] * V2[0] + dm1[1] * V2[1] + dm1[2] * V2[2];
  const float dtx = xmul * -fx * iz2 * dj02;
  const float dty = ymul * -fy * iz2 * dj12;
  const float dtz = -fx * iz2 * dj00 - fy * iz2 * dj11 + (2.f * fx * tx) * iz3 * dj02 + (2.f * fy * ty) * iz3 * dj12;
  float3 dmean = make_float3(vm[0] * dtx + vm[1] * dty + vm[2] * dtz, vm[4] * dtx + vm[5] * dty + vm[6] * dtz,
                             vm[8] * dtx + vm[9] * dty + vm[10] * dtz);
  // projection (means2D gradient is in NDC)
  const float4 ph = xform_point_4x4(p, pm);
  const float mw = __builtin_amdgcn_rcpf(ph.w + 0.0000001f);
  const float mul1 = ph.x * mw * mw, mul2 = ph.y * mw * mw;
  dmean.x += (pm[0] * mw - pm[3] * mul1) * dmx + (pm[1] * mw - pm[3] * mul2) * dmy;
  dmean.y += (pm[4] * mw - pm[7] * mul1) * dmx + (pm[5] * mw - pm[7] * mul2) * dmy;
  dmean.z += (pm[8] * mw - pm[11] * mul1) * dmx + (pm[9] * mw - pm[11] * mul2) * dmy;
  return dmean;
}

// SH backward for one view: recomputes the colour (for the clamp mask), sets
// the 16 basis values b (zero past the active degree) and the masked colour
// gradient g, and returns dL/dmean through the normalised view direction.
// The basis partials are those of the polynomials in sh_to_rgb.
template <typename SH>
__device__ __forceinline__ float3 sh_backward(int deg, const SH& sh, float3 pos, const float* campos,
                                              float3 dL_dRGB, float (&b)[16], float3& g) {
  const float3 v = make_float3(pos.x - campos[0], pos.y - campos[1], pos.z - campos[2]);
  const float inv_len = 1.0f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  const float x = v.x * inv_len, y = v.y * inv_len, z = v.z * inv_len;
#pragma unroll
  for (int k = 0; k < 16; ++k) b[k] = 0.f;
  float3 rgb = make_float3(0.5f, 0.5f, 0.5f);
  float3 dRx = make_float3(0.f, 0.f, 0.f), dRy = dRx, dRz = dRx;
  // basis value k, and each nonzero partial of it along x, y, z; the sums
  // keep the order over k (a partial that is identically zero adds nothing)
  auto basis = [&](int k, float bk) {
    b[k] = bk;
    rgb.x += bk * sh[3 * k]; rgb.y += bk * sh[3 * k + 1]; rgb.z += bk * sh[3 * k + 2];
  };
  auto part = [&](float3& dR, int k, float d) {
    dR.x += d * sh[3 * k]; dR.y += d * sh[3 * k + 1]; dR.z += d * sh[3 * k + 2];
  };
  basis(0, kSH_C0);
  if (deg > 0) {
    basis(1, -kSH_C1 * y);
    part(dRy, 1, -kSH_C1);
    basis(2, kSH_C1 * z);
    part(dRz, 2, kSH_C1);
    basis(3, -kSH_C1 * x);
    part(dRx, 3, -kSH_C1);
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      basis(4, kSH_C2[0] * xy);
      part(dRx, 4, kSH_C2[0] * y);
      part(dRy, 4, kSH_C2[0] * x);
      basis(5, kSH_C2[1] * yz);
      part(dRy, 5, kSH_C2[1] * z);
      part(dRz, 5, kSH_C2[1] * y);
      basis(6, kSH_C2[2] * (2.f * zz - xx - yy));
      part(dRx, 6, -2.f * kSH_C2[2] * x);
      part(dRy, 6, -2.f * kSH_C2[2] * y);
      part(dRz, 6, 4.f * kSH_C2[2] * z);
      basis(7, kSH_C2[3] * xz);
      part(dRx, 7, kSH_C2[3] * z);
      part(dRz, 7, kSH_C2[3] * x);
      basis(8, kSH_C2[4] * (xx - yy));
      part(dRx, 8, 2.f * kSH_C2[4] * x);
      part(dRy, 8, -2.f * kSH_C2[4] * y);
      if (deg > 2) {
        basis(9, kSH_C3[0] * y * (3.f * xx - yy));
        part(dRx, 9, kSH_C3[0] * 6.f * xy);
        part(dRy, 9, kSH_C3[0] * 3.f * (xx - yy));
        basis(10, kSH_C3[1] * xy * z);
        part(dRx, 10, kSH_C3[1] * yz);
        part(dRy, 10, kSH_C3[1] * xz);
        part(dRz, 10, kSH_C3[1] * xy);
        basis(11, kSH_C3[2] * y * (4.f * zz - xx - yy));
        part(dRx, 11, kSH_C3[2] * (-2.f * xy));
        part(dRy, 11, kSH_C3[2] * (4.f * zz - xx - 3.f * yy));
        part(dRz, 11, kSH_C3[2] * 8.f * yz);
        basis(12, kSH_C3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy));
        part(dRx, 12, kSH_C3[3] * (-6.f * xz));
        part(dRy, 12, kSH_C3[3] * (-6.f * yz));
        part(dRz, 12, kSH_C3[3] * 3.f * (2.f * zz - xx - yy));
        basis(13, kSH_C3[4] * x * (4.f * zz - xx - yy));
        part(dRx, 13, kSH_C3[4] * (4.f * zz - 3.f * xx - yy));
        part(dRy, 13, kSH_C3[4] * (-2.f * xy));
        part(dRz, 13, kSH_C3[4] * 8.f * xz);
        basis(14, kSH_C3[5] * z * (xx - yy));
        part(dRx, 14, kSH_C3[5] * 2.f * xz);
        part(dRy, 14, kSH_C3[5] * (-2.f * yz));
        part(dRz, 14, kSH_C3[5] * (xx - yy));
        basis(15, kSH_C3[6] * x * (xx - 3.f * yy));
        part(dRx, 15, kSH_C3[6] * 3.f * (xx - yy));
        part(dRy, 15, kSH_C3[6] * (-6.f * xy));
      }
    }
  }
  g = make_float3(rgb.x < 0.f ? 0.f : dL_dRGB.x, rgb.y < 0.f ? 0.f : dL_dRGB.y, rgb.z < 0.f ? 0.f : dL_dRGB.z);
  const float3 ddir = make_float3(dRx.x * g.x + dRx.y * g.y + dRx.z * g.z, dRy.x * g.x + dRy.y * g.y + dRy.z * g.z,
                                  dRz.x * g.x + dRz.y * g.y + dRz.z * g.z);
  // d normalize(v) / dv applied to ddir (|v|^-3 only scales the gradient:
  // hardware reciprocal square root, 1 ulp)
  const float s2 = v.x * v.x + v.y * v.y + v.z * v.z;
  const float inv32 = __builtin_amdgcn_rsqf(s2 * s2 * s2);
  return make_float3(((s2 - v.x * v.x) * ddir.x - v.y * v.x * ddir.y - v.z * v.x * ddir.z) * inv32,
                     (-v.x * v.y * ddir.x + (s2 - v.y * v.y) * ddir.y - v.z * v.y * ddir.z) * inv32,
                     (-v.x * v.z * ddir.x - v.y * v.z * ddir.y + (s2 - v.z * v.z) * ddir.z) * inv32);
}

// dL/dSigma (6-vector, off-diagonals counted once) -> dL/ds (s = mod * scale,
// no extra factor mod) and dL/dq for the un-normalised quaternion formula.
__device__ __forceinline__ void cov3d_backward(const float dc[6], float3 s3, const float R[9], float4 q, float3& ds_out,
                                               float4& dq_out) {
  // G = dL/dSigma (symmetric), dL/dM = 2 G M with M = R diag(s)
  const float G[9] = {dc[0], 0.5f * dc[1], 0.5f * dc[2], 0.5f * dc[1], dc[3], 0.5f * dc[4],
                      0.5f * dc[2], 0.5f * dc[4], dc[5]};
  const float sv[3] = {s3.x, s3.y, s3.z};
  float dM[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      dM[3 * i + j] = 2.f * sv[j] * (G[3 * i] * R[j] + G[3 * i + 1] * R[3 + j] + G[3 * i + 2] * R[6 + j]);
  float ds[3], gR[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    ds[j] = dM[j] * R[j] + dM[3 + j] * R[3 + j] + dM[6 + j] * R[6 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i) gR[3 * i + j] = dM[3 * i + j] * sv[j];
  }
  ds_out = make_float3(ds[0], ds[1], ds[2]);
  const float r = q.x, x = q.y, y = q.z, z = q.w;
  const float dr = 2.f * (-z * gR[1] + y * gR[2] + z * gR[3] - x * gR[5] - y * gR[6] + x * gR[7]);
  const float dx = 2.f * (y * gR[1] + z * gR[2] + y * gR[3] - 2.f * x * gR[4] - r * gR[5] + z * gR[6] + r * gR[7] -
                          2.f * x * gR[8]);
  const float dy = 2.f * (-2.f * y * gR[0] + x * gR[1] + r * gR[2] + x * gR[3] + z * gR[5] - r * gR[6] + z * gR[7] -
                          2.f * y * gR[8]);
  const float dz = 2.f * (-2.f * z * gR[0] - r * gR[1] + x * gR[2] + r * gR[3] - 2.f * z * gR[4] + y * gR[5] +
                          x * gR[6] + y * gR[7]);
  dq_out = make_float4(dr, dx, dy, dz);
}

}  // namespace dmgs
