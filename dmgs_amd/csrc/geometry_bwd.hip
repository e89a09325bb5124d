// Per-Gaussian backward for gfx950: one fused kernel for what the reference
// splits into computeCov2DCUDA and preprocessCUDA backward [U]
// (diff-gaussian-rasterization, reached from gaussian_renderer/__init__.py:86-94).
//
// Input is the blend backward's per-Gaussian accumulator (dL/dmean2D in NDC,
// dL/dconic with the half-valued off-diagonal, dL/dopacity, dL/dcolour).  The
// 3D covariance and the SH clamp flags are recomputed here rather than
// stored by the forward.
//
// Gradient conventions kept from the reference: 1 / (det^2 + 1e-7) in the
// conic -> cov2D step; the tan(fov) clamp zeroes d/dtx (d/dty) and the
// clamped t.x is a constant inside d/dtz; the mean gradient from the cov2D
// path is assigned and the projection term added; no quaternion-normalisation
// term; dL/dscale is w.r.t. s = mod * scale; per-channel SH gradient zeroed
// where the colour was clamped.
#include "dmgs_common.h"
#include "dmgs_kernels.h"
#include "gs_math.h"

namespace dmgs {

// dL/dsh_k = b_k * g for the first 16 coefficients, zero beyond.
__device__ __forceinline__ void store_dsh(float* __restrict__ dsh, int M, const float (&b)[16], float3 g) {
  if (M == 16 && ((reinterpret_cast<uintptr_t>(dsh) & 15) == 0)) {
    float4* d4 = reinterpret_cast<float4*>(dsh);
    const float gg[3] = {g.x, g.y, g.z};
#pragma unroll
    for (int q = 0; q < 12; ++q)
      d4[q] = make_float4(b[(4 * q) / 3] * gg[(4 * q) % 3], b[(4 * q + 1) / 3] * gg[(4 * q + 1) % 3],
                          b[(4 * q + 2) / 3] * gg[(4 * q + 2) % 3], b[(4 * q + 3) / 3] * gg[(4 * q + 3) % 3]);
  } else {
    for (int k = 0; k < M; ++k) {
      float bk = 0.f;
#pragma unroll
      for (int j = 0; j < 16; ++j) bk = (j == k) ? b[j] : bk;
      dsh[3 * k] = bk * g.x;
      dsh[3 * k + 1] = bk * g.y;
      dsh[3 * k + 2] = bk * g.z;
    }
  }
}

__global__ __launch_bounds__(256) void geometry_bwd_kernel(GeometryBwdArgs a) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.P) return;
  const float2* gg = reinterpret_cast<const float2*>(a.gacc + gacc_index(a.P, idx, 0));  // 8-byte aligned rows
  const float2 r0 = gg[0], r1 = gg[1], r2 = gg[2];
  const float* gc = a.gacc + gacc_index(a.P, idx, DMGS_GACC_GEO);
  const float4 g03 = make_float4(r0.x, r0.y, r1.x, r1.y);
  const float4 g47 = make_float4(r2.x, r2.y, gc[0], gc[1]);
  const float g8 = gc[2];
  const float dmx = g03.x, dmy = g03.y, dqa = g03.z, dqb = g03.w, dqc = g47.x, dop = g47.y;
  if (a.dL_dmeans2D) {
    a.dL_dmeans2D[3 * idx] = dmx;
    a.dL_dmeans2D[3 * idx + 1] = dmy;
    a.dL_dmeans2D[3 * idx + 2] = 0.f;
  }
  if (a.dL_dcolors) {
    a.dL_dcolors[3 * idx] = g47.z;
    a.dL_dcolors[3 * idx + 1] = g47.w;
    a.dL_dcolors[3 * idx + 2] = g8;
  }
  if (a.dL_dopacity) a.dL_dopacity[idx] = dop;
  if (a.dL_dconic) reinterpret_cast<float4*>(a.dL_dconic)[idx] = make_float4(dqa, dqb, 0.f, dqc);

  const bool vis = a.radii[idx] > 0;
  if (!vis) {
    if (a.dL_dmeans3D) { a.dL_dmeans3D[3 * idx] = 0.f; a.dL_dmeans3D[3 * idx + 1] = 0.f; a.dL_dmeans3D[3 * idx + 2] = 0.f; }
    if (a.dL_dcov3D) for (int k = 0; k < 6; ++k) a.dL_dcov3D[6 * idx + k] = 0.f;
    if (a.dL_dsh) for (int k = 0; k < 3 * a.sh_coeffs; ++k) a.dL_dsh[(size_t)idx * 3 * a.sh_coeffs + k] = 0.f;
    if (a.dL_dscales) { a.dL_dscales[3 * idx] = 0.f; a.dL_dscales[3 * idx + 1] = 0.f; a.dL_dscales[3 * idx + 2] = 0.f; }
    if (a.dL_drotations) reinterpret_cast<float4*>(a.dL_drotations)[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const float3 p = make_float3(a.means3D[3 * idx], a.means3D[3 * idx + 1], a.means3D[3 * idx + 2]);

  // --- 3D covariance (recomputed) ---
  float c[6], R[9];
  float3 s3 = make_float3(0.f, 0.f, 0.f);
  float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
  if (a.cov3D_precomp) {
    for (int k = 0; k < 6; ++k) c[k] = a.cov3D_precomp[6 * idx + k];
  } else {
    const float m = a.scale_modifier;
    s3 = make_float3(m * a.scales[3 * idx], m * a.scales[3 * idx + 1], m * a.scales[3 * idx + 2]);
    q = reinterpret_cast<const float4*>(a.rotations)[idx];
    quat_to_R(q, R);
    cov3d_from_sR(s3, R, c);
  }
  const ViewParams v = view_params(a.viewmatrix, a.projmatrix, a.campos, a.W, a.H, 0, 0, a.focal_x, a.focal_y,
                                   a.tanfovx, a.tanfovy);
  float dc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float3 dmean = ewa_backward(p, c, v, dmx, dmy, dqa, dqb, dqc, dc);
  if (a.dL_dcov3D)
    for (int k = 0; k < 6; ++k) a.dL_dcov3D[6 * idx + k] = dc[k];

  if (a.shs) {
    float sh[48], b[16];
    float3 gm;
    load_sh16(a.shs + (size_t)idx * 3 * a.sh_coeffs, a.sh_coeffs, sh);
    const float3 dm = sh_backward(a.sh_degree, sh, p, v.campos, make_float3(g47.z, g47.w, g8), b, gm);
    store_dsh(a.dL_dsh + (size_t)idx * 3 * a.sh_coeffs, a.sh_coeffs, b, gm);
    dmean.x += dm.x;
    dmean.y += dm.y;
    dmean.z += dm.z;
  }
  if (a.dL_dmeans3D) {
    a.dL_dmeans3D[3 * idx] = dmean.x;
    a.dL_dmeans3D[3 * idx + 1] = dmean.y;
    a.dL_dmeans3D[3 * idx + 2] = dmean.z;
  }
  if (!a.cov3D_precomp && (a.dL_dscales || a.dL_drotations)) {
    float3 ds;
    float4 dq;
    cov3d_backward(dc, s3, R, q, ds, dq);
    if (a.dL_dscales) {
      a.dL_dscales[3 * idx] = ds.x;
      a.dL_dscales[3 * idx + 1] = ds.y;
      a.dL_dscales[3 * idx + 2] = ds.z;
    }
    if (a.dL_drotations) reinterpret_cast<float4*>(a.dL_drotations)[idx] = dq;
  }
}

void launch_geometry_bwd(const GeometryBwdArgs& a, hipStream_t s) {
  if (a.P <= 0) return;
  hipLaunchKernelGGL(geometry_bwd_kernel, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
}

}  // namespace dmgs
