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

namespace dmgs {

__constant__ float kB_C0 = 0.28209479177387814f;
__constant__ float kB_C1 = 0.4886025119029199f;
__constant__ float kB_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                               -1.0925484305920792f, 0.5462742152960396f};
__constant__ float kB_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                               0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                               -0.5900435899266435f};

// SH backward: recomputes the colour (for the clamp mask), writes dL/dsh and
// returns dL/d(mean) through the normalised view direction.  The basis
// partials along x, y, z are those of the polynomials in sh_to_rgb.
__device__ __forceinline__ float3 sh_backward(int deg, int M, const float* __restrict__ shp, float3 pos,
                                              const float* campos, float3 dL_dRGB, float* __restrict__ dsh) {
  const float3 v = make_float3(pos.x - campos[0], pos.y - campos[1], pos.z - campos[2]);
  const float inv_len = 1.0f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  const float x = v.x * inv_len, y = v.y * inv_len, z = v.z * inv_len;
  float sh[48];
  load_sh16(shp, M, sh);
  const int n = (deg + 1) * (deg + 1);

  float b[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) b[k] = 0.f;
  float3 rgb = make_float3(0.5f, 0.5f, 0.5f);
  float3 dRx = make_float3(0.f, 0.f, 0.f), dRy = dRx, dRz = dRx;
  // accumulate colour and the three directional partials for coefficient k
  auto acc = [&](int k, float bk, float bxk, float byk, float bzk) {
    const float s0 = sh[3 * k], s1 = sh[3 * k + 1], s2 = sh[3 * k + 2];
    b[k] = bk;
    rgb.x += bk * s0; rgb.y += bk * s1; rgb.z += bk * s2;
    dRx.x += bxk * s0; dRx.y += bxk * s1; dRx.z += bxk * s2;
    dRy.x += byk * s0; dRy.y += byk * s1; dRy.z += byk * s2;
    dRz.x += bzk * s0; dRz.y += bzk * s1; dRz.z += bzk * s2;
  };
  acc(0, kB_C0, 0.f, 0.f, 0.f);
  if (deg > 0) {
    acc(1, -kB_C1 * y, 0.f, -kB_C1, 0.f);
    acc(2, kB_C1 * z, 0.f, 0.f, kB_C1);
    acc(3, -kB_C1 * x, -kB_C1, 0.f, 0.f);
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      acc(4, kB_C2[0] * xy, kB_C2[0] * y, kB_C2[0] * x, 0.f);
      acc(5, kB_C2[1] * yz, 0.f, kB_C2[1] * z, kB_C2[1] * y);
      acc(6, kB_C2[2] * (2.f * zz - xx - yy), -2.f * kB_C2[2] * x, -2.f * kB_C2[2] * y, 4.f * kB_C2[2] * z);
      acc(7, kB_C2[3] * xz, kB_C2[3] * z, 0.f, kB_C2[3] * x);
      acc(8, kB_C2[4] * (xx - yy), 2.f * kB_C2[4] * x, -2.f * kB_C2[4] * y, 0.f);
      if (deg > 2) {
        acc(9, kB_C3[0] * y * (3.f * xx - yy), kB_C3[0] * 6.f * xy, kB_C3[0] * 3.f * (xx - yy), 0.f);
        acc(10, kB_C3[1] * xy * z, kB_C3[1] * yz, kB_C3[1] * xz, kB_C3[1] * xy);
        acc(11, kB_C3[2] * y * (4.f * zz - xx - yy), kB_C3[2] * (-2.f * xy), kB_C3[2] * (4.f * zz - xx - 3.f * yy),
            kB_C3[2] * 8.f * yz);
        acc(12, kB_C3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy), kB_C3[3] * (-6.f * xz), kB_C3[3] * (-6.f * yz),
            kB_C3[3] * 3.f * (2.f * zz - xx - yy));
        acc(13, kB_C3[4] * x * (4.f * zz - xx - yy), kB_C3[4] * (4.f * zz - 3.f * xx - yy), kB_C3[4] * (-2.f * xy),
            kB_C3[4] * 8.f * xz);
        acc(14, kB_C3[5] * z * (xx - yy), kB_C3[5] * 2.f * xz, kB_C3[5] * (-2.f * yz), kB_C3[5] * (xx - yy));
        acc(15, kB_C3[6] * x * (xx - 3.f * yy), kB_C3[6] * 3.f * (xx - yy), kB_C3[6] * (-6.f * xy), 0.f);
      }
    }
  }
  (void)n;
  const float gx = rgb.x < 0.f ? 0.f : dL_dRGB.x;
  const float gy = rgb.y < 0.f ? 0.f : dL_dRGB.y;
  const float gz = rgb.z < 0.f ? 0.f : dL_dRGB.z;
  // dL/dsh_k = b_k * g (zero past the active degree: b_k = 0 there)
  if (M == 16 && ((reinterpret_cast<uintptr_t>(dsh) & 15) == 0)) {
    float4* d4 = reinterpret_cast<float4*>(dsh);
    const float gg[3] = {gx, gy, gz};
#pragma unroll
    for (int q = 0; q < 12; ++q)
      d4[q] = make_float4(b[(4 * q) / 3] * gg[(4 * q) % 3], b[(4 * q + 1) / 3] * gg[(4 * q + 1) % 3],
                          b[(4 * q + 2) / 3] * gg[(4 * q + 2) % 3], b[(4 * q + 3) / 3] * gg[(4 * q + 3) % 3]);
  } else {
    for (int k = 0; k < M; ++k) {
      float bk = 0.f;
#pragma unroll
      for (int j = 0; j < 16; ++j) bk = (j == k) ? b[j] : bk;
      dsh[3 * k] = bk * gx;
      dsh[3 * k + 1] = bk * gy;
      dsh[3 * k + 2] = bk * gz;
    }
  }
  const float3 ddir = make_float3(dRx.x * gx + dRx.y * gy + dRx.z * gz, dRy.x * gx + dRy.y * gy + dRy.z * gz,
                                  dRz.x * gx + dRz.y * gy + dRz.z * gz);
  // d normalize(v) / dv applied to ddir
  const float s2 = v.x * v.x + v.y * v.y + v.z * v.z;
  const float inv32 = 1.0f / sqrtf(s2 * s2 * s2);
  return make_float3(((s2 - v.x * v.x) * ddir.x - v.y * v.x * ddir.y - v.z * v.x * ddir.z) * inv32,
                     (-v.x * v.y * ddir.x + (s2 - v.y * v.y) * ddir.y - v.z * v.y * ddir.z) * inv32,
                     (-v.x * v.z * ddir.x - v.y * v.z * ddir.y + (s2 - v.z * v.z) * ddir.z) * inv32);
}

__global__ __launch_bounds__(256) void geometry_bwd_kernel(GeometryBwdArgs a) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.P) return;
  const float* g = a.gacc + (size_t)idx * DMGS_GACC_STRIDE;
  const float4 g03 = *reinterpret_cast<const float4*>(g);
  const float4 g47 = *reinterpret_cast<const float4*>(g + 4);
  const float g8 = g[8];
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
  const float* vm = a.viewmatrix;
  const float* pm = a.projmatrix;
  const float3 p = make_float3(a.means3D[3 * idx], a.means3D[3 * idx + 1], a.means3D[3 * idx + 2]);

  // --- 3D covariance (recomputed) ---
  float c[6];
  float R[9];
  float3 s3 = make_float3(0.f, 0.f, 0.f);
  float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
  if (a.cov3D_precomp) {
    for (int k = 0; k < 6; ++k) c[k] = a.cov3D_precomp[6 * idx + k];
  } else {
    s3 = make_float3(a.scale_modifier * a.scales[3 * idx], a.scale_modifier * a.scales[3 * idx + 1],
                     a.scale_modifier * a.scales[3 * idx + 2]);
    q = reinterpret_cast<const float4*>(a.rotations)[idx];
    const float r = q.x, x = q.y, y = q.z, z = q.w;
    R[0] = 1.f - 2.f * (y * y + z * z); R[1] = 2.f * (x * y - r * z); R[2] = 2.f * (x * z + r * y);
    R[3] = 2.f * (x * y + r * z); R[4] = 1.f - 2.f * (x * x + z * z); R[5] = 2.f * (y * z - r * x);
    R[6] = 2.f * (x * z - r * y); R[7] = 2.f * (y * z + r * x); R[8] = 1.f - 2.f * (x * x + y * y);
    const float M00 = R[0] * s3.x, M01 = R[1] * s3.y, M02 = R[2] * s3.z;
    const float M10 = R[3] * s3.x, M11 = R[4] * s3.y, M12 = R[5] * s3.z;
    const float M20 = R[6] * s3.x, M21 = R[7] * s3.y, M22 = R[8] * s3.z;
    c[0] = M00 * M00 + M01 * M01 + M02 * M02;
    c[1] = M00 * M10 + M01 * M11 + M02 * M12;
    c[2] = M00 * M20 + M01 * M21 + M02 * M22;
    c[3] = M10 * M10 + M11 * M11 + M12 * M12;
    c[4] = M10 * M20 + M11 * M21 + M12 * M22;
    c[5] = M20 * M20 + M21 * M21 + M22 * M22;
  }

  // --- EWA forward quantities ---
  const float3 t = xform_point_4x3(p, vm);
  const float limx = 1.3f * a.tanfovx, limy = 1.3f * a.tanfovy;
  const float txtz = t.x / t.z, tytz = t.y / t.z;
  const float tx = fminf(limx, fmaxf(-limx, txtz)) * t.z;
  const float ty = fminf(limy, fmaxf(-limy, tytz)) * t.z;
  const float xmul = (txtz < -limx || txtz > limx) ? 0.f : 1.f;
  const float ymul = (tytz < -limy || tytz > limy) ? 0.f : 1.f;
  const float iz = 1.0f / t.z, iz2 = iz * iz, iz3 = iz2 * iz;
  const float fx = a.focal_x, fy = a.focal_y;
  const float j00 = fx * iz, j02 = -(fx * tx) * iz2, j11 = fy * iz, j12 = -(fy * ty) * iz2;
  // view rotation rows V[r][k] = vm[4 k + r]
  const float V0[3] = {vm[0], vm[4], vm[8]}, V1[3] = {vm[1], vm[5], vm[9]}, V2[3] = {vm[2], vm[6], vm[10]};
  float m0[3], m1[3];
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

  // --- conic -> cov2D ---
  const float denom = ca * cc - cb * cb;
  const float d2inv = 1.0f / ((denom * denom) + 0.0000001f);
  float dLa = 0.f, dLb = 0.f, dLc = 0.f;
  if (d2inv != 0.f) {
    dLa = d2inv * (-cc * cc * dqa + 2.f * cb * cc * dqb + (denom - ca * cc) * dqc);
    dLc = d2inv * (-ca * ca * dqc + 2.f * ca * cb * dqb + (denom - ca * cc) * dqa);
    dLb = d2inv * 2.f * (cb * cc * dqa - (denom + 2.f * cb * cb) * dqb + ca * cb * dqc);
  }
  // --- cov2D -> cov3D (6-vector) ---
  float dc[6];
  dc[0] = dLa * m0[0] * m0[0] + dLb * m0[0] * m1[0] + dLc * m1[0] * m1[0];
  dc[3] = dLa * m0[1] * m0[1] + dLb * m0[1] * m1[1] + dLc * m1[1] * m1[1];
  dc[5] = dLa * m0[2] * m0[2] + dLb * m0[2] * m1[2] + dLc * m1[2] * m1[2];
  dc[1] = 2.f * dLa * m0[0] * m0[1] + dLb * (m0[0] * m1[1] + m0[1] * m1[0]) + 2.f * dLc * m1[0] * m1[1];
  dc[2] = 2.f * dLa * m0[0] * m0[2] + dLb * (m0[0] * m1[2] + m0[2] * m1[0]) + 2.f * dLc * m1[0] * m1[2];
  dc[4] = 2.f * dLa * m0[1] * m0[2] + dLb * (m0[1] * m1[2] + m0[2] * m1[1]) + 2.f * dLc * m1[1] * m1[2];
  if (a.dL_dcov3D)
    for (int k = 0; k < 6; ++k) a.dL_dcov3D[6 * idx + k] = dc[k];

  // --- cov2D -> J -> t -> mean ---
  float dm0[3], dm1[3];
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

  // --- projection (means2D gradient is in NDC) ---
  const float4 ph = xform_point_4x4(p, pm);
  const float mw = 1.0f / (ph.w + 0.0000001f);
  const float mul1 = ph.x * mw * mw, mul2 = ph.y * mw * mw;
  dmean.x += (pm[0] * mw - pm[3] * mul1) * dmx + (pm[1] * mw - pm[3] * mul2) * dmy;
  dmean.y += (pm[4] * mw - pm[7] * mul1) * dmx + (pm[5] * mw - pm[7] * mul2) * dmy;
  dmean.z += (pm[8] * mw - pm[11] * mul1) * dmx + (pm[9] * mw - pm[11] * mul2) * dmy;

  // --- SH ---
  if (a.shs) {
    const float3 dm = sh_backward(a.sh_degree, a.sh_coeffs, a.shs + (size_t)idx * 3 * a.sh_coeffs, p, a.campos,
                                  make_float3(g47.z, g47.w, g8), a.dL_dsh + (size_t)idx * 3 * a.sh_coeffs);
    dmean.x += dm.x;
    dmean.y += dm.y;
    dmean.z += dm.z;
  }
  if (a.dL_dmeans3D) {
    a.dL_dmeans3D[3 * idx] = dmean.x;
    a.dL_dmeans3D[3 * idx + 1] = dmean.y;
    a.dL_dmeans3D[3 * idx + 2] = dmean.z;
  }

  // --- cov3D -> scale, rotation ---
  if (!a.cov3D_precomp && (a.dL_dscales || a.dL_drotations)) {
    // G = dL/dSigma (symmetric), dL/dM = 2 G M with M = R diag(s)
    const float G[9] = {dc[0], 0.5f * dc[1], 0.5f * dc[2], 0.5f * dc[1], dc[3], 0.5f * dc[4],
                        0.5f * dc[2], 0.5f * dc[4], dc[5]};
    const float sv[3] = {s3.x, s3.y, s3.z};
    float dM[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        dM[3 * i + j] = 2.f * (G[3 * i] * R[j] * sv[j] + G[3 * i + 1] * R[3 + j] * sv[j] + G[3 * i + 2] * R[6 + j] * sv[j]);
    float ds[3];
    float gR[9];
    for (int j = 0; j < 3; ++j) {
      ds[j] = dM[j] * R[j] + dM[3 + j] * R[3 + j] + dM[6 + j] * R[6 + j];
      for (int i = 0; i < 3; ++i) gR[3 * i + j] = dM[3 * i + j] * sv[j];
    }
    if (a.dL_dscales) {
      a.dL_dscales[3 * idx] = ds[0];
      a.dL_dscales[3 * idx + 1] = ds[1];
      a.dL_dscales[3 * idx + 2] = ds[2];
    }
    if (a.dL_drotations) {
      const float r = q.x, x = q.y, y = q.z, z = q.w;
      const float dr = 2.f * (-z * gR[1] + y * gR[2] + z * gR[3] - x * gR[5] - y * gR[6] + x * gR[7]);
      const float dx = 2.f * (y * gR[1] + z * gR[2] + y * gR[3] - 2.f * x * gR[4] - r * gR[5] + z * gR[6] + r * gR[7] -
                              2.f * x * gR[8]);
      const float dy = 2.f * (-2.f * y * gR[0] + x * gR[1] + r * gR[2] + x * gR[3] + z * gR[5] - r * gR[6] + z * gR[7] -
                              2.f * y * gR[8]);
      const float dz = 2.f * (-2.f * z * gR[0] - r * gR[1] + x * gR[2] + r * gR[3] - 2.f * z * gR[4] + y * gR[5] +
                              x * gR[6] + y * gR[7]);
      reinterpret_cast<float4*>(a.dL_drotations)[idx] = make_float4(dr, dx, dy, dz);
    }
  }
}

void launch_geometry_bwd(const GeometryBwdArgs& a, hipStream_t s) {
  if (a.P <= 0) return;
  hipLaunchKernelGGL(geometry_bwd_kernel, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
}

}  // namespace dmgs
