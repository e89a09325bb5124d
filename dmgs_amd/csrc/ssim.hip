// Fused (1 - lambda) L1 + lambda (1 - SSIM) loss, forward and backward, for
// gfx950.
//
// Same objective as utils/loss_utils.py:17-62 with train_geo_stage3.py:122-124:
// SSIM with an 11x11 Gaussian window (sigma 1.5), zero padding 5, per channel
// (depthwise), C1 = 0.01^2, C2 = 0.03^2, mean over channels and pixels; L1 is
// the mean absolute difference.  The 2-D window is the outer product of the
// normalised 1-D window, so both passes blur separably through LDS.
//
// Forward: one 16x16 output tile per workgroup (channel in blockIdx.z) stages
// a 26x26 patch of both images, blurs x, y, x^2, y^2, xy, evaluates the SSIM
// map, writes per-block partial sums (deterministic two-level reduction) and
// the three per-pixel coefficient maps the backward needs:
//   a = dS/dmu1, b = dS/d(E[x^2]), c = dS/d(E[xy])   (each times dL/dS).
// Backward: dL/dx = G*a + 2 x (G*b) + y (G*c) + (1 - lambda) sign(x - y) / N.
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

constexpr int kT = 16;            // output tile edge
constexpr int kR = 5;             // window radius
constexpr int kP = kT + 2 * kR;   // 26, staged patch edge

struct Window11 {
  float w[11];
};

// One loss per group of `group` consecutive planes (group = C: one loss over
// the whole image; group = 3: one per RGB view of a batch).
__global__ __launch_bounds__(256) void dssim_fwd_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                        int H, int W, int group, Window11 g, float lam,
                                                        float* __restrict__ coef, float* __restrict__ partials) {
  __shared__ float sx[kP][kP + 1], sy[kP][kP + 1];
  __shared__ float hq[5][kP][kT + 1];
  __shared__ float red[2][4];
  const int c = blockIdx.z;
  const int ox = blockIdx.x * kT, oy = blockIdx.y * kT;
  const size_t plane = (size_t)H * W;
  const float* xc = x + c * plane;
  const float* yc = y + c * plane;
  for (int i = threadIdx.x; i < kP * kP; i += 256) {
    const int r = i / kP, q = i % kP;
    const int gy = oy + r - kR, gx = ox + q - kR;
    const bool in = gy >= 0 && gy < H && gx >= 0 && gx < W;
    sx[r][q] = in ? xc[(size_t)gy * W + gx] : 0.f;
    sy[r][q] = in ? yc[(size_t)gy * W + gx] : 0.f;
  }
  __syncthreads();
  // horizontal pass over all 26 rows for the 16 output columns
  for (int i = threadIdx.x; i < kP * kT; i += 256) {
    const int r = i / kT, q = i % kT;
    float m1 = 0.f, m2 = 0.f, s11 = 0.f, s22 = 0.f, s12 = 0.f;
#pragma unroll
    for (int k = 0; k < 11; ++k) {
      const float a = sx[r][q + k], b = sy[r][q + k], w = g.w[k];
      m1 += w * a;
      m2 += w * b;
      s11 += w * (a * a);
      s22 += w * (b * b);
      s12 += w * (a * b);
    }
    hq[0][r][q] = m1;
    hq[1][r][q] = m2;
    hq[2][r][q] = s11;
    hq[3][r][q] = s22;
    hq[4][r][q] = s12;
  }
  __syncthreads();
  const int lx = threadIdx.x % kT, ly = threadIdx.x / kT;
  const int px = ox + lx, py = oy + ly;
  float v[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 11; ++k) {
    const float w = g.w[k];
#pragma unroll
    for (int q = 0; q < 5; ++q) v[q] += w * hq[q][ly + k][lx];
  }
  const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
  const float mu1 = v[0], mu2 = v[1];
  const float mu1_sq = mu1 * mu1, mu2_sq = mu2 * mu2, mu12 = mu1 * mu2;
  const float sig1 = v[2] - mu1_sq, sig2 = v[3] - mu2_sq, sig12 = v[4] - mu12;
  const float A1 = 2.f * mu12 + C1, A2 = 2.f * sig12 + C2;
  const float B1 = mu1_sq + mu2_sq + C1, B2 = sig1 + sig2 + C2;
  const bool inside = px < W && py < H;
  float S = 0.f, l1 = 0.f;
  if (inside) {
    S = (A1 * A2) / (B1 * B2);
    l1 = fabsf(sx[ly + kR][lx + kR] - sy[ly + kR][lx + kR]);
    // dL/dS for the mean over the group's group*H*W pixels
    const float dS = -lam / (float)((size_t)group * plane);
    const float inv = 1.f / (B1 * B2);
    const float dmu1 = (2.f * mu2 * A2 - 2.f * mu2 * A1) * inv - S * (2.f * mu1 / B1 - 2.f * mu1 / B2);
    const float ds11 = -S / B2;
    const float ds12 = 2.f * A1 * inv;
    const size_t o = c * plane + (size_t)py * W + px;
    const size_t all = (size_t)gridDim.z * plane;
    coef[o] = dS * dmu1;
    coef[all + o] = dS * ds11;
    coef[2 * all + o] = dS * ds12;
  }
  // block sums of S and |x - y|
  const float s_sum = wave_sum(S), l_sum = wave_sum(l1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = s_sum;
    red[1][wave] = l_sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    partials[2 * b] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    partials[2 * b + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// One workgroup per loss group (its partials are contiguous: `nblocks`
// blocks per group): loss = (1-lam) * L1 + lam * (1 - SSIM);
// out[3 g ..] = (loss, l1, ssim).
__global__ __launch_bounds__(256) void dssim_reduce_kernel(const float* __restrict__ partials, int nblocks,
                                                           float inv_n, float lam, float* __restrict__ out) {
  __shared__ float red[2][4];
  partials += (size_t)2 * nblocks * blockIdx.x;
  out += 3 * blockIdx.x;
  float s = 0.f, l = 0.f;
  for (int i = threadIdx.x; i < nblocks; i += 256) {
    s += partials[2 * i];
    l += partials[2 * i + 1];
  }
  s = wave_sum(s);
  l = wave_sum(l);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = s;
    red[1][wave] = l;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const float ssim = ((red[0][0] + red[0][1]) + (red[0][2] + red[0][3])) * inv_n;
    const float l1 = ((red[1][0] + red[1][1]) + (red[1][2] + red[1][3])) * inv_n;
    out[0] = (1.f - lam) * l1 + lam * (1.f - ssim);
    out[1] = l1;
    out[2] = ssim;
  }
}

__global__ __launch_bounds__(256) void dssim_bwd_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                        const float* __restrict__ coef, int H, int W, int C,
                                                        int group, Window11 g, float lam,
                                                        const float* __restrict__ dloss, float* __restrict__ dx) {
  __shared__ float sa[3][kP][kP + 1];
  __shared__ float hq[3][kP][kT + 1];
  const int c = blockIdx.z;
  const int ox = blockIdx.x * kT, oy = blockIdx.y * kT;
  const size_t plane = (size_t)H * W;
  const size_t all = (size_t)C * plane;
  for (int i = threadIdx.x; i < kP * kP; i += 256) {
    const int r = i / kP, q = i % kP;
    const int gy = oy + r - kR, gx = ox + q - kR;
    const bool in = gy >= 0 && gy < H && gx >= 0 && gx < W;
    const size_t o = c * plane + (size_t)gy * W + gx;
#pragma unroll
    for (int m = 0; m < 3; ++m) sa[m][r][q] = in ? coef[m * all + o] : 0.f;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kP * kT; i += 256) {
    const int r = i / kT, q = i % kT;
    float h0 = 0.f, h1 = 0.f, h2 = 0.f;
#pragma unroll
    for (int k = 0; k < 11; ++k) {
      const float w = g.w[k];
      h0 += w * sa[0][r][q + k];
      h1 += w * sa[1][r][q + k];
      h2 += w * sa[2][r][q + k];
    }
    hq[0][r][q] = h0;
    hq[1][r][q] = h1;
    hq[2][r][q] = h2;
  }
  __syncthreads();
  const int lx = threadIdx.x % kT, ly = threadIdx.x / kT;
  const int px = ox + lx, py = oy + ly;
  if (px >= W || py >= H) return;
  float b0 = 0.f, b1 = 0.f, b2 = 0.f;
#pragma unroll
  for (int k = 0; k < 11; ++k) {
    const float w = g.w[k];
    b0 += w * hq[0][ly + k][lx];
    b1 += w * hq[1][ly + k][lx];
    b2 += w * hq[2][ly + k][lx];
  }
  const size_t o = c * plane + (size_t)py * W + px;
  const float xv = x[o], yv = y[o];
  const float d = xv - yv;
  const float sgn = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
  const float g_l1 = (1.f - lam) * sgn / (float)((size_t)group * plane);
  dx[o] = dloss[c / group] * (b0 + 2.f * xv * b1 + yv * b2 + g_l1);
}

void launch_dssim_forward(const float* x, const float* y, int C, int group, int H, int W, const float* w11,
                          float lam, float* coef, float* partials, float* out, hipStream_t s) {
  Window11 g;
  for (int k = 0; k < 11; ++k) g.w[k] = w11[k];
  const dim3 grid((W + kT - 1) / kT, (H + kT - 1) / kT, C);
  hipLaunchKernelGGL(dssim_fwd_kernel, grid, dim3(256), 0, s, x, y, H, W, group, g, lam, coef, partials);
  const int per_group = (int)(grid.x * grid.y) * group;
  hipLaunchKernelGGL(dssim_reduce_kernel, dim3(C / group), dim3(256), 0, s, (const float*)partials, per_group,
                     1.0f / (float)((size_t)group * H * W), lam, out);
}

void launch_dssim_backward(const float* x, const float* y, const float* coef, int C, int group, int H, int W,
                           const float* w11, float lam, const float* dloss, float* dx, hipStream_t s) {
  Window11 g;
  for (int k = 0; k < 11; ++k) g.w[k] = w11[k];
  const dim3 grid((W + kT - 1) / kT, (H + kT - 1) / kT, C);
  hipLaunchKernelGGL(dssim_bwd_kernel, grid, dim3(256), 0, s, x, y, coef, H, W, C, group, g, lam, dloss, dx);
}

size_t dssim_partials(int C, int H, int W) {
  return (size_t)2 * ((W + kT - 1) / kT) * ((H + kT - 1) / kT) * C;
}

}  // namespace dmgs
