// Fused (1 - lambda) L1 + lambda (1 - SSIM) loss, forward and backward, for
// gfx950.
//
// Same objective as utils/loss_utils.py:17-62 with train_geo_stage3.py:122-124:
// SSIM with an 11x11 Gaussian window (sigma 1.5), zero padding 5, per channel
// (depthwise), C1 = 0.01^2, C2 = 0.03^2, mean over channels and pixels; L1 is
// the mean absolute difference.  The 2-D window is the outer product of the
// normalised 1-D window, so both passes blur separably through LDS.
//
// Forward: one 32x32 output tile per workgroup (the grid's x, y, plane;
// xcd_tile maps work-group ids to tiles so that each XCD works through one
// contiguous run of them) stages the 42x42 patch of both images in LDS.
// The horizontal blur gives each
// task four neighbouring outputs of one patch row from 14 loaded values
// (x, y and the three products formed once per value); the vertical blur
// gives each thread four outputs of one column from 14 blurred rows.  The
// kernel evaluates the SSIM map, writes per-block partial sums
// (deterministic two-level reduction) and the three per-pixel coefficient
// maps the backward needs:
//   a = dS/dmu1, b = dS/d(E[x^2]), c = dS/d(E[xy])   (each times dL/dS).
// Backward: dL/dx = G*a + 2 x (G*b) + y (G*c) + (1 - lambda) sign(x - y) / N,
// with the same tiling over the three coefficient planes.
//
// Fixed targets: the two target maps G*y and G*y^2 do not change between
// steps; dssim_target_kernel computes them once (the forward's own loops, so
// the values are bit-identical) and the forward then blurs three maps
// (x, x^2, xy) instead of five.
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

constexpr int kTW = 32, kTH = 32;              // output tile
constexpr int kR = 5;                          // window radius
constexpr int kPW = kTW + 2 * kR;              // 42 staged columns
constexpr int kPH = kTH + 2 * kR;              // 42 staged rows
constexpr int kPS = 44;                        // staged row stride (16-byte rows)
constexpr int kHS = 32;                        // blurred row stride (32 floats: one more work-group per CU)
constexpr int kQuads = kTW / 4;                // horizontal tasks per row
constexpr int kHTasks = kPH * kQuads;          // 336
static_assert(kTW == 32 && kTH == 32, "the vertical pass maps 256 threads to 32 columns x 8 row quads");

struct Window11 {
  float w[11];
};

// Stage zero-padded kPH x kPW patches of NP planes (origin at (oy - 5, ox - 5)).
// Thread t < 252 owns patch column t % 42 and rows t / 42 + 6k (k = 0..6).
// Every element is loaded from its coordinates clamped into the plane, so no
// load is conditional, and replaced by zero outside the image; all loads of
// the thread are issued before the first LDS store, so the workgroup waits
// for one memory latency, not one per element.  Offsets within a plane are
// 32-bit products of 24-bit factors (sides below 2^24, planes of at most
// 2^31 pixels).
constexpr int kStageGroups = 6;                    // row groups
constexpr int kStageIters = kPH / kStageGroups;    // 7
static_assert(kPH % kStageGroups == 0 && kPW * kStageGroups <= 256, "one patch column per thread and group");
template <int NP>
__device__ __forceinline__ void stage_patches(float (*const dst[NP])[kPS], const float* const src[NP], int H, int W,
                                              int ox, int oy) {
  const int t = threadIdx.x;
  if (t >= kPW * kStageGroups) return;
  const int q = t % kPW, r0 = t / kPW;
  const int gx = ox + q - kR;
  const uint32_t cx = (uint32_t)min(max(gx, 0), W - 1);
  const bool in_x = (uint32_t)gx < (uint32_t)W;
  float v[NP][kStageIters];
#pragma unroll
  for (int k = 0; k < kStageIters; ++k) {
    const int gy = oy + r0 + kStageGroups * k - kR;
    const uint32_t o = __umul24((uint32_t)min(max(gy, 0), H - 1), (uint32_t)W) + cx;
#pragma unroll
    for (int m = 0; m < NP; ++m) v[m][k] = src[m][o];
  }
#pragma unroll
  for (int k = 0; k < kStageIters; ++k) {
    const int gy = oy + r0 + kStageGroups * k - kR;
    const bool in = in_x && (uint32_t)gy < (uint32_t)H;
#pragma unroll
    for (int m = 0; m < NP; ++m) dst[m][r0 + kStageGroups * k][q] = in ? v[m][k] : 0.f;
  }
}

// Offset of pixel (px, py) in its plane (both inside the image).
__device__ __forceinline__ uint32_t pix(int px, int py, int W) {
  return __umul24((uint32_t)py, (uint32_t)W) + (uint32_t)px;
}

// The tile of this work-group.  Work-groups are dealt to the GPU's 8 XCDs
// in turn (linear id mod 8) and each XCD has its own L2, so consecutive ids
// would put neighbouring tiles, whose 5-pixel margins overlap, on different
// L2s and every margin line would come from HBM once per tile.  Instead XCD
// k takes one contiguous run of tiles in (plane, row, column) order: a
// tile's left neighbour ran just before it on the same XCD and the tile
// above it one row of tiles earlier, so most margin lines are L2 hits.  A
// bijection of the grid onto itself for any grid size.
struct Tile {
  int bx, by, bz;
};
__device__ __forceinline__ Tile xcd_tile() {
  constexpr uint32_t kXcds = 8;
  const uint32_t gx = gridDim.x, gy = gridDim.y;
  const uint32_t n = gx * gy * gridDim.z;
  const uint32_t L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const uint32_t k = L % kXcds, j = L / kXcds;
  const uint32_t T = k * (n / kXcds) + min(k, n % kXcds) + j;
  Tile t;
  t.bx = (int)(T % gx);
  t.by = (int)((T / gx) % gy);
  t.bz = (int)(T / (gx * gy));
  return t;
}

// The 14 staged values of row r starting at column 4j (16-byte aligned).
__device__ __forceinline__ void load14(const float (*src)[kPS], int r, int j, float v[14]) {
  const float4* p = reinterpret_cast<const float4*>(&src[r][4 * j]);
  const float4 a = p[0], b = p[1], c = p[2];
  const float2 d = *reinterpret_cast<const float2*>(&src[r][4 * j + 12]);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  v[8] = c.x; v[9] = c.y; v[10] = c.z; v[11] = c.w;
  v[12] = d.x; v[13] = d.y;
}

// One loss per group of `group` consecutive planes (group = C: one loss over
// the whole image; group = 3: one per RGB view of a batch).
// STATS: ystats holds G*y (planes 0..C-1) and G*y^2 (planes C..2C-1).
template <bool STATS>
__global__ __launch_bounds__(256) void dssim_fwd_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                        const float* __restrict__ ystats, int H, int W, int group,
                                                        Window11 g, float lam, float* __restrict__ coef,
                                                        float* __restrict__ partials) {
  constexpr int NQ = STATS ? 3 : 5;  // blurred maps: x, x^2, xy (, y, y^2)
  // the x and y patches until every row blur has read them, then the
  // row-blurred maps (one buffer: more work-groups per CU)
  __shared__ __align__(16) union {
    float sxy[2][kPH][kPS];
    float hq[NQ][kPH][kHS];
  } u;
  __shared__ float red[2][4];
  const Tile tl = xcd_tile();
  const int c = tl.bz;
  const int ox = tl.bx * kTW, oy = tl.by * kTH;
  const size_t plane = (size_t)H * W;
  // the target's maps at this thread's four outputs (computed once by
  // dssim_target_kernel), requested together with the patches
  float ys[4][2];
  if constexpr (STATS) {
    const int lx = threadIdx.x & 31, r0 = (threadIdx.x >> 5) * 4;
    const float* y0 = ystats + c * plane;
    const float* y1 = y0 + (size_t)gridDim.z * plane;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t o = pix(min(ox + lx, W - 1), min(oy + r0 + i, H - 1), W);
      ys[i][0] = y0[o];
      ys[i][1] = y1[o];
    }
  }
  {
    float(*const dst[2])[kPS] = {u.sxy[0], u.sxy[1]};
    const float* const src[2] = {x + c * plane, y + c * plane};
    stage_patches<2>(dst, src, H, W, ox, oy);
  }
  __syncthreads();
  const int lx = threadIdx.x & 31, r0 = (threadIdx.x >> 5) * 4;
  const int px = ox + lx;
  // |x - y| at the thread's four outputs, before the patches are overwritten
  float l_acc = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (px < W && oy + r0 + i < H) l_acc += fabsf(u.sxy[0][r0 + i + kR][lx + kR] - u.sxy[1][r0 + i + kR][lx + kR]);
  // horizontal: four outputs of one staged row per task, held in registers
  // until every task has read the patches; map order x, x^2, xy (, y, y^2)
  static_assert(kHTasks <= 2 * 256, "two row tasks per thread at most");
  float res[2][NQ][4];
#pragma unroll
  for (int rep = 0; rep < 2; ++rep) {
    const int t = threadIdx.x + rep * 256;
    if (t >= kHTasks) break;
    const int r = t / kQuads, j = t - r * kQuads;
    float a[14], b[14];
    load14(u.sxy[0], r, j, a);
    load14(u.sxy[1], r, j, b);
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int i = 0; i < 4; ++i) res[rep][q][i] = 0.f;
#pragma unroll
    for (int e = 0; e < 14; ++e) {
      const float aa = a[e] * a[e], bb = b[e] * b[e], ab = a[e] * b[e];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = e - i;
        if (k < 0 || k > 10) continue;
        const float w = g.w[k];
        res[rep][0][i] = fmaf(w, a[e], res[rep][0][i]);
        res[rep][1][i] = fmaf(w, aa, res[rep][1][i]);
        res[rep][2][i] = fmaf(w, ab, res[rep][2][i]);
        if constexpr (!STATS) {
          res[rep][3][i] = fmaf(w, b[e], res[rep][3][i]);
          res[rep][4][i] = fmaf(w, bb, res[rep][4][i]);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int rep = 0; rep < 2; ++rep) {
    const int t = threadIdx.x + rep * 256;
    if (t >= kHTasks) break;
    const int r = t / kQuads, j = t - r * kQuads;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      *reinterpret_cast<float4*>(&u.hq[q][r][4 * j]) =
          make_float4(res[rep][q][0], res[rep][q][1], res[rep][q][2], res[rep][q][3]);
  }
  __syncthreads();
  // vertical: four outputs of one column per thread
  float v[4][5];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int q = 0; q < 5; ++q) v[i][q] = 0.f;
#pragma unroll
  for (int e = 0; e < 14; ++e) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const float h = u.hq[q][r0 + e][lx];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = e - i;
        if (k < 0 || k > 10) continue;
        v[i][q] = fmaf(g.w[k], h, v[i][q]);
      }
    }
  }
  if constexpr (STATS) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i][3] = ys[i][0];
      v[i][4] = ys[i][1];
    }
  }
  const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
  const float dS = -lam / (float)((size_t)group * plane);  // dL/dS for the mean over the group's pixels
  const size_t all = (size_t)gridDim.z * plane;
  float* const c0 = coef + c * plane;
  float* const c1 = c0 + all;
  float* const c2 = c1 + all;
  float s_acc = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int py = oy + r0 + i;
    if (px >= W || py >= H) continue;
    const float mu1 = v[i][0], mu2 = v[i][3];
    const float mu1_sq = mu1 * mu1, mu2_sq = mu2 * mu2, mu12 = mu1 * mu2;
    const float sig1 = v[i][1] - mu1_sq, sig2 = v[i][4] - mu2_sq, sig12 = v[i][2] - mu12;
    const float A1 = 2.f * mu12 + C1, A2 = 2.f * sig12 + C2;
    const float B1 = mu1_sq + mu2_sq + C1, B2 = sig1 + sig2 + C2;
    const float inv = 1.f / (B1 * B2);
    const float S = (A1 * A2) * inv;
    s_acc += S;
    // 1/B1 = B2 inv and 1/B2 = B1 inv
    const float dmu1 = inv * (2.f * mu2 * (A2 - A1) - S * (2.f * mu1) * (B2 - B1));
    const float ds11 = -S * B1 * inv;
    const float ds12 = 2.f * A1 * inv;
    const uint32_t o = pix(px, py, W);
    c0[o] = dS * dmu1;
    c1[o] = dS * ds11;
    c2[o] = dS * ds12;
  }
  // block sums of S and |x - y|
  const float s_sum = wave_sum(s_acc), l_sum = wave_sum(l_acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = s_sum;
    red[1][wave] = l_sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const size_t b = ((size_t)tl.bz * gridDim.y + tl.by) * gridDim.x + tl.bx;
    partials[2 * b] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    partials[2 * b + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// The target's maps G*y and G*y^2 (stats planes 0..C-1 and C..2C-1), with
// the forward's tiling and loops.
__global__ __launch_bounds__(256) void dssim_target_kernel(const float* __restrict__ y, int H, int W, Window11 g,
                                                           float* __restrict__ stats) {
  __shared__ __align__(16) float sy[kPH][kPS];
  __shared__ __align__(16) float hq[2][kPH][kHS];
  const Tile tl = xcd_tile();
  const int c = tl.bz;
  const int ox = tl.bx * kTW, oy = tl.by * kTH;
  const size_t plane = (size_t)H * W;
  {
    float(*const dst[1])[kPS] = {sy};
    const float* const src[1] = {y + c * plane};
    stage_patches<1>(dst, src, H, W, ox, oy);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kHTasks; t += 256) {
    const int r = t / kQuads, j = t - r * kQuads;
    float b[14];
    load14(sy, r, j, b);
    float m2[4] = {0.f, 0.f, 0.f, 0.f}, s22[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int e = 0; e < 14; ++e) {
      const float bb = b[e] * b[e];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = e - i;
        if (k < 0 || k > 10) continue;
        const float w = g.w[k];
        m2[i] = fmaf(w, b[e], m2[i]);
        s22[i] = fmaf(w, bb, s22[i]);
      }
    }
    *reinterpret_cast<float4*>(&hq[0][r][4 * j]) = make_float4(m2[0], m2[1], m2[2], m2[3]);
    *reinterpret_cast<float4*>(&hq[1][r][4 * j]) = make_float4(s22[0], s22[1], s22[2], s22[3]);
  }
  __syncthreads();
  const int lx = threadIdx.x & 31, r0 = (threadIdx.x >> 5) * 4;
  float v[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i][0] = v[i][1] = 0.f;
#pragma unroll
  for (int e = 0; e < 14; ++e) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const float h = hq[q][r0 + e][lx];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = e - i;
        if (k < 0 || k > 10) continue;
        v[i][q] = fmaf(g.w[k], h, v[i][q]);
      }
    }
  }
  const int px = ox + lx;
  if (px >= W) return;
  float* const s0 = stats + c * plane;
  float* const s1 = s0 + (size_t)gridDim.z * plane;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int py = oy + r0 + i;
    if (py >= H) break;
    const uint32_t o = pix(px, py, W);
    s0[o] = v[i][0];
    s1[o] = v[i][1];
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
  // eight pairs per thread requested at once (indices past the end read the
  // last pair and add nothing); each thread still adds its pairs in order
  const float2* __restrict__ pp = reinterpret_cast<const float2*>(partials);
  for (int i0 = threadIdx.x; i0 < nblocks; i0 += 256 * 8) {
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = pp[min(i0 + 256 * u, nblocks - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (i0 + 256 * u < nblocks) {
        s += v[u].x;
        l += v[u].y;
      }
    }
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
  // the coefficient patches until every row blur has read them, then the
  // row-blurred coefficients (one buffer: more work-groups per CU)
  __shared__ __align__(16) union {
    float sa[3][kPH][kPS];
    float hq[3][kPH][kHS];
  } u;
  const Tile tl = xcd_tile();
  const int c = tl.bz;
  const int ox = tl.bx * kTW, oy = tl.by * kTH;
  const size_t plane = (size_t)H * W;
  const size_t all = (size_t)C * plane;
  const int lx = threadIdx.x & 31, r0 = (threadIdx.x >> 5) * 4;
  const int px = ox + lx;
  // the thread's own pixels of x and y, loaded while the patches stage
  float xv[4], yv[4];
  {
    const float* const xc = x + c * plane;
    const float* const yc = y + c * plane;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t o = pix(min(px, W - 1), min(oy + r0 + i, H - 1), W);
      xv[i] = xc[o];
      yv[i] = yc[o];
    }
  }
  const float scale = dloss[c / group];
  {
    float(*const dst[3])[kPS] = {u.sa[0], u.sa[1], u.sa[2]};
    const float* const src[3] = {coef + c * plane, coef + all + c * plane, coef + 2 * all + c * plane};
    stage_patches<3>(dst, src, H, W, ox, oy);
  }
  __syncthreads();
  // horizontal: four outputs per task and map, held in registers until every
  // task has read the patches
  static_assert(kHTasks <= 2 * 256, "two row tasks per thread at most");
  float res[2][3][4];
#pragma unroll
  for (int rep = 0; rep < 2; ++rep) {
    const int t = threadIdx.x + rep * 256;
    if (t >= kHTasks) break;
    const int r = t / kQuads, j = t - r * kQuads;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      float a[14];
      load14(u.sa[m], r, j, a);
#pragma unroll
      for (int i = 0; i < 4; ++i) res[rep][m][i] = 0.f;
#pragma unroll
      for (int e = 0; e < 14; ++e)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = e - i;
          if (k < 0 || k > 10) continue;
          res[rep][m][i] = fmaf(g.w[k], a[e], res[rep][m][i]);
        }
    }
  }
  __syncthreads();
#pragma unroll
  for (int rep = 0; rep < 2; ++rep) {
    const int t = threadIdx.x + rep * 256;
    if (t >= kHTasks) break;
    const int r = t / kQuads, j = t - r * kQuads;
#pragma unroll
    for (int m = 0; m < 3; ++m)
      *reinterpret_cast<float4*>(&u.hq[m][r][4 * j]) =
          make_float4(res[rep][m][0], res[rep][m][1], res[rep][m][2], res[rep][m][3]);
  }
  __syncthreads();
  float b[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i) b[i][0] = b[i][1] = b[i][2] = 0.f;
#pragma unroll
  for (int e = 0; e < 14; ++e)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const float h = u.hq[m][r0 + e][lx];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = e - i;
        if (k < 0 || k > 10) continue;
        b[i][m] = fmaf(g.w[k], h, b[i][m]);
      }
    }
  if (px >= W) return;
  const float g_l1 = (1.f - lam) / (float)((size_t)group * plane);
  float* const dxc = dx + c * plane;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int py = oy + r0 + i;
    if (py >= H) break;
    const uint32_t o = pix(px, py, W);
    const float d = xv[i] - yv[i];
    const float sgn = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
    dxc[o] = scale * (b[i][0] + 2.f * xv[i] * b[i][1] + yv[i] * b[i][2] + g_l1 * sgn);
  }
}

void launch_dssim_target(const float* y, int C, int H, int W, const float* w11, float* stats, hipStream_t s) {
  Window11 g;
  for (int k = 0; k < 11; ++k) g.w[k] = w11[k];
  const dim3 grid((W + kTW - 1) / kTW, (H + kTH - 1) / kTH, C);
  hipLaunchKernelGGL(dssim_target_kernel, grid, dim3(256), 0, s, y, H, W, g, stats);
}

void launch_dssim_forward(const float* x, const float* y, const float* ystats, int C, int group, int H, int W,
                          const float* w11, float lam, float* coef, float* partials, float* out, hipStream_t s) {
  Window11 g;
  for (int k = 0; k < 11; ++k) g.w[k] = w11[k];
  const dim3 grid((W + kTW - 1) / kTW, (H + kTH - 1) / kTH, C);
  if (ystats)
    hipLaunchKernelGGL(dssim_fwd_kernel<true>, grid, dim3(256), 0, s, x, y, ystats, H, W, group, g, lam, coef,
                       partials);
  else
    hipLaunchKernelGGL(dssim_fwd_kernel<false>, grid, dim3(256), 0, s, x, y, ystats, H, W, group, g, lam, coef,
                       partials);
  const int per_group = (int)(grid.x * grid.y) * group;
  hipLaunchKernelGGL(dssim_reduce_kernel, dim3(C / group), dim3(256), 0, s, (const float*)partials, per_group,
                     1.0f / (float)((size_t)group * H * W), lam, out);
}

void launch_dssim_backward(const float* x, const float* y, const float* coef, int C, int group, int H, int W,
                           const float* w11, float lam, const float* dloss, float* dx, hipStream_t s) {
  Window11 g;
  for (int k = 0; k < 11; ++k) g.w[k] = w11[k];
  const dim3 grid((W + kTW - 1) / kTW, (H + kTH - 1) / kTH, C);
  hipLaunchKernelGGL(dssim_bwd_kernel, grid, dim3(256), 0, s, x, y, coef, H, W, C, group, g, lam, dloss, dx);
}

size_t dssim_partials(int C, int H, int W) {
  return (size_t)2 * ((W + kTW - 1) / kTW) * ((H + kTH - 1) / kTH) * C;
}

}  // namespace dmgs
