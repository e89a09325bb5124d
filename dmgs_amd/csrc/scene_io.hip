// Scene and frame conversions at the edges of the path, for gfx950.
//
// * PLY vertex rows <-> component-major scene parameters.  A 3DGS PLY
//   (scene/gaussian_model.py:191-256) stores one row of float32 properties
//   per Gaussian; the batched kernels read 59 component-major channels
//   (xyz, log-scale, raw quaternion, opacity logit, SH coefficient k of
//   colour c at 11 + 3k + c).  Both directions are a gather through a column
//   map, done as an LDS tile transpose: a workgroup reads 64 whole rows
//   (contiguous) and writes 64 consecutive Gaussians of every channel
//   (256-byte runs), so both sides stay coalesced.
// * Frames for the viewer wire format (train.py:60, network_gui.py:51-55):
//   a (3, H, W) float image becomes H x W x 3 bytes of
//   uint8(clamp(x, 0, 1) * 255) (truncation, as torch's .byte()).
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

constexpr int kIoRows = 64;  // Gaussians per workgroup

__global__ __launch_bounds__(256) void rows_to_channels_kernel(const float* __restrict__ rows, int64_t P, int row_len,
                                                               ChannelMap map, float* __restrict__ params) {
  __shared__ float tile[kIoRows * kMaxRowFloats];
  const int64_t g0 = (int64_t)blockIdx.x * kIoRows;
  const int n = (int)min((int64_t)kIoRows, P - g0);
  const float* src = rows + g0 * row_len;
  for (int i = threadIdx.x; i < n * row_len; i += 256) tile[i] = src[i];
  __syncthreads();
  for (int i = threadIdx.x; i < map.channels * kIoRows; i += 256) {
    const int c = i / kIoRows, j = i - c * kIoRows;
    if (j >= n) continue;
    const int col = map.col[c];
    params[(size_t)c * P + g0 + j] = col >= 0 ? tile[j * row_len + col] : 0.f;
  }
}

__global__ __launch_bounds__(256) void channels_to_rows_kernel(const float* __restrict__ params, int64_t P,
                                                               int row_len, ChannelMap map, float* __restrict__ rows) {
  __shared__ float tile[kIoRows * kMaxRowFloats];
  const int64_t g0 = (int64_t)blockIdx.x * kIoRows;
  const int n = (int)min((int64_t)kIoRows, P - g0);
  // here map.col[j] is the parameter channel of row column j (-1: zero)
  for (int i = threadIdx.x; i < row_len * kIoRows; i += 256) {
    const int col = i / kIoRows, j = i - col * kIoRows;
    if (j >= n) continue;
    const int c = map.col[col];
    tile[j * row_len + col] = c >= 0 ? params[(size_t)c * P + g0 + j] : 0.f;
  }
  __syncthreads();
  float* dst = rows + g0 * row_len;
  for (int i = threadIdx.x; i < n * row_len; i += 256) dst[i] = tile[i];
}

void launch_rows_to_channels(const float* rows, int64_t P, int row_len, const ChannelMap& map, float* params,
                             hipStream_t s) {
  if (P <= 0) return;
  hipLaunchKernelGGL(rows_to_channels_kernel, dim3((unsigned)((P + kIoRows - 1) / kIoRows)), dim3(256), 0, s, rows, P,
                     row_len, map, params);
}

void launch_channels_to_rows(const float* params, int64_t P, int row_len, const ChannelMap& map, float* rows,
                             hipStream_t s) {
  if (P <= 0) return;
  hipLaunchKernelGGL(channels_to_rows_kernel, dim3((unsigned)((P + kIoRows - 1) / kIoRows)), dim3(256), 0, s,
                     params, P, row_len, map, rows);
}

// One thread per pixel: three planar loads, one packed 3-byte store.
__global__ __launch_bounds__(256) void rgb8_kernel(const float* __restrict__ chw, int64_t plane,
                                                   uint8_t* __restrict__ hwc) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= plane) return;
  uint8_t* o = hwc + 3 * p;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float x = chw[c * plane + p];
    const float v = fminf(fmaxf(x, 0.f), 1.f) * 255.f;  // NaN -> 0
    o[c] = (uint8_t)(int)v;
  }
}

void launch_rgb8(const float* chw, int64_t plane, uint8_t* hwc, hipStream_t s) {
  if (plane <= 0) return;
  hipLaunchKernelGGL(rgb8_kernel, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, chw, plane, hwc);
}

}  // namespace dmgs
