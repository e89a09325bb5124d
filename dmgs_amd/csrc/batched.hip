// Multi-view forward/backward of the per-Gaussian stages for gfx950.
//
// A batch renders B views of one scene.  The scene is held component-major
// (59 channels x P, see DMGS_MV_* below): each thread owns one Gaussian, reads
// its raw parameters once with fully coalesced 4-byte lanes, applies the
// model activations of scene/gaussian_model.py:26-41,95-118 (exp scale,
// normalised quaternion, sigmoid opacity, SH = [dc, rest]) and produces the
// splat records of all B views.  The backward walks the same B views,
// accumulates the gradient of every view in registers (3D covariance, SH and
// opacity gradients are view-independent sums) and writes the raw-parameter
// gradients once, through the activation derivatives.
//
// Per-view math is the single-view kernels' (gs_math.h), so a batch of B
// views gives the same images and the same summed gradients as B single-view
// launches followed by autograd through the activations.
#include "dmgs_common.h"
#include "dmgs_kernels.h"
#include "gs_math.h"

namespace dmgs {

// Activations with the accurate expf and IEEE division, as torch evaluates
// exp / sigmoid / normalize in FP32, so the activated values match the
// single-view path fed by the model getters.
__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

struct MVGaussian {
  float3 p, logs, s;  // position, log-scale, modifier * exp(log-scale)
  float4 rq, q;       // raw and normalised quaternion
  float rq_inv;       // 1 / max(|rq|, 1e-12)
  float ol, o;        // opacity logit and sigmoid
  float R[9];
  float c3[6];
  float sh[48];
};

__device__ __forceinline__ void load_mv(const MVArgs& a, int g, MVGaussian& G) {
  const size_t P = (size_t)a.P;
  const float* pr = a.params;
  auto ch = [&](int c) { return pr[(size_t)c * P + g]; };
  G.p = make_float3(ch(DMGS_MV_XYZ), ch(DMGS_MV_XYZ + 1), ch(DMGS_MV_XYZ + 2));
  G.logs = make_float3(ch(DMGS_MV_SCALE), ch(DMGS_MV_SCALE + 1), ch(DMGS_MV_SCALE + 2));
  const float m = a.scale_modifier;
  G.s = make_float3(m * expf(G.logs.x), m * expf(G.logs.y), m * expf(G.logs.z));
  G.rq = make_float4(ch(DMGS_MV_ROT), ch(DMGS_MV_ROT + 1), ch(DMGS_MV_ROT + 2), ch(DMGS_MV_ROT + 3));
  const float nrm = fmaxf(sqrtf(G.rq.x * G.rq.x + G.rq.y * G.rq.y + G.rq.z * G.rq.z + G.rq.w * G.rq.w), 1e-12f);
  G.rq_inv = 1.0f / nrm;
  G.q = make_float4(G.rq.x / nrm, G.rq.y / nrm, G.rq.z / nrm, G.rq.w / nrm);
  G.ol = ch(DMGS_MV_OPACITY);
  G.o = sigmoidf(G.ol);
#pragma unroll
  for (int k = 0; k < 48; ++k) G.sh[k] = ch(DMGS_MV_SH + k);
  quat_to_R(G.q, G.R);
  cov3d_from_sR(G.s, G.R, G.c3);
}

// Depth sort key of a visible entry.  View-major keys: the top byte of a
// positive float's bits is its exponent field over two, so a visible depth
// d > 0.2 (>= 2^-3) has a top byte of at least 0x3E, and d < 2^27 keeps it at
// most 0x4C: the top byte fits 4 bits (0..14) after subtracting 0x3E and the
// freed 4 bits carry the view.  Sorting (view << 28 | (top - 0x3E) << 24 |
// low 24 bits) orders by view, then by depth exactly as the float bits do.
// A depth outside that range gets the far bucket (bits 24-27 all set, low
// bits zero, so never kSortSkip), which the sort reports in bit 0 of its
// status word.
__device__ __forceinline__ uint32_t entry_sort_key(float depth, int view, bool view_major) {
  const uint32_t bits = __float_as_uint(depth);
  if (!view_major) return bits;
  const uint32_t top = bits >> 24;
  if (top < 0x3Eu || top - 0x3Eu > 14u) return ((uint32_t)view << 28) | kFarBucketBits;
  return ((uint32_t)view << 28) | ((top - 0x3Eu) << 24) | (bits & 0xFFFFFFu);
}

__global__ __launch_bounds__(256) void mv_preprocess_kernel(MVArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (a.status && g == 0) *a.status = 0u;  // the depth sort sets its flags after this kernel
  const bool live = g < a.P;  // every lane stays for the per-view ballots
  MVGaussian G;
  if (live) load_mv(a, g, G);
  for (int v = 0; v < a.B; ++v) {
    const ViewParams vp = view_params(const_ptr(a.views)[__builtin_amdgcn_readfirstlane(v)]);
    const size_t i = (size_t)v * a.P + g;
    SplatGeom o;
    const bool visible = live && project_splat(G.p, G.c3, vp, o);
    if (!live) continue;
    if (!visible) {  // the radius and the sort's skip key are all that is read of a culled entry
      a.radii[i] = 0;
      a.depth_keys[i] = kSortSkip;
      continue;
    }
    float3 dir = make_float3(G.p.x - vp.campos[0], G.p.y - vp.campos[1], G.p.z - vp.campos[2]);
    const float inv_len = 1.0f / sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
    dir = make_float3(dir.x * inv_len, dir.y * inv_len, dir.z * inv_len);
    float3 rgb = sh_to_rgb(a.sh_degree, G.sh, dir);
    rgb = make_float3(fmaxf(rgb.x, 0.0f), fmaxf(rgb.y, 0.0f), fmaxf(rgb.z, 0.0f));
    float4* sp = a.splats + (size_t)DMGS_SPLAT_F4 * i;
    sp[0] = make_float4(o.pix.x, o.pix.y, o.conic.x, o.conic.y);
    sp[1] = make_float4(o.conic.z, G.o, rgb.x, rgb.y);
    sp[2] = make_float4(rgb.z, o.depth, __int_as_float(o.radius), 0.0f);
    a.radii[i] = o.radius;
    a.rects[i] = pack_rect(o.rmin, o.rmax);  // binning reads rectangles in depth order
    a.depth_keys[i] = entry_sort_key(o.depth, v, a.view_major != 0);
  }
}

// The backward is two kernels over the Gaussians, each reading only its plane
// of the accumulator rows and few enough registers for several waves per
// SIMD: mv_geometry_bwd_kernel (projection, EWA and 3D-covariance chain,
// opacity, densification statistics; writes the position gradient of those
// paths and the scale, rotation and opacity channels) and then
// mv_sh_bwd_kernel (SH coefficients and the view-direction term of the
// position gradient, added to what the first kernel wrote).
__global__ __launch_bounds__(256) void mv_geometry_bwd_kernel(MVArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.P) return;
  const size_t P = (size_t)a.P;
  const size_t N = (size_t)a.B * P;
  const float* pr = a.params;
  auto ch = [&](int c) { return pr[(size_t)c * P + g]; };
  const float3 p = make_float3(ch(DMGS_MV_XYZ), ch(DMGS_MV_XYZ + 1), ch(DMGS_MV_XYZ + 2));
  const float3 logs = make_float3(ch(DMGS_MV_SCALE), ch(DMGS_MV_SCALE + 1), ch(DMGS_MV_SCALE + 2));
  const float4 rq = make_float4(ch(DMGS_MV_ROT), ch(DMGS_MV_ROT + 1), ch(DMGS_MV_ROT + 2), ch(DMGS_MV_ROT + 3));
  const float ol = ch(DMGS_MV_OPACITY);
  // the activations exactly as load_mv computes them
  const float m = a.scale_modifier;
  const float3 sc = make_float3(m * expf(logs.x), m * expf(logs.y), m * expf(logs.z));
  const float nrm = fmaxf(sqrtf(rq.x * rq.x + rq.y * rq.y + rq.z * rq.z + rq.w * rq.w), 1e-12f);
  const float rq_inv = 1.0f / nrm;
  const float4 q = make_float4(rq.x / nrm, rq.y / nrm, rq.z / nrm, rq.w / nrm);
  const float o = sigmoidf(ol);
  float R[9], c3[6];
  quat_to_R(q, R);
  cov3d_from_sR(sc, R, c3);
  float3 dmean = make_float3(0.f, 0.f, 0.f);
  float dc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float dop = 0.f;
  bool any = false;
  // densification statistics (gaussian_model.py:405-407, train.py:116-118)
  float d_acc = 0.f, d_cnt = 0.f, d_maxr = 0.f;
  if (a.dens) {
    d_acc = a.dens[g];
    d_cnt = a.dens[P + g];
    d_maxr = a.dens[2 * P + g];
  }
  // the next view's radius and row are requested before this view is
  // computed (rows of culled entries hold zeros and are skipped)
  auto row = [&](int v, float2 (&r)[3]) {
    const float2* gr = reinterpret_cast<const float2*>(a.gacc + gacc_index(N, (size_t)v * P + g, 0));
    r[0] = gr[0];
    r[1] = gr[1];
    r[2] = gr[2];
  };
  int rad_next = a.radii[g];
  float2 nx[3];
  row(0, nx);
  for (int v = 0; v < a.B; ++v) {
    const int radius = rad_next;
    const float2 r0 = nx[0], r1 = nx[1], r2 = nx[2];
    if (v + 1 < a.B) {
      rad_next = a.radii[(size_t)(v + 1) * P + g];
      row(v + 1, nx);
    }
    if (radius <= 0) continue;
    any = true;
    const ViewParams vp = view_params(const_ptr(a.views)[__builtin_amdgcn_readfirstlane(v)]);
    if (a.dens) {
      d_acc += sqrtf(r0.x * r0.x + r0.y * r0.y);
      d_cnt += 1.0f;
      d_maxr = fmaxf(d_maxr, (float)radius);
    }
    const float3 dm = ewa_backward(p, c3, vp, r0.x, r0.y, r1.x, r1.y, r2.x, dc);
    dmean.x += dm.x;
    dmean.y += dm.y;
    dmean.z += dm.z;
    dop += r2.y;
  }
  float3 dlogs = make_float3(0.f, 0.f, 0.f);
  float4 drq = make_float4(0.f, 0.f, 0.f, 0.f);
  if (any) {
    float3 ds;
    float4 dq;
    cov3d_backward(dc, sc, R, q, ds, dq);
    // scale = exp(log-scale): d/dlog = dL/dscale * exp(log) (dL/dscale is taken w.r.t. mod * scale)
    dlogs = make_float3(ds.x * expf(logs.x), ds.y * expf(logs.y), ds.z * expf(logs.z));
    // q = rq / |rq|: d/drq = (dq - q (q . dq)) / |rq|
    const float qd = q.x * dq.x + q.y * dq.y + q.z * dq.z + q.w * dq.w;
    drq = make_float4((dq.x - q.x * qd) * rq_inv, (dq.y - q.y * qd) * rq_inv, (dq.z - q.z * qd) * rq_inv,
                      (dq.w - q.w * qd) * rq_inv);
  }
  if (a.dens) {
    a.dens[g] = d_acc;
    a.dens[P + g] = d_cnt;
    a.dens[2 * P + g] = d_maxr;
  }
  float* gd = a.grads;
  auto put = [&](int c, float val) { gd[(size_t)c * P + g] = val; };
  put(DMGS_MV_XYZ, dmean.x);
  put(DMGS_MV_XYZ + 1, dmean.y);
  put(DMGS_MV_XYZ + 2, dmean.z);
  put(DMGS_MV_SCALE, dlogs.x);
  put(DMGS_MV_SCALE + 1, dlogs.y);
  put(DMGS_MV_SCALE + 2, dlogs.z);
  put(DMGS_MV_ROT, drq.x);
  put(DMGS_MV_ROT + 1, drq.y);
  put(DMGS_MV_ROT + 2, drq.z);
  put(DMGS_MV_ROT + 3, drq.w);
  put(DMGS_MV_OPACITY, dop * o * (1.0f - o));
}

// One thread's 48 SH coefficients in a [48][256] LDS array.
struct ShColumn {
  const float* base;
  __device__ __forceinline__ float operator[](int k) const { return base[k * 256]; }
};

__global__ __launch_bounds__(256) void mv_sh_bwd_kernel(MVArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.P) return;
  const size_t P = (size_t)a.P;
  const size_t N = (size_t)a.B * P;
  const float* pr = a.params;
  auto ch = [&](int c) { return pr[(size_t)c * P + g]; };
  const float3 p = make_float3(ch(DMGS_MV_XYZ), ch(DMGS_MV_XYZ + 1), ch(DMGS_MV_XYZ + 2));
  // the coefficients wait in LDS (column threadIdx.x: conflict-free) rather
  // than in 48 registers, which leaves room for a third wave per SIMD
  __shared__ float s_sh[48][256];
  float dsh[48];
#pragma unroll
  for (int k = 0; k < 48; ++k) {
    s_sh[k][threadIdx.x] = ch(DMGS_MV_SH + k);
    dsh[k] = 0.f;
  }
  ShColumn sh{&s_sh[0][threadIdx.x]};
  float3 dmean = make_float3(0.f, 0.f, 0.f);
  auto col = [&](int v) {
    const float* gc = a.gacc + gacc_index(N, (size_t)v * P + g, DMGS_GACC_GEO);
    return make_float3(gc[0], gc[1], gc[2]);
  };
  int rad_next = a.radii[g];
  float3 col_next = col(0);
  for (int v = 0; v < a.B; ++v) {
    const int radius = rad_next;
    const float3 dcol = col_next;
    if (v + 1 < a.B) {  // requested before this view is computed
      rad_next = a.radii[(size_t)(v + 1) * P + g];
      col_next = col(v + 1);
    }
    if (radius <= 0) continue;
    const ViewParams vp = view_params(const_ptr(a.views)[__builtin_amdgcn_readfirstlane(v)]);
    float b[16];
    float3 gm;
    const float3 dd = sh_backward(a.sh_degree, sh, p, vp.campos, dcol, b, gm);
    dmean.x += dd.x;
    dmean.y += dd.y;
    dmean.z += dd.z;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      dsh[3 * k] += b[k] * gm.x;
      dsh[3 * k + 1] += b[k] * gm.y;
      dsh[3 * k + 2] += b[k] * gm.z;
    }
  }
  float* gd = a.grads;
  gd[(size_t)DMGS_MV_XYZ * P + g] += dmean.x;
  gd[(size_t)(DMGS_MV_XYZ + 1) * P + g] += dmean.y;
  gd[(size_t)(DMGS_MV_XYZ + 2) * P + g] += dmean.z;
#pragma unroll
  for (int k = 0; k < 48; ++k) gd[(size_t)(DMGS_MV_SH + k) * P + g] = dsh[k];
}

__global__ __launch_bounds__(256) void gather_rects_kernel(const uint2* __restrict__ rects,
                                                           const uint32_t* __restrict__ order,
                                                           uint2* __restrict__ rects_sorted,
                                                           uint32_t* __restrict__ counts, size_t n,
                                                           const uint32_t* __restrict__ n_dev) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (n_dev && i >= *n_dev) {  // past the sorted entries: no instances
    counts[i] = 0u;
    return;
  }
  const uint2 r = rects[order[i]];
  rects_sorted[i] = r;
  int2 rmin, rmax;
  unpack_rect(r, rmin, rmax);
  counts[i] = (rmax.x > rmin.x && rmax.y > rmin.y) ? (uint32_t)cell_count(rmin, rmax) : 0u;
}

void launch_gather_rects(const uint2* rects, const uint32_t* order, uint2* rects_sorted, uint32_t* counts, size_t n,
                         hipStream_t s, const uint32_t* n_dev) {
  if (n == 0) return;
  hipLaunchKernelGGL(gather_rects_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rects, order,
                     rects_sorted, counts, n, n_dev);
}

void launch_mv_preprocess(const MVArgs& a, hipStream_t s) {
  if (a.P <= 0 || a.B <= 0) return;
  hipLaunchKernelGGL(mv_preprocess_kernel, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
}

void launch_mv_geometry_bwd(const MVArgs& a, hipStream_t s) {
  if (a.P <= 0 || a.B <= 0) return;
  hipLaunchKernelGGL(mv_geometry_bwd_kernel, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(mv_sh_bwd_kernel, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
}

}  // namespace dmgs
