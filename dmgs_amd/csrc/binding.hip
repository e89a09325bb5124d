// Mesh -> Gaussian binding for gfx950, forward and backward into the mesh
// vertices.
//
// Stage 3 (scene/gaussian_geo_model_finetune.py:396-482): per face the
// tangent frame R_t2w = [x y n] with x = normalize(v1 - v0),
// n = normalize((v1 - v0) x (v2 - v0)), y = normalize(n x x); per Gaussian
// xyz = sum_i bc_i v_i, R = R_t2w R_t with R_t the in-plane rotation by
// (a, b) = normalize(_rotation), q = normalize(matrix_to_quaternion(R)),
// scale = (exp(s0), exp(s1), thin_z).
// Stage 2 (scene/gaussian_geo_model_mlp_flex.py:258-335, 370-385): the same
// frame and means, and one covariance per face Sigma = (R_t2w L k)(R_t2w L k)^T
// with L the affine triangle matrix built without gradient and
// k = tanh(scale_factor) * max_scale, shared by the face's Gaussians.
//
// Backward: one thread per face walks its G Gaussians, sums their
// contributions to the face frame and the three vertices, back-propagates
// through the frame construction once, and adds into the vertex gradient
// with 9 atomics per face.  The quaternion gradient is mapped to the rotation
// through the body-frame angular gradient g_w = 1/2 Im(conj(q) * dL/dq): any
// variation of R here is a rotation, so a matrix gradient that agrees with the
// true one on rotation tangents gives exact vertex gradients.
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {

namespace {

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 load3(const float* p, int i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

// F.normalize semantics: v / max(|v|, 1e-12)
__device__ __forceinline__ float inv_norm(V3 v) { return 1.0f / fmaxf(sqrtf(dot(v, v)), 1e-12f); }

// d normalize(u) / du applied to g: (g - y (y.g)) / |u|  (y = normalize(u))
__device__ __forceinline__ V3 normalize_bwd(V3 y, float inv, V3 g) { return inv * (g - dot(y, g) * y); }

struct Frame {
  V3 e1, e2, w, u;  // e1 = v1 - v0, e2 = v2 - v0, w = e1 x e2, u = n x x
  float ie1, iw, iu;
  V3 x, y, n;       // columns of R_t2w
};

__device__ __forceinline__ Frame make_frame(V3 v0, V3 v1, V3 v2) {
  Frame f;
  f.e1 = v1 - v0;
  f.e2 = v2 - v0;
  f.ie1 = inv_norm(f.e1);
  f.x = f.ie1 * f.e1;
  f.w = cross(f.e1, f.e2);
  f.iw = inv_norm(f.w);
  f.n = f.iw * f.w;
  f.u = cross(f.n, f.x);
  f.iu = inv_norm(f.u);
  f.y = f.iu * f.u;
  return f;
}

// Back-propagate column gradients (gx, gy, gn) of R_t2w to (de1, de2).
__device__ __forceinline__ void frame_bwd(const Frame& f, V3 gx, V3 gy, V3 gn, V3& de1, V3& de2) {
  const V3 gu = normalize_bwd(f.y, f.iu, gy);      // y = normalize(u)
  gn = gn + cross(f.x, gu);                         // u = n x x
  gx = gx + cross(gu, f.n);
  const V3 gw = normalize_bwd(f.n, f.iw, gn);      // n = normalize(w)
  de1 = normalize_bwd(f.x, f.ie1, gx) + cross(f.e2, gw);  // x = normalize(e1), w = e1 x e2
  de2 = cross(gw, f.e1);
}

// matrix_to_quaternion (published algorithm of pytorch3d, real part first):
// candidates from the four "1 +- m00 +- m11 +- m22" square roots, the one with
// the largest root is kept (first on ties), then the real part is made >= 0.
__device__ __forceinline__ float4 mat_to_quat(const float R[3][3]) {
  const float m00 = R[0][0], m01 = R[0][1], m02 = R[0][2];
  const float m10 = R[1][0], m11 = R[1][1], m12 = R[1][2];
  const float m20 = R[2][0], m21 = R[2][1], m22 = R[2][2];
  float qa[4] = {1.f + m00 + m11 + m22, 1.f + m00 - m11 - m22, 1.f - m00 + m11 - m22, 1.f - m00 - m11 + m22};
  for (int i = 0; i < 4; ++i) qa[i] = qa[i] > 0.f ? sqrtf(qa[i]) : 0.f;
  int best = 0;
  for (int i = 1; i < 4; ++i)
    if (qa[i] > qa[best]) best = i;
  float4 c;
  if (best == 0) c = make_float4(qa[0] * qa[0], m21 - m12, m02 - m20, m10 - m01);
  else if (best == 1) c = make_float4(m21 - m12, qa[1] * qa[1], m10 + m01, m02 + m20);
  else if (best == 2) c = make_float4(m02 - m20, m10 + m01, qa[2] * qa[2], m12 + m21);
  else c = make_float4(m10 - m01, m20 + m02, m21 + m12, qa[3] * qa[3]);
  const float d = 1.f / (2.f * fmaxf(qa[best], 0.1f));
  c = make_float4(c.x * d, c.y * d, c.z * d, c.w * d);
  if (c.x < 0.f) c = make_float4(-c.x, -c.y, -c.z, -c.w);
  return c;
}

__device__ __forceinline__ void rotation_of(const Frame& f, float a, float b, float R[3][3]) {
  // columns: a x + b y, -b x + a y, n
  const V3 c0 = a * f.x + b * f.y, c1 = (-b) * f.x + a * f.y;
  R[0][0] = c0.x; R[1][0] = c0.y; R[2][0] = c0.z;
  R[0][1] = c1.x; R[1][1] = c1.y; R[2][1] = c1.z;
  R[0][2] = f.n.x; R[1][2] = f.n.y; R[2][2] = f.n.z;
}

__device__ __forceinline__ int face_vertex(const int32_t* faces, int f, int k) { return faces[3 * f + k]; }

}  // namespace

// --------------------------------------------------------------- stage 3 --
__global__ __launch_bounds__(256) void bind_fwd_kernel(BindArgs a) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= a.F * a.G) return;
  const int f = gi / a.G, j = gi - f * a.G;
  const V3 v0 = load3(a.verts, face_vertex(a.faces, f, 0));
  const V3 v1 = load3(a.verts, face_vertex(a.faces, f, 1));
  const V3 v2 = load3(a.verts, face_vertex(a.faces, f, 2));
  const float* bc = a.bc + 3 * (a.bc_per_gaussian ? gi : j);
  const V3 p = bc[0] * v0 + bc[1] * v1 + bc[2] * v2;
  a.xyz[3 * gi] = p.x;
  a.xyz[3 * gi + 1] = p.y;
  a.xyz[3 * gi + 2] = p.z;
  if (a.rot2d) {
    const Frame fr = make_frame(v0, v1, v2);
    const float ra = a.rot2d[2 * gi], rb = a.rot2d[2 * gi + 1];
    const float ic = 1.0f / fmaxf(sqrtf(ra * ra + rb * rb), 1e-12f);
    float R[3][3];
    rotation_of(fr, ra * ic, rb * ic, R);
    float4 q = mat_to_quat(R);
    const float iq = 1.0f / fmaxf(sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w), 1e-12f);
    reinterpret_cast<float4*>(a.rot)[gi] = make_float4(q.x * iq, q.y * iq, q.z * iq, q.w * iq);
  }
  if (a.scale2d) {
    a.scale[3 * gi] = __expf(a.scale2d[2 * gi]);
    a.scale[3 * gi + 1] = __expf(a.scale2d[2 * gi + 1]);
    a.scale[3 * gi + 2] = a.thin_z;
  }
}

__global__ __launch_bounds__(256) void bind_bwd_kernel(BindArgs a) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.F) return;
  const int i0 = face_vertex(a.faces, f, 0), i1 = face_vertex(a.faces, f, 1), i2 = face_vertex(a.faces, f, 2);
  const V3 v0 = load3(a.verts, i0), v1 = load3(a.verts, i1), v2 = load3(a.verts, i2);
  const Frame fr = make_frame(v0, v1, v2);
  V3 dv0 = v3(0.f, 0.f, 0.f), dv1 = dv0, dv2 = dv0;
  V3 gx = dv0, gy = dv0, gn = dv0;  // gradient of the frame columns
  for (int j = 0; j < a.G; ++j) {
    const int gi = f * a.G + j;
    const float* bc = a.bc + 3 * (a.bc_per_gaussian ? gi : j);
    const V3 dp = load3(a.dxyz, gi);
    dv0 = dv0 + bc[0] * dp;
    dv1 = dv1 + bc[1] * dp;
    dv2 = dv2 + bc[2] * dp;
    if (a.rot2d && a.drot) {
      const float ra = a.rot2d[2 * gi], rb = a.rot2d[2 * gi + 1];
      const float rr = ra * ra + rb * rb;
      const float ic = 1.0f / fmaxf(sqrtf(rr), 1e-12f);
      const float ca = ra * ic, cb = rb * ic;
      float R[3][3];
      rotation_of(fr, ca, cb, R);
      float4 q = mat_to_quat(R);
      const float iq = 1.0f / fmaxf(sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w), 1e-12f);
      q = make_float4(q.x * iq, q.y * iq, q.z * iq, q.w * iq);
      const float4 g = reinterpret_cast<const float4*>(a.drot)[gi];
      // g_w = 1/2 Im(conj(q) * g), quaternions (w, x, y, z)
      const V3 qv = v3(q.y, q.z, q.w), gv = v3(g.y, g.z, g.w);
      const V3 gw = 0.5f * (q.x * gv - g.x * qv - cross(qv, gv));
      // rotation about z by theta = atan2(rb, ra): dL/dtheta = gw.z
      if (a.drot2d) {
        const float s = rr > 0.f ? gw.z / rr : 0.f;
        a.drot2d[2 * gi] = -rb * s;
        a.drot2d[2 * gi + 1] = ra * s;
      }
      // frame angular gradient g1 = R_t gw; column gradients 1/2 R_t2w (g1 x e_j)
      const V3 g1 = v3(ca * gw.x - cb * gw.y, cb * gw.x + ca * gw.y, gw.z);
      // R_t2w (g1 x e_x) = R_t2w (0, g1.z, -g1.y) = g1.z y - g1.y n, etc.
      gx = gx + 0.5f * (g1.z * fr.y - g1.y * fr.n);
      gy = gy + 0.5f * (g1.x * fr.n - g1.z * fr.x);
      gn = gn + 0.5f * (g1.y * fr.x - g1.x * fr.y);
    }
    if (a.scale2d && a.dscale && a.dscale2d) {
      a.dscale2d[2 * gi] = a.dscale[3 * gi] * __expf(a.scale2d[2 * gi]);
      a.dscale2d[2 * gi + 1] = a.dscale[3 * gi + 1] * __expf(a.scale2d[2 * gi + 1]);
    }
  }
  if (a.dcov) {
    // stage 2: Sigma = A A^T, A = R_t2w L k, shared by the face's Gaussians
    float Gs[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int j = 0; j < a.G; ++j)
      for (int k = 0; k < 6; ++k) Gs[k] += a.dcov[6 * (f * a.G + j) + k];
    const float Gm[3][3] = {{Gs[0], 0.5f * Gs[1], 0.5f * Gs[2]},
                            {0.5f * Gs[1], Gs[3], 0.5f * Gs[4]},
                            {0.5f * Gs[2], 0.5f * Gs[4], Gs[5]}};
    const float* L = a.Lmat + 9 * f;
    const float k = a.k[0];
    const float Rw[3][3] = {{fr.x.x, fr.y.x, fr.n.x}, {fr.x.y, fr.y.y, fr.n.y}, {fr.x.z, fr.y.z, fr.n.z}};
    float B[3][3];  // B = L L^T k^2
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        B[r][c] = k * k * (L[3 * r] * L[3 * c] + L[3 * r + 1] * L[3 * c + 1] + L[3 * r + 2] * L[3 * c + 2]);
    // dL/dR = 2 G R B ; dL/dk = 2/k <G, R L L^T R^T k^2>
    float RB[3][3], dR[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) RB[r][c] = Rw[r][0] * B[0][c] + Rw[r][1] * B[1][c] + Rw[r][2] * B[2][c];
    float gk = 0.f;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        dR[r][c] = 2.f * (Gm[r][0] * RB[0][c] + Gm[r][1] * RB[1][c] + Gm[r][2] * RB[2][c]);
        const float S = RB[r][0] * Rw[c][0] + RB[r][1] * Rw[c][1] + RB[r][2] * Rw[c][2];
        gk += Gm[r][c] * S;
      }
    gx = gx + v3(dR[0][0], dR[1][0], dR[2][0]);
    gy = gy + v3(dR[0][1], dR[1][1], dR[2][1]);
    gn = gn + v3(dR[0][2], dR[1][2], dR[2][2]);
    if (a.dk && k != 0.f) unsafeAtomicAdd(a.dk, 2.f * gk / k);
  }
  V3 de1, de2;
  frame_bwd(fr, gx, gy, gn, de1, de2);
  dv0 = dv0 - de1 - de2;
  dv1 = dv1 + de1;
  dv2 = dv2 + de2;
  unsafeAtomicAdd(a.dverts + 3 * i0, dv0.x);
  unsafeAtomicAdd(a.dverts + 3 * i0 + 1, dv0.y);
  unsafeAtomicAdd(a.dverts + 3 * i0 + 2, dv0.z);
  unsafeAtomicAdd(a.dverts + 3 * i1, dv1.x);
  unsafeAtomicAdd(a.dverts + 3 * i1 + 1, dv1.y);
  unsafeAtomicAdd(a.dverts + 3 * i1 + 2, dv1.z);
  unsafeAtomicAdd(a.dverts + 3 * i2, dv2.x);
  unsafeAtomicAdd(a.dverts + 3 * i2 + 1, dv2.y);
  unsafeAtomicAdd(a.dverts + 3 * i2 + 2, dv2.z);
}

// --------------------------------------------------------------- stage 2 --
// Per face: L (affine triangle matrix, no gradient) and Sigma for its Gaussians.
__global__ __launch_bounds__(256) void bind_cov_fwd_kernel(BindArgs a) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.F) return;
  const V3 v0 = load3(a.verts, face_vertex(a.faces, f, 0));
  const V3 v1 = load3(a.verts, face_vertex(a.faces, f, 1));
  const V3 v2 = load3(a.verts, face_vertex(a.faces, f, 2));
  const Frame fr = make_frame(v0, v1, v2);
  // v1v2_len is clamped at 1e-12 (mlp_flex.py:268-270)
  const float len = fmaxf(sqrtf(dot(fr.e1, fr.e1)), 1e-12f);
  float m01 = 0.f, m11 = 1.f;
  if (a.adaptive) {
    // M = [[1, (A0 - E0) / E1], [0, A1 / E1]] with A = (e2.x, e2.y) in the face
    // frame and E = (len / 2, len sqrt(3) / 2) (mlp_flex.py:289-303)
    const float A0 = dot(fr.e2, fr.x), A1 = dot(fr.e2, fr.y);
    const float E0 = len * 0.5f, E1 = len * 0.8660254037844386f;
    m01 = (A0 - E0) / E1;
    m11 = A1 / E1;
  }
  const float s = len * a.rad_base;
  float L[9] = {s, m01 * s, 0.f, 0.f, m11 * s, 0.f, 0.f, 0.f, a.flat_scale};
  float* Lo = a.Lmat + 9 * f;
  for (int t = 0; t < 9; ++t) Lo[t] = L[t];
  const float k = a.k[0];
  const float Rw[3][3] = {{fr.x.x, fr.y.x, fr.n.x}, {fr.x.y, fr.y.y, fr.n.y}, {fr.x.z, fr.y.z, fr.n.z}};
  float A[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = k * (Rw[r][0] * L[c] + Rw[r][1] * L[3 + c] + Rw[r][2] * L[6 + c]);
  float S[6];
  S[0] = A[0][0] * A[0][0] + A[0][1] * A[0][1] + A[0][2] * A[0][2];
  S[1] = A[0][0] * A[1][0] + A[0][1] * A[1][1] + A[0][2] * A[1][2];
  S[2] = A[0][0] * A[2][0] + A[0][1] * A[2][1] + A[0][2] * A[2][2];
  S[3] = A[1][0] * A[1][0] + A[1][1] * A[1][1] + A[1][2] * A[1][2];
  S[4] = A[1][0] * A[2][0] + A[1][1] * A[2][1] + A[1][2] * A[2][2];
  S[5] = A[2][0] * A[2][0] + A[2][1] * A[2][1] + A[2][2] * A[2][2];
  for (int j = 0; j < a.G; ++j) {
    const int gi = f * a.G + j;
    for (int t = 0; t < 6; ++t) a.cov[6 * gi + t] = S[t];
    const float* bc = a.bc + 3 * (a.bc_per_gaussian ? gi : j);
    const V3 p = bc[0] * v0 + bc[1] * v1 + bc[2] * v2;
    a.xyz[3 * gi] = p.x;
    a.xyz[3 * gi + 1] = p.y;
    a.xyz[3 * gi + 2] = p.z;
  }
}

// Frustum test of face centroids (finetune.py:33-48): |ndc| < 1.05 on all
// three axes and w > 0, with p_hom = c @ P (row-vector convention).
__global__ __launch_bounds__(256) void faces_in_frustum_kernel(int F, const float* __restrict__ verts,
                                                               const int32_t* __restrict__ faces,
                                                               const float* __restrict__ proj,
                                                               uint8_t* __restrict__ mask) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const V3 c = (1.0f / 3.0f) * (load3(verts, faces[3 * f]) + load3(verts, faces[3 * f + 1]) +
                                load3(verts, faces[3 * f + 2]));
  const float4 ph = xform_point_4x4(make_float3(c.x, c.y, c.z), proj);
  const float w = ph.w + 1e-6f;
  const bool in = fabsf(ph.x / w) < 1.05f && fabsf(ph.y / w) < 1.05f && fabsf(ph.z / w) < 1.05f && w > 0.f;
  mask[f] = in ? 1 : 0;
}

void launch_bind_fwd(const BindArgs& a, hipStream_t s) {
  const int n = a.F * a.G;
  if (n <= 0) return;
  hipLaunchKernelGGL(bind_fwd_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a);
}

void launch_bind_cov_fwd(const BindArgs& a, hipStream_t s) {
  if (a.F <= 0) return;
  hipLaunchKernelGGL(bind_cov_fwd_kernel, dim3((a.F + 255) / 256), dim3(256), 0, s, a);
}

void launch_bind_bwd(const BindArgs& a, hipStream_t s) {
  if (a.F <= 0) return;
  hipLaunchKernelGGL(bind_bwd_kernel, dim3((a.F + 255) / 256), dim3(256), 0, s, a);
}

void launch_faces_in_frustum(int F, const float* verts, const int32_t* faces, const float* proj, uint8_t* mask,
                             hipStream_t s) {
  if (F <= 0) return;
  hipLaunchKernelGGL(faces_in_frustum_kernel, dim3((F + 255) / 256), dim3(256), 0, s, F, verts, faces, proj, mask);
}

}  // namespace dmgs
