"""Mesh -> Gaussian binding in FP64 torch (test infrastructure only).

Restates
* stage 3: scene/gaussian_geo_model_finetune.py:396-421 (renew_gaussian: face
  normals via geo/mesh_utils.py:43-57, face_x = normalize(v1 - v0),
  face_y = normalize(n x x), R_t2w = [x y n], barycentric means),
  :445-482 (get_scaling = (exp(_scaling), thin_z), get_rot_matrix
  R = R_t2w R_t(normalize(_rotation)), get_rotation =
  normalize(matrix_to_quaternion(R)));
* stage 2: scene/gaussian_geo_model_mlp_flex.py:258-310 (frame, means, affine
  triangle matrix M and L under no_grad) and :370-385 (get_covariance_dyn);
* in_frustum: scene/gaussian_geo_model_finetune.py:33-48.

matrix_to_quaternion is pytorch3d's published algorithm (pytorch3d is
installed from its git head by the reference, README.md:16; no version is
pinned): four candidates from sqrt(max(0, 1 +- m00 +- m11 +- m22)), the one
with the largest root divided by 2 max(root, 0.1), real part made
non-negative.  The sign choice is "parity unpinned" (no pytorch3d here); it
does not affect R(q), the covariance or any gradient, which the tests pin.
Gradients are torch autograd of these forwards.

Pins (tests/test_oracle_binding.py): R^T R = I, det R = +1, n orthogonal to
both edges, x along v1 - v0; means inside the triangle with the barycentrics
recovered; R(q) = R; for an equilateral triangle M = I; the covariance's
eigenvalues are (k s)^2, (k m11 s)^2 ... through L L^T; autograd equals
central finite differences.
"""
import torch
import torch.nn.functional as F

F64 = torch.float64


def face_frame(verts, faces):
    fv = verts[faces]
    e1 = fv[:, 1] - fv[:, 0]
    e2 = fv[:, 2] - fv[:, 0]
    n = F.normalize(torch.cross(e1, e2, dim=-1))
    x = F.normalize(e1)
    y = F.normalize(torch.cross(n, x, dim=-1))
    return torch.stack([x, y, n], dim=2), fv


def means(fv, bc, G):
    F_ = fv.shape[0]
    if bc.shape[0] == G:
        b = bc.view(1, G, 3).expand(F_, G, 3)
    else:
        b = bc.view(F_, G, 3)
    return torch.matmul(b, fv).reshape(-1, 3)


def matrix_to_quaternion(R):
    m = R.reshape(R.shape[:-2] + (9,))
    m00, m01, m02, m10, m11, m12, m20, m21, m22 = m.unbind(-1)
    raw = torch.stack([1.0 + m00 + m11 + m22, 1.0 + m00 - m11 - m22, 1.0 - m00 + m11 - m22,
                       1.0 - m00 - m11 + m22], -1)
    q_abs = torch.where(raw > 0, torch.sqrt(torch.clamp(raw, min=1e-300)), torch.zeros_like(raw))
    cand = torch.stack([
        torch.stack([q_abs[..., 0] ** 2, m21 - m12, m02 - m20, m10 - m01], -1),
        torch.stack([m21 - m12, q_abs[..., 1] ** 2, m10 + m01, m02 + m20], -1),
        torch.stack([m02 - m20, m10 + m01, q_abs[..., 2] ** 2, m12 + m21], -1),
        torch.stack([m10 - m01, m20 + m02, m21 + m12, q_abs[..., 3] ** 2], -1)], -2)
    cand = cand / (2.0 * torch.clamp(q_abs, min=0.1))[..., None]
    pick = torch.nn.functional.one_hot(q_abs.argmax(-1), 4).to(torch.bool)
    out = cand[pick].reshape(R.shape[:-2] + (4,))
    return torch.where(out[..., :1] < 0, -out, out)


def quat_to_rot(q):
    q = F.normalize(q, dim=-1)
    r, x, y, z = q.unbind(-1)
    return torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y),
                        2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x),
                        2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)],
                       -1).reshape(q.shape[:-1] + (3, 3))


def stage3(verts, faces, bc, rot2d, scale2d, thin_z):
    """Returns (xyz, q, scale, R) for F*G Gaussians (FP64, differentiable)."""
    F_ = faces.shape[0]
    G = rot2d.shape[0] // F_
    R_t2w, fv = face_frame(verts, faces)
    xyz = means(fv, bc, G)
    c = F.normalize(rot2d, dim=1)
    zero = torch.zeros_like(c[:, 0])
    one = torch.ones_like(c[:, 0])
    R_t = torch.stack([c[:, 0], -c[:, 1], zero, c[:, 1], c[:, 0], zero, zero, zero, one], -1).view(F_, G, 3, 3)
    R = torch.matmul(R_t2w.view(F_, 1, 3, 3), R_t).view(-1, 3, 3)
    q = F.normalize(matrix_to_quaternion(R), dim=-1)
    scale = torch.cat([torch.exp(scale2d), torch.full((scale2d.shape[0], 1), thin_z, dtype=scale2d.dtype)], 1)
    return xyz, q, scale, R


def triangle_L(verts, faces, rad_base, flat_scale, adaptive=True):
    """cov3D_L of mlp_flex.py:285-310, computed without gradient."""
    with torch.no_grad():
        R_t2w, fv = face_frame(verts, faces)
        v1v2 = fv[:, 1] - fv[:, 0]
        length = torch.norm(v1v2, dim=-1).clamp_min(1e-12)
        v1v3 = fv[:, 2] - fv[:, 0]
        n_face = faces.shape[0]
        M = torch.zeros(n_face, 2, 2, dtype=verts.dtype)
        M[:, 0, 0] = 1.0
        if adaptive:
            A0 = (v1v3 * R_t2w[:, :, 0]).sum(-1)
            A1 = (v1v3 * R_t2w[:, :, 1]).sum(-1)
            E0, E1 = length / 2.0, length * (3 ** 0.5) / 2
            M[:, 0, 1] = (A0 - E0) / E1
            M[:, 1, 1] = A1 / E1
        else:
            M[:, 1, 1] = 1.0
        L = torch.zeros(n_face, 3, 3, dtype=verts.dtype)
        L[:, :2, :2] = M * (length * rad_base).view(-1, 1, 1)
        L[:, 2, 2] = flat_scale
    return L


def stage2(verts, faces, bc, G, k, rad_base, flat_scale, adaptive=True):
    """Returns (xyz, cov6, L); k is the scalar tanh(scale_factor) * max_scale."""
    R_t2w, fv = face_frame(verts, faces)
    xyz = means(fv, bc, G)
    L = triangle_L(verts, faces, rad_base, flat_scale, adaptive)
    A = torch.matmul(R_t2w, L * k)
    S = A @ A.transpose(1, 2)
    cov = torch.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], -1)
    return xyz, cov.view(-1, 1, 6).expand(-1, G, 6).reshape(-1, 6), L


def in_frustum(proj, verts, faces):
    c = verts[faces].mean(dim=1)
    p = c @ proj[:3, :] + proj[3:, :]
    w = p[:, 3:] + 1e-6
    q = p[:, :3] / w
    return (q.abs() < 1.05).all(dim=-1) & (w.squeeze(-1) > 0)
