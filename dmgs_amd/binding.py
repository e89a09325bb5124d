"""Mesh -> Gaussian binding on gfx950 (forward and backward into the vertices).

* ``bind_stage3`` / ``FaceBoundGaussians`` — the stage-3 model
  (scene/gaussian_geo_model_finetune.py:396-482): vertices are parameters;
  per face a tangent frame, per Gaussian barycentric means, rotation
  R_t2w R_t(_rotation) as a quaternion and scale (exp(_scaling), thin_z).
  ``FaceBoundGaussians`` exposes the reference model protocol (get_xyz,
  get_rotation, get_scaling, get_opacity, get_features, active_sh_degree,
  max_sh_degree) so the drop-in ``render`` takes it unchanged, and supports
  the optional frustum face culling of ``renew_gaussian`` (:33-48, :405-409).
* ``bind_stage2`` — the stage-2 covariance binding
  (scene/gaussian_geo_model_mlp_flex.py:258-310, 370-385): means plus one
  covariance (R_t2w L k)(R_t2w L k)^T per face for ``render_dyn``'s
  precomputed-covariance path; L carries no gradient, k = tanh(scale_factor)
  * max_scale does.
"""
import math

import torch

from . import _native as N


def _i32(faces):
    return faces.to(dtype=torch.int32).contiguous()


def _f32(t):
    t = t.to(dtype=torch.float32).contiguous()
    return t.clone() if t.data_ptr() % 16 else t


class _Stage3(torch.autograd.Function):
    @staticmethod
    def forward(ctx, verts, rot2d, scale2d, faces, bc, G, thin_z):
        lib = N.load()
        dev = verts.device
        v, f, b = _f32(verts), _i32(faces), _f32(bc)
        r2, s2 = _f32(rot2d), _f32(scale2d)
        F = f.shape[0]
        n = F * G
        per_g = int(b.shape[0] == n and n != G)
        xyz = torch.empty(n, 3, device=dev)
        rot = torch.empty(n, 4, device=dev)
        scale = torch.empty(n, 3, device=dev)
        N.check(lib.dmgs_bind_forward(v.shape[0], N.ptr(v), F, N.ptr(f), G, N.ptr(b), per_g, N.ptr(r2), N.ptr(s2),
                                      float(thin_z), N.ptr(xyz), N.ptr(rot), N.ptr(scale), N.stream_handle(dev)),
                "dmgs_bind_forward")
        ctx.save_for_backward(v, f, b, r2, s2)
        ctx.G, ctx.per_g = G, per_g
        return xyz, rot, scale

    @staticmethod
    def backward(ctx, dxyz, drot, dscale):
        v, f, b, r2, s2 = ctx.saved_tensors
        dev = v.device
        n = f.shape[0] * ctx.G
        z = lambda *s: torch.zeros(*s, device=dev)
        dxyz = z(n, 3) if dxyz is None else _f32(dxyz)
        drot = z(n, 4) if drot is None else _f32(drot)
        dscale = z(n, 3) if dscale is None else _f32(dscale)
        dverts = z(*v.shape)
        drot2d = torch.empty(n, 2, device=dev)
        dscale2d = torch.empty(n, 2, device=dev)
        N.check(N.load().dmgs_bind_backward(v.shape[0], N.ptr(v), f.shape[0], N.ptr(f), ctx.G, N.ptr(b), ctx.per_g,
                                            N.ptr(r2), N.ptr(s2), N.ptr(dxyz), N.ptr(drot), N.ptr(dscale),
                                            N.ptr(dverts), N.ptr(drot2d), N.ptr(dscale2d), N.stream_handle(dev)),
                "dmgs_bind_backward")
        return dverts, drot2d, dscale2d, None, None, None, None


def bind_stage3(verts, faces, bc, rot2d, scale2d, thin_z):
    """Returns (xyz, rotation quaternion (r, x, y, z), scale) for F*G Gaussians."""
    if faces.shape[0] == 0 or rot2d.shape[0] % faces.shape[0]:
        raise ValueError("rot2d must hold G entries per face")
    G = rot2d.shape[0] // faces.shape[0]
    if bc.shape[0] not in (G, faces.shape[0] * G):
        raise ValueError("bc must be (G, 3) or (F*G, 3)")
    return _Stage3.apply(verts, rot2d, scale2d, faces, bc, G, thin_z)


class _Stage2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, verts, k, faces, bc, G, rad_base, flat_scale, adaptive):
        lib = N.load()
        dev = verts.device
        v, f, b = _f32(verts), _i32(faces), _f32(bc)
        kk = _f32(k.reshape(1))
        F = f.shape[0]
        n = F * G
        per_g = int(b.shape[0] == n and n != G)
        xyz = torch.empty(n, 3, device=dev)
        cov = torch.empty(n, 6, device=dev)
        L = torch.empty(F, 3, 3, device=dev)
        N.check(lib.dmgs_bind_cov_forward(v.shape[0], N.ptr(v), F, N.ptr(f), G, N.ptr(b), per_g, N.ptr(kk),
                                          float(rad_base), float(flat_scale), int(bool(adaptive)), N.ptr(xyz),
                                          N.ptr(cov), N.ptr(L), N.stream_handle(dev)), "dmgs_bind_cov_forward")
        ctx.save_for_backward(v, f, b, kk, L)
        ctx.G, ctx.per_g = G, per_g
        ctx.mark_non_differentiable(L)
        return xyz, cov, L

    @staticmethod
    def backward(ctx, dxyz, dcov, _dL):
        v, f, b, kk, L = ctx.saved_tensors
        dev = v.device
        n = f.shape[0] * ctx.G
        dxyz = torch.zeros(n, 3, device=dev) if dxyz is None else _f32(dxyz)
        dcov = torch.zeros(n, 6, device=dev) if dcov is None else _f32(dcov)
        dverts = torch.zeros_like(v)
        dk = torch.zeros(1, device=dev)
        N.check(N.load().dmgs_bind_cov_backward(v.shape[0], N.ptr(v), f.shape[0], N.ptr(f), ctx.G, N.ptr(b),
                                                ctx.per_g, N.ptr(kk), N.ptr(L), N.ptr(dxyz), N.ptr(dcov),
                                                N.ptr(dverts), N.ptr(dk), N.stream_handle(dev)),
                "dmgs_bind_cov_backward")
        return dverts, dk.reshape(()), None, None, None, None, None, None


def bind_stage2(verts, faces, bc, scale_factor, rad_base, max_scale=2.0, spatial_lr_scale=1.0, adaptive=True,
                gaussians_per_face=None):
    """Returns (xyz, covariance 6-vectors, L (F,3,3)).  bc is (G,3) shared by
    every face (as in the reference) unless ``gaussians_per_face`` is given
    with a per-Gaussian (F*G,3) bc."""
    G = bc.shape[0] if gaussians_per_face is None else int(gaussians_per_face)
    if bc.shape[0] not in (G, faces.shape[0] * G):
        raise ValueError("bc must be (G, 3) or (F*G, 3)")
    k = torch.tanh(scale_factor.reshape(())) * max_scale
    return _Stage2.apply(verts, k, faces, bc, G, rad_base, spatial_lr_scale * 1e-6, adaptive)


def faces_in_frustum(verts, faces, full_proj_transform):
    lib = N.load()
    v, f = _f32(verts.detach()), _i32(faces)
    p = _f32(full_proj_transform.to(v.device))
    mask = torch.empty(f.shape[0], dtype=torch.uint8, device=v.device)
    N.check(lib.dmgs_faces_in_frustum(v.shape[0], N.ptr(v), f.shape[0], N.ptr(f), N.ptr(p), N.ptr(mask),
                                      N.stream_handle(v.device)), "dmgs_faces_in_frustum")
    return mask.bool()


def barycentric_v2(num, device="cpu"):
    """generate_barycentric_v2 (geo/mesh_utils.py:15-40): fixed barycentric
    sets for 1, 3 or 6 Gaussians per face and the matching radius base."""
    s3 = 3 ** 0.5
    if num == 1:
        bc, r = [[1 / 3, 1 / 3, 1 / 3]], 1.0 / (2.0 * s3)
    elif num == 3:
        a, c = (3 - s3) / 6, s3 / 3
        bc, r = [[a, a, c], [a, c, a], [c, a, a]], 1.0 / (2.0 + 2.0 * s3)
    elif num == 6:
        bc = [[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3], [1 / 6, 5 / 12, 5 / 12],
              [5 / 12, 1 / 6, 5 / 12], [5 / 12, 5 / 12, 1 / 6]]
        r = 1 / (4.0 + 2.0 * s3)
    else:
        raise NotImplementedError
    return torch.tensor(bc, device=device), r


class FaceBoundGaussians:
    """Stage-3 face-bound Gaussians with the reference model protocol."""

    def __init__(self, verts, faces, bc, rot2d, scale2d, thin_z, features_dc, features_rest, opacity,
                 sh_degree=3, use_frustum=False):
        self.verts = verts
        self.faces = faces
        self.bc_coords = bc
        self._rotation = rot2d
        self._scaling = scale2d
        self.thin_z_scale = float(thin_z)
        self._features_dc = features_dc
        self._features_rest = features_rest
        self._opacity = opacity
        self.active_sh_degree = sh_degree
        self.max_sh_degree = sh_degree
        self.use_frustum = use_frustum
        self.G = rot2d.shape[0] // faces.shape[0]
        self.gs_mask = None

    def parameters(self):
        return [self.verts, self._rotation, self._scaling, self._features_dc, self._features_rest, self._opacity]

    def requires_grad_(self, flag=True):
        for p in self.parameters():
            p.requires_grad_(flag)
        return self

    def renew_gaussian(self, viewpoint_cam=None):
        faces, rot, scl, bc = self.faces, self._rotation, self._scaling, self.bc_coords
        self.gs_mask = None
        if self.use_frustum and viewpoint_cam is not None:
            fm = faces_in_frustum(self.verts, faces, viewpoint_cam.full_proj_transform)
            self.gs_mask = torch.repeat_interleave(fm, self.G)
            faces, rot, scl = faces[fm], rot[self.gs_mask], scl[self.gs_mask]
            if bc.shape[0] != self.G:
                bc = bc[self.gs_mask]
        self._xyz, self._rot, self._scale = bind_stage3(self.verts, faces, bc, rot, scl, self.thin_z_scale)

    def _masked(self, t):
        return t if self.gs_mask is None else t[self.gs_mask]

    @property
    def get_xyz(self):
        return self._xyz

    @property
    def get_rotation(self):
        return self._rot

    @property
    def get_scaling(self):
        return self._scale

    @property
    def get_features(self):
        return torch.cat((self._masked(self._features_dc), self._masked(self._features_rest)), dim=1)

    @property
    def get_opacity(self):
        return torch.sigmoid(self._masked(self._opacity))


def mesh_scene(subdivisions=6, gaussians_per_face=6, jitter=0.005, seed=0, sh_degree=3, device="cpu"):
    """The mesh-bound workload: icosphere, vertices jittered N(0, jitter),
    Dirichlet(1,1,1) barycentrics per Gaussian, R_t = I (rotation (1, 0)),
    in-plane scale 0.3 sqrt(face area), normal scale 1e-4, SH degree 3
    (DC N(0, 0.5), rest N(0, 0.1)), opacity sigmoid(N(0, 1))."""
    from .scenes import icosphere
    g = torch.Generator().manual_seed(seed)
    verts, faces = icosphere(subdivisions)
    verts = verts + torch.randn(verts.shape, generator=g) * jitter
    F, G = faces.shape[0], gaussians_per_face
    e = -torch.log(torch.rand(F * G, 3, generator=g))            # Dirichlet(1,1,1) via exponentials
    bc = e / e.sum(1, keepdim=True)
    v = verts[faces]
    area = 0.5 * torch.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], dim=-1).norm(dim=-1)
    s = torch.log(0.3 * area.sqrt()).repeat_interleave(G)
    scale2d = torch.stack([s, s], 1)
    rot2d = torch.tensor([[1.0, 0.0]]).repeat(F * G, 1)
    m = (sh_degree + 1) ** 2
    dc = torch.randn(F * G, 1, 3, generator=g) * 0.5
    rest = torch.randn(F * G, m - 1, 3, generator=g) * 0.1
    opacity = torch.randn(F * G, 1, generator=g)
    to = lambda t: t.to(device)
    return FaceBoundGaussians(to(verts), to(faces), to(bc), to(rot2d), to(scale2d), 1e-4, to(dc), to(rest),
                              to(opacity), sh_degree)
