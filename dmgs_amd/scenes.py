"""Seeded synthetic scenes for the benchmark workloads (all FP32, generated on
the CPU with ``torch.Generator`` so every device sees identical bytes).

* ``FreeGaussians``  — the reference model protocol (scene/gaussian_model.py:26-41,
  95-118): raw parameters ``_xyz, _features_dc, _features_rest, _scaling (log),
  _rotation (unnormalised), _opacity (logit)`` with activations exp / normalize /
  sigmoid in the getters.
* ``free_scene``     — means U[-1,1]^3 ("uniform") or a 60/40 mixture of the unit
  ball interior and a shell ("mixture"); log-scales N(-3.5, 0.4); quaternions
  N(0,1)^4 (normalised by the getter); opacity sigmoid(N(0,1)); SH degree 3
  with DC N(0, 0.5) and the rest N(0, 0.1).
* ``icosphere``      — subdivided icosahedron on the unit sphere.
"""
import math

import numpy as np
import torch


class FreeGaussians:
    """Free 3D Gaussians with the reference getters."""

    def __init__(self, xyz, features_dc, features_rest, scaling, rotation, opacity, sh_degree=3):
        self._xyz = xyz
        self._features_dc = features_dc
        self._features_rest = features_rest
        self._scaling = scaling
        self._rotation = rotation
        self._opacity = opacity
        self.active_sh_degree = sh_degree
        self.max_sh_degree = sh_degree

    def parameters(self):
        return [self._xyz, self._features_dc, self._features_rest, self._scaling, self._rotation, self._opacity]

    def requires_grad_(self, flag=True):
        for p in self.parameters():
            p.requires_grad_(flag)
        return self

    def to(self, device):
        return FreeGaussians(*[p.detach().to(device) for p in self.parameters()], sh_degree=self.max_sh_degree)

    @property
    def num(self):
        return self._xyz.shape[0]

    # Renders of an unchanged scene without autograd (viewers, serving,
    # evaluation) reuse the activated tensors: a getter called under no_grad
    # returns the value it computed last for the same parameter tensors at the
    # same version (any in-place update or replacement of a parameter
    # recomputes it).  With autograd on, every call builds its own graph, as
    # the reference getters do.
    def _memo(self, name, fn, *params):
        if torch.is_grad_enabled():
            return fn()
        key = tuple((p.data_ptr(), p._version, tuple(p.shape)) for p in params)
        memo = self.__dict__.setdefault("_memo_cache", {})
        hit = memo.get(name)
        if hit is not None and hit[0] == key:
            return hit[1]
        value = fn()
        memo[name] = (key, value)
        return value

    @property
    def get_scaling(self):
        return self._memo("scaling", lambda: torch.exp(self._scaling), self._scaling)

    @property
    def get_rotation(self):
        return self._memo("rotation", lambda: torch.nn.functional.normalize(self._rotation), self._rotation)

    @property
    def get_xyz(self):
        return self._xyz

    @property
    def get_features(self):
        return self._memo("features", lambda: torch.cat((self._features_dc, self._features_rest), dim=1),
                          self._features_dc, self._features_rest)

    @property
    def get_opacity(self):
        return self._memo("opacity", lambda: torch.sigmoid(self._opacity), self._opacity)

    # ---- densification statistics (the consumer of radii and the NDC
    # means2D gradient; scene/gaussian_model.py:405-407, train.py:116-118) ----
    def _stats(self):
        if getattr(self, "xyz_gradient_accum", None) is None or self.xyz_gradient_accum.shape[0] != self.num:
            dev = self._xyz.device
            self.xyz_gradient_accum = torch.zeros((self.num, 1), device=dev)
            self.denom = torch.zeros((self.num, 1), device=dev)
            self.max_radii2D = torch.zeros((self.num,), device=dev)

    def add_densification_stats(self, viewspace_point_tensor, update_filter):
        """Accumulate |dL/dmean2D[:2]| (NDC) and the view count of the visible
        Gaussians of one rendered view."""
        self._stats()
        self.xyz_gradient_accum[update_filter] += torch.norm(viewspace_point_tensor.grad[update_filter, :2], dim=-1,
                                                             keepdim=True)
        self.denom[update_filter] += 1

    def update_max_radii2D(self, radii, visibility_filter):
        """Keep the largest screen radius of each visible Gaussian."""
        self._stats()
        self.max_radii2D[visibility_filter] = torch.max(self.max_radii2D[visibility_filter], radii[visibility_filter])

    def get_covariance(self, scaling_modifier=1.0):
        """build_covariance_from_scaling_rotation (scene/gaussian_model.py:27-31)
        with the raw rotation, as the reference passes ``self._rotation``."""
        s = self.get_scaling * scaling_modifier
        q = self._rotation / self._rotation.norm(dim=1, keepdim=True)
        r, x, y, z = q.unbind(-1)
        R = torch.stack([
            1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y),
            2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x),
            2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)], -1).reshape(-1, 3, 3)
        L = R * s.unsqueeze(-2)
        S = L @ L.transpose(1, 2)
        return torch.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], -1)


def _means(n, g, layout, shell_radius):
    if layout == "uniform":
        return torch.rand(n, 3, generator=g) * 2 - 1
    # 60/40 mixture: unit-ball interior, then a shell of radius shell_radius
    n_in = int(round(0.6 * n))
    d = torch.randn(n, 3, generator=g)
    d = d / d.norm(dim=1, keepdim=True)
    r = torch.empty(n, 1)
    r[:n_in] = torch.rand(n_in, 1, generator=g) ** (1.0 / 3.0)
    r[n_in:] = shell_radius
    return d * r


def free_scene(n, seed=0, layout="uniform", shell_radius=3.0, sh_degree=3, device="cpu"):
    g = torch.Generator().manual_seed(seed)
    xyz = _means(n, g, layout, shell_radius)
    scaling = torch.randn(n, 3, generator=g) * 0.4 - 3.5
    rotation = torch.randn(n, 4, generator=g)
    opacity = torch.randn(n, 1, generator=g)
    m = (sh_degree + 1) ** 2
    dc = torch.randn(n, 1, 3, generator=g) * 0.5
    rest = torch.randn(n, m - 1, 3, generator=g) * 0.1
    model = FreeGaussians(xyz, dc, rest, scaling, rotation, opacity, sh_degree)
    return model.to(device) if device != "cpu" else model


def target_image(H, W, seed=1, device="cpu"):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(3, H, W, generator=g).to(device)


def icosphere(subdivisions):
    """Unit icosphere: returns (verts (V,3) float32, faces (F,3) int64);
    F = 20 * 4^subdivisions."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, dtype=np.int64)
    for _ in range(subdivisions):
        edges = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]], 0)
        edges = np.sort(edges, axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        mids = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = verts.shape[0]
        verts = np.concatenate([verts, mids], 0)
        nf = faces.shape[0]
        inv = inv.reshape(-1)
        m01, m12, m20 = base + inv[:nf], base + inv[nf:2 * nf], base + inv[2 * nf:]
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        faces = np.concatenate([np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1),
                                np.stack([c, m20, m12], 1), np.stack([m01, m12, m20], 1)], 0)
    return torch.tensor(verts, dtype=torch.float32), torch.tensor(faces, dtype=torch.int64)
