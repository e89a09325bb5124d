"""Camera matrices with the reference's conventions.

* ``getWorld2View2``      — utils/graphics_utils.py:38-49
* ``getProjectionMatrix`` — utils/graphics_utils.py:51-71 (znear / zfar, P[3,2] = 1)
* ``fov2focal`` / ``focal2fov`` — utils/graphics_utils.py:73-77
* ``MiniCam``             — scene/cameras.py:59-70
* ``SimpleCamera``        — the fields of scene/cameras.py:17-57 that the renderer
  reads (no image): world_view_transform = W2C^T, full_proj_transform =
  world_view @ P^T, camera_center = inverse(world_view)[3, :3], znear 0.01,
  zfar 100 (scene/cameras.py:48-57).
* ``look_at`` / ``orbit_cameras`` — seeded orbit cameras for the multi-view
  and serving workloads (COLMAP axes: x right, y down, z forward).
"""
import math

import numpy as np
import torch


def getWorld2View2(R, t, translate=np.array([0.0, 0.0, 0.0]), scale=1.0):
    Rt = np.zeros((4, 4))
    Rt[:3, :3] = np.asarray(R).transpose()
    Rt[:3, 3] = t
    Rt[3, 3] = 1.0
    C2W = np.linalg.inv(Rt)
    C2W[:3, 3] = (C2W[:3, 3] + translate) * scale
    return np.float32(np.linalg.inv(C2W))


def getProjectionMatrix(znear, zfar, fovX, fovY):
    tan_y, tan_x = math.tan(fovY / 2), math.tan(fovX / 2)
    top, right = tan_y * znear, tan_x * znear
    bottom, left = -top, -right
    P = torch.zeros(4, 4)
    P[0, 0] = 2.0 * znear / (right - left)
    P[1, 1] = 2.0 * znear / (top - bottom)
    P[0, 2] = (right + left) / (right - left)
    P[1, 2] = (top + bottom) / (top - bottom)
    P[3, 2] = 1.0
    P[2, 2] = zfar / (zfar - znear)
    P[2, 3] = -(zfar * znear) / (zfar - znear)
    return P


def fov2focal(fov, pixels):
    return pixels / (2 * math.tan(fov / 2))


def focal2fov(focal, pixels):
    return 2 * math.atan(pixels / (2 * focal))


class MiniCam:
    def __init__(self, width, height, fovy, fovx, znear, zfar, world_view_transform, full_proj_transform):
        self.image_width = width
        self.image_height = height
        self.FoVy = fovy
        self.FoVx = fovx
        self.znear = znear
        self.zfar = zfar
        self.world_view_transform = world_view_transform
        self.full_proj_transform = full_proj_transform
        view_inv = torch.inverse(self.world_view_transform)
        self.camera_center = view_inv[3][:3]


class SimpleCamera:
    """Renderer-facing fields of the reference Camera, built on ``device``."""

    def __init__(self, R, T, FoVx, FoVy, width, height, device="cuda", znear=0.01, zfar=100.0,
                 trans=np.array([0.0, 0.0, 0.0]), scale=1.0):
        self.R, self.T = np.asarray(R), np.asarray(T)
        self.FoVx, self.FoVy = float(FoVx), float(FoVy)
        self.image_width, self.image_height = int(width), int(height)
        self.znear, self.zfar = znear, zfar
        wv = torch.tensor(getWorld2View2(self.R, self.T, trans, scale)).transpose(0, 1)
        proj = getProjectionMatrix(znear, zfar, self.FoVx, self.FoVy).transpose(0, 1)
        full = (wv.unsqueeze(0).bmm(proj.unsqueeze(0))).squeeze(0)
        self.world_view_transform = wv.to(device)
        self.projection_matrix = proj.to(device)
        self.full_proj_transform = full.to(device)
        self.camera_center = wv.inverse()[3, :3].to(device)


def look_at(position, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """Camera-to-world rotation for a camera at ``position`` looking at
    ``target``; (0, 0, -d) gives the identity."""
    p = np.asarray(position, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - p
    z /= np.linalg.norm(z)
    x = np.cross(np.asarray(up, dtype=np.float64), z)
    if np.linalg.norm(x) < 1e-8:
        x = np.cross(np.array([1.0, 0.0, 0.0]), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=1)


def camera_at(position, width, height, fov_deg=50.0, device="cuda"):
    """Pinhole camera at ``position`` looking at the origin; fovx = fov_deg and
    fovy from the same focal length."""
    R = look_at(position)
    T = -R.T @ np.asarray(position, dtype=np.float64)
    fovx = math.radians(fov_deg)
    fovy = focal2fov(fov2focal(fovx, width), height)
    return SimpleCamera(R, T, fovx, fovy, width, height, device=device)


def front_camera(width, height, fov_deg=50.0, distance=4.0, device="cuda"):
    """R = I, T = (0, 0, distance): centre (0, 0, -distance) looking along +z."""
    f = math.radians(fov_deg)
    return SimpleCamera(np.eye(3), np.array([0.0, 0.0, distance]), f, f, width, height, device=device)


def sphere_positions(n, radius, seed):
    """n seeded camera positions uniformly on a sphere of the given radius
    (radius may be a (lo, hi) range, drawn uniformly per camera)."""
    g = np.random.default_rng(seed)
    v = g.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    if isinstance(radius, (tuple, list)):
        r = g.uniform(radius[0], radius[1], size=(n, 1))
    else:
        r = np.full((n, 1), float(radius))
    return v * r
