"""Camera matrices, restated from the reference (test infrastructure only).

* ``world_to_view``   — utils/graphics_utils.py:38-49 (getWorld2View2)
* ``projection``      — utils/graphics_utils.py:51-71 (getProjectionMatrix)
* ``fov2focal``       — utils/graphics_utils.py:73-74
* ``make_camera``     — scene/cameras.py:48-57: world_view_transform = W2C^T,
                        full_proj = world_view @ P^T, camera_center =
                        inverse(world_view)[3, :3]; znear 0.01, zfar 100.

Pins (tests/test_oracle_sh_camera.py): frustum corners map to NDC +-1, FoV 90 deg
gives focal = W/2, the camera centre maps to the view-space origin.
"""
import math

import numpy as np
import torch


def world_to_view(R, t, translate=(0.0, 0.0, 0.0), scale=1.0):
    """getWorld2View2 (utils/graphics_utils.py:38-49): 4x4 world->camera, float32."""
    Rt = np.zeros((4, 4))
    Rt[:3, :3] = np.asarray(R, dtype=np.float64).T
    Rt[:3, 3] = np.asarray(t, dtype=np.float64)
    Rt[3, 3] = 1.0
    c2w = np.linalg.inv(Rt)
    centre = (c2w[:3, 3] + np.asarray(translate, dtype=np.float64)) * scale
    c2w[:3, 3] = centre
    return np.float32(np.linalg.inv(c2w))


def projection(znear, zfar, fovx, fovy):
    """getProjectionMatrix (utils/graphics_utils.py:51-71), float32 torch (4,4)."""
    th_y = math.tan(fovy / 2)
    th_x = math.tan(fovx / 2)
    top, right = th_y * znear, th_x * znear
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


class OracleCamera:
    """Same fields as the reference Camera (scene/cameras.py:17-57) that the
    renderer reads: FoVx, FoVy, image_width, image_height,
    world_view_transform, full_proj_transform, camera_center (all float32, CPU)."""

    def __init__(self, R, T, fovx, fovy, width, height, znear=0.01, zfar=100.0):
        self.R, self.T = np.asarray(R), np.asarray(T)
        self.FoVx, self.FoVy = float(fovx), float(fovy)
        self.image_width, self.image_height = int(width), int(height)
        self.znear, self.zfar = znear, zfar
        self.world_view_transform = torch.tensor(world_to_view(self.R, self.T)).transpose(0, 1)
        self.projection_matrix = projection(znear, zfar, self.FoVx, self.FoVy).transpose(0, 1)
        self.full_proj_transform = (self.world_view_transform.unsqueeze(0)
                                    .bmm(self.projection_matrix.unsqueeze(0))).squeeze(0)
        self.camera_center = self.world_view_transform.inverse()[3, :3]


def look_at_R(position, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """Camera-to-world rotation (COLMAP axes: x right, y down, z forward) for a
    camera at ``position`` looking at ``target``; the R argument of
    getWorld2View2 is this matrix (it stores R^T as the world->view rotation).
    A camera at (0, 0, -d) gets R = I, matching ``front_camera``."""
    p = np.asarray(position, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - p
    z /= np.linalg.norm(z)
    u = np.asarray(up, dtype=np.float64)
    x = np.cross(u, z)
    if np.linalg.norm(x) < 1e-8:
        x = np.cross(np.array([1.0, 0.0, 0.0]), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=1)


def orbit_camera(position, fov_deg, width, height):
    """Pinhole camera at ``position`` looking at the origin; fovx = fov_deg,
    fovy from the aspect ratio at the same focal length."""
    R = look_at_R(position)
    T = -R.T @ np.asarray(position, dtype=np.float64)
    fovx = math.radians(fov_deg)
    focal = fov2focal(fovx, width)
    fovy = 2 * math.atan(height / (2 * focal))
    return OracleCamera(R, T, fovx, fovy, width, height)


def front_camera(width, height, fov_deg=50.0, distance=4.0):
    """The benchmark camera: R = I, T = (0, 0, distance) — centre at
    (0, 0, -distance) looking along +z at the origin, FoVx = FoVy = fov_deg."""
    f = math.radians(fov_deg)
    return OracleCamera(np.eye(3), np.array([0.0, 0.0, distance]), f, f, width, height)
