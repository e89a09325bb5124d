"""Drop-in ``render`` / ``render_dyn`` (gaussian_renderer/__init__.py:18-101,
104-193), same signatures and return dicts, backed by the gfx950 rasterizer.

``pipe`` only needs the attributes ``PipelineParams`` defines
(arguments/__init__.py:64-69): ``convert_SHs_python``, ``compute_cov3D_python``
and ``debug``.  The model protocol is the reference's (get_xyz, get_opacity,
get_scaling, get_rotation, get_features, get_covariance, active_sh_degree,
max_sh_degree, scene/gaussian_model.py:95-118).  The Python SH path applies
``sigmoid`` to the SH value, as the reference adapter does
(gaussian_renderer/__init__.py:74-78); the kernel path adds 0.5 and clamps.
"""
import math

import torch

from .rasterizer import GaussianRasterizationSettings, GaussianRasterizer
from .sh_utils import eval_sh


class PipelineParams:
    """Defaults of the reference's PipelineParams (arguments/__init__.py:64-69)."""

    def __init__(self, convert_SHs_python=False, compute_cov3D_python=False, debug=False):
        self.convert_SHs_python = convert_SHs_python
        self.compute_cov3D_python = compute_cov3D_python
        self.debug = debug


def _settings(cam, bg_color, sh_degree, scaling_modifier, debug):
    dev = bg_color.device
    return GaussianRasterizationSettings(
        image_height=int(cam.image_height), image_width=int(cam.image_width),
        tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5), bg=bg_color,
        scale_modifier=scaling_modifier, viewmatrix=cam.world_view_transform.to(dev),
        projmatrix=cam.full_proj_transform.to(dev), sh_degree=sh_degree,
        campos=cam.camera_center.to(dev), prefiltered=False, debug=debug)


def _screenspace(xyz):
    pts = torch.zeros_like(xyz, dtype=xyz.dtype, requires_grad=True, device=xyz.device)
    try:
        pts.retain_grad()
    except Exception:
        pass
    return pts


def _python_colors(features_view, xyz, cam, active_sh_degree):
    dirs = xyz - cam.camera_center.to(xyz.device).repeat(features_view.shape[0], 1)
    dirs = dirs / dirs.norm(dim=1, keepdim=True)
    return torch.sigmoid(eval_sh(active_sh_degree, features_view, dirs))


def render(viewpoint_camera, pc, pipe, bg_color: torch.Tensor, scaling_modifier=1.0, override_color=None):
    """Render the scene (gaussian_renderer/__init__.py:18-101)."""
    screenspace_points = _screenspace(pc.get_xyz)
    rs = _settings(viewpoint_camera, bg_color, pc.active_sh_degree, scaling_modifier, pipe.debug)
    rasterizer = GaussianRasterizer(raster_settings=rs)
    means3D = pc.get_xyz
    opacity = pc.get_opacity
    scales = rotations = cov3D_precomp = None
    if pipe.compute_cov3D_python:
        cov3D_precomp = pc.get_covariance(scaling_modifier)
    else:
        scales = pc.get_scaling
        rotations = pc.get_rotation
    shs = colors_precomp = None
    if override_color is None:
        if pipe.convert_SHs_python:
            feats = pc.get_features
            shs_view = feats.transpose(1, 2).view(-1, 3, (pc.max_sh_degree + 1) ** 2)
            colors_precomp = _python_colors(shs_view, pc.get_xyz, viewpoint_camera, pc.active_sh_degree)
        else:
            shs = pc.get_features
    else:
        colors_precomp = override_color
    rendered_image, radii = rasterizer(means3D=means3D, means2D=screenspace_points, shs=shs,
                                       colors_precomp=colors_precomp, opacities=opacity, scales=scales,
                                       rotations=rotations, cov3D_precomp=cov3D_precomp)
    return {"render": rendered_image, "viewspace_points": screenspace_points, "visibility_filter": radii > 0,
            "radii": radii}


def render_dyn(viewpoint_camera, gs_info: dict, pipe, bg_color: torch.Tensor, scaling_modifier=1.0,
               override_color=None):
    """Render dynamically generated Gaussians (gaussian_renderer/__init__.py:104-193).
    Requires ``pipe.compute_cov3D_python`` like the reference (NotImplementedError otherwise)."""
    xyz = gs_info['xyz']
    opacity = gs_info['opacity']
    covariance = gs_info['covariance']
    features = gs_info['features']
    active_sh_degree = gs_info['active_sh_degree']
    max_sh_degree = gs_info['max_sh_degree']
    screenspace_points = _screenspace(xyz)
    rs = _settings(viewpoint_camera, bg_color, active_sh_degree, scaling_modifier, pipe.debug)
    rasterizer = GaussianRasterizer(raster_settings=rs)
    if not pipe.compute_cov3D_python:
        raise NotImplementedError
    cov3D_precomp = covariance
    shs = colors_precomp = None
    if override_color is None:
        if pipe.convert_SHs_python:
            shs_view = features.view(-1, 3, (max_sh_degree + 1) ** 2)
            colors_precomp = _python_colors(shs_view, xyz, viewpoint_camera, active_sh_degree)
        else:
            shs = features
    else:
        colors_precomp = override_color
    rendered_image, radii = rasterizer(means3D=xyz, means2D=screenspace_points, shs=shs,
                                       colors_precomp=colors_precomp, opacities=opacity, scales=None,
                                       rotations=None, cov3D_precomp=cov3D_precomp)
    return {"render": rendered_image, "viewspace_points": screenspace_points, "visibility_filter": radii > 0,
            "radii": radii}
