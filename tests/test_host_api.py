"""Host-side behaviour of the drop-in API that needs no GPU: argument errors,
settings fields, pipeline defaults, scene generators."""
import pytest
import torch

from dmgs_amd import gaussian_renderer as GR
from dmgs_amd import rasterizer as RZ
from dmgs_amd import scenes as S


def _settings():
    return RZ.GaussianRasterizationSettings(
        image_height=8, image_width=8, tanfovx=0.5, tanfovy=0.5, bg=torch.ones(3), scale_modifier=1.0,
        viewmatrix=torch.eye(4), projmatrix=torch.eye(4), sh_degree=0, campos=torch.zeros(3), prefiltered=False,
        debug=False)


def test_settings_fields_match_reference():
    assert RZ.GaussianRasterizationSettings._fields == (
        "image_height", "image_width", "tanfovx", "tanfovy", "bg", "scale_modifier", "viewmatrix", "projmatrix",
        "sh_degree", "campos", "prefiltered", "debug")


def test_rasterizer_argument_errors():
    r = RZ.GaussianRasterizer(_settings())
    m = torch.zeros(4, 3)
    o = torch.ones(4, 1)
    with pytest.raises(Exception, match="SHs or precomputed colors"):
        r(means3D=m, means2D=m, opacities=o, scales=m, rotations=torch.zeros(4, 4))
    with pytest.raises(Exception, match="SHs or precomputed colors"):
        r(means3D=m, means2D=m, opacities=o, shs=torch.zeros(4, 1, 3), colors_precomp=m, scales=m,
          rotations=torch.zeros(4, 4))
    with pytest.raises(Exception, match="scale/rotation pair"):
        r(means3D=m, means2D=m, opacities=o, shs=torch.zeros(4, 1, 3), scales=m)
    with pytest.raises(Exception, match="scale/rotation pair"):
        r(means3D=m, means2D=m, opacities=o, shs=torch.zeros(4, 1, 3), scales=m, rotations=torch.zeros(4, 4),
          cov3D_precomp=torch.zeros(4, 6))


def test_render_dyn_requires_python_covariance():
    cam = type("Cam", (), {})()
    cam.image_height = cam.image_width = 8
    cam.FoVx = cam.FoVy = 1.0
    cam.world_view_transform = cam.full_proj_transform = torch.eye(4)
    cam.camera_center = torch.zeros(3)
    info = dict(xyz=torch.zeros(2, 3), opacity=torch.ones(2, 1), covariance=torch.zeros(2, 6),
                features=torch.zeros(2, 3, 16), active_sh_degree=0, max_sh_degree=3)
    with pytest.raises(NotImplementedError):
        GR.render_dyn(cam, info, GR.PipelineParams(compute_cov3D_python=False), torch.ones(3))


def test_pipeline_defaults():
    p = GR.PipelineParams()
    assert (p.convert_SHs_python, p.compute_cov3D_python, p.debug) == (False, False, False)


def test_scene_generators():
    m = S.free_scene(10000, seed=0, layout="mixture", shell_radius=3.0)
    r = m.get_xyz.norm(dim=1)
    assert int((r <= 1.0 + 1e-6).sum()) == 6000 and torch.allclose(r[6000:], torch.full((4000,), 3.0))
    assert m.get_features.shape == (10000, 16, 3)
    assert torch.allclose(m.get_rotation.norm(dim=1), torch.ones(10000), atol=1e-6)
    assert 0 < m.get_opacity.min() and m.get_opacity.max() < 1
    assert abs(m._scaling.mean().item() + 3.5) < 0.02 and abs(m._scaling.std().item() - 0.4) < 0.02
    u = S.free_scene(1000, seed=1, layout="uniform")
    assert u.get_xyz.abs().max() <= 1.0
    # the same seed gives the same bytes
    assert torch.equal(S.free_scene(100, seed=5).get_xyz, S.free_scene(100, seed=5).get_xyz)
    v, f = S.icosphere(2)
    assert f.shape == (320, 3) and v.shape == (162, 3)
    assert torch.allclose(v.norm(dim=1), torch.ones(162), atol=1e-6)
    # closed manifold: every edge is shared by exactly two faces
    e = torch.cat([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]).sort(1).values
    _, counts = torch.unique(e, dim=0, return_counts=True)
    assert (counts == 2).all()
    # the model covariance equals the oracle's R S S^T R^T for the normalised rotation
    from oracle import raster as OR
    c = m.get_covariance()[:50].double()
    ref = OR.cov3d_from_scale_rot(m.get_scaling[:50].double(), m.get_rotation[:50].double(), 1.0)
    assert torch.allclose(c, ref, rtol=1e-5, atol=1e-9)
