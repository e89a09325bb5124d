"""Pins of the oracle's SH basis and camera matrices (CPU)."""
import math

import numpy as np
import torch
from scipy import special

from dmgs_amd import cameras as PC
from dmgs_amd import sh_utils as PS
from oracle import camera as OC
from oracle import sh as OS


def _scipy_real_sh(l, m, d):
    """Real SH from scipy's complex Y_l^m (Condon-Shortley phase included):
    sqrt(2) Re Y_l^m for m > 0, sqrt(2) Im Y_l^|m| for m < 0, Y_l^0 for m = 0."""
    theta = np.arccos(np.clip(d[:, 2], -1, 1))
    phi = np.arctan2(d[:, 1], d[:, 0])
    Y = special.sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return np.sqrt(2) * Y.real
    if m < 0:
        return np.sqrt(2) * Y.imag
    return Y.real


def test_sh_basis_matches_scipy():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(200, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    B = OS.sh_basis(3, torch.tensor(d)).numpy()
    for l in range(4):
        for m in range(-l, l + 1):
            k = l * l + l + m
            np.testing.assert_allclose(B[:, k], _scipy_real_sh(l, m, d), rtol=1e-10, atol=1e-12)


def test_sh_special_cases():
    # degree 0 is the constant C0; along +z only the m = 0 terms survive
    d = torch.tensor([[0.0, 0.0, 1.0]], dtype=torch.float64)
    B = OS.sh_basis(3, d)[0]
    assert abs(B[0] - 0.28209479177387814) < 1e-15
    nonzero = {k for k in range(16) if abs(B[k]) > 1e-12}
    assert nonzero == {0, 2, 6, 12}


def test_product_eval_sh_matches_oracle():
    g = torch.Generator().manual_seed(1)
    sh = torch.randn(50, 3, 16, generator=g, dtype=torch.float64)
    d = torch.randn(50, 3, generator=g, dtype=torch.float64)
    d = d / d.norm(dim=1, keepdim=True)
    for deg in range(4):
        np.testing.assert_allclose(PS.eval_sh(deg, sh, d).numpy(), OS.eval_sh(deg, sh, d).numpy(), rtol=1e-12,
                                   atol=1e-12)
    assert torch.allclose(PS.SH2RGB(PS.RGB2SH(torch.tensor(0.3))), torch.tensor(0.3))


def _ndc(cam, p):
    ph = torch.cat([torch.as_tensor(p, dtype=torch.float32), torch.ones(1)]) @ cam.full_proj_transform
    return (ph[:3] / ph[3]).numpy()


def test_camera_known_answers():
    cam = OC.front_camera(200, 100, fov_deg=90.0, distance=4.0)
    # the camera centre is (0, 0, -4) and maps to the view-space origin
    np.testing.assert_allclose(cam.camera_center.numpy(), [0, 0, -4], atol=1e-6)
    pv = torch.tensor([0.0, 0.0, -4.0, 1.0]) @ cam.world_view_transform
    np.testing.assert_allclose(pv[:3].numpy(), 0, atol=1e-6)
    # FoV 90 deg: focal = W / 2
    assert abs(OC.fov2focal(math.pi / 2, 200) - 100.0) < 1e-9
    # frustum corners at depth z map to NDC x, y = +-1, and near/far to z = 0 / 1
    z = 3.0
    tx, ty = math.tan(cam.FoVx / 2) * z, math.tan(cam.FoVy / 2) * z
    for sx in (-1, 1):
        for sy in (-1, 1):
            n = _ndc(cam, [sx * tx, sy * ty, -4.0 + z])
            np.testing.assert_allclose(n[:2], [sx, sy], atol=1e-5)
    assert abs(_ndc(cam, [0, 0, -4.0 + 0.01])[2]) < 1e-4
    assert abs(_ndc(cam, [0, 0, -4.0 + 100.0])[2] - 1.0) < 1e-4


def test_product_cameras_match_oracle():
    oc = OC.orbit_camera((1.0, -2.0, 3.5), 50.0, 320, 180)
    pc = PC.camera_at((1.0, -2.0, 3.5), 320, 180, fov_deg=50.0, device="cpu")
    for a in ("world_view_transform", "full_proj_transform", "camera_center"):
        np.testing.assert_allclose(getattr(oc, a).numpy(), getattr(pc, a).numpy(), rtol=1e-6, atol=1e-6)
    assert abs(oc.FoVy - pc.FoVy) < 1e-12
    # the orbit camera looks at the origin: it projects to the image centre
    np.testing.assert_allclose(_ndc(oc, [0, 0, 0])[:2], 0, atol=1e-5)
    f = PC.front_camera(64, 64, device="cpu")
    o = OC.front_camera(64, 64)
    np.testing.assert_allclose(f.full_proj_transform.numpy(), o.full_proj_transform.numpy())
